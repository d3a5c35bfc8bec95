"""The C ABI: every function include/fdlbm.h declares is exported by the
built libfdlbm.so with matching ctypes declarations, and the product refuses
to run (loudly) where there is no GPU.  No compute calls."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fdlbm.h")
LIB = os.path.join(ROOT, "fingering_dynamics_amd", "_lib", "libfdlbm.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fd_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib_path():
    if not os.path.exists(LIB):
        from fingering_dynamics_amd import build
        build.build_hip()
    return LIB


def test_header_declares_the_expected_surface():
    names = declared_functions()
    for n in ("fd_create", "fd_init", "fd_step", "fd_get_fields", "fd_get_populations",
              "fd_phase_mass", "fd_destroy", "fd_last_error", "fd_save_state", "fd_load_state"):
        assert n in names


def test_every_declared_symbol_is_exported(lib_path):
    lib = ctypes.CDLL(lib_path)
    for n in declared_functions():
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fd_[a-z_0-9]+)", out))
    assert set(declared_functions()) <= exported


def test_python_binding_covers_the_header(lib_path):
    from fingering_dynamics_amd import _native
    assert set(_native.EXPORTS) == set(declared_functions())
    _native.load()


def test_params_struct_layout_matches_header():
    """ctypes mirror of fd_params: same field order and sizes as the C struct."""
    from fingering_dynamics_amd._native import FdParams
    text = open(HEADER).read()
    body = re.search(r"typedef struct fd_params \{(.*?)\} fd_params;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body)
    fields = []
    for line in body.split(";"):
        line = line.strip()
        if not line:
            continue
        ctype, names = line.split(None, 1)
        for nm in names.split(","):
            fields.append((nm.strip(), ctype))
    assert [f[0] for f in FdParams._fields_] == [f[0] for f in fields]
    size = {"int32_t": 4, "double": 8}
    for (name, ct), (pname, pct) in zip(fields, FdParams._fields_):
        assert ctypes.sizeof(pct) == size[ct], name


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="this machine has a GPU")
def test_no_gpu_means_an_error_not_a_fallback(lib_path):
    """Without a GPU the solver raises; nothing runs on a CPU path."""
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import numpy as np\n"
        "from fingering_dynamics_amd.lbm import Compute\n"
        "from fingering_dynamics_amd.params import LbmParams\n"
        "from fingering_dynamics_amd._native import FdError\n"
        "try:\n"
        "    Compute(np.ones((8, 8), bool), np.zeros((8, 8), np.uint8), LbmParams.periodic(8, 8))\n"
        "except FdError as e:\n"
        "    print('raised', e.code); sys.exit(0)\n"
        "print('no error'); sys.exit(1)\n" % ROOT)
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "raised" in p.stdout


SURR_HEADER = os.path.join(ROOT, "include", "fdsurr.h")
SURR_LIB = os.path.join(ROOT, "fingering_dynamics_amd", "_lib", "libfdsurr.so")


def test_surrogate_library_exports_its_header():
    """libfdsurr.so exports every function include/fdsurr.h declares, and
    the binding in surrogate_ops declares exactly those."""
    if not os.path.exists(SURR_LIB):
        from fingering_dynamics_amd import build
        build.build_surrogate()
    text = re.sub(r"/\*.*?\*/", "", open(SURR_HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(fds_[a-z_0-9]+)\s*\(", text)))
    assert "fds_conv3x3_fwd" in names and "fds_conv3x3_bwd_weight" in names
    out = subprocess.run(["nm", "-D", "--defined-only", SURR_LIB], capture_output=True, text=True).stdout
    assert set(names) <= set(re.findall(r"\bT (fds_[a-z_0-9]+)", out))
    from fingering_dynamics_amd import surrogate_ops
    assert set(surrogate_ops.EXPORTS) == set(names)
    surrogate_ops.load()


def test_surrogate_library_rejects_bad_arguments_before_any_launch():
    """Shape and pointer checks of libfdsurr.so return -1 with a message
    before any HIP call, so they run without a GPU.  Every call here is
    rejected; none reaches a kernel."""
    if not os.path.exists(SURR_LIB):
        from fingering_dynamics_amd import build
        build.build_surrogate()
    from fingering_dynamics_amd import surrogate_ops
    lib = surrogate_ops.load()
    err = lambda: lib.fds_last_error().decode()
    bogus = ctypes.c_void_p(4096)
    cases = [
        (lambda: lib.fds_conv3x3_fwd(None, None, None, None, None, 0, 8, 8, 8, 8, 1, None), "bad shape"),
        (lambda: lib.fds_conv3x3_fwd(None, None, None, None, None, 1, 8, 8, 12, 8, 1, None), "Cin must be"),
        (lambda: lib.fds_conv3x3_fwd(None, None, None, None, None, 1, 8, 8, 8, 12, 1, None), "Cout must be"),
        (lambda: lib.fds_conv3x3_fwd(None, None, None, None, None, 1, 8, 8, 8, 8, 1, None), "NULL pointer"),
        (lambda: lib.fds_conv3x3_fwd(None, None, None, None, None, 1 << 20, 1 << 10, 1 << 10, 8, 8, 1, None),
         "too many pixels"),
        (lambda: lib.fds_conv3x3_bwd_data(None, None, None, None, None, 1, 8, 8, 16, 8, None), "NULL pointer"),
        (lambda: lib.fds_conv3x3_bwd_data(None, None, None, None, None, 1, 8, 8, 8, 4, None), "bad shape"),
        (lambda: lib.fds_conv3x3_bwd_weight(None, None, None, None, None, None, 1, 8, 8, 8, 12, None),
         "Cout must be"),
        (lambda: lib.fds_conv3x3_bwd_weight(None, None, None, None, None, None, 1, 8, 8, 8, 8, None),
         "NULL pointer"),
        (lambda: lib.fds_flip_weights(None, None, 8, 8, None), "bad arguments"),
        (lambda: lib.fds_cast_flip(0, None, None, None, None, None, None), "1 to 16 layers"),
        (lambda: lib.fds_cast_flip(17, None, None, None, None, None, None), "1 to 16 layers"),
        (lambda: lib.fds_cast_flip(1, None, None, None, None, None, None), "NULL pointer"),
        # a non-NULL address that no call below dereferences: the shape
        # checks reject it first
        (lambda: lib.fds_up_cat(None, bogus, bogus, 1, 8, 8, 8, 8, None), "NULL pointer"),
        (lambda: lib.fds_up_cat(bogus, bogus, bogus, 1, 7, 8, 8, 8, None), "even H, W"),
        (lambda: lib.fds_up_cat(bogus, bogus, bogus, 1, 8, 8, 12, 8, None), "multiples of 8"),
        (lambda: lib.fds_up_cat_bwd(bogus, bogus, None, 1, 8, 8, 8, 8, None), "NULL pointer"),
        (lambda: lib.fds_up_cat_bwd(bogus, bogus, bogus, 0, 8, 8, 8, 8, None), "even H, W"),
        (lambda: lib.fds_up_cat_bwd(bogus, bogus, bogus, 1, 8, 8, 8, 0, None), "multiples of 8"),
        (lambda: lib.fds_relu_mask(bogus, None, bogus, 8, None), "NULL pointer"),
        (lambda: lib.fds_relu_mask(bogus, bogus, bogus, 12, None), "multiple of 8"),
        (lambda: lib.fds_up_cat_bwd_masked(bogus, bogus, None, bogus, 1, 8, 8, 8, 8, None), "NULL pointer"),
        (lambda: lib.fds_up_cat_bwd_masked(bogus, bogus, bogus, bogus, 1, 8, 9, 8, 8, None), "even H, W"),
        (lambda: lib.fds_pool_skip_bwd(bogus, bogus, 8, None, bogus, 1, 8, 8, 8, None), "NULL pointer"),
        (lambda: lib.fds_pool_skip_bwd(bogus, bogus, 8, bogus, bogus, 1, 8, 7, 8, None), "even H, W"),
        (lambda: lib.fds_pool_skip_bwd(bogus, bogus, 16, bogus, bogus, 1, 8, 8, 12, None), "multiple of 8"),
        (lambda: lib.fds_pool_skip_bwd(bogus, bogus, 12, bogus, bogus, 1, 8, 8, 8, None), "dskip_ld"),
        (lambda: lib.fds_up_cat_bwd(bogus, bogus, None, 1, 8, 8, 8, 8, None), "NULL pointer"),
    ]
    for call, msg in cases:
        assert call() == -1
        assert msg in err(), (msg, err())
    # the weight-gradient scratch holds S slices of Cout x (9 Cin + 1) floats
    for shape in [(64, 128, 256, 8, 32), (64, 32, 64, 128, 128), (1, 8, 8, 8, 8)]:
        n = lib.fds_bwd_weight_scratch(*shape)
        per = shape[4] * (9 * shape[3] + 1)
        assert n >= per and n % per == 0


def test_solver_rejects_bad_parameters_before_touching_the_device(lib_path):
    """fd_create checks its parameters first (-1 and a message, no HIP
    call), so the checks run without a GPU.  Every case is invalid; none
    creates a handle."""
    import numpy as np
    from fingering_dynamics_amd import _native
    from fingering_dynamics_amd.lbm import _to_fd_params
    from fingering_dynamics_amd.params import LbmParams
    lib = _native.load()
    solid = np.zeros(64 * 64, np.uint8)
    links = np.zeros(64 * 64, np.uint8)
    cases = [
        (dict(H=2), "H and W must be >= 3"),
        (dict(W=2), "H and W must be >= 3"),
        (dict(H=8192, W=4096), "domain too large"),
        (dict(y_bc=2), "bad y_bc"),
        (dict(x_bc=-1), "bad x_bc"),
        (dict(profile=2), "bad profile"),
        (dict(outlet=2), "bad outlet"),
        (dict(init_mode=2), "bad init_mode"),
        (dict(H=63, x_bc=0, profile=0), "Gaussian inlet profile needs an even H"),
        (dict(y_bc=1, wall_lo=-1), "wall rows outside"),
        (dict(y_bc=1, wall_hi=64), "wall rows outside"),
        (dict(tau=0.5), "bad tau/eta_n/M"),
        (dict(eta_n=0.0), "bad tau/eta_n/M"),
        (dict(M=float("nan")), "bad tau/eta_n/M"),
    ]
    for change, msg in cases:
        fdp = _to_fd_params(LbmParams.periodic(64, 64))
        for k, v in change.items():
            setattr(fdp, k, v)
        h = ctypes.c_void_p()
        rc = lib.fd_create(ctypes.byref(fdp), _native.ptr(solid), _native.ptr(links), ctypes.byref(h))
        msg_got = lib.fd_last_error().decode()
        assert rc == -1 and msg in msg_got, (change, rc, msg_got)
        assert not h.value
    h = ctypes.c_void_p()
    assert lib.fd_create(None, _native.ptr(solid), _native.ptr(links), ctypes.byref(h)) == -1
    assert "params is NULL" in lib.fd_last_error().decode()
