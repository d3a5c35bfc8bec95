"""Host-side argument and lifetime guards that must hold before any native
call: run without a GPU, with a recording stub in place of the library so a
guard that fails shows up as a recorded call instead of a device access."""
import numpy as np
import pytest
import torch


class _RecordingLib:
    """Stands in for the ctypes library: records every call, returns 0."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def fn(*args):
            self.calls.append(name)
            return 0
        return fn


def _bare_compute(H=8, W=16):
    from fingering_dynamics_amd.lbm import Compute
    c = Compute.__new__(Compute)
    c._lib = _RecordingLib()
    c._h = object()
    c._ensembles = 0
    c.H, c.W = H, W
    c._cache = None
    return c


def test_close_refuses_a_solver_that_belongs_to_an_ensemble():
    c = _bare_compute()
    c._ensembles = 1
    with pytest.raises(RuntimeError, match="Ensemble"):
        c.close()
    assert c._lib.calls == []          # fd_destroy was not reached
    c._ensembles = 0
    c.close()
    assert c._lib.calls == ["fd_destroy"] and c._h is None


def test_ensemble_close_releases_its_members():
    from fingering_dynamics_amd.ensemble import Ensemble
    a, b = _bare_compute(), _bare_compute()
    a._ensembles = b._ensembles = 1
    e = Ensemble.__new__(Ensemble)
    e._lib = _RecordingLib()
    e._h = object()
    e.computes = [a, b]
    e.close()
    assert e._lib.calls == ["fd_ens_destroy"]
    assert a._ensembles == 0 and b._ensembles == 0
    a.close()
    b.close()


def test_init_rejects_fields_of_the_wrong_shape():
    c = _bare_compute(8, 16)
    for psi0, rho0 in ((np.zeros((16, 8)), None), (np.zeros((8, 16)), np.zeros((8, 15))),
                       (np.zeros(128), None)):
        with pytest.raises(ValueError, match="must be \\[H\\]\\[W\\]"):
            c.init(psi0, rho0)
    assert c._lib.calls == []
    c.init(np.zeros((8, 16)), np.ones((8, 16)))
    assert c._lib.calls == ["fd_init"]


def _bf(*shape):
    return torch.zeros(shape, dtype=torch.bfloat16)


def test_convolution_wrappers_check_shapes_before_any_launch():
    from fingering_dynamics_amd import surrogate_ops as so
    x = _bf(2, 8, 8, 16)
    w = _bf(32, 3, 3, 16)
    # forward: bias of the wrong length / dtype, weights of the wrong layout
    with pytest.raises(ValueError, match="bias"):
        so.conv3x3(x, w, torch.zeros(16, dtype=torch.float32))
    with pytest.raises(ValueError, match="bias"):
        so.conv3x3(x, w, torch.zeros(32, dtype=torch.bfloat16))
    with pytest.raises(ValueError, match="w must be"):
        so.conv3x3(x, _bf(32, 3, 3, 8))
    # host tensors of the right shape never reach the kernels
    with pytest.raises(ValueError, match="same GPU device"):
        so.conv3x3(x, w, torch.zeros(32, dtype=torch.float32))
    # input gradient: wt must be [Cin, 3, 3, Cout] bf16 for dy's Cout
    dy = _bf(2, 8, 8, 32)
    with pytest.raises(ValueError, match="wt must be"):
        so.conv3x3_bwd_data(dy, _bf(16, 3, 3, 16))
    with pytest.raises(ValueError, match="wt must be"):
        so.conv3x3_bwd_data(dy, torch.zeros((16, 3, 3, 32), dtype=torch.float32))
    with pytest.raises(ValueError, match="wt must be"):
        so.conv3x3_bwd_data(dy, _bf(16, 3, 32, 3).transpose(2, 3))
    with pytest.raises(ValueError, match="same GPU device"):
        so.conv3x3_bwd_data(dy, _bf(16, 3, 3, 32))
    # weight gradient: dy must match x in N, H, W
    for bad in (_bf(1, 8, 8, 32), _bf(2, 8, 4, 32), _bf(2, 4, 8, 32)):
        with pytest.raises(ValueError, match="N, H, W of x"):
            so.conv3x3_bwd_weight(x, bad)
    with pytest.raises(ValueError, match="same GPU device"):
        so.conv3x3_bwd_weight(x, dy)
    with pytest.raises(ValueError, match="w must be"):
        so.flip_weights(_bf(32, 3, 2, 16))


def test_plumbing_wrappers_check_arguments_before_any_launch():
    """cast_flip, up_cat_bwd's coarse mask and pool_skip_bwd reject bad
    arguments (and host tensors) with ValueError before any native call."""
    from fingering_dynamics_amd import surrogate_ops as so
    w = torch.zeros(32, 3, 3, 16)
    with pytest.raises(ValueError, match="1 to 16"):
        so.cast_flip([])
    with pytest.raises(ValueError, match="1 to 16"):
        so.cast_flip([w] * 17)
    with pytest.raises(ValueError, match="float32"):
        so.cast_flip([w.double()])
    with pytest.raises(ValueError, match="float32"):
        so.cast_flip([torch.zeros(32, 3, 2, 16)])
    with pytest.raises(ValueError, match="same GPU device"):
        so.cast_flip([w])
    dout = _bf(2, 8, 8, 48)
    with pytest.raises(ValueError, match="lo_relu"):
        so.up_cat_bwd(dout, 32, lo_relu=_bf(2, 4, 4, 16))
    with pytest.raises(ValueError, match="same GPU device"):
        so.up_cat_bwd(dout, 32, lo_relu=_bf(2, 4, 4, 32))
    s = _bf(2, 8, 8, 16)
    with pytest.raises(ValueError, match="dskip must be shaped"):
        so.pool_skip_bwd(_bf(2, 4, 4, 16), _bf(2, 8, 8, 8), s)
    with pytest.raises(ValueError, match="dskip must be shaped"):
        so.pool_skip_bwd(_bf(2, 8, 8, 16), s, s)
    with pytest.raises(ValueError, match="contiguous channels"):
        so.pool_skip_bwd(_bf(2, 4, 4, 16), _bf(2, 8, 8, 20)[..., 2:18], s)
    with pytest.raises(ValueError, match="same GPU device"):
        so.pool_skip_bwd(_bf(2, 4, 4, 16), _bf(2, 8, 8, 32)[..., 8:24], s)
