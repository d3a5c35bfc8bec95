"""Drop-in for ``Compute`` of lattice_boltzmann/fingering_periodic.py.

The reference driver does

    cm = Compute(mask)                                   # :452
    for i in range(MAX_T):                               # :454-479
        <collide>; stream(cm.f, cm.g); bb.halfway_bounceback_...(...)
        cm.zou_he_boundary_inlet(); cm.zou_he_boundary_outlet(); <moments>
    plt.pcolor(x, y, cm.psi)                             # :486-487

With this class the loop body is one call, ``cm.step(MAX_T)``, that runs
the fused gfx950 kernel through libfdlbm.so; ``cm.psi``, ``cm.rho``,
``cm.ux``, ``cm.uy``, ``cm.f`` and ``cm.g`` read the state back with the
reference's shapes (psi, f, g full; rho, ux, uy compressed over the fluid
mask in C order).

There is no CPU fallback: constructing a Compute without the built library
or without a gfx950 device raises.
"""
import numpy as np

from . import _native
from .params import LbmParams


def _to_fd_params(p: LbmParams) -> _native.FdParams:
    return _native.FdParams(
        H=p.H, W=p.W, tau=p.tau, eta_n=p.eta_n, M=p.M, a=p.a, kappa=p.kappa, gamma=p.gamma,
        psi_wall=p.psi_wall, u0=p.u0, rho_out=p.rho_out, outlet_f3_coef=p.outlet_f3_coef,
        gx=p.gx, gy=p.gy, y_bc=p.y_bc, x_bc=p.x_bc, profile=p.profile, outlet=p.outlet,
        init_mode=p.init_mode, wall_lo=p.wall_lo, wall_hi=p.wall_hi_row, psi_visc_bound=p.psi_visc_bound)


def initial_psi(params: LbmParams, mask):
    """psi = -1, +1 in the first n_inv_cols columns, psi_wall on solids
    (fingering_periodic.py:90-93)."""
    psi = np.full((params.H, params.W), -1.0)
    psi[:, :params.n_inv_cols] = 1.0
    psi[~np.asarray(mask, dtype=bool)] = params.psi_wall
    return psi


class Compute:
    """GPU solver state for one H x W domain.

    mask    bool[H][W], True on fluid nodes (fingering_periodic.py:450-451)
    links   uint8[H][W] bounce-back bits (geometry.Bounce_back.link_bits_*)
    params  LbmParams
    psi0    optional initial psi [H][W] (default: initial_psi)
    rho0    optional initial rho [H][W] (default: params.rho0 on fluid)
    """

    def __init__(self, mask, links, params: LbmParams, psi0=None, rho0=None):
        self._ensembles = 0
        self._h = None
        self._lib = _native.load()
        self.params = params
        self.mask = np.ascontiguousarray(mask, dtype=bool)
        H, W = self.mask.shape
        if (H, W) != (params.H, params.W):
            raise ValueError("mask shape %s != params (%d, %d)" % (self.mask.shape, params.H, params.W))
        self.H, self.W = H, W
        solid = np.ascontiguousarray(~self.mask, dtype=np.uint8)
        bits = np.ascontiguousarray(links, dtype=np.uint8)
        if bits.shape != (H, W):
            raise ValueError("links must be uint8[H][W]")
        self.links = bits
        self._fdp = _to_fd_params(params)
        h = _native.ctypes.c_void_p()
        _native.check(self._lib.fd_create(_native.ctypes.byref(self._fdp), _native.ptr(solid),
                                          _native.ptr(bits), _native.ctypes.byref(h)))
        self._h = h
        self._tstream = None
        if psi0 is None:
            psi0 = initial_psi(params, self.mask)
        if rho0 is None:
            rho0 = np.full((H, W), float(params.rho0))
        self.init(psi0, rho0)

    # ------------------------------------------------------------ control
    def init(self, psi0, rho0=None):
        psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
        rho0 = None if rho0 is None else np.ascontiguousarray(rho0, dtype=np.float64)
        for name, a in (("psi0", psi0), ("rho0", rho0)):
            if a is not None and a.shape != (self.H, self.W):
                raise ValueError("%s must be [H][W] = (%d, %d), got %s" % (name, self.H, self.W, a.shape))
        _native.check(self._lib.fd_init(self._h, _native.ptr(psi0), _native.ptr(rho0)))
        self._cache = None

    def init_fields(self, psi, rho, ux, uy):
        """Restart from fields [H][W] (fd_init_fields): the flow keeps its
        density and velocity, the phase field is replaced; t restarts at 0."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (psi, rho, ux, uy)]
        for a in arrs:
            if a.shape != (self.H, self.W):
                raise ValueError("fields must be [H][W]")
        _native.check(self._lib.fd_init_fields(self._h, *[_native.ptr(a) for a in arrs]))
        self._cache = None

    def step(self, n=1):
        """Run n iterations of fingering_periodic.py:455-479 (asynchronous)."""
        _native.check(self._lib.fd_step(self._h, int(n)))
        self._cache = None

    @property
    def solid_inert(self):
        """True when solid nodes are neither read nor written (leak-free geometry)."""
        return bool(self._lib.fd_solid_inert(self._h))

    @property
    def t(self):
        return int(self._lib.fd_steps_done(self._h))

    def synchronize(self):
        _native.check(self._lib.fd_synchronize(self._h))

    def last_step_ms(self):
        return float(self._lib.fd_last_step_ms(self._h))

    def set_tile(self, waves_x, rows):
        """Launch shape: workgroups of waves_x (1, 2, 4 or 8) waves, each wave
        marching `rows` rows of its own 64-column strip."""
        _native.check(self._lib.fd_set_tile(self._h, int(waves_x), int(rows)))

    def use_graph(self, on=True):
        """Replay long step loops from a captured HIP graph, or launch every
        iteration eagerly (the default); identical results."""
        _native.check(self._lib.fd_use_graph(self._h, 1 if on else 0))

    def set_fusion(self, depth, waves_x=0, rows=0):
        """Iterations per launch: 3 or 2 keep the intermediate states of
        each group of iterations on chip; 1 runs one launch per iteration.
        Identical results.  waves_x / rows (> 0) set that depth's launch
        shape.  The default depends on the domain size (``fusion``)."""
        _native.check(self._lib.fd_set_fusion(self._h, int(depth), int(waves_x), int(rows)))

    @property
    def fusion(self):
        """(iterations per launch, waves per workgroup, rows per wave) that
        step() uses now."""
        d, w, r = (_native.ctypes.c_int32() for _ in range(3))
        _native.check(self._lib.fd_get_fusion(self._h, _native.ctypes.byref(d), _native.ctypes.byref(w),
                                              _native.ctypes.byref(r)))
        return d.value, w.value, r.value

    TUNE_CANDIDATES = tuple((3, 1, r) for r in (16, 20, 26, 39)) + \
        tuple((2, 1, r) for r in (1, 2, 4, 8, 12, 16, 24)) + ((2, 4, 16),) + \
        tuple((1, wx, r) for wx, r in ((1, 1), (4, 1), (4, 4), (4, 16)))

    def autotune(self, candidates=TUNE_CANDIDATES, steps=20):
        """Time ``steps`` real iterations per launch configuration
        (depth, waves, rows) and keep the fastest; the results do not depend
        on it, so the trial iterations are part of the run.  Returns
        {(depth, waves, rows): ms per iteration}."""
        times = {}
        steps -= steps % 6
        steps = max(steps, 6)
        for depth, wx, rows in candidates:
            if depth == 1:
                self.set_fusion(1)
                self.set_tile(wx, rows)
            else:
                self.set_fusion(depth, wx, rows)
            self.step(6)
            self.step(steps)
            times[(depth, wx, rows)] = self.last_step_ms()
        best = min(times, key=times.get)
        if best[0] == 1:
            self.set_fusion(1)
            self.set_tile(*best[1:])
        else:
            self.set_fusion(*best)
        return times

    # ------------------------------------------------------------ outputs
    def fields(self):
        """(psi, rho, ux, uy) as float32 [H][W] (rho/u are 0 on solid nodes)."""
        if self._cache is None:
            out = [np.empty((self.H, self.W), dtype=np.float32) for _ in range(4)]
            _native.check(self._lib.fd_get_fields(self._h, *[_native.ptr(a) for a in out]))
            self._cache = tuple(out)
        return self._cache

    def fields_window(self, x0, x1):
        """(psi, rho, ux, uy) of columns [x0, x1) as float32 [H][x1-x0]; only
        the strips holding the window are evaluated."""
        w = int(x1) - int(x0)
        out = [np.empty((self.H, w), dtype=np.float32) for _ in range(4)]
        _native.check(self._lib.fd_get_fields_window(self._h, int(x0), int(x1), *[_native.ptr(a) for a in out]))
        return tuple(out)

    @property
    def psi(self):
        return self.fields()[0].astype(np.float64)

    @property
    def rho(self):
        return self.fields()[1][self.mask].astype(np.float64)

    @property
    def ux(self):
        return self.fields()[2][self.mask].astype(np.float64)

    @property
    def uy(self):
        return self.fields()[3][self.mask].astype(np.float64)

    def populations(self):
        f = np.empty((9, self.H, self.W), dtype=np.float32)
        g = np.empty((9, self.H, self.W), dtype=np.float32)
        _native.check(self._lib.fd_get_populations(self._h, _native.ptr(f), _native.ptr(g)))
        return f, g

    def set_populations(self, f, g):
        """Replace the state by post-stream populations f, g [9][H][W] of the
        current iteration (fd_set_populations): the next iteration collides
        exactly these.  Stored in FP32."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (f, g)]
        for name, a in zip("fg", arrs):
            if a.shape != (9, self.H, self.W):
                raise ValueError("%s must be [9][H][W] = (9, %d, %d), got %s" % (name, self.H, self.W, a.shape))
        _native.check(self._lib.fd_set_populations(self._h, _native.ptr(arrs[0]), _native.ptr(arrs[1])))
        self._cache = None

    @property
    def f(self):
        return self.populations()[0].astype(np.float64)

    @f.setter
    def f(self, value):
        # cm.f = ... as in the reference's in-place writes (fingering_periodic.py:459)
        self.set_populations(value, self.populations()[1])

    @property
    def g(self):
        return self.populations()[1].astype(np.float64)

    @g.setter
    def g(self, value):
        self.set_populations(self.populations()[0], value)

    def mass(self):
        """FP64 sums over fluid nodes: (sum rho, sum psi, non-finite count)."""
        out = np.zeros(3, dtype=np.float64)
        _native.check(self._lib.fd_phase_mass(self._h, _native.ptr(out)))
        return float(out[0]), float(out[1]), int(out[2])

    def watchdog(self):
        """(non-finite fluid nodes, min over fluid nodes of tau_mix - 1/2),
        FP64 (fd_watchdog).  Healthy: 0 and a positive minimum."""
        out = np.zeros(2, dtype=np.float64)
        _native.check(self._lib.fd_watchdog(self._h, _native.ptr(out)))
        return int(out[0]), float(out[1])

    # --------------------------------------------------------- checkpoint
    def save_state(self):
        n = int(self._lib.fd_state_bytes(self._h))
        buf = np.empty(n, dtype=np.uint8)
        _native.check(self._lib.fd_save_state(self._h, _native.ptr(buf), n))
        return buf

    def load_state(self, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        _native.check(self._lib.fd_load_state(self._h, _native.ptr(buf), buf.size))
        self._cache = None

    # -------------------------------------------------- device arrays
    def torch_stream(self):
        """The handle's stream (fd_stream) as a torch stream.  Torch work on
        the device arrays below goes inside
        ``with torch.cuda.stream(cm.torch_stream()):`` to stay ordered with
        the solver's launches (and its allocations with their use)."""
        import torch
        if self._tstream is None:
            self._tstream = torch.cuda.ExternalStream(self._lib.fd_stream(self._h))
        return self._tstream

    def fields_device(self, out=None):
        """(psi, rho, ux, uy) of the last completed iteration as a float32
        [4][H][W] device tensor, enqueued on the handle's stream
        (fd_fields_device).  `out` is reused when given."""
        import torch
        if out is None:
            with torch.cuda.stream(self.torch_stream()):
                out = torch.empty((4, self.H, self.W), dtype=torch.float32, device="cuda")
        self._check_dev(out[0])
        _native.check(self._lib.fd_fields_device(self._h, *[out[i].data_ptr() for i in range(4)]))
        return out

    def init_fields_device(self, psi, rho, ux, uy):
        """init_fields from float32 [H][W] device tensors (fd_init_fields_device)."""
        for a in (psi, rho, ux, uy):
            self._check_dev(a)
        _native.check(self._lib.fd_init_fields_device(self._h, psi.data_ptr(), rho.data_ptr(), ux.data_ptr(),
                                                      uy.data_ptr()))
        self._cache = None

    def save_state_device(self, buf=None):
        """(device blob, iteration) of the current state (fd_save_state_device)."""
        import torch
        n = int(self._lib.fd_state_device_bytes(self._h))
        if buf is None or buf.numel() < n:
            with torch.cuda.stream(self.torch_stream()):
                buf = torch.empty(n, dtype=torch.uint8, device="cuda")
        _native.check(self._lib.fd_save_state_device(self._h, buf.data_ptr(), n))
        return buf, self.t

    def load_state_device(self, state):
        buf, t = state
        _native.check(self._lib.fd_load_state_device(self._h, buf.data_ptr(), buf.numel(), int(t)))
        self._cache = None

    def _check_dev(self, a):
        import torch
        if (not isinstance(a, torch.Tensor) or a.dtype != torch.float32 or not a.is_cuda
                or tuple(a.shape) != (self.H, self.W) or not a.is_contiguous()):
            raise ValueError("expected a contiguous float32 [H][W] device tensor")

    def close(self):
        if getattr(self, "_ensembles", 0):
            raise RuntimeError("this solver belongs to an open Ensemble; close the ensemble first")
        if getattr(self, "_h", None):
            self._lib.fd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
