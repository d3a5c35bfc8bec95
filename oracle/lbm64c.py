"""ctypes front-end of the C restatement (oracle/c/lbm64.c) - TEST INFRASTRUCTURE ONLY.

Same constructor arguments and outputs as oracle.lbm64.Lbm64, for long runs.
"""
import ctypes
import os

import numpy as np

_LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")


class _P(ctypes.Structure):
    _fields_ = [("H", ctypes.c_int), ("W", ctypes.c_int)] + \
        [(n, ctypes.c_double) for n in ("tau", "eta_n", "M", "a", "kappa", "gamma", "psi_wall", "u0",
                                        "rho_out", "c3", "gx", "gy", "rho0")] + \
        [(n, ctypes.c_int) for n in ("y_walls", "x_periodic", "gauss", "pressure", "init_walled",
                                     "wall_lo", "wall_hi")] + \
        [("psi_visc_bound", ctypes.c_double)]


_libs = {}


def _load(bits=64):
    if bits not in _libs:
        path = os.path.join(_LIBDIR, "liboracle%d.so" % bits)
        if not os.path.exists(path):
            from oracle import build_oracle
            build_oracle.build()
        lib = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        lib.o64_create.restype = vp
        lib.o64_create.argtypes = [ctypes.POINTER(_P), vp, vp, vp, vp]
        lib.o64_step.argtypes = [vp, ctypes.c_int]
        lib.o64_fields.argtypes = [vp, vp, vp, vp, vp]
        lib.o64_pops.argtypes = [vp, vp, vp]
        lib.o64_mass.argtypes = [vp, vp]
        lib.o64_destroy.argtypes = [vp]
        lib.o64_perturb.argtypes = [vp, ctypes.c_double, ctypes.c_ulonglong]
        lib.o64_round32.argtypes = [vp]
        _libs[bits] = lib
    return _libs[bits]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Lbm64C:
    """FP64 C oracle; ``bits=32`` selects the same code built with float
    arithmetic and storage (the FP32 yardstick)."""

    def __init__(self, mask, links, *, bits=64, tau, eta_n, M, kappa, a, gamma, u0, psi_wall, rho0=1.0, n_inv_cols=5,
                 inlet="gauss", outlet="velocity", outlet_f3_coef=2.0 / 3.0, y_walls=False, x_periodic=False,
                 init="mu0", body_force=(0.0, 0.0), rho_out=1.0, psi0=None, rho_init=None, wall_rows="edge",
                 psi_visc_bound=0.0):
        lib = self._lib = _load(bits)
        H, W = mask.shape
        self.H, self.W = H, W
        self.mask = np.ascontiguousarray(mask, dtype=bool)
        if psi0 is None:
            psi0 = np.full((H, W), -1.0)
            psi0[:, :n_inv_cols] = 1.0
        psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
        rho_init = None if rho_init is None else np.ascontiguousarray(rho_init, dtype=np.float64)
        lo, hi = (0, H - 1) if wall_rows == "edge" else (1, H - 2)
        p = _P(H=H, W=W, tau=tau, eta_n=eta_n, M=M, a=a, kappa=kappa, gamma=gamma, psi_wall=psi_wall, u0=u0,
               rho_out=rho_out, c3=outlet_f3_coef, gx=body_force[0], gy=body_force[1], rho0=rho0,
               y_walls=int(y_walls), x_periodic=int(x_periodic), gauss=int(inlet == "gauss"),
               pressure=int(outlet == "pressure"), init_walled=int(init == "walled"), wall_lo=lo, wall_hi=hi,
               psi_visc_bound=psi_visc_bound)
        fl = np.ascontiguousarray(self.mask, dtype=np.uint8)
        lk = np.ascontiguousarray(links, dtype=np.uint8)
        self._h = lib.o64_create(ctypes.byref(p), _ptr(fl), _ptr(lk), _ptr(psi0),
                                 None if rho_init is None else _ptr(rho_init))
        self.t = 0

    @classmethod
    def from_params(cls, p, mask, links, psi0=None, rho0=None, bits=64):
        wall_hi = p.H - 1 if p.wall_hi < 0 else p.wall_hi
        return cls(mask, links, bits=bits, tau=p.tau, eta_n=p.eta_n, M=p.M, kappa=p.kappa, a=p.a, gamma=p.gamma, u0=p.u0,
                   psi_wall=p.psi_wall, rho0=p.rho0, n_inv_cols=p.n_inv_cols,
                   inlet="const" if p.profile == 1 else "gauss", outlet="pressure" if p.outlet == 1 else "velocity",
                   outlet_f3_coef=p.outlet_f3_coef, y_walls=p.y_bc == 1, x_periodic=p.x_bc == 1,
                   init="walled" if p.init_mode == 1 else "mu0", body_force=(p.gx, p.gy), rho_out=p.rho_out,
                   psi0=psi0, rho_init=rho0,
                   wall_rows="edge" if (p.wall_lo == 0 and wall_hi == p.H - 1) else "shipped",
                   psi_visc_bound=getattr(p, "psi_visc_bound", 0.0))

    def step(self, n=1):
        self._lib.o64_step(self._h, int(n))
        self.t += n

    def perturb(self, eps, seed):
        """Rounding-sized multiplicative noise on every fluid population."""
        self._lib.o64_perturb(self._h, float(eps), int(seed))

    def round32(self):
        """Round the stored populations to FP32."""
        self._lib.o64_round32(self._h)

    def fields(self):
        out = [np.empty((self.H, self.W)) for _ in range(4)]
        self._lib.o64_fields(self._h, *[_ptr(a) for a in out])
        return tuple(out)

    def populations(self):
        f = np.empty((9, self.H, self.W))
        g = np.empty((9, self.H, self.W))
        self._lib.o64_pops(self._h, _ptr(f), _ptr(g))
        return f, g

    def mass(self):
        out = np.zeros(2)
        self._lib.o64_mass(self._h, _ptr(out))
        return float(out[0]), float(out[1])

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.o64_destroy(self._h)
            self._h = None
