"""Lattice constants and boundary selection of the two-phase D2Q9 solver.

The reference keeps these as module globals that ``Compute`` reads at call
time (lattice_boltzmann/fingering_periodic.py:15-44; the walled sibling
lattice_boltzmann/fingering.py:16-50).  Here they are one dataclass that is
copied into the C ``fd_params`` struct of include/fdlbm.h.
"""
import math
from dataclasses import dataclass, replace

Y_PERIODIC, Y_WALLS = 0, 1
X_ZOUHE, X_PERIODIC = 0, 1
PROFILE_GAUSS, PROFILE_CONST = 0, 1
OUTLET_VELOCITY, OUTLET_PRESSURE = 0, 1
INIT_MU0, INIT_WALLED = 0, 1


def unit_conversion(Pe=15.0, C_W=1.5e-7, Ca=2.0 * 7.33e-3, Eta=0.001, R_sigma=0.045,
                    C_rho=1.0e3, rho0=1.0, xi=2.0, gamma_length=20.0):
    """Physical -> lattice units of fingering_periodic.py:20-40.

    ``gamma_length`` is the length scale in the mobility: 20 in
    fingering_periodic.py:40, the channel height H in fingering.py:43.
    Returns a dict with tau, eta_n, sigma, u0, kappa, a, gamma.
    """
    tau = 1.0 / (3.0 - math.sqrt(3.0))          # relaxation time of g (and of the psi=+1 fluid)
    nu0 = (tau - 0.5) / 3.0                     # lattice kinematic viscosity
    R_Nu = Eta / 1000.0                         # physical kinematic viscosity
    C_t = nu0 / R_Nu * C_W * C_W                # time conversion
    eta_n = Eta / (C_rho * C_W * C_W / C_t)     # lattice dynamic viscosity
    sigma = R_sigma * C_t * C_t / (C_rho * C_W ** 3)
    u0 = Ca * sigma / (rho0 * nu0)
    kappa = 0.75 * sigma * xi
    a = -2.0 * kappa / (xi * xi)
    gamma = u0 * gamma_length / ((-a * Pe) * (tau - 0.5))
    return dict(tau=tau, eta_n=eta_n, sigma=sigma, u0=u0, kappa=kappa, a=a, gamma=gamma)


@dataclass
class LbmParams:
    """Everything the kernel needs besides the masks and initial fields."""
    H: int
    W: int
    tau: float
    eta_n: float
    M: float
    a: float
    kappa: float
    gamma: float
    psi_wall: float
    u0: float
    rho0: float = 1.0
    rho_out: float = 1.0
    outlet_f3_coef: float = 2.0 / 3.0
    gx: float = 0.0
    gy: float = 0.0
    y_bc: int = Y_PERIODIC
    x_bc: int = X_ZOUHE
    profile: int = PROFILE_GAUSS
    outlet: int = OUTLET_VELOCITY
    init_mode: int = INIT_MU0
    wall_lo: int = 0
    wall_hi: int = -1          # -1 -> H-1
    n_inv_cols: int = 5        # psi = +1 in the first columns (fingering_periodic.py:91)
    psi_visc_bound: float = 0.0  # > 0: |psi| bound of the tau_mix interpolation (0: reference formula)

    @classmethod
    def periodic(cls, H, W, M=20.0, psi_wall=-0.5, Pe=15.0, **kw):
        """The configuration of fingering_periodic.py:15-44 (periodic y,
        Zou-He velocity inlet/outlet with the Gaussian profile)."""
        c = unit_conversion(Pe=Pe)
        base = dict(H=H, W=W, tau=c["tau"], eta_n=c["eta_n"], M=M, a=c["a"], kappa=c["kappa"],
                    gamma=c["gamma"], psi_wall=psi_wall, u0=c["u0"])
        base.update(kw)
        return cls(**base)

    @classmethod
    def walled(cls, H, W, M=20.0, psi_wall=-1.0, Pe=400.0, shipped_walls=True, **kw):
        """The configuration of fingering.py:16-50: walls at the bottom and
        top, constant inlet velocity, outlet f3 coefficient 1.5 (:378), mu
        and uy evaluated at init (:121-127).  ``shipped_walls`` places the
        walls where fingering.py:573 actually applies them (rows 1 and H-2,
        because its [:, 1:-1] slices rows); False uses rows 0 and H-1."""
        c = unit_conversion(Pe=Pe, Ca=3.0 * 7.33e-3, gamma_length=float(H))
        # fingering.py:39 uses u0 = Ca*sigma/Eta_n, equal to Ca*sigma/(rho0*nu0)
        # because Eta_n == nu0 in this unit system.
        base = dict(H=H, W=W, tau=c["tau"], eta_n=c["eta_n"], M=M, a=c["a"], kappa=c["kappa"],
                    gamma=c["gamma"], psi_wall=psi_wall, u0=c["u0"], y_bc=Y_WALLS,
                    profile=PROFILE_CONST, outlet_f3_coef=1.5, init_mode=INIT_WALLED,
                    wall_lo=1 if shipped_walls else 0, wall_hi=H - 2 if shipped_walls else H - 1)
        base.update(kw)
        return cls(**base)

    def with_(self, **kw):
        return replace(self, **kw)

    @property
    def wall_hi_row(self):
        return self.H - 1 if self.wall_hi < 0 else self.wall_hi
