"""The C restatement of the oracle (oracle/c/lbm64.c) against the numpy
oracle, node for node, in every boundary mode."""
import numpy as np
import pytest

from fingering_dynamics_amd.params import LbmParams, X_PERIODIC, PROFILE_CONST, OUTLET_PRESSURE
from oracle import lbm64
from oracle.lbm64c import Lbm64C
from tests.helpers import rect_domain, circle_domain

MODES = {
    "periodic_gauss": (lambda H, W: LbmParams.periodic(H, W, M=20.0, psi_wall=-0.5), "rect"),
    "periodic_circles": (lambda H, W: LbmParams.periodic(H, W, M=10.0, psi_wall=0.2, n_inv_cols=9), "circle"),
    "const_pressure": (lambda H, W: LbmParams.periodic(H, W, M=5.0, psi_wall=0.0, profile=PROFILE_CONST,
                                                       outlet=OUTLET_PRESSURE, rho_out=1.01), "rect"),
    "walled_shipped": (lambda H, W: LbmParams.walled(H, W, gamma=1.0027), "rect"),
    "walled_edge": (lambda H, W: LbmParams.walled(H, W, shipped_walls=False, psi_wall=-0.5, M=10.0,
                                                  gamma=1.0027), "rect"),
    "xperiodic_force": (lambda H, W: LbmParams.walled(H, W, shipped_walls=False, psi_wall=0.0, M=5.0,
                                                      x_bc=X_PERIODIC, gx=1e-5, gamma=1.0027, n_inv_cols=9), "rect"),
}


@pytest.mark.parametrize("mode", sorted(MODES))
def test_c_oracle_matches_numpy_oracle(mode):
    H, W = 30, 62
    make, kind = MODES[mode]
    if kind == "rect":
        mask, _, links = rect_domain(H, W, 5, size=4, seed=4, x_min=12)
    else:
        mask, _, links = circle_domain(H, W, 3, 5, 0.88, seed=5, x_lo=12)
    p = make(H, W)
    a = lbm64.Lbm64.from_params(p, mask, links)
    b = Lbm64C.from_params(p, mask, links)
    for n in (1, 4, 25):
        a.step(n - a.t)
        b.step(n - b.t)
        psi, rho, ux, uy = b.fields()
        assert np.allclose(psi, a.psi, rtol=0, atol=1e-12), (mode, n)
        assert np.allclose(rho, a.field("rho"), rtol=0, atol=1e-12), (mode, n)
        assert np.allclose(ux, a.field("ux"), rtol=0, atol=1e-13), (mode, n)
        assert np.allclose(uy, a.field("uy"), rtol=0, atol=1e-13), (mode, n)
        f, g = b.populations()
        assert np.allclose(f, a.f, rtol=0, atol=1e-12) and np.allclose(g, a.g, rtol=0, atol=1e-12), (mode, n)
        ma, mb = a.mass(), b.mass()
        assert abs(ma[0] - mb[0]) < 1e-9 and abs(ma[1] - mb[1]) < 1e-9


def test_visc_bound_keeps_overshoot_finite():
    """psi below -(1+M)/(M-1) on a few nodes: the reference formula's tau_mix
    turns negative and the run goes non-finite within a few iterations; with
    the |psi| bound it stays finite, and numpy and C oracles agree."""
    from oracle import lbm64
    from fingering_dynamics_amd.params import LbmParams, PROFILE_CONST, OUTLET_PRESSURE
    H, W = 24, 48
    mask = np.ones((H, W), dtype=bool)
    links = np.zeros((9, H, W), dtype=bool)
    psi0 = np.full((H, W), -1.0)
    psi0[:, :8] = 1.0
    psi0[10:13, 30:33] = -1.25
    out = {}
    for b in (0.0, 1.0):
        p = LbmParams.periodic(H, W, M=15.0, psi_wall=0.0, profile=PROFILE_CONST, outlet=OUTLET_PRESSURE,
                               psi_visc_bound=b)
        o = Lbm64C.from_params(p, mask, links, psi0=psi0)
        with np.errstate(all="ignore"):
            o.step(40)
            psi = o.fields()[0]
        out[b] = bool(np.isfinite(psi).all())
        if b:
            n = lbm64.Lbm64.from_params(p, mask, links, psi0=psi0)
            with np.errstate(all="ignore"):
                for _ in range(40):
                    n.step()
            assert np.abs(n.psi - psi).max() < 1e-9
    assert out == {0.0: False, 1.0: True}
