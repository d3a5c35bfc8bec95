"""The per-node arithmetic the gfx950 kernel runs (csrc/lbm_node.h), built
for the host in FP64 (libfdnode.so), against the FP64 oracle on random
states.  Agreement to ~1e-13 shows the kernel's algebraic rearrangements
(e.g. the closed form of tau_mix, the folded weights of the Zou-He unknowns)
are the reference's formulas."""
import ctypes
import os

import numpy as np
import pytest

from oracle import lbm64, params64
from oracle.lbm64 import E

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "fingering_dynamics_amd", "_lib", "libfdnode.so")


@pytest.fixture(scope="module")
def node():
    if not os.path.exists(LIB):
        from fingering_dynamics_amd import build
        build.build_node_host()
    lib = ctypes.CDLL(LIB)
    d = ctypes.POINTER(ctypes.c_double)
    lib.fdn_stencil_f64.argtypes = [d, d]
    lib.fdn_moments_f64.argtypes = [d, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, d, d]
    lib.fdn_equilibria_f64.argtypes = [d, d, d, d]
    lib.fdn_collide_f64.argtypes = [d, d, d, d]
    lib.fdn_inlet_f64.argtypes = [d, d, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.fdn_outlet_f64.argtypes = [d, d, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double]
    lib.fdn_inlet_corner_f64.argtypes = [d, d, ctypes.c_double, ctypes.c_int]
    lib.fdn_inlet_rho_f64.argtypes = [d, ctypes.c_double, ctypes.c_double]
    lib.fdn_inlet_rho_f64.restype = ctypes.c_double
    return lib


def P(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def consts(c, gx=0.0, gy=0.0, psi_wall=-0.5):
    return np.array([1.0 / c["tau"], c["eta_n"], c["M"], c["a"], c["kappa"], c["gamma"], psi_wall, gx, gy])


def oracle_node(rng, M=20.0, gx=0.0, gy=0.0):
    """A 5 x 5 all-fluid oracle with random post-stream populations and psi;
    returns it after its moments stage (the state the collision uses)."""
    c = params64.derived_constants(M=M)
    o = lbm64.Lbm64(np.ones((5, 5), bool), np.zeros((9, 5, 5), bool), tau=c["tau"], eta_n=c["eta_n"],
                    M=M, kappa=c["kappa"], a=c["a"], gamma=c["gamma"], u0=c["u0"], psi_wall=-0.5,
                    body_force=(gx, gy), x_periodic=True)
    o.f = lbm64.WGT[:, None, None] * (1 + 0.05 * rng.standard_normal((9, 5, 5)))
    o.g = 0.05 * rng.standard_normal((9, 5, 5))
    o.g[0] += rng.uniform(-1, 1, (5, 5))
    o.moments()
    return o, c


@pytest.mark.parametrize("M,gx,gy", [(20.0, 0.0, 0.0), (5.0, 0.0, 0.0), (10.0, 3e-4, -2e-4)])
def test_moments_equilibria_collision(node, M, gx, gy):
    rng = np.random.default_rng(int(M))
    o, c = oracle_node(rng, M, gx, gy)
    k = consts(c, gx, gy)
    f_post = o.f.copy()
    g_post = o.g.copy()
    o.collide()
    for n, (y, x) in enumerate(np.ndindex(5, 5)):
        f = np.ascontiguousarray(f_post[:, y, x])
        g = np.ascontiguousarray(g_post[:, y, x])
        mom = np.zeros(9)
        node.fdn_moments_f64(P(f), o.psi[y, x], o.nabla_psix[y, x], o.nabla_psiy[y, x], o.nabla_psi2[y, x],
                             P(k), P(mom))
        rho, psi, mu, ux, uy, p, itau = mom[:7]
        assert np.isclose(rho, o.rho[n], rtol=0, atol=1e-14)
        assert np.isclose(mu, o.mu[n], rtol=0, atol=1e-14)
        assert np.isclose(ux, o.ux[n], rtol=0, atol=1e-14)
        assert np.isclose(uy, o.uy[n], rtol=0, atol=1e-14)
        assert np.isclose(p, o.p[n], rtol=0, atol=1e-14)
        assert np.isclose(1.0 / itau, o.mix_tau[n], rtol=1e-13)
        feq = np.zeros(9)
        geq = np.zeros(9)
        node.fdn_equilibria_f64(P(mom), P(k), P(feq), P(geq))
        assert np.allclose(feq, o.feq[:, n], rtol=0, atol=1e-14)
        assert np.allclose(geq, o.geq[:, n], rtol=0, atol=1e-14)
        node.fdn_collide_f64(P(f), P(g), P(mom), P(k))
        assert np.allclose(f, o.f[:, y, x], rtol=0, atol=1e-14)
        assert np.allclose(g, o.g[:, y, x], rtol=0, atol=1e-14)


def test_stencil_matches_oracle(node):
    rng = np.random.default_rng(3)
    o, _ = oracle_node(rng)
    o.psi = rng.uniform(-1, 1, (5, 5))
    dx, dy, lap = o.getNabla_psix(), o.getNabla_psiy(), o.getNabla_psi2()
    for y in range(5):
        for x in range(5):
            v = lambda yy, xx: o.psi[yy % 5, xx % 5]
            n = np.array([v(y, x), v(y, x + 1), v(y + 1, x), v(y, x - 1), v(y - 1, x),
                          v(y + 1, x + 1), v(y + 1, x - 1), v(y - 1, x - 1), v(y - 1, x + 1)])
            out = np.zeros(3)
            node.fdn_stencil_f64(P(n), P(out))
            assert np.allclose(out, [dx[y, x], dy[y, x], lap[y, x]], rtol=0, atol=1e-14)


def _zou_oracle(rng, **kw):
    c = params64.derived_constants()
    H, W = 6, 5
    o = lbm64.Lbm64(np.ones((H, W), bool), np.zeros((9, H, W), bool), tau=c["tau"], eta_n=c["eta_n"],
                    M=20.0, kappa=c["kappa"], a=c["a"], gamma=c["gamma"], u0=c["u0"], psi_wall=-0.5, **kw)
    o.f = 0.1 + 0.02 * rng.standard_normal((9, H, W))
    o.g = 0.1 * rng.standard_normal((9, H, W))
    o.psi = rng.uniform(-1, 1, (H, W))
    o.nabla_psi2 = 0.1 * rng.standard_normal((H, W))
    return o


@pytest.mark.parametrize("kw", [dict(), dict(inlet="const"), dict(outlet="pressure", rho_out=0.98),
                                dict(outlet_f3_coef=1.5), dict(y_walls=True, inlet="const")])
def test_zou_he_matches_oracle(node, kw):
    rng = np.random.default_rng(4)
    o = _zou_oracle(rng, **kw)
    f0, g0 = o.f.copy(), o.g.copy()
    Pin = o._zou_he_P(0)
    Pout = o._zou_he_P(-1)
    u = lbm64.inlet_profile(o.H, o.u0, o.inlet)
    o.zou_he_boundary_inlet()
    o.zou_he_boundary_outlet()
    for y in range(o.H):
        f = np.ascontiguousarray(f0[:, y, 0])
        g = np.ascontiguousarray(g0[:, y, 0])
        if o.y_walls and y in (0, o.H - 1):
            r = 1 if y == 0 else o.H - 2
            fr = np.ascontiguousarray(f0[:, r, 0])
            rho_nb = node.fdn_inlet_rho_f64(P(fr), u[r], Pin[0][r])
            node.fdn_inlet_corner_f64(P(f), P(g), rho_nb, 1 if y else 0)
        else:
            node.fdn_inlet_f64(P(f), P(g), u[y], Pin[0][y], Pin[1][y])
        assert np.allclose(f, o.f[:, y, 0], rtol=0, atol=1e-14), y
        assert np.allclose(g, o.g[:, y, 0], rtol=0, atol=1e-14), y
        f = np.ascontiguousarray(f0[:, y, -1])
        g = np.ascontiguousarray(g0[:, y, -1])
        node.fdn_outlet_f64(P(f), P(g), u[y], Pout[0][y], Pout[1][y], 1 if o.outlet == "pressure" else 0,
                            o.rho_out, o.outlet_f3_coef)
        if o.y_walls and y == 0:
            f[2], g[2] = f[4], g[4]
        if o.y_walls and y == o.H - 1:
            f[4], g[4] = f[2], g[2]
        assert np.allclose(f, o.f[:, y, -1], rtol=0, atol=1e-14), y
        assert np.allclose(g, o.g[:, y, -1], rtol=0, atol=1e-14), y


def test_tau_mix_bound(node):
    """With the |psi| bound the interpolation is the reference's inside
    [-1, 1] (bit for bit) and its value at the pure phase outside; without
    it the denominator changes sign at psi = -(1+M)/(M-1)."""
    node.fdn_inv_tau_mix_f64.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                         ctypes.c_double]
    node.fdn_inv_tau_mix_f64.restype = ctypes.c_double
    M, eta = 15.0, 0.1
    c = np.array([1.0, eta, M, -0.05, 0.1, 1.0, -0.5, 0.0, 0.0])
    f = lambda rho, psi, b: node.fdn_inv_tau_mix_f64(rho, psi, P(c), b)
    closed = lambda rho, psi: 1.0 / (6 * M * eta / (rho * ((1 - psi) + M * (1 + psi))) + 0.5)
    for psi in np.linspace(-1.0, 1.0, 41):
        assert f(1.02, psi, 1.0) == f(1.02, psi, 0.0)
        assert abs(f(1.02, psi, 0.0) - closed(1.02, psi)) < 1e-14
    assert f(1.0, -1.3, 1.0) == f(1.0, -1.0, 0.0)
    assert f(1.0, 1.2, 1.0) == f(1.0, 1.0, 0.0)
    sing = -(1 + M) / (M - 1)
    assert f(1.0, sing - 0.01, 0.0) < 0 < f(1.0, sing + 0.01, 0.0)    # 1/tau flips sign past the pole
