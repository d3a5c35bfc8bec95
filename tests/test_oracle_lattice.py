"""Pins of the FP64 oracle (oracle/lbm64.py) by closed forms, invariants and
brute force - none of these retypes the oracle's formulas; each is chosen so
a dropped term, a wrong sign or a transposed index fails it."""
import math

import numpy as np
import pytest

from oracle import lbm64, params64
from oracle.lbm64 import E, WGT, OPP

C = params64.derived_constants(M=20.0)


def make(H=6, W=7, mask=None, links=None, **kw):
    mask = np.ones((H, W), dtype=bool) if mask is None else mask
    links = np.zeros((9, H, W), dtype=bool) if links is None else links
    args = dict(tau=C["tau"], eta_n=C["eta_n"], M=C["M"], kappa=C["kappa"], a=C["a"],
                gamma=C["gamma"], u0=C["u0"], psi_wall=-0.5)
    args.update(kw)
    return lbm64.Lbm64(mask, links, **args)


def random_state(o, rng):
    n = o.mask.sum()
    o.rho = 1.0 + 0.1 * rng.standard_normal(n)
    o.ux = 0.05 * rng.standard_normal(n)
    o.uy = 0.05 * rng.standard_normal(n)
    o.mu = 0.1 * rng.standard_normal(n)
    o.psi[o.mask] = rng.uniform(-1, 1, n)
    o.p = o.getP()
    o.mix_tau = o.getMix_tau()
    o.nabla_psix = rng.standard_normal(o.psi.shape)
    o.nabla_psiy = rng.standard_normal(o.psi.shape)


# ------------------------------------------------------------------ lattice
def test_d2q9_moment_identities():
    assert math.isclose(WGT.sum(), 1.0, rel_tol=0, abs_tol=1e-15)
    assert np.allclose((WGT[:, None] * E).sum(0), 0.0, atol=1e-15)
    second = np.einsum("i,ia,ib->ab", WGT, E, E)
    assert np.allclose(second, np.eye(2) / 3.0, atol=1e-15)
    third = np.einsum("i,ia,ib,ic->abc", WGT, E, E, E)
    assert np.allclose(third, 0.0, atol=1e-15)
    for i in range(9):
        assert tuple(E[OPP[i]]) == tuple(-E[i])


def test_constants_closed_form():
    # fingering_periodic.py:15-44 with the reference's default inputs
    assert math.isclose(C["tau"], 1 / (3 - math.sqrt(3)), rel_tol=1e-15)
    assert math.isclose(C["eta_n"], (C["tau"] - 0.5) / 3, rel_tol=1e-12)  # Eta_n == nu0
    assert math.isclose(C["sigma"], 0.0625, rel_tol=1e-12)
    assert math.isclose(C["kappa"], 0.09375, rel_tol=1e-12)
    assert math.isclose(C["a"], -0.046875, rel_tol=1e-12)
    assert math.isclose(C["u0"], 2 * 7.33e-3 * 0.0625 / C["eta_n"], rel_tol=1e-12)
    assert math.isclose(C["gamma"], C["u0"] * 20 / (0.046875 * 15 * (C["tau"] - 0.5)), rel_tol=1e-12)


# --------------------------------------------------------------- equilibria
def test_equilibrium_moments():
    rng = np.random.default_rng(0)
    o = make()
    random_state(o, rng)
    feq = np.array([o.getfeq(i) for i in range(9)])
    geq = np.array([o.getgeq(i) for i in range(9)])
    psi = o.psi[o.mask]
    u = np.stack([o.ux, o.uy])
    assert np.allclose(feq.sum(0), o.rho, atol=1e-13)
    assert np.allclose(np.einsum("in,ia->an", feq, E), o.rho * u, atol=1e-13)
    P = np.einsum("in,ia,ib->abn", feq, E, E)
    for a in range(2):
        for b in range(2):
            exp = o.p * (a == b) + o.rho * u[a] * u[b]
            assert np.allclose(P[a, b], exp, atol=1e-13)
    assert np.allclose(geq.sum(0), psi, atol=1e-13)
    assert np.allclose(np.einsum("in,ia->an", geq, E), psi * u, atol=1e-13)
    G = np.einsum("in,ia,ib->abn", geq, E, E)
    for a in range(2):
        for b in range(2):
            exp = o.gamma * o.mu * (a == b) + psi * u[a] * u[b]
            assert np.allclose(G[a, b], exp, atol=1e-13)


def test_forcing_moments():
    # Guo forcing: sum F = 0, sum F e = (1 - 1/(2 tau)) * mu * grad psi
    rng = np.random.default_rng(1)
    o = make()
    random_state(o, rng)
    F = np.array([o.getLarge_F(i) for i in range(9)])
    fac = 1 - 1 / (2 * o.mix_tau)
    assert np.allclose(F.sum(0), 0.0, atol=1e-14)
    Fe = np.einsum("in,ia->an", F, E)
    assert np.allclose(Fe[0], fac * o.mu * o.nabla_psix[o.mask], atol=1e-14)
    assert np.allclose(Fe[1], fac * o.mu * o.nabla_psiy[o.mask], atol=1e-14)


def test_collision_conserves_and_adds_force():
    """Per fluid node: sum f* = rho, sum g* = psi, and the momentum gains the
    full interfacial force, sum f* e = sum f e + mu grad psi, given that u is
    the half-force-corrected velocity (fingering_periodic.py:126-139)."""
    rng = np.random.default_rng(2)
    H, W = 8, 9
    o = make(H, W, psi0=np.clip(rng.uniform(-1.2, 1.2, (H, W)), -1, 1))
    o.f = o.f + 0.01 * rng.standard_normal(o.f.shape)
    o.g = o.g + 0.01 * rng.standard_normal(o.g.shape)
    o.moments()
    f0 = o.f[:, o.mask].copy()
    g0 = o.g[:, o.mask].copy()
    o.collide()
    f1 = o.f[:, o.mask]
    g1 = o.g[:, o.mask]
    assert np.allclose(f1.sum(0), f0.sum(0), atol=1e-13)
    assert np.allclose(g1.sum(0), g0.sum(0), atol=1e-13)
    force = np.stack([o.mu * o.nabla_psix[o.mask], o.mu * o.nabla_psiy[o.mask]])
    assert np.allclose(np.einsum("in,ia->an", f1, E), np.einsum("in,ia->an", f0, E) + force, atol=1e-13)


def test_tau_mix_limits():
    o = make(3, 4)
    o.rho = np.full(12, 1.0)
    for psi, nu in ((1.0, C["eta_n"]), (-1.0, C["M"] * C["eta_n"])):
        o.psi[:] = psi
        assert np.allclose(o.getMix_tau(), 3 * nu + 0.5, rtol=1e-14)
    assert math.isclose(3 * C["eta_n"] + 0.5, C["tau"], rel_tol=1e-12)  # psi=+1 fluid relaxes with tau


# -------------------------------------------------------------- streaming
def test_stream_matches_brute_force_roll():
    rng = np.random.default_rng(3)
    H, W = 5, 7
    f = rng.standard_normal((9, H, W))
    g = rng.standard_normal((9, H, W))
    f0, g0 = f.copy(), g.copy()
    lbm64.stream(f, g)
    for i in range(9):
        for y in range(H):
            for x in range(W):
                ys, xs = (y - E[i][1]) % H, (x - E[i][0]) % W
                assert f[i, y, x] == f0[i, ys, xs]
                assert g[i, y, x] == g0[i, ys, xs]


# --------------------------------------------------------------- stencils
def _brute_stencil(psi, ghost):
    H, W = psi.shape
    dx = np.zeros_like(psi)
    dy = np.zeros_like(psi)
    lap = np.zeros_like(psi)
    for y in range(H):
        for x in range(W):
            v = lambda yy, xx: ghost(yy, xx)
            E_, W_, N_, S_ = v(y, x + 1), v(y, x - 1), v(y + 1, x), v(y - 1, x)
            NE, NW, SW, SE = v(y + 1, x + 1), v(y + 1, x - 1), v(y - 1, x - 1), v(y - 1, x + 1)
            dx[y, x] = (4 * (E_ - W_) + NE - NW - SW + SE) / 12
            dy[y, x] = (4 * (N_ - S_) + NE + NW - SW - SE) / 12
            lap[y, x] = (-20 * psi[y, x] + 4 * (E_ + W_ + N_ + S_) + NE + NW + SW + SE) / 6
    return dx, dy, lap


@pytest.mark.parametrize("mode", ["periodic", "walled", "xperiodic_walled"])
def test_stencil_ghosts_brute_force(mode):
    rng = np.random.default_rng(4)
    H, W = 6, 8
    mask = rng.random((H, W)) > 0.2
    kw = dict(psi_wall=0.37)
    if mode != "periodic":
        kw["y_walls"] = True
    if mode == "xperiodic_walled":
        kw["x_periodic"] = True
    o = make(H, W, mask=mask, **kw)
    o.psi = rng.uniform(-1, 1, (H, W))
    o.psi[~mask] = 0.37

    def ghost(y, x):
        if o.y_walls and (y < 0 or y >= H):
            return 0.37
        y %= H
        if not o.x_periodic and x < 0:
            return 1.0
        if not o.x_periodic and x >= W:
            return -1.0
        x %= W
        return o.psi[y, x]

    bx, by, bl = _brute_stencil(o.psi, ghost)
    assert np.allclose(o.getNabla_psix(), bx, atol=1e-14)
    assert np.allclose(o.getNabla_psiy(), by, atol=1e-14)
    assert np.allclose(o.getNabla_psi2(), bl, atol=1e-14)


def test_stencil_exact_on_quadratics():
    H, W = 12, 14
    o = make(H, W)
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    inner = (slice(2, H - 2), slice(2, W - 2))
    o.psi = 0.3 * xx - 0.7 * yy + 0.1
    assert np.allclose(o.getNabla_psix()[inner], 0.3, atol=1e-13)
    assert np.allclose(o.getNabla_psiy()[inner], -0.7, atol=1e-13)
    assert np.allclose(o.getNabla_psi2()[inner], 0.0, atol=1e-13)
    o.psi = xx ** 2 + 2 * yy ** 2 - xx * yy
    assert np.allclose(o.getNabla_psix()[inner], (2 * xx - yy)[inner], atol=1e-12)
    assert np.allclose(o.getNabla_psiy()[inner], (4 * yy - xx)[inner], atol=1e-12)
    assert np.allclose(o.getNabla_psi2()[inner], 6.0, atol=1e-12)


def test_chemical_potential_vanishes_on_tanh_interface():
    """mu = a psi (1 - psi^2) - kappa lap psi with a = -2 kappa / xi^2
    (fingering_periodic.py:37-39) vanishes on psi = tanh(x/xi) in the
    continuum; on the lattice (xi = 8 here) only an O(1/xi^2) remainder of
    the two cancelling terms is left."""
    H, W, xi = 4, 120, 8.0
    o = make(H, W)
    o.a = -2.0 * o.kappa / xi ** 2
    x = np.arange(W, dtype=float)
    o.psi = np.tile(np.tanh((x - 60.3) / xi), (H, 1))
    o.nabla_psi2 = o.getNabla_psi2()
    mu = o.getMu_plain()[:, 20:100]
    bulk = np.abs(o.a * o.psi * (1 - o.psi ** 2))[:, 20:100]
    assert bulk.max() > 1e-4
    assert np.abs(mu).max() < 0.01 * bulk.max()


# ----------------------------------------------------------------- Zou-He
def _zouhe_case(rng, **kw):
    H, W = 8, 6
    o = make(H, W, **kw)
    o.f = 0.1 + 0.02 * rng.standard_normal((9, H, W))
    o.g = 0.1 * rng.standard_normal((9, H, W))
    o.psi = rng.uniform(-1, 1, (H, W))
    o.nabla_psi2 = rng.standard_normal((H, W)) * 0.1
    return o


@pytest.mark.parametrize("profile", ["gauss", "const"])
def test_zou_he_inlet_invariants(profile):
    rng = np.random.default_rng(5)
    o = _zouhe_case(rng, inlet=profile)
    Px, Py = o._zou_he_P(0)
    f, g = o.f, o.g
    u = lbm64.inlet_profile(o.H, o.u0, profile)
    rho_in = (f[0, :, 0] + f[2, :, 0] + f[4, :, 0] + 2 * (f[3, :, 0] + f[6, :, 0] + f[7, :, 0]) - Px / 2) / (1 - u)
    o.zou_he_boundary_inlet()
    col_f, col_g = o.f[:, :, 0], o.g[:, :, 0]
    assert np.allclose(col_f.sum(0), rho_in, atol=1e-14)
    assert np.allclose(np.einsum("iy,i->y", col_f, E[:, 0]) + Px / 2, rho_in * u, atol=1e-14)
    assert np.allclose(np.einsum("iy,i->y", col_f, E[:, 1]) + Py / 2, 0.0, atol=1e-14)
    assert np.allclose(col_g.sum(0), 1.0, atol=1e-14)
    # the unknowns g1:g5:g8 keep the lattice-weight ratios 4:1:1
    assert np.allclose(col_g[1], 4 * col_g[5], atol=1e-15) and np.allclose(col_g[5], col_g[8], atol=1e-15)


@pytest.mark.parametrize("c3", [2.0 / 3.0, 1.5])
def test_zou_he_outlet_invariants(c3):
    rng = np.random.default_rng(6)
    o = _zouhe_case(rng, outlet_f3_coef=c3)
    Px, Py = o._zou_he_P(-1)
    f = o.f
    u = lbm64.inlet_profile(o.H, o.u0, "gauss")
    known = f[0, :, -1] + f[2, :, -1] + f[4, :, -1] + 2 * (f[1, :, -1] + f[5, :, -1] + f[8, :, -1]) + Px / 2
    rho = known / (1 + u)
    o.zou_he_boundary_outlet()
    col = o.f[:, :, -1]
    # mass: exact for the 2/3 coefficient; fingering.py's 1.5 leaves (2/3 - 1.5) u rho
    assert np.allclose(col.sum(0), rho + (2.0 / 3.0 - c3) * u * rho, atol=1e-14)
    assert np.allclose(np.einsum("iy,i->y", col, E[:, 0]) + Px / 2, rho * u + (c3 - 2.0 / 3.0) * u * rho, atol=1e-14)
    # transverse momentum: the outlet's P_y terms enter with the opposite sign
    # of the inlet's (fingering_periodic.py:319-323), so sum f e_y = +P_y/2
    assert np.allclose(np.einsum("iy,i->y", col, E[:, 1]) - Py / 2, 0.0, atol=1e-14)
    assert np.allclose(o.g[:, :, -1].sum(0), -1.0, atol=1e-14)


def test_pressure_outlet_fixes_density():
    rng = np.random.default_rng(7)
    o = _zouhe_case(rng, outlet="pressure", rho_out=1.02)
    o.zou_he_boundary_outlet()
    assert np.allclose(o.f[:, :, -1].sum(0), 1.02, atol=1e-14)


def test_walled_inlet_corners():
    """fingering.py:334-364: the corner nodes take the neighbouring row's
    inlet density and make sum g = 1."""
    rng = np.random.default_rng(8)
    o = _zouhe_case(rng, y_walls=True, inlet="const")
    Px, _ = o._zou_he_P(0)
    f = o.f
    u = o.u0
    rho_in = (f[0, :, 0] + f[2, :, 0] + f[4, :, 0] + 2 * (f[3, :, 0] + f[6, :, 0] + f[7, :, 0]) - Px / 2) / (1 - u)
    o.zou_he_boundary_inlet()
    assert math.isclose(o.f[:, 0, 0].sum(), rho_in[1], abs_tol=1e-14)
    assert math.isclose(o.f[:, -1, 0].sum(), rho_in[-2], abs_tol=1e-14)
    assert math.isclose(o.g[:, 0, 0].sum(), 1.0, abs_tol=1e-14)
    assert math.isclose(o.g[:, -1, 0].sum(), 1.0, abs_tol=1e-14)


# ------------------------------------------------------------ whole steps
def test_first_collision_is_identity_with_mu0_init():
    rng = np.random.default_rng(9)
    H, W = 10, 12
    mask = rng.random((H, W)) > 0.15
    o = make(H, W, mask=mask, n_inv_cols=4)
    f0, g0 = o.f.copy(), o.g.copy()
    o.collide()
    assert np.array_equal(o.f, f0) and np.array_equal(o.g, g0)
    # and the init state itself: g0 = psi, g_i = 0, f = w_i rho (p = rho/3)
    assert np.array_equal(g0[0][mask], o.psi[mask]) and not g0[1:].any()
    assert np.allclose(f0[:, mask], WGT[:, None] * 1.0, atol=1e-15)
    assert not f0[:, ~mask].any() and not g0[:, ~mask].any()


def test_bounce_back_uses_same_node_post_collision_opposite():
    rng = np.random.default_rng(10)
    H, W = 5, 6
    links = rng.random((9, H, W)) > 0.7
    links[0] = False
    fb = rng.standard_normal((9, H, W))
    gb = rng.standard_normal((9, H, W))
    f = rng.standard_normal((9, H, W))
    g = rng.standard_normal((9, H, W))
    f1, g1 = f.copy(), g.copy()
    lbm64.bounce_back(links, fb, gb, f1, g1)
    for i in range(1, 9):
        for y in range(H):
            for x in range(W):
                exp_f = fb[OPP[i], y, x] if links[i, y, x] else f[i, y, x]
                assert f1[i, y, x] == exp_f
                exp_g = gb[OPP[i], y, x] if links[i, y, x] else g[i, y, x]
                assert g1[i, y, x] == exp_g


def test_shipped_walls_act_on_rows_1_and_H_minus_2():
    """fingering.py:573 slices the (9, H, W) stacks with [:, 1:-1], i.e.
    rows; the wall copies therefore land on rows 1 and H-2."""
    rng = np.random.default_rng(11)
    H, W = 7, 5
    fb = rng.standard_normal((9, H, W))
    gb = rng.standard_normal((9, H, W))
    f = rng.standard_normal((9, H, W))
    g = rng.standard_normal((9, H, W))
    f1, g1 = f.copy(), g.copy()
    lbm64.bottom_top_wall(fb[:, 1:-1], gb[:, 1:-1], f1[:, 1:-1], g1[:, 1:-1])
    changed = np.argwhere(f1 != f)
    assert set(changed[:, 1]) == {1, H - 2}
    assert np.array_equal(f1[2, 1], fb[4, 1]) and np.array_equal(f1[7, H - 2], fb[5, H - 2])


def test_body_force_channel_poiseuille():
    """Build-original mode: periodic x, half-way walls at rows 0 and H-1,
    single fluid (psi = psi_wall = +1 everywhere, so mu = 0), constant body
    force G.  Steady state: u(y) = G/(2 rho nu) (y + 1/2)(H - 1/2 - y)."""
    H, W = 12, 3
    G = 1e-6
    o = make(H, W, y_walls=True, x_periodic=True, psi_wall=1.0, body_force=(G, 0.0),
             psi0=np.ones((H, W)), M=1.0)
    o.step(6000)
    nu = C["eta_n"]
    y = np.arange(H)
    exact = G / (2 * nu) * (y + 0.5) * (H - 0.5 - y)
    ux = o.field("ux")[:, 1]
    assert np.abs(ux - exact).max() < 0.01 * exact.max()
    assert np.abs(o.field("uy")).max() < 1e-12
