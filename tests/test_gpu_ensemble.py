"""Batched ensembles: batched launches over many domains (two iterations
per launch by default, one when asked or when a member's geometry does not
allow it) must give exactly the states the domains reach when stepped one
by one."""
import numpy as np
import pytest

from fingering_dynamics_amd.params import (LbmParams, X_PERIODIC, PROFILE_CONST, OUTLET_PRESSURE, Y_WALLS)
from tests.helpers import rect_domain, circle_domain

pytestmark = pytest.mark.gpu


def _specs():
    """Five domains of different sizes (some not multiples of 64), boundary
    families, viscosity ratios and wall values."""
    out = []
    H, W = 40, 150
    m, b, _ = rect_domain(H, W, 10, size=5, seed=3, x_min=20)
    out.append((m, b, LbmParams.periodic(H, W, M=5.0, psi_wall=0.0, n_inv_cols=20)))
    H, W = 34, 70
    m, b, _ = rect_domain(H, W, 6, size=4, seed=5, x_min=15)
    out.append((m, b, LbmParams.walled(H, W, M=10.0, psi_wall=-0.5, gamma=1.0027, shipped_walls=False)))
    H, W = 64, 128
    m, b, _ = circle_domain(H, W, 3, 5, 0.88, seed=7, x_lo=20)
    out.append((m, b, LbmParams.periodic(H, W, M=3.0, psi_wall=-0.3, profile=PROFILE_CONST,
                                         outlet=OUTLET_PRESSURE, y_bc=Y_WALLS, wall_lo=0, wall_hi=H - 1,
                                         n_inv_cols=12)))
    H, W = 33, 96
    m, b, _ = rect_domain(H, W, 5, size=4, seed=8, x_min=10)
    out.append((m, b, LbmParams.walled(H, W, M=2.0, psi_wall=-0.2, gamma=1.0027, shipped_walls=False,
                                       x_bc=X_PERIODIC, gx=1e-5)))
    H, W = 20, 300
    m, b, _ = rect_domain(H, W, 12, size=3, seed=9, x_min=30)
    out.append((m, b, LbmParams.periodic(H, W, M=12.0, psi_wall=0.2, n_inv_cols=25)))
    return out


def _solvers():
    from fingering_dynamics_amd.lbm import Compute
    return [Compute(m, b, p) for m, b, p in _specs()]


def _state(cs):
    out = []
    for c in cs:
        f, g = c.populations()
        out.append((c.fields(), f, g))
    return out


def _same(a, b):
    for (fa, pa, ga), (fb, pb, gb) in zip(a, b):
        for x, y in zip(fa, fb):
            if not np.array_equal(x, y):
                return False
        if not (np.array_equal(pa, pb) and np.array_equal(ga, gb)):
            return False
    return True


@pytest.mark.parametrize("depth", [3, 2, 1])
def test_batched_equals_individual(depth):
    from fingering_dynamics_amd.ensemble import Ensemble
    ref = _solvers()
    cs = _solvers()
    ens = Ensemble(cs)
    ens.set_fusion(depth)
    assert ens.depth == depth and ens.fused == (depth > 1)
    done = 0
    # graph replays (even chunks), odd chunks, eager launches, per-handle
    # reads between chunks, and a sequential stretch
    for n, mode in ((1, "graph"), (2, "graph"), (7, "graph"), (4, "eager"), (20, "graph"), (3, "seq"),
                    (9, "graph")):
        if mode == "eager":
            ens.use_graph(False)
        else:
            ens.use_graph(True)
        if mode == "seq":
            ens.step_sequential(n)
        else:
            ens.step(n)
        for r in ref:
            r.step(n)
        done += n
        assert all(c.t == done for c in cs)
        assert _same(_state(cs), _state(ref)), (done, mode)
    ens.close()
    for c in cs + ref:
        c.close()


def test_tile_independent_and_mass():
    from fingering_dynamics_amd.ensemble import Ensemble
    a, b = _solvers(), _solvers()
    ea, eb = Ensemble(a), Ensemble(b)
    ea.set_tile(4, 16)
    eb.set_tile(1, 5)
    ea.step(13)
    eb.step(13)
    assert _same(_state(a), _state(b))
    for c in a:
        s_rho, s_psi, bad = c.mass()
        assert bad == 0 and np.isfinite(s_rho)
    ea.set_fusion(2)
    times = ea.autotune(candidates=((4, 16), (2, 8)), steps=4)
    assert set(times) == {(4, 16), (2, 8)} and all(t > 0 for t in times.values())
    assert ea.fusion[1:] in times
    eb.step(a[0].t - b[0].t)         # the autotuner's trial iterations
    assert _same(_state(a), _state(b))
    # three iterations per launch
    ea.set_fusion(3)
    times = ea.autotune(candidates=((1, 16), (2, 5)), steps=6)
    assert ea.fusion[1:] in times
    eb.step(a[0].t - b[0].t)
    assert _same(_state(a), _state(b))
    # the same with one iteration per launch: the single-iteration shapes
    ea.set_fusion(1)
    times = ea.autotune(candidates=((4, 16), (1, 5)), steps=3)
    assert ea.tile in times
    eb.step(a[0].t - b[0].t)
    assert _same(_state(a), _state(b))
    ea.close()
    eb.close()
    for c in a + b:
        c.close()


def test_misuse_raises():
    from fingering_dynamics_amd.ensemble import Ensemble
    from fingering_dynamics_amd._native import FdError
    cs = _solvers()[:2]
    with pytest.raises(FdError):
        Ensemble([cs[0], cs[0]])
    cs[0].step(3)
    e = Ensemble(cs)
    with pytest.raises(FdError):
        e.step(1)                    # handles at different iterations
    cs[1].step(3)
    e.step(2)
    assert cs[0].t == cs[1].t == 5
    e.close()
    for c in cs:
        c.close()


def test_member_without_two_iteration_geometry():
    """A member with the reference's wall rows 1 and H-2 (fingering.py:573)
    or H < 8 keeps the batch at one iteration per launch; still exact."""
    from fingering_dynamics_amd.ensemble import Ensemble
    from fingering_dynamics_amd.lbm import Compute
    specs = _specs()[:2]
    H, W = 34, 70
    m, b, _ = rect_domain(H, W, 6, size=4, seed=5, x_min=15)
    specs.append((m, b, LbmParams.walled(H, W, M=10.0, psi_wall=-0.5, gamma=1.0027)))
    assert specs[-1][2].wall_lo == 1
    m, b, _ = rect_domain(6, 80, 2, size=2, seed=6, x_min=15)
    specs.append((m, b, LbmParams.periodic(6, 80, M=4.0, psi_wall=0.1, n_inv_cols=10, profile=PROFILE_CONST)))
    for drop in (None, 2, 3):
        use = [sp for i, sp in enumerate(specs) if i != drop]
        cs = [Compute(*sp) for sp in use]
        ref = [Compute(*sp) for sp in use]
        ens = Ensemble(cs)
        assert not ens.fused
        for n in (3, 8):
            ens.step(n)
            ens.step_sequential(1)
            for r in ref:
                r.step(n + 1)
            assert _same(_state(cs), _state(ref)), (drop, n)
        ens.close()
        for c in cs + ref:
            c.close()
    # without both: pairs run batched
    cs = [Compute(*sp) for sp in specs[:2]]
    ens = Ensemble(cs)
    assert ens.fused
    ens.close()
    for c in cs:
        c.close()
