"""GPU: writable populations (fd_set_populations) and the failure watchdog
(fd_watchdog), through the C ABI.

1. Handing the solver its own populations changes nothing, bit for bit:
   the collision of the staged state equals the step kernel's.
2. An oracle state from the middle of a trajectory (FP64, rounded to FP32
   and given to both) is stepped by the GPU as by the oracle, at the parity
   tolerances of test_gpu_parity - single-step parity from non-equilibrium
   states instead of only from the initial state.
3. The watchdog's min(tau_mix - 1/2) matches the oracle's, and a NaN
   population and a psi = -1.2 patch at M = 20 (past the pole
   -(1+M)/(M-1) of the viscosity interpolation) trip its two counters.
   This is NaN and negative-tau arithmetic in the data; every kernel runs
   to completion.
"""
import numpy as np
import pytest

import bench
from fingering_dynamics_amd import geometry
from fingering_dynamics_amd.params import LbmParams
from tests.helpers import rect_domain, oracle_for, max_err
from tests.test_gpu_parity import _compare, _setup, TOL_POP, TOL_U

pytestmark = pytest.mark.gpu


def _compute(*a, **k):
    from fingering_dynamics_amd.lbm import Compute
    return Compute(*a, **k)


@pytest.mark.parametrize("case", ["periodic_circles", "periodic_cfg1", "walled_pressure"])
def test_set_own_populations_is_bit_exact(case):
    params, mask, bits, links = _setup(case)
    a = _compute(mask, bits, params)
    ref = _compute(mask, bits, params)
    a.step(13)
    ref.step(13)
    f, g = a.populations()
    fields = [x.copy() for x in a.fields()]
    mass = a.mass()
    a.set_populations(f, g)
    assert a.t == 13
    f2, g2 = a.populations()
    assert np.array_equal(f, f2) and np.array_equal(g, g2)
    for x, y in zip(fields, a.fields()):
        assert np.array_equal(x, y)
    # same node values; the FP64 sums are reduced in another block order
    m2 = a.mass()
    assert m2[2] == mass[2] and abs(m2[0] - mass[0]) <= 1e-12 * abs(mass[0])
    assert abs(m2[1] - mass[1]) <= 1e-12 * mask.size
    a.step(7)
    ref.step(7)
    for x, y in zip(ref.fields(), a.fields()):
        assert np.array_equal(x, y)
    fr, gr = ref.populations()
    fa, ga = a.populations()
    assert np.array_equal(fr, fa) and np.array_equal(gr, ga)
    # a checkpoint taken right after the hand-over restores it
    a.set_populations(fa, ga)
    blob = a.save_state()
    a.step(5)
    after = [x.copy() for x in a.fields()]
    a.load_state(blob)
    assert a.t == 20
    a.step(5)
    for x, y in zip(after, a.fields()):
        assert np.array_equal(x, y)
    a.close()
    ref.close()


def _oracle_state_pair(params, mask, bits, links, t0, checkpoints):
    o = oracle_for(params, mask, links)
    o.step(t0)
    f32, g32 = o.f.astype(np.float32), o.g.astype(np.float32)
    o.set_populations(f32.astype(np.float64), g32.astype(np.float64))
    cm = _compute(mask, bits, params)
    cm.set_populations(f32, g32)
    _compare(cm, o, mask, "t0+0")
    done = 0
    for t in checkpoints:
        cm.step(t - done)
        o.step(t - done)
        done = t
        _compare(cm, o, mask, "t0+%d" % t)
    cm.close()


@pytest.mark.parametrize("case", ["periodic_cfg1", "walled_pressure", "visc_bound"])
def test_oracle_mid_trajectory_state(case):
    params, mask, bits, links = _setup(case)
    _oracle_state_pair(params, mask, bits, links, 40, [1, 2, 5, 20])


def test_oracle_mid_trajectory_state_bench_workload():
    """The bench's boundary family (walls, pressure outlet, complete links,
    calibrated psi_wall, M = 10) on a 256 x 512 copy, from iteration 60."""
    mask, bits, params, _ = bench.workload(256, 512)
    _oracle_state_pair(params, mask, bits, geometry.unpack_links(bits), 60, [1, 2, 5, 20])


def test_watchdog_matches_oracle():
    params, mask, bits, links = _setup("periodic_cfg1")
    cm = _compute(mask, bits, params)
    o = oracle_for(params, mask, links)
    for t in (0, 1, 20):
        cm.step(t - cm.t)
        o.step(t - o.t)
        bad, tex = cm.watchdog()
        obad, otex = o.health()
        print("t=%d watchdog gpu (%d, %.6f) oracle (%d, %.6f)" % (t, bad, tex, obad, otex))
        assert bad == obad == 0
        assert tex > 0 and abs(tex - otex) < 1e-5 * otex
    cm.close()


def test_watchdog_trips_on_nan_and_on_the_tau_pole():
    H, W = 34, 70
    mask, bits, links = rect_domain(H, W, 6, size=4, seed=7, x_min=14)
    params = LbmParams.periodic(H, W, M=20.0, psi_wall=-0.3)     # no |psi| bound
    psi0 = -np.tanh((np.arange(W)[None, :] - 8.0) / 2.0) * np.ones((H, 1))
    o = oracle_for(params, mask, links, psi0=psi0)
    o.step(10)
    f, g = o.f.astype(np.float32), o.g.astype(np.float32)
    cm = _compute(mask, bits, params, psi0=psi0)
    cm.set_populations(f, g)
    bad, tex = cm.watchdog()
    assert bad == 0 and tex > 0, (bad, tex)
    # one NaN population on a fluid node: exactly that node is non-finite
    yx = np.argwhere(mask)[len(np.argwhere(mask)) // 2]
    fn = f.copy()
    fn[4, yx[0], yx[1]] = np.nan
    cm.set_populations(fn, g)
    bad, tex = cm.watchdog()
    o.set_populations(fn.astype(np.float64), g.astype(np.float64))
    assert bad == 1 == o.health()[0]
    assert tex > 0
    # psi = -1.2 on a patch: past -21/19, tau_mix - 1/2 < 0 there
    patch = np.zeros_like(mask)
    patch[20:23, 40:44] = True
    patch &= mask
    gp = g.copy()
    gp[:, patch] = 0.0
    gp[0, patch] = -1.2
    cm.set_populations(fn, gp)
    bad, tex = cm.watchdog()
    o.set_populations(fn.astype(np.float64), gp.astype(np.float64))
    obad, otex = o.health()
    print("tripped: gpu (%d, %.4f) oracle (%d, %.4f)" % (bad, tex, obad, otex))
    assert bad == 1 and tex < 0 and otex < 0
    assert abs(tex - otex) < 1e-4 * abs(otex)
    # the solver keeps running on such data (NaN arithmetic, no fault)
    cm.step(3)
    bad, tex = cm.watchdog()
    assert bad >= 1
    cm.close()
