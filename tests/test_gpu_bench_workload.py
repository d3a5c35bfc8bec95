"""GPU correctness of the benchmarked workload itself: walls on rows 0 and
H-1, constant Zou-He inlet, pressure outlet, seeded circles r 8-24 at
porosity 0.75 with the complete half-way link set (inert solids), psi_wall
of the 60-degree calibration, M = 10 (bench.workload; reference:
fingering.py:298-390, 432-451, 573 and fingering_periodic.py:454-479).

1. A 512 x 256 copy built by the same generator against the FP64 numpy
   oracle at iterations 1, 2, 5, 20, 50 (tolerances of test_gpu_parity).
2. The full 4096 x 2048 domain against the FP64 C oracle at iterations 1, 5
   and 10, fields and populations.
3. The full domain for 1,000 iterations: every node finite, tau_mix > 1/2
   everywhere (fd_watchdog), density drift small.
4. A 400 x 100 copy for 2,000 iterations against the FP64 C oracle, with
   the same algorithm in FP32 as the yardstick (as tests/test_gpu_trajectory)
   and the invading front position; then on to 4,000 iterations, by which
   the scheme itself has diverged at this size (the FP64 oracle goes
   non-finite between 3,000 and 3,500 as psi overshoots towards the
   tau_mix pole at -11/9): the watchdog must report it.
"""
import numpy as np
import pytest

import bench
from fingering_dynamics_amd import geometry
from oracle.lbm64c import Lbm64C
from tests.helpers import oracle_for, max_err
from tests.test_gpu_parity import _compare, TOL_POP, TOL_U

pytestmark = pytest.mark.gpu


def _compute(*a, **k):
    from fingering_dynamics_amd.lbm import Compute
    return Compute(*a, **k)


def test_bench_workload_copy_512x256():
    mask, bits, params, _ = bench.workload(256, 512)
    cm = _compute(mask, bits, params)
    assert cm.solid_inert
    o = oracle_for(params, mask, geometry.unpack_links(bits))
    _compare(cm, o, mask, "t=0")
    done = 0
    for t in (1, 2, 5, 20, 50):
        cm.step(t - done)
        o.step(t - done)
        done = t
        _compare(cm, o, mask, "t=%d" % t)
    r_gpu, p_gpu, bad = cm.mass()
    r_o, p_o = o.mass()
    assert bad == 0
    assert abs(r_gpu - r_o) < 1e-6 * abs(r_o)
    assert abs(p_gpu - p_o) < 1e-6 * mask.sum()
    cm.close()


def test_bench_workload_full_size_against_fp64():
    mask, bits, params, _ = bench.workload()
    cm = _compute(mask, bits, params)
    o = Lbm64C.from_params(params, mask, geometry.unpack_links(bits))
    for t in (1, 5, 10):
        cm.step(t - cm.t)
        o.step(t - o.t)
        gp, op = cm.fields(), o.fields()
        e = dict(psi=max_err(gp[0], op[0]), rho=max_err(gp[1], op[1], mask),
                 ux=max_err(gp[2], op[2], mask), uy=max_err(gp[3], op[3], mask))
        print("t=%d" % t, {k: "%.2e" % v for k, v in e.items()}, flush=True)
        assert e["psi"] < TOL_POP and e["rho"] < TOL_POP, (t, e)
        assert e["ux"] < TOL_U and e["uy"] < TOL_U, (t, e)
    f, g = cm.populations()
    of, og = o.populations()
    ef = float(np.abs(f[:, mask] - of[:, mask]).max())
    eg = float(np.abs(g[:, mask] - og[:, mask]).max())
    print("t=10 populations f %.2e g %.2e" % (ef, eg))
    assert ef < TOL_POP and eg < TOL_POP
    assert not f[:, ~mask].any() and not g[:, ~mask].any()
    r, p, bad = cm.mass()
    ro, po = o.mass()
    assert bad == 0 and abs(r - ro) < 1e-7 * abs(ro) and abs(p - po) < 1e-6 * mask.sum()
    cm.close()


def test_bench_workload_full_size_1000_iterations_healthy():
    mask, bits, params, _ = bench.workload()
    cm = _compute(mask, bits, params)
    r0, p0, _ = cm.mass()
    cm.step(1000)
    r1, p1, bad = cm.mass()
    nbad, tex = cm.watchdog()
    psi = cm.fields()[0]
    print("after 1000: sum rho %.6e (%.2e rel), sum psi %.6e, non-finite %d, min tau-1/2 %.4f, psi [%.4f, %.4f]"
          % (r1, (r1 - r0) / r0, p1, bad, tex, psi[mask].min(), psi[mask].max()))
    assert bad == 0 and nbad == 0
    assert tex > 0
    assert abs(r1 - r0) < 1e-3 * r0
    assert np.isfinite(psi).all() and psi[mask].min() > -1.2 and psi[mask].max() < 1.2
    cm.close()


def _front(psi, mask):
    """Mean over rows of the furthest column with psi > 0 on a fluid node."""
    cols = np.where((psi > 0) & mask, np.arange(psi.shape[1])[None, :], -1)
    return float(cols.max(axis=1).mean())


def _errs(a, b, m):
    d = np.abs(a[m] - b[m])
    return float(d.max()), float(np.sqrt(np.mean(d ** 2)))


def test_bench_family_trajectory_400x100():
    mask, bits, p, _ = bench.workload(100, 400)
    links = geometry.unpack_links(bits)
    cm = _compute(mask, bits, p)
    o64 = Lbm64C.from_params(p, mask, links)
    o32 = Lbm64C.from_params(p, mask, links, bits=32)
    for t in range(250, 2001, 250):
        cm.step(t - cm.t)
        o64.step(t - o64.t)
        o32.step(t - o32.t)
        gp, fa, fb = cm.fields(), o64.fields(), o32.fields()
        parts = []
        for k, i in (("psi", 0), ("ux", 2), ("uy", 3)):
            m = np.ones_like(mask) if k == "psi" else mask
            g_inf, g_l2 = _errs(gp[i], fa[i], m)
            r_inf, r_l2 = _errs(fb[i], fa[i], m)
            parts.append("%s gpu %.1e/%.1e fp32 %.1e/%.1e" % (k, g_inf, g_l2, r_inf, r_l2))
            assert g_inf <= 4 * r_inf + (1e-5 if k == "psi" else 1e-7), (t, k, g_inf, r_inf)
            assert g_l2 <= 4 * r_l2 + (1e-6 if k == "psi" else 1e-8), (t, k, g_l2, r_l2)
        fg, f64, f32 = _front(gp[0], mask), _front(fa[0], mask), _front(fb[0], mask)
        assert abs(fg - f64) <= max(0.5, 4 * abs(f32 - f64)), (t, fg, f64, f32)
        print("t=%d" % t, "; ".join(parts), "; front gpu %.2f fp64 %.2f fp32 %.2f" % (fg, f64, f32),
              "; watchdog", cm.watchdog(), flush=True)
        assert cm.watchdog()[1] > 0
    # on to 4,000: the scheme diverges at this size, and the watchdog says so
    cm.step(4000 - cm.t)
    o64.step(4000 - o64.t)
    fa = o64.fields()
    bad, tex = cm.watchdog()
    print("t=4000 gpu watchdog (%d, %g); fp64 oracle finite: %s" % (bad, tex, bool(np.isfinite(fa[0]).all())))
    assert not np.isfinite(fa[0]).all()
    assert bad > 0 or tex <= 0
    cm.close()
