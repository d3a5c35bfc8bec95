"""GPU side of the output formats: the column-window field read, the
breakthrough detector against the FP64 oracle, and checkpoint files."""
import numpy as np
import pytest

from fingering_dynamics_amd import output
from fingering_dynamics_amd.geometry import unpack_links
from fingering_dynamics_amd.params import LbmParams, PROFILE_CONST, OUTLET_PRESSURE, Y_WALLS
from oracle.lbm64c import Lbm64C
from tests.helpers import rect_domain

pytestmark = pytest.mark.gpu


def _compute(*a, **k):
    from fingering_dynamics_amd.lbm import Compute
    return Compute(*a, **k)


@pytest.mark.parametrize("walled", [False, True])
def test_fields_window_matches_full_fields(walled):
    """The window read evaluates only the strips holding [x0, x1); its
    values are the same kernel's, so they equal the full read bit for bit."""
    H, W = 40, 150
    mask, bits, _ = rect_domain(H, W, 12, size=5, seed=4, x_min=20)
    if walled:
        p = LbmParams.walled(H, W, M=10.0, psi_wall=-0.5, gamma=1.0027)
    else:
        p = LbmParams.periodic(H, W, M=10.0, psi_wall=-0.3, n_inv_cols=20)
    cm = _compute(mask, bits, p)
    for t in (0, 1, 17):
        if t:
            cm.step(t - cm.t)
        full = cm.fields()
        for x0, x1 in ((0, 1), (W - 1, W), (30, 100), (63, 65), (0, W)):
            win = cm.fields_window(x0, x1)
            for a, b in zip(win, full):
                assert np.array_equal(a, b[:, x0:x1]), (t, x0, x1)
    with pytest.raises(Exception):
        cm.fields_window(5, 5)
    cm.close()


def test_breakthrough_agrees_with_oracle():
    """A 48 x 32 channel, invading fluid in the first 30 columns: the front
    crosses the probe column (39) at about iteration 2,300 (both).  At
    every 100 iterations the GPU detector must agree with the oracle's psi
    wherever the oracle's column maximum is not within 5e-3 of the
    threshold, and both must report breakthrough by the end."""
    H, W = 32, 48
    mask = np.ones((H, W), dtype=bool)
    bits = np.zeros((H, W), dtype=np.uint8)
    p = LbmParams.periodic(H, W, M=5.0, psi_wall=0.0, n_inv_cols=30)
    cm = _compute(mask, bits, p)
    o = Lbm64C.from_params(p, mask, unpack_links(bits))
    c = output.probe_column(W)
    first_gpu = first_orc = None
    for t in range(100, 4001, 100):
        cm.step(100)
        o.step(100)
        reached, rows = output.breakthrough(cm)
        opsi = o.fields()[0][:, c]
        if abs(opsi.max()) > 5e-3:
            assert reached == bool(opsi.max() > 0), (t, opsi.max())
            assert np.array_equal(rows, np.flatnonzero(opsi > 0)), t
        if reached and first_gpu is None:
            first_gpu = t
        if opsi.max() > 0 and first_orc is None:
            first_orc = t
    print("breakthrough at probe column %d: gpu %s oracle %s" % (c, first_gpu, first_orc))
    assert first_gpu is not None and first_orc is not None
    assert abs(first_gpu - first_orc) <= 100
    cm.close()


def test_checkpoint_file_resume_bit_exact(tmp_path):
    H, W = 34, 90
    mask, bits, _ = rect_domain(H, W, 8, size=4, seed=6, x_min=15)
    p = LbmParams.periodic(H, W, M=10.0, psi_wall=-0.2, n_inv_cols=10, profile=PROFILE_CONST,
                           outlet=OUTLET_PRESSURE, y_bc=Y_WALLS, wall_lo=0, wall_hi=H - 1)
    a = _compute(mask, bits, p)
    a.step(37)
    path = str(tmp_path / "ck.npz")
    output.save_checkpoint(a, path)
    b = output.load_checkpoint(path)
    assert b.t == 37
    a.step(13)
    b.step(13)
    for x, y in zip(a.fields(), b.fields()):
        assert np.array_equal(x, y)
    fa, ga = a.populations()
    fb, gb = b.populations()
    assert np.array_equal(fa, fb) and np.array_equal(ga, gb)
    a.close()
    b.close()


def test_run_with_series_and_breakthrough(tmp_path):
    """The driver on the GPU: frames land at the snapshot iterations and
    equal the fields read at those iterations by a second solver."""
    H, W = 32, 48
    mask = np.ones((H, W), dtype=bool)
    bits = np.zeros((H, W), dtype=np.uint8)
    p = LbmParams.periodic(H, W, M=5.0, psi_wall=0.0, n_inv_cols=30)
    cm = _compute(mask, bits, p)
    series = output.SnapshotSeries(H, W, p)
    log = output.MetricsLog(str(tmp_path / "m.jsonl"))
    res = output.run(cm, 4000, mabiki=1000, series=series, check_every=500, metrics=log,
                     stop_at_breakthrough=True)
    log.close()
    assert res["breakthrough_t"] is not None and res["t"] == res["breakthrough_t"]
    assert series.steps.tolist() == list(range(0, res["t"] + 1, 1000))
    path = str(tmp_path / "s.npz")
    series.save(path)
    back = output.SnapshotSeries.load(path)
    ref = _compute(mask, bits, p)
    for t, fr in zip(back.steps, back.frames):
        if t:
            ref.step(int(t) - ref.t)
        assert np.array_equal(fr, ref.fields()[0]), t
    recs = output.MetricsLog.read(str(tmp_path / "m.jsonl"))
    assert recs[-1]["breakthrough"] and recs[-1]["non_finite"] == 0
    cm.close()
    ref.close()
