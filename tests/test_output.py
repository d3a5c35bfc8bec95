"""Output formats and the driver loop, on the CPU.

The driver is exercised with a stand-in solver whose psi field is a known
function of the iteration count, so the snapshot times, the breakthrough
iteration and the checkpoint cadence are fixed by arithmetic.
"""
import json

import numpy as np
import pytest

from fingering_dynamics_amd import output
from fingering_dynamics_amd.params import LbmParams


class FakeCompute:
    """psi = -1 everywhere except the probe column (W-1-OUTLET_MARGIN), where
    it rises linearly and crosses 0 at iteration t_bt; rows 0 and 1 are
    solid there.  The outlet column holds +1 (never looked at)."""

    def __init__(self, H=6, W=12, t_bt=730):
        self.H, self.W, self.t_bt = H, W, t_bt
        self.c = W - 1 - output.OUTLET_MARGIN
        self.mask = np.ones((H, W), dtype=bool)
        self.mask[:2, self.c] = False
        self.links = np.zeros((H, W), dtype=np.uint8)
        self.params = LbmParams.periodic(H, W)
        self.t = 0
        self.loaded = None
        self.fields_calls = 0
        self.window_calls = 0

    def _psi(self):
        psi = -np.ones((self.H, self.W), dtype=np.float32)
        psi[:, self.c] = (self.t - self.t_bt) / 1000.0
        psi[0, self.c] = 5.0                        # solid node: never counts
        psi[:, self.W - 1] = 1.0
        return psi

    def step(self, n=1):
        assert n >= 1
        self.t += n

    def fields(self):
        self.fields_calls += 1
        z = np.zeros((self.H, self.W), dtype=np.float32)
        return self._psi(), z, z, z

    def fields_window(self, x0, x1):
        self.window_calls += 1
        p = self._psi()[:, x0:x1]
        z = np.zeros_like(p)
        return p, z, z, z

    def mass(self):
        return float(self.H * self.W), float(self._psi().sum()), 0

    def last_step_ms(self):
        return 0.25

    def synchronize(self):
        pass

    def save_state(self):
        return np.arange(40, dtype=np.uint8) ^ np.uint8(self.t & 0xFF)

    def load_state(self, buf):
        self.loaded = buf


def test_series_round_trip(tmp_path):
    p = LbmParams.periodic(4, 5, M=7.0)
    s = output.SnapshotSeries(4, 5, p)
    rng = np.random.default_rng(0)
    frames = [rng.standard_normal((4, 5)).astype(np.float32) for _ in range(3)]
    for t, fr in zip((0, 50, 100), frames):
        s.append(t, fr)
    path = str(tmp_path / "series.npz")
    s.save(path)
    r = output.SnapshotSeries.load(path)
    assert np.array_equal(r.frames, np.stack(frames))
    assert r.steps.tolist() == [0, 50, 100]
    assert r.params == p
    with np.load(path, allow_pickle=False) as z:       # plain arrays, no pickle
        assert z["frames"].dtype == np.float32 and z["frames"].shape == (3, 4, 5)


def test_series_rejects_bad_frames():
    s = output.SnapshotSeries(3, 3)
    s.append(10, np.zeros((3, 3)))
    with pytest.raises(ValueError):
        s.append(5, np.zeros((3, 3)))
    with pytest.raises(ValueError):
        s.append(20, np.zeros((3, 4)))


def test_checkpoint_file_round_trip(tmp_path):
    cm = FakeCompute()
    cm.links[2, 3] = 0x5A
    cm.step(37)
    path = str(tmp_path / "ck.npz")
    output.save_checkpoint(cm, path)
    state, mask, links, params = output.read_checkpoint(path)
    assert np.array_equal(state, cm.save_state())
    assert np.array_equal(mask, cm.mask) and np.array_equal(links, cm.links)
    assert params == cm.params


def test_checkpoint_needs_a_step(tmp_path):
    with pytest.raises(ValueError):
        output.save_checkpoint(FakeCompute(), str(tmp_path / "ck.npz"))


def test_probe_column():
    assert output.probe_column(48) == 48 - 1 - output.OUTLET_MARGIN
    assert output.probe_column(48, 47) == 47
    with pytest.raises(ValueError):
        output.probe_column(5)
    with pytest.raises(ValueError):
        output.probe_column(48, 48)


def test_outlet_hits():
    psi = np.array([0.5, -0.2, 0.0, 1e-3, 0.9])
    fluid = np.array([False, True, True, True, True])
    assert output.outlet_hits(psi, fluid).tolist() == [3, 4]
    assert output.outlet_hits(psi, fluid, threshold=0.5).tolist() == [4]
    assert output.outlet_hits(psi, np.zeros(5, bool)).size == 0


def test_metrics_log(tmp_path):
    path = str(tmp_path / "m.jsonl")
    log = output.MetricsLog(path)
    log.write(t=1, a=0.5)
    log.write(t=2, b=[1, 2])
    log.close()
    assert output.MetricsLog.read(path) == [{"a": 0.5, "t": 1}, {"b": [1, 2], "t": 2}]
    for line in open(path):
        json.loads(line)


def test_run_snapshots_breakthrough_and_checkpoints(tmp_path):
    cm = FakeCompute(t_bt=730)
    series = output.SnapshotSeries(cm.H, cm.W)
    log = output.MetricsLog(str(tmp_path / "m.jsonl"))
    ck = str(tmp_path / "ck.npz")
    res = output.run(cm, 1000, mabiki=250, series=series, check_every=100, metrics=log,
                     checkpoint_every=400, checkpoint_path=ck)
    log.close()
    # frames at every multiple of the interval, the final iteration included
    assert series.steps.tolist() == [0, 250, 500, 750, 1000]
    ref = FakeCompute(t_bt=730)
    ref.t = 500
    assert np.array_equal(series.frames[2], ref._psi())
    # psi at the outlet crosses 0 at 730: first positive check is 800
    # (the solid node's psi = 5 never counts)
    assert res["breakthrough_t"] == 800 and res["t"] == 1000
    recs = output.MetricsLog.read(str(tmp_path / "m.jsonl"))
    assert [r["t"] for r in recs] == list(range(100, 1001, 100))
    assert [r["breakthrough"] for r in recs] == [False] * 7 + [True] * 3
    # the last checkpoint is the one at 800
    state, _, _, _ = output.read_checkpoint(ck)
    probe = FakeCompute()
    probe.t = 800
    assert np.array_equal(state, probe.save_state())


def test_run_stops_at_breakthrough():
    cm = FakeCompute(t_bt=730)
    res = output.run(cm, 5000, check_every=50, stop_at_breakthrough=True)
    assert res["breakthrough_t"] == 750 and res["t"] == 750 and cm.t == 750


def test_run_argument_checks():
    with pytest.raises(ValueError):
        output.run(FakeCompute(), 10, series=output.SnapshotSeries(6, 12))
    with pytest.raises(ValueError):
        output.run(FakeCompute(), 10, checkpoint_every=5)
