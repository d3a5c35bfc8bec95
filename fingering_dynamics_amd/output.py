"""Output and on-disk formats of a run: the psi snapshot series, checkpoint
files, the breakthrough detector, a JSON-lines metrics log, and the driver
loop that ties them to a ``Compute``.

The reference keeps psi frames by appending ``cm.psi`` to an array every
``mabiki`` iterations (lattice_boltzmann/fingering_periodic.py:453, 462-463)
and pickles that array for the movie writer (openmovie.py:14-15).  Here the
series is a compressed ``.npz``: frames float32 [n][H][W], the iteration
each frame was taken at, and the run's parameters as JSON - loadable with
``numpy.load`` without unpickling anything.

Breakthrough is the first iteration at which the invading phase (psi > 0)
reaches the outlet.  The outlet itself cannot show it: the reference's
outlet condition sets the unknown g populations so that psi = -1 on column
W-1 (fingering_periodic.py:312-317), and the columns next to it sit in that
Dirichlet layer (an invading front stalls about three columns short).  So
the detector watches a probe column OUTLET_MARGIN columns upstream of the
outlet by default: any fluid node there with psi above the threshold.  The
check reads only the strip holding that column on the GPU
(``Compute.fields_window``).
"""
import dataclasses
import json
import os
import time

import numpy as np

from .params import LbmParams

SERIES_FORMAT = "fingering_dynamics_amd.psi_series/1"
OUTLET_MARGIN = 8
CHECKPOINT_FORMAT = "fingering_dynamics_amd.checkpoint/1"


def _params_json(params):
    return json.dumps(dataclasses.asdict(params), sort_keys=True) if params is not None else "null"


def _params_from_json(text):
    d = json.loads(text)
    return None if d is None else LbmParams(**d)


class SnapshotSeries:
    """psi frames of one run, [H][W] float32 each, with their iterations."""

    def __init__(self, H, W, params=None):
        self.H, self.W = int(H), int(W)
        self.params = params
        self._frames = []
        self._steps = []

    def append(self, t, psi):
        psi = np.asarray(psi, dtype=np.float32)
        if psi.shape != (self.H, self.W):
            raise ValueError("frame shape %s != (%d, %d)" % (psi.shape, self.H, self.W))
        if self._steps and int(t) <= self._steps[-1]:
            raise ValueError("frames must be appended in increasing iteration order")
        self._frames.append(psi.copy())
        self._steps.append(int(t))

    def __len__(self):
        return len(self._frames)

    @property
    def frames(self):
        """float32 [n][H][W] (the reference's ``cc``)."""
        if not self._frames:
            return np.zeros((0, self.H, self.W), dtype=np.float32)
        return np.stack(self._frames)

    @property
    def steps(self):
        return np.asarray(self._steps, dtype=np.int64)

    def save(self, path):
        np.savez_compressed(path, format=np.array(SERIES_FORMAT), frames=self.frames, steps=self.steps,
                            params=np.array(_params_json(self.params)))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            if str(z["format"]) != SERIES_FORMAT:
                raise ValueError("%s is not a psi series (%s)" % (path, z["format"]))
            frames, steps = z["frames"], z["steps"]
            params = _params_from_json(str(z["params"]))
        s = cls(frames.shape[1], frames.shape[2], params)
        for t, fr in zip(steps, frames):
            s.append(int(t), fr)
        return s


def save_checkpoint(cm, path):
    """Write the solver's resumable state to ``path`` (.npz): the bit-exact
    state blob of fd_save_state plus the masks and parameters needed to
    rebuild the solver.  Requires at least one completed iteration."""
    if cm.t < 1:
        raise ValueError("a checkpoint needs at least one completed iteration (t >= 1)")
    np.savez(path, format=np.array(CHECKPOINT_FORMAT), state=cm.save_state(),
             mask=np.packbits(cm.mask, axis=None), shape=np.array(cm.mask.shape, dtype=np.int64),
             links=cm.links, params=np.array(_params_json(cm.params)))


def read_checkpoint(path):
    """(state blob, mask, links, params) of a checkpoint file."""
    with np.load(path, allow_pickle=False) as z:
        if str(z["format"]) != CHECKPOINT_FORMAT:
            raise ValueError("%s is not a checkpoint (%s)" % (path, z["format"]))
        H, W = (int(v) for v in z["shape"])
        mask = np.unpackbits(z["mask"], count=H * W).reshape(H, W).astype(bool)
        return z["state"].copy(), mask, z["links"].copy(), _params_from_json(str(z["params"]))


def load_checkpoint(path):
    """A new ``Compute`` holding the state saved in ``path``."""
    from .lbm import Compute
    state, mask, links, params = read_checkpoint(path)
    cm = Compute(mask, links, params)
    cm.load_state(state)
    return cm


def probe_column(W, column=None):
    """The breakthrough probe column: ``column`` if given, else W-1-OUTLET_MARGIN."""
    c = W - 1 - OUTLET_MARGIN if column is None else int(column)
    if not 0 <= c < W:
        raise ValueError("probe column %d outside [0, %d)" % (c, W))
    return c


def outlet_hits(psi_col, fluid_col, threshold=0.0):
    """Rows of the probe column where a fluid node holds psi > threshold."""
    psi_col = np.asarray(psi_col).reshape(-1)
    fluid_col = np.asarray(fluid_col, dtype=bool).reshape(-1)
    return np.flatnonzero(fluid_col & (psi_col > threshold))


def breakthrough(cm, threshold=0.0, column=None):
    """(reached, rows): whether any fluid node of the probe column has psi
    above ``threshold`` after the last completed iteration."""
    c = probe_column(cm.W, column)
    psi_col = cm.fields_window(c, c + 1)[0][:, 0]
    rows = outlet_hits(psi_col, cm.mask[:, c], threshold)
    return rows.size > 0, rows


class MetricsLog:
    """Append-only JSON-lines log, one object per record."""

    def __init__(self, path):
        self.path = path
        self._f = open(path, "a", buffering=1)

    def write(self, **rec):
        self._f.write(json.dumps(rec, sort_keys=True) + "\n")

    def close(self):
        if self._f:
            self._f.close()
            self._f = None

    @staticmethod
    def read(path):
        with open(path) as f:
            return [json.loads(line) for line in f if line.strip()]


def _next_event(t, max_t, *periods):
    nxt = max_t
    for p in periods:
        if p:
            nxt = min(nxt, (t // p + 1) * p)
    return nxt


def run(cm, max_t, mabiki=None, series=None, check_every=None, metrics=None, stop_at_breakthrough=False,
        checkpoint_every=None, checkpoint_path=None, threshold=0.0, column=None):
    """The reference's main loop (fingering_periodic.py:454-479) around a
    ``Compute``, up to ``max_t`` completed iterations.

    mabiki            append psi to ``series`` whenever the iteration count is a
                      multiple of it (the reference's snapshot interval; the
                      reference also stores the initial frame twice, here once)
    check_every       every so many iterations: FP64 mass, breakthrough (probe
                      ``column``, default W-1-OUTLET_MARGIN), and a record in
                      ``metrics`` (a MetricsLog)
    stop_at_breakthrough  end the run at the first check that sees breakthrough
    checkpoint_every  write ``checkpoint_path`` every so many iterations

    Returns a dict: t (completed iterations), breakthrough_t (first check
    that saw it, or None), wall_s, ms_per_step (GPU time of the last chunk).
    """
    if series is not None and not mabiki:
        raise ValueError("a snapshot series needs mabiki")
    if checkpoint_every and not checkpoint_path:
        raise ValueError("checkpoint_every needs checkpoint_path")
    t = cm.t
    bt = None
    t0 = time.perf_counter()
    while True:
        if series is not None and t % mabiki == 0 and (len(series) == 0 or series.steps[-1] < t):
            series.append(t, cm.fields()[0])
        if t >= max_t:
            break
        nxt = _next_event(t, max_t, mabiki if series is not None else None, check_every, checkpoint_every)
        cm.step(nxt - t)
        t = nxt
        if check_every and t % check_every == 0:
            reached, rows = breakthrough(cm, threshold, column)
            if reached and bt is None:
                bt = t
            if metrics is not None:
                s_rho, s_psi, bad = cm.mass()
                metrics.write(t=t, sum_rho=s_rho, sum_psi=s_psi, non_finite=bad, breakthrough=bool(reached),
                              outlet_rows=int(rows.size), ms_per_step=cm.last_step_ms(),
                              wall_s=round(time.perf_counter() - t0, 4))
            if reached and stop_at_breakthrough:
                break
        if checkpoint_every and t % checkpoint_every == 0:
            tmp = checkpoint_path + ".tmp.npz"
            save_checkpoint(cm, tmp)
            os.replace(tmp, checkpoint_path if checkpoint_path.endswith(".npz") else checkpoint_path + ".npz")
    cm.synchronize()
    return {"t": t, "breakthrough_t": bt, "wall_s": time.perf_counter() - t0, "ms_per_step": cm.last_step_ms()}
