"""Ensembles of independent domains advanced together by batched launches,
and the seeded generator of the many-domain workload.

The reference runs one domain per process (fingering_periodic.py:454-479).
Here any number of ``Compute`` solvers on one GPU - different sizes,
viscosity ratios, wall values, obstacles and boundary conditions - are
batched: one kernel launch walks a device table of the domains and advances
all of them by three iterations (the three-iteration kernel of fd_step, the
default; two with set_fusion(2)) when every domain's geometry allows it, or
by one, and step loops are replayed from a captured HIP graph.  ``step_sequential`` runs the same iterations as
one launch per domain on the same stream, the baseline the batched launch
is measured against; both give bit-identical states.

``autotune`` picks the launch shape by timing a few real iterations per
candidate; the results do not depend on the shape (every node's update is
the same arithmetic whatever the tiling), so the trial iterations are simply
part of the run.

The many-domain workload is periodic in y, as fingering_periodic.py, and
bounds |psi| by 1 inside the viscosity interpolation
(fd_params.psi_visc_bound).  Without the bound the domains with M above
about 12 cross the pole psi = -(1+M)/(M-1) of tau_mix
(fingering_periodic.py:201-208) near the sharp initial interface: 13 of the
256 seeded domains are non-finite within 50 iterations, in the FP64 oracle
as on the GPU.  With it, the domains that still diverge within 10,000
iterations (21 of 256; 38 with half-way walls on rows 0 and H-1) all have
M < 1.8: with both fluids at tau near 0.79 nothing damps the start-up
pressure waves, and the FP64 oracle diverges on them as well.
``Compute.mass()`` counts the non-finite nodes of each domain and
``Compute.watchdog()`` reports min(tau_mix - 1/2).
"""
import math

import numpy as np

from . import _native, geometry, wetting
from .params import LbmParams, PROFILE_CONST, OUTLET_PRESSURE, Y_WALLS

TILE_CANDIDATES = ((4, 16), (4, 8), (4, 32), (8, 16), (2, 16), (1, 16))
FUSED_CANDIDATES = ((1, 8), (1, 6), (1, 12), (1, 16), (2, 8), (4, 8))
FUSED3_CANDIDATES = ((1, 16), (1, 20), (1, 26), (1, 32), (1, 12), (2, 16))


class Ensemble:
    """Batched stepping of several ``Compute`` solvers (same device)."""

    def __init__(self, computes):
        self._lib = _native.load()
        self.computes = list(computes)
        if not self.computes:
            raise ValueError("an ensemble needs at least one solver")
        arr = (_native.ctypes.c_void_p * len(self.computes))(*[c._h.value for c in self.computes])
        h = _native.ctypes.c_void_p()
        _native.check(self._lib.fd_ens_create(arr, len(self.computes), _native.ctypes.byref(h)))
        self._h = h
        self.tile = (4, 16)
        self.fusion = (3, 1, 26)
        self._shapes = {2: (1, 8), 3: (1, 26)}
        # the device table points at every member's buffers: a member may not
        # be closed while the ensemble exists (Compute.close raises)
        for c in self.computes:
            c._ensembles += 1

    @property
    def nodes(self):
        return sum(c.H * c.W for c in self.computes)

    def step(self, n=1):
        """n iterations of every domain, batched launches (asynchronous)."""
        _native.check(self._lib.fd_ens_step(self._h, int(n)))
        self._invalidate()

    def step_sequential(self, n=1):
        """The same iterations as one launch per domain per iteration."""
        _native.check(self._lib.fd_ens_step_sequential(self._h, int(n)))
        self._invalidate()

    def _invalidate(self):
        for c in self.computes:
            c._cache = None

    def set_tile(self, waves_x, rows):
        """Launch shape of the single-iteration batched kernel."""
        _native.check(self._lib.fd_ens_set_tile(self._h, int(waves_x), int(rows)))
        self.tile = (int(waves_x), int(rows))

    def set_fusion(self, depth, waves_x=0, rows=0):
        """Iterations per batched launch (3, 2 or 1) and the launch shape of
        that depth's kernel (<= 0 keeps the current one)."""
        _native.check(self._lib.fd_ens_set_fusion(self._h, int(depth), int(waves_x), int(rows)))
        d = int(depth)
        if d >= 2:
            wx0, r0 = self._shapes[d]
            self._shapes[d] = (int(waves_x) if waves_x > 0 else wx0, int(rows) if rows > 0 else r0)
            self.fusion = (d,) + self._shapes[d]
        else:
            self.fusion = (1,) + self.fusion[1:]

    @property
    def depth(self):
        """Iterations per batched launch that step() uses: the fusion depth
        when every domain's geometry allows the fused kernels, else 1."""
        ok = all(c.H >= 8 and (c.params.y_bc != Y_WALLS or (c.params.wall_lo == 0 and c.params.wall_hi == c.H - 1))
                 for c in self.computes)
        return self.fusion[0] if ok else 1

    @property
    def fused(self):
        """True when groups of iterations run as two- or three-iteration launches."""
        return self.depth >= 2

    def use_graph(self, on=True):
        _native.check(self._lib.fd_ens_use_graph(self._h, 1 if on else 0))

    def synchronize(self):
        _native.check(self._lib.fd_ens_synchronize(self._h))

    def last_step_ms(self):
        return float(self._lib.fd_ens_last_step_ms(self._h))

    def autotune(self, candidates=None, steps=20):
        """Time ``steps`` real iterations per launch shape of the kernel that
        runs (two-iteration when ``fused``) and keep the fastest.  Returns
        {shape: ms per iteration}."""
        depth = self.depth
        if candidates is None:
            candidates = {3: FUSED3_CANDIDATES, 2: FUSED_CANDIDATES}.get(depth, TILE_CANDIDATES)
        steps = max(6, steps - steps % 6)
        times = {}
        for wx, rows in candidates:
            if depth >= 2:
                self.set_fusion(depth, wx, rows)
            else:
                self.set_tile(wx, rows)
            self.step(6)                         # table rebuild and graph capture
            self.step(steps)
            times[(wx, rows)] = self.last_step_ms()
        best = min(times, key=times.get)
        if depth >= 2:
            self.set_fusion(depth, *best)
        else:
            self.set_tile(*best)
        return times

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fd_ens_destroy(self._h)
            self._h = None
            for c in self.computes:
                c._ensembles -= 1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def domain_spec(k, seed=3):
    """Domain k of the many-domain workload: W in {256, 288, ..., 1024} and
    H in {128, 160, ..., 512} drawn uniformly and independently; porosity
    U(0.6, 0.85) as the target (random placement of non-overlapping
    obstacles jams below some coverage in narrow domains; the porosity
    reached is what the mask holds); squares or circles with probability
    1/2 (seed 1000*seed+k);
    viscosity ratio M = exp(U(0, ln 20)); contact angle U(31, 90) degrees
    through the defending fluid, turned into psi_wall by the droplet
    calibration (60..90 degrees through the invading fluid stays inside the
    calibrated, stable range; see bench.wall_value).  The flow: constant
    Zou-He inlet u = 0.001, pressure outlet, periodic y, complete
    bounce-back links; |psi| <= 1 in the viscosity interpolation (see the
    module docstring)."""
    rng = np.random.default_rng([seed, k])
    W = int(rng.integers(256 // 32, 1024 // 32 + 1)) * 32
    H = int(rng.integers(128 // 32, 512 // 32 + 1)) * 32
    porosity = float(rng.uniform(0.6, 0.85))
    circles = bool(rng.random() < 0.5)
    M = float(math.exp(rng.uniform(0.0, math.log(20.0))))
    theta = float(rng.uniform(31.0, 90.0))
    return dict(k=k, H=H, W=W, porosity=porosity, circles=circles, M=M, theta=theta,
                obstacle_seed=1000 * seed + k, psi_visc_bound=1.0)


def build_domain(spec):
    """(mask, links, params) of one domain_spec."""
    H, W = spec["H"], spec["W"]
    x_lo, x_hi = 24, W - 25
    if spec["circles"]:
        rmax = max(4, min(12, H // 12))
        circles = geometry.random_circles(H, W, 3, rmax, spec["porosity"], seed=spec["obstacle_seed"],
                                          margin=1, x_lo=x_lo, x_hi=x_hi, max_tries=20000, strict=False)
        bpa, _, _, _ = geometry.Createblock(H, W).setCirleblock(circles)
    else:
        bpa = _square_bed(H, W, spec["porosity"], spec["obstacle_seed"], x_lo, x_hi)
    mask = geometry.fluid_mask(bpa)
    links = geometry.link_bits_complete(mask)
    params = LbmParams.periodic(H, W, M=spec["M"], psi_wall=wetting.psi_wall_for_angle(180.0 - spec["theta"]),
                                u0=0.001, profile=PROFILE_CONST, outlet=OUTLET_PRESSURE, rho_out=1.0,
                                n_inv_cols=16, psi_visc_bound=spec.get("psi_visc_bound", 0.0))
    return mask, links, params


def _square_bed(H, W, porosity, seed, x_lo, x_hi, size=6):
    """Seeded non-overlapping size x size squares between columns x_lo and
    x_hi, rows 2..H-3, until the solid fraction of the whole domain reaches
    1 - porosity (or no further square fits)."""
    rng = np.random.default_rng(seed)
    block = np.zeros((H, W), dtype=np.uint8)
    target = (1.0 - porosity) * H * W
    solid = 0
    tries = 0
    while solid + size * size <= target and tries < 20000:
        tries += 1
        x0 = int(rng.integers(x_lo, x_hi - size))
        y0 = int(rng.integers(2, H - 2 - size))
        if block[y0 - 1:y0 + size + 1, x0 - 1:x0 + size + 1].any():
            continue
        block[y0:y0 + size, x0:x0 + size] = 1
        solid += size * size
    return block
