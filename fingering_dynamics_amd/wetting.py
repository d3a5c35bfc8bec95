"""Contact angle <-> wall order parameter (theta <-> psi_wall).

The reference sets wettability only through psi_wall, the order parameter
of the ghost rows beyond the walls (fingering_periodic.py:18, fingering.py:19,
224); it has no map to a contact angle.  Its droplet scenario
(lattice_boltzmann/validation.py:14-40, 323-326, 387-425) is the template
used here to build one: a droplet of the psi = +1 phase on the bottom wall
of a channel periodic in x, with half-way walls on rows 0 and H-1 and the
psi_wall ghost rows beyond them, relaxed on the GPU until its shape is
stationary.

The angle is measured through the droplet (psi > 0) phase: a circle is
fitted to the psi = 0 contour away from the wall layer, and with the wall
line at y_w = -1/2 (half-way between row 0 and the ghost row)

    cos theta = (y_w - y_c) / R.

theta(psi_wall) is tabulated for a set of psi_wall values and inverted by
monotone interpolation.  The relation depends on the interface constants
only through the width xi = sqrt(2 kappa / -a) (not on M, Pe, sigma or the
flow); xi = 2 in both reference configurations.  A table records xi and
refuses other values.
"""
import json
import math
import os

import numpy as np

from .params import LbmParams, X_PERIODIC, Y_WALLS, PROFILE_CONST

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "wetting_table.json")
Y_WALL = -0.5


def droplet_params(H, W, psi_wall, M=1.0, Pe=15.0, soften=8.0):
    """A closed channel, periodic in x, walls on rows 0 and H-1, no flow:
    the validation.py scenario with the fingering_periodic.py constants,
    except that a and kappa are divided by ``soften`` and the mobility
    gamma multiplied by it.  The equilibrium (mu = const, u = 0) depends on
    a and kappa only through xi = sqrt(2 kappa / -a), and gamma * mu is
    unchanged, so the psi dynamics and the contact angle are the same while
    the capillary forces are ``soften`` times weaker: the full-strength
    constants (sigma = 0.0625) drive strongly wetting droplets unstable."""
    p = LbmParams.periodic(H, W, M=M, psi_wall=psi_wall, Pe=Pe, x_bc=X_PERIODIC, y_bc=Y_WALLS,
                           wall_lo=0, wall_hi=H - 1, profile=PROFILE_CONST, u0=0.0)
    return p.with_(a=p.a / soften, kappa=p.kappa / soften, gamma=p.gamma * soften)


def interface_width(params):
    """xi of the equilibrium profile psi = tanh(x / xi): mu = a psi (1 - psi^2)
    - kappa psi'' = 0 needs xi^2 = 2 kappa / (-a)."""
    return math.sqrt(2.0 * params.kappa / -params.a)


def droplet_psi(H, W, R, xi, yc=Y_WALL):
    """psi of a disk of radius R centred at (W/2, yc) with the equilibrium
    tanh profile of width xi (a half disk on the wall for yc = -1/2).
    validation.py:323-326 starts from a sharp +-1 disk tangent to the wall;
    a sharp circular interface drives the scheme unstable within 20
    iterations (the FP64 oracle as well), the equilibrium profile does not."""
    y, x = np.mgrid[0:H, 0:W]
    d = np.sqrt((x - (W - 1) / 2.0) ** 2 + (y - yc) ** 2)
    return np.tanh((R - d) / xi)


def contour_points(psi, y_min=3, y_max=None):
    """Sub-grid points of the psi = 0 contour (linear interpolation between
    neighbouring nodes of opposite sign), along columns and rows, with
    y_min <= y <= y_max."""
    H, W = psi.shape
    y_max = H - 4 if y_max is None else y_max
    pts = []
    a, b = psi[:-1, :], psi[1:, :]                       # vertical neighbours
    yy, xx = np.nonzero((a > 0) != (b > 0))
    t = a[yy, xx] / (a[yy, xx] - b[yy, xx])
    pts.append(np.stack([xx.astype(float), yy + t], axis=1))
    a, b = psi[:, :-1], psi[:, 1:]                       # horizontal neighbours
    yy, xx = np.nonzero((a > 0) != (b > 0))
    t = a[yy, xx] / (a[yy, xx] - b[yy, xx])
    pts.append(np.stack([xx + t, yy.astype(float)], axis=1))
    p = np.concatenate(pts)
    return p[(p[:, 1] >= y_min) & (p[:, 1] <= y_max)]


def fit_circle(pts):
    """Algebraic least-squares circle through pts [n][2] -> (xc, yc, R)."""
    x, y = pts[:, 0], pts[:, 1]
    A = np.stack([x, y, np.ones_like(x)], axis=1)
    rhs = -(x * x + y * y)
    (D, E, F), *_ = np.linalg.lstsq(A, rhs, rcond=None)
    xc, yc = -D / 2.0, -E / 2.0
    return xc, yc, math.sqrt(max(xc * xc + yc * yc - F, 0.0))


def contact_angle(psi, y_min=3, y_wall=Y_WALL):
    """Contact angle (degrees, through the psi > 0 phase) of a droplet on the
    bottom wall of psi [H][W]."""
    pts = contour_points(np.asarray(psi, dtype=np.float64), y_min=y_min)
    if len(pts) < 8:
        raise ValueError("no droplet contour found")
    _, yc, R = fit_circle(pts)
    return math.degrees(math.acos(min(1.0, max(-1.0, (y_wall - yc) / R))))


MAX_STEPS = 100000


def relax_droplet(psi_wall, H=128, W=320, R=40, M=1.0, chunk=5000, max_steps=MAX_STEPS, tol_deg=0.05, phase=1.0):
    """Run the droplet on the GPU until its angle moves by less than tol_deg
    over a chunk.  ``phase`` = -1 runs the mirror problem (a psi = -1
    droplet in psi = +1) and measures through the psi < 0 phase.  Returns
    (theta_deg, steps, psi); steps == max_steps means it did not settle."""
    from .lbm import Compute
    p = droplet_params(H, W, psi_wall, M=M)
    mask = np.ones((H, W), dtype=bool)
    links = np.zeros((H, W), dtype=np.uint8)
    cm = Compute(mask, links, p, psi0=phase * droplet_psi(H, W, R, interface_width(p)))
    try:
        prev = None
        while True:
            cm.step(chunk)
            psi = cm.fields()[0]
            th = contact_angle(phase * psi)
            if (prev is not None and abs(th - prev) < tol_deg) or cm.t >= max_steps:
                s_rho, s_psi, bad = cm.mass()
                if bad:
                    raise FloatingPointError("droplet at psi_wall=%g went non-finite" % psi_wall)
                return th, cm.t, psi
            prev = th
    finally:
        cm.close()


# The two droplet sizes of a calibration point.  The apparent angle of a
# diffuse droplet carries a curvature bias ~ c / R (about 48 degrees / R
# here: theta(psi_wall = 0) is 92.7, 91.2, 90.8 at R = 18, 40, 60), removed
# by Richardson extrapolation in 1 / R.
SIZES = (dict(H=128, W=320, R=40), dict(H=192, W=448, R=60))


def extrapolate(theta1, R1, theta2, R2):
    """theta at R -> infinity from two sizes, assuming theta(R) = theta_inf + c / R."""
    return (R2 * theta2 - R1 * theta1) / (R2 - R1)


def calibrate(psi_walls, sizes=SIZES):
    """theta(psi_wall) for each value (GPU runs): the droplet at two sizes,
    extrapolated to R -> infinity.  Values whose droplet does not settle
    into a measurable cap (a film, or a drop lifting off) are skipped.
    Returns a table dict."""
    c = LbmParams.periodic(8, 8)
    rows = []
    for pw in psi_walls:
        try:
            runs = [relax_droplet(float(pw), **kw) for kw in sizes]
        except (ValueError, FloatingPointError) as e:
            print("psi_wall %g skipped: %s" % (pw, e), flush=True)
            continue
        (t1, n1, _), (t2, n2, _) = runs
        if n1 >= MAX_STEPS or n2 >= MAX_STEPS or not (1.0 < t1 < 179.0 and 1.0 < t2 < 179.0):
            print("psi_wall %g skipped: not settled into a cap (%.2f, %.2f deg after %d, %d steps)"
                  % (pw, t1, t2, n1, n2), flush=True)
            continue
        th = extrapolate(t1, sizes[0]["R"], t2, sizes[1]["R"])
        rows.append({"psi_wall": float(pw), "theta_deg": round(th, 3), "theta_small": round(t1, 3),
                     "theta_large": round(t2, 3), "steps": [int(n1), int(n2)]})
        print(json.dumps(rows[-1]), flush=True)
    return {"xi": interface_width(c), "rows": rows,
            "scenario": "droplet of psi=+1 on the bottom wall, periodic x, walls on rows 0 and H-1 "
                        "(validation.py), sizes %s, extrapolated in 1/R" % json.dumps(list(sizes))}


def load_table(path=TABLE_PATH):
    with open(path) as f:
        return json.load(f)


def _check_constants(table, params):
    if params is None:
        return
    if not math.isclose(table["xi"], interface_width(params), rel_tol=1e-9):
        raise ValueError("the wetting table was made for interface width xi=%g, not %g"
                         % (table["xi"], interface_width(params)))


def psi_wall_for_angle(theta_deg, table=None, params=None):
    """psi_wall giving contact angle theta_deg (through the psi = +1 phase),
    by monotone interpolation of the table; raises outside its range."""
    table = load_table() if table is None else table
    _check_constants(table, params)
    rows = sorted(table["rows"], key=lambda r: r["theta_deg"])
    th = np.array([r["theta_deg"] for r in rows])
    pw = np.array([r["psi_wall"] for r in rows])
    if np.any(np.diff(th) <= 0) or np.any(np.diff(pw) >= 0):
        raise ValueError("wetting table is not monotone")
    if not th[0] <= theta_deg <= th[-1]:
        raise ValueError("theta %.1f outside the calibrated range [%.1f, %.1f]" % (theta_deg, th[0], th[-1]))
    return float(np.interp(theta_deg, th, pw))


def angle_for_psi_wall(psi_wall, table=None, params=None):
    """Inverse of psi_wall_for_angle."""
    table = load_table() if table is None else table
    _check_constants(table, params)
    rows = sorted(table["rows"], key=lambda r: r["psi_wall"])
    pw = np.array([r["psi_wall"] for r in rows])
    th = np.array([r["theta_deg"] for r in rows])
    if not pw[0] <= psi_wall <= pw[-1]:
        raise ValueError("psi_wall %.3f outside the calibrated range" % psi_wall)
    return float(np.interp(psi_wall, pw, th))


if __name__ == "__main__":
    # python -m fingering_dynamics_amd.wetting [OUT.json]  (needs a GPU);
    # writes the table to OUT.json (default: the packaged table)
    import sys
    out = sys.argv[1] if len(sys.argv) > 1 else TABLE_PATH
    values = [round(v, 2) for v in np.arange(-0.7, 0.71, 0.05)]
    table = calibrate(values)
    table["command"] = "python -m fingering_dynamics_amd.wetting"
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    with open(out, "w") as f:
        json.dump(table, f, indent=1)
