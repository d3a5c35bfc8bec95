"""Shared set-up for the parity tests: one domain description drives both
the GPU solver (fingering_dynamics_amd) and the FP64 oracle."""
import numpy as np

from fingering_dynamics_amd import geometry
from fingering_dynamics_amd.params import LbmParams
from oracle import geometry64, lbm64


def rect_domain(H, W, n, size=6, seed=1, x_min=None):
    """Seeded squares -> (fluid mask, product link bits, oracle link masks)."""
    rects = geometry.random_squares(H, W, n, size=size, seed=seed,
                                    x_min=(min(41, W // 3) if x_min is None else x_min))
    bpa, corners = geometry.Createblock(H, W).setblock(rects)
    mask = geometry.fluid_mask(bpa)
    bits = geometry.Bounce_back(H, W).link_bits_rec(corners)
    ob, oc = geometry64.setblock(H, W, rects)
    assert np.array_equal(ob, bpa)
    links = geometry64.rect_link_masks(H, W, oc)
    return mask, bits, links


def circle_domain(H, W, r_min, r_max, porosity, seed=2, x_lo=None):
    circles = geometry.random_circles(H, W, r_min, r_max, porosity, seed=seed, x_lo=x_lo)
    bpa, side, cave, vex = geometry.Createblock(H, W).setCirleblock(circles)
    mask = geometry.fluid_mask(bpa)
    bits = geometry.Bounce_back(H, W).link_bits_circle(side, cave, vex)
    ob, oside, ocave, ovex = geometry64.set_circle_block(H, W, circles)
    assert np.array_equal(ob, bpa)
    links = geometry64.circle_link_masks(oside, ocave, ovex)
    return mask, bits, links


def oracle_for(params: LbmParams, mask, links, psi0=None, rho0=None):
    """An Lbm64 configured exactly like ``params``."""
    return lbm64.Lbm64.from_params(params, mask, links, psi0=psi0, rho0=rho0)


def oracle_fields(o):
    """(psi, rho, ux, uy) of the oracle as full [H][W] arrays (0 on solids)."""
    return o.psi.copy(), o.field("rho"), o.field("ux"), o.field("uy")


def max_err(a, b, mask=None):
    d = np.abs(np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64))
    if mask is not None:
        d = d[mask]
    return float(d.max()) if d.size else 0.0
