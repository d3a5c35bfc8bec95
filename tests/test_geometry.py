"""Obstacle front-end: oracle pins (known answers, invariants) and bit-exact
agreement of the product's packed link bits with the oracle's masks."""
import numpy as np
import pytest

from fingering_dynamics_amd import geometry
from oracle import geometry64


def test_rect_filled_box_closed_form():
    H, W = 20, 30
    bpa, corners = geometry64.setblock(H, W, [((4, 3), (9, 8)), ((20, 10), (25, 15))])
    exp = np.zeros((H, W), dtype=int)
    exp[3:9, 4:10] = 1
    exp[10:16, 20:26] = 1
    assert np.array_equal(bpa, exp)
    assert corners[0] == {"top_left": (4, 8), "bottom_left": (4, 3),
                          "top_right": (9, 8), "bottom_right": (9, 3)}


def test_rect_links_known_answer():
    """One 6x6 block occupying x 4..9, y 3..8 in a 14 x 16 grid.  Link sets
    enumerated by hand from bounce_back.py:37-86:
      N row y=9, x 4..9 ; S row y=2, x 4..9 ; W col x=3, y 3..8 ; E col x=10, y 3..8
      corners NW (3,9) NE (10,9) SW (3,2) SE (10,2)   [(x, y)]
      dir1 {E} dir2 {N} dir3 {W} dir4 {S} dir5 {NE,N,E} dir6 {NW,N,W}
      dir7 {SW,S,W} dir8 {SE,S,E}"""
    H, W = 14, 16
    _, corners = geometry64.setblock(H, W, [((4, 3), (9, 8))])
    links = geometry64.rect_link_masks(H, W, corners)
    N = {(x, 9) for x in range(4, 10)}
    S = {(x, 2) for x in range(4, 10)}
    Wc = {(3, y) for y in range(3, 9)}
    Ec = {(10, y) for y in range(3, 9)}
    expect = {1: Ec, 2: N, 3: Wc, 4: S,
              5: {(10, 9)} | N | Ec, 6: {(3, 9)} | N | Wc,
              7: {(3, 2)} | S | Wc, 8: {(10, 2)} | S | Ec}
    for i in range(1, 9):
        got = {(int(x), int(y)) for y, x in np.argwhere(links[i])}
        assert got == expect[i], i
    # and the product packs the same sets
    bits = geometry.Bounce_back(H, W).link_bits_rec(corners)
    assert np.array_equal(geometry.unpack_links(bits), links)


def test_rect_index_semantics():
    """A block at x0 = 0 puts its west barrier at column -1 (wraps to W-1);
    a block whose top row is H-1 raises IndexError (row H)."""
    H, W = 12, 14
    _, corners = geometry64.setblock(H, W, [((0, 3), (4, 6))])
    links = geometry64.rect_link_masks(H, W, corners)
    assert links[3][3:7, W - 1].all()
    bits = geometry.Bounce_back(H, W).link_bits_rec(geometry.Createblock(H, W).setblock([((0, 3), (4, 6))])[1])
    assert np.array_equal(geometry.unpack_links(bits), links)
    _, c2 = geometry64.setblock(H, W, [((3, 5), (6, H - 1))])
    with pytest.raises(IndexError):
        geometry64.rect_link_masks(H, W, c2)
    with pytest.raises(IndexError):
        geometry.Bounce_back(H, W).link_bits_rec(c2)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rect_product_matches_oracle(seed):
    H, W = 100, 400
    rects = geometry.random_squares(H, W, 56, size=6, seed=seed)
    bpa, corners = geometry.Createblock(H, W).setblock(rects)
    obpa, ocorners = geometry64.setblock(H, W, rects)
    assert np.array_equal(bpa, obpa)
    assert corners == ocorners
    bits = geometry.Bounce_back(H, W).link_bits_rec(corners)
    assert np.array_equal(geometry.unpack_links(bits), geometry64.rect_link_masks(H, W, ocorners))
    mask = geometry.fluid_mask(bpa)
    assert np.array_equal(mask, geometry64.fluid_mask(obpa))
    assert abs((1 - mask.mean()) - 0.0504) < 1e-9          # 56 * 36 / 40000


def test_overlapping_blocks_become_fluid():
    H, W = 20, 20
    bpa, _ = geometry.Createblock(H, W).setblock([((2, 2), (7, 7)), ((5, 5), (10, 10))])
    mask = geometry.fluid_mask(bpa)
    assert mask[6, 6] and not mask[3, 3] and not mask[9, 9]


def test_random_squares_constraints():
    H, W = 100, 400
    rects = geometry.random_squares(H, W, 56, size=6, seed=1)
    occ = np.zeros((H, W), dtype=int)
    for (x0, y0), (x1, y1) in rects:
        assert x1 - x0 == 5 and y1 - y0 == 5
        assert x0 >= 41 and y0 >= 1 and x1 <= W - 3 and y1 <= H - 3
        occ[y0 - 1:y1 + 2, x0 - 1:x1 + 2] += 1
    # a one-cell gap: grown boxes may touch but blocks never share a node or a side
    assert geometry.random_squares(H, W, 56, size=6, seed=1) == rects


# ----------------------------------------------------------------- circles
def test_circle_outline_properties():
    for r in range(1, 30):
        pts = geometry64.circle_outline(2 * r + 5, 2 * r + 5, r + 2, r + 2, r)
        ys, xs = np.nonzero(pts)
        d = np.hypot(xs - (r + 2), ys - (r + 2))
        assert np.all(np.abs(d - r) < 1.0), r
        assert np.array_equal(pts, pts[::-1, :]) and np.array_equal(pts, pts[:, ::-1])
        assert np.array_equal(pts, pts.T)
        # the product rasteriser draws the same pixels
        prod = geometry.Createblock(2 * r + 5, 2 * r + 5).getCicleblock((r + 2, r + 2), r)
        assert np.array_equal(prod, pts)


def _neighbour_class_ok(block, x, y, cls):
    """The neighbour pattern each class of create_block.py:100-112 names,
    for a first-quadrant node (x, y) of the outline `block`."""
    left, below, diag = block[y, x - 1], block[y - 1, x], block[y - 1, x - 1]
    return {"cave": left and below, "left": left and not below,
            "top": below and not left, "vex": diag and not left and not below}[cls]


def test_circle_classifier_invariants():
    H, W = 80, 90
    for (cx, cy, r) in ((40, 40, 10), (45, 38, 17), (30, 30, 5), (50, 44, 23)):
        bpa, side, cave, vex = geometry64.set_circle_block(H, W, [((cx, cy), r)])
        block = geometry64.circle_outline(H, W, cx, cy, r)
        solid = bpa == 1
        top, bottom, right, left = side
        classes = {"top": top, "left": left, "cave": cave[0], "vex": vex[0]}
        for name, m in classes.items():
            for y, x in np.argwhere(m):
                assert not solid[y, x], (name, x, y)       # every boundary node is fluid
                if x > cx and y > cy:                       # first quadrant: the stated pattern
                    assert _neighbour_class_ok(block, x, y, name), (name, x, y)
        # four-fold mirror symmetry of the class sets
        def mirror_x(m):
            out = np.zeros_like(m)
            ys, xs = np.nonzero(m)
            out[ys, 2 * cx - xs] = True
            return out

        def mirror_y(m):
            out = np.zeros_like(m)
            ys, xs = np.nonzero(m)
            out[2 * cy - ys, xs] = True
            return out
        assert np.array_equal(mirror_y(top), bottom)
        assert np.array_equal(mirror_x(left), right)
        assert np.array_equal(mirror_x(cave[0]), cave[1])
        assert np.array_equal(mirror_y(cave[0]), cave[2])
        assert np.array_equal(mirror_x(mirror_y(vex[0])), vex[3])
        # every solid node next to fluid has its fluid neighbours bounced in the
        # axis directions (no leak through a side)
        links = geometry64.circle_link_masks(side, cave, vex)
        for y, x in np.argwhere(solid):
            for i, (ex, ey) in ((1, (1, 0)), (2, (0, 1)), (3, (-1, 0)), (4, (0, -1))):
                ny, nx = y + ey, x + ex
                if not solid[ny, nx]:
                    assert links[i][ny, nx], (i, x, y)


@pytest.mark.parametrize("seed", [2, 5, 9])
def test_circle_product_matches_oracle(seed):
    H, W = 160, 240
    circles = geometry.random_circles(H, W, 8, 24, 0.75, seed=seed, x_lo=20)
    a = geometry.Createblock(H, W).setCirleblock(circles)
    b = geometry64.set_circle_block(H, W, circles)
    assert np.array_equal(a[0], b[0])
    for x, y in zip(a[1] + a[2] + a[3], b[1] + b[2] + b[3]):
        assert np.array_equal(x, y)
    bits = geometry.Bounce_back(H, W).link_bits_circle(*a[1:])
    assert np.array_equal(geometry.unpack_links(bits), geometry64.circle_link_masks(*b[1:]))


def test_circle_edge_wrap_and_index_error():
    H, W = 120, 160
    wrap_cases = [((3, 40), 7), ((60, 4), 9), ((0, 60), 5)]
    a = geometry.Createblock(H, W).setCirleblock(wrap_cases)
    b = geometry64.set_circle_block(H, W, wrap_cases)
    assert np.array_equal(a[0], b[0])
    for x, y in zip(a[1] + a[2] + a[3], b[1] + b[2] + b[3]):
        assert np.array_equal(x, y)
    for bad in ([((150, 100), 9)], [((60, 112), 8)]):
        with pytest.raises(IndexError):
            geometry64.set_circle_block(H, W, bad)
        with pytest.raises(IndexError):
            geometry.Createblock(H, W).setCirleblock(bad)


def test_random_circles_porosity_and_margins():
    H, W = 256, 512
    circles = geometry.random_circles(H, W, 8, 24, 0.75, seed=2, x_lo=30)
    bpa, *_ = geometry.Createblock(H, W).setCirleblock(circles)
    mask = geometry.fluid_mask(bpa)
    assert abs(mask.mean() - 0.75) < 0.01
    assert bpa.max() == 1                                  # no overlaps
    for (cx, cy), r in circles:
        assert 8 <= r <= 24 and cx - r - 1 >= 30 and cy - r - 1 >= 2 and cy + r + 1 <= H - 3
