"""Obstacle front-end: solid masks and packed bounce-back link bits.

Keeps the reference's set-up API so a driver written against
lattice_boltzmann/create_block.py and bounce_back.py switches over unchanged:

  Createblock(H, W).setblock(rect_corner_list)  -> (block_psi_all, corner_list)
                                                   create_block.py:395-407
  Createblock(H, W).setCirleblock(circle_list)  -> (block_psi_all, side_list,
                                                   concave_list, convex_list)
                                                   create_block.py:51-220
  fluid_mask(block_psi_all)                     -> fingering_periodic.py:450-451

and replaces the per-step mask work of bounce_back.py (which rebuilds or
re-applies up to 12 boolean masks per step, bounce_back.py:25-167) with one
uint8 per node, packed once: bit (i-1) set means direction i is overwritten
by the node's own post-collision opposite population.

  Bounce_back(H, W).link_bits_rec(corner_list)                  bounce_back.py:25-86
  Bounce_back(H, W).link_bits_circle(side, concave, convex)     bounce_back.py:89-167

Index semantics follow the reference's numpy assignments: a row or column
index of -1 wraps to the last one, an index equal to H or W raises
IndexError (bounce_back.py:37-44).

The circle raster stands in for cv2.circle (create_block.py:21), which is not
available here; it is the integer midpoint circle of radius r (octant march
keeping dx^2 + dy^2 - r^2 <= 0 after each inward step).
"""
import numpy as np
from scipy.ndimage import binary_fill_holes

# link-bit classes of the circle classifier
_TOP, _BOTTOM, _RIGHT, _LEFT = 0, 1, 2, 3
_CAVE = {"tr": 4, "tl": 5, "br": 6, "bl": 7}
_VEX = {"tr": 8, "tl": 9, "br": 10, "bl": 11}

# direction i (1..8) <- classes whose nodes bounce it (bounce_back.py:103-167,
# with N=top, S=bottom, E=right, W=left, nw=tr, ne=tl, sw=br, se=bl)
_DIR_CLASSES = {
    1: (_LEFT, _CAVE["tr"], _CAVE["br"]),
    2: (_TOP, _CAVE["tr"], _CAVE["tl"]),
    3: (_RIGHT, _CAVE["tl"], _CAVE["bl"]),
    4: (_BOTTOM, _CAVE["br"], _CAVE["bl"]),
    5: (_VEX["tr"], _CAVE["tr"], _TOP, _LEFT),
    6: (_VEX["tl"], _CAVE["tl"], _TOP, _RIGHT),
    7: (_VEX["bl"], _CAVE["bl"], _BOTTOM, _RIGHT),
    8: (_VEX["br"], _CAVE["br"], _BOTTOM, _LEFT),
}


def _idx(v, n):
    """numpy scalar-index semantics for -n <= v < n."""
    if v >= n or v < -n:
        raise IndexError("index %d is out of bounds for axis with size %d" % (v, n))
    return v % n


class Createblock:
    """Obstacle rasteriser with create_block.py's interface."""

    def __init__(self, H, W):
        self.H = H
        self.W = W

    def getRectangleblock(self, bottom_left, top_right):
        """Outline of the box with corners bottom_left / top_right, inclusive
        (cv2.rectangle thickness 1, create_block.py:13-16).  The box must lie
        inside the image."""
        xa, xb = sorted((int(bottom_left[0]), int(top_right[0])))
        ya, yb = sorted((int(bottom_left[1]), int(top_right[1])))
        if xa < 0 or ya < 0 or xb >= self.W or yb >= self.H:
            raise ValueError("rectangle %s-%s outside the %dx%d image" % (bottom_left, top_right, self.H, self.W))
        block = np.zeros((self.H, self.W), dtype=np.uint8)
        block[[ya, yb], xa:xb + 1] = 1
        block[ya:yb + 1, [xa, xb]] = 1
        return block

    def getCicleblock(self, center, radius):
        """Outline of a circle (stand-in for cv2.circle, create_block.py:19-21)."""
        block = np.zeros((self.H, self.W), dtype=np.uint8)
        cx, cy = int(center[0]), int(center[1])
        for X, Y in _circle_points(cx, cy, int(radius)):
            if 0 <= X < self.W and 0 <= Y < self.H:
                block[Y, X] = 1
        return block

    @staticmethod
    def getCorner(block):
        """Corners of the bounding box of the non-zero pixels (create_block.py:39-48)."""
        ys, xs = np.nonzero(block)
        x0, x1, y0, y1 = int(xs.min()), int(xs.max()), int(ys.min()), int(ys.max())
        return {"top_left": (x0, y1), "bottom_left": (x0, y0),
                "top_right": (x1, y1), "bottom_right": (x1, y0)}

    def setblock(self, rect_corner_list):
        """create_block.py:395-407: per-rectangle filled boxes summed into an
        int count array, and the corner dicts."""
        block_psi_all = np.zeros((self.H, self.W), dtype=int)
        corner_list = []
        for rect in rect_corner_list:
            block = self.getRectangleblock(rect[0], rect[1])
            corner_list.append(self.getCorner(block))
            _add_filled(block_psi_all, block)
        return block_psi_all, corner_list

    def setCirleblock(self, circle_list):
        """create_block.py:51-220: filled discs summed into an int count
        array, and the fluid nodes next to each outline classified from a
        scan of the first quadrant, mirrored to the other three.

        A circle whose scan window (centre +- (r+1)) lies inside the image is
        rasterised, filled and scanned on that window only; the result is the
        same as on the full image.  Others take the full-image path, which
        keeps numpy's wrap-around and IndexError behaviour."""
        H, W = self.H, self.W
        block_psi_all = np.zeros((H, W), dtype=int)
        pts = {k: [] for k in ("top", "bottom", "right", "left")}
        cave = {k: [] for k in ("tr", "tl", "br", "bl")}
        vex = {k: [] for k in ("tr", "tl", "br", "bl")}
        for (center, r) in circle_list:
            x, y, r = int(center[0]), int(center[1]), int(r)
            if x - r - 1 >= 0 and y - r - 1 >= 0 and x + r + 1 < W and y + r + 1 < H:
                oy, ox = y - r - 1, x - r - 1
                n = 2 * r + 3
                block = np.zeros((n, n), dtype=np.uint8)
                for X, Y in _circle_points(r + 1, r + 1, r):
                    block[Y, X] = 1
                block_psi_all[oy:oy + n, ox:ox + n] += binary_fill_holes(block).astype(int)
                q = _classify_quadrant(block, r + 1, r + 1, r)
                q = [[(px + ox, py + oy) for px, py in lst] for lst in q]
            else:
                block = self.getCicleblock((x, y), r)
                _add_filled(block_psi_all, block)
                q = _classify_quadrant(block, x, y, r)
            q_top, q_left, q_cave, q_vex = q
            # mirror images: (2x - px, py), (px, 2y - py), (2x - px, 2y - py)
            for px, py in q_top:
                pts["top"].append((2 * x - px, py))
                pts["bottom"].append((px, 2 * y - py))
                pts["bottom"].append((2 * x - px, 2 * y - py))
            pts["top"].append((x, y + r + 1))
            pts["top"].extend(q_top)
            pts["bottom"].append((x, y - r - 1))
            for px, py in q_left:
                pts["right"].append((2 * x - px, py))
                pts["left"].append((px, 2 * y - py))
                pts["right"].append((2 * x - px, 2 * y - py))
            pts["left"].extend(q_left)
            pts["left"].append((x + r + 1, y))
            pts["right"].append((x - r - 1, y))
            for store, quad in ((cave, q_cave), (vex, q_vex)):
                for px, py in quad:
                    store["tl"].append((2 * x - px, py))
                    store["br"].append((px, 2 * y - py))
                    store["bl"].append((2 * x - px, 2 * y - py))
                store["tr"].extend(quad)

        def mask(points):
            m = np.zeros((H, W), dtype=bool)
            if points:
                a = np.asarray(points, dtype=np.int64)
                px, py = a[:, 0], a[:, 1]
                if (py >= H).any() or (py < -H).any() or (px >= W).any() or (px < -W).any():
                    bad = a[(py >= H) | (py < -H) | (px >= W) | (px < -W)][0]
                    raise IndexError("boundary node %s is out of bounds for a %dx%d grid"
                                     % (tuple(bad), H, W))
                m[py % H, px % W] = True
            return m

        side_list = [mask(pts[k]) for k in ("top", "bottom", "right", "left")]
        concave_list = [mask(cave[k]) for k in ("tr", "tl", "br", "bl")]
        convex_list = [mask(vex[k]) for k in ("tr", "tl", "br", "bl")]
        return block_psi_all, side_list, concave_list, convex_list


def _circle_points(cx, cy, r):
    """Pixels of the integer midpoint circle: march dy = 0, 1, ... in the
    octant dx >= dy keeping err = dx^2 + dy^2 - r^2, step dx inward once
    whenever err turns positive, and take the 8 mirror images."""
    octant = []
    dx, dy, err = r, 0, 0
    while dx >= dy:
        octant.append((dx, dy))
        dy += 1
        err += 2 * dy - 1
        if err > 0:
            err -= 2 * dx - 1
            dx -= 1
    pts = set()
    for a, b in octant:
        for sx in (-1, 1):
            for sy in (-1, 1):
                pts.add((cx + sx * a, cy + sy * b))
                pts.add((cx + sx * b, cy + sy * a))
    return pts


def _add_filled(acc, block):
    """acc += binary_fill_holes(block) computed on the outline's bounding box
    grown by one pixel and clipped to the image: the enclosed region and its
    connection to the image border are the same as on the full image, at a
    cost proportional to the obstacle instead of H*W."""
    ys, xs = np.nonzero(block)
    if ys.size == 0:
        return
    H, W = block.shape
    y0, y1 = max(int(ys.min()) - 1, 0), min(int(ys.max()) + 2, H)
    x0, x1 = max(int(xs.min()) - 1, 0), min(int(xs.max()) + 2, W)
    acc[y0:y1, x0:x1] += binary_fill_holes(block[y0:y1, x0:x1]).astype(acc.dtype)


def _classify_quadrant(block, x, y, r):
    """First-quadrant scan of create_block.py:81-112 over the outline
    ``block``: rows y+1 .. y+r+1, each from column x+jj inward until the
    outline is met; the next row starts one column outside that hit."""
    H, W = block.shape

    def at(yy, xx):
        return block[_idx(yy, H), _idx(xx, W)]

    top, left, cave, vex = [], [], [], []
    start = r
    for i in range(1, r + 2):
        for j in range(start, 0, -1):
            if at(y + i, x + j):
                start = j + 1
                break
            below = at(y + i - 1, x + j)
            west = at(y + i, x + j - 1)
            diag = at(y + i - 1, x + j - 1)
            node = (x + j, y + i)
            if west and below:
                cave.append(node)
            elif west:
                left.append(node)
            elif below:
                top.append(node)
            elif diag:
                vex.append(node)
    return top, left, cave, vex


def fluid_mask(block_psi_all):
    """fingering_periodic.py:450-451: a node is solid iff exactly one block covers it."""
    return np.asarray(block_psi_all) != 1


class Bounce_back:
    """Packs the reference's bounce-back masks into one uint8 per node."""

    def __init__(self, H, W):
        self.H = H
        self.W = W

    def link_bits_rec(self, corner_list):
        """Bits of halfway_bounceback_rec (bounce_back.py:28-86).  Per block:
        the row above / below and the columns left / right of the box, plus
        the four diagonal corner nodes."""
        H, W = self.H, self.W
        bits = np.zeros((H, W), dtype=np.uint8)
        b = lambda i: np.uint8(1 << (i - 1))
        for cor in corner_list:
            x0, y1 = cor["top_left"]
            x1 = cor["top_right"][0]
            y0 = cor["bottom_left"][1]
            xs = slice(x0, x1 + 1)
            ys = slice(y0, y1 + 1)
            north, south = _idx(y1 + 1, H), _idx(y0 - 1, H)
            west, east = _idx(x0 - 1, W), _idx(x1 + 1, W)
            bits[north, xs] |= b(2) | b(5) | b(6)        # n_barrier: 2, 5, 6
            bits[south, xs] |= b(4) | b(7) | b(8)        # s_barrier: 4, 7, 8
            bits[ys, west] |= b(3) | b(6) | b(7)         # w_barrier: 3, 6, 7
            bits[ys, east] |= b(1) | b(5) | b(8)         # e_barrier: 1, 5, 8
            bits[north, west] |= b(6)                    # nw corner
            bits[north, east] |= b(5)                    # ne corner
            bits[south, west] |= b(7)                    # sw corner
            bits[south, east] |= b(8)                    # se corner
        return bits

    def link_bits_circle(self, side_list, concave_list, convex_list):
        """Bits of halfway_bounceback_circle (bounce_back.py:89-167)."""
        cls = list(side_list) + list(concave_list) + list(convex_list)
        # side_list = [top, bottom, right, left]; concave/convex = [tr, tl, br, bl]
        bits = np.zeros((self.H, self.W), dtype=np.uint8)
        for d, members in _DIR_CLASSES.items():
            m = np.zeros((self.H, self.W), dtype=bool)
            for c in members:
                m |= np.asarray(cls[c], dtype=bool)
            bits[m] |= np.uint8(1 << (d - 1))
        return bits


def link_bits_complete(mask):
    """Build-original: the complete half-way bounce-back set.  Bit (i-1) is
    set at every fluid node whose upstream neighbour x - e_i (periodic
    indexing, as in the streaming step) is solid.

    The reference's circle classifier (create_block.py:81-112) leaves some
    diagonal solid -> fluid links out of its masks; there the solid node's
    streamed, never-collided populations (copies of what bounced off it)
    re-enter the fluid and add mass.  With this set no population ever
    leaves a solid node into fluid, so solid nodes need no update at all.
    For the rectangles of halfway_bounceback_rec it differs from the
    reference only by the reference's extra corner-adjacent bounces
    (bounce_back.py:62-86)."""
    fluid = np.asarray(mask, dtype=bool)
    solid = ~fluid
    bits = np.zeros(fluid.shape, dtype=np.uint8)
    E = ((1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1))
    for i, (ex, ey) in enumerate(E, start=1):
        upstream_solid = np.roll(np.roll(solid, ey, axis=0), ex, axis=1)
        bits[fluid & upstream_solid] |= np.uint8(1 << (i - 1))
    return bits


def unpack_links(bits):
    """uint8 bits [H][W] -> bool[9][H][W] per-direction masks (index 0 False)."""
    bits = np.asarray(bits, dtype=np.uint8)
    out = np.zeros((9,) + bits.shape, dtype=bool)
    for i in range(1, 9):
        out[i] = (bits >> (i - 1)) & 1
    return out


# ------------------------------------------------------------ generators
def random_squares(H, W, n, size=6, seed=1, x_min=41, x_max=None, y_min=1, y_max=None, gap=1,
                   max_tries=200000):
    """Seeded, non-overlapping size x size squares as setblock corner pairs.

    Lower-left corners are drawn uniformly with x0 in [x_min, x_max] and y0 in
    [y_min, y_max] (defaults W-size-2, H-size-2) and a square is kept only if
    it stays ``gap`` cells clear of every kept square.
    """
    rng = np.random.default_rng(seed)
    x_max = W - size - 2 if x_max is None else x_max
    y_max = H - size - 2 if y_max is None else y_max
    occ = np.zeros((H, W), dtype=bool)
    rects = []
    tries = 0
    while len(rects) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place %d squares" % n)
        x0 = int(rng.integers(x_min, x_max + 1))
        y0 = int(rng.integers(y_min, y_max + 1))
        ya, yb = max(y0 - gap, 0), min(y0 + size + gap, H)
        xa, xb = max(x0 - gap, 0), min(x0 + size + gap, W)
        if occ[ya:yb, xa:xb].any():
            continue
        occ[y0:y0 + size, x0:x0 + size] = True
        rects.append(((x0, y0), (x0 + size - 1, y0 + size - 1)))
    return rects


def random_circles(H, W, r_min, r_max, porosity, seed=2, margin=2, x_lo=None, x_hi=None,
                   max_tries=1000000, strict=True):
    """Seeded circles with integer radius U{r_min..r_max} placed until the
    solid fraction reaches 1 - porosity.  Circles keep ``margin`` cells from
    the domain edges (and from x_lo / x_hi) and from each other, so their
    discs never overlap and every classified boundary node is in range.
    Random sequential placement jams below some coverage; if the target is
    not reached after ``max_tries`` placements this raises, or with
    strict=False returns the circles placed so far.
    Returns the circle_list of setCirleblock: [((cx, cy), r), ...]."""
    rng = np.random.default_rng(seed)
    x_lo = margin if x_lo is None else x_lo
    x_hi = W - 1 - margin if x_hi is None else x_hi
    target = (1.0 - porosity) * H * W
    occ = np.zeros((H, W), dtype=bool)
    yy, xx = np.mgrid[0:H, 0:W]
    circles = []
    solid = 0
    tries = 0
    while solid < target:
        tries += 1
        if tries > max_tries:
            if not strict:
                break
            raise RuntimeError("porosity %.3f not reached" % porosity)
        r = int(rng.integers(r_min, r_max + 1))
        lo_x, hi_x = x_lo + r + 1, x_hi - r - 1
        lo_y, hi_y = margin + r + 1, H - 2 - margin - r
        if lo_x > hi_x or lo_y > hi_y:
            continue
        cx = int(rng.integers(lo_x, hi_x + 1))
        cy = int(rng.integers(lo_y, hi_y + 1))
        R = r + 1 + margin
        ya, yb = max(cy - R, 0), min(cy + R + 1, H)
        xa, xb = max(cx - R, 0), min(cx + R + 1, W)
        sub = (yy[ya:yb, xa:xb] - cy) ** 2 + (xx[ya:yb, xa:xb] - cx) ** 2 <= R * R
        if (occ[ya:yb, xa:xb] & sub).any():
            continue
        disc = (yy[ya:yb, xa:xb] - cy) ** 2 + (xx[ya:yb, xa:xb] - cx) ** 2 <= r * r
        occ[ya:yb, xa:xb] |= disc
        solid += int(disc.sum())
        circles.append(((cx, cy), r))
    return circles
