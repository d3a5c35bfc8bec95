"""Oracle restatement of the obstacle front-end and bounce-back masks (TEST INFRASTRUCTURE ONLY).

Follows
  lattice_boltzmann/create_block.py:13-16   getRectangleblock (cv2.rectangle outline)
  lattice_boltzmann/create_block.py:19-31   getCicleblock     (cv2.circle outline)
  lattice_boltzmann/create_block.py:39-48   getCorner         (cv2.boundingRect)
  lattice_boltzmann/create_block.py:51-220  setCirleblock     (first-quadrant classifier + mirroring)
  lattice_boltzmann/create_block.py:395-407 setblock
  lattice_boltzmann/bounce_back.py:25-86    halfway_bounceback_rec  (mask construction + unions)
  lattice_boltzmann/bounce_back.py:89-167   halfway_bounceback_circle (mask unions)
  lattice_boltzmann/fingering_periodic.py:450-451 solid rule ``block_psi_all == 1``

numpy indexing semantics are kept on purpose: an index of -1 wraps to the
last row/column and an index equal to H or W raises IndexError, exactly as
the reference's mask assignments do (bounce_back.py:37-44,
create_block.py:158-204).

Pinning
-------
rectangles  closed form: cv2.rectangle + binary_fill_holes of an in-image
            rectangle is the filled inclusive box; getCorner returns its four
            corners.  The link unions are pinned by a hand-enumerated known
            answer (a single 6x6 block) in tests/test_geometry.py.
circles     *parity unpinned* for the raster: ``circle_outline`` is this
            project's restatement of an integer midpoint circle standing in
            for cv2.circle (cv2 is absent and the reference holds no golden
            masks).  The classifier on top of it is pinned by invariants:
            every classified node is fluid, carries exactly the neighbour
            pattern its class names, and the class sets are mirror-symmetric.
"""
import numpy as np
from scipy.ndimage import binary_fill_holes


# ----------------------------------------------------------------- rectangles
def rect_outline(H, W, p0, p1):
    """cv2.rectangle(block, p0, p1, 1) with thickness 1 on an HxW image
    (create_block.py:13-16).  Rectangles must lie inside the image."""
    x0, x1 = sorted((int(p0[0]), int(p1[0])))
    y0, y1 = sorted((int(p0[1]), int(p1[1])))
    if x0 < 0 or y0 < 0 or x1 >= W or y1 >= H:
        raise ValueError("rectangle outside the image")
    block = np.zeros((H, W), dtype=np.uint8)
    block[y0, x0:x1 + 1] = 1
    block[y1, x0:x1 + 1] = 1
    block[y0:y1 + 1, x0] = 1
    block[y0:y1 + 1, x1] = 1
    return block


def get_corner(block):
    """getCorner (create_block.py:39-48): bounding rectangle of the non-zero pixels."""
    ys, xs = np.nonzero(block)
    x, y = int(xs.min()), int(ys.min())
    w, h = int(xs.max()) - x + 1, int(ys.max()) - y + 1
    return {"top_left": (x, y + h - 1), "bottom_left": (x, y),
            "top_right": (x + w - 1, y + h - 1), "bottom_right": (x + w - 1, y)}


def setblock(H, W, rect_corner_list):
    """setblock (create_block.py:395-407) -> (block_psi_all int[H][W], corner_list)."""
    corner_list = []
    block_psi_all = np.zeros((H, W), dtype=int)
    for rect in rect_corner_list:
        block = rect_outline(H, W, rect[0], rect[1])
        block_psi = binary_fill_holes(block).astype(int)
        corner_list.append(get_corner(block))
        block_psi_all = block_psi_all + block_psi
    return block_psi_all, corner_list


def rect_link_masks(H, W, corner_list):
    """Per-direction bounce-back masks of halfway_bounceback_rec (bounce_back.py:28-86).

    Returns bool[9][H][W]; entry i is the union of the masks on which
    direction i is overwritten by the post-collision opposite population.
    """
    n_b = np.zeros((H, W), dtype=bool)
    s_b = np.zeros((H, W), dtype=bool)
    e_b = np.zeros((H, W), dtype=bool)
    w_b = np.zeros((H, W), dtype=bool)
    nw = np.zeros((H, W), dtype=bool)
    ne = np.zeros((H, W), dtype=bool)
    sw = np.zeros((H, W), dtype=bool)
    se = np.zeros((H, W), dtype=bool)
    for cor in corner_list:
        tl, bl, tr, br = cor["top_left"], cor["bottom_left"], cor["top_right"], cor["bottom_right"]
        n_b[tl[1] + 1, tl[0]:tr[0] + 1] = True          # bounce_back.py:37
        s_b[bl[1] - 1, tl[0]:tr[0] + 1] = True          # :38
        w_b[bl[1]:tl[1] + 1, tl[0] - 1] = True          # :39
        e_b[bl[1]:tl[1] + 1, tr[0] + 1] = True          # :40
        nw[tl[1] + 1, tl[0] - 1] = True                 # :41
        ne[tl[1] + 1, tr[0] + 1] = True                 # :42
        sw[bl[1] - 1, bl[0] - 1] = True                 # :43
        se[bl[1] - 1, br[0] + 1] = True                 # :44
    links = np.zeros((9, H, W), dtype=bool)
    links[1] = e_b                                       # :47-49
    links[2] = n_b                                       # :50-52
    links[3] = w_b                                       # :53-55
    links[4] = s_b                                       # :56-58
    links[5] = ne | n_b | e_b                            # :59-65
    links[6] = nw | n_b | w_b                            # :66-72
    links[7] = sw | s_b | w_b                            # :73-79
    links[8] = se | s_b | e_b                            # :80-86
    return links


# -------------------------------------------------------------------- circles
def circle_outline(H, W, cx, cy, r):
    """Stand-in for cv2.circle(block, (cx, cy), r, 1) (create_block.py:19-21).

    Integer midpoint circle: march dy = 0, 1, ... in the octant dx >= dy,
    keeping err = dx^2 + dy^2 - r^2 and stepping dx inward once whenever err
    turns positive; plot the 8 mirror images, clipped to the image.
    PARITY UNPINNED against cv2 (absent here).
    """
    block = np.zeros((H, W), dtype=np.uint8)
    dx, dy = int(r), 0
    err = 0
    while dx >= dy:
        for px, py in ((dx, dy), (dy, dx), (-dy, dx), (-dx, dy),
                       (-dx, -dy), (-dy, -dx), (dy, -dx), (dx, -dy)):
            X, Y = cx + px, cy + py
            if 0 <= X < W and 0 <= Y < H:
                block[Y, X] = 1
        dy += 1
        err += 2 * dy - 1
        if err > 0:
            err -= 2 * dx - 1
            dx -= 1
    return block


def set_circle_block(H, W, circle_list):
    """setCirleblock (create_block.py:51-220).

    Returns (block_psi_all int[H][W], side_list, concave_list, convex_list)
    with the reference's list orders:
      side_list    = [top, bottom, right, left]          (:211-214)
      concave_list = [tr, tl, br, bl]                     (:207-210)
      convex_list  = [tr, tl, br, bl]                     (:215-218)
    """
    side_top, side_left, side_right, side_bottom = [], [], [], []
    vex = {"tr": [], "tl": [], "br": [], "bl": []}
    cave = {"tr": [], "tl": [], "br": [], "bl": []}
    block_psi_all = np.zeros((H, W), dtype=int)
    for (center, r) in circle_list:
        x, y = int(center[0]), int(center[1])
        r = int(r)
        block = circle_outline(H, W, x, y, r)
        block_psi_all = binary_fill_holes(block).astype(int) + block_psi_all
        t_top, t_left, t_vex, t_cave = [], [], [], []
        jj = r
        for i in range(1, r + 2):                        # :82
            for j in range(jj, 0, -1):                   # :86
                if block[y + i, x + j]:                  # :90-92
                    jj = j + 1
                    break
                fb = block[y + i - 1, x + j]             # :94
                fl = block[y + i, x - 1 + j]             # :95
                flb = block[y + i - 1, x - 1 + j]        # :96
                if fl and fb:
                    t_cave.append((x + j, y + i))
                elif fl:
                    t_left.append((x + j, y + i))
                elif (not fl) and fb:
                    t_top.append((x + j, y + i))
                elif flb:
                    t_vex.append((x + j, y + i))
        for t in t_top:                                  # :114-124
            side_top.append((2 * x - t[0], t[1]))
            side_bottom.append((t[0], 2 * y - t[1]))
            side_bottom.append((2 * x - t[0], 2 * y - t[1]))
        side_top.append((x, y + r + 1))
        side_top.extend(t_top)
        side_bottom.append((x, y - r - 1))
        for l in t_left:                                 # :129-135
            side_right.append((2 * x - l[0], l[1]))
            side_left.append((l[0], 2 * y - l[1]))
            side_right.append((2 * x - l[0], 2 * y - l[1]))
        side_left.extend(t_left)
        side_left.append((x + r + 1, y))
        side_right.append((x - r - 1, y))
        for name, tmp in (("cave", t_cave), ("vex", t_vex)):   # :139-152
            d = cave if name == "cave" else vex
            for c in tmp:
                d["tl"].append((2 * x - c[0], c[1]))
                d["br"].append((c[0], 2 * y - c[1]))
                d["bl"].append((2 * x - c[0], 2 * y - c[1]))
            d["tr"].extend(tmp)

    def to_mask(pts):                                    # :158-204
        m = np.zeros((H, W), dtype=bool)
        for p in pts:
            m[p[1], p[0]] = True
        return m

    concave_list = [to_mask(cave[k]) for k in ("tr", "tl", "br", "bl")]
    side_list = [to_mask(side_top), to_mask(side_bottom), to_mask(side_right), to_mask(side_left)]
    convex_list = [to_mask(vex[k]) for k in ("tr", "tl", "br", "bl")]
    return block_psi_all, side_list, concave_list, convex_list


def circle_link_masks(side_list, concave_list, convex_list):
    """Per-direction unions of halfway_bounceback_circle (bounce_back.py:89-167)."""
    n_b, s_b, e_b, w_b = side_list                       # :90-93
    nw_vex, ne_vex, sw_vex, se_vex = convex_list         # :94-97
    nw_cave, ne_cave, sw_cave, se_cave = concave_list    # :98-101
    H, W = n_b.shape
    links = np.zeros((9, H, W), dtype=bool)
    links[1] = w_b | nw_cave | sw_cave                   # :104-110
    links[2] = n_b | nw_cave | ne_cave                   # :111-117
    links[3] = e_b | ne_cave | se_cave                   # :118-124
    links[4] = s_b | sw_cave | se_cave                   # :125-131
    links[5] = nw_vex | nw_cave | n_b | w_b              # :132-140
    links[6] = ne_vex | ne_cave | n_b | e_b              # :141-149
    links[7] = se_vex | se_cave | s_b | e_b              # :150-158
    links[8] = sw_vex | sw_cave | s_b | w_b              # :159-167
    return links


def fluid_mask(block_psi_all):
    """Fluid mask of fingering_periodic.py:450-451 (solid iff the count is exactly 1)."""
    return np.logical_not(block_psi_all == 1)
