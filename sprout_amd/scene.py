"""Input generators for the fluid step: the solid geometry of a tank and its seeded liquid.

Restates the parts of the reference's scene and map code that produce the step's inputs:
  * Scene.rasterize  (src/scene.rs:215-325): node SDF, open face fractions, smoothed solid
    occupancy, open volume and the sealed-region table (union_depth :326-389, regions
    :391-432, seal_pockets :434-477), for the tank frame and boxes;
  * fill_rect + Rng  (src/maps.rs:456-490): rest-density particles on a jittered 2 x 2
    lattice per cell;
  * Particle::new    (src/sim/params.rs:60-62).
Arithmetic is float32 like the reference's f32 code.
"""
import math

import numpy as np

from ._native import PARTICLE_DTYPE, SPAWN_S, SPAWN_TEMP, PPC

WALL = 3.0          # scene.rs:8, thickness of the solid frame in cells
MIN_OPEN = 0.1      # scene.rs:125
MIN_POCKET = 16     # scene.rs:127
REGION_MAX = 31     # scene.rs:145
REGION_VENT = 256   # scene.rs:146
F32 = np.float32


def _rnd(x):
    """f32::round: half away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


class Box:
    """Shape::Box (scene.rs:66-70): centre, half extents, rotation in radians."""

    def __init__(self, c, half, angle=0.0):
        self.c = (F32(c[0]), F32(c[1]))
        self.half = (F32(half[0]), F32(half[1]))
        self.angle = F32(angle)

    def sdf(self, px, py):
        s, co = np.sin(self.angle, dtype=F32), np.cos(self.angle, dtype=F32)
        dx = (np.asarray(px, F32) - self.c[0]).astype(F32)
        dy = (np.asarray(py, F32) - self.c[1]).astype(F32)
        qx = (co * dx + s * dy).astype(F32)
        qy = (-s * dx + co * dy).astype(F32)
        ex = (np.abs(qx) - self.half[0]).astype(F32)
        ey = (np.abs(qy) - self.half[1]).astype(F32)
        out = np.hypot(np.maximum(ex, F32(0)), np.maximum(ey, F32(0))).astype(F32)
        return (out + np.minimum(np.maximum(ex, ey), F32(0))).astype(F32)

    def bound(self):
        return self.c, float(np.hypot(self.half[0], self.half[1]))


NO_TEMP = -1.0e9    # scene.rs: NO_TEMP, an insulating solid


class SolidField:
    """SolidField (scene.rs:110-122)."""

    def __init__(self, sdf, solid, region, temp=None):
        self.sdf = sdf        # float32[(H+1)*(W+1)]
        self.solid = solid    # float32[H*W, 4]
        self.region = region  # uint32[H*W]
        self.temp = temp      # float32[H*W] deg C, NO_TEMP where insulating; None: no solid has one


class Scene:
    def __init__(self, grid, wall=WALL):
        self.grid = (int(grid[0]), int(grid[1]))
        self.wall = F32(wall)
        self.shapes = []
        self.temps = []  # per shape: the temperature it is held at (deg C), or None

    def add(self, shape, temp=None):
        self.shapes.append(shape)
        self.temps.append(temp)
        return self

    def wall_sdf(self, px, py):
        W, H = F32(self.grid[0]), F32(self.grid[1])
        px = np.asarray(px, F32)
        py = np.asarray(py, F32)
        a = (px - self.wall).astype(F32)
        b = (W - self.wall - px).astype(F32)
        c = (py - self.wall).astype(F32)
        d = (H - self.wall - py).astype(F32)
        return np.minimum(np.minimum(np.minimum(a, b), c), d).astype(F32)

    def sdf(self, px, py):
        d = self.wall_sdf(px, py)
        for s in self.shapes:
            d = np.minimum(d, s.sdf(px, py))
        return d

    def rasterize(self):
        w, h = self.grid
        nw, nh = w + 1, h + 1
        jj, ii = np.mgrid[0:nh, 0:nw]
        sdf = self.wall_sdf(ii.astype(F32), jj.astype(F32))
        REACH = 6.0
        for s in self.shapes:
            c, r = s.bound()
            r = r + REACH
            i0 = max(int(math.floor(c[0] - r)), 0)
            i1 = min(max(int(math.ceil(c[0] + r)), 0), nw - 1)
            j0 = max(int(math.floor(c[1] - r)), 0)
            j1 = min(max(int(math.ceil(c[1] + r)), 0), nh - 1)
            i1 = min(max(i1, i0), nw - 1)
            j1 = min(max(j1, j0), nh - 1)
            sub = sdf[j0:j1 + 1, i0:i1 + 1]
            sdf[j0:j1 + 1, i0:i1 + 1] = np.minimum(sub, s.sdf(ii[j0:j1 + 1, i0:i1 + 1].astype(F32),
                                                                 jj[j0:j1 + 1, i0:i1 + 1].astype(F32)))
        sdf = _union_depth(sdf)
        # open volume from four bilinear sub-samples (scene.rs:245-259)
        d00, d10, d01, d11 = sdf[:-1, :-1], sdf[:-1, 1:], sdf[1:, :-1], sdf[1:, 1:]
        acc = np.zeros((h, w), F32)
        for fx, fy in ((0.25, 0.25), (0.75, 0.25), (0.25, 0.75), (0.75, 0.75)):
            fx, fy = F32(fx), F32(fy)
            v = ((d00 * (F32(1) - fx) + d10 * fx) * (F32(1) - fy) + (d01 * (F32(1) - fx) + d11 * fx) * fy).astype(F32)
            acc = (acc + np.clip(v / F32(0.5) + F32(0.5), F32(0), F32(1))).astype(F32)
        open_ = (acc * F32(0.25)).astype(F32)
        open_[open_ < MIN_OPEN] = 0.0
        _seal_pockets(open_, sdf, w, h)
        au, av = _faces(open_, sdf, w, h)
        # kernel-smoothed solid occupancy (scene.rs:277-296)
        K = (F32(1.0 / 6.0), F32(2.0 / 3.0), F32(1.0 / 6.0))
        pad = np.ones((h + 2, w + 2), F32)
        pad[1:-1, 1:-1] = (F32(1) - open_).astype(F32)
        occ = np.zeros((h, w), F32)
        for b in range(3):
            for a in range(3):
                occ = (occ + (K[a] * K[b]) * pad[b:b + h, a:a + w]).astype(F32)
        solid = np.stack([au, av, occ, open_], axis=-1).reshape(h * w, 4).astype(F32)
        region = _regions(solid, w, h)
        # cells in or against a shape held at a temperature (scene.rs:295-313)
        temp = None
        temps = list(self.temps) + [None] * (len(self.shapes) - len(self.temps))
        if any(t is not None for t in temps):
            temp = np.full(w * h, NO_TEMP, F32)
            for s, t in zip(self.shapes, temps):
                if t is None:
                    continue
                c, r = s.bound()
                r = r + 2.0
                i0 = max(int(math.floor(c[0] - r)), 0)
                i1 = min(max(int(math.ceil(c[0] + r)), 0), w)
                j0 = max(int(math.floor(c[1] - r)), 0)
                j1 = min(max(int(math.ceil(c[1] + r)), 0), h)
                for j in range(j0, j1):
                    xs = np.arange(i0, i1, dtype=F32) + F32(0.5)
                    d = s.sdf(xs, np.full(len(xs), F32(j) + F32(0.5), F32))
                    row = temp[j * w + i0:j * w + i1]
                    row[d < F32(0.75)] = F32(t)
        return SolidField(np.ascontiguousarray(sdf.reshape(-1), F32), np.ascontiguousarray(solid), region, temp)


def _open_fraction(d0, d1):
    """scene.rs:130-141, vectorised."""
    hi = np.maximum(d0, d1)
    lo = np.minimum(d0, d1)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = np.where((d0 > 0) & (d1 > 0), F32(1),
                     np.where((d0 <= 0) & (d1 <= 0), F32(0), (hi / (hi - lo)).astype(F32)))
    f = f.astype(F32)
    f[f < 0.05] = 0.0
    return f


def _faces(open_, sdf, w, h):
    """Open fraction of each cell's left and bottom face (scene.rs:263-273)."""
    here = open_ > 0
    left = np.zeros_like(here)
    left[:, 1:] = here[:, 1:] & (open_[:, :-1] > 0)
    below = np.zeros_like(here)
    below[1:, :] = here[1:, :] & (open_[:-1, :] > 0)
    fu = _open_fraction(sdf[:-1, :-1], sdf[1:, :-1])   # nodes (i, j) and (i, j+1)
    fv = _open_fraction(sdf[:-1, :-1], sdf[:-1, 1:])   # nodes (i, j) and (i+1, j)
    return np.where(left, fu, F32(0)).astype(F32), np.where(below, fv, F32(0)).astype(F32)


def _union_depth(sdf):
    """scene.rs:326-389: chamfer depth inside the solids."""
    nh, nw = sdf.shape
    FAR = 1e9
    inside = sdf < 0
    pad = np.zeros((nh + 2, nw + 2), bool)
    pad[1:-1, 1:-1] = ~inside
    edge = np.zeros_like(inside)
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            if di == 0 and dj == 0:
                continue
            edge |= pad[1 + dj:1 + dj + nh, 1 + di:1 + di + nw]
    depth = np.zeros((nh, nw), np.float32)
    depth[inside] = np.where(edge[inside], -sdf[inside], FAR)
    if not (inside & ~edge).any():
        return sdf
    D = np.float32(math.sqrt(2.0))
    one = np.float32(1.0)
    d = depth.tolist()
    for j in range(nh):
        row = d[j]
        up = d[j - 1] if j > 0 else None
        for i in range(nw):
            v = row[i]
            if v == 0.0:
                continue
            if i > 0 and row[i - 1] + 1.0 < v:
                v = float(np.float32(row[i - 1]) + one)
            if up is not None:
                if up[i] + 1.0 < v:
                    v = float(np.float32(up[i]) + one)
                if i > 0 and up[i - 1] + D < v:
                    v = float(np.float32(up[i - 1]) + D)
                if i + 1 < nw and up[i + 1] + D < v:
                    v = float(np.float32(up[i + 1]) + D)
            row[i] = v
    for j in range(nh - 1, -1, -1):
        row = d[j]
        dn = d[j + 1] if j + 1 < nh else None
        for i in range(nw - 1, -1, -1):
            v = row[i]
            if v == 0.0:
                continue
            if i + 1 < nw and row[i + 1] + 1.0 < v:
                v = float(np.float32(row[i + 1]) + one)
            if dn is not None:
                if dn[i] + 1.0 < v:
                    v = float(np.float32(dn[i]) + one)
                if i + 1 < nw and dn[i + 1] + D < v:
                    v = float(np.float32(dn[i + 1]) + D)
                if i > 0 and dn[i - 1] + D < v:
                    v = float(np.float32(dn[i - 1]) + D)
            row[i] = v
    depth = np.asarray(d, np.float32)
    out = sdf.copy()
    m = inside & (depth < FAR)
    out[m] = np.minimum(out[m], -depth[m])
    return out.astype(F32)


def _components(au, av, open_, w, h):
    """Groups of open cells connected through open faces: (labels numbered in scan order of
    their first cell, with -1 for closed cells, and the group count)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    n = w * h
    idx = np.arange(n).reshape(h, w)
    lu = au > 0
    lu[:, 0] = False
    lv = av > 0
    lv[0, :] = False
    rows = np.concatenate([idx[lu], idx[lv]])
    cols = np.concatenate([idx[lu] - 1, idx[lv] - w])
    g = coo_matrix((np.ones(len(rows), np.int8), (rows, cols)), shape=(n, n))
    _, lab = connected_components(g, directed=False)
    is_open = (open_.reshape(-1) > 0)
    lab = np.where(is_open, lab, -1)
    # renumber in order of each group's first cell
    first = {}
    order = []
    for k in np.flatnonzero(is_open):
        l = lab[k]
        if l not in first:
            first[l] = len(order)
            order.append(l)
    remap = np.full(lab.max() + 2 if len(lab) else 1, -1, np.int64)
    for l, r in first.items():
        remap[l] = r
    out = np.where(lab >= 0, remap[lab], -1)
    return out, len(order)


def _seal_pockets(open_, sdf, w, h):
    """scene.rs:434-477: enclosed groups of fewer than MIN_POCKET open cells become solid,
    and their node distances are pushed inside the solid."""
    au, av = _faces(open_, sdf, w, h)
    lab, n = _components(au, av, open_, w, h)
    if n == 0:
        return
    sizes = np.bincount(lab[lab >= 0], minlength=n)
    small = np.flatnonzero(sizes < MIN_POCKET)
    if len(small) == 0:
        return
    cells = np.flatnonzero(np.isin(lab, small))
    i, j = cells % w, cells // w
    open_[j, i] = 0.0
    for dj, di in ((0, 0), (0, 1), (1, 0), (1, 1)):
        sdf[j + dj, i + di] = np.minimum(sdf[j + dj, i + di], np.float32(-0.5))


def _regions(solid, w, h):
    """scene.rs:391-432: number the open regions (at most REGION_MAX), and mark each
    region's vent: the leftmost cell of its top row."""
    au = solid[:, 0].reshape(h, w)
    av = solid[:, 1].reshape(h, w)
    open_ = solid[:, 3].reshape(h, w)
    lab, n = _components(au.copy(), av.copy(), open_, w, h)
    region = np.where(lab >= 0, np.minimum(lab, REGION_MAX), 0).astype(np.uint32)
    for r in range(n):
        cells = np.flatnonzero(lab == r)
        top = cells // w
        hi = top.max()
        vent = cells[top == hi].min()
        region[vent] |= REGION_VENT
    return region


class Rng:
    """maps.rs:456-465: xorshift32, 24-bit uniform floats."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFF

    def next(self):
        x = self.s
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        self.s = x
        return np.float32(x >> 8) / np.float32(1 << 24)


def _xorshift(x):
    """One xorshift32 step (maps.rs:456-465) of every state in the uint32 array x."""
    x = x ^ (x << np.uint32(13))
    x = x ^ (x >> np.uint32(17))
    return x ^ (x << np.uint32(5))


_LANE = 4096
_JUMP = None


def _jump_columns():
    """The xorshift32 step is linear over GF(2): the images of the 32 unit vectors under
    _LANE steps are the columns of its _LANE-th power."""
    global _JUMP
    if _JUMP is None:
        x = (np.uint32(1) << np.arange(32, dtype=np.uint32)).astype(np.uint32)
        for _ in range(_LANE):
            x = _xorshift(x)
        _JUMP = [int(v) for v in x]
    return _JUMP


def _rng_stream(seed, n):
    """n successive Rng(seed).next() values: (x >> 8) / 2^24 is exact in float32. The
    stream is cut into lanes of _LANE values; each lane's starting state is the seed moved
    ahead by whole lanes (the step's _LANE-th power applied as a GF(2) matrix), and the
    lanes then step side by side."""
    x = seed & 0xFFFFFFFF
    lanes = max((n + _LANE - 1) // _LANE, 1)
    cols = _jump_columns()
    starts = np.empty(lanes, np.uint32)
    for j in range(lanes):
        starts[j] = x
        y = 0
        b = 0
        while x:
            if x & 1:
                y ^= cols[b]
            x >>= 1
            b += 1
        x = y
    out = np.empty((_LANE, lanes), np.uint32)
    v = starts
    for k in range(_LANE):
        v = _xorshift(v)
        out[k] = v
    vals = out.T.reshape(-1)[:n] >> np.uint32(8)
    return (vals.astype(np.float64) * (1.0 / (1 << 24))).astype(np.float32)


def new_particles(row, pos):
    """Particle::new(material, pos, 0) for every position (params.rs:60-62)."""
    out = np.zeros(len(pos), PARTICLE_DTYPE)
    out["pos"] = pos
    out["s"] = SPAWN_S
    out["kind"] = row + 1
    out["t"] = SPAWN_TEMP
    return out


def fill_rect(scene, rect, row, n_before=0):
    """fill_rect (maps.rs:468-490): rest-density particles in `rect` (cells), skipping
    solids. `n_before` is the number of particles already made by earlier fills (it seeds
    the jitter)."""
    sub = int(_rnd(math.sqrt(PPC)))
    step = np.float32(1.0 / sub)
    w, h = np.float32(scene.grid[0]), np.float32(scene.grid[1])
    W = scene.wall
    x0 = int(_rnd(float(np.float32(max(np.float32(rect[0]), W)) / step)))
    x1 = int(_rnd(float(np.float32(min(np.float32(rect[2]), np.float32(w - W))) / step)))
    y0 = int(_rnd(float(np.float32(max(np.float32(rect[1]), W)) / step)))
    y1 = int(_rnd(float(np.float32(min(np.float32(rect[3]), np.float32(h - W))) / step)))
    nx, ny = max(x1 - x0, 0), max(y1 - y0, 0)
    r = _rng_stream(0x9E3779B9 ^ ((n_before * 747796405 + 1) & 0xFFFFFFFF), 2 * nx * ny).reshape(ny, nx, 2)
    I, J = np.meshgrid(np.arange(x0, x1, dtype=np.float32), np.arange(y0, y1, dtype=np.float32))
    half, q = np.float32(0.5), np.float32(0.5)
    px = ((I + half + (r[..., 0] - half) * q) * step).astype(np.float32).reshape(-1)
    py = ((J + half + (r[..., 1] - half) * q) * step).astype(np.float32).reshape(-1)
    keep = scene.sdf(px, py) > np.float32(0.1)
    pos = np.stack([px[keep], py[keep]], axis=-1)
    return new_particles(row, pos)


def dam_break(grid, frac=(0.4, 0.6), row=0, wall=WALL):
    """Dam break: a block x < frac[0] W, y < frac[1] H of one liquid in the tank frame."""
    scene = Scene(grid, wall=wall)
    W, H = grid
    parts = fill_rect(scene, (0.0, 0.0, frac[0] * W, frac[1] * H), row)
    return scene, parts
