"""Block topology and obstacle masks on the host (numpy), in the reference's BlockLevel conventions.

A level is a list of 16^3-cell blocks.  Block b has 1-based block coordinates (map_x, map_y, map_z)
(active_coords of src/geometry/blocks.jl) and a 27-entry row of the neighbour table, stored
column-major n_blocks x 27 with 1-based entries and 0 = none (src/geometry/domain_topology.jl:919-955).
Cell (x, y, z) of block b (0-based inside the block) is element b*bs^3 + z*bs^2 + y*bs + x of a field,
and direction k of the populations adds k*n_cells: the column-major f[x,y,z,b,k] of the reference.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BS = 16


def lattice_d3q27():
    """D3Q27 velocity set in the reference order k = (cx+1) + 3(cy+1) + 9(cz+1) (lattice.jl:10-59).

    Returns (c [27,3] int, w [27] float64, opp, mirror_y, mirror_z).
    """
    c = np.array([(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)], dtype=np.int64)
    d2 = (c * c).sum(1)
    w = np.select([d2 == 0, d2 == 1, d2 == 2, d2 == 3], [8 / 27, 2 / 27, 1 / 54, 1 / 216])
    idx = {tuple(v): i for i, v in enumerate(c)}
    opp = np.array([idx[tuple(-v)] for v in c])
    mir_y = np.array([idx[(v[0], -v[1], v[2])] for v in c])
    mir_z = np.array([idx[(v[0], v[1], -v[2])] for v in c])
    return c, w, opp, mir_y, mir_z


def build_neighbor_table(coords: np.ndarray, nblk: tuple[int, int, int],
                         periodic: tuple[bool, bool, bool] = (False, False, False)) -> np.ndarray:
    """27-neighbour table of a block list, restating build_neighbor_table (domain_topology.jl:919-955).

    coords: [nb, 3] 1-based block coordinates; nblk: blocks per axis.  Returns int32 [nb, 27] with
    1-based block indices and 0 where no block exists; a periodic axis wraps (the reference wraps y
    only, under periodic_y).  The caller passes the Fortran-ordered (column-major) array to the ABI.
    """
    coords = np.asarray(coords, dtype=np.int64)
    nb = coords.shape[0]
    ptr = np.zeros(tuple(nblk), dtype=np.int32)
    ptr[coords[:, 0] - 1, coords[:, 1] - 1, coords[:, 2] - 1] = np.arange(1, nb + 1, dtype=np.int32)
    table = np.zeros((nb, 27), dtype=np.int32)
    k = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                n = coords + np.array([dx, dy, dz])
                ok = np.ones(nb, dtype=bool)
                for a in range(3):
                    if periodic[a]:
                        n[:, a] = (n[:, a] - 1) % nblk[a] + 1
                    else:
                        ok &= (n[:, a] >= 1) & (n[:, a] <= nblk[a])
                vals = np.zeros(nb, dtype=np.int32)
                vals[ok] = ptr[n[ok, 0] - 1, n[ok, 1] - 1, n[ok, 2] - 1]
                table[:, k] = vals
                k += 1
    return table


@dataclass
class BlockGrid:
    """One level of blocks covering the extent `n` (cells) of one or more domains."""
    n: tuple[int, int, int]
    periodic: tuple[bool, bool, bool]
    coords: np.ndarray                 # [nb, 3] int32, 1-based block coordinates
    neighbor: np.ndarray               # [nb, 27] int32, 1-based, 0 = none
    solid: np.ndarray | None = None    # [nb*bs^3] uint8 or None
    bs: int = BS
    block_domain: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    @property
    def n_blocks(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_cells(self) -> int:
        return self.n_blocks * self.bs ** 3

    def cell_coords(self) -> np.ndarray:
        """[n_cells, 3] 0-based global cell coordinates of every stored cell."""
        bs = self.bs
        z, y, x = np.meshgrid(np.arange(bs), np.arange(bs), np.arange(bs), indexing="ij")
        loc = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
        org = (self.coords.astype(np.int64) - 1) * bs
        return (org[:, None, :] + loc[None, :, :]).reshape(-1, 3)

    def inside(self) -> np.ndarray:
        """[n_cells] bool: the cell lies inside the extent (blocks may overhang it)."""
        g = self.cell_coords()
        return (g[:, 0] < self.n[0]) & (g[:, 1] < self.n[1]) & (g[:, 2] < self.n[2])

    def to_blocks(self, dense: np.ndarray) -> np.ndarray:
        """Dense field [nx, ny, nz, ...] (x first) -> block layout [n_cells, ...]; overhang cells get 0."""
        g = self.cell_coords()
        ins = self.inside()
        out = np.zeros((self.n_cells,) + dense.shape[3:], dtype=dense.dtype)
        out[ins] = dense[g[ins, 0], g[ins, 1], g[ins, 2]]
        return out

    def to_dense(self, blocks: np.ndarray) -> np.ndarray:
        """Block layout [n_cells, ...] -> dense [nx, ny, nz, ...]."""
        g = self.cell_coords()
        ins = self.inside()
        out = np.zeros(tuple(self.n) + blocks.shape[1:], dtype=blocks.dtype)
        out[g[ins, 0], g[ins, 1], g[ins, 2]] = blocks[ins]
        return out


def uniform_grid(n, periodic=(False, False, False), bs: int = BS) -> BlockGrid:
    """Every block of a box of n = (nx, ny, nz) cells, block x fastest; blocks overhang a non-multiple extent."""
    n = tuple(int(v) for v in n)
    periodic = tuple(bool(p) for p in periodic)
    for a in range(3):
        if periodic[a] and n[a] % bs:
            raise ValueError("a periodic axis needs an extent that is a multiple of the block size")
    nblk = tuple(-(-v // bs) for v in n)
    bz, by, bx = np.meshgrid(np.arange(nblk[2]), np.arange(nblk[1]), np.arange(nblk[0]), indexing="ij")
    coords = np.stack([bx.ravel(), by.ravel(), bz.ravel()], axis=1).astype(np.int32) + 1
    table = build_neighbor_table(coords, nblk, periodic)
    return BlockGrid(n=n, periodic=periodic, coords=coords, neighbor=table, bs=bs)


def sphere_mask(grid: BlockGrid, center, radius: float) -> np.ndarray:
    """Solid iff the cell centre lies within the sphere: apply_analytic_sphere_obstacle!
    (src/geometry/domain_generation.jl:111-154) in lattice units, cell i (1-based) centred at i - 1/2,
    i.e. 0-based cell j centred at j + 1/2."""
    g = grid.cell_coords().astype(np.float64) + 0.5
    d = g - np.asarray(center, dtype=np.float64)[None, :]
    m = (d * d).sum(1) <= radius * radius
    return (m & grid.inside()).astype(np.uint8)


def icosphere(center, radius: float, subdivisions: int = 3):
    """Triangulated sphere surface, the mesh the reference reads from an STL file (gpu_mesh of
    src/forces/surface.jl): an icosahedron whose edges are split `subdivisions` times (20 * 4^s
    triangles), vertices on the sphere.  Returns Float32 centres [3, n] (triangle centroids),
    unit outward normals [3, n] and areas [n], computed in FP64."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(int(subdivisions)):
        mid: dict = {}

        def midpoint(a: int, b: int) -> int:
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                mid[key] = len(v) - 1
            return mid[key]
        nf = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(v) * float(radius)
    F = np.array(faces)
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    cr = np.cross(p1 - p0, p2 - p0)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    nrm = cr / np.linalg.norm(cr, axis=1)[:, None]
    cen = (p0 + p1 + p2) / 3.0
    flip = (nrm * cen).sum(1) < 0.0
    nrm[flip] *= -1.0
    cen = cen + np.asarray(center, dtype=np.float64)[None, :]
    return (np.ascontiguousarray(cen.T, dtype=np.float32), np.ascontiguousarray(nrm.T, dtype=np.float32),
            area.astype(np.float32))


def link_solid(grid: BlockGrid) -> np.ndarray:
    """[27, n_cells] bool: the cell x + c_k is a solid cell of the level (inside the extent, in a present
    block; periodic axes wrap through the neighbour table).  Only blocks holding solid cells and their
    neighbours can have such links; the others are skipped."""
    c, _, _, _, _ = lattice_d3q27()
    bs = grid.bs
    bs3 = bs ** 3
    out = np.zeros((27, grid.n_cells), dtype=bool)
    if grid.solid is None:
        return out
    solid = np.asarray(grid.solid) != 0
    nbr = np.asarray(grid.neighbor, dtype=np.int64)
    has = solid.reshape(grid.n_blocks, bs3).any(1)
    cand = has.copy()
    for col in range(27):
        nb = nbr[has, col]
        cand[nb[nb > 0] - 1] = True
    blocks = np.nonzero(cand)[0]
    if blocks.size == 0:
        return out
    z, y, x = np.meshgrid(np.arange(bs), np.arange(bs), np.arange(bs), indexing="ij")
    loc0 = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    b = np.repeat(blocks, bs3)
    loc = np.tile(loc0, (blocks.size, 1))
    g = (np.asarray(grid.coords, dtype=np.int64)[b] - 1) * bs + loc
    cells = b * bs3 + (loc[:, 2] * bs + loc[:, 1]) * bs + loc[:, 0]
    for k in range(27):
        t = loc + c[k]
        o = np.floor_divide(t, bs)
        col = (o[:, 0] + 1) + 3 * (o[:, 1] + 1) + 9 * (o[:, 2] + 1)
        nb = np.where(col == 13, b + 1, nbr[b, col])
        ok = nb > 0
        gt = g + c[k]
        for a in range(3):
            if not grid.periodic[a]:
                ok &= (gt[:, a] >= 0) & (gt[:, a] < grid.n[a])
        tl = t - o * bs
        cell = (nb - 1) * bs3 + (tl[:, 2] * bs + tl[:, 1]) * bs + tl[:, 0]
        out[k, cells[ok]] = solid[cell[ok]]
    return out


def sphere_qmap(grid: BlockGrid, center, radius: float) -> np.ndarray:
    """Wall distances q of the links from fluid cells to solid cells of a sphere (float16 [27, n_cells],
    0 elsewhere), for interpolated bounce-back.  q is the distance from the cell centre to the sphere
    along c_k in units of |c_k|: the first ray-sphere hit beyond 1e-10 max(R, 1)
    (ray_sphere_first_hit, bouzidi_math.jl:127-150; compute_q_for_sphere_cell :159-204, without the
    symmetry-plane rule), kept for 0 < q <= 1 on links whose end cell is solid
    (_bouzidi_qmap_block!, bouzidi_setup.jl:91-139), rounded to float16 as the reference's q-map."""
    c, _, _, _, _ = lattice_d3q27()
    nc = grid.n_cells
    q = np.zeros((27, nc), dtype=np.float16)
    ends = link_solid(grid)
    fluid = grid.inside() & ((np.asarray(grid.solid) == 0) if grid.solid is not None else True)
    sel_any = ends.any(0) & fluid
    idx = np.nonzero(sel_any)[0]
    if idx.size == 0:
        return q
    bs = grid.bs
    bidx = idx // bs ** 3
    r = idx % bs ** 3
    loc = np.stack([r % bs, (r // bs) % bs, r // bs ** 2], 1)
    p = ((np.asarray(grid.coords, dtype=np.int64)[bidx] - 1) * bs + loc).astype(np.float64) + 0.5
    oc = p - np.asarray(center, dtype=np.float64)[None, :]
    cc = (oc * oc).sum(1) - radius * radius
    eps = 1.0e-10 * max(radius, 1.0)
    for k in range(27):
        if k == 13:
            continue
        sel = ends[k, idx]
        if not sel.any():
            continue
        cm = float(np.sqrt((c[k] * c[k]).sum()))
        d = c[k].astype(np.float64) / cm
        bq = oc[sel] @ d
        disc = bq * bq - cc[sel]
        with np.errstate(invalid="ignore"):
            sq = np.sqrt(disc)
        t1, t2 = -bq - sq, -bq + sq
        t = np.where(t1 > eps, t1, np.where(t2 > eps, t2, np.inf))
        t = np.where(disc < 0.0, np.inf, t)
        qk = t / cm
        keep = (qk > 0.0) & (qk <= 1.0)
        vals = np.zeros(int(sel.sum()), dtype=np.float64)
        vals[keep] = qk[keep]
        q[k, idx[sel]] = vals.astype(np.float16)
    return q


def ellipsoid_mask(grid: BlockGrid, center, semi_axes, rotation: np.ndarray) -> np.ndarray:
    """Solid iff the cell centre lies inside the ellipsoid with the given semi-axes along the columns
    of `rotation` (a 3x3 rotation matrix)."""
    g = grid.cell_coords().astype(np.float64) + 0.5
    d = (g - np.asarray(center, dtype=np.float64)[None, :]) @ np.asarray(rotation, dtype=np.float64)
    q = ((d / np.asarray(semi_axes, dtype=np.float64)[None, :]) ** 2).sum(1)
    return ((q <= 1.0) & grid.inside()).astype(np.uint8)


def batch_grids(grids: list[BlockGrid]) -> BlockGrid:
    """Independent boxes as one block list: the blocks of grids[i] follow those of grids[i - 1], each
    box's neighbour table is offset to the combined numbering (no links between boxes), and
    block_domain gives each block's box (the domain index of a batched level)."""
    coords, nbrs, solids, bdom = [], [], [], []
    off = 0
    for i, g in enumerate(grids):
        coords.append(np.asarray(g.coords, dtype=np.int32))
        nb = np.asarray(g.neighbor, dtype=np.int32).copy()
        nb[nb > 0] += off
        nbrs.append(nb)
        solids.append(np.zeros(g.n_cells, dtype=np.uint8) if g.solid is None else np.asarray(g.solid, dtype=np.uint8))
        bdom.append(np.full(g.n_blocks, i, dtype=np.int32))
        off += g.n_blocks
    return BlockGrid(n=(0, 0, 0), periodic=(False, False, False), coords=np.concatenate(coords),
                     neighbor=np.concatenate(nbrs), solid=np.concatenate(solids), bs=grids[0].bs,
                     block_domain=np.concatenate(bdom))


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """A rotation matrix from a uniformly distributed unit quaternion (Shoemake's method)."""
    u1, u2, u3 = rng.uniform(0.0, 1.0, 3)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    w, x, y, z = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def ellipsoid_frontal_area(semi_axes, rotation: np.ndarray, direction=(1.0, 0.0, 0.0)) -> float:
    """Area of the ellipsoid's projection on the plane normal to `direction`: pi a b c sqrt(n^T M n) with
    M = R diag(1/a^2, 1/b^2, 1/c^2) R^T."""
    a = np.asarray(semi_axes, dtype=np.float64)
    R = np.asarray(rotation, dtype=np.float64)
    n = np.asarray(direction, dtype=np.float64)
    M = R @ np.diag(1.0 / a ** 2) @ R.T
    return float(np.pi * a.prod() * np.sqrt(n @ M @ n))


def channel_walls(grid: BlockGrid, axis: int = 1) -> np.ndarray:
    """Solid layers at the first and last cell along `axis` (a plane channel)."""
    g = grid.cell_coords()
    m = (g[:, axis] == 0) | (g[:, axis] == grid.n[axis] - 1)
    return (m & grid.inside()).astype(np.uint8)


def block_octree(coarse_extent, refined, periodic=(False, False, False), bs: int = BS) -> list[BlockGrid]:
    """Levels of a block octree from the refined blocks of each level but the finest.

    coarse_extent: level-0 extent in cells; level l has extent coarse_extent * 2^l.  refined[l] is a set of
    0-based block coordinates of level l whose eight children exist on level l + 1 (level 0 holds every
    block of its extent).  The sets are first closed for the nesting a refined plan needs: the blocks next
    to a block of level l >= 1 exist on level l - 1's refined set's children, i.e. their parents are
    refined too.  Returns the levels coarsest first, blocks in lexicographic order (z slowest), each with
    its neighbour table and no obstacle cells.
    """
    ext0 = np.asarray(coarse_extent, dtype=np.int64)
    if np.any(ext0 % bs):
        raise ValueError("the coarse extent must be a multiple of the block size")
    L = len(refined) + 1
    nblk = [tuple(int(v) for v in (ext0 << l) // bs) for l in range(L)]
    ref = [set(map(tuple, r)) for r in refined]
    for l in range(L - 2, 0, -1):
        extra = set()
        for (x, y, z) in ref[l]:
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        n = [x + dx, y + dy, z + dz]
                        ok = True
                        for a in range(3):
                            if periodic[a]:
                                n[a] %= nblk[l][a]
                            elif not 0 <= n[a] < nblk[l][a]:
                                ok = False
                        if ok:
                            extra.add((n[0] // 2, n[1] // 2, n[2] // 2))
        ref[l - 1] |= extra
    # each finer level only refines blocks that exist on it
    for l in range(1, L - 1):
        ref[l] = {b for b in ref[l] if (b[0] // 2, b[1] // 2, b[2] // 2) in ref[l - 1]}
    levels = []
    for l in range(L):
        if l == 0:
            blocks = np.stack(np.meshgrid(*[np.arange(v) for v in nblk[0]], indexing="ij"), -1).reshape(-1, 3)
        else:
            par = np.array(sorted(ref[l - 1]), dtype=np.int64).reshape(-1, 3)
            ch = np.array([(i, j, k) for k in (0, 1) for j in (0, 1) for i in (0, 1)], dtype=np.int64)
            blocks = (2 * par[:, None, :] + ch[None, :, :]).reshape(-1, 3)
        order = np.lexsort((blocks[:, 0], blocks[:, 1], blocks[:, 2]))     # z slowest, x fastest
        coords = blocks[order].astype(np.int32) + 1
        n = tuple(int(v) for v in ext0 << l)
        table = build_neighbor_table(coords, nblk[l], tuple(bool(p) for p in periodic))
        levels.append(BlockGrid(n=n, periodic=tuple(bool(p) for p in periodic), coords=coords, neighbor=table,
                                bs=bs))
    return levels


def sphere_octree(coarse_extent, n_levels: int, center_finest, radius_finest: float, thresholds,
                  periodic=(False, False, False), bs: int = BS) -> list[BlockGrid]:
    """Block octree refined around a sphere by distance to its surface (block-AMR set-up).

    coarse_extent: level-0 extent in cells; level l has extent coarse_extent * 2^l.  center_finest and
    radius_finest are in cells of the finest level.  thresholds[l - 1] (finest cells) selects the
    blocks of level l >= 1: a level-(l-1) block is refined when its box comes within thresholds[l - 1]
    of the sphere's surface (or lies inside the sphere).  The refined sets are closed for nesting
    (block_octree).  Returns the levels coarsest first, each with its sphere mask (cell centres j + 1/2
    at that level's spacing).
    """
    L = n_levels
    ext0 = np.asarray(coarse_extent, dtype=np.int64)
    cf = np.asarray(center_finest, dtype=np.float64)
    nblk = [tuple(int(v) for v in (ext0 << l) // bs) for l in range(L)]

    def box_dist(coords_0b: np.ndarray, l: int) -> np.ndarray:
        """distance (finest cells) from the sphere centre to the nearest point of each level-l block"""
        sc = float(2 ** (L - 1 - l) * bs)            # block edge in finest cells
        lo = coords_0b * sc
        hi = lo + sc
        q = np.clip(cf[None, :], lo, hi)
        return np.sqrt(((q - cf[None, :]) ** 2).sum(1))

    ref = []
    for l in range(L - 1):
        nb = nblk[l]
        g = np.stack(np.meshgrid(*[np.arange(v) for v in nb], indexing="ij"), -1).reshape(-1, 3)
        sel = box_dist(g, l) <= radius_finest + float(thresholds[l])
        ref.append({tuple(v) for v in g[sel]})
    levels = block_octree(coarse_extent, ref, periodic, bs)
    for l, g in enumerate(levels):
        s = 2.0 ** (L - 1 - l)
        g.solid = sphere_mask(g, cf / s, radius_finest / s)
    return levels
