"""Refined (block-octree) runs of the oracle: level transfers and the recursive step.

TEST INFRASTRUCTURE ONLY (the checker of the library's refined plans; see oracle.py).

What it restates:
  covered / ghost blocks, fill and        src/octree_q19/src/generator.jl:692-780 (ghost donors: cubic
  restriction stencils                    4x4x4 Lagrange where complete, else trilinear renormalised over
                                          the admissible donors; restriction: cubic at the coarse centre,
                                          else the admissible children averaged), octree.jl:82-86
  transfer of a cell                      src/octree_q19/src/coupling.jl:86-233 in its role (time blend,
                                          stencil sum, non-equilibrium rescaled by the levels' relaxation
                                          times and 1/(1 - omega), positivity fallback to the equilibrium,
                                          kernel.jl:94-100), on the donors' full post-collision
                                          populations with the second-order non-equilibrium moments
                                          corrected through the Hermite term (the library's form for the
                                          cumulant collision: a rebuild from rho, u and Pi_neq alone would
                                          impose the polynomial equilibrium's higher moments)
  recursive step, initial coupling        coupling.jl:236-268 (step_recursive!, couple_initial!)
  level relaxation times                  src/octree_q19/src/config.jl:328-335 (acoustic scaling)
  WALE on refined levels                  kernel_cumulant.jl:496-595 per level (amr_sgs_scale per level,
                                          src/collision/timestep.jl:1088-1128); a transferred cell's
                                          velocity is that of its interpolated state, the role the
                                          parent-level velocities play in the reference's refined gather
                                          (pinned: tests/test_oracle_octree.py, the recorded velocity
                                          equals j / rho of the stored populations)

Floating-point operations are written one by one in numpy scalars of the working type (float32, or
float64 for the FP64 form), in the order the library's kernels use, so that results agree bit for bit.
Levels are numbered from 0 (coarsest).  A level's grid here holds the caller's blocks followed by its
ghost blocks; the caller's blocks carry the same cell indices as in the library.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import oracle as O

BS = 16
STEPPED, COVERED, GHOST = 0, 1, 2


def cubic_weights(t: float):
    """octree.jl:84"""
    return (-t * (1 - t) * (2 - t) / 6, (1 + t) * (1 - t) * (2 - t) / 2, (1 + t) * t * (2 - t) / 2,
            -(1 + t) * t * (1 - t) / 6)


@dataclass
class AugGrid:
    """A level with its ghost blocks (the attributes oracle.Grid reads)."""
    n: tuple
    periodic: tuple
    coords: np.ndarray      # [nb, 3] 1-based
    neighbor: np.ndarray    # [nb, 27] 1-based, 0 = none
    solid: np.ndarray       # [nb * 4096] uint8
    active: np.ndarray      # [nb] uint8: stepped
    status: np.ndarray      # [nb]
    n_caller: int
    bs: int = BS


def _neighbors(coords0: np.ndarray, nblk, periodic) -> np.ndarray:
    """[nb, 27] 1-based neighbour table of 0-based block coordinates (0 = none)."""
    look = {tuple(v): i for i, v in enumerate(coords0.tolist())}
    nb = coords0.shape[0]
    t = np.zeros((nb, 27), dtype=np.int32)
    for b in range(nb):
        x, y, z = coords0[b]
        for k in range(27):
            n = [x + k % 3 - 1, y + (k // 3) % 3 - 1, z + k // 9 - 1]
            ok = True
            for a in range(3):
                if periodic[a]:
                    n[a] %= nblk[a]
                elif not 0 <= n[a] < nblk[a]:
                    ok = False
            if ok:
                t[b, k] = look.get(tuple(n), -1) + 1
    return t


def build_levels(grids) -> list[AugGrid]:
    """Covered blocks, ghost blocks and neighbour tables of a block octree (callers' levels, coarsest first)."""
    L = len(grids)
    sets = [{tuple(v): i for i, v in enumerate((np.asarray(g.coords) - 1).tolist())} for g in grids]
    out = []
    for l, g in enumerate(grids):
        nblk = tuple(int(v) // BS for v in g.n)
        c0 = (np.asarray(g.coords, dtype=np.int64) - 1)
        nc = c0.shape[0]
        status = np.zeros(nc, dtype=np.uint8)
        if l + 1 < L:
            for b in range(nc):
                if tuple(2 * c0[b]) in sets[l + 1]:
                    status[b] = COVERED
        ghosts = []
        if l > 0:
            seen = set(sets[l])
            for b in range(nc):
                x, y, z = c0[b]
                for k in range(27):
                    if k == 13:
                        continue
                    n = [x + k % 3 - 1, y + (k // 3) % 3 - 1, z + k // 9 - 1]
                    ok = True
                    for a in range(3):
                        if g.periodic[a]:
                            n[a] %= nblk[a]
                        elif not 0 <= n[a] < nblk[a]:
                            ok = False
                    if ok and tuple(n) not in seen:
                        seen.add(tuple(n))
                        ghosts.append(n)
        gc = np.array(ghosts, dtype=np.int64).reshape(-1, 3)
        allc = np.concatenate([c0, gc]) if len(gc) else c0
        solid = np.zeros(allc.shape[0] * BS ** 3, dtype=np.uint8)
        if g.solid is not None:
            solid[:nc * BS ** 3] = np.asarray(g.solid, dtype=np.uint8)
        st = np.concatenate([status, np.full(len(gc), GHOST, dtype=np.uint8)])
        out.append(AugGrid(n=tuple(int(v) for v in g.n), periodic=tuple(bool(p) for p in g.periodic),
                           coords=(allc + 1).astype(np.int32), neighbor=_neighbors(allc, nblk, g.periodic),
                           solid=solid, active=(st == STEPPED).astype(np.uint8), status=st, n_caller=nc))
    return out


class _Dense:
    """Dense per-cell maps of a level: cell index, block status, solid."""

    def __init__(self, A: AugGrid):
        self.n = A.n
        self.per = A.periodic
        self.idx = np.full(A.n, -1, dtype=np.int64)
        self.st = np.full(A.n, 255, dtype=np.uint8)
        self.solid = np.zeros(A.n, dtype=bool)
        g = (np.asarray(A.coords, dtype=np.int64) - 1) * BS
        z, y, x = np.meshgrid(np.arange(BS), np.arange(BS), np.arange(BS), indexing="ij")
        loc = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
        for b in range(A.coords.shape[0]):
            p = g[b] + loc
            self.idx[p[:, 0], p[:, 1], p[:, 2]] = b * BS ** 3 + np.arange(BS ** 3)
            self.st[p[:, 0], p[:, 1], p[:, 2]] = A.status[b]
            self.solid[p[:, 0], p[:, 1], p[:, 2]] = A.solid[b * BS ** 3:(b + 1) * BS ** 3] != 0

    def at(self, p: np.ndarray):
        """p [..., 3] global cell coordinates -> (cell index or -1, admissible donor)."""
        p = p.copy()
        ok = np.ones(p.shape[:-1], dtype=bool)
        for a in range(3):
            if self.per[a]:
                p[..., a] %= self.n[a]
            else:
                ok &= (p[..., a] >= 0) & (p[..., a] < self.n[a])
        q = np.where(ok[..., None], p, 0)
        idx = np.where(ok, self.idx[q[..., 0], q[..., 1], q[..., 2]], -1)
        st = np.where(ok, self.st[q[..., 0], q[..., 1], q[..., 2]], 255)
        sol = np.where(ok, self.solid[q[..., 0], q[..., 1], q[..., 2]], True)
        adm = (idx >= 0) & (st != GHOST) & (st != 255) & ~sol
        return idx, st, adm


@dataclass
class Xfer:
    src: np.ndarray        # donor cells
    nob: np.ndarray        # donor not blended in time
    dst: np.ndarray        # written cells
    slot: np.ndarray       # [n_dst, 64] -1 = unused
    w: np.ndarray          # [n_dst, 64] float32


def _slots(cells: np.ndarray, used: np.ndarray):
    """unique donor list and the slot of every (cell, node)"""
    flat = cells[used]
    uniq, inv = np.unique(flat, return_inverse=True)
    slot = np.full(cells.shape, -1, dtype=np.int64)
    slot[used] = inv
    return uniq, slot


def build_fill(F: AugGrid, C: AugGrid) -> Xfer:
    """Ghost cells of F next to a stepped cell, donors on C (generator.jl:692-736)."""
    DF, DC = _Dense(F), _Dense(C)
    stepped = DF.st == STEPPED
    near = np.zeros_like(stepped)
    for d in range(27):
        sh = (d % 3 - 1, (d // 3) % 3 - 1, d // 9 - 1)
        s = stepped
        for a in range(3):
            if sh[a]:
                if F.periodic[a]:
                    s = np.roll(s, -sh[a], axis=a)
                else:
                    s = np.roll(s, -sh[a], axis=a)
                    sl = [slice(None)] * 3
                    sl[a] = slice(-1, None) if sh[a] > 0 else slice(0, 1)
                    s = s.copy()
                    s[tuple(sl)] = False
        near |= s
    ghost = DF.st == GHOST
    pts = np.argwhere(near & ghost)                  # x, y, z
    dst = DF.idx[pts[:, 0], pts[:, 1], pts[:, 2]]
    order = np.argsort(dst)
    pts, dst = pts[order], dst[order]
    n = len(dst)
    fx = pts / 2.0 - 0.25
    i0 = np.floor(fx).astype(np.int64)
    t = fx - i0
    wc = [np.array([cubic_weights(v) for v in t[:, a]]) for a in range(3)]      # [n, 4] per axis
    nodes = np.array([(q & 3, (q >> 2) & 3, q >> 4) for q in range(64)])
    cub_p = i0[:, None, :] - 1 + nodes[None, :, :]
    cidx, cst, cadm = DC.at(cub_p)
    complete = cadm.all(1)
    cells = np.full((n, 64), -1, dtype=np.int64)
    cov = np.zeros((n, 64), dtype=bool)
    w = np.zeros((n, 64), dtype=np.float32)
    keep = np.ones(n, dtype=bool)
    for i in range(n):
        if complete[i]:
            for q in range(64):
                di, dj, dk = nodes[q]
                cells[i, q] = cidx[i, q]
                cov[i, q] = cst[i, q] == COVERED
                w[i, q] = np.float32(wc[0][i, di] * wc[1][i, dj] * wc[2][i, dk])
        else:
            wt = [0.0] * 8
            tc = [-1] * 8
            tcov = [False] * 8
            wsum = 0.0
            for m in range(8):
                di, dj, dk = m & 1, (m >> 1) & 1, m >> 2
                c, s, a = DC.at(np.array([[i0[i, 0] + di, i0[i, 1] + dj, i0[i, 2] + dk]]))
                if not a[0]:
                    continue
                tc[m] = int(c[0])
                tcov[m] = s[0] == COVERED
                wt[m] = ((t[i, 0] if di else 1 - t[i, 0]) * (t[i, 1] if dj else 1 - t[i, 1])) * (t[i, 2] if dk else 1 - t[i, 2])
                wsum += wt[m]
            if not wsum > 0.0:
                keep[i] = False
                continue
            for m in range(8):
                if tc[m] < 0:
                    continue
                di, dj, dk = m & 1, (m >> 1) & 1, m >> 2
                q = (di + 1) + 4 * (dj + 1) + 16 * (dk + 1)
                cells[i, q] = tc[m]
                cov[i, q] = tcov[m]
                w[i, q] = np.float32(wt[m] / wsum)
    cells, cov, w, dst = cells[keep], cov[keep], w[keep], dst[keep]
    used = cells >= 0
    src, slot = _slots(cells, used)
    nob = np.zeros(len(src), dtype=bool)
    nob[slot[used]] = cov[used]
    return Xfer(src=src, nob=nob, dst=dst, slot=slot, w=w)


def build_restrict(C: AugGrid, F: AugGrid) -> Xfer:
    """Covered fluid cells of C, sources on F (octree.jl:86, coupling.jl:52-76)."""
    DC, DF = _Dense(C), _Dense(F)
    pts = np.argwhere((DC.st == COVERED) & ~DC.solid)
    dst = DC.idx[pts[:, 0], pts[:, 1], pts[:, 2]]
    order = np.argsort(dst)
    pts, dst = pts[order], dst[order]
    n = len(dst)
    wr = cubic_weights(0.5)
    nodes = np.array([(q & 3, (q >> 2) & 3, q >> 4) for q in range(64)])
    fidx, _, fadm = DF.at(2 * pts[:, None, :] - 1 + nodes[None, :, :])
    complete = fadm.all(1)
    cells = np.full((n, 64), -1, dtype=np.int64)
    w = np.zeros((n, 64), dtype=np.float32)
    wcub = np.array([np.float32(wr[di] * wr[dj] * wr[dk]) for di, dj, dk in nodes], dtype=np.float32)
    inner = np.array([(m & 1, (m >> 1) & 1, m >> 2) for m in range(8)])
    cidx, _, cadm = DF.at(2 * pts[:, None, :] + inner[None, :, :])
    keep = np.ones(n, dtype=bool)
    for i in range(n):
        if complete[i]:
            cells[i] = fidx[i]
            w[i] = wcub
        else:
            cnt = int(cadm[i].sum())
            if cnt == 0:
                keep[i] = False
                continue
            for m in range(8):
                if not cadm[i, m]:
                    continue
                di, dj, dk = inner[m]
                q = (di + 1) + 4 * (dj + 1) + 16 * (dk + 1)
                cells[i, q] = cidx[i, m]
                w[i, q] = np.float32(1.0 / cnt)
    cells, w, dst = cells[keep], w[keep], dst[keep]
    used = cells >= 0
    src, slot = _slots(cells, used)
    return Xfer(src=src, nob=np.zeros(len(src), dtype=bool), dst=dst, slot=slot, w=w)


# -------------------------------------------------------------------------------------------------
# Transfers in the working type T, operation by operation
# -------------------------------------------------------------------------------------------------
def _work(storage, f):
    """stored populations [27, ncell] -> working values (T)"""
    T = O.WORK_DTYPE[storage]
    if storage in ("f32", "f64"):
        return f.astype(T)
    return O.to_unshifted(storage, f).astype(T)


def gather(fw: np.ndarray, cells: np.ndarray) -> np.ndarray:
    """working populations of cells: [n, 27]"""
    return np.ascontiguousarray(fw[:, cells].T)


def _moments(T, acc):
    """rho, u and the second-order non-equilibrium moments of populations acc [n, 27], in the kernel's order"""
    lat = O.lattice()
    c = np.stack([lat[0], lat[1], lat[2]], 1).astype(int)
    z = np.zeros(acc.shape[0], dtype=T)
    v = [acc[:, k] for k in range(27)]
    rho = z.copy()
    for k in range(27):
        rho = rho + v[k]

    def signed(sel):
        a = z.copy()
        for k in range(27):
            s = sel(k)
            if s > 0:
                a = a + v[k]
            elif s < 0:
                a = a - v[k]
        return a
    jx, jy, jz = signed(lambda k: c[k][0]), signed(lambda k: c[k][1]), signed(lambda k: c[k][2])
    pxx, pyy, pzz = signed(lambda k: abs(c[k][0])), signed(lambda k: abs(c[k][1])), signed(lambda k: abs(c[k][2]))
    pxy = signed(lambda k: c[k][0] * c[k][1])
    pxz = signed(lambda k: c[k][0] * c[k][2])
    pyz = signed(lambda k: c[k][1] * c[k][2])
    ux, uy, uz = jx / rho, jy / rho, jz / rho
    third = T(1.0 / 3.0)
    rx, ry = rho * ux, rho * uy
    q = ((pxx - rho * third) - rx * ux, (pyy - rho * third) - ry * uy, (pzz - rho * third) - (rho * uz) * uz,
         pxy - rx * uy, pxz - rx * uz, pyz - ry * uz)
    return rho, ux, uy, uz, q


def apply(storage, f_dst: np.ndarray, X: Xfer, m: np.ndarray, scale, keep):
    """stencil sums of the gathered populations m [n_src, 27], the second-order non-equilibrium moments set
    to scale * keep times theirs, stored into f_dst (stored form, to nearest); returns the velocity
    (ux, uy, uz) of each written cell's interpolated state (what a WALE run records for the cell)"""
    T = m.dtype.type
    n = len(X.dst)
    if n == 0:
        return None
    # the library's order: 16 partial sums of 4 nodes each (nodes 4j .. 4j + 3 from zero, in node
    # order), then added as a pairwise tree (partial j plus partial j xor d for d = 1, 2, 4, 8)
    part = np.zeros((16, n, 27), dtype=T)
    for j in range(16):
        for q in range(4 * j, 4 * j + 4):
            sl = X.slot[:, q]
            ok = sl >= 0
            part[j, ok] = part[j, ok] + X.w[ok, q].astype(T)[:, None] * m[sl[ok]]
    for d in (1, 2, 4, 8):
        part = part + part[np.arange(16) ^ d]
    acc = part[0]
    rho, ux, uy, uz, (qxx, qyy, qzz, qxy, qxz, qyz) = _moments(T, acc)
    sk = T(scale) * T(keep)
    dxx, dyy, dzz = sk * qxx - qxx, sk * qyy - qyy, sk * qzz - qzz
    dxy, dxz, dyz = sk * qxy - qxy, sk * qxz - qxz, sk * qyz - qyz
    tr = (dxx + dyy) + dzz
    usq = (ux * ux + uy * uy) + uz * uz
    third, two, three, h = T(1.0 / 3.0), T(2), T(3), T(4.5)
    lat = O.lattice()
    w = lat[7] if T == np.float64 else lat[6]
    out = np.empty((27, n), dtype=T)
    eq = np.empty((27, n), dtype=T)
    bad = np.zeros(n, dtype=bool)
    for k in range(27):
        cx, cy, cz = T(lat[0][k]), T(lat[1][k]), T(lat[2][k])
        wk = T(w[k])
        cac = ((cx * cx * dxx + cy * cy * dyy) + cz * cz * dzz) + two * ((cx * cy * dxy + cx * cz * dxz) + cy * cz * dyz)
        out[k] = acc[:, k] + (wk * h) * (cac - tr * third)
        cu = (cx * ux + cy * uy) + cz * uz
        eq[k] = (wk * rho) * (((T(1) + three * cu) + h * cu * cu) - T(1.5) * usq)
        bad |= out[k] < T(1e-6) * wk
    out[:, bad] = eq[:, bad]
    if storage in ("f32", "f64"):
        f_dst[:, X.dst] = out.astype(f_dst.dtype)
    else:
        f_dst[:, X.dst] = O.from_unshifted(storage, out.astype(np.float64))
    return ux, uy, uz


# -------------------------------------------------------------------------------------------------
# The recursive run
# -------------------------------------------------------------------------------------------------
class OctreeRun:
    """An oracle run of a refined plan: per level the two buffers (A = read, B = other) of the
    augmented grid, the transfer lists and scales, and the recursive step."""

    def __init__(self, grids, storage: str, taus, u_inlet=0.0, amr_scales=None, **kw):
        self.storage = storage
        self.T = O.WORK_DTYPE[storage]
        self.levels = build_levels(grids)
        # WALE: each level keeps the velocity its gradient gather reads (the last step's for stepped
        # cells, the transfer's for the ghost and covered cells it writes) and its amr_sgs_scale
        self.wale = kw.get("c_wale", 0.0) != 0.0
        self.amr = [1.0] * len(grids) if amr_scales is None else [float(a) for a in amr_scales]
        self.V = [None] * len(grids)
        # a level collects the force when its caller gave an obstacle array (the library's rule)
        self.has_obstacle = [g.solid is not None for g in grids]
        self.G = [O.Grid(A) for A in self.levels]
        self.taus = [float(np.float32(t)) for t in taus]
        self.kw = dict(kw, u_inlet=u_inlet)
        L = len(grids)
        self.fill = [None] + [build_fill(self.levels[l], self.levels[l - 1]) for l in range(1, L)]
        self.restr = [build_restrict(self.levels[l], self.levels[l + 1]) for l in range(L - 1)] + [None]
        # relaxation rate omega1 of each level as the collision computes it (FP32 or FP64)
        self.omega = [self._omega(t) for t in self.taus]
        T = self.T
        self.fill_scale, self.fill_keep, self.rest_scale, self.rest_keep = [None] * L, [None] * L, [None] * L, [None] * L
        tau = [T(t) for t in self.taus]
        om = [T(o) for o in self.omega]

        def pcf(o):
            return (T(1) / (T(1) - o)) if abs(T(1) - o) > T(0.05) else T(0)
        for l in range(L):
            if l > 0:
                self.fill_scale[l] = (tau[l] / (T(2) * tau[l - 1])) * pcf(om[l - 1])
                self.fill_keep[l] = T(1) - om[l]
            if l + 1 < L:
                self.rest_scale[l] = ((T(2) * tau[l]) / tau[l + 1]) * pcf(om[l + 1])
                self.rest_keep[l] = T(1) - om[l]
        self.A = [None] * L
        self.B = [None] * L

    def _omega(self, tau):
        """omega1 without WALE: 1 / max(1/2 + 3 nu, 0.500001) with nu = (tau - 1/2)/3 in the working type"""
        T = self.T
        nu = (T(tau) - T(0.5)) / T(3)
        te = T(0.5) + T(3) * nu
        lo = T(np.float32(0.500001))
        return T(1) / (lo if te < lo else te)

    def init(self, rho_list, u_list, rho0=1.0, u0=(0.0, 0.0, 0.0)):
        for l, A in enumerate(self.levels):
            nc = A.n_caller * BS ** 3
            ncell = A.coords.shape[0] * BS ** 3
            rho = np.full(ncell, rho0, dtype=np.float64)
            u = np.zeros((3, ncell))
            u[:] = np.asarray(u0, dtype=np.float64)[:, None]
            if rho_list is not None:
                rho[:nc] = rho_list[l]
            if u_list is not None:
                u[:, :nc] = u_list[l]
            self.A[l] = O.init_eq(self.storage, self.G[l], rho, u)
            self.B[l] = self.A[l].copy()
            if self.wale:
                self.V[l] = u.astype(self.T)
        self.coupled = False

    def set_f(self, f_list):
        """unshifted populations [27, n_caller * 4096] (float64) of every level; ghost cells start at rest"""
        if self.wale:
            raise ValueError("a WALE run starts from init(): its velocity field is the initial one")
        w = O.lattice()[7]
        for l, A in enumerate(self.levels):
            ncell = A.coords.shape[0] * BS ** 3
            full = np.empty((27, ncell))
            nc = A.n_caller * BS ** 3
            full[:, :nc] = f_list[l]
            full[:, nc:] = w[:, None]
            self.A[l] = O.from_unshifted(self.storage, full)
            self.B[l] = self.A[l].copy()
        self.coupled = False

    def _restrict(self, l):
        X = self.restr[l]
        if X is None or len(X.dst) == 0:
            return
        m = gather(_work(self.storage, self.A[l + 1]), X.src)
        u = apply(self.storage, self.A[l], X, m, self.rest_scale[l], self.rest_keep[l])
        if self.wale:
            self.V[l][:, X.dst] = np.stack(u)

    def _fill(self, l, w1):
        X = self.fill[l]
        if X is None or len(X.dst) == 0:
            return
        mn = gather(_work(self.storage, self.A[l - 1]), X.src)
        if w1 != 1.0:
            T = self.T
            mo = gather(_work(self.storage, self.B[l - 1]), X.src)
            w1 = T(w1)
            w0 = T(1) - w1
            blend = mo * w0 + mn * w1
            m = np.where(X.nob[:, None], mn, blend)
        else:
            m = mn
        u = apply(self.storage, self.A[l], X, m, self.fill_scale[l], self.fill_keep[l])
        if self.wale:
            self.V[l][:, X.dst] = np.stack(u)

    def couple(self):
        L = len(self.levels)
        for l in range(L - 2, -1, -1):
            self._restrict(l)
            self.B[l] = self.A[l].copy()
        for l in range(1, L):
            self._fill(l, 1.0)
            self.B[l] = self.A[l].copy()
        self.coupled = True

    def _step_level(self, l, t, want_force):
        p = O.make_params(self.taus[l], time_step=t, arith=O.FAST, amr_sgs_scale=self.amr[l], **self.kw)
        _, F, mac = O.step(self.storage, self.G[l], p, self.A[l], out=self.B[l], want_force=want_force,
                           vel_in=self.V[l], want_macro=self.wale)
        self.A[l], self.B[l] = self.B[l], self.A[l]
        if self.wale:
            # the stepped cells' velocities; the others are rewritten by a transfer before they are read
            self.V[l] = mac[1]
        # ghost cells of the new read buffer are refilled before the level's next sub-step reads them
        return F

    def step_rec(self, l, t, last, force):
        has_solid = self.has_obstacle[l]
        F = self._step_level(l, t, force and last and has_solid)
        if force and last and has_solid:
            self.force[l] = F
        if l + 1 < len(self.levels):
            self.step_rec(l + 1, 2 * t, False, force)
            self._restrict(l)
            self._fill(l + 1, 0.5)
            self.step_rec(l + 1, 2 * t + 1, last, force)
            self._restrict(l)
            self._fill(l + 1, 1.0)

    def step(self, n, first=0, force=False):
        if not self.coupled:
            self.couple()
        self.force = {}
        for i in range(n):
            self.step_rec(0, first + i, i == n - 1, force)

    def caller_f(self, l):
        """the level's populations of the caller's blocks, stored form [27, n_caller * 4096]"""
        nc = self.levels[l].n_caller * BS ** 3
        return np.ascontiguousarray(self.A[l][:, :nc])
