// Device code of the fused stream-collide step (templates), included by the per-storage translation
// units oal_step_<storage>.hip so that the four storage forms compile in parallel.
//
// Work decomposition of the step kernels: one work-item per cell, one workgroup per 16 x 16 plane of
// a 16^3 block (256 lanes = 4 wavefronts of 16 x 4 cells), or per several planes in the pipelined
// kernels.  Populations are block-major: each block holds its 27 direction arrays of 4096 cells
// contiguously (pop_index), so the 64 lanes of a wavefront touch 4 consecutive 64-byte rows of each
// of the block's direction arrays.  A workgroup's block, plane, domain and neighbour-block indices are
// uniform, which keeps them in scalar registers.
//
// Reference: src/collision/kernel_cumulant.jl:29-912 (stream_collide_cumulant!).
#pragma once

#include <hip/hip_runtime.h>

#include "oal_kernels.h"

namespace oal {

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X.  Remap the workgroup index so
// that each XCD receives a contiguous range of items: neighbouring planes and blocks, which share
// halo cache lines, then meet in the same L2.
__device__ __forceinline__ uint32_t xcd_item(uint32_t i, uint32_t n) {
    const uint32_t q = n >> 3, r = n & 7u, x = i & 7u, j = i >> 3;
    return x < r ? x * (q + 1u) + j : r * (q + 1u) + (x - r) * q + j;
}

// The list position and first plane of a pipelined workgroup (the grid holds n_blocks * groups
// workgroups, groups = planes per block / planes per workgroup).  BLOCKS_FASTEST walks the blocks
// fastest, so that x-neighbour blocks, which share the rows of each other's x-shifted accesses, run at
// the same time on the same L2: on a 256^3 level the two-cells-per-lane FP32 kernel 627 -> 615 us and
// the AA odd half-step's reads 1.14 -> 1.06 of the algorithmic bytes.  The others walk each block's
// planes fastest (the WALE kernel 835 us against 859, the AA even half-step faster too).
struct PipeItem {
    uint32_t pos;
    int z0;
};
template <bool BLOCKS_FASTEST>
__device__ __forceinline__ PipeItem pipe_item(uint32_t groups, int np) {
    const uint32_t litem = xcd_item(blockIdx.x, gridDim.x);
    if constexpr (BLOCKS_FASTEST) {
        const uint32_t nblk = gridDim.x / groups;
        return PipeItem{litem % nblk, (int)(litem / nblk) * np};
    } else {
        return PipeItem{litem / groups, (int)(litem % groups) * np};
    }
}

// how the population arriving in one direction is obtained
enum : int {
    SRC_REG = 0,     // pulled from the source cell
    SRC_INLET = 1,   // equilibrium at (1, u_inlet)                       kernel_cumulant.jl:235-239
    SRC_OWN = 2,     // the cell's own population, same direction (outlet) kernel_cumulant.jl:240-245
    SRC_OPP = 3,     // own opposite population: no-slip wall            kernel_cumulant.jl:270-271, 340-341
    SRC_MIRY = 4,    // own y-mirrored population: free slip             kernel_cumulant.jl:342-343
    SRC_MIRZ = 5,    // own z-mirrored population: free slip             kernel_cumulant.jl:344-345
    SRC_REST = 6,    // rest weight (no neighbour block on level 1)      kernel_cumulant.jl:389-390
    SRC_SOLID = 7,   // halfway bounce-back from a solid source cell
    SRC_ZIMG = 8,    // exact specular z wall: the image cell's mirror_z population  kernel_cumulant.jl:272-321
    SRC_TFS = 9      // transverse freestream: own population blended toward feq(1, u_inlet)  :323-339
};

enum : int { PAT_AB = 0, PAT_AA_EVEN = 1, PAT_AA_ODD = 2 };

// Uniform (per-workgroup) and per-lane state of one cell
struct CellCtx {
    uint32_t b, flags, own;
    uint32_t lmask;   // bit k: the source of direction k is solid; bit 27: the cell is solid
    int lx, ly, z, gx, gy, gz;
    int ext0, ext1, ext2, per0, per1, per2;
    int zwall, ymin_wall;
};

// The source of direction K for this cell: classification and, for SRC_REG / SRC_SOLID, the cell index.
// GEN = false is the interior path (block flags all clear): every source is a regular cell of this
// block or of one of its 26 neighbours, which all exist.
// FX: the exact specular z wall and the transverse freestream are compiled in (levels that use neither
// run a form without them)
template <int K, bool GEN, bool REG = false, bool FX = true>
__device__ __forceinline__ int classify(const StepArgs& a, const CellCtx& c, uint32_t& src) {
    constexpr int cx = dcx(K), cy = dcy(K), cz = dcz(K);
    if constexpr (GEN) {
        if (c.flags & BF_EDGE) {
            const int sgx = c.gx - cx, sgy = c.gy - cy, sgz = c.gz - cz;
            const bool xl = cx > 0 && !c.per0 && sgx < 0;
            const bool xh = cx < 0 && !c.per0 && sgx >= c.ext0;
            const bool yl = cy > 0 && !c.per1 && sgy < 0;
            const bool yh = cy < 0 && !c.per1 && sgy >= c.ext1;
            const bool zl = cz > 0 && !c.per2 && sgz < 0;
            const bool zh = cz < 0 && !c.per2 && sgz >= c.ext2;
            if (xl) return SRC_INLET;
            if (xh) return SRC_OWN;
            if (FX && c.zwall == 1 && (zl || zh)) {
                // the source beyond the wall mirrors the wall-row cell (x - cx, y - cy, z): that cell's
                // population mirror_z(K) arrives; with no such cell on the level (beyond a non-periodic
                // y face, or an absent block) the cell's own mirror_z(K)
                const int ix = c.lx - cx, iy = c.ly - cy, igy = c.gy - cy;
                const int ox = ix < 0 ? -1 : (ix >= BS ? 1 : 0), oy = iy < 0 ? -1 : (iy >= BS ? 1 : 0);
                int32_t nbk = (int32_t)c.b;
                if (ox | oy) nbk = a.nbr_bm[(size_t)c.b * 32 + (ox + 1) + 3 * (oy + 1) + 9];
                if ((!c.per1 && (igy < 0 || igy >= c.ext1)) || nbk < 0) return SRC_MIRZ;
                src = (uint32_t)nbk * BS3 + ((uint32_t)c.z << 8) + ((uint32_t)(iy & (BS - 1)) << 4) + (uint32_t)(ix & (BS - 1));
                return SRC_ZIMG;
            }
            if (c.zwall != 0 && (zl || zh)) return SRC_OPP;
            if (FX && a.cp.tfs && (yl || yh || zl || zh)) return SRC_TFS;
            if (yl && c.ymin_wall == 1) return SRC_OPP;
            if (yl || yh) return SRC_MIRY;
            if (zl || zh) return SRC_MIRZ;
        }
    }
    if constexpr (REG) {
        // regular addressing: the caller forms the address; only the solid test remains
        if constexpr (GEN) {
            if ((c.flags & BF_SOLID) && a.cp.bb_mode == 1 && ((c.lmask >> K) & 1u)) return SRC_SOLID;
        }
        return SRC_REG;
    }
    const int sx = c.lx - cx, sy = c.ly - cy, sz = c.z - cz;
    int32_t nbk = (int32_t)c.b;
    if constexpr (cx != 0 || cy != 0 || cz != 0) {
        // candidate neighbour blocks are uniform (scalar) values; the lane picks one by its crossing
        const int oz = cz == 0 ? 0 : (sz < 0 ? -1 : (sz >= BS ? 1 : 0));
        const int dz = 9 * (oz + 1);
        const int32_t* nt = a.nbr_bm + (size_t)c.b * 32;
        int32_t n_c = oz == 0 ? (int32_t)c.b : nt[4 + dz];
        if constexpr (!GEN) n_c = max(n_c, 0);
        int32_t pick = n_c;
        if constexpr (cx != 0 && cy != 0) {
            int32_t n_x = nt[(1 - cx) + 3 + dz], n_y = nt[1 + 3 * (1 - cy) + dz], n_xy = nt[(1 - cx) + 3 * (1 - cy) + dz];
            if constexpr (!GEN) { n_x = max(n_x, 0); n_y = max(n_y, 0); n_xy = max(n_xy, 0); }
            const bool crx = (unsigned)sx >= (unsigned)BS, cry = (unsigned)sy >= (unsigned)BS;
            pick = crx ? (cry ? n_xy : n_x) : (cry ? n_y : n_c);
        } else if constexpr (cx != 0) {
            int32_t n_x = nt[(1 - cx) + 3 + dz];
            if constexpr (!GEN) n_x = max(n_x, 0);
            pick = ((unsigned)sx >= (unsigned)BS) ? n_x : n_c;
        } else if constexpr (cy != 0) {
            int32_t n_y = nt[1 + 3 * (1 - cy) + dz];
            if constexpr (!GEN) n_y = max(n_y, 0);
            pick = ((unsigned)sy >= (unsigned)BS) ? n_y : n_c;
        }
        nbk = pick;
    }
    if constexpr (GEN) {
        if (nbk < 0) return SRC_REST;   // absent block inside the extent (BF_MISSING); never an address
    }
    src = (uint32_t)nbk * BS3 + ((uint32_t)(sz & (BS - 1)) << 8) + ((uint32_t)(sy & (BS - 1)) << 4) + (uint32_t)(sx & (BS - 1));
    if constexpr (GEN) {
        if ((c.flags & BF_SOLID) && a.cp.bb_mode == 1 && ((c.lmask >> K) & 1u)) return SRC_SOLID;
    }
    return SRC_REG;
}

// Regular interior blocks: the block's neighbours sit at lexicographic offsets (b +- 1, +- nbx,
// +- nbx*nby), so a source cell's address is arithmetic.  Each lane forms, once, the byte offsets of
// its nine in-plane source classes (cx, cy) - own cell shifted by -c in x and y, into the neighbour
// block where it crosses - relative to the start of its own block; every direction's access adds a
// uniform part (the block, the direction slot, the z plane) to its scalar base pointer.
template <class SP>
struct RegAddr {
    int32_t v[3][3];           // [cx+1][cy+1]: byte offset of cell (x - cx, y - cy, zref) from the own block
    int64_t blk0;              // element offset of the own block: b * 27 * BS3
    int64_t plane_elems;       // elements between consecutive block layers (nbx * nby * 27 * BS3)
    int zref;                  // the plane the offsets were formed for (other planes: uniform shift)
    __device__ __forceinline__ void init(const CellCtx& c, int nbx, int nby) {
        blk0 = (int64_t)c.b * (27 * BS3);
        plane_elems = (int64_t)nbx * nby * (27 * BS3);
        zref = c.z;
        const int32_t local = (int32_t)(c.own - c.b * BS3);   // zref * BS2 + ly * BS + lx
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int cx = i - 1, cy = j - 1;
                const int sx = c.lx - cx, sy = c.ly - cy;
                const int dx = sx < 0 ? -1 : (sx >= BS ? 1 : 0);
                const int dy = sy < 0 ? -1 : (sy >= BS ? 1 : 0);
                const int32_t idx = (dx + dy * nbx) * (27 * BS3) + local + (sx - BS * dx - c.lx) + BS * (sy - BS * dy - c.ly);
                v[i][j] = idx * (int32_t)sizeof(typename SP::st);
            }
    }
    // uniform base of direction-slot `slot` for plane z of the block and a source shifted by -cz in z
    __device__ __forceinline__ const char* base(const void* f, int slot, int z, int cz) const {
        const int sz = z - cz;
        const int dz = sz < 0 ? -1 : (sz >= BS ? 1 : 0);
        const int64_t e = blk0 + (int64_t)slot * BS3 + dz * plane_elems + (int64_t)(sz - BS * dz - zref) * BS2;
        return (const char*)f + e * (int64_t)sizeof(typename SP::st);
    }
};

// Faces of non-periodic axes that a block of the regular list touches.  Such blocks are whole (not
// partial), so a face is where the block's origin is 0 or its far side is the extent.  Cells on a face
// have sources outside the extent: the pipelined kernel applies their rules itself (FACE_RULES) or
// skips these cells (FACE_SKIP), the one-plane regular kernel skips them; the face-cell kernel steps
// the skipped cells.
struct FaceBits {
    bool xlo, xhi, ylo, yhi, zlo, zhi;
};

__device__ __forceinline__ FaceBits face_bits(const StepArgs& a, uint32_t b, const DevDomain* dm) {
    const int ox = a.borg[3 * b], oy = a.borg[3 * b + 1], oz = a.borg[3 * b + 2];
    FaceBits f;
    f.xlo = !dm->per[0] && ox == 0;
    f.xhi = !dm->per[0] && ox + BS == dm->ext[0];
    f.ylo = !dm->per[1] && oy == 0;
    f.yhi = !dm->per[1] && oy + BS == dm->ext[1];
    f.zlo = !dm->per[2] && oz == 0;
    f.zhi = !dm->per[2] && oz + BS == dm->ext[2];
    return f;
}

__device__ __forceinline__ bool face_lane(const FaceBits& f, int lx, int ly) {
    return (f.xlo && lx == 0) || (f.xhi && lx == BS - 1) || (f.ylo && ly == 0) || (f.yhi && ly == BS - 1);
}

__device__ __forceinline__ bool face_plane(const FaceBits& f, int z) {
    return (f.zlo && z == 0) || (f.zhi && z == BS - 1);
}

// Population J of this cell's own post-collision state of the previous step: slot (x, J) after an
// AB step, slot (x, opp J) after an AA even step.
template <class SP, int PAT, int J>
__device__ __forceinline__ typename SP::T own_prev(const StepArgs& a, const CellCtx& c) {
    constexpr int slot = PAT == PAT_AB ? J : dopp(J);
    return SP::ld((const typename SP::st*)a.fin + pop_index(c.own, slot), wgt<typename SP::T>(J));
}

template <int J>
using dir_c = std::integral_constant<int, J>;

// transverse freestream: (1 - b) f_K + b feq_K(1, u_inlet) (kernel_cumulant.jl:333-339)
template <int K, class T>
__device__ __forceinline__ T tfs_value(const StepArgs& a, T donor, float u_in) {
    const T b = T(a.cp.tfs_blend);
    const T feq = feq_auto(T(1), T(u_in), T(0), T(0), wgt<T>(K), T(dcx(K)), T(dcy(K)), T(dcz(K)), a.cp.compressible);
    return (T(1) - b) * donor + b * feq;
}

// The value of a boundary rule for direction K.  `own(dir_c<j>)` yields population j of the cell's
// own post-collision state of the step the rule refers to.
template <int K, class T, class OWN>
__device__ __forceinline__ T boundary_value(int kind, const StepArgs& a, const CellCtx& c, float u_in, int32_t step,
                                            OWN&& own) {
    switch (kind) {
        case SRC_INLET: {
            // the noise amplitude in the working precision (FP32 for the FP32-arithmetic forms, as the reference)
            const float turb = a.cp.inlet_turb;
            const T noise = turb > 0.0f ? T(inlet_noise(c.gy + 1, c.gz + 1, step, 1234)) * T(turb) * T(u_in) : T(0);
            const T u_inst = T(u_in) + noise;
            return feq_auto(T(1), u_inst, T(0), T(0), wgt<T>(K), T(dcx(K)), T(dcy(K)), T(dcz(K)), a.cp.compressible);
        }
        case SRC_OWN: return own(dir_c<K>{});
        case SRC_OPP:
        case SRC_SOLID: return own(dir_c<dopp(K)>{});
        case SRC_MIRY: return own(dir_c<dmir_y(K)>{});
        case SRC_MIRZ:
        case SRC_ZIMG: return own(dir_c<dmir_z(K)>{});   // (the pull gather reads the image cell itself)
        case SRC_TFS: return tfs_value<K, T>(a, own(dir_c<K>{}), u_in);
        default: return wgt<T>(K);
    }
}

// The population of the cell's own post-collision state that a boundary rule for direction K returns
// (the rest weight, SRC_REST, does not depend on it)
template <int K>
__device__ __forceinline__ int rule_dir(int kind) {
    switch (kind) {
        case SRC_OPP:
        case SRC_SOLID: return dopp(K);
        case SRC_MIRY: return dmir_y(K);
        case SRC_MIRZ:
        case SRC_ZIMG: return dmir_z(K);
        default: return K;
    }
}

// workgroup sum of 3 doubles into part[item*4 + 0..2]; every lane of the workgroup must call it
__device__ __forceinline__ void wg_sum3(double v0, double v1, double v2, double* part, uint32_t item) {
    __shared__ double red[3][WG / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v0 += __shfl_xor(v0, off, 64);
        v1 += __shfl_xor(v1, off, 64);
        v2 += __shfl_xor(v2, off, 64);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { red[0][wave] = v0; red[1][wave] = v1; red[2][wave] = v2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int w = 0; w < WG / 64; ++w) { s0 += red[0][w]; s1 += red[1][w]; s2 += red[2][w]; }
        part[(size_t)item * 4 + 0] = s0;
        part[(size_t)item * 4 + 1] = s1;
        part[(size_t)item * 4 + 2] = s2;
    }
}

// WALE: central differences of the previous step's velocity over the six face neighbours; a solid or
// absent neighbour contributes the cell's own velocity (physics_utils.jl:1981-2035, 2217-2233)
template <class T>
__device__ __forceinline__ T wale_of(const StepArgs& a, const CellCtx& c) {
    const T* v = (const T*)a.vel_in;
    const uint32_t n = a.ncell;
    T un[6][3];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int ddx = q == 0 ? 1 : (q == 1 ? -1 : 0);
        const int ddy = q == 2 ? 1 : (q == 3 ? -1 : 0);
        const int ddz = q == 4 ? 1 : (q == 5 ? -1 : 0);
        bool ok = true;
        if (c.flags & BF_EDGE) {
            const int ngx = c.gx + ddx, ngy = c.gy + ddy, ngz = c.gz + ddz;
            if (!c.per0 && (ngx < 0 || ngx >= c.ext0)) ok = false;
            if (!c.per1 && (ngy < 0 || ngy >= c.ext1)) ok = false;
            if (!c.per2 && (ngz < 0 || ngz >= c.ext2)) ok = false;
        }
        const int sx = c.lx + ddx, sy = c.ly + ddy, sz = c.z + ddz;
        const int ox = sx < 0 ? -1 : (sx >= BS ? 1 : 0);
        const int oy = sy < 0 ? -1 : (sy >= BS ? 1 : 0);
        const int oz = sz < 0 ? -1 : (sz >= BS ? 1 : 0);
        int32_t nbk = (int32_t)c.b;
        if (ok && (ox | oy | oz)) nbk = a.nbr_bm[(size_t)c.b * 32 + (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)];
        if (nbk < 0) ok = false;
        // the face neighbour x + e is the source of the direction -e: its solid bit in the link mask
        constexpr int kq[6] = {12, 14, 10, 16, 4, 22};
        uint32_t idx = c.own;
        if (ok) {
            idx = (uint32_t)nbk * BS3 + ((uint32_t)(sz & (BS - 1)) << 8) + ((uint32_t)(sy & (BS - 1)) << 4) + (uint32_t)(sx & (BS - 1));
            if ((c.lmask >> kq[q]) & 1u) idx = c.own;
        }
        un[q][0] = v[idx]; un[q][1] = v[n + idx]; un[q][2] = v[2 * (size_t)n + idx];
    }
    const T h = T(0.5);
    return wale_nu(h * (un[0][0] - un[1][0]), h * (un[2][0] - un[3][0]), h * (un[4][0] - un[5][0]),
                   h * (un[0][1] - un[1][1]), h * (un[2][1] - un[3][1]), h * (un[4][1] - un[5][1]),
                   h * (un[0][2] - un[1][2]), h * (un[2][2] - un[3][2]), h * (un[4][2] - un[5][2]), T(a.cp.c_wale));
}

// The fused stream-collide step of one cell.  PAT_AB pulls from fin into fout (the reference's
// scheme); PAT_AA_EVEN / PAT_AA_ODD are the two half-steps of the in-place AA pattern on one buffer:
//   even: read f_k(x) from slot (x, k), write f*_k(x) to slot (x, opp k);
//   odd:  read f_k(x) from slot (x - c_k, opp k), write f*_k(x) to slot (x + c_k, k).
// Boundary rules of the odd step's writes: where x + c_k is not a cell that steps, the rule of the
// opposite direction is evaluated from the fresh post-collision state and stored in (x, opp k), the
// slot the following even step reads, so the even step needs no boundary logic.
//
// Momentum exchange (OAL_STEP_FORCE): summed over the links whose populations this step's collision
// produced, i.e. the exchange of the streaming that follows it: 2 (f_i* - w_i) c_i for halfway
// bounce-back, (f_i*(t) + f_i*(t-1) - 2 w_i) c_i for the reference's solid reflection (the solid cell
// returns last step's population), i toward a solid cell (src/octree_q19/src/stream.jl:248-256).
// Interpolated bounce-back (the linear BFL rule of Bouzidi, Firdaouss and Lallemand 2001;
// bouzidi_kernel.jl:204-370 with its D3Q27 settings) in pull form, on the populations of one cell whose
// sources are solid (bits of solid_links).  For such a direction K the link toward the wall is
// I = opp(K), f[K] holds its halfway value f*_I(x) and q is the wall distance of link I:
//   q < 1/2:  2q f*_I(x) + (1 - 2q) f*_I(x - c_I)      (x - c_I: the pull source of direction I,
//             used when it is a fluid cell of the level; otherwise the fallback of a.bfl_fallback)
//   q >= 1/2: f*_I(x) / (2q) + (2q - 1) / (2q) f*_K(x)
// so the reflected population arrives at x itself.  q outside (q_min, 1] keeps halfway bounce-back.
// The link's momentum exchange (f*_I(x) + arriving value) c_I is summed here, at the gather.
template <class SP>
__device__ __forceinline__ void bfl_links(const StepArgs& a, const CellCtx& c, typename SP::T (&f)[27],
                                          uint32_t solid_links, uint32_t reg_links, bool want_force,
                                          double& F0, double& F1, double& F2) {
    using T = typename SP::T;
    using st = typename SP::st;
    const st* fi = (const st*)a.fin;
    const uint16_t* qb = a.qmap + (size_t)a.lslot[c.b] * (27 * BS3) + (c.own - c.b * BS3);
    T rho_f = T(0);
    if (a.bfl_fallback == 1) {
        // density of the cell's own post-collision populations (the wall-equilibrium fallback)
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            rho_f += SP::ld(fi + pop_index(c.own, K), wgt<T>(K));
        });
    }
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int I = dopp(K);
        constexpr T w = wgt<T>(K);
        if ((solid_links >> K) & 1u) {
            const T fk = f[K];
            T val = fk;
            const float q = (float)__builtin_bit_cast(_Float16, qb[(size_t)I * BS3]);
            if (q > a.q_min && q <= 1.0f) {
                const T tq = T(q);
                if (q < 0.5f) {
                    T fff = fk;
                    if ((reg_links >> I) & 1u) fff = f[I];
                    else if (a.bfl_fallback == 1) fff = rho_f * w;
                    else if (a.bfl_fallback == 2) fff = SP::ld(fi + pop_index(c.own, K), w);
                    const T c1 = T(2) * tq;
                    val = c1 * fk + (T(1) - c1) * fff;
                } else {
                    const T inv = T(1) / (T(2) * tq);
                    const T c2 = (T(2) * tq - T(1)) * inv;
                    val = inv * fk + c2 * SP::ld(fi + pop_index(c.own, K), w);
                }
            }
            if (want_force) {
                const double m = ((double)fk - (double)w) + ((double)val - (double)w);
                F0 += m * dcx(I); F1 += m * dcy(I); F2 += m * dcz(I);
            }
            f[K] = val;
        }
    });
}

// BFL: the level has a q-map (interpolated bounce-back); a separate instantiation, so that the boundary
// kernels of levels without one keep their register budget.
template <class SP, int PAT, bool WALE, bool GEN, bool REG = false, bool BFL = false, bool SPONGE = false, bool FX = true>
__device__ __forceinline__ void step_cell(const StepArgs& a, const CellCtx& c, const DevDomain* dm, bool active,
                                          bool solid_own, double& F0, double& F1, double& F2) {
    using T = typename SP::T;
    using st = typename SP::st;
    if (!active) return;
    const uint32_t b = c.b;
    const float u_in = dm->u_in;
    st* fo = (st*)a.fout;
    const st* fi = (const st*)a.fin;
    const uint32_t n = a.ncell;
    RegAddr<SP> ra;
    if constexpr (REG) ra.init(c, dm->nbx, dm->nby);

    T f[27];
    typename SP::raw raw[27];     // the loaded populations, decoded after all loads
    uint32_t inlet_links = 0, rest_links = 0, solid_links = 0, reg_links = 0, tfs_links = 0;
    // AA odd half-step: the even step's copy of the populations with cx = -1 of the cell's row on the
    // outlet face (they sit in slots that the west neighbour overwrites during this half-step)
    const st* ocopy = nullptr;
    if constexpr (PAT == PAT_AA_ODD && GEN) {
        if (c.flags & BF_EDGE) {
            const int32_t s = a.oslot[b];
            if (s >= 0) ocopy = (const st*)a.obuf + (size_t)s * 9 * BS2 + c.ly * BS + c.z;
        }
    }
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr T w = wgt<T>(K);
        T val;
        if constexpr (REG && !GEN && PAT == PAT_AA_EVEN) {
            val = SP::ld((const st*)(ra.base(fi, K, c.z, 0) + ra.v[1][1]), w);
        } else if constexpr (REG && !GEN) {
            constexpr int slot = PAT == PAT_AB ? K : dopp(K);
            val = SP::ld((const st*)(ra.base(fi, slot, c.z, dcz(K)) + ra.v[dcx(K) + 1][dcy(K) + 1]), w);
        } else if constexpr (PAT == PAT_AA_EVEN) {
            // every direction loads the own slot K, so the 27 loads are in flight together; an inlet
            // link's value replaces it afterwards, evaluated in full precision as the pull step does
            // rather than read back from a slot that went through the storage rounding
            raw[K] = SP::ld_raw(fi + pop_index(c.own, K));
            if constexpr (GEN && dcx(K) > 0)
                inlet_links |= ((c.flags & BF_EDGE) && !c.per0 && c.gx == 0 ? 1u : 0u) << K;
            val = T(0);   // set below
        } else if constexpr (PAT == PAT_AB) {
            // branch-free gather: each direction reads one population - its source, or the own-cell
            // population its boundary rule returns (an inlet or rest link reads its own slot, unused) -
            // so the 27 loads are all in flight together; the inlet and rest rules replace the value
            // afterwards.  The rules' directions (opposite, mirrors) have the weight of K.
            uint32_t src = 0;
            const int kind = classify<K, GEN, REG, FX>(a, c, src);
            const int jd = kind == SRC_REG ? K : rule_dir<K>(kind);
            const st* p;
            if constexpr (REG)
                p = kind == SRC_REG ? (const st*)(ra.base(fi, K, c.z, dcz(K)) + ra.v[dcx(K) + 1][dcy(K) + 1])
                                    : fi + pop_index(FX && kind == SRC_ZIMG ? src : c.own, jd);
            else
                p = fi + pop_index(kind == SRC_REG || (FX && kind == SRC_ZIMG) ? src : c.own, jd);
            raw[K] = SP::ld_raw(p);
            inlet_links |= (kind == SRC_INLET ? 1u : 0u) << K;
            rest_links |= (kind == SRC_REST ? 1u : 0u) << K;
            if constexpr (FX) tfs_links |= (kind == SRC_TFS ? 1u : 0u) << K;
            solid_links |= (kind == SRC_SOLID ? 1u : 0u) << K;
            reg_links |= (kind == SRC_REG ? 1u : 0u) << K;
            val = T(0);   // set below
        } else {
            // AA odd half-step, branch-free as the AB gather: each direction loads one slot - the
            // source cell's slot opp K, the outlet copy (the own population of direction K, cx = -1),
            // or the own slot opp j of the population j its boundary rule returns; an inlet or rest
            // link reads the own slot K, unused, which no other cell writes in this half-step (its
            // writer would be the absent cell x - c_K) - and the inlet, rest and freestream rules
            // replace the value afterwards
            static_assert(PAT == PAT_AA_ODD, "the AA odd half-step");
            uint32_t src = 0;
            const int kind = classify<K, GEN, REG, FX>(a, c, src);
            const st* p = fi + pop_index(c.own, kind == SRC_INLET || kind == SRC_REST ? K : dopp(rule_dir<K>(kind)));
            if (kind == SRC_REG) {
                if constexpr (REG) p = (const st*)(ra.base(fi, dopp(K), c.z, dcz(K)) + ra.v[dcx(K) + 1][dcy(K) + 1]);
                else p = fi + pop_index(src, dopp(K));
            }
            if constexpr (GEN && dcx(K) < 0)
                if (kind == SRC_OWN) p = ocopy + (K / 3) * BS2;
            raw[K] = SP::ld_raw(p);
            inlet_links |= (kind == SRC_INLET ? 1u : 0u) << K;
            rest_links |= (kind == SRC_REST ? 1u : 0u) << K;
            if constexpr (FX) tfs_links |= (kind == SRC_TFS ? 1u : 0u) << K;
            val = T(0);   // set below
        }
        f[K] = val;
    });
    if constexpr (!(REG && !GEN)) {
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            constexpr T w = wgt<T>(K);
            T val = SP::from_raw(raw[K], w);
            if constexpr (GEN) {
                if ((rest_links >> K) & 1u) val = w;
                if constexpr (dcx(K) > 0)
                    if ((inlet_links >> K) & 1u)
                        val = boundary_value<K, T>(SRC_INLET, a, c, u_in, a.step(), [&](auto) { return T(0); });
                if constexpr (FX && (dcy(K) != 0 || dcz(K) != 0))
                    if ((tfs_links >> K) & 1u) val = tfs_value<K, T>(a, val, u_in);
            }
            f[K] = val;
        });
    }
    if constexpr (PAT == PAT_AB && GEN && BFL) {
        if ((c.flags & BF_SOLID) && solid_links)
            bfl_links<SP>(a, c, f, solid_links, reg_links, a.fpart != nullptr, F0, F1, F2);
    }

    T rho_o, ux_o, uy_o, uz_o;
    if (GEN && solid_own) {
        // reference solid cell: full-way reflection of the gathered populations (kernel_cumulant.jl:405-417)
        T g[27];
#pragma unroll
        for (int k = 0; k < 27; ++k) g[k] = f[26 - k];
#pragma unroll
        for (int k = 0; k < 27; ++k) f[k] = g[k];
        rho_o = T(1); ux_o = uy_o = uz_o = T(0);
    } else {
        CellRates<T> q;
        if constexpr (WALE) {
            const T nu_eddy = wale_of<T>(a, c);
            cell_rates<T>(nu_eddy, T(dm->tau), T(dm->amr_scale), T(a.cp.nu_sgs_bg), a.cp.adaptive_omega4,
                          T(a.cp.lambda), T(a.cp.omega_4), q.omega1, q.omega4_diag);
        } else if constexpr (sizeof(T) == 8) {
            q.omega1 = dm->omega1_d; q.omega4_diag = dm->omega4d_d;
        } else {
            q.omega1 = dm->omega1; q.omega4_diag = dm->omega4d;
        }
        q.forced = dm->forced != 0;
        q.Fx = T(dm->F[0]); q.Fy = T(dm->F[1]); q.Fz = T(dm->F[2]);
        if constexpr (SPONGE) {
            q.sponge = (c.flags & BF_SPONGE) ? T(a.sponge[(size_t)a.sslot[b] * BS3 + (c.own - b * BS3)]) : T(0);
            q.sponge_u = T(u_in);
            q.sponge_blend = a.sponge_blend;
        }
        collide_cell<T, false, SPONGE>(f, q, a.cp, rho_o, ux_o, uy_o, uz_o);
    }
    if (a.vel_out) {
        T* vo = (T*)a.vel_out;
        vo[c.own] = ux_o; vo[n + c.own] = uy_o; vo[2 * (size_t)n + c.own] = uz_o;
    }
    if (a.rho_out) ((T*)a.rho_out)[c.own] = rho_o;

    const bool want_force = a.fpart != nullptr;
    // (with a q-map the links' momentum exchange was taken at the gather)
    const bool link_force = GEN && !BFL && want_force && (c.flags & BF_SOLID) && !solid_own;
    const bool outlet_cell = GEN && PAT == PAT_AA_EVEN && (c.flags & BF_EDGE) && !c.per0 && c.gx == c.ext0 - 1;
    // rounding of the stored post-collision populations: population j of this cell takes the random
    // bits of (cell, j, step) whichever slot it goes to, so every pattern stores the same values
    const bool dith = SP::dithered && a.cp.dither;
    const uint32_t hc = SP::dithered ? round_cell(round_key(a.step()), c.own - dm->cell0) : 0u;
    auto rb = [&](int j) { return round_bits(hc, j); };
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int J = dopp(K);   // the source of direction J is x + c_K, where f*_K goes
        constexpr T wK = wgt<T>(K);
        if constexpr (REG && PAT == PAT_AB) {
            SP::sto_r((st*)(ra.base(fo, K, c.z, 0) + ra.v[1][1]), f[K], wK, rb(K), dith);
        } else if constexpr (REG && PAT == PAT_AA_EVEN) {
            SP::sto_r((st*)(ra.base(fo, J, c.z, 0) + ra.v[1][1]), f[K], wK, rb(K), dith);
            if constexpr (GEN && dcx(K) < 0) {
                if (outlet_cell)
                    SP::sto_r((st*)a.obuf + ((size_t)a.oslot[b] * 9 + K / 3) * BS2 + c.ly * BS + c.z, f[K], wK, rb(K), dith);
            }
        } else if constexpr (REG && !GEN) {
            SP::sto_r((st*)(ra.base(fo, K, c.z, -dcz(K)) + ra.v[1 - dcx(K)][1 - dcy(K)]), f[K], wK, rb(K), dith);
        } else if constexpr (REG) {
            uint32_t dst = 0;
            const int kind = classify<J, GEN, REG, FX>(a, c, dst);
            if (kind == SRC_REG) {
                SP::sto_r((st*)(ra.base(fo, K, c.z, -dcz(K)) + ra.v[1 - dcx(K)][1 - dcy(K)]), f[K], wK, rb(K), dith);
            } else if (kind != SRC_INLET) {
                const T v = boundary_value<J, T>(kind, a, c, u_in, a.step() + 1,
                                                 [&](auto jc) { return f[decltype(jc)::value]; });
                SP::sto_r(fo + pop_index(c.own, J), v, wgt<T>(J), rb(rule_dir<J>(kind)), dith);
            }
        } else if constexpr (PAT == PAT_AB) {
            SP::sto_r(fo + pop_index(c.own, K), f[K], wK, rb(K), dith);
        } else if constexpr (PAT == PAT_AA_EVEN) {
            SP::sto_r(fo + pop_index(c.own, J), f[K], wK, rb(K), dith);
            if constexpr (GEN && dcx(K) < 0) {
                if (outlet_cell)
                    SP::sto_r((st*)a.obuf + ((size_t)a.oslot[b] * 9 + K / 3) * BS2 + c.ly * BS + c.z, f[K], wK, rb(K), dith);
            }
        } else {
            uint32_t dst = 0;
            const int kind = classify<J, GEN, false, FX>(a, c, dst);
            if (kind == SRC_REG) {
                SP::sto_r(fo + pop_index(dst, K), f[K], wK, rb(K), dith);
            } else if (kind != SRC_INLET) {   // inlet links are evaluated by the even step itself
                const T v = boundary_value<J, T>(kind, a, c, u_in, a.step() + 1,
                                                 [&](auto jc) { return f[decltype(jc)::value]; });
                SP::sto_r(fo + pop_index(c.own, J), v, wgt<T>(J), rb(rule_dir<J>(kind)), dith);
            }
        }
        if constexpr (GEN) {
            if (link_force) {
                uint32_t dst = 0;
                const int kind = classify<J, GEN, REG, FX>(a, c, dst);
                double m = 0.0;
                if (kind == SRC_SOLID) {
                    m = 2.0 * ((double)SP::rt_r(f[K], wK, rb(K), dith) - (double)wK);
                } else if (PAT == PAT_AB && a.cp.bb_mode == 0 && kind == SRC_REG && ((c.lmask >> J) & 1u)) {
                    m = ((double)SP::rt_r(f[K], wK, rb(K), dith) - (double)wK) +
                        ((double)own_prev<SP, PAT, K>(a, c) - (double)wK);
                }
                F0 += m * dcx(K); F1 += m * dcy(K); F2 += m * dcz(K);
            }
        }
    });
}

// SEL: which blocks a launch covers.  SEL_ALL dispatches on the block flags inside one launch;
// SEL_INTERIOR / SEL_BOUNDARY run the interior and the boundary block lists as separate launches, so
// that the interior kernel is compiled for its own (smaller) register budget.
enum : int { SEL_ALL = 0, SEL_INTERIOR = 1, SEL_BOUNDARY = 2, SEL_REGULAR = 3, SEL_BOUNDARY_REGULAR = 4 };

#ifndef OAL_STEP_WAVES
#define OAL_STEP_WAVES 3
#endif
// register budget of the one-plane kernels: 3 waves per SIMD, 2 for the forms that do not fit 168 VGPRs
// without spilling (FP64 arithmetic, the AA pattern's two-sided slots, WALE's velocity gather), 1 for the
// FP64 AA forms
template <class SP, int PAT, bool WALE>
constexpr int step_waves() {
    return sizeof(typename SP::T) == 8 && PAT != PAT_AB ? 1 : sizeof(typename SP::T) == 8 || PAT != PAT_AB || WALE ? 2 : OAL_STEP_WAVES;
}

template <class SP, int PAT, bool WALE, int SEL = SEL_ALL, bool BFL = false, bool SPONGE = false, bool FX = true>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(step_waves<SP, PAT, WALE>()))) k_step(const StepArgs a) {
    const uint32_t litem = xcd_item(blockIdx.x, gridDim.x);
    const uint32_t lb = litem / PLANES;
    const uint32_t b = a.blist ? (uint32_t)a.blist[lb] : lb;
    const uint32_t item = b * PLANES + litem % PLANES;
    CellCtx c;
    c.b = b;
    c.z = (int)(item % PLANES);
    c.lx = threadIdx.x & (BS - 1);
    c.ly = threadIdx.x >> 4;
    c.flags = a.bflags[b];
    const DevDomain* dm = a.doms + a.bdom[b];
    c.ext0 = dm->ext[0]; c.ext1 = dm->ext[1]; c.ext2 = dm->ext[2];
    c.per0 = dm->per[0]; c.per1 = dm->per[1]; c.per2 = dm->per[2];
    c.zwall = dm->zwall; c.ymin_wall = dm->ymin_wall;
    c.gx = a.borg[3 * b] + c.lx;
    c.gy = a.borg[3 * b + 1] + c.ly;
    c.gz = a.borg[3 * b + 2] + c.z;
    c.own = b * BS3 + (uint32_t)c.z * BS2 + (uint32_t)c.ly * BS + (uint32_t)c.lx;
    double F0 = 0.0, F1 = 0.0, F2 = 0.0;
    c.lmask = 0;
    if constexpr (SEL == SEL_REGULAR) {
        bool live = true;
        if constexpr (PAT == PAT_AB) {
            if (a.regular_face) {
                const FaceBits fb = face_bits(a, b, dm);
                live = !(face_lane(fb, c.lx, c.ly) || face_plane(fb, c.z));
            }
        }
        step_cell<SP, PAT, WALE, false, true>(a, c, dm, live, false, F0, F1, F2);
    } else if constexpr (SEL == SEL_INTERIOR) {
        step_cell<SP, PAT, WALE, false>(a, c, dm, true, false, F0, F1, F2);
    } else {
        const bool inside = !(c.flags & BF_PARTIAL) || (c.gx < c.ext0 && c.gy < c.ext1 && c.gz < c.ext2);
        if (c.flags & BF_SOLID) c.lmask = a.lmask[(size_t)a.lslot[b] * BS3 + (c.own - b * BS3)];
        const bool solid_own = (c.lmask >> 27) & 1u;
        const bool active = inside && !(solid_own && a.cp.bb_mode == 1);
        if constexpr (SEL == SEL_BOUNDARY_REGULAR)
            step_cell<SP, PAT, WALE, true, true, BFL, SPONGE, FX>(a, c, dm, active, solid_own, F0, F1, F2);
        else
            step_cell<SP, PAT, WALE, true, false, BFL, SPONGE, FX>(a, c, dm, active, solid_own, F0, F1, F2);
    }
    if (a.fpart) wg_sum3(F0 * dm->force_scale, F1 * dm->force_scale, F2 * dm->force_scale, a.fpart, item);
}

// Regular interior blocks, pull/AB pattern, software-pipelined over z: a workgroup walks `a.pipe_planes`
// consecutive planes of one block and issues the loads of the next plane before the collision of the
// current one, so every wave keeps a plane of loads in flight while it computes.
//
// Vector-memory loads and stores share one completion counter on gfx950, and with both pending the
// compiler may only wait for all of them.  The schedule is therefore: wait for everything at the top of
// an iteration (the current plane's loads and the previous plane's stores, both issued one collision
// earlier), issue the next plane's loads, then collide and store the current plane from registers that
// are already complete.  Prefetched values stay in their storage form until the collision uses them.
template <class SP>
__device__ __forceinline__ void reg_load_raw(const StepArgs& a, const RegAddr<SP>& ra, int z, typename SP::raw (&r)[27]) {
    using st = typename SP::st;
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        r[K] = SP::ld_raw((const st*)(ra.base(a.fin, K, z, dcz(K)) + ra.v[dcx(K) + 1][dcy(K) + 1]));
    });
}

__device__ __forceinline__ void wait_all_vmem() {
    __builtin_amdgcn_s_waitcnt(0x0f70);   // vmcnt(0), expcnt and lgkmcnt left open
    __builtin_amdgcn_sched_barrier(0);
}

// inlet: the equilibrium at (1, u_inlet + noise) for the directions entering through x = 0
struct InletRule {
    bool on;          // this lane's cell lies on the inlet face
    float u_in;
    int gy, gz;       // the cell's global y and z (the noise position)
};

// WALE in the pipelined kernel (AB pattern, regular blocks: no solid cell in the block or its halo, the
// neighbour blocks at lexicographic offsets).  The previous step's velocity at the six face neighbours
// (x+, x-, y+, y-, z+, z-) of a lane's cell is loaded before the next plane's populations, so that the
// collision waits for these loads only; a neighbour across a non-periodic domain face is the cell
// itself (physics_utils.jl:1981-2035, 2217-2233, the rule of wale_of).
template <class T>
struct PipeWale {
    T u[6][3];
};
template <class T>
struct PipeWaleRates {
    T tau, amr_scale;           // of the domain, read before the first store
};
template <class SP, bool UNIT, bool WALE = false, bool SPONGE = false>
__device__ __forceinline__ void reg_collide_store(const StepArgs& a, const RegAddr<SP>& ra, const CellRates<typename SP::T>& q0,
                                                  int z, uint32_t cell, const typename SP::raw (&r)[27],
                                                  const InletRule* inlet = nullptr, bool store = true,
                                                  const PipeWale<typename SP::T>* wv = nullptr,
                                                  const PipeWaleRates<typename SP::T>* wr = nullptr, uint32_t vidx = 0,
                                                  typename SP::T sponge = typename SP::T(0)) {
    using T = typename SP::T;
    using st = typename SP::st;
    CellRates<T> q = q0;
    if constexpr (SPONGE) q.sponge = sponge;
    if constexpr (WALE) {
        const T h = T(0.5);
        const auto& u = wv->u;
        const T nu_eddy = wale_nu(h * (u[0][0] - u[1][0]), h * (u[2][0] - u[3][0]), h * (u[4][0] - u[5][0]),
                                  h * (u[0][1] - u[1][1]), h * (u[2][1] - u[3][1]), h * (u[4][1] - u[5][1]),
                                  h * (u[0][2] - u[1][2]), h * (u[2][2] - u[3][2]), h * (u[4][2] - u[5][2]), T(a.cp.c_wale));
        cell_rates<T>(nu_eddy, wr->tau, wr->amr_scale, T(a.cp.nu_sgs_bg), a.cp.adaptive_omega4, T(a.cp.lambda),
                      T(a.cp.omega_4), q.omega1, q.omega4_diag);
    }
    T f[27];
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        f[K] = SP::from_raw(r[K], wgt<T>(K));
    });
    if (inlet && inlet->on) {
        CellCtx c{};
        c.gy = inlet->gy;
        c.gz = inlet->gz;
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            if constexpr (dcx(K) > 0)
                f[K] = boundary_value<K, T>(SRC_INLET, a, c, inlet->u_in, a.step(), [&](auto) { return T(0); });
        });
    }
    T rho_o, ux_o, uy_o, uz_o;
    collide_cell<T, UNIT, SPONGE>(f, q, a.cp, rho_o, ux_o, uy_o, uz_o);
    const bool dith = SP::dithered && a.cp.dither;
    const uint32_t hc = SP::dithered ? round_cell(round_key(a.step()), cell) : 0u;
    if (store) {
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            SP::sto_r((st*)(ra.base(a.fout, K, z, 0) + ra.v[1][1]), f[K], wgt<T>(K), round_bits(hc, K), dith);
        });
        if constexpr (WALE) {
            T* vo = (T*)a.vel_out;
            vo[vidx] = ux_o; vo[a.ncell + vidx] = uy_o; vo[2 * (size_t)a.ncell + vidx] = uz_o;
        }
    }
}

// Domain-face rules inside the pipelined kernel (AB pattern).  A source outside the extent is replaced
// by a population of the cell's own state (kernel_cumulant.jl:233-345, the order of classify): the
// rules only change the address a lane reads, except at the inlet, whose equilibrium replaces the
// loaded value after conversion.
struct FaceRules {
    bool xl, xh, yl, yh;          // per lane: the cell lies on the x-min / x-max / y-min / y-max face
    bool zlo, zhi;                // the block touches the z-min / z-max face (uniform)
    bool zwall, ymin_wall;        // no-slip variants of the z faces and of the y-min face (uniform)
    // blocks that overhang the extent (FACE_PARTIAL): the planes of the z faces inside the block
    // (-1: none) and, per lane, a cell beyond the extent in x or y (its in-plane source offsets all
    // point at its own cell; never stored)
    int zlo_plane, zhi_plane;
    bool dead;
};

// Gather with the domain-face rules.  Every read is addressed by its byte offset from the start of
// the own block, formed in 32-bit arithmetic modulo 2^32 and read as a signed offset: the host uses
// this form only where a layer of blocks spans less than 2 GiB, so every offset read is exact.  The rules pick, per lane, one
// of a few candidate offsets that are all computed (uniform parts in scalar registers), so that the
// choice is a select rather than a branch.
template <class SP, bool PARTIAL = false>
__device__ __forceinline__ void reg_load_raw_face(const StepArgs& a, const RegAddr<SP>& ra, const FaceRules& fr, int z,
                                                  typename SP::raw (&r)[27]) {
    using st = typename SP::st;
    const char* f0 = (const char*)a.fin + ra.blk0 * (int64_t)sizeof(st);   // the own block (uniform)
    const uint32_t nbytes = BS3 * (uint32_t)sizeof(st);                    // one direction slot
    uint32_t zs[3];                                           // plane shift of sources with cz = -1, 0, 1
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int sz = z - (q - 1);
        const int dz = sz < 0 ? -1 : (sz >= BS ? 1 : 0);
        zs[q] = (uint32_t)((dz * ra.plane_elems + (int64_t)(sz - BS * dz - ra.zref) * BS2) * (int64_t)sizeof(st));
    }
    const bool zlo = PARTIAL ? z == fr.zlo_plane : (fr.zlo && z == 0);       // uniform
    const bool zhi = PARTIAL ? z == fr.zhi_plane : (fr.zhi && z == BS - 1);
    // uniform parts pass through readfirstlane, which the optimiser cannot see through: it must not
    // fold them into per-lane sums and hoist 27 of those out of the plane loop
    auto uni = [](uint32_t u) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)u); };
    const uint32_t v11 = (uint32_t)ra.v[1][1];
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int cx = dcx(K), cy = dcy(K), cz = dcz(K);
        uint32_t o = (uint32_t)ra.v[cx + 1][cy + 1] + uni((uint32_t)K * nbytes + zs[cz + 1]);
        auto own = [&](uint32_t slot) { return v11 + uni(slot * nbytes + zs[1]); };
        // lowest precedence first; each later rule overrides (kernel_cumulant.jl:233-345)
        if constexpr (cz != 0) {
            const bool zo = cz > 0 ? zlo : zhi;
            const uint32_t o_z = own(fr.zwall ? dopp(K) : dmir_z(K));
            o = zo ? o_z : o;                                               // z faces
            if constexpr (cy != 0) {
                const bool yo = cy > 0 ? fr.yl : fr.yh;
                const uint32_t o_y = own(cy > 0 && fr.ymin_wall ? dopp(K) : dmir_y(K));
                o = (yo && !(zo && fr.zwall)) ? o_y : o;                    // y faces, below no-slip z walls
            }
        } else if constexpr (cy != 0) {
            const bool yo = cy > 0 ? fr.yl : fr.yh;
            const uint32_t o_y = own(cy > 0 && fr.ymin_wall ? dopp(K) : dmir_y(K));
            o = yo ? o_y : o;
        }
        if constexpr (cx != 0) {
            const uint32_t o_x = own(K);
            o = (cx > 0 ? fr.xl : fr.xh) ? o_x : o;    // outlet: own population; inlet: any valid address
        }
        r[K] = SP::ld_raw((const st*)(f0 + (int32_t)o));   // within +-2 GiB of the own block
    });
}

// FACE: how blocks on a domain face are handled.  FACE_PARTIAL: blocks that overhang the extent, with
// the face rules at the extent's last cells and the cells beyond it neither read through nor stored.
// FACE_SOLID: blocks of the regular grid whose only flag is a solid cell in or next to them, away from
// the domain faces: halfway bounce-back (the source of a solid link is the cell's own slot opp K of the
// previous step) chosen per lane by address, the link masks loaded two planes ahead
enum : int { FACE_NONE = 0, FACE_RULES = 1, FACE_SKIP = 2, FACE_PARTIAL = 3, FACE_SOLID = 4 };

// Register budget: with one plane of populations in flight beside the collision of another the kernel
// fits the 168 VGPRs of 3 waves per SIMD; the in-kernel face rules add per-lane address selects and
// scalar offsets that only fit the 256 VGPRs of 2 waves.
#ifndef OAL_FACE_RULES_WAVES
#define OAL_FACE_RULES_WAVES 2
#endif
// (the 16-bit forms' conversions and rounding need more registers: 2 waves for every form of theirs)
// (WALE's velocity gather as well)
template <class SP, int FACE, bool WALE = false>
constexpr int pipe_waves() {
    return FACE == FACE_RULES || FACE == FACE_PARTIAL || FACE == FACE_SOLID || SP::dithered || WALE ? OAL_FACE_RULES_WAVES : 3;
}

// UNIT: the collision form for the unit rates omega_bulk = omega_3 = omega_4 = 1 (a separate kernel, so
// that each form keeps its own register allocation).  WALE: the eddy viscosity from the previous step's
// velocity, and this step's velocity written.
// SPONGE: blocks with sponge cells (kernel_cumulant.jl:428-446), each cell's strength loaded with its
// populations
template <class SP, int NP, int FACE, bool UNIT, bool WALE = false, bool SPONGE = false>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(pipe_waves<SP, FACE, WALE>(), pipe_waves<SP, FACE, WALE>())))
k_step_pipe(const StepArgs a) {
    using T = typename SP::T;
    using st = typename SP::st;
    // the loop bound stays a run-time value (equal to NP): with a constant bound the prefetch
    // conditions fold away, the scheduler mixes the next plane's loads into the collision and the
    // kernel spills
    const int np = a.pipe_planes;
    constexpr uint32_t groups = (uint32_t)(PLANES / NP);
    const PipeItem it = pipe_item<false>(groups, np);
    const uint32_t b = (uint32_t)a.blist[it.pos];
    const int z0 = it.z0;
    // FACE_PARTIAL: planes from zmax on lie beyond the extent and are neither read nor stepped (the
    // loop bound itself stays z0 + np: a bound the compiler can narrow lets it reorder the prefetch)
    int zmax = BS;
    if constexpr (FACE == FACE_PARTIAL || FACE == FACE_RULES) {
        // (FACE_RULES: whole blocks, zmax >= BS; the guard it adds to every plane keeps the compiler from
        // hoisting the next plane's address arithmetic, which spills at 2 to 8 planes)
        const DevDomain* d0 = a.doms + a.bdom[b];
        zmax = d0->ext[2] - a.borg[3 * b + 2];
        if (FACE == FACE_PARTIAL && z0 >= zmax) return;
    }
    CellCtx c;
    c.b = b;
    c.lx = threadIdx.x & (BS - 1);
    c.ly = threadIdx.x >> 4;
    c.z = z0;
    c.own = b * BS3 + (uint32_t)z0 * BS2 + (uint32_t)c.ly * BS + (uint32_t)c.lx;
    const DevDomain* dm = a.doms + a.bdom[b];
    RegAddr<SP> ra;
    ra.init(c, dm->nbx, dm->nby);   // offsets of plane z0
    // every domain parameter is read before the first store: reads issued after stores to the
    // population arrays could not be proven independent of them
    CellRates<T> q;
    if constexpr (sizeof(T) == 8) { q.omega1 = dm->omega1_d; q.omega4_diag = dm->omega4d_d; }
    else { q.omega1 = dm->omega1; q.omega4_diag = dm->omega4d; }
    q.forced = dm->forced != 0;
    q.Fx = T(dm->F[0]); q.Fy = T(dm->F[1]); q.Fz = T(dm->F[2]);
    const float* sp_cells = nullptr;
    if constexpr (SPONGE) {
        q.sponge = T(0);
        q.sponge_u = T(dm->u_in);
        q.sponge_blend = a.sponge_blend;
        if (a.bflags[b] & BF_SPONGE) sp_cells = a.sponge + (size_t)a.sslot[b] * BS3 + (c.own - b * BS3 - (uint32_t)z0 * BS2);
    }
    // the strength of the lane's cell in plane z (0 outside the sponge)
    auto sp_load = [&](int z) -> T {
        if constexpr (SPONGE) return sp_cells ? T(sp_cells[(uint32_t)z * BS2]) : T(0);
        else return T(0);
    };
    // blocks on a domain face: FACE_RULES applies the face rules here, FACE_SKIP leaves the face cells
    // to the face-cell kernel
    FaceRules fr{};
    InletRule in{};
    bool lane_face = false;
    FaceBits fskip{};
    if constexpr (FACE == FACE_SKIP) {
        fskip = face_bits(a, b, dm);
        lane_face = face_lane(fskip, c.lx, c.ly);
    }
    auto live = [&](int z) { return FACE != FACE_SKIP || !(lane_face || face_plane(fskip, z)); };
    if constexpr (FACE == FACE_PARTIAL) {
        const int gx = a.borg[3 * b] + c.lx, gy = a.borg[3 * b + 1] + c.ly, oz = a.borg[3 * b + 2];
        fr.xl = !dm->per[0] && gx == 0;
        fr.xh = !dm->per[0] && gx == dm->ext[0] - 1;
        fr.yl = !dm->per[1] && gy == 0;
        fr.yh = !dm->per[1] && gy == dm->ext[1] - 1;
        fr.zlo_plane = (!dm->per[2] && oz == 0) ? 0 : -1;
        fr.zhi_plane = (!dm->per[2] && dm->ext[2] - 1 - oz < BS) ? dm->ext[2] - 1 - oz : -1;
        fr.dead = gx >= dm->ext[0] || gy >= dm->ext[1];
        if (fr.dead) {
            // a cell beyond the extent reads only its own column (its x / y sources may lie in blocks
            // that do not exist); the z rules keep its reads inside the level as for every lane
            const int32_t v11 = ra.v[1][1];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) ra.v[i][j] = v11;
        }
        fr.zwall = dm->zwall != 0;
        fr.ymin_wall = dm->ymin_wall == 1;
        in.on = fr.xl;
        in.u_in = dm->u_in;
        in.gy = gy;
        in.gz = oz;      // + z per plane
    }
    if constexpr (FACE == FACE_RULES) {
        const FaceBits fb = face_bits(a, b, dm);
        fr.xl = fb.xlo && c.lx == 0;
        fr.xh = fb.xhi && c.lx == BS - 1;
        fr.yl = fb.ylo && c.ly == 0;
        fr.yh = fb.yhi && c.ly == BS - 1;
        fr.zlo = fb.zlo;
        fr.zhi = fb.zhi;
        fr.zwall = dm->zwall != 0;
        fr.ymin_wall = dm->ymin_wall == 1;
        in.on = fr.xl;
        in.u_in = dm->u_in;
        in.gy = a.borg[3 * b + 1] + c.ly;
        in.gz = a.borg[3 * b + 2];      // + z per plane
    }
    // WALE: the x / y neighbours' cell offsets of the lane (0: across a domain face, the cell itself),
    // the block layer, and the domain's rates
    int32_t wxp = 0, wxm = 0, wyp = 0, wym = 0, wlayer = 0;
    PipeWaleRates<T> wr{};
    if constexpr (WALE) {
        const int32_t row = dm->nbx * BS3 - (BS - 1) * BS;
        wxp = fr.xh ? 0 : (c.lx < BS - 1 ? 1 : BS3 - (BS - 1));
        wxm = fr.xl ? 0 : (c.lx > 0 ? -1 : -(BS3 - (BS - 1)));
        wyp = fr.yh ? 0 : (c.ly < BS - 1 ? BS : row);
        wym = fr.yl ? 0 : (c.ly > 0 ? -BS : -row);
        if constexpr (FACE == FACE_PARTIAL) {
            if (fr.dead) { wxp = 0; wxm = 0; wyp = 0; wym = 0; }
        }
        wlayer = dm->nbx * dm->nby * BS3;
        wr.tau = T(dm->tau);
        wr.amr_scale = T(dm->amr_scale);
    }
    // the six neighbours' velocities of the lane's cell in plane z
    auto wale_load = [&](int z, PipeWale<T>& w) {
        const T* v = (const T*)a.vel_in;
        const uint32_t own = c.own + (uint32_t)(z - z0) * BS2;
        bool zl = false, zh = false;
        if constexpr (FACE == FACE_RULES) { zl = fr.zlo && z == 0; zh = fr.zhi && z == BS - 1; }
        if constexpr (FACE == FACE_PARTIAL) { zl = z == fr.zlo_plane; zh = z == fr.zhi_plane; }
        const int32_t zp = zh ? 0 : (z < BS - 1 ? BS2 : wlayer - (BS - 1) * BS2);
        const int32_t zm = zl ? 0 : (z > 0 ? -BS2 : -(wlayer - (BS - 1) * BS2));
        const int32_t off[6] = {wxp, wxm, wyp, wym, zp, zm};
#pragma unroll
        for (int q6 = 0; q6 < 6; ++q6) {
            const uint32_t i = (uint32_t)((int32_t)own + off[q6]);
            w.u[q6][0] = v[i]; w.u[q6][1] = v[a.ncell + i]; w.u[q6][2] = v[2 * (size_t)a.ncell + i];
        }
    };
    // FACE_SOLID: the solid-link masks of the lane's cell (bit 27: the cell itself is solid)
    const uint32_t* lm_cells = nullptr;
    if constexpr (FACE == FACE_SOLID) lm_cells = a.lmask + (size_t)a.lslot[b] * BS3 + (c.own - b * BS3 - (uint32_t)z0 * BS2);
    auto lm_load = [&](int z) -> uint32_t {
        if constexpr (FACE == FACE_SOLID) return z < z0 + np ? lm_cells[(uint32_t)z * BS2] : 0u;
        else return 0u;
    };
    uint32_t lm_cur = 0;   // FACE_SOLID: the mask of the plane the next load() reads
    auto load = [&](const RegAddr<SP>& r, int z, typename SP::raw (&dst)[27]) {
        if constexpr (FACE == FACE_RULES) { if (z < zmax) reg_load_raw_face<SP>(a, r, fr, z, dst); }
        else if constexpr (FACE == FACE_PARTIAL) { if (z < zmax) reg_load_raw_face<SP, true>(a, r, fr, z, dst); }
        else if constexpr (FACE == FACE_SOLID) {
            using st = typename SP::st;
            const uint32_t m = lm_cur;
            for_dirs([&](auto kc) {
                constexpr int K = decltype(kc)::value;
                const char* pr = r.base(a.fin, K, z, dcz(K)) + r.v[dcx(K) + 1][dcy(K) + 1];
                const char* ps = r.base(a.fin, dopp(K), z, 0) + r.v[1][1];
                dst[K] = SP::ld_raw((const st*)(((m >> K) & 1u) ? ps : pr));
            });
        }
        else if (live(z)) reg_load_raw<SP>(a, r, z, dst);
    };
    const uint32_t cell0 = dm->cell0;   // read before the first store, as every domain parameter
    auto collide = [&](const RegAddr<SP>& r, int z, const typename SP::raw (&src)[27], const PipeWale<T>& w, T sp,
                       uint32_t lm) {
        const uint32_t vidx = c.own + (uint32_t)(z - z0) * BS2;
        const uint32_t cell = vidx - cell0;
        if constexpr (FACE == FACE_SOLID) {
            // a solid cell is not stepped (halfway bounce-back)
            reg_collide_store<SP, UNIT, WALE, SPONGE>(a, r, q, z, cell, src, nullptr, !((lm >> 27) & 1u), &w, &wr, vidx, sp);
            return;
        }
        if constexpr (FACE == FACE_RULES) {
            if (z >= zmax) return;
            InletRule iz = in;
            iz.gz += z;
            reg_collide_store<SP, UNIT, WALE, SPONGE>(a, r, q, z, cell, src, &iz, true, &w, &wr, vidx, sp);
        } else if constexpr (FACE == FACE_PARTIAL) {
            if (z >= zmax) return;
            InletRule iz = in;
            iz.gz += z;
            reg_collide_store<SP, UNIT, WALE, SPONGE>(a, r, q, z, cell, src, &iz, !fr.dead, &w, &wr, vidx, sp);
        } else if (live(z)) {
            reg_collide_store<SP, UNIT, WALE, SPONGE>(a, r, q, z, cell, src, nullptr, true, &w, &wr, vidx, sp);
        }
    };
    // WALE: the plane's neighbour velocities, issued before the next plane's populations
    auto wale_pre = [&](int z, PipeWale<T>& w) {
        if constexpr (WALE) {
            if (z < zmax && live(z)) wale_load(z, w);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // the per-lane offsets stay those of plane z0; later planes shift the uniform base
    typename SP::raw fa[27], fb[27];
    T spa = T(0), spb = T(0);
    uint32_t lma = 0, lmb = 0, lmn = 0;   // FACE_SOLID: the masks of the planes in fa / fb, and the next one's
    PipeWale<T> wv;
    if constexpr (FACE == FACE_SOLID) {
        lma = lm_load(z0);
        wait_all_vmem();
        lm_cur = lma;
    }
    load(ra, z0, fa);
    if constexpr (SPONGE) spa = sp_load(z0);
    if constexpr (FACE == FACE_SOLID) lmn = lm_load(z0 + 1);
    for (int z = z0; z < z0 + np; z += 2) {
        wait_all_vmem();
        wale_pre(z, wv);
        if (z + 1 < z0 + np) {
            if constexpr (FACE == FACE_SOLID) { lmb = lmn; lm_cur = lmb; }
            load(ra, z + 1, fb);
            if constexpr (SPONGE) spb = sp_load(z + 1);
            if constexpr (FACE == FACE_SOLID) lmn = lm_load(z + 2);
        }
        collide(ra, z, fa, wv, spa, lma);
        if (z + 1 >= z0 + np) break;
        wait_all_vmem();
        wale_pre(z + 1, wv);
        if (z + 2 < z0 + np) {
            if constexpr (FACE == FACE_SOLID) { lma = lmn; lm_cur = lma; }
            load(ra, z + 2, fa);
            if constexpr (SPONGE) spa = sp_load(z + 2);
            if constexpr (FACE == FACE_SOLID) lmn = lm_load(z + 3);
        }
        collide(ra, z + 1, fb, wv, spb, lmb);
    }
}

// The in-place AA pattern on the regular blocks (no domain face, no solid near; neighbour blocks at
// lexicographic offsets), software-pipelined over z as k_step_pipe.  The even half-step reads each
// population from the cell's own slot k and writes f*_k to its own slot opp k; the odd half-step reads
// slot opp k of x - c_k and writes f*_k to slot k of x + c_k.  Either way a lane reads and writes the
// same 27 slots and no other lane touches them, so the next plane's prefetch never meets a store of the
// current plane.  The storage rounding of f*_k takes the bits of (cell, k, step) as in every pattern.
// (the odd half-step's two-sided addressing needs the 2-wave register budget)
template <class SP, int PAT>
constexpr int pipe_aa_waves() { return PAT == PAT_AA_ODD ? OAL_FACE_RULES_WAVES : pipe_waves<SP, FACE_NONE>(); }

template <class SP, int NP, int PAT, bool UNIT>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(pipe_aa_waves<SP, PAT>(), pipe_aa_waves<SP, PAT>())))
k_step_pipe_aa(const StepArgs a) {
    using T = typename SP::T;
    using st = typename SP::st;
    const int np = a.pipe_planes;   // (a run-time bound, as in k_step_pipe)
    constexpr uint32_t groups = (uint32_t)(PLANES / NP);
    const PipeItem it = pipe_item<PAT == PAT_AA_ODD>(groups, np);
    const uint32_t b = (uint32_t)a.blist[it.pos];
    const int z0 = it.z0;
    CellCtx c;
    c.b = b;
    c.lx = threadIdx.x & (BS - 1);
    c.ly = threadIdx.x >> 4;
    c.z = z0;
    c.own = b * BS3 + (uint32_t)z0 * BS2 + (uint32_t)c.ly * BS + (uint32_t)c.lx;
    const DevDomain* dm = a.doms + a.bdom[b];
    RegAddr<SP> ra;
    ra.init(c, dm->nbx, dm->nby);
    CellRates<T> q;
    if constexpr (sizeof(T) == 8) { q.omega1 = dm->omega1_d; q.omega4_diag = dm->omega4d_d; }
    else { q.omega1 = dm->omega1; q.omega4_diag = dm->omega4d; }
    q.forced = dm->forced != 0;
    q.Fx = T(dm->F[0]); q.Fy = T(dm->F[1]); q.Fz = T(dm->F[2]);
    const uint32_t cell0 = dm->cell0;
    const uint32_t key = SP::dithered ? round_key(a.step()) : 0u;
    auto load = [&](int z, typename SP::raw (&r)[27]) {
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            if constexpr (PAT == PAT_AA_EVEN)
                r[K] = SP::ld_raw((const st*)(ra.base(a.fin, K, z, 0) + ra.v[1][1]));
            else
                r[K] = SP::ld_raw((const st*)(ra.base(a.fin, dopp(K), z, dcz(K)) + ra.v[dcx(K) + 1][dcy(K) + 1]));
        });
    };
    auto collide = [&](int z, const typename SP::raw (&r)[27]) {
        T f[27];
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            f[K] = SP::from_raw(r[K], wgt<T>(K));
        });
        T rho_o, ux_o, uy_o, uz_o;
        collide_cell<T, UNIT>(f, q, a.cp, rho_o, ux_o, uy_o, uz_o);
        const bool dith = SP::dithered && a.cp.dither;
        const uint32_t hc = SP::dithered ? round_cell(key, c.own + (uint32_t)(z - z0) * BS2 - cell0) : 0u;
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            st* p;
            if constexpr (PAT == PAT_AA_EVEN) p = (st*)(ra.base(a.fout, dopp(K), z, 0) + ra.v[1][1]);
            else p = (st*)(ra.base(a.fout, K, z, -dcz(K)) + ra.v[1 - dcx(K)][1 - dcy(K)]);
            SP::sto_r(p, f[K], wgt<T>(K), round_bits(hc, K), dith);
        });
    };
    typename SP::raw fa[27], fb[27];
    load(z0, fa);
    for (int z = z0; z < z0 + np; z += 2) {
        wait_all_vmem();
        if (z + 1 < z0 + np) load(z + 1, fb);
        collide(z, fa);
        if (z + 1 >= z0 + np) break;
        wait_all_vmem();
        if (z + 2 < z0 + np) load(z + 2, fa);
        collide(z + 1, fb);
    }
}

template <class SP, int NP, int PAT>
static void launch_pipe_aa(dim3 grid, hipStream_t s, const StepArgs& r) {
    if (r.cp.unit) hipLaunchKernelGGL((k_step_pipe_aa<SP, NP, PAT, true>), grid, dim3(WG), 0, s, r);
    else hipLaunchKernelGGL((k_step_pipe_aa<SP, NP, PAT, false>), grid, dim3(WG), 0, s, r);
}

// Face cells of the regular-list blocks (AB pattern): one lane per listed cell, the general path with
// the domain-face rules.  These blocks hold no solid cell and no solid link, so no force is summed.
template <class SP, bool WALE>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(step_waves<SP, PAT_AB, WALE>()))) k_face(const StepArgs a) {
    const uint32_t i = blockIdx.x * WG + threadIdx.x;
    if (i >= a.n_face) return;
    const uint32_t own = a.face_cells[i];
    const uint32_t b = own / BS3, loc = own % BS3;
    CellCtx c;
    c.b = b;
    c.own = own;
    c.lx = (int)(loc & (BS - 1));
    c.ly = (int)((loc >> 4) & (BS - 1));
    c.z = (int)(loc >> 8);
    c.flags = a.bflags[b];
    const DevDomain* dm = a.doms + a.bdom[b];
    c.ext0 = dm->ext[0]; c.ext1 = dm->ext[1]; c.ext2 = dm->ext[2];
    c.per0 = dm->per[0]; c.per1 = dm->per[1]; c.per2 = dm->per[2];
    c.zwall = dm->zwall; c.ymin_wall = dm->ymin_wall;
    c.gx = a.borg[3 * b] + c.lx;
    c.gy = a.borg[3 * b + 1] + c.ly;
    c.gz = a.borg[3 * b + 2] + c.z;
    c.lmask = 0;
    double F0 = 0.0, F1 = 0.0, F2 = 0.0;
    step_cell<SP, PAT_AB, WALE, true>(a, c, dm, true, false, F0, F1, F2);
}

}  // namespace oal

#include "oal_step_pair.h"
#include "oal_step_pair32.h"

namespace oal {

// (WALE: the face-rule form for every whole block; the forms without rules spill with WALE's gather)
template <class SP, int NP, bool UNIT, bool WALE>
static void launch_pipe_u(int face, dim3 grid, hipStream_t s, const StepArgs& r) {
    if constexpr (WALE) {
        if (face == FACE_PARTIAL) {
            if constexpr (NP <= 8) hipLaunchKernelGGL((k_step_pipe<SP, NP, FACE_PARTIAL, UNIT, true>), grid, dim3(WG), 0, s, r);
        } else {
            hipLaunchKernelGGL((k_step_pipe<SP, NP, FACE_RULES, UNIT, true>), grid, dim3(WG), 0, s, r);
        }
    } else {
        if (face == FACE_RULES) hipLaunchKernelGGL((k_step_pipe<SP, NP, FACE_RULES, UNIT, false>), grid, dim3(WG), 0, s, r);
        else if (face == FACE_SKIP) hipLaunchKernelGGL((k_step_pipe<SP, NP, FACE_SKIP, UNIT, false>), grid, dim3(WG), 0, s, r);
        else if (face == FACE_PARTIAL) {
            if constexpr (NP <= 8) hipLaunchKernelGGL((k_step_pipe<SP, NP, FACE_PARTIAL, UNIT, false>), grid, dim3(WG), 0, s, r);
        }
        else if (face == FACE_SOLID) {
            if constexpr (NP <= 8) hipLaunchKernelGGL((k_step_pipe<SP, NP, FACE_SOLID, UNIT, false>), grid, dim3(WG), 0, s, r);
        }
        else hipLaunchKernelGGL((k_step_pipe<SP, NP, FACE_NONE, UNIT, false>), grid, dim3(WG), 0, s, r);
    }
}
// (the WALE forms come in 2, 4 and 8 planes; the launch caps the planes at 8)
template <class SP, int NP>
static void launch_pipe_sponge(dim3 grid, hipStream_t s, const StepArgs& r) {
    if (r.cp.unit) hipLaunchKernelGGL((k_step_pipe<SP, NP, FACE_PARTIAL, true, false, true>), grid, dim3(WG), 0, s, r);
    else hipLaunchKernelGGL((k_step_pipe<SP, NP, FACE_PARTIAL, false, false, true>), grid, dim3(WG), 0, s, r);
}

template <class SP, int NP, bool WALE = false>
static void launch_pipe(int face, dim3 grid, hipStream_t s, const StepArgs& r) {
    if constexpr (!WALE || NP <= 8) {
        if (r.cp.unit) launch_pipe_u<SP, NP, true, WALE>(face, grid, s, r);
        else launch_pipe_u<SP, NP, false, WALE>(face, grid, s, r);
    }
}

// One step of a level: launches over disjoint block lists (regular, other interior, regular boundary,
// boundary), each compiled for its own register budget, plus, in the AB pattern, the face cells that
// the regular kernels skip.  The boundary and face launches go to `side` (when given) so that they
// overlap the bandwidth-bound bulk launch; the caller joins the two streams.
template <class SP, int PAT, bool WALE>
static hipError_t launch_pat(const StepArgs& a, hipStream_t stream, hipStream_t side, hipEvent_t bulk0, hipEvent_t bulk1) {
    StepArgs r = a;
    hipStream_t bs = side ? side : stream;
    // (a q-map exists only in the AB pattern)
    const bool bfl = PAT == PAT_AB && a.qmap != nullptr;
    // regular-grid blocks whose only flag is a solid cell near (at the head of the list): the pipelined
    // kernel's solid-link form where it applies (AB pattern, FP32 arithmetic, halfway bounce-back, no
    // q-map, WALE or force), the rest on the one-plane kernel
    uint32_t br_first = 0;
    if constexpr (PAT == PAT_AB && !WALE && sizeof(typename SP::T) == 4) {
        if (a.n_boundary_regular_solid && !bfl && !a.fpart && a.cp.bb_mode == 1 && a.pipe_planes > 1) {
            StepArgs rs = r;
            rs.blist = a.blist_boundary_regular;
            rs.pipe_planes = a.pipe_planes > 8 ? 8 : a.pipe_planes;
            const dim3 g(a.n_boundary_regular_solid * (PLANES / rs.pipe_planes));
            switch (rs.pipe_planes) {
                case 2: launch_pipe<SP, 2>(FACE_SOLID, g, bs, rs); break;
                case 4: launch_pipe<SP, 4>(FACE_SOLID, g, bs, rs); break;
                default: launch_pipe<SP, 8>(FACE_SOLID, g, bs, rs); break;
            }
            br_first = a.n_boundary_regular_solid;
        }
    }
    if (a.n_boundary_regular > br_first) {
        r.blist = a.blist_boundary_regular + br_first;
        const dim3 g((a.n_boundary_regular - br_first) * PLANES);
        if constexpr (PAT == PAT_AB) {
            if (bfl && a.face_x) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR, true>), g, dim3(WG), 0, bs, r);
            else if (bfl) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR, true, false, false>), g, dim3(WG), 0, bs, r);
            else if (a.sponge && a.face_x) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR, false, true>), g, dim3(WG), 0, bs, r);
            else if (a.sponge) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR, false, true, false>), g, dim3(WG), 0, bs, r);
            else if (a.face_x) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR>), g, dim3(WG), 0, bs, r);
            else hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR, false, false, false>), g, dim3(WG), 0, bs, r);
        } else {
            if (a.face_x) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR>), g, dim3(WG), 0, bs, r);
            else hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR, false, false, false>), g, dim3(WG), 0, bs, r);
        }
    }
    if (a.n_boundary) {
        r.blist = a.blist_boundary;
        const dim3 g(a.n_boundary * PLANES);
        if constexpr (PAT == PAT_AB) {
            if (bfl && a.face_x) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY, true>), g, dim3(WG), 0, bs, r);
            else if (bfl) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY, true, false, false>), g, dim3(WG), 0, bs, r);
            else if (a.sponge && a.face_x) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY, false, true>), g, dim3(WG), 0, bs, r);
            else if (a.sponge) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY, false, true, false>), g, dim3(WG), 0, bs, r);
            else if (a.face_x) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY>), g, dim3(WG), 0, bs, r);
            else hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY, false, false, false>), g, dim3(WG), 0, bs, r);
        } else {
            if (a.face_x) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY>), g, dim3(WG), 0, bs, r);
            else hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY, false, false, false>), g, dim3(WG), 0, bs, r);
        }
    }
    // face cells: the pipelined kernel applies the face rules itself, the one-plane regular kernel
    // leaves them to the face-cell kernel
    constexpr bool pipe_capable = PAT == PAT_AB && sizeof(typename SP::T) == 4;
    if constexpr (PAT == PAT_AB) {
        // (the pair kernel applies the rules too; with WALE the pipelined kernel always does)
        const bool piped = pipe_capable && a.pipe_planes > 1;
        if (a.n_face && (!piped || (a.face_kernel && !WALE)))
            hipLaunchKernelGGL((k_face<SP, WALE>), dim3((a.n_face + WG - 1) / WG), dim3(WG), 0, bs, a);
    }
    if (bulk0) {
        const hipError_t e = hipEventRecord(bulk0, stream);
        if (e != hipSuccess) return e;
    }
    // regular blocks (no force on them): one launch in lexicographic order, so that neighbouring blocks,
    // which share halo rows, stay close in the sweep; the 16-bit storage forms step them two cells per
    // lane (the face rules applied in the kernel)
    auto launch_regular = [&](const int32_t* list, uint32_t n, bool face, bool pair) {
        r.blist = list;
        r.regular_face = face ? 1 : 0;
        if constexpr (pipe_capable) {
            // (WALE: at most 8 planes per workgroup)
            r.pipe_planes = WALE && a.pipe_planes > 8 ? 8 : a.pipe_planes;
            const dim3 g(n * (PLANES / (r.pipe_planes > 1 ? r.pipe_planes : 1)));
            if constexpr (SP::dithered && !WALE) {
                if (pair) {
                    launch_pair<SP>(a.pipe_planes, face, g, stream, r);
                    return;
                }
            } else if constexpr (SP::code == StF32::code && !WALE) {
                if (pair) {
                    launch_pair32<SP>(a.pipe_planes, face, g, stream, r);
                    return;
                }
            }
            const int fc = !face ? FACE_NONE : (a.face_kernel && !WALE ? FACE_SKIP : FACE_RULES);
            switch (r.pipe_planes) {
                case 2: launch_pipe<SP, 2, WALE>(fc, g, stream, r); break;
                case 4: launch_pipe<SP, 4, WALE>(fc, g, stream, r); break;
                case 8: launch_pipe<SP, 8, WALE>(fc, g, stream, r); break;
                case 16: launch_pipe<SP, 16, WALE>(fc, g, stream, r); break;
                default: hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_REGULAR>), g, dim3(WG), 0, stream, r); break;
            }
        } else if constexpr (PAT != PAT_AB && !WALE && sizeof(typename SP::T) == 4) {
            // the AA pattern's regular blocks (none on a domain face): pipelined over z, 2 to 8 planes
            r.pipe_planes = a.pipe_planes > 8 ? 8 : a.pipe_planes;
            const dim3 g(n * (PLANES / (r.pipe_planes > 1 ? r.pipe_planes : 1)));
            switch (r.pipe_planes) {
                case 2: launch_pipe_aa<SP, 2, PAT>(g, stream, r); break;
                case 4: launch_pipe_aa<SP, 4, PAT>(g, stream, r); break;
                case 8: launch_pipe_aa<SP, 8, PAT>(g, stream, r); break;
                default: hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_REGULAR>), g, dim3(WG), 0, stream, r); break;
            }
        } else {
            hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_REGULAR>), dim3(n * PLANES), dim3(WG), 0, stream, r);
        }
    };
    const bool pair = pipe_capable && (SP::dithered || SP::code == StF32::code) && !WALE && a.pair_kernel &&
                      a.pipe_planes > 1 && !(a.regular_face && a.face_kernel);
    if (pair && a.pair_kernel == 2) {
        // the blocks on a domain face and the others apart (the latter with the smaller register budget)
        if (a.n_regular_face) launch_regular(a.blist_regular_face, a.n_regular_face, true, true);
        if (a.n_regular_plain) launch_regular(a.blist_regular_plain, a.n_regular_plain, false, true);
    } else if (a.n_regular) {
        launch_regular(a.blist_regular, a.n_regular, a.regular_face != 0, pair);
    }
    if (a.n_partial) {
        // regular blocks that overhang the extent (listed only where the pipelined kernel applies); at
        // most 8 planes per workgroup (the 16-plane form of this kernel spills); on the side stream
        // with the boundary launches, overlapping the bulk launch
        r.blist = a.blist_partial;
        if constexpr (pipe_capable) {
            StepArgs rp = r;
            rp.pipe_planes = a.pipe_planes > 8 ? 8 : a.pipe_planes;
            const dim3 g(a.n_partial * (PLANES / (rp.pipe_planes > 1 ? rp.pipe_planes : 1)));
            switch (rp.pipe_planes) {
                case 2: launch_pipe<SP, 2, WALE>(FACE_PARTIAL, g, bs, rp); break;
                case 4: launch_pipe<SP, 4, WALE>(FACE_PARTIAL, g, bs, rp); break;
                case 8: launch_pipe<SP, 8, WALE>(FACE_PARTIAL, g, bs, rp); break;
                default: hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR>), g, dim3(WG), 0, bs, rp); break;
            }
        } else {
            hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR>), dim3(a.n_partial * PLANES), dim3(WG), 0,
                               bs, r);
        }
    }
    if (a.n_sponge) {
        // regular blocks with sponge cells and no solid near (AB pattern): the pipelined kernel's sponge
        // form (without WALE; at most 8 planes per workgroup), else the one-plane kernel's; on the side
        // stream with the boundary launches
        StepArgs rs = r;
        rs.blist = a.blist_sponge;
        bool piped = false;
        if constexpr (pipe_capable && !WALE) {
            if (a.pipe_planes > 1) {
                rs.pipe_planes = a.pipe_planes > 8 ? 8 : a.pipe_planes;
                const dim3 g(a.n_sponge * (PLANES / rs.pipe_planes));
                switch (rs.pipe_planes) {
                    case 2: launch_pipe_sponge<SP, 2>(g, bs, rs); break;
                    case 4: launch_pipe_sponge<SP, 4>(g, bs, rs); break;
                    default: launch_pipe_sponge<SP, 8>(g, bs, rs); break;
                }
                piped = true;
            }
        }
        if (!piped) {
            const dim3 g(a.n_sponge * PLANES);
            if (a.face_x) hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR, false, true>), g, dim3(WG), 0, bs, rs);
            else hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR, false, true, false>), g, dim3(WG), 0, bs, rs);
        }
    }
    if (a.n_interior) {
        r.blist = a.blist_interior;
        hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_INTERIOR>), dim3(a.n_interior * PLANES), dim3(WG), 0, stream, r);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess || !bulk1) return e;
    return hipEventRecord(bulk1, stream);
}

template <class SP>
static hipError_t launch_sp(int stream_mode, bool wale, const StepArgs& a, hipStream_t s, hipStream_t side, hipEvent_t e0,
                            hipEvent_t e1) {
    if (stream_mode == 0)
        return wale ? launch_pat<SP, PAT_AB, true>(a, s, side, e0, e1) : launch_pat<SP, PAT_AB, false>(a, s, side, e0, e1);
    if (a.parity == 0)
        return wale ? launch_pat<SP, PAT_AA_EVEN, true>(a, s, side, e0, e1)
                    : launch_pat<SP, PAT_AA_EVEN, false>(a, s, side, e0, e1);
    return wale ? launch_pat<SP, PAT_AA_ODD, true>(a, s, side, e0, e1) : launch_pat<SP, PAT_AA_ODD, false>(a, s, side, e0, e1);
}

}  // namespace oal
