// Kernels of liboal for gfx950: the fused D3Q27 cumulant stream-collide step (pull/AB and in-place
// AA pattern), initialisation, layout conversion, macroscopic output, FP64 reductions and the
// stability guard.
//
// Work decomposition of the step kernels: one work-item per cell, one workgroup per 16 x 16 plane of
// a 16^3 block (256 lanes = 4 wavefronts of 16 x 4 cells).  Populations are SoA per direction with
// every block's 4096 cells contiguous (the reference's f[x,y,z,b,k] layout), so the 64 lanes of a
// wavefront touch 4 consecutive 64-byte rows of every direction array.  A workgroup's block, plane,
// domain and neighbour-block indices are uniform, which keeps them in scalar registers.
//
// Reference: src/collision/kernel_cumulant.jl:29-912 (stream_collide_cumulant!).
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

// how the population arriving in one direction is obtained
enum : int {
    SRC_REG = 0,     // pulled from the source cell
    SRC_INLET = 1,   // equilibrium at (1, u_inlet)                       kernel_cumulant.jl:235-239
    SRC_OWN = 2,     // the cell's own population, same direction (outlet) kernel_cumulant.jl:240-245
    SRC_OPP = 3,     // own opposite population: no-slip wall            kernel_cumulant.jl:270-271, 340-341
    SRC_MIRY = 4,    // own y-mirrored population: free slip             kernel_cumulant.jl:342-343
    SRC_MIRZ = 5,    // own z-mirrored population: free slip             kernel_cumulant.jl:344-345
    SRC_REST = 6,    // rest weight (no neighbour block on level 1)      kernel_cumulant.jl:389-390
    SRC_SOLID = 7    // halfway bounce-back from a solid source cell
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
template <int K, bool GEN, bool REG = false>
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
            if (c.zwall != 0 && (zl || zh)) return SRC_OPP;
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
// +- nbx*nby), so a source cell's index is arithmetic.  Each lane forms, once, the byte offsets of its
// nine in-plane source classes (cx, cy) - own cell shifted by -c in x and y, into the neighbour block
// where it crosses - and every direction's access adds a uniform z part to its scalar base pointer.
template <class SP>
struct RegAddr {
    uint32_t v[3][3];          // [cx+1][cy+1]: byte offset of cell (x - cx, y - cy, z)
    int64_t plane_cells;       // cells between consecutive block layers (nbx * nby * BS3)
    __device__ __forceinline__ void init(const CellCtx& c, int nbx, int nby) {
        plane_cells = (int64_t)nbx * nby * BS3;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int cx = i - 1, cy = j - 1;
                const int sx = c.lx - cx, sy = c.ly - cy;
                const int dx = sx < 0 ? -1 : (sx >= BS ? 1 : 0);
                const int dy = sy < 0 ? -1 : (sy >= BS ? 1 : 0);
                const int32_t idx = (int32_t)c.own + dx * BS3 + dy * nbx * BS3 + (sx - BS * dx - c.lx) + BS * (sy - BS * dy - c.ly);
                v[i][j] = (uint32_t)idx * (uint32_t)sizeof(typename SP::st);
            }
    }
    // uniform base of direction-slot `slot` for a source shifted by -cz in z
    __device__ __forceinline__ const char* base(const void* f, uint32_t n, int slot, int z, int cz) const {
        const int sz = z - cz;
        const int dz = sz < 0 ? -1 : (sz >= BS ? 1 : 0);
        const int64_t cells = (int64_t)slot * n + dz * plane_cells + (int64_t)(sz - BS * dz - z) * BS2;
        return (const char*)f + cells * (int64_t)sizeof(typename SP::st);
    }
};

// Population J of this cell's own post-collision state of the previous step: slot (x, J) after an
// AB step, slot (x, opp J) after an AA even step.
template <class SP, int PAT, int J>
__device__ __forceinline__ typename SP::T own_prev(const StepArgs& a, const CellCtx& c) {
    constexpr int slot = PAT == PAT_AB ? J : dopp(J);
    return SP::ld((const typename SP::st*)a.fin + (size_t)slot * a.ncell + c.own, wgt<typename SP::T>(J));
}

template <int J>
using dir_c = std::integral_constant<int, J>;

// The value of a boundary rule for direction K.  `own(dir_c<j>)` yields population j of the cell's
// own post-collision state of the step the rule refers to.
template <int K, class T, class OWN>
__device__ __forceinline__ T boundary_value(int kind, const StepArgs& a, const CellCtx& c, float u_in, int32_t step,
                                            OWN&& own) {
    switch (kind) {
        case SRC_INLET: {
            const float turb = a.cp.inlet_turb;
            const float noise = turb > 0.0f ? inlet_noise(c.gy + 1, c.gz + 1, step, 1234) * turb * u_in : 0.0f;
            const T u_inst = T(u_in) + T(noise);
            return feq_auto(T(1), u_inst, T(0), T(0), wgt<T>(K), T(dcx(K)), T(dcy(K)), T(dcz(K)), a.cp.compressible);
        }
        case SRC_OWN: return own(dir_c<K>{});
        case SRC_OPP:
        case SRC_SOLID: return own(dir_c<dopp(K)>{});
        case SRC_MIRY: return own(dir_c<dmir_y(K)>{});
        case SRC_MIRZ: return own(dir_c<dmir_z(K)>{});
        default: return wgt<T>(K);
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
template <class SP, int PAT, bool WALE, bool GEN, bool REG = false>
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
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr T w = wgt<T>(K);
        T val;
        if constexpr (REG && !GEN && PAT == PAT_AA_EVEN) {
            val = SP::ld((const st*)(ra.base(fi, n, K, c.z, 0) + ra.v[1][1]), w);
        } else if constexpr (REG && !GEN) {
            constexpr int slot = PAT == PAT_AB ? K : dopp(K);
            val = SP::ld((const st*)(ra.base(fi, n, slot, c.z, dcz(K)) + ra.v[dcx(K) + 1][dcy(K) + 1]), w);
        } else if constexpr (PAT == PAT_AA_EVEN) {
            // inlet links are evaluated here in full precision, as the pull step does, rather than
            // read back from a slot that went through the storage rounding
            bool inlet = false;
            if constexpr (GEN && dcx(K) > 0) inlet = (c.flags & BF_EDGE) && !c.per0 && c.gx == 0;
            if (inlet)
                val = boundary_value<K, T>(SRC_INLET, a, c, u_in, a.time_step, [&](auto) { return T(0); });
            else
                val = SP::ld(fi + (size_t)K * n + c.own, w);
        } else {
            uint32_t src = 0;
            const int kind = classify<K, GEN, REG>(a, c, src);
            if (kind == SRC_REG) {
                constexpr int slot = PAT == PAT_AB ? K : dopp(K);
                if constexpr (REG)
                    val = SP::ld((const st*)(ra.base(fi, n, slot, c.z, dcz(K)) + ra.v[dcx(K) + 1][dcy(K) + 1]), w);
                else
                    val = SP::ld(fi + (size_t)slot * n + src, w);
            } else if (PAT == PAT_AA_ODD && kind == SRC_OWN) {
                // outlet: the own population of direction K (cx = -1) sits in slot (x, opp K), which
                // the west neighbour overwrites during this step; the even step saved a copy
                val = SP::ld((const st*)a.obuf + ((size_t)a.oslot[b] * 9 + K / 3) * BS2 + c.ly * BS + c.z, w);
            } else {
                val = boundary_value<K, T>(kind, a, c, u_in, a.time_step,
                                           [&](auto jc) { return own_prev<SP, PAT, decltype(jc)::value>(a, c); });
            }
        }
        f[K] = val;
    });

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
        collide_cell<T>(f, q, a.cp, rho_o, ux_o, uy_o, uz_o);
    }
    if (a.vel_out) {
        T* vo = (T*)a.vel_out;
        vo[c.own] = ux_o; vo[n + c.own] = uy_o; vo[2 * (size_t)n + c.own] = uz_o;
    }
    if (a.rho_out) ((T*)a.rho_out)[c.own] = rho_o;

    const bool want_force = a.fpart != nullptr;
    const bool link_force = GEN && want_force && (c.flags & BF_SOLID) && !solid_own;
    const bool outlet_cell = GEN && PAT == PAT_AA_EVEN && (c.flags & BF_EDGE) && !c.per0 && c.gx == c.ext0 - 1;
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int J = dopp(K);   // the source of direction J is x + c_K, where f*_K goes
        if constexpr (REG && PAT == PAT_AB) {
            SP::sto((st*)(ra.base(fo, n, K, c.z, 0) + ra.v[1][1]), f[K], wgt<T>(K));
        } else if constexpr (REG && PAT == PAT_AA_EVEN) {
            SP::sto((st*)(ra.base(fo, n, J, c.z, 0) + ra.v[1][1]), f[K], wgt<T>(K));
            if constexpr (GEN && dcx(K) < 0) {
                if (outlet_cell) SP::sto((st*)a.obuf + ((size_t)a.oslot[b] * 9 + K / 3) * BS2 + c.ly * BS + c.z, f[K], wgt<T>(K));
            }
        } else if constexpr (REG && !GEN) {
            SP::sto((st*)(ra.base(fo, n, K, c.z, -dcz(K)) + ra.v[1 - dcx(K)][1 - dcy(K)]), f[K], wgt<T>(K));
        } else if constexpr (REG) {
            uint32_t dst = 0;
            const int kind = classify<J, GEN, REG>(a, c, dst);
            if (kind == SRC_REG) {
                SP::sto((st*)(ra.base(fo, n, K, c.z, -dcz(K)) + ra.v[1 - dcx(K)][1 - dcy(K)]), f[K], wgt<T>(K));
            } else if (kind != SRC_INLET) {
                const T v = boundary_value<J, T>(kind, a, c, u_in, a.time_step + 1,
                                                 [&](auto jc) { return f[decltype(jc)::value]; });
                SP::sto(fo + (size_t)J * n + c.own, v, wgt<T>(J));
            }
        } else if constexpr (PAT == PAT_AB) {
            SP::sto(fo + (size_t)K * n + c.own, f[K], wgt<T>(K));
        } else if constexpr (PAT == PAT_AA_EVEN) {
            SP::sto(fo + (size_t)J * n + c.own, f[K], wgt<T>(K));
            if constexpr (GEN && dcx(K) < 0) {
                if (outlet_cell) SP::sto((st*)a.obuf + ((size_t)a.oslot[b] * 9 + K / 3) * BS2 + c.ly * BS + c.z, f[K], wgt<T>(K));
            }
        } else {
            uint32_t dst = 0;
            const int kind = classify<J, GEN>(a, c, dst);
            if (kind == SRC_REG) {
                SP::sto(fo + (size_t)K * n + dst, f[K], wgt<T>(K));
            } else if (kind != SRC_INLET) {   // inlet links are evaluated by the even step itself
                const T v = boundary_value<J, T>(kind, a, c, u_in, a.time_step + 1,
                                                 [&](auto jc) { return f[decltype(jc)::value]; });
                SP::sto(fo + (size_t)J * n + c.own, v, wgt<T>(J));
            }
        }
        if constexpr (GEN) {
            if (link_force) {
                uint32_t dst = 0;
                const int kind = classify<J, GEN, REG>(a, c, dst);
                double m = 0.0;
                if (kind == SRC_SOLID) {
                    m = 2.0 * ((double)SP::rt(f[K], wgt<T>(K)) - (double)wgt<T>(K));
                } else if (PAT == PAT_AB && a.cp.bb_mode == 0 && kind == SRC_REG && ((c.lmask >> J) & 1u)) {
                    m = ((double)SP::rt(f[K], wgt<T>(K)) - (double)wgt<T>(K)) +
                        ((double)own_prev<SP, PAT, K>(a, c) - (double)wgt<T>(K));
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
#ifndef OAL_PIPE_WAVES
#define OAL_PIPE_WAVES 2
#endif
template <class SP, int PAT, bool WALE, int SEL = SEL_ALL>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(OAL_STEP_WAVES))) k_step(const StepArgs a) {
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
        step_cell<SP, PAT, WALE, false, true>(a, c, dm, true, false, F0, F1, F2);
    } else if constexpr (SEL == SEL_INTERIOR) {
        step_cell<SP, PAT, WALE, false>(a, c, dm, true, false, F0, F1, F2);
    } else {
        const bool inside = !(c.flags & BF_PARTIAL) || (c.gx < c.ext0 && c.gy < c.ext1 && c.gz < c.ext2);
        if (c.flags & BF_SOLID) c.lmask = a.lmask[(size_t)a.lslot[b] * BS3 + (c.own - b * BS3)];
        const bool solid_own = (c.lmask >> 27) & 1u;
        const bool active = inside && !(solid_own && a.cp.bb_mode == 1);
        if constexpr (SEL == SEL_BOUNDARY_REGULAR)
            step_cell<SP, PAT, WALE, true, true>(a, c, dm, active, solid_own, F0, F1, F2);
        else
            step_cell<SP, PAT, WALE, true>(a, c, dm, active, solid_own, F0, F1, F2);
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
        r[K] = SP::ld_raw((const st*)(ra.base(a.fin, a.ncell, K, z, dcz(K)) + ra.v[dcx(K) + 1][dcy(K) + 1]));
    });
}

__device__ __forceinline__ void wait_all_vmem() {
    __builtin_amdgcn_s_waitcnt(0x0f70);   // vmcnt(0), expcnt and lgkmcnt left open
    __builtin_amdgcn_sched_barrier(0);
}

template <class SP>
__device__ __forceinline__ void reg_collide_store(const StepArgs& a, const RegAddr<SP>& ra, const CellRates<typename SP::T>& q,
                                                  int z, const typename SP::raw (&r)[27]) {
    using T = typename SP::T;
    using st = typename SP::st;
    T f[27];
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        f[K] = SP::from_raw(r[K], wgt<T>(K));
    });
    T rho_o, ux_o, uy_o, uz_o;
    collide_cell<T>(f, q, a.cp, rho_o, ux_o, uy_o, uz_o);
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        SP::sto((st*)(ra.base(a.fout, a.ncell, K, z, 0) + ra.v[1][1]), f[K], wgt<T>(K));
    });
}

template <class SP>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(OAL_PIPE_WAVES))) k_step_pipe(const StepArgs a) {
    using T = typename SP::T;
    using st = typename SP::st;
    const int np = a.pipe_planes;
    const uint32_t groups = (uint32_t)(PLANES / np);
    const uint32_t litem = xcd_item(blockIdx.x, gridDim.x);
    const uint32_t b = (uint32_t)a.blist[litem / groups];
    const int z0 = (int)(litem % groups) * np;
    CellCtx c;
    c.b = b;
    c.lx = threadIdx.x & (BS - 1);
    c.ly = threadIdx.x >> 4;
    c.z = z0;
    c.own = b * BS3 + (uint32_t)z0 * BS2 + (uint32_t)c.ly * BS + (uint32_t)c.lx;
    const DevDomain* dm = a.doms + a.bdom[b];
    RegAddr<SP> ra;
    ra.init(c, dm->nbx, dm->nby);   // offsets of plane z0; plane z adds (z - z0) * BS2 cells
    // every domain parameter is read before the first store: reads issued after stores to the
    // population arrays could not be proven independent of them
    CellRates<T> q;
    if constexpr (sizeof(T) == 8) { q.omega1 = dm->omega1_d; q.omega4_diag = dm->omega4d_d; }
    else { q.omega1 = dm->omega1; q.omega4_diag = dm->omega4d; }
    q.forced = dm->forced != 0;
    q.Fx = T(dm->F[0]); q.Fy = T(dm->F[1]); q.Fz = T(dm->F[2]);
    const uint32_t step = BS2 * (uint32_t)sizeof(st);
    typename SP::raw fa[27], fb[27];
    reg_load_raw<SP>(a, ra, z0, fa);
    for (int z = z0; z < z0 + np; z += 2) {
        RegAddr<SP> rn = ra;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) rn.v[i][j] += step;
        wait_all_vmem();
        if (z + 1 < z0 + np) reg_load_raw<SP>(a, rn, z + 1, fb);
        reg_collide_store<SP>(a, ra, q, z, fa);
        if (z + 1 >= z0 + np) break;
        RegAddr<SP> rm = rn;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) rm.v[i][j] += step;
        wait_all_vmem();
        if (z + 2 < z0 + np) reg_load_raw<SP>(a, rm, z + 2, fa);
        reg_collide_store<SP>(a, rn, q, z + 1, fb);
        ra = rm;
    }
}

// One step of a level: up to four launches over disjoint block lists (regular interior, other
// interior, regular boundary, boundary), each compiled for its own register budget.  The boundary
// launches go to `side` (when given) so that they overlap the bandwidth-bound interior launch; the
// caller joins the two streams.
template <class SP, int PAT, bool WALE>
static hipError_t launch_pat(const StepArgs& a, hipStream_t stream, hipStream_t side) {
    StepArgs r = a;
    hipStream_t bs = side ? side : stream;
    if (a.n_boundary_regular) {
        r.blist = a.blist_boundary_regular;
        hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY_REGULAR>), dim3(a.n_boundary_regular * PLANES), dim3(WG), 0,
                           bs, r);
    }
    if (a.n_boundary) {
        r.blist = a.blist_boundary;
        hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_BOUNDARY>), dim3(a.n_boundary * PLANES), dim3(WG), 0, bs, r);
    }
    if (a.n_regular) {
        r.blist = a.blist_regular;
        if (PAT == PAT_AB && !WALE && a.pipe_planes > 1 && sizeof(typename SP::T) == 4)   // (no force on interior blocks)
            hipLaunchKernelGGL((k_step_pipe<SP>), dim3(a.n_regular * (PLANES / a.pipe_planes)), dim3(WG), 0, stream, r);
        else
            hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_REGULAR>), dim3(a.n_regular * PLANES), dim3(WG), 0, stream, r);
    }
    if (a.n_interior) {
        r.blist = a.blist_interior;
        hipLaunchKernelGGL((k_step<SP, PAT, WALE, SEL_INTERIOR>), dim3(a.n_interior * PLANES), dim3(WG), 0, stream, r);
    }
    return hipGetLastError();
}

template <class SP>
static hipError_t launch_sp(int stream_mode, bool wale, const StepArgs& a, hipStream_t s, hipStream_t side) {
    if (stream_mode == 0) return wale ? launch_pat<SP, PAT_AB, true>(a, s, side) : launch_pat<SP, PAT_AB, false>(a, s, side);
    if (a.parity == 0)
        return wale ? launch_pat<SP, PAT_AA_EVEN, true>(a, s, side) : launch_pat<SP, PAT_AA_EVEN, false>(a, s, side);
    return wale ? launch_pat<SP, PAT_AA_ODD, true>(a, s, side) : launch_pat<SP, PAT_AA_ODD, false>(a, s, side);
}

hipError_t launch_step(int storage, int stream_mode, bool wale, const StepArgs& a, hipStream_t stream, hipStream_t side) {
    switch (storage) {
        case 0: return launch_sp<StF32>(stream_mode, wale, a, stream, side);
        case 1: return launch_sp<StF64>(stream_mode, wale, a, stream, side);
        case 2: return launch_sp<StF16S>(stream_mode, wale, a, stream, side);
        case 3: return launch_sp<StBF16S>(stream_mode, wale, a, stream, side);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------------------------------------
// Bandwidth probe: a streaming 16-byte copy, the measured HBM ceiling of the roofline
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(WG) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * WG;
    for (size_t i = (size_t)blockIdx.x * WG + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    const unsigned grid = (unsigned)((n + WG - 1) / WG < 65536 * 8 ? (n + WG - 1) / WG : 65536 * 8);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(WG), 0, s, (const uint4*)src, (uint4*)dst, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Initialisation: init_eq! (src/main.jl:546-572), per-cell or uniform (rho, u)
// ------------------------------------------------------------------------------------------------
template <class SP>
__global__ void __launch_bounds__(WG) k_init_eq(typename SP::st* f, uint32_t ncell, const double* rho, const double* u,
                                                double rho0, double ux0, double uy0, double uz0, int permute,
                                                typename SP::T* vel) {
    using T = typename SP::T;
    const uint32_t cidx = blockIdx.x * WG + threadIdx.x;
    if (cidx >= ncell) return;
    const T r = rho ? T(rho[cidx]) : T(rho0);
    const T ux = u ? T(u[cidx]) : T(ux0);
    const T uy = u ? T(u[ncell + cidx]) : T(uy0);
    const T uz = u ? T(u[2 * (size_t)ncell + cidx]) : T(uz0);
    const T cs2 = T(1) / T(3);
    const T cs4 = cs2 * cs2;
    const T usq = ux * ux + uy * uy + uz * uz;
    const T inv_cs2 = T(1) / cs2, inv_2cs4 = T(0.5) / cs4, inv_2cs2 = T(0.5) / cs2;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const T cu = T(dcx(k)) * ux + T(dcy(k)) * uy + T(dcz(k)) * uz;
        const T feq = r * wgt<T>(k) * (T(1) + inv_cs2 * cu + inv_2cs4 * cu * cu - inv_2cs2 * usq);
        SP::sto(f + (size_t)(permute ? 26 - k : k) * ncell + cidx, feq, wgt<T>(k));
    }
    if (vel) { vel[cidx] = ux; vel[ncell + cidx] = uy; vel[2 * (size_t)ncell + cidx] = uz; }
}

template <class SP>
static hipError_t init_sp(void* f, uint32_t ncell, const double* rho, const double* u, double rho0, double ux0,
                          double uy0, double uz0, int permute, void* vel, hipStream_t s) {
    hipLaunchKernelGGL((k_init_eq<SP>), dim3((ncell + WG - 1) / WG), dim3(WG), 0, s, (typename SP::st*)f, ncell, rho, u,
                       rho0, ux0, uy0, uz0, permute, (typename SP::T*)vel);
    return hipGetLastError();
}

hipError_t launch_init_eq(int storage, void* f, uint32_t ncell, const double* rho, const double* u, double rho0,
                          double ux0, double uy0, double uz0, int permute, void* vel, hipStream_t s) {
    switch (storage) {
        case 0: return init_sp<StF32>(f, ncell, rho, u, rho0, ux0, uy0, uz0, permute, vel, s);
        case 1: return init_sp<StF64>(f, ncell, rho, u, rho0, ux0, uy0, uz0, permute, vel, s);
        case 2: return init_sp<StF16S>(f, ncell, rho, u, rho0, ux0, uy0, uz0, permute, vel, s);
        case 3: return init_sp<StBF16S>(f, ncell, rho, u, rho0, ux0, uy0, uz0, permute, vel, s);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------------------------------------
// Layout conversion between host populations (unshifted f32/f64) and the storage form
// ------------------------------------------------------------------------------------------------
template <class SP, class H>
__global__ void __launch_bounds__(WG) k_convert_in(typename SP::st* f, const H* src, uint32_t ncell, int permute) {
    using T = typename SP::T;
    const uint32_t cidx = blockIdx.x * WG + threadIdx.x;
    const int k = blockIdx.y;
    if (cidx >= ncell) return;
    SP::sto(f + (size_t)(permute ? 26 - k : k) * ncell + cidx, T(src[(size_t)k * ncell + cidx]), wgt<T>(k));
}

template <class SP, class H>
__global__ void __launch_bounds__(WG) k_convert_out(const typename SP::st* f, H* dst, uint32_t ncell, int permute) {
    using T = typename SP::T;
    const uint32_t cidx = blockIdx.x * WG + threadIdx.x;
    const int k = blockIdx.y;
    if (cidx >= ncell) return;
    dst[(size_t)k * ncell + cidx] = H(SP::ld(f + (size_t)(permute ? 26 - k : k) * ncell + cidx, wgt<T>(k)));
}

template <class SP>
static hipError_t conv_in_sp(void* f, const void* src, int dt, uint32_t ncell, int permute, hipStream_t s) {
    dim3 grid((ncell + WG - 1) / WG, 27);
    if (dt == 0) hipLaunchKernelGGL((k_convert_in<SP, float>), grid, dim3(WG), 0, s, (typename SP::st*)f, (const float*)src, ncell, permute);
    else hipLaunchKernelGGL((k_convert_in<SP, double>), grid, dim3(WG), 0, s, (typename SP::st*)f, (const double*)src, ncell, permute);
    return hipGetLastError();
}
template <class SP>
static hipError_t conv_out_sp(const void* f, void* dst, int dt, uint32_t ncell, int permute, hipStream_t s) {
    dim3 grid((ncell + WG - 1) / WG, 27);
    if (dt == 0) hipLaunchKernelGGL((k_convert_out<SP, float>), grid, dim3(WG), 0, s, (const typename SP::st*)f, (float*)dst, ncell, permute);
    else hipLaunchKernelGGL((k_convert_out<SP, double>), grid, dim3(WG), 0, s, (const typename SP::st*)f, (double*)dst, ncell, permute);
    return hipGetLastError();
}

hipError_t launch_convert_in(int storage, void* f, const void* src, int dt, uint32_t ncell, int permute, hipStream_t s) {
    switch (storage) {
        case 0: return conv_in_sp<StF32>(f, src, dt, ncell, permute, s);
        case 1: return conv_in_sp<StF64>(f, src, dt, ncell, permute, s);
        case 2: return conv_in_sp<StF16S>(f, src, dt, ncell, permute, s);
        case 3: return conv_in_sp<StBF16S>(f, src, dt, ncell, permute, s);
        default: return hipErrorInvalidValue;
    }
}
hipError_t launch_convert_out(int storage, const void* f, void* dst, int dt, uint32_t ncell, int permute, hipStream_t s) {
    switch (storage) {
        case 0: return conv_out_sp<StF32>(f, dst, dt, ncell, permute, s);
        case 1: return conv_out_sp<StF64>(f, dst, dt, ncell, permute, s);
        case 2: return conv_out_sp<StF16S>(f, dst, dt, ncell, permute, s);
        case 3: return conv_out_sp<StBF16S>(f, dst, dt, ncell, permute, s);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------------------------------------
// Macroscopic output in FP64: rho = 1 + sum (f_k - w_k), u = sum c_k (f_k - w_k) / rho
// ------------------------------------------------------------------------------------------------
template <class SP>
__global__ void __launch_bounds__(WG) k_macro(const typename SP::st* f, double* rho, double* u, uint32_t ncell,
                                              int permute, typename SP::T* vel) {
    using T = typename SP::T;
    const uint32_t cidx = blockIdx.x * WG + threadIdx.x;
    if (cidx >= ncell) return;
    double dr = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const double s = (double)SP::ld(f + (size_t)(permute ? 26 - k : k) * ncell + cidx, wgt<T>(k)) - (double)wgt<T>(k);
        dr += s; jx += s * dcx(k); jy += s * dcy(k); jz += s * dcz(k);
    }
    const double r = 1.0 + dr;
    if (rho) rho[cidx] = r;
    if (u) { u[cidx] = jx / r; u[ncell + cidx] = jy / r; u[2 * (size_t)ncell + cidx] = jz / r; }
    if (vel) { vel[cidx] = T(jx / r); vel[ncell + cidx] = T(jy / r); vel[2 * (size_t)ncell + cidx] = T(jz / r); }
}

template <class SP>
static hipError_t macro_sp(const void* f, double* rho, double* u, uint32_t ncell, int permute, void* vel, hipStream_t s) {
    hipLaunchKernelGGL((k_macro<SP>), dim3((ncell + WG - 1) / WG), dim3(WG), 0, s, (const typename SP::st*)f, rho, u, ncell,
                       permute, (typename SP::T*)vel);
    return hipGetLastError();
}

hipError_t launch_macro(int storage, const void* f, double* rho, double* u, uint32_t ncell, int permute, void* vel,
                        hipStream_t s) {
    switch (storage) {
        case 0: return macro_sp<StF32>(f, rho, u, ncell, permute, vel, s);
        case 1: return macro_sp<StF64>(f, rho, u, ncell, permute, vel, s);
        case 2: return macro_sp<StF16S>(f, rho, u, ncell, permute, vel, s);
        case 3: return macro_sp<StBF16S>(f, rho, u, ncell, permute, vel, s);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------------------------------------
// AA pattern: outlet side buffer from a state in stored form (after upload / initialisation), the
// copy the even step makes of every outlet-face population with cx = -1
// ------------------------------------------------------------------------------------------------
template <class SP>
__global__ void __launch_bounds__(WG) k_aa_outlet_prep(const StepArgs a) {
    using st = typename SP::st;
    const uint32_t b = blockIdx.x;
    const int32_t os = a.oslot[b];
    if (os < 0) return;
    const DevDomain* dm = a.doms + a.bdom[b];
    const int lx = dm->ext[0] - 1 - a.borg[3 * b];   // the outlet plane of this block
    const int ly = threadIdx.x & (BS - 1), z = threadIdx.x >> 4;
    const uint32_t own = b * BS3 + (uint32_t)z * BS2 + (uint32_t)ly * BS + (uint32_t)lx;
    const st* f = (const st*)a.fin;
    st* o = (st*)a.obuf;
#pragma unroll
    for (int q = 0; q < 9; ++q) o[((size_t)os * 9 + q) * BS2 + ly * BS + z] = f[(size_t)(26 - 3 * q) * a.ncell + own];
}

hipError_t launch_aa_outlet_prep(int storage, const StepArgs& a, hipStream_t s) {
    switch (storage) {
        case 0: hipLaunchKernelGGL((k_aa_outlet_prep<StF32>), dim3(a.nb), dim3(WG), 0, s, a); break;
        case 1: hipLaunchKernelGGL((k_aa_outlet_prep<StF64>), dim3(a.nb), dim3(WG), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_aa_outlet_prep<StF16S>), dim3(a.nb), dim3(WG), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_aa_outlet_prep<StBF16S>), dim3(a.nb), dim3(WG), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// FP64 mass and momentum partials, one workgroup per block (16 cells per lane); then a per-domain
// reduction over contiguous item ranges in a fixed order, so results are bit-reproducible.
// ------------------------------------------------------------------------------------------------
template <class SP>
__global__ void __launch_bounds__(WG) k_mass(const typename SP::st* f, const uint8_t* solid, const int32_t* borg,
                                             const int32_t* bdom, const DevDomain* doms, uint32_t nb, int permute,
                                             double* part) {
    using T = typename SP::T;
    const uint32_t b = blockIdx.x;
    const DevDomain* dm = doms + bdom[b];
    const uint32_t ncell = nb * BS3;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int z = 0; z < BS; ++z) {
        const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
        const uint32_t cidx = b * BS3 + z * BS2 + ly * BS + lx;
        const bool inside = borg[3 * b] + lx < dm->ext[0] && borg[3 * b + 1] + ly < dm->ext[1] && borg[3 * b + 2] + z < dm->ext[2];
        if (!inside || (solid && solid[cidx])) continue;
        double dr = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
        for (int k = 0; k < 27; ++k) {
            const double s = (double)SP::ld(f + (size_t)(permute ? 26 - k : k) * ncell + cidx, wgt<T>(k)) - (double)wgt<T>(k);
            dr += s; jx += s * dcx(k); jy += s * dcy(k); jz += s * dcz(k);
        }
        acc[0] += 1.0 + dr; acc[1] += jx; acc[2] += jy; acc[3] += jz;
    }
    __shared__ double red[4][WG / 64];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = acc[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < WG / 64; ++w) s += red[threadIdx.x][w];
        part[(size_t)b * 4 + threadIdx.x] = s;
    }
}

hipError_t launch_mass(int storage, const void* f, const uint8_t* solid, const int32_t* borg, const int32_t* bdom,
                       const DevDomain* doms, uint32_t nb, int permute, double* part, hipStream_t s) {
    switch (storage) {
        case 0: hipLaunchKernelGGL((k_mass<StF32>), dim3(nb), dim3(WG), 0, s, (const float*)f, solid, borg, bdom, doms, nb, permute, part); break;
        case 1: hipLaunchKernelGGL((k_mass<StF64>), dim3(nb), dim3(WG), 0, s, (const double*)f, solid, borg, bdom, doms, nb, permute, part); break;
        case 2: hipLaunchKernelGGL((k_mass<StF16S>), dim3(nb), dim3(WG), 0, s, (const _Float16*)f, solid, borg, bdom, doms, nb, permute, part); break;
        case 3: hipLaunchKernelGGL((k_mass<StBF16S>), dim3(nb), dim3(WG), 0, s, (const uint16_t*)f, solid, borg, bdom, doms, nb, permute, part); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// one workgroup per domain: lanes stride over the domain's items, then a fixed-order tree
__global__ void __launch_bounds__(WG) k_reduce(const ReduceArgs r) {
    const int d = blockIdx.x;
    const int i0 = r.dom_item0[d], i1 = r.dom_item0[d + 1];
    __shared__ double sh[4][WG];
    for (int q = 0; q < 4; ++q) {
        double s = 0.0;
        if (q < r.width)
            for (int i = i0 + (int)threadIdx.x; i < i1; i += WG) s += r.part[(size_t)i * 4 + q];
        sh[q][threadIdx.x] = s;
    }
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) r.out[(size_t)d * 4 + threadIdx.x] = sh[threadIdx.x][0];
}

hipError_t launch_reduce(const ReduceArgs& r, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3(r.n_domains), dim3(WG), 0, s, r);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Stability guard: rho <= rho_min, |u| >= u_max or a non-finite population in a fluid cell
// (src/octree_q19/src/io_run.jl:31-65)
// ------------------------------------------------------------------------------------------------
template <class SP>
__global__ void __launch_bounds__(WG) k_guard(const typename SP::st* f, const uint8_t* solid, const int32_t* borg,
                                              const int32_t* bdom, const DevDomain* doms, uint32_t nb, float rho_min,
                                              float u_max, unsigned long long* count) {
    using T = typename SP::T;
    const uint32_t b = blockIdx.x;
    const DevDomain* dm = doms + bdom[b];
    const uint32_t ncell = nb * BS3;
    unsigned bad = 0;
    for (int z = 0; z < BS; ++z) {
        const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
        const uint32_t cidx = b * BS3 + z * BS2 + ly * BS + lx;
        const bool inside = borg[3 * b] + lx < dm->ext[0] && borg[3 * b + 1] + ly < dm->ext[1] && borg[3 * b + 2] + z < dm->ext[2];
        if (!inside || (solid && solid[cidx])) continue;
        float r = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
        bool finite = true;
#pragma unroll
        for (int k = 0; k < 27; ++k) {
            const float v = (float)SP::ld(f + (size_t)k * ncell + cidx, wgt<T>(k));
            finite = finite && isfinite(v);
            r += v; jx += v * dcx(k); jy += v * dcy(k); jz += v * dcz(k);
        }
        const float u2 = (jx * jx + jy * jy + jz * jz) / (r * r);
        if (!finite || !(r > rho_min) || !(u2 < u_max * u_max)) ++bad;
    }
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_xor(bad, off, 64);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(count, (unsigned long long)bad);
}

hipError_t launch_guard(int storage, const void* f, const uint8_t* solid, const int32_t* borg, const int32_t* bdom,
                        const DevDomain* doms, uint32_t nb, float rho_min, float u_max, unsigned long long* count,
                        hipStream_t s) {
    switch (storage) {
        case 0: hipLaunchKernelGGL((k_guard<StF32>), dim3(nb), dim3(WG), 0, s, (const float*)f, solid, borg, bdom, doms, nb, rho_min, u_max, count); break;
        case 1: hipLaunchKernelGGL((k_guard<StF64>), dim3(nb), dim3(WG), 0, s, (const double*)f, solid, borg, bdom, doms, nb, rho_min, u_max, count); break;
        case 2: hipLaunchKernelGGL((k_guard<StF16S>), dim3(nb), dim3(WG), 0, s, (const _Float16*)f, solid, borg, bdom, doms, nb, rho_min, u_max, count); break;
        case 3: hipLaunchKernelGGL((k_guard<StBF16S>), dim3(nb), dim3(WG), 0, s, (const uint16_t*)f, solid, borg, bdom, doms, nb, rho_min, u_max, count); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace oal
