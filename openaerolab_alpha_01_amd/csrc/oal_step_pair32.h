// The pipelined step of regular blocks for FP32 storage, two cells per lane.  Included by oal_step.h
// after oal_step_pair.h, whose face-rule helpers (PairFaces, face_rule, PairInlet) it shares.
//
// A lane steps the two cells (2j, 2j + 1) of a row of 16, so a workgroup of 128 lanes covers a 16 x 16
// plane and every population moves as a 64-bit word holding both cells: one 8-byte load and one 8-byte
// store per direction and lane instead of two 4-byte ones of each, every one of them aligned.
//
// Gather (pull): the word a lane loads for direction k is the aligned pair (2j, 2j + 1) of the source
// row (y - cy, z - cz).  For cx = 0 that is the pair of sources itself.  For cx = +1 the sources are
// (2j - 1, 2j): the upper cell of the left lane's word (DPP row_shr:1) and the lower cell of its own;
// the lane at j = 0 takes the left cell from the x-neighbour block instead (its cell 15: one extra
// 4-byte load for those lanes only).  cx = -1 mirrors this with the right lane (row_shl:1) and the
// neighbour block's cell 0.  Rows of 8 lanes sit in the 16-lane DPP rows two at a time; the lanes where
// a shift would cross between them are the j = 0 / j = 7 lanes that read the neighbour block.
//
// Domain-face rules (FACE_RULES): as the 16-bit pair kernel's (kernel_cumulant.jl:233-345): a y or z face
// rule takes the lane's own pair unshifted; the inlet replaces cell 0's arriving populations with
// cx = +1 by the inlet equilibrium after the load; the outlet gives cell 15 its own population of the
// same direction.  Each cell goes through the arithmetic of the one-cell kernels, so the results are
// the same bit for bit.
#pragma once

namespace oal {

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Addresses: one buffer resource per workgroup based `bias` bytes below the own block's start (one
// layer of blocks, one row of blocks and one block), so that every offset is non-negative.  v[j] is the
// per-lane offset of the aligned source pair in row y - cy (cy = j - 1); the uniform part (scalar
// offset) is the layer of blocks, the direction slot and the plane.
struct Pair32Addr {
    uint32_t v[3];
    __amdgpu_buffer_rsrc_t res;
    uint32_t plane;   // bytes between consecutive block layers
    __device__ __forceinline__ void init(const void* f, uint32_t b, int jx, int ly, int nbx, int nby) {
        const int64_t blk = 27 * BS3 * 4;
        plane = (uint32_t)((int64_t)nbx * nby * blk);
        const int64_t base = (int64_t)b * blk - (int64_t)(nbx + 1) * blk - (int64_t)plane;
        res = __builtin_amdgcn_make_buffer_rsrc((char*)f + base, (short)0, (int)0xffffffff, 0x00020000);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int ys = ly - (j - 1);
            const int dy = ys < 0 ? -1 : (ys >= BS ? 1 : 0);
            v[j] = (uint32_t)((dy + 1) * nbx + 1) * (uint32_t)blk + (uint32_t)(((ys - BS * dy) * BS + 2 * jx) * 4);
        }
    }
    // uniform offset of direction slot `slot` at plane z, for sources shifted by -cz in z
    __device__ __forceinline__ uint32_t soff(int slot, int z, int cz) const {
        const int sz = z - cz;
        const int dz = sz < 0 ? -1 : (sz >= BS ? 1 : 0);
        return (uint32_t)(dz + 1) * plane + (uint32_t)(slot * (BS3 * 4) + (sz - BS * dz) * (BS2 * 4));
    }
    __device__ __forceinline__ u32x2 load2(uint32_t voff, uint32_t so) const {
        return __builtin_amdgcn_raw_buffer_load_b64(res, (int)voff, (int)so, 0);
    }
    __device__ __forceinline__ uint32_t load1(uint32_t voff, uint32_t so) const {
        return __builtin_amdgcn_raw_buffer_load_b32(res, (int)voff, (int)so, 0);
    }
};

// the x-neighbour blocks' cells adjoining the row: the -x block's cell 15 from pair 0, the +x block's
// cell 0 from pair 7 (byte offsets)
constexpr int32_t PAIR32_EDGE_M = (-27 * BS3 + 15) * 4;
constexpr int32_t PAIR32_EDGE_P = (27 * BS3 - 14) * 4;

struct Pair32Raw {
    u32x2 p[27];
    uint32_t e[18];   // the x-neighbour cell of the directions with cx != 0 on the row's end lanes
};

template <int FACE>
__device__ __forceinline__ void pair32_load(const Pair32Addr& pa, const PairFaces& fc, int jx, int z, Pair32Raw& r) {
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int cx = dcx(K), cy = dcy(K), cz = dcz(K);
        const uint32_t so = pa.soff(K, z, cz);
        if constexpr (FACE == FACE_NONE) {
            r.p[K] = pa.load2(pa.v[cy + 1], so);
            if constexpr (cx > 0) {
                if (jx == 0) r.e[edge_slot<K>()] = pa.load1(pa.v[cy + 1] + (uint32_t)PAIR32_EDGE_M, so);
            } else if constexpr (cx < 0) {
                if (jx == 7) r.e[edge_slot<K>()] = pa.load1(pa.v[cy + 1] + (uint32_t)PAIR32_EDGE_P, so);
            }
        } else {
            // a face lane reads the rule's population from its own row and plane; the regular source
            // beyond the face is never addressed
            uint32_t voff = pa.v[cy + 1] + so;
            bool face = false;
            if constexpr (cy != 0 || cz != 0) {
                int j;
                face = face_rule<K>(fc, z, j);
                const uint32_t own = pa.v[1] + pa.soff(j, z, 0);
                voff = face ? own : voff;
            }
            r.p[K] = pa.load2(voff, 0);
            if constexpr (cx > 0) {
                // (the inlet's cell 0 is replaced after the load: no x-neighbour cell there)
                if (jx == 0 && !fc.xlo && !face) r.e[edge_slot<K>()] = pa.load1(pa.v[cy + 1] + (uint32_t)PAIR32_EDGE_M + so, 0);
            } else if constexpr (cx < 0) {
                if (jx == 7) {
                    if (fc.xhi)   // outlet: cell 15's own population of direction K (the upper cell of the own pair)
                        r.e[edge_slot<K>()] = pa.load1(pa.v[1] + pa.soff(K, z, 0) + 4u, 0);
                    else if (!face)
                        r.e[edge_slot<K>()] = pa.load1(pa.v[cy + 1] + (uint32_t)PAIR32_EDGE_P + so, 0);
                }
            }
        }
    });
}

// the two sources of (cell 2j, cell 2j + 1) for every direction: .x cell 2j's, .y cell 2j + 1's
template <int FACE>
__device__ __forceinline__ void pair32_sources(const Pair32Raw& r, const PairFaces& fc, int jx, int z, u32x2 (&s)[27]) {
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int cx = dcx(K), cy = dcy(K), cz = dcz(K);
        u32x2 v;
        if constexpr (cx == 0) {
            v = r.p[K];
        } else if constexpr (cx > 0) {
            // (2j - 1, 2j): the left lane's upper cell, the own lower cell
            const uint32_t left = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r.p[K].y, 0x111, 0xf, 0xf, false);
            v.x = jx == 0 ? r.e[edge_slot<K>()] : left;
            v.y = r.p[K].x;
        } else {
            // (2j + 1, 2j + 2): the own upper cell, the right lane's lower cell
            const uint32_t right = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r.p[K].x, 0x101, 0xf, 0xf, false);
            v.x = r.p[K].y;
            v.y = jx == 7 ? r.e[edge_slot<K>()] : right;
        }
        if constexpr (FACE == FACE_RULES) {
            if constexpr (cx != 0 && (cy != 0 || cz != 0)) {
                int j;
                if (face_rule<K>(fc, z, j)) v = r.p[K];   // the own pair of the rule, unshifted
            }
            if constexpr (cx < 0) {
                // outlet: cell 15 keeps its own population
                if (fc.xhi && jx == 7) v.y = r.e[edge_slot<K>()];
            }
        }
        s[K] = v;
    });
}

// collide the pair and leave its post-collision populations in s
template <bool UNIT>
__device__ __forceinline__ void pair32_collide(const StepArgs& a, const CellRates<float>& q, const PairInlet* inlet,
                                               u32x2 (&s)[27]) {
    float rho_o, ux_o, uy_o, uz_o;
    float f[27];
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        f[K] = __uint_as_float(s[K].x);
    });
    if (inlet && inlet->on) {
        CellCtx c{};
        c.gy = inlet->gy;
        c.gz = inlet->gz;
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            if constexpr (dcx(K) > 0)
                f[K] = boundary_value<K, float>(SRC_INLET, a, c, inlet->u_in, a.step(), [&](auto) { return 0.0f; });
        });
    }
    collide_cell<float, UNIT>(f, q, a.cp, rho_o, ux_o, uy_o, uz_o);
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        s[K].x = __float_as_uint(f[K]);
        f[K] = __uint_as_float(s[K].y);
    });
    collide_cell<float, UNIT>(f, q, a.cp, rho_o, ux_o, uy_o, uz_o);
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        s[K].y = __float_as_uint(f[K]);
    });
}

template <bool UNIT, int FACE>
constexpr int pair32_waves() { return 2; }

// Regular blocks (no solid near, whole; FACE_RULES: on domain faces too), AB pattern, FP32 storage: a
// workgroup of 128 lanes walks `a.pipe_planes` planes of one block with the next plane's loads in flight
// during the current plane's collision (the schedule of k_step_pipe).
template <int NP, bool UNIT, int FACE>
__global__ void __launch_bounds__(PWG)
__attribute__((amdgpu_waves_per_eu(pair32_waves<UNIT, FACE>(), pair32_waves<UNIT, FACE>())))
k_step_pair32(const StepArgs a) {
    const int np = a.pipe_planes;   // run-time bound equal to NP (see k_step_pipe)
    constexpr uint32_t groups = (uint32_t)(PLANES / NP);
    const PipeItem it = pipe_item<true>(groups, np);
    const uint32_t b = (uint32_t)a.blist[it.pos];
    const int z0 = it.z0;
    const int jx = (int)(threadIdx.x & 7u), ly = (int)(threadIdx.x >> 3);
    const DevDomain* dm = a.doms + a.bdom[b];
    Pair32Addr pa, po;   // the populations read and those written
    pa.init(a.fin, b, jx, ly, dm->nbx, dm->nby);
    po.init(a.fout, b, jx, ly, dm->nbx, dm->nby);
    CellRates<float> q;
    q.omega1 = dm->omega1;
    q.omega4_diag = dm->omega4d;
    q.forced = dm->forced != 0;
    q.Fx = dm->F[0]; q.Fy = dm->F[1]; q.Fz = dm->F[2];
    PairFaces fc{};
    PairInlet in{};
    if constexpr (FACE == FACE_RULES) {
        const FaceBits fb = face_bits(a, b, dm);
        fc.yl = fb.ylo && ly == 0;
        fc.yh = fb.yhi && ly == BS - 1;
        fc.zlo = fb.zlo;
        fc.zhi = fb.zhi;
        fc.xlo = fb.xlo;
        fc.xhi = fb.xhi;
        fc.zwall = dm->zwall != 0;
        fc.ymin_wall = dm->ymin_wall == 1;
        in.on = fb.xlo && jx == 0;
        in.u_in = dm->u_in;
        in.gy = a.borg[3 * b + 1] + ly;
        in.gz = a.borg[3 * b + 2];   // + z per plane
    }
    Pair32Raw raw{};   // (the neighbour cells are loaded on the row's end lanes only)
    pair32_load<FACE>(pa, fc, jx, z0, raw);
    for (int z = z0; z < z0 + np; ++z) {
        wait_all_vmem();
        u32x2 s[27];
        pair32_sources<FACE>(raw, fc, jx, z, s);
        __builtin_amdgcn_sched_barrier(0);
        if (z + 1 < z0 + np) pair32_load<FACE>(pa, fc, jx, z + 1, raw);
        if constexpr (FACE == FACE_RULES) {
            PairInlet iz = in;
            iz.gz += z;
            pair32_collide<UNIT>(a, q, &iz, s);
        } else {
            pair32_collide<UNIT>(a, q, nullptr, s);
        }
        const uint32_t so = po.soff(0, z, 0);
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            // (aux 2: the non-temporal hint, as the other step kernels' stores)
            __builtin_amdgcn_raw_buffer_store_b64(s[K], po.res, (int)po.v[1], (int)(so + K * (BS3 * 4)), OAL_PAIR_STORE_AUX);
        });
    }
}

template <bool UNIT, int FACE>
static void launch_pair32_u(int np, dim3 grid, hipStream_t s, const StepArgs& r) {
    switch (np) {
        case 2: hipLaunchKernelGGL((k_step_pair32<2, UNIT, FACE>), grid, dim3(PWG), 0, s, r); break;
        case 4: hipLaunchKernelGGL((k_step_pair32<4, UNIT, FACE>), grid, dim3(PWG), 0, s, r); break;
        case 8: hipLaunchKernelGGL((k_step_pair32<8, UNIT, FACE>), grid, dim3(PWG), 0, s, r); break;
        default: hipLaunchKernelGGL((k_step_pair32<16, UNIT, FACE>), grid, dim3(PWG), 0, s, r); break;
    }
}
// (a template, so that only the FP32 translation unit instantiates the kernels)
template <class SP>
static void launch_pair32(int np, bool face, dim3 grid, hipStream_t s, const StepArgs& r) {
    static_assert(SP::code == StF32::code, "the two-cells-per-lane FP32 kernel steps FP32 storage");
    if (r.cp.unit) {
        if (face) launch_pair32_u<true, FACE_RULES>(np, grid, s, r);
        else launch_pair32_u<true, FACE_NONE>(np, grid, s, r);
    } else {
        if (face) launch_pair32_u<false, FACE_RULES>(np, grid, s, r);
        else launch_pair32_u<false, FACE_NONE>(np, grid, s, r);
    }
}

}  // namespace oal
