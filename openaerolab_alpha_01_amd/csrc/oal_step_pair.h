// The pipelined step of regular blocks for the 16-bit storage forms, two cells per lane.  Included by
// oal_step.h.
//
// A lane steps the two cells (2j, 2j + 1) of a row of 16, so a workgroup of 128 lanes covers a
// 16 x 16 plane and every population moves as a 32-bit word holding both cells: one load and one store
// per direction and lane instead of two of each, and half the address and bookkeeping work per cell.
//
// Gather (pull): the word a lane loads for direction k is the aligned pair (2j, 2j + 1) of the source
// row (y - cy, z - cz).  For cx = 0 that is the pair of sources itself.  For cx = +1 the sources are
// (2j - 1, 2j): the upper cell of the left lane's word and the lower cell of its own, brought together
// with a lane shift (DPP row_shr:1) and a byte alignment; the lane at j = 0 takes the left cell from the
// x-neighbour block instead (x = 15 there: one extra load for those lanes only).  cx = -1 mirrors this
// with the right lane (row_shl:1) and the neighbour block's x = 0.  Rows of 8 lanes sit in the 16-lane
// DPP rows two at a time; the lanes where a shift would cross between them are the j = 0 / j = 7 lanes
// that read the neighbour block.
//
// Domain-face rules (FACE_RULES, whole blocks on a non-periodic face; kernel_cumulant.jl:233-345 in
// the order of classify): on a y or z face the rule replaces a direction's source by a population of
// the cell's own state, the same for both cells of a lane (they share y and z), so the lane loads that
// pair from its own row and plane instead and takes it unshifted.  The x faces touch one cell of the
// row's end lanes: the inlet (x = 0) replaces the value of the cell 2j = 0 by the inlet equilibrium, the
// outlet (x = extent - 1) gives cell 15 its own population of the same direction, which the end lane
// loads in place of the x-neighbour word.
//
// Each cell goes through the same arithmetic as the one-cell kernels (decode, collision, rounding with
// the cell's own random bits), so the results are the same bit for bit.
#pragma once

namespace oal {

constexpr int PWG = 128;   // lanes of a pair workgroup: 8 pairs x 16 rows

// Addresses: buffer loads and stores on one buffer resource per workgroup, based `bias` bytes below the
// own block's start (one layer of blocks, one row of blocks and one block), so that every offset is
// non-negative and no lane forms a 64-bit address.  v[j] is the per-lane offset of the aligned source pair
// in row y - cy (cy = j - 1), which may lie in a y-neighbour block (b -+ nbx); the uniform part (scalar
// offset) is the layer of blocks, the direction slot and the plane.
struct PairAddr {
    uint32_t v[3];
    __amdgpu_buffer_rsrc_t res;
    uint32_t plane;   // bytes between consecutive block layers
    __device__ __forceinline__ void init(const void* f, uint32_t b, int jx, int ly, int nbx, int nby) {
        const int64_t blk = 27 * BS3 * 2;
        plane = (uint32_t)((int64_t)nbx * nby * blk);
        const int64_t base = (int64_t)b * blk - (int64_t)(nbx + 1) * blk - (int64_t)plane;
        res = __builtin_amdgcn_make_buffer_rsrc((char*)f + base, (short)0, (int)0xffffffff, 0x00020000);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int ys = ly - (j - 1);
            const int dy = ys < 0 ? -1 : (ys >= BS ? 1 : 0);
            // dy nbx blocks from the own block, plus the bias of nbx + 1 blocks
            v[j] = (uint32_t)((dy + 1) * nbx + 1) * (uint32_t)blk + (uint32_t)(((ys - BS * dy) * BS + 2 * jx) * 2);
        }
    }
    // uniform offset of direction slot `slot` at plane z, for sources shifted by -cz in z
    __device__ __forceinline__ uint32_t soff(int slot, int z, int cz) const {
        const int sz = z - cz;
        const int dz = sz < 0 ? -1 : (sz >= BS ? 1 : 0);
        return (uint32_t)(dz + 1) * plane + (uint32_t)(slot * (BS3 * 2) + (sz - BS * dz) * (BS2 * 2));
    }
    __device__ __forceinline__ uint32_t load(uint32_t voff, uint32_t so) const {
        return __builtin_amdgcn_raw_buffer_load_b32(res, (int)voff, (int)so, 0);
    }
};

// the x-neighbour block's cell adjoining the row: (-27 BS3 + 14) elements from pair 0 (the -x block's
// pair 14, 15), (27 BS3 - 14) from pair 7 (the +x block's pair 0, 1)
constexpr int32_t PAIR_EDGE_M = (-27 * BS3 + 14) * 2;
constexpr int32_t PAIR_EDGE_P = (27 * BS3 - 14) * 2;

// the loads of one plane: the aligned pair of every direction, and the x-neighbour word of the
// directions with cx != 0 on the row's end lanes (index: 0..8 for cx = +1 in direction order, 9..17 cx = -1)
struct PairRaw {
    uint32_t p[27];
    uint32_t e[18];
};

template <int K>
constexpr int edge_slot() {
    // directions with cx = +1 are K % 3 == 2, cx = -1 are K % 3 == 0
    return dcx(K) > 0 ? K / 3 : 9 + K / 3;
}

// the domain faces a block touches and a lane's part in them (FACE_RULES)
struct PairFaces {
    bool yl, yh;            // per lane: the row lies on the y-min / y-max face
    bool zlo, zhi;          // the block touches the z-min / z-max face
    bool xlo, xhi;          // the block touches the inlet (x-min) / outlet (x-max) face
    bool zwall, ymin_wall;  // no-slip variants of the z faces and of the y-min face
};

// Is direction K of this lane's cells (those not on an x face) given by a y / z face rule at plane z, and
// which own population does the rule take (precedence: no-slip z wall, no-slip y-min wall, y mirror,
// z mirror)?
template <int K>
__device__ __forceinline__ bool face_rule(const PairFaces& fc, int z, int& j) {
    constexpr int cy = dcy(K), cz = dcz(K);
    const bool zf = (cz > 0 && fc.zlo && z == 0) || (cz < 0 && fc.zhi && z == BS - 1);   // uniform
    const bool yf = (cy > 0 && fc.yl) || (cy < 0 && fc.yh);
    j = (fc.zwall && zf) ? dopp(K) : (fc.ymin_wall && cy > 0 && fc.yl) ? dopp(K) : yf ? dmir_y(K) : dmir_z(K);
    return zf || yf;
}

template <int FACE>
__device__ __forceinline__ void pair_load(const PairAddr& pa, const PairFaces& fc, int jx, int z, PairRaw& r) {
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int cx = dcx(K), cy = dcy(K), cz = dcz(K);
        const uint32_t so = pa.soff(K, z, cz);
        if constexpr (FACE == FACE_NONE) {
            r.p[K] = pa.load(pa.v[cy + 1], so);
            if constexpr (cx > 0) {
                if (jx == 0) r.e[edge_slot<K>()] = pa.load(pa.v[cy + 1] + (uint32_t)PAIR_EDGE_M, so);
            } else if constexpr (cx < 0) {
                if (jx == 7) r.e[edge_slot<K>()] = pa.load(pa.v[cy + 1] + (uint32_t)PAIR_EDGE_P, so);
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
            r.p[K] = pa.load(voff, 0);
            if constexpr (cx > 0) {
                // (the inlet's cell 0 is replaced after decoding: no x-neighbour word there)
                if (jx == 0 && !fc.xlo && !face) r.e[edge_slot<K>()] = pa.load(pa.v[cy + 1] + (uint32_t)PAIR_EDGE_M + so, 0);
            } else if constexpr (cx < 0) {
                if (jx == 7) {
                    if (fc.xhi)   // outlet: cell 15's own population of direction K (the own pair 14, 15)
                        r.e[edge_slot<K>()] = pa.load(pa.v[1] + pa.soff(K, z, 0), 0);
                    else if (!face)
                        r.e[edge_slot<K>()] = pa.load(pa.v[cy + 1] + (uint32_t)PAIR_EDGE_P + so, 0);
                }
            }
        }
    });
}

// the word of the two sources of (cell 2j, cell 2j + 1) for every direction: low half cell 2j's
template <int FACE>
__device__ __forceinline__ void pair_sources(const PairRaw& r, const PairFaces& fc, int jx, int z, uint32_t (&s)[27]) {
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int cx = dcx(K), cy = dcy(K), cz = dcz(K);
        uint32_t v;
        if constexpr (cx == 0) {
            v = r.p[K];
        } else if constexpr (cx > 0) {
            // (2j - 1, 2j): the left lane's upper cell, the own lower cell
            const uint32_t left = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r.p[K], 0x111, 0xf, 0xf, false);
            const uint32_t l = jx == 0 ? r.e[edge_slot<K>()] : left;
            v = __builtin_amdgcn_alignbyte(r.p[K], l, 2);
        } else {
            // (2j + 1, 2j + 2): the own upper cell, the right lane's lower cell
            const uint32_t right = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r.p[K], 0x101, 0xf, 0xf, false);
            const uint32_t rr = jx == 7 ? r.e[edge_slot<K>()] : right;
            v = __builtin_amdgcn_alignbyte(rr, r.p[K], 2);
        }
        if constexpr (FACE == FACE_RULES) {
            if constexpr (cx != 0 && (cy != 0 || cz != 0)) {
                int j;
                if (face_rule<K>(fc, z, j)) v = r.p[K];   // the own pair of the rule, unshifted
            }
            if constexpr (cx < 0) {
                // outlet: cell 15 keeps its own population (the upper half of the own pair 14, 15)
                if (fc.xhi && jx == 7) v = (v & 0xffffu) | (r.e[edge_slot<K>()] & 0xffff0000u);
            }
        }
        s[K] = v;
    });
}

// the inlet equilibrium that replaces the arriving populations with cx = +1 of a cell at x = 0
struct PairInlet {
    bool on;       // this lane's cell 2j lies on the inlet face
    float u_in;
    int gy, gz;    // the cell's global y and z (the noise position)
};

// collide the pair and leave its post-collision populations in s (cell 2j in the low half)
template <class SP, bool UNIT>
__device__ __forceinline__ void pair_collide(const StepArgs& a, const CellRates<float>& q, uint32_t hca, uint32_t hcb,
                                             const PairInlet* inlet, uint32_t (&s)[27]) {
    const bool dith = a.cp.dither;
    float rho_o, ux_o, uy_o, uz_o;
    float f[27];
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        f[K] = SP::dec_lo(s[K], wgt<float>(K));
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
    // the rounding mode is uniform: one branch for the 27 stores rather than one per direction
    if (dith) {
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            s[K] = SP::template put<false>(s[K], f[K], wgt<float>(K), round_bits(hca, K), true);
        });
    } else {
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            s[K] = SP::template put<false>(s[K], f[K], wgt<float>(K), 0u, false);
        });
    }
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        f[K] = SP::dec_hi(s[K], wgt<float>(K));
    });
    collide_cell<float, UNIT>(f, q, a.cp, rho_o, ux_o, uy_o, uz_o);
    if (dith) {
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            s[K] = SP::template put<true>(s[K], f[K], wgt<float>(K), round_bits(hcb, K), true);
        });
    } else {
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            s[K] = SP::template put<true>(s[K], f[K], wgt<float>(K), 0u, false);
        });
    }
}

// register budget: 3 waves per SIMD (168 VGPRs) where the kernel fits it without spilling, 2 otherwise
template <class SP, bool UNIT, int FACE>
constexpr int pair_waves() { return FACE == FACE_NONE && (UNIT || SP::code == StBF16S::code) ? 3 : 2; }

#ifndef OAL_PAIR_STORE_AUX
#define OAL_PAIR_STORE_AUX 2
#endif

// Regular blocks (no solid near, whole; FACE_RULES: on domain faces too), AB pattern, 16-bit storage: a
// workgroup of 128 lanes walks `a.pipe_planes` planes of one block with the next plane's loads in flight
// during the current plane's collision (the schedule of k_step_pipe).
template <class SP, int NP, bool UNIT, int FACE>
__global__ void __launch_bounds__(PWG)
__attribute__((amdgpu_waves_per_eu(pair_waves<SP, UNIT, FACE>(), pair_waves<SP, UNIT, FACE>())))
k_step_pair(const StepArgs a) {
    const int np = a.pipe_planes;   // run-time bound equal to NP (see k_step_pipe)
    constexpr uint32_t groups = (uint32_t)(PLANES / NP);
    const PipeItem it = pipe_item<true>(groups, np);
    const uint32_t b = (uint32_t)a.blist[it.pos];
    const int z0 = it.z0;
    const int jx = (int)(threadIdx.x & 7u), ly = (int)(threadIdx.x >> 3);
    const DevDomain* dm = a.doms + a.bdom[b];
    PairAddr pa, po;   // the populations read and those written
    pa.init(a.fin, b, jx, ly, dm->nbx, dm->nby);
    po.init(a.fout, b, jx, ly, dm->nbx, dm->nby);
    CellRates<float> q;
    q.omega1 = dm->omega1;
    q.omega4_diag = dm->omega4d;
    q.forced = dm->forced != 0;
    q.Fx = dm->F[0]; q.Fy = dm->F[1]; q.Fz = dm->F[2];
    const uint32_t cell0 = dm->cell0;   // read before the first store, as every domain parameter
    const uint32_t key_step = round_key(a.step());
    const uint32_t own0 = b * BS3 + (uint32_t)ly * BS + 2u * (uint32_t)jx;   // cell 2j of plane 0
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
    PairRaw raw{};   // (the neighbour words are loaded on the row's end lanes only)
    pair_load<FACE>(pa, fc, jx, z0, raw);
    for (int z = z0; z < z0 + np; ++z) {
        wait_all_vmem();
        uint32_t s[27];
        pair_sources<FACE>(raw, fc, jx, z, s);
        __builtin_amdgcn_sched_barrier(0);
        if (z + 1 < z0 + np) pair_load<FACE>(pa, fc, jx, z + 1, raw);
        const uint32_t ca = own0 + (uint32_t)z * BS2 - cell0;
        const uint32_t hca = SP::dithered ? round_cell(key_step, ca) : 0u;
        const uint32_t hcb = SP::dithered ? round_cell(key_step, ca + 1u) : 0u;
        if constexpr (FACE == FACE_RULES) {
            PairInlet iz = in;
            iz.gz += z;
            pair_collide<SP, UNIT>(a, q, hca, hcb, &iz, s);
        } else {
            pair_collide<SP, UNIT>(a, q, hca, hcb, nullptr, s);
        }
        const uint32_t so = po.soff(0, z, 0);
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            // (aux 2: the non-temporal hint, as the other step kernels' stores)
            __builtin_amdgcn_raw_buffer_store_b32(s[K], po.res, (int)po.v[1], (int)(so + K * (BS3 * 2)), OAL_PAIR_STORE_AUX);
        });
    }
}

template <class SP, bool UNIT, int FACE>
static void launch_pair_u(int np, dim3 grid, hipStream_t s, const StepArgs& r) {
    switch (np) {
        case 2: hipLaunchKernelGGL((k_step_pair<SP, 2, UNIT, FACE>), grid, dim3(PWG), 0, s, r); break;
        case 4: hipLaunchKernelGGL((k_step_pair<SP, 4, UNIT, FACE>), grid, dim3(PWG), 0, s, r); break;
        case 8: hipLaunchKernelGGL((k_step_pair<SP, 8, UNIT, FACE>), grid, dim3(PWG), 0, s, r); break;
        default: hipLaunchKernelGGL((k_step_pair<SP, 16, UNIT, FACE>), grid, dim3(PWG), 0, s, r); break;
    }
}
template <class SP>
static void launch_pair(int np, bool face, dim3 grid, hipStream_t s, const StepArgs& r) {
    if (r.cp.unit) {
        if (face) launch_pair_u<SP, true, FACE_RULES>(np, grid, s, r);
        else launch_pair_u<SP, true, FACE_NONE>(np, grid, s, r);
    } else {
        if (face) launch_pair_u<SP, false, FACE_RULES>(np, grid, s, r);
        else launch_pair_u<SP, false, FACE_NONE>(np, grid, s, r);
    }
}

}  // namespace oal
