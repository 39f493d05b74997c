// Kernel argument structs and launchers shared by the plan code and the kernels of liboal.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oal_device.h"

namespace oal {

constexpr int BS = 16;            // cells per block edge
constexpr int BS2 = BS * BS;
constexpr int BS3 = BS * BS * BS;
constexpr int WG = 256;           // threads per workgroup: one 16 x 16 plane of a block
constexpr int PLANES = BS;        // workgroups per block in the step kernels

// per-block flags, computed on the host when the plan is built
enum : uint32_t {
    BF_EDGE = 1u,      // a source of some cell may lie outside the domain extent (domain-face rules)
    BF_SOLID = 2u,     // the block or its one-cell halo holds a solid cell
    BF_MISSING = 4u,   // some neighbour block is absent inside the extent (rest-weight fallback)
    BF_PARTIAL = 8u,   // some cells of the block lie beyond the extent
    BF_INACTIVE = 16u, // refined level: a covered block (written by restriction) or a ghost block (written by
                       // the fill from the coarser level); never stepped
    BF_SPONGE = 32u    // some cell of the block has a nonzero sponge strength
};

// one domain of a level on the device
struct DevDomain {
    int32_t ext[3];
    int32_t per[3];
    float tau, u_in, amr_scale, force_scale;
    float F[3];
    int32_t ymin_wall, zwall, forced;
    float omega1, omega4d;          // collision rates without WALE, FP32 forms
    double omega1_d, omega4d_d;     // and FP64 forms
    int32_t nbx, nby;               // blocks per row / per layer of the domain's lexicographic grid
    uint32_t cell0;                 // first cell of the domain: cells are numbered from it for the
                                    // rounding hash, so a domain's results do not depend on its place
    int32_t pad;
};

// arguments of the fused stream-collide kernels
struct StepArgs {
    const void* fin;             // populations read this step
    void* fout;                  // populations written (== fin for the AA pattern)
    const int32_t* nbr_bm;       // [32 * nb] block-major neighbour rows, 0-based, -1 = none
    const int32_t* borg;         // [3 * nb] block origin in cells
    const int32_t* bdom;         // [nb] domain index
    const uint32_t* bflags;      // [nb]
    const uint8_t* solid;        // [ncell] or null
    const DevDomain* doms;
    const void* vel_in;          // [3 * ncell] WALE velocity of the previous step (T)
    void* vel_out;               // [3 * ncell] velocity written (WALE, macro output)
    void* rho_out;               // [ncell] density written (macro output) or null
    double* fpart;               // [n_items * 4] per-workgroup force partials or null
    const int32_t* blist;        // blocks of this launch (item / PLANES -> block), null = all in order
    const int32_t* blist_regular;    // [n_regular] interior blocks whose neighbours sit at lexicographic offsets
    const int32_t* blist_interior;   // [n_interior] other blocks with all flags clear
    const int32_t* blist_boundary_regular;  // [n_boundary_regular] boundary blocks with lexicographic neighbours
    const int32_t* blist_boundary;   // [n_boundary] the other blocks
    uint32_t n_regular, n_interior, n_boundary_regular, n_boundary;
    uint32_t n_boundary_regular_solid;   // blocks at the head of blist_boundary_regular whose only flag is BF_SOLID
    const int32_t* blist_partial;    // [n_partial] AB pattern: regular blocks that overhang the extent, no solid near
    uint32_t n_partial;
    const int32_t* blist_sponge;     // [n_sponge] AB pattern: regular blocks with sponge cells, no solid near
    uint32_t n_sponge;
    // the regular list split (AB pattern): blocks on no domain face, and whole blocks on a domain face
    const int32_t* blist_regular_plain;
    uint32_t n_regular_plain;
    const int32_t* blist_regular_face;
    uint32_t n_regular_face;
    const int32_t* lslot;        // [nb] link-mask slot of a block with solid cells near, -1 = none
    const uint32_t* lmask;       // [n_lslots * BS3] per cell: bit k source of direction k solid, bit 27 cell solid
    const uint32_t* face_cells;  // [n_face] AB pattern: cells of regular-list blocks with a source outside the extent
    uint32_t n_face;
    int32_t regular_face;        // AB pattern: the blocks of this launch may lie on a domain face (set per launch)
    int32_t face_kernel;         // 1: the pipelined kernel skips face cells and the face-cell kernel steps
                                 // them; 0: it applies the face rules itself
    const int32_t* oslot;        // [nb] AA pattern: outlet side-buffer slot of the block, -1 = none
    void* obuf;                  // [n_oslots * 9 * BS2] AA pattern: outlet-face populations with cx = -1
    uint32_t ncell;              // cells of the level (the stride of the velocity fields' components)
    uint32_t nb;                 // blocks
    uint32_t n_items;            // workgroups launched (nb * PLANES)
    int32_t time_step;
    // a step replayed from a captured graph reads its index from device memory: ts_dev[0] * ts_mul + ts_add
    // (the coarse step, 2^level, the sub-step within it); null: time_step
    const int32_t* ts_dev;
    int32_t ts_mul, ts_add;
    int32_t parity;              // AA pattern: 1 odd step (neighbour slots), 0 even step (own slots)
    int32_t pipe_planes;         // planes per workgroup of the pipelined interior kernel (1 = off; divides 16)
    int32_t face_x;              // a domain of the level has the exact specular z wall, or the transverse freestream is on
    int32_t pair_kernel;         // FP32 / 16-bit storage: regular blocks two cells per lane (pipelined shapes)
    // interpolated bounce-back: per link-mask slot, [27][BS3] Float16 bits of the wall distance q of the
    // link from a cell along c_k (units of |c_k|); null = halfway bounce-back on every solid link
    const uint16_t* qmap;
    float q_min;                 // links with q <= q_min (or q > 1) keep halfway bounce-back
    // sponge (blocks flagged BF_SPONGE): per sponge slot [BS3] strengths; slot of a block; 1 = the
    // populations are blended too
    const float* sponge;
    const int32_t* sslot;
    int32_t sponge_blend;
    int32_t bfl_fallback;        // q < 1/2 without a fluid cell behind: 0 halfway, 1 wall equilibrium, 2 own opposite
    CollParams cp;
    __device__ __forceinline__ int32_t step() const { return ts_dev ? ts_dev[0] * ts_mul + ts_add : time_step; }
};

// Whether a domain's block grid fits the 32-bit offsets of the two-cells-per-lane kernels (populations
// of `esize` bytes, one buffer resource per block): two layers of blocks, two rows and three blocks
// below 4 GiB
inline bool pair_fits(int nbx, int nby, int esize = 2) {
    const uint64_t blk = 27ull * BS3 * (uint64_t)esize;
    return 2ull * (uint64_t)nbx * (uint64_t)nby * blk + (2ull * (uint64_t)nbx + 3ull) * blk < (1ull << 32);
}

struct ReduceArgs {
    const double* part;          // [n_items * 4]
    const int32_t* dom_item0;    // [n_domains + 1] first item of each domain
    double* out;                 // [n_domains * 4]
    int32_t n_domains;
    int32_t width;               // values per item used (3 force, 4 mass)
};

// launchers (oal_kernels.hip); all asynchronous on `stream`
// When bulk0/bulk1 are given they are recorded on `stream` around the bulk launches (regular and
// interior block lists), the kernels that carry most of the cells.
hipError_t launch_step(int storage, int stream_mode, bool wale, const StepArgs& a, hipStream_t stream, hipStream_t side,
                       hipEvent_t bulk0 = nullptr, hipEvent_t bulk1 = nullptr);
// values per gathered donor record of the level transfers: 27 populations padded to 28 (16-byte rows)
constexpr int XFER_STRIDE = 28;

// p[0] = v, in stream order (the step index of a replayed step graph)
hipError_t launch_set_i32(int32_t* p, int32_t v, hipStream_t stream);
// per-storage step launchers (oal_step_<storage>.hip)
hipError_t launch_step_f32(int stream_mode, bool wale, const StepArgs& a, hipStream_t stream, hipStream_t side,
                           hipEvent_t bulk0, hipEvent_t bulk1);
hipError_t launch_step_f64(int stream_mode, bool wale, const StepArgs& a, hipStream_t stream, hipStream_t side,
                           hipEvent_t bulk0, hipEvent_t bulk1);
hipError_t launch_step_f16s(int stream_mode, bool wale, const StepArgs& a, hipStream_t stream, hipStream_t side,
                            hipEvent_t bulk0, hipEvent_t bulk1);
hipError_t launch_step_bf16s(int stream_mode, bool wale, const StepArgs& a, hipStream_t stream, hipStream_t side,
                             hipEvent_t bulk0, hipEvent_t bulk1);
hipError_t launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t stream);
// Host arrays of init / conversion / macro output cover the first nhost cells of a level (the caller's
// blocks); the device arrays have ncell cells per direction.
hipError_t launch_init_eq(int storage, void* f, uint32_t ncell, uint32_t nhost, const double* rho, const double* u,
                          double rho0, double ux0, double uy0, double uz0, int permute, void* vel, hipStream_t stream);
hipError_t launch_convert_in(int storage, void* f, const void* src, int src_dtype, uint32_t ncell, uint32_t nhost,
                             int permute, hipStream_t stream);
hipError_t launch_convert_out(int storage, const void* f, void* dst, int dst_dtype, uint32_t ncell, uint32_t nhost,
                              int permute, hipStream_t stream);
hipError_t launch_macro(int storage, const void* f, double* rho, double* u, uint32_t ncell, uint32_t nhost, int permute,
                        void* vel, hipStream_t stream);
// Level transfers of a refined plan.  gather: the working-precision populations of each listed cell
// (bit 31 set = not blended in time) of fnew, or of (1 - w1) fold + w1 fnew when fold is given, in
// records of XFER_STRIDE values; for the 16-bit forms without fold, records of the stored codes
// (raw, decoded by the rebuild, which must then be told `raw`).  rebuild: for each target cell, the weighted sum of the gathered populations over its
// 64 stencil nodes (slot[64 * i + node], -1 = unused; weight w[...]: cell-major), its second-order
// non-equilibrium moments set to scale * keep times their value, written to f; with vel (WALE), the
// velocity of the interpolated state written to vel ([3][ncell], working precision) too.
hipError_t launch_xfer_gather(int storage, const void* fnew, const void* fold, const uint32_t* cells, uint32_t n,
                               uint32_t ncell, double w1, void* out, hipStream_t stream);
hipError_t launch_xfer_rebuild(int storage, bool raw, void* f, uint32_t ncell, const uint32_t* tcells, uint32_t n,
                               const int32_t* slot, const float* w, const void* m, double scale, double keep, void* vel,
                               hipStream_t stream);
hipError_t launch_aa_outlet_prep(int storage, const StepArgs& a, hipStream_t stream);
hipError_t launch_mass(int storage, const void* f, const uint8_t* solid, const int32_t* borg, const int32_t* bdom,
                       const uint32_t* bflags, const DevDomain* doms, uint32_t nb, int permute, double* part, hipStream_t stream);
hipError_t launch_reduce(const ReduceArgs& r, hipStream_t stream);
hipError_t launch_guard(int storage, const void* f, const uint8_t* solid, const int32_t* borg, const int32_t* bdom,
                        const uint32_t* bflags, const DevDomain* doms, uint32_t nb, float rho_min, float u_max,
                        unsigned long long* count, hipStream_t stream);

// Surface-stress map and its integration over a triangle mesh (oal_surface.hip)
constexpr int SURF_WG = 256;      // triangles per workgroup; one FP64 partial record per workgroup
constexpr int SURF_SUMS = 10;     // Fp[3], Fv[3], M[3], mapped count
struct SurfaceArgs {
    const void* f;               // populations in stored form
    int permute;                 // AA pattern: directions stored reversed
    const uint8_t* solid;        // [ncell] or null
    const int32_t* bmap;         // [nbx * nby * nbz] block of each block position of the domain, -1 = absent
    int32_t nbx, nby, nbz;
    int32_t ext[3];              // domain extent in cells
    int64_t n_tri;
    const float* cen;            // [3 * n_tri] centres
    const float* nrm;            // [3 * n_tri] normals
    const float* area;           // [n_tri]
    float dx, off[3], pressure_scale, stress_scale, tau, c_wale, nu_sgs_bg, cs2;
    double ref[3];               // moment centre
    int32_t search_radius, recover_normal;
    float* p_map;                // [n_tri]
    float* s_map;                // [3 * n_tri]
    float* sign_map;             // [n_tri]
    float* wd_map;               // [n_tri]
    double* part;                // [SURF_SUMS * n_groups]
    double* out;                 // [SURF_SUMS]
};
hipError_t launch_surface(int storage, const SurfaceArgs& a, hipStream_t stream);

}  // namespace oal
