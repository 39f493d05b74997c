// liboal host side: the C ABI of include/oal.h.  A plan owns the device state of a solver run
// (every level's block tables, populations, reductions) and launches the kernels of
// oal_kernels.hip on its stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <algorithm>
#include <array>
#include <numeric>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/oal.h"
#include "oal_kernels.h"

using namespace oal;

// the scratch array of oal_mass_noalloc (see scratch_init): the per-block partials of one level
static __device__ __align__(256) double scratch[(size_t)OAL_NOALLOC_MAX_BLOCKS * 4];

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define OAL_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return fail(OAL_ERR_HIP, std::string(#call " : ") + hipGetErrorString(e_)); \
    } while (0)

// Scratch of oal_mass_noalloc: the per-block FP64 partials (mass, jx, jy, jz) of one level, in a
// file-scope device array so that the entry point allocates nothing.  Every device holds its own copy
// at its own address, which hipGetSymbolAddress gives once per device when a plan is created there
// (scratch_addr); host code passes that address to kernels, never the array's name.
constexpr uint32_t NOALLOC_MAX_BLOCKS = OAL_NOALLOC_MAX_BLOCKS;

std::mutex g_scratch_mu;
std::unordered_map<int, double*> g_scratch_addr;   // device ordinal -> address of `scratch` there

// call with `device` current
int scratch_init(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr.count(device)) return OAL_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail(OAL_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[device] = (double*)ptr;
    return OAL_OK;
}

double* scratch_addr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    return it == g_scratch_addr.end() ? nullptr : it->second;
}

size_t storage_bytes(int storage) { return storage == OAL_F64 ? 8 : storage == OAL_F32 ? 4 : 2; }
size_t work_bytes(int storage) { return storage == OAL_F64 ? 8 : 4; }

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        n = count;
        return count ? hipMalloc((void**)&p, count * sizeof(T)) : hipSuccess;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct Level {
    uint32_t nb = 0, ncell = 0, n_items = 0;
    int n_domains = 0;
    std::vector<DevDomain> doms_h;
    std::vector<int32_t> dom_item0_h, dom_block0_h;
    DevBuf<int32_t> nbr, borg, bdom, dom_item0, dom_block0;
    DevBuf<uint32_t> bflags;
    DevBuf<uint8_t> solid;
    DevBuf<DevDomain> doms;
    DevBuf<unsigned char> fA, fB;    // populations (AB: read A write B then swap; AA: A only)
    DevBuf<unsigned char> velA, velB;  // WALE velocity, working precision
    DevBuf<double> fpart, mpart, red_out;
    DevBuf<int32_t> oslot;             // AA: outlet side-buffer slot per block
    DevBuf<int32_t> blist_rg, blist_in, blist_br, blist_bd;  // block lists per kernel kind
    uint32_t n_regular = 0, n_interior = 0, n_boundary_regular = 0, n_boundary = 0;
    uint32_t n_br_solid = 0;   // blocks at the head of the regular boundary list whose only flag is BF_SOLID
    DevBuf<int32_t> blist_pr;          // regular blocks that overhang the extent (pipelined, AB)
    DevBuf<int32_t> blist_sp;          // regular blocks with sponge cells, no solid near (pipelined, AB)
    uint32_t n_sponge = 0;
    uint32_t n_partial = 0;
    // the regular blocks split in two (AB), for the kernels that step the ones on a domain face apart
    DevBuf<int32_t> blist_rp, blist_rf;   // plain (no face) and on a domain face, each in lexicographic order
    uint32_t n_regular_face = 0;          // (n_regular counts them too)
    DevBuf<uint32_t> face_cells;       // AB: face cells of the regular-list blocks
    uint32_t n_face = 0;
    bool regular_face = false;         // AB: the regular list holds blocks on a domain face
    DevBuf<int32_t> lslot;             // link-mask slot per block
    DevBuf<uint32_t> lmask;            // per-cell solid-link masks of blocks near solid cells
    std::vector<int32_t> lslot_h;      // host copy of lslot
    std::vector<int32_t> map_h;        // [3 * nb] 1-based block coordinates (caller's blocks first)
    // a level the caller gave in 8^3 blocks: caller cell (b8 * 512 + local) -> library cell, and the
    // caller's 1-based 8^3 block coordinates and obstacle (empty for 16^3 levels)
    std::vector<uint32_t> c8map;
    std::vector<int32_t> map8_h;
    std::vector<uint8_t> obst8_h;
    std::vector<uint8_t> obst_h;       // [nb * BS3] obstacle of the caller's layout, empty = none
    int32_t n_lslots = 0;
    // interpolated bounce-back: per link-mask slot [27][BS3] Float16 q codes (empty: halfway only)
    DevBuf<uint16_t> qmap;
    int32_t bfl_fallback = 0;
    float q_min = 0.001f;
    DevBuf<unsigned char> obuf;        // AA: outlet-face populations with cx = -1
    DevBuf<unsigned long long> guard;
    // refined plans: the caller's blocks come first (ncell_host cells), the level's ghost blocks after them
    uint32_t ncell_host = 0;
    uint32_t n_caller_blocks = 0, n_ghost_blocks = 0, n_covered_blocks = 0;
    // fill of this level's ghost cells from the coarser level: moments of g_src cells (coarser level),
    // stencils (g_slot, g_w: [n_gdst][64], cell-major) onto g_dst cells (this level)
    DevBuf<uint32_t> g_src, g_dst;
    DevBuf<int32_t> g_slot;
    DevBuf<float> g_w;
    DevBuf<unsigned char> g_m;
    uint32_t n_gsrc = 0, n_gdst = 0;
    double fill_scale = 0.0, fill_keep = 0.0;
    // restriction of the finer level into this level's covered cells
    DevBuf<uint32_t> r_src, r_dst;
    DevBuf<int32_t> r_slot;
    DevBuf<float> r_w;
    DevBuf<unsigned char> r_m;
    uint32_t n_rsrc = 0, n_rdst = 0;
    // r_*: the covered cells the stepping reads (next to stepped cells, fill donors, donors of the
    // coarser level's restriction); rf_*: every covered fluid cell (before the state is read out)
    DevBuf<uint32_t> rf_src, rf_dst;
    DevBuf<int32_t> rf_slot;
    DevBuf<float> rf_w;
    uint32_t n_rfsrc = 0, n_rfdst = 0;
    double restrict_scale = 0.0, restrict_keep = 0.0;
    bool has_solid = false;
    int64_t n_fluid = 0;
    // host tables the block lists are built from (build_lists)
    std::vector<uint32_t> flags_h;
    std::vector<char> regular_h, near_h;
    std::vector<int32_t> borg_h, bdom_h;
    bool ab = true, pipe_ok = true;
    bool tfs = false;                // transverse freestream on the y and z faces
    // sponge (kernel_cumulant.jl:428-446): per sponge slot [BS3] strengths of the blocks that hold a
    // nonzero one (empty: no sponge)
    DevBuf<float> sponge;
    DevBuf<int32_t> sslot;
    int32_t sponge_blend = 0;
    bool has_sponge = false;
    // surface mesh of oal_set_surface (centres and normals [3][n_tri], areas), the block of each block
    // position of the domain, and the per-triangle maps and FP64 records of the last oal_surface_forces
    int64_t n_tri = 0;
    bool surface_mapped = false;
    DevBuf<float> s_cen, s_nrm, s_area, s_p, s_shear, s_sign, s_wd;
    DevBuf<int32_t> s_bmap;
    DevBuf<double> s_part, s_out;
    int32_t s_nblk[3] = {0, 0, 0};
    void release_surface() {
        s_cen.release(); s_nrm.release(); s_area.release(); s_p.release(); s_shear.release(); s_sign.release();
        s_wd.release(); s_bmap.release(); s_part.release(); s_out.release();
        n_tri = 0; surface_mapped = false;
    }
    bool have_f = false;
    bool force_valid = false;
    int aa_phase = 0;                // AA: 0 = populations in stored form (next step odd), 1 = after an odd step
    void release() {
        nbr.release(); borg.release(); bdom.release(); dom_item0.release(); dom_block0.release();
        bflags.release(); solid.release(); doms.release(); fA.release(); fB.release(); velA.release(); velB.release();
        fpart.release(); mpart.release(); red_out.release(); guard.release(); oslot.release(); obuf.release(); blist_rg.release(); blist_in.release(); blist_br.release(); blist_bd.release(); blist_pr.release(); blist_sp.release(); blist_rp.release(); blist_rf.release(); face_cells.release();
        lslot.release(); lmask.release(); qmap.release(); sponge.release(); sslot.release();
        g_src.release(); g_dst.release(); g_slot.release(); g_w.release(); g_m.release();
        r_src.release(); r_dst.release(); r_slot.release(); r_w.release(); r_m.release();
        rf_src.release(); rf_dst.release(); rf_slot.release(); rf_w.release();
        release_surface();
    }
};

}  // namespace

struct OalPlan {
    int device = 0;
    int pipe_planes = 4;   // planes per workgroup of the pipelined interior kernel
    bool overlap = true;   // boundary and face launches on the side stream
    bool face_kernel = false;   // face cells by the face-cell kernel instead of rules in the pipelined one
    int pair_kernel = 2;        // FP32 / 16-bit storage: regular blocks two cells per lane (1 one launch, 2 face blocks apart)
    int storage = OAL_F32;
    int stream_mode = OAL_STREAM_AB;
    int bb_mode = OAL_BB_HALFWAY;
    bool wale = false;
    CollParams cp{};
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;             // boundary launches, overlapping the interior ones
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<hipEvent_t> bulk_ev;        // pairs around the bulk launches of timed steps
    bool multilevel = false;                // a refined plan: levels coupled by ghost fill and restriction
    bool coupled = false;                   // the initial restriction and fill have run since the last upload
    bool covered_stale = false;             // steps restricted only the covered cells the stepping reads
    std::vector<Level> levels;
    // captured steps: oal_step replays a hipGraph of one (coarse) step; the step index comes from ts_dev
    bool graph = false;
    bool capturing = false;
    int32_t* ts_dev = nullptr;
    struct LevelState { const void* fA; const void* velA; int aa_phase; };
    struct StepGraph { hipGraphExec_t exec = nullptr; std::vector<LevelState> before, after; };
    std::vector<StepGraph> graphs;          // one per buffer arrangement met so far
};

namespace {

hipStream_t pick_stream(OalPlan* p, void* s) { return s ? (hipStream_t)s : p->stream; }

int check_level(const OalPlan* p, int level) {
    if (!p) return fail(OAL_ERR_ARG, "plan is NULL");
    if (level < 0 || level >= (int)p->levels.size()) return fail(OAL_ERR_ARG, "level index out of range");
    return OAL_OK;
}

template <class T>
int upload(DevBuf<T>& d, const std::vector<T>& h) {
    OAL_HIP(d.alloc(h.size()));
    if (!h.empty()) OAL_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return OAL_OK;
}


// Host tables of one level, built and validated before any device call
struct HostLevel {
    std::vector<int32_t> borg, bdom, nbr, nbr_bm, oslot, blist_rg, blist_in, blist_br, blist_bd, blist_pr, blist_sp, blist_rp, blist_rf,
        lslot;
    std::vector<uint32_t> lmask;
    std::vector<uint32_t> face_cells;
    int32_t n_oslots = 0;
    std::vector<uint32_t> flags;
    std::vector<uint8_t> solid;
};

// The block lists of a level's step kernels from its block flags (L.flags_h) and the host tables kept
// for that purpose; rebuilt when the flags change (a sponge field marks its blocks).
void build_lists(Level& L, HostLevel& H) {
    H.blist_rg.clear(); H.blist_rp.clear(); H.blist_rf.clear(); H.blist_in.clear();
    H.blist_br.clear(); H.blist_bd.clear(); H.blist_pr.clear(); H.blist_sp.clear(); H.face_cells.clear();
    L.regular_face = false;
    const std::vector<uint32_t>& flags = L.flags_h;
    // blocks on a face whose rule the pipelined kernels do not apply (the exact specular z wall reads
    // another cell of the wall row; the transverse freestream blends): the general kernels step them
    auto special_face = [&](uint32_t b, const DevDomain& dm) {
        bool lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            const int o = L.borg_h[3 * b + a];
            lo[a] = !dm.per[a] && o == 0;
            hi[a] = !dm.per[a] && o + BS >= dm.ext[a];
        }
        const bool zface = lo[2] || hi[2], yface = lo[1] || hi[1];
        return (dm.zwall == 1 && zface) || (L.tfs && (zface || yface));
    };
    for (uint32_t b = 0; b < L.nb; ++b) {
        const DevDomain& dm = L.doms_h[L.bdom_h[b]];
        if (flags[b] & BF_INACTIVE) continue;   // never stepped
        const bool regular = L.regular_h[b] != 0;
        const bool special = (flags[b] & BF_EDGE) && special_face(b, dm);
        if (flags[b] == 0) {
            (regular ? H.blist_rg : H.blist_in).push_back((int32_t)b);
            if (regular) H.blist_rp.push_back((int32_t)b);
        } else if (L.ab && regular && flags[b] == BF_EDGE && !special) {
            // AB pattern: a whole block on a domain face, no solid near, joins the regular list; the
            // cells on its faces (sources outside the extent) go to the face-cell list
            H.blist_rg.push_back((int32_t)b);
            H.blist_rf.push_back((int32_t)b);
            L.regular_face = true;
            const int o[3] = {L.borg_h[3 * b], L.borg_h[3 * b + 1], L.borg_h[3 * b + 2]};
            bool lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                lo[a] = !dm.per[a] && o[a] == 0;
                hi[a] = !dm.per[a] && o[a] + BS == dm.ext[a];
            }
            for (int z = 0; z < BS; ++z)
                for (int y = 0; y < BS; ++y)
                    for (int x = 0; x < BS; ++x) {
                        const int l[3] = {x, y, z};
                        bool face = false;
                        for (int a = 0; a < 3; ++a) face = face || (lo[a] && l[a] == 0) || (hi[a] && l[a] == BS - 1);
                        if (face) H.face_cells.push_back(b * BS3 + (uint32_t)(x + BS * (y + BS * z)));
                    }
        } else if (L.ab && L.pipe_ok && regular && flags[b] == (BF_EDGE | BF_PARTIAL) && L.near_h[L.bdom_h[b]] && !special) {
            // AB pattern: a block that overhangs the extent, no solid near: the pipelined kernel with the
            // face rules at the extent's last cells (in a domain whose block layers span 2 GiB or more the
            // one-plane boundary kernel, whose addresses are 64-bit)
            H.blist_pr.push_back((int32_t)b);
        } else if (L.ab && L.pipe_ok && regular && (flags[b] & BF_SPONGE) && !(flags[b] & ~(BF_SPONGE | BF_EDGE | BF_PARTIAL)) &&
                   L.near_h[L.bdom_h[b]] && !special) {
            // AB pattern: a block with sponge cells and no solid near (whole or overhanging the extent):
            // the pipelined kernel's sponge form, with the face rules at the extent's last cells
            H.blist_sp.push_back((int32_t)b);
        } else {
            (regular ? H.blist_br : H.blist_bd).push_back((int32_t)b);
        }
    }
    // the regular boundary blocks whose only flag is a solid cell near first (the pipelined kernel's
    // solid-link form takes them where it applies)
    {
        const auto mid = std::stable_partition(H.blist_br.begin(), H.blist_br.end(),
                                               [&](int32_t b) { return flags[b] == BF_SOLID; });
        L.n_br_solid = (uint32_t)(mid - H.blist_br.begin());
    }
    L.n_face = (uint32_t)H.face_cells.size();
    L.n_regular_face = (uint32_t)H.blist_rf.size();
    L.n_regular = (uint32_t)H.blist_rg.size();
    L.n_interior = (uint32_t)H.blist_in.size();
    L.n_boundary_regular = (uint32_t)H.blist_br.size();
    L.n_partial = (uint32_t)H.blist_pr.size();
    L.n_sponge = (uint32_t)H.blist_sp.size();
    L.n_boundary = (uint32_t)H.blist_bd.size();
}

// Host-side construction of one level's tables from the reference-convention descriptor.
int prepare_level(const OalLevelDesc& d, const OalCollision& coll, bool ab, bool pipe_ok, size_t sbytes, Level& L, HostLevel& H,
                  const std::vector<uint8_t>* inactive = nullptr) {
    if (d.n_blocks <= 0) return fail(OAL_ERR_ARG, "level with no blocks");
    if (!d.map_x || !d.map_y || !d.map_z || !d.neighbor_table) return fail(OAL_ERR_ARG, "level block arrays missing");
    if (d.n_domains <= 0 || !d.domains) return fail(OAL_ERR_ARG, "level has no domain descriptors");
    const uint64_t ncell64 = (uint64_t)d.n_blocks * BS3;
    if (ncell64 * 27ull >= (1ull << 32)) return fail(OAL_ERR_UNSUPPORTED, "level too large for 32-bit cell indexing (n_blocks*4096*27 >= 2^32)");
    L.nb = (uint32_t)d.n_blocks;
    L.ncell = (uint32_t)ncell64;
    L.ncell_host = L.ncell;
    L.n_caller_blocks = L.nb;
    L.n_items = L.nb * PLANES;
    L.n_domains = d.n_domains;
    const uint32_t nb = L.nb;

    std::vector<int32_t>& borg = H.borg;
    std::vector<int32_t>& bdom = H.bdom;
    std::vector<int32_t>& nbr = H.nbr;
    borg.assign(3 * (size_t)nb, 0);
    bdom.assign(nb, 0);
    nbr.assign(27 * (size_t)nb, -1);
    L.map_h.assign(3 * (size_t)nb, 0);
    L.obst_h.clear();
    if (d.obstacle) {
        L.obst_h.resize((size_t)nb * BS3);
        for (size_t i = 0; i < L.obst_h.size(); ++i) L.obst_h[i] = d.obstacle[i] ? 1 : 0;
    }
    for (uint32_t b = 0; b < nb; ++b) {
        if (d.map_x[b] < 1 || d.map_y[b] < 1 || d.map_z[b] < 1) return fail(OAL_ERR_ARG, "block coordinates are 1-based");
        L.map_h[3 * b] = d.map_x[b];
        L.map_h[3 * b + 1] = d.map_y[b];
        L.map_h[3 * b + 2] = d.map_z[b];
        borg[3 * b] = (d.map_x[b] - 1) * BS;
        borg[3 * b + 1] = (d.map_y[b] - 1) * BS;
        borg[3 * b + 2] = (d.map_z[b] - 1) * BS;
        if (d.block_domain) {
            bdom[b] = d.block_domain[b];
            if (bdom[b] < 0 || bdom[b] >= d.n_domains) return fail(OAL_ERR_ARG, "block_domain out of range");
            if (b > 0 && bdom[b] < bdom[b - 1]) return fail(OAL_ERR_ARG, "blocks of one domain must be contiguous and in domain order");
        }
        for (int k = 0; k < 27; ++k) {
            const int32_t v = d.neighbor_table[(size_t)k * nb + b];
            if (v < 0 || v > (int32_t)nb) return fail(OAL_ERR_ARG, "neighbor_table entry out of range");
            nbr[(size_t)k * nb + b] = v - 1;
        }
        if (nbr[(size_t)13 * nb + b] != -1 && nbr[(size_t)13 * nb + b] != (int32_t)b)
            return fail(OAL_ERR_ARG, "neighbor_table centre entry must be 0 or the block itself");
    }

    L.doms_h.resize(d.n_domains);
    for (int i = 0; i < d.n_domains; ++i) {
        const OalDomainDesc& s = d.domains[i];
        DevDomain& t = L.doms_h[i];
        for (int a = 0; a < 3; ++a) {
            if (s.extent[a] <= 0) return fail(OAL_ERR_ARG, "domain extent must be positive");
            if (s.periodic[a] && s.extent[a] % BS) return fail(OAL_ERR_UNSUPPORTED, "a periodic axis needs an extent that is a multiple of the block size");
            t.ext[a] = s.extent[a];
            t.per[a] = s.periodic[a] ? 1 : 0;
            t.F[a] = s.body_force[a];
        }
        if (!(s.tau > 0.5f)) return fail(OAL_ERR_ARG, "tau must exceed 1/2");
        if (s.zwall_bc < 0 || s.zwall_bc > 2) return fail(OAL_ERR_ARG, "zwall_bc must be 0, 1 or 2");
        if (s.zwall_bc == 1 && !ab) return fail(OAL_ERR_UNSUPPORTED, "zwall_bc 1 (exact specular slip) needs the AB pattern");
        if (s.zwall_bc != 0 && coll.transverse_freestream)
            return fail(OAL_ERR_ARG, "z walls and the transverse freestream act on the same faces (the reference rejects the pair)");
        t.tau = s.tau;
        t.u_in = s.u_inlet;
        t.amr_scale = s.amr_sgs_scale;
        t.force_scale = s.force_scale;
        t.ymin_wall = s.ymin_wall_bc;
        t.zwall = s.zwall_bc;
        t.forced = (s.body_force[0] != 0.0f || s.body_force[1] != 0.0f || s.body_force[2] != 0.0f) ? 1 : 0;
        cell_rates<float>(0.0f, s.tau, s.amr_sgs_scale, coll.nu_sgs_background, coll.adaptive_omega4, coll.lambda,
                          coll.omega_4, t.omega1, t.omega4d);
        cell_rates<double>(0.0, (double)s.tau, (double)s.amr_sgs_scale, (double)coll.nu_sgs_background,
                           coll.adaptive_omega4, (double)coll.lambda, (double)coll.omega_4, t.omega1_d, t.omega4d_d);
    }
    // first block / first item of every domain (blocks sorted by domain)
    L.dom_block0_h.assign(d.n_domains + 1, 0);
    {
        std::vector<int32_t> count(d.n_domains, 0);
        for (uint32_t b = 0; b < nb; ++b) count[bdom[b]]++;
        for (int i = 0; i < d.n_domains; ++i) L.dom_block0_h[i + 1] = L.dom_block0_h[i] + count[i];
    }
    for (int i = 0; i < d.n_domains; ++i) L.doms_h[i].cell0 = (uint32_t)L.dom_block0_h[i] * BS3;
    L.dom_item0_h.resize(d.n_domains + 1);
    for (int i = 0; i <= d.n_domains; ++i) L.dom_item0_h[i] = L.dom_block0_h[i] * PLANES;

    // per-block flags
    std::vector<uint8_t> block_solid(nb, 0);
    if (d.obstacle) {
        for (uint32_t b = 0; b < nb; ++b)
            for (int c = 0; c < BS3; ++c)
                if (d.obstacle[(size_t)b * BS3 + c]) { block_solid[b] = 1; break; }
    }
    std::vector<uint32_t>& flags = H.flags;
    flags.assign(nb, 0);
    for (uint32_t b = 0; b < nb; ++b) {
        const DevDomain& dm = L.doms_h[bdom[b]];
        uint32_t fl = 0;
        for (int a = 0; a < 3; ++a) {
            const int o = borg[3 * b + a];
            if (!dm.per[a] && (o == 0 || o + BS >= dm.ext[a])) fl |= BF_EDGE;
            if (o + BS > dm.ext[a]) fl |= BF_PARTIAL;
            if (o >= dm.ext[a]) return fail(OAL_ERR_ARG, "a block lies entirely outside its domain extent");
        }
        if (block_solid[b]) fl |= BF_SOLID;
        if (inactive && (*inactive)[b]) fl |= BF_INACTIVE;
        for (int k = 0; k < 27; ++k) {
            if (k == 13) continue;
            const int ox = k % 3 - 1, oy = (k / 3) % 3 - 1, oz = k / 9 - 1;
            const int32_t n = nbr[(size_t)k * nb + b];
            if (n >= 0) {
                if (block_solid[n]) fl |= BF_SOLID;
                if (bdom[n] != bdom[b]) return fail(OAL_ERR_ARG, "neighbour block belongs to another domain");
                continue;
            }
            // absent neighbour: is its position inside the domain extent?
            bool inside = true;
            const int off[3] = {ox, oy, oz};
            for (int a = 0; a < 3; ++a) {
                const int o = borg[3 * b + a] + off[a] * BS;
                if (dm.per[a]) continue;
                if (o < 0 || o >= dm.ext[a]) inside = false;
            }
            if (inside) fl |= BF_MISSING;
        }
        flags[b] = fl;
    }
    // block lists: interior blocks (all flags clear) whose 26 neighbours sit at the lexicographic offsets
    // of their domain's block grid take the arithmetic-address kernel; other interior blocks the
    // table kernel; the rest the boundary kernel
    for (auto& dm : L.doms_h) {
        dm.nbx = (dm.ext[0] + BS - 1) / BS;
        dm.nby = (dm.ext[1] + BS - 1) / BS;
    }
    // the in-kernel face rules address by signed 32-bit byte offsets from a block: only domains whose
    // layer of blocks (plus a row and a block) spans less than 2 GiB may use them
    std::vector<char> near(L.doms_h.size());
    for (size_t i = 0; i < L.doms_h.size(); ++i) {
        const DevDomain& dm = L.doms_h[i];
        near[i] = ((uint64_t)dm.nbx * dm.nby + dm.nbx + 1) * 27u * BS3 * (uint64_t)sbytes < (1ull << 31);
    }
    // (a boundary block qualifies when no neighbour inside the extent is absent and every present one is
    // at its lexicographic offset: its out-of-extent sources take the face rules and are never read)
    L.regular_h.assign(nb, 0);
    for (uint32_t b = 0; b < nb; ++b) {
        const DevDomain& dm = L.doms_h[bdom[b]];
        bool regular = !(flags[b] & BF_MISSING);
        for (int k = 0; k < 27 && regular; ++k) {
            const int ox = k % 3 - 1, oy = (k / 3) % 3 - 1, oz = k / 9 - 1;
            const int64_t want = (int64_t)b + ox + (int64_t)oy * dm.nbx + (int64_t)oz * dm.nbx * dm.nby;
            const int32_t have = k == 13 ? (int32_t)b : nbr[(size_t)k * nb + b];
            if (have < 0) continue;   // absent: outside the extent (BF_MISSING is clear)
            regular = have == want;
        }
        L.regular_h[b] = regular ? 1 : 0;
    }
    L.near_h = near;
    L.flags_h = flags;
    L.borg_h = borg;
    L.bdom_h = bdom;
    L.ab = ab;
    L.pipe_ok = pipe_ok;
    L.tfs = coll.transverse_freestream != 0;
    build_lists(L, H);
    // block-major copy of the neighbour table: one 32-entry row per block (scalar loads)
    H.nbr_bm.assign((size_t)nb * 32, -1);
    for (uint32_t b = 0; b < nb; ++b)
        for (int k = 0; k < 27; ++k) H.nbr_bm[(size_t)b * 32 + k] = k == 13 ? (int32_t)b : nbr[(size_t)k * nb + b];
    // blocks holding an outlet face (x = extent - 1 of a non-periodic x axis): AA side-buffer slots
    H.oslot.assign(nb, -1);
    for (uint32_t b = 0; b < nb; ++b) {
        const DevDomain& dm = L.doms_h[bdom[b]];
        const int o = borg[3 * b];
        if (!dm.per[0] && o <= dm.ext[0] - 1 && dm.ext[0] - 1 < o + BS) H.oslot[b] = H.n_oslots++;
    }
    L.has_solid = d.obstacle != nullptr;
    if (d.obstacle) {
        H.solid.resize(L.ncell);
        for (size_t i = 0; i < L.ncell; ++i) H.solid[i] = d.obstacle[i] ? 1 : 0;
    }
    // fluid cells: inside the extent of their domain and not solid
    L.n_fluid = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        const DevDomain& dm = L.doms_h[bdom[b]];
        if (inactive && (*inactive)[b]) continue;
        for (int z = 0; z < BS; ++z)
            for (int y = 0; y < BS; ++y)
                for (int x = 0; x < BS; ++x) {
                    if (borg[3 * b] + x >= dm.ext[0] || borg[3 * b + 1] + y >= dm.ext[1] || borg[3 * b + 2] + z >= dm.ext[2])
                        continue;
                    const size_t c = (size_t)b * BS3 + x + BS * (y + BS * z);
                    if (!d.obstacle || !d.obstacle[c]) ++L.n_fluid;
                }
    }
    // solid-link masks of the blocks flagged solid: bit k when the source x - c_k of a cell is a solid
    // cell inside the extent, bit 27 when the cell itself is solid
    H.lslot.assign(nb, -1);
    int32_t nls = 0;
    for (uint32_t b = 0; b < nb; ++b)
        if (flags[b] & BF_SOLID) H.lslot[b] = nls++;
    H.lmask.assign((size_t)nls * BS3, 0u);
    L.lslot_h = H.lslot;
    L.n_lslots = nls;
    for (uint32_t b = 0; b < nb; ++b) {
        if (H.lslot[b] < 0) continue;
        const DevDomain& dm = L.doms_h[bdom[b]];
        uint32_t* m = H.lmask.data() + (size_t)H.lslot[b] * BS3;
        for (int z = 0; z < BS; ++z)
            for (int y = 0; y < BS; ++y)
                for (int x = 0; x < BS; ++x) {
                    uint32_t bits = H.solid[(size_t)b * BS3 + (z * BS + y) * BS + x] ? (1u << 27) : 0u;
                    for (int k = 0; k < 27; ++k) {
                        const int cx = k % 3 - 1, cy = (k / 3) % 3 - 1, cz = k / 9 - 1;
                        const int g[3] = {borg[3 * b] + x - cx, borg[3 * b + 1] + y - cy, borg[3 * b + 2] + z - cz};
                        bool in = true;
                        for (int a = 0; a < 3; ++a)
                            if (!dm.per[a] && (g[a] < 0 || g[a] >= dm.ext[a])) in = false;
                        if (!in) continue;
                        const int sx = x - cx, sy = y - cy, sz = z - cz;
                        const int ox = sx < 0 ? -1 : (sx >= BS ? 1 : 0), oy = sy < 0 ? -1 : (sy >= BS ? 1 : 0),
                                  oz = sz < 0 ? -1 : (sz >= BS ? 1 : 0);
                        const int32_t nbk = (ox | oy | oz) ? nbr[(size_t)((ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)) * nb + b] : (int32_t)b;
                        if (nbk < 0) continue;
                        const size_t src = (size_t)nbk * BS3 + ((sz - BS * oz) * BS + (sy - BS * oy)) * BS + (sx - BS * ox);
                        if (H.solid[src]) bits |= 1u << k;
                    }
                    m[(z * BS + y) * BS + x] = bits;
                }
    }
    return OAL_OK;
}

// ------------------------------------------------------------------------------------------------
// Refined plans: a block octree of 16^3 blocks.  Level l + 1 refines level l by 2 in every direction;
// a refined block is replaced by its 8 children.  From the callers' block lists the library derives
//   covered blocks  level-l blocks whose children exist: not stepped, written by restriction;
//   ghost blocks    level-(l+1) blocks next to the level's blocks that the level does not hold: not
//                   stepped, their cells next to stepped cells written by the fill from level l;
//   transfer lists  the cells each fill / restriction writes, the donor cells it reads and the 64-node
//                   stencils (cubic Lagrange where complete, else trilinear over the admissible donors;
//                   restriction: cubic at the coarse centre, else the children averaged), as the octree
//                   generator builds them (src/octree_q19/src/generator.jl:692-780, octree.jl:82-86).
// ------------------------------------------------------------------------------------------------
enum : uint8_t { OCT_STEPPED = 0, OCT_COVERED = 1, OCT_GHOST = 2 };

struct OctLevel {
    int ext[3] = {0, 0, 0}, per[3] = {0, 0, 0};
    int nbx = 0, nby = 0, nbz = 0;            // blocks per axis of the level's extent
    int n_caller = 0;
    std::vector<int32_t> mx, my, mz;          // 1-based block coordinates: the caller's blocks, then ghosts
    std::vector<uint8_t> status;
    std::vector<uint8_t> solid;               // per cell
    std::vector<int32_t> nbr;                 // column-major nb x 27, 1-based, 0 = none
    std::unordered_map<int64_t, int32_t> index;
    int64_t key(int bx, int by, int bz) const { return ((int64_t)bz * nby + by) * nbx + bx; }
    static int wrap(int v, int n) { return ((v % n) + n) % n; }
    // block at 0-based block coordinates (periodic axes wrap), -1 = none
    int32_t find(int bx, int by, int bz) const {
        if (per[0]) bx = wrap(bx, nbx);
        if (per[1]) by = wrap(by, nby);
        if (per[2]) bz = wrap(bz, nbz);
        if (bx < 0 || by < 0 || bz < 0 || bx >= nbx || by >= nby || bz >= nbz) return -1;
        const auto it = index.find(key(bx, by, bz));
        return it == index.end() ? -1 : it->second;
    }
    // cell index of global cell coordinates, -1 = outside the extent or no block; *blk = its block
    int64_t cell(int gx, int gy, int gz, int32_t* blk) const {
        if (per[0]) gx = wrap(gx, ext[0]);
        if (per[1]) gy = wrap(gy, ext[1]);
        if (per[2]) gz = wrap(gz, ext[2]);
        if (gx < 0 || gy < 0 || gz < 0 || gx >= ext[0] || gy >= ext[1] || gz >= ext[2]) return -1;
        const int32_t b = find(gx / BS, gy / BS, gz / BS);
        if (b < 0) return -1;
        *blk = b;
        return (int64_t)b * BS3 + (gz % BS) * BS2 + (gy % BS) * BS + gx % BS;
    }
    int nb() const { return (int)mx.size(); }
};

struct XferHost {
    std::vector<uint32_t> src, dst;   // donor cells (bit 31: not blended in time), written cells
    std::vector<int32_t> slot;        // [64][n_dst] index into src, -1 = unused node
    std::vector<float> w;             // [64][n_dst]
};

// [64][n] (node-major, as the stencils are built) -> [n][64] (cell-major, as the rebuild reads them)
template <class V>
static std::vector<V> node_to_cell_major(const std::vector<V>& a, size_t n) {
    std::vector<V> b(a.size());
    for (size_t q = 0; q < 64; ++q)
        for (size_t i = 0; i < n; ++i) b[i * 64 + q] = a[q * n + i];
    return b;
}

// written cells inside the destination level, donors inside the source level, slots inside the donor list
static bool xfer_in_bounds(const XferHost& X, uint64_t ncell_dst, uint64_t ncell_src) {
    for (uint32_t c : X.dst)
        if (c >= ncell_dst) return false;
    for (uint32_t c : X.src)
        if ((c & 0x7fffffffu) >= ncell_src) return false;
    if (X.slot.size() != 64 * X.dst.size() || X.w.size() != X.slot.size()) return false;
    for (int32_t s : X.slot)
        if (s < -1 || s >= (int64_t)X.src.size()) return false;
    return true;
}

static void cubic_weights(double t, double w[4]) {   // octree.jl:84
    w[0] = -t * (1 - t) * (2 - t) / 6;
    w[1] = (1 + t) * (1 - t) * (2 - t) / 2;
    w[2] = (1 + t) * t * (2 - t) / 2;
    w[3] = -(1 + t) * t * (1 - t) / 6;
}

// append one written cell with its stencil (per-cell arrays; transposed to node-major at the end)
struct StencilBuilder {
    std::unordered_map<uint32_t, int32_t> slot_of;
    std::vector<std::array<int32_t, 64>> sl;
    std::vector<std::array<float, 64>> ww;
    XferHost& X;
    explicit StencilBuilder(XferHost& x) : X(x) {}
    int32_t slot(uint32_t cell, bool no_blend) {
        const auto it = slot_of.find(cell);
        if (it != slot_of.end()) return it->second;
        const int32_t s = (int32_t)X.src.size();
        X.src.push_back(cell | (no_blend ? 0x80000000u : 0u));
        slot_of.emplace(cell, s);
        return s;
    }
    void finish() {
        // donors in ascending cell order: the gather then reads each direction's array in order
        std::vector<int32_t> order(X.src.size()), remap(X.src.size());
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(),
                  [&](int32_t a, int32_t b) { return (X.src[a] & 0x7fffffffu) < (X.src[b] & 0x7fffffffu); });
        std::vector<uint32_t> src(X.src.size());
        for (size_t r = 0; r < order.size(); ++r) {
            remap[order[r]] = (int32_t)r;
            src[r] = X.src[order[r]];
        }
        X.src.swap(src);
        const size_t n = X.dst.size();
        X.slot.assign(64 * n, -1);
        X.w.assign(64 * n, 0.0f);
        for (size_t i = 0; i < n; ++i)
            for (int q = 0; q < 64; ++q) {
                X.slot[(size_t)q * n + i] = sl[i][q] < 0 ? -1 : remap[sl[i][q]];
                X.w[(size_t)q * n + i] = ww[i][q];
            }
    }
};

// The entries of X whose written cell is marked in keep (indexed by entry), with only their donors.
static XferHost xfer_subset(const XferHost& X, const std::vector<char>& keep) {
    XferHost Y;
    const size_t n = X.dst.size();
    std::vector<size_t> rows;
    for (size_t i = 0; i < n; ++i)
        if (keep[i]) rows.push_back(i);
    std::vector<char> used(X.src.size(), 0);
    for (size_t r : rows)
        for (int q = 0; q < 64; ++q) {
            const int32_t sl = X.slot[(size_t)q * n + r];
            if (sl >= 0) used[sl] = 1;
        }
    std::vector<int32_t> remap(X.src.size(), -1);
    for (size_t i = 0; i < X.src.size(); ++i)
        if (used[i]) {
            remap[i] = (int32_t)Y.src.size();
            Y.src.push_back(X.src[i]);
        }
    const size_t m = rows.size();
    Y.slot.assign(64 * m, -1);
    Y.w.assign(64 * m, 0.0f);
    for (size_t j = 0; j < m; ++j) {
        const size_t r = rows[j];
        Y.dst.push_back(X.dst[r]);
        for (int q = 0; q < 64; ++q) {
            const int32_t sl = X.slot[(size_t)q * n + r];
            Y.slot[(size_t)q * m + j] = sl < 0 ? -1 : remap[sl];
            Y.w[(size_t)q * m + j] = X.w[(size_t)q * n + r];
        }
    }
    return Y;
}

// Fill of level F's ghost cells (those next to a stepped cell) from level C = F - 1.  A ghost cell at
// global fine coordinate g sits at coarse coordinate g / 2 - 1/4 (cell centres); its donors are the 4^3
// coarse cells around it with cubic weights when all are stepped or covered fluid cells, else the
// admissible ones of the inner 2^3 with trilinear weights renormalised; none: the cell is left out.
static int build_fill(const OctLevel& F, const OctLevel& C, XferHost& X, int64_t* n_cubic) {
    StencilBuilder B(X);
    *n_cubic = 0;
    for (int gb = F.n_caller; gb < F.nb(); ++gb) {
        const int bx = F.mx[gb] - 1, by = F.my[gb] - 1, bz = F.mz[gb] - 1;
        bool stepped[27];
        bool any = false;
        for (int k = 0; k < 27; ++k) {
            const int32_t nbk = k == 13 ? -1 : F.find(bx + k % 3 - 1, by + (k / 3) % 3 - 1, bz + k / 9 - 1);
            stepped[k] = nbk >= 0 && F.status[nbk] == OCT_STEPPED;
            any = any || stepped[k];
        }
        if (!any) continue;
        for (int lz = 0; lz < BS; ++lz)
            for (int ly = 0; ly < BS; ++ly)
                for (int lx = 0; lx < BS; ++lx) {
                    if (lx > 0 && lx < BS - 1 && ly > 0 && ly < BS - 1 && lz > 0 && lz < BS - 1) continue;
                    bool need = false;
                    for (int d = 0; d < 27 && !need; ++d) {
                        const int nx = lx + d % 3 - 1, ny = ly + (d / 3) % 3 - 1, nz = lz + d / 9 - 1;
                        const int ox = nx < 0 ? -1 : (nx >= BS ? 1 : 0), oy = ny < 0 ? -1 : (ny >= BS ? 1 : 0),
                                  oz = nz < 0 ? -1 : (nz >= BS ? 1 : 0);
                        if ((ox | oy | oz) && stepped[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)]) need = true;
                    }
                    if (!need) continue;
                    const int g[3] = {bx * BS + lx, by * BS + ly, bz * BS + lz};
                    int i0[3];
                    double t[3], wc[3][4];
                    for (int a = 0; a < 3; ++a) {
                        const double fx = g[a] / 2.0 - 0.25;
                        i0[a] = (int)std::floor(fx);
                        t[a] = fx - i0[a];
                        cubic_weights(t[a], wc[a]);
                    }
                    std::array<int32_t, 64> sl;
                    std::array<float, 64> ww;
                    sl.fill(-1);
                    ww.fill(0.0f);
                    int64_t dc[64];
                    bool cov[64];
                    bool complete = true;
                    for (int n = 0; n < 64 && complete; ++n) {
                        const int di = n & 3, dj = (n >> 2) & 3, dk = n >> 4;
                        int32_t blk = -1;
                        const int64_t c = C.cell(i0[0] - 1 + di, i0[1] - 1 + dj, i0[2] - 1 + dk, &blk);
                        complete = c >= 0 && C.status[blk] != OCT_GHOST && !C.solid[c];
                        if (complete) { dc[n] = c; cov[n] = C.status[blk] == OCT_COVERED; }
                    }
                    if (complete) {
                        ++*n_cubic;
                        for (int n = 0; n < 64; ++n) {
                            const int di = n & 3, dj = (n >> 2) & 3, dk = n >> 4;
                            sl[n] = B.slot((uint32_t)dc[n], cov[n]);
                            ww[n] = (float)(wc[0][di] * wc[1][dj] * wc[2][dk]);
                        }
                    } else {
                        double wt[8], wsum = 0.0;
                        int64_t tc[8];
                        bool tcov[8];
                        for (int n = 0; n < 8; ++n) {
                            const int di = n & 1, dj = (n >> 1) & 1, dk = n >> 2;
                            int32_t blk = -1;
                            const int64_t c = C.cell(i0[0] + di, i0[1] + dj, i0[2] + dk, &blk);
                            wt[n] = 0.0;
                            tc[n] = -1;
                            if (c < 0 || C.status[blk] == OCT_GHOST || C.solid[c]) continue;
                            tc[n] = c;
                            tcov[n] = C.status[blk] == OCT_COVERED;
                            wt[n] = (di ? t[0] : 1 - t[0]) * (dj ? t[1] : 1 - t[1]) * (dk ? t[2] : 1 - t[2]);
                            wsum += wt[n];
                        }
                        if (!(wsum > 0.0)) continue;
                        for (int n = 0; n < 8; ++n) {
                            if (tc[n] < 0) continue;
                            const int di = n & 1, dj = (n >> 1) & 1, dk = n >> 2;
                            const int q = (di + 1) + 4 * (dj + 1) + 16 * (dk + 1);
                            sl[q] = B.slot((uint32_t)tc[n], tcov[n]);
                            ww[q] = (float)(wt[n] / wsum);
                        }
                    }
                    X.dst.push_back((uint32_t)gb * BS3 + (uint32_t)(lz * BS2 + ly * BS + lx));
                    B.sl.push_back(sl);
                    B.ww.push_back(ww);
                }
    }
    B.finish();
    return OAL_OK;
}

// Restriction of level F = C + 1 into the covered fluid cells of C: the coarse centre lies between fine
// cells 2i and 2i + 1; cubic weights (-1, 9, 9, -1)/16 per axis over fine cells 2i - 1 .. 2i + 2 when all
// are stepped or covered fluid cells, else the average of the admissible children (octree.jl:86,
// coupling.jl:52-76).
static int build_restrict(const OctLevel& C, const OctLevel& F, XferHost& X, int64_t* n_cubic) {
    StencilBuilder B(X);
    *n_cubic = 0;
    double wr[4];
    cubic_weights(0.5, wr);
    for (int cb = 0; cb < C.n_caller; ++cb) {
        if (C.status[cb] != OCT_COVERED) continue;
        const int bx = C.mx[cb] - 1, by = C.my[cb] - 1, bz = C.mz[cb] - 1;
        for (int lz = 0; lz < BS; ++lz)
            for (int ly = 0; ly < BS; ++ly)
                for (int lx = 0; lx < BS; ++lx) {
                    const uint32_t c = (uint32_t)cb * BS3 + (uint32_t)(lz * BS2 + ly * BS + lx);
                    if (C.solid[c]) continue;
                    const int i[3] = {bx * BS + lx, by * BS + ly, bz * BS + lz};
                    std::array<int32_t, 64> sl;
                    std::array<float, 64> ww;
                    sl.fill(-1);
                    ww.fill(0.0f);
                    int64_t fc[64];
                    bool complete = true;
                    for (int n = 0; n < 64 && complete; ++n) {
                        const int di = n & 3, dj = (n >> 2) & 3, dk = n >> 4;
                        int32_t blk = -1;
                        const int64_t f = F.cell(2 * i[0] - 1 + di, 2 * i[1] - 1 + dj, 2 * i[2] - 1 + dk, &blk);
                        complete = f >= 0 && F.status[blk] != OCT_GHOST && !F.solid[f];
                        fc[n] = f;
                    }
                    if (complete) {
                        ++*n_cubic;
                        for (int n = 0; n < 64; ++n) {
                            const int di = n & 3, dj = (n >> 2) & 3, dk = n >> 4;
                            sl[n] = B.slot((uint32_t)fc[n], false);
                            ww[n] = (float)(wr[di] * wr[dj] * wr[dk]);
                        }
                    } else {
                        int64_t ch[8];
                        int cnt = 0;
                        for (int n = 0; n < 8; ++n) {
                            const int di = n & 1, dj = (n >> 1) & 1, dk = n >> 2;
                            int32_t blk = -1;
                            const int64_t f = F.cell(2 * i[0] + di, 2 * i[1] + dj, 2 * i[2] + dk, &blk);
                            ch[n] = (f >= 0 && F.status[blk] != OCT_GHOST && !F.solid[f]) ? f : -1;
                            if (ch[n] >= 0) ++cnt;
                        }
                        if (!cnt) continue;
                        for (int n = 0; n < 8; ++n) {
                            if (ch[n] < 0) continue;
                            const int di = n & 1, dj = (n >> 1) & 1, dk = n >> 2;
                            const int q = (di + 1) + 4 * (dj + 1) + 16 * (dk + 1);
                            sl[q] = B.slot((uint32_t)ch[n], false);
                            ww[q] = (float)(1.0 / cnt);
                        }
                    }
                    X.dst.push_back(c);
                    B.sl.push_back(sl);
                    B.ww.push_back(ww);
                }
    }
    B.finish();
    return OAL_OK;
}

// The octree of a refined plan from the callers' level descriptors (coarsest first).
static int build_octree(const OalPlanDesc* desc, std::vector<OctLevel>& O) {
    const int L = desc->n_levels;
    O.assign(L, OctLevel{});
    for (int l = 0; l < L; ++l) {
        const OalLevelDesc& d = desc->levels[l];
        if (d.n_domains != 1 || !d.domains) return fail(OAL_ERR_UNSUPPORTED, "each level of a refined plan holds one domain");
        if (d.block_domain) return fail(OAL_ERR_UNSUPPORTED, "block_domain must be NULL on a refined plan");
        if (d.n_blocks <= 0 || !d.map_x || !d.map_y || !d.map_z) return fail(OAL_ERR_ARG, "level block arrays missing");
        OctLevel& o = O[l];
        for (int a = 0; a < 3; ++a) {
            o.ext[a] = d.domains[0].extent[a];
            o.per[a] = d.domains[0].periodic[a] ? 1 : 0;
            if (o.ext[a] <= 0 || o.ext[a] % BS) return fail(OAL_ERR_UNSUPPORTED, "refined plans need extents that are multiples of 16");
            if (l > 0 && (o.ext[a] != 2 * O[l - 1].ext[a] || o.per[a] != O[l - 1].per[a]))
                return fail(OAL_ERR_ARG, "level l + 1 must have twice the extent of level l and the same periodicity");
        }
        o.nbx = o.ext[0] / BS; o.nby = o.ext[1] / BS; o.nbz = o.ext[2] / BS;
        o.n_caller = d.n_blocks;
        o.mx.assign(d.map_x, d.map_x + d.n_blocks);
        o.my.assign(d.map_y, d.map_y + d.n_blocks);
        o.mz.assign(d.map_z, d.map_z + d.n_blocks);
        for (int b = 0; b < d.n_blocks; ++b) {
            const int bx = o.mx[b] - 1, by = o.my[b] - 1, bz = o.mz[b] - 1;
            if (bx < 0 || by < 0 || bz < 0 || bx >= o.nbx || by >= o.nby || bz >= o.nbz)
                return fail(OAL_ERR_ARG, "block coordinates outside the level extent");
            if (!o.index.emplace(o.key(bx, by, bz), b).second) return fail(OAL_ERR_ARG, "duplicate block in a level");
        }
        o.status.assign(d.n_blocks, OCT_STEPPED);
        o.solid.assign((size_t)d.n_blocks * BS3, 0);
        if (d.obstacle)
            for (size_t i = 0; i < o.solid.size(); ++i) o.solid[i] = d.obstacle[i] ? 1 : 0;
    }
    // nesting: every block of a finer level has its parent, and parents have all 8 children
    for (int l = 1; l < L; ++l) {
        OctLevel& o = O[l];
        for (int b = 0; b < o.n_caller; ++b) {
            const int bx = o.mx[b] - 1, by = o.my[b] - 1, bz = o.mz[b] - 1;
            if (O[l - 1].find(bx / 2, by / 2, bz / 2) < 0) return fail(OAL_ERR_ARG, "a refined block has no parent block");
            for (int c = 0; c < 8; ++c)
                if (o.find((bx & ~1) + (c & 1), (by & ~1) + ((c >> 1) & 1), (bz & ~1) + (c >> 2)) < 0)
                    return fail(OAL_ERR_ARG, "a refined block lacks one of its 7 siblings");
        }
        // covered parents
        OctLevel& c = O[l - 1];
        for (int b = 0; b < c.n_caller; ++b)
            if (o.find(2 * (c.mx[b] - 1), 2 * (c.my[b] - 1), 2 * (c.mz[b] - 1)) >= 0) c.status[b] = OCT_COVERED;
    }
    // ghost blocks around every finer level; their parents must exist (a margin of one coarse block)
    for (int l = 1; l < L; ++l) {
        OctLevel& o = O[l];
        for (int b = 0; b < o.n_caller; ++b) {
            const int bx = o.mx[b] - 1, by = o.my[b] - 1, bz = o.mz[b] - 1;
            for (int k = 0; k < 27; ++k) {
                if (k == 13) continue;
                int nx = bx + k % 3 - 1, ny = by + (k / 3) % 3 - 1, nz = bz + k / 9 - 1;
                if (o.per[0]) nx = OctLevel::wrap(nx, o.nbx);
                if (o.per[1]) ny = OctLevel::wrap(ny, o.nby);
                if (o.per[2]) nz = OctLevel::wrap(nz, o.nbz);
                if (nx < 0 || ny < 0 || nz < 0 || nx >= o.nbx || ny >= o.nby || nz >= o.nbz) continue;
                if (o.find(nx, ny, nz) >= 0) continue;
                const int32_t pb = O[l - 1].find(nx / 2, ny / 2, nz / 2);
                if (pb < 0 || O[l - 1].status[pb] == OCT_GHOST)
                    return fail(OAL_ERR_ARG, "a refined region needs a margin of one coarser block around it");
                const int32_t g = o.nb();
                o.index.emplace(o.key(nx, ny, nz), g);
                o.mx.push_back(nx + 1);
                o.my.push_back(ny + 1);
                o.mz.push_back(nz + 1);
                o.status.push_back(OCT_GHOST);
                o.solid.resize(o.solid.size() + BS3, 0);
            }
        }
    }
    for (int l = 0; l < L; ++l) {
        OctLevel& o = O[l];
        const int nb = o.nb();
        o.nbr.assign((size_t)27 * nb, 0);
        for (int b = 0; b < nb; ++b)
            for (int k = 0; k < 27; ++k) {
                const int32_t n = o.find(o.mx[b] - 1 + k % 3 - 1, o.my[b] - 1 + (k / 3) % 3 - 1, o.mz[b] - 1 + k / 9 - 1);
                o.nbr[(size_t)k * nb + b] = n + 1;
            }
    }
    return OAL_OK;
}

// Device state of one level from its validated host tables.
int upload_level(OalPlan* p, Level& L, const HostLevel& H) {
    const uint32_t nb = L.nb;
    int rc;
    if ((rc = upload(L.nbr, H.nbr_bm)) || (rc = upload(L.borg, H.borg)) || (rc = upload(L.bdom, H.bdom)) ||
        (rc = upload(L.bflags, H.flags)) || (rc = upload(L.doms, L.doms_h)) || (rc = upload(L.dom_item0, L.dom_item0_h)) ||
        (rc = upload(L.dom_block0, L.dom_block0_h)))
        return rc;
    if (!H.solid.empty() && (rc = upload(L.solid, H.solid))) return rc;
    if ((rc = upload(L.blist_rg, H.blist_rg)) || (rc = upload(L.blist_in, H.blist_in)) ||
        (rc = upload(L.blist_br, H.blist_br)) || (rc = upload(L.blist_bd, H.blist_bd)) || (rc = upload(L.blist_pr, H.blist_pr)) || (rc = upload(L.blist_sp, H.blist_sp)) || (rc = upload(L.blist_rp, H.blist_rp)) || (rc = upload(L.blist_rf, H.blist_rf)) ||
        (rc = upload(L.lslot, H.lslot)) || (rc = upload(L.lmask, H.lmask)) || (rc = upload(L.face_cells, H.face_cells)))
        return rc;
    if (p->stream_mode == OAL_STREAM_AA) {
        if ((rc = upload(L.oslot, H.oslot))) return rc;
        if (H.n_oslots && L.obuf.alloc((size_t)H.n_oslots * 9 * BS2 * storage_bytes(p->storage)) != hipSuccess)
            return fail(OAL_ERR_NOMEM, "outlet buffer allocation failed");
    }
    const size_t fbytes = (size_t)L.ncell * 27 * storage_bytes(p->storage);
    if (L.fA.alloc(fbytes) != hipSuccess) return fail(OAL_ERR_NOMEM, "population buffer allocation failed");
    if (p->stream_mode == OAL_STREAM_AB && L.fB.alloc(fbytes) != hipSuccess)
        return fail(OAL_ERR_NOMEM, "population buffer allocation failed");
    if (p->wale) {
        const size_t vbytes = (size_t)L.ncell * 3 * work_bytes(p->storage);
        if (L.velA.alloc(vbytes) != hipSuccess || L.velB.alloc(vbytes) != hipSuccess)
            return fail(OAL_ERR_NOMEM, "velocity buffer allocation failed");
        OAL_HIP(hipMemset(L.velA.p, 0, vbytes));
        OAL_HIP(hipMemset(L.velB.p, 0, vbytes));
    }
    if (L.fpart.alloc((size_t)L.n_items * 4) != hipSuccess || L.mpart.alloc((size_t)nb * 4) != hipSuccess ||
        L.red_out.alloc((size_t)L.n_domains * 4) != hipSuccess || L.guard.alloc(1) != hipSuccess)
        return fail(OAL_ERR_NOMEM, "reduction buffer allocation failed");
    // force partials start at zero: the pipelined interior kernel (no solid links) never writes them
    OAL_HIP(hipMemset(L.fpart.p, 0, (size_t)L.n_items * 4 * sizeof(double)));
    return OAL_OK;
}

StepArgs make_args(OalPlan* p, Level& L, int32_t time_step, bool want_force) {
    StepArgs a{};
    a.fin = L.fA.p;
    a.fout = p->stream_mode == OAL_STREAM_AB ? (void*)L.fB.p : (void*)L.fA.p;
    a.nbr_bm = L.nbr.p;
    a.borg = L.borg.p;
    a.bdom = L.bdom.p;
    a.bflags = L.bflags.p;
    a.solid = L.solid.p;
    a.doms = L.doms.p;
    a.vel_in = p->wale ? L.velA.p : nullptr;
    a.vel_out = p->wale ? L.velB.p : nullptr;
    a.rho_out = nullptr;
    a.fpart = want_force ? L.fpart.p : nullptr;
    a.oslot = L.oslot.p;
    a.blist = nullptr;
    a.blist_regular = L.blist_rg.p;
    a.blist_interior = L.blist_in.p;
    a.blist_boundary_regular = L.blist_br.p;
    a.blist_boundary = L.blist_bd.p;
    a.n_regular = L.n_regular;
    a.blist_regular_plain = L.blist_rp.p;
    a.n_regular_plain = L.n_regular - L.n_regular_face;
    a.blist_regular_face = L.blist_rf.p;
    a.n_regular_face = L.n_regular_face;
    a.blist_partial = L.blist_pr.p;
    a.blist_sponge = L.blist_sp.p;
    a.n_sponge = L.n_sponge;
    a.n_partial = L.n_partial;
    a.n_interior = L.n_interior;
    a.n_boundary_regular = L.n_boundary_regular;
    a.n_boundary = L.n_boundary;
    a.n_boundary_regular_solid = L.pipe_ok ? L.n_br_solid : 0;
    a.lslot = L.lslot.p;
    a.lmask = L.lmask.p;
    a.face_cells = L.face_cells.p;
    a.n_face = L.n_face;
    a.regular_face = L.regular_face ? 1 : 0;
    // the in-kernel face rules address by signed 32-bit byte offsets from a block: a domain whose
    // layer of blocks (plus a row and a block) spans 2 GiB or more takes the face-cell kernel
    bool near = true;
    for (const DevDomain& dm : L.doms_h)
        near = near && ((uint64_t)dm.nbx * dm.nby + dm.nbx + 1) * 27u * BS3 * (uint64_t)storage_bytes(p->storage) < (1ull << 31);
    a.face_kernel = (p->face_kernel || !near) ? 1 : 0;
    a.face_x = p->cp.tfs ? 1 : 0;
    for (const DevDomain& dm : L.doms_h) a.face_x |= dm.zwall == 1 ? 1 : 0;
    a.obuf = L.obuf.p;
    a.ncell = L.ncell;
    a.nb = L.nb;
    a.n_items = L.n_items;
    a.time_step = time_step;
    if (p->capturing) {
        // replayed steps: the coarse step from device memory, times 2^level, plus the sub-step
        a.ts_dev = p->ts_dev;
        a.ts_mul = 1 << (int)(&L - p->levels.data());
        a.ts_add = time_step;
    } else {
        a.ts_dev = nullptr;
        a.ts_mul = 0;
        a.ts_add = 0;
    }
    a.parity = p->stream_mode == OAL_STREAM_AA ? (L.aa_phase == 0 ? 1 : 0) : 0;
    a.pipe_planes = p->pipe_planes;
    // (two cells per lane where the level's block grid fits its 32-bit offsets)
    bool pair_ok = p->pair_kernel != 0;
    for (const DevDomain& dm : L.doms_h) pair_ok = pair_ok && pair_fits(dm.nbx, dm.nby, p->storage == OAL_F32 ? 4 : 2);
    a.pair_kernel = pair_ok ? p->pair_kernel : 0;
    a.qmap = L.qmap.p;
    a.sponge = L.has_sponge ? L.sponge.p : nullptr;
    a.sslot = L.sslot.p;
    a.sponge_blend = L.sponge_blend;
    a.q_min = L.q_min;
    a.bfl_fallback = L.bfl_fallback;
    a.cp = p->cp;
    return a;
}

int level_step(OalPlan* p, int level, int32_t time_step, int flags, hipStream_t s, hipEvent_t bulk0 = nullptr,
               hipEvent_t bulk1 = nullptr) {
    Level& L = p->levels[level];
    if (!L.have_f) return fail(OAL_ERR_STATE, "no populations: call oal_upload_f or oal_init_equilibrium first");
    const bool want_force = (flags & OAL_STEP_FORCE) != 0;
    if (want_force && !L.has_solid) return fail(OAL_ERR_STATE, "force requested on a level without obstacle cells");
    StepArgs a = make_args(p, L, time_step, want_force);
    // side-stream work: boundary blocks, face cells and the blocks that overhang the extent
    const bool split = p->overlap && L.n_regular + L.n_interior > 0 &&
                       L.n_boundary_regular + L.n_boundary + L.n_face + L.n_partial + L.n_sponge > 0;
    if (split) {
        // fork: the side stream starts after everything queued on s, then s waits for the side stream
        OAL_HIP(hipEventRecord(p->fork, s));
        OAL_HIP(hipStreamWaitEvent(p->side, p->fork, 0));
    }
    OAL_HIP(launch_step(p->storage, p->stream_mode, p->wale, a, s, split ? p->side : nullptr, bulk0, bulk1));
    if (split) {
        OAL_HIP(hipEventRecord(p->join, p->side));
        OAL_HIP(hipStreamWaitEvent(s, p->join, 0));
    }
    if (p->stream_mode == OAL_STREAM_AB) std::swap(L.fA, L.fB);
    else L.aa_phase ^= 1;
    if (p->wale) std::swap(L.velA, L.velB);
    if (want_force) {
        ReduceArgs r{L.fpart.p, L.dom_item0.p, L.red_out.p, L.n_domains, 3};
        OAL_HIP(launch_reduce(r, s));
        L.force_valid = true;
    } else {
        L.force_valid = false;
    }
    return OAL_OK;
}

int permute_of(const OalPlan* p) { return p->stream_mode == OAL_STREAM_AA ? 1 : 0; }

// ---- refined plans: transfers and the recursive step (src/octree_q19/src/coupling.jl:134-268) ----
// Fill level l's ghost cells from level l - 1 at weight w1 along the coarse step: 1 = its current state
// (fA), else the blend with its previous one (fB); covered donors (the finer level's restriction)
// enter unblended.
int ml_fill(OalPlan* p, int l, double w1, hipStream_t s) {
    Level& F = p->levels[l];
    Level& C = p->levels[l - 1];
    if (!F.n_gdst) return OAL_OK;
    OAL_HIP(launch_xfer_gather(p->storage, C.fA.p, w1 == 1.0 ? nullptr : C.fB.p, F.g_src.p, F.n_gsrc, C.ncell, w1,
                                F.g_m.p, s));
    OAL_HIP(launch_xfer_rebuild(p->storage, w1 == 1.0, F.fA.p, F.ncell, F.g_dst.p, F.n_gdst, F.g_slot.p, F.g_w.p, F.g_m.p,
                                F.fill_scale, F.fill_keep, p->wale ? (void*)F.velA.p : nullptr, s));
    return OAL_OK;
}

// Restrict level l + 1's current state into level l's covered cells: those the stepping reads, or
// (full) every covered fluid cell.
int ml_restrict(OalPlan* p, int l, hipStream_t s, bool full = false) {
    Level& C = p->levels[l];
    Level& F = p->levels[l + 1];
    const uint32_t nd = full ? C.n_rfdst : C.n_rdst, ns = full ? C.n_rfsrc : C.n_rsrc;
    if (!nd) return OAL_OK;
    OAL_HIP(launch_xfer_gather(p->storage, F.fA.p, nullptr, full ? C.rf_src.p : C.r_src.p, ns, F.ncell, 1.0, C.r_m.p, s));
    OAL_HIP(launch_xfer_rebuild(p->storage, true, C.fA.p, C.ncell, full ? C.rf_dst.p : C.r_dst.p, nd,
                                full ? C.rf_slot.p : C.r_slot.p, full ? C.rf_w.p : C.r_w.p, C.r_m.p, C.restrict_scale,
                                C.restrict_keep, p->wale ? (void*)C.velA.p : nullptr, s));
    return OAL_OK;
}

// Before the state of a refined plan is read out: every covered cell restricted, finest first.
int ml_sync_covered(OalPlan* p, hipStream_t s) {
    if (!p->multilevel || !p->covered_stale) return OAL_OK;
    for (int l = (int)p->levels.size() - 2; l >= 0; --l)
        if (int rc = ml_restrict(p, l, s, true)) return rc;
    p->covered_stale = false;
    return OAL_OK;
}

// Initial consistency (couple_initial!, coupling.jl:253-268): restrict every level into its parent,
// then fill every level's ghosts; both buffers of a level hold the result.
int ml_couple(OalPlan* p, hipStream_t s) {
    const int nl = (int)p->levels.size();
    for (int l = 0; l < nl; ++l)
        if (!p->levels[l].have_f) return fail(OAL_ERR_STATE, "every level of a refined plan needs populations");
    for (int l = nl - 2; l >= 0; --l) {
        if (int rc = ml_restrict(p, l, s, true)) return rc;
        OAL_HIP(hipMemcpyAsync(p->levels[l].fB.p, p->levels[l].fA.p, p->levels[l].fA.n, hipMemcpyDeviceToDevice, s));
    }
    for (int l = 1; l < nl; ++l) {
        if (int rc = ml_fill(p, l, 1.0, s)) return rc;
        OAL_HIP(hipMemcpyAsync(p->levels[l].fB.p, p->levels[l].fA.p, p->levels[l].fA.n, hipMemcpyDeviceToDevice, s));
    }
    p->coupled = true;
    p->covered_stale = false;
    return OAL_OK;
}

// One step of level l and, recursively, two of each finer level (step_recursive!, coupling.jl:236-251).
// Level l's sub-steps are numbered t = 2^l (coarse step) + k.  The force is collected on the last
// sub-step of every level that holds obstacle cells when `last`.
int ml_step_rec(OalPlan* p, int l, int32_t t, int flags, bool last, hipStream_t s) {
    Level& L = p->levels[l];
    const int f = (last && L.has_solid) ? flags : (flags & ~OAL_STEP_FORCE);
    if (int rc = level_step(p, l, t, f, s)) return rc;
    // the ghost cells of level l in its new read buffer are refilled by the parent before level l's
    // next sub-step reads them; nothing reads them in between (transfers take stepped or covered
    // donors only), so they need no copy from the other buffer
    if (l + 1 < (int)p->levels.size()) {
        if (int rc = ml_step_rec(p, l + 1, 2 * t, flags, false, s)) return rc;
        if (int rc = ml_restrict(p, l, s)) return rc;
        if (int rc = ml_fill(p, l + 1, 0.5, s)) return rc;
        if (int rc = ml_step_rec(p, l + 1, 2 * t + 1, flags, last, s)) return rc;
        if (int rc = ml_restrict(p, l, s)) return rc;
        if (int rc = ml_fill(p, l + 1, 1.0, s)) return rc;
    }
    return OAL_OK;
}

// ---- captured steps ----
// The host state a step changes: which buffer of each level is read next, the AA phase.
std::vector<OalPlan::LevelState> level_states(const OalPlan* p) {
    std::vector<OalPlan::LevelState> v;
    for (const Level& L : p->levels) v.push_back({L.fA.p, L.velA.p, L.aa_phase});
    return v;
}

void set_level_states(OalPlan* p, const std::vector<OalPlan::LevelState>& v) {
    for (size_t l = 0; l < v.size(); ++l) {
        Level& L = p->levels[l];
        if (L.fA.p != v[l].fA) std::swap(L.fA, L.fB);
        if (L.velA.p != v[l].velA) std::swap(L.velA, L.velB);
        L.aa_phase = v[l].aa_phase;
        L.force_valid = false;
    }
}

bool same_states(const std::vector<OalPlan::LevelState>& a, const std::vector<OalPlan::LevelState>& b) {
    if (a.size() != b.size()) return false;
    for (size_t l = 0; l < a.size(); ++l)
        if (a[l].fA != b[l].fA || a[l].velA != b[l].velA || a[l].aa_phase != b[l].aa_phase) return false;
    return true;
}

void drop_graphs(OalPlan* p) {
    for (auto& g : p->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    p->graphs.clear();
}

// Capture one (coarse) step from the current buffer arrangement into a graph; the host state is left
// as it was (the capture only records launches).
int capture_step(OalPlan* p, OalPlan::StepGraph& g) {
    g.before = level_states(p);
    hipStream_t s = p->stream;
    OAL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    p->capturing = true;
    int rc = p->multilevel ? ml_step_rec(p, 0, 0, 0, false, s) : level_step(p, 0, 0, 0, s);
    p->capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    g.after = level_states(p);
    set_level_states(p, g.before);
    if (rc == OAL_OK && e != hipSuccess) rc = fail(OAL_ERR_HIP, std::string("step capture: ") + hipGetErrorString(e));
    if (rc == OAL_OK) {
        const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) rc = fail(OAL_ERR_HIP, std::string("step graph instantiation: ") + hipGetErrorString(ei));
    }
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
}

// One (coarse) step with index t by graph replay.
int replay_step(OalPlan* p, int32_t t) {
    const auto now = level_states(p);
    OalPlan::StepGraph* g = nullptr;
    for (auto& c : p->graphs)
        if (same_states(c.before, now)) g = &c;
    if (!g) {
        if (p->graphs.size() >= 4) drop_graphs(p);
        OalPlan::StepGraph ng;
        if (int rc = capture_step(p, ng)) return rc;
        p->graphs.push_back(ng);
        g = &p->graphs.back();
    }
    OAL_HIP(launch_set_i32(p->ts_dev, t, p->stream));
    OAL_HIP(hipGraphLaunch(g->exec, p->stream));
    set_level_states(p, g->after);
    return OAL_OK;
}

// AA pattern: the outlet side buffer mirrors the freshly set state (the first step is an odd one)
hipError_t aa_prep(OalPlan* p, Level& L) {
    if (p->stream_mode != OAL_STREAM_AA || !L.obuf.p) return hipSuccess;
    StepArgs a = make_args(p, L, 0, false);
    return launch_aa_outlet_prep(p->storage, a, p->stream);
}

int require_plain_state(const OalPlan* p, const Level& L) {
    if (!L.have_f) return fail(OAL_ERR_STATE, "no populations on this level");
    if (p->stream_mode == OAL_STREAM_AA && L.aa_phase != 0)
        return fail(OAL_ERR_STATE, "AA pattern: the state is readable after an even number of steps only");
    return OAL_OK;
}

// WALE needs the previous step's velocity: rebuild it from the populations just set
int refresh_velocity(OalPlan* p, Level& L) {
    if (!p->wale) return OAL_OK;
    OAL_HIP(launch_macro(p->storage, L.fA.p, nullptr, nullptr, L.ncell, L.ncell, permute_of(p), L.velA.p, p->stream));
    return OAL_OK;
}

// ---- The reference's 8^3 BlockLevel (src/geometry/blocks.jl:14, BLOCK_SIZE = 8) -----------------------
// A level given in 8^3 blocks is regrouped into the library's 16^3 blocks: each 16^3 block holds the
// eight 8^3 blocks (octants) 2B + (0|1, 0|1, 0|1).  Every octant of a 16^3 block that reaches into the
// domain extent must be present (no holes: the kernels' neighbour rules work per 16^3 block); octants
// wholly beyond the extent may be absent and their cells are never stepped or read through.  The 16^3
// neighbour table follows the caller's 8^3 table (so periodic wraps and separate domains carry over).
constexpr int BS8 = 8, BS8_3 = 512;
struct Regroup8 {
    std::vector<int32_t> mx, my, mz, nbr, bdom;   // the 16^3 level; nbr column-major nb x 27, 1-based
    std::vector<uint8_t> obst;                    // [nb * BS3] or empty
    std::vector<uint32_t> cmap;                   // caller cell -> library cell
};

int regroup8(const OalLevelDesc& d, Regroup8& R) {
    if (d.n_blocks <= 0) return fail(OAL_ERR_ARG, "level with no blocks");
    if (!d.map_x || !d.map_y || !d.map_z || !d.neighbor_table) return fail(OAL_ERR_ARG, "level block arrays missing");
    if (d.n_domains <= 0 || !d.domains) return fail(OAL_ERR_ARG, "level has no domain descriptors");
    const int n8 = d.n_blocks;
    struct Key { int x, y, z, dom; };
    std::unordered_map<uint64_t, int> idx16;     // 16^3 block coordinate -> 16^3 block
    std::vector<Key> keys;
    std::vector<std::array<int32_t, 8>> oct;      // 8^3 block of each octant, -1 = absent
    auto pack = [](int x, int y, int z) { return ((uint64_t)(uint32_t)x << 42) | ((uint64_t)(uint32_t)y << 21) | (uint64_t)(uint32_t)z; };
    std::vector<int32_t> b16(n8), o8(n8);
    for (int b = 0; b < n8; ++b) {
        if (d.map_x[b] < 1 || d.map_y[b] < 1 || d.map_z[b] < 1) return fail(OAL_ERR_ARG, "block coordinates are 1-based");
        if (d.map_x[b] > (1 << 20) || d.map_y[b] > (1 << 20) || d.map_z[b] > (1 << 20)) return fail(OAL_ERR_ARG, "block coordinate too large");
        const int cx = d.map_x[b] - 1, cy = d.map_y[b] - 1, cz = d.map_z[b] - 1;
        const int dom = d.block_domain ? d.block_domain[b] : 0;
        if (dom < 0 || dom >= d.n_domains) return fail(OAL_ERR_ARG, "block_domain out of range");
        const uint64_t k = pack(cx >> 1, cy >> 1, cz >> 1);
        auto it = idx16.find(k);
        int B;
        if (it == idx16.end()) {
            B = (int)keys.size();
            idx16.emplace(k, B);
            keys.push_back({cx >> 1, cy >> 1, cz >> 1, dom});
            std::array<int32_t, 8> a;
            a.fill(-1);
            oct.push_back(a);
        } else {
            B = it->second;
            if (keys[B].dom != dom) return fail(OAL_ERR_UNSUPPORTED, "8^3 blocks: the octants of a 16^3 block must share one domain");
        }
        const int o = (cx & 1) + 2 * (cy & 1) + 4 * (cz & 1);
        if (oct[B][o] >= 0) return fail(OAL_ERR_ARG, "two blocks with the same coordinates");
        oct[B][o] = b;
        b16[b] = B;
        o8[b] = o;
    }
    const int n16 = (int)keys.size();
    // no holes inside the extent
    for (int B = 0; B < n16; ++B) {
        const OalDomainDesc& dm = d.domains[keys[B].dom];
        for (int o = 0; o < 8; ++o) {
            if (oct[B][o] >= 0) continue;
            const int ox = (2 * keys[B].x + (o & 1)) * BS8, oy = (2 * keys[B].y + ((o >> 1) & 1)) * BS8,
                      oz = (2 * keys[B].z + (o >> 2)) * BS8;
            if (ox < dm.extent[0] && oy < dm.extent[1] && oz < dm.extent[2])
                return fail(OAL_ERR_UNSUPPORTED, "8^3 blocks: a 16^3 block of the library would miss an octant inside the "
                                                 "domain extent (levels with holes at that scale are not supported)");
        }
    }
    // library order: by domain, then z, y, x
    std::vector<int> order(n16);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        const Key &p = keys[a], &q = keys[b];
        if (p.dom != q.dom) return p.dom < q.dom;
        if (p.z != q.z) return p.z < q.z;
        if (p.y != q.y) return p.y < q.y;
        return p.x < q.x;
    });
    std::vector<int> rank(n16);
    for (int i = 0; i < n16; ++i) rank[order[i]] = i;
    R.mx.resize(n16); R.my.resize(n16); R.mz.resize(n16); R.bdom.resize(n16);
    for (int B = 0; B < n16; ++B) {
        const int i = rank[B];
        R.mx[i] = keys[B].x + 1; R.my[i] = keys[B].y + 1; R.mz[i] = keys[B].z + 1; R.bdom[i] = keys[B].dom;
    }
    // neighbour table: for direction k, the octants on that side of the block must agree on the 16^3
    // block their 8^3 neighbours fall in (or all have none)
    R.nbr.assign((size_t)27 * n16, 0);
    for (int B = 0; B < n16; ++B) {
        for (int k = 0; k < 27; ++k) {
            const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
            if (k == 13) { R.nbr[(size_t)13 * n16 + rank[B]] = rank[B] + 1; continue; }
            int found = -2;   // -2: no octant looked at yet
            for (int o = 0; o < 8; ++o) {
                const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
                if ((dx > 0 && ox != 1) || (dx < 0 && ox != 0) || (dy > 0 && oy != 1) || (dy < 0 && oy != 0) ||
                    (dz > 0 && oz != 1) || (dz < 0 && oz != 0))
                    continue;
                const int b = oct[B][o];
                if (b < 0) continue;
                const int32_t v = d.neighbor_table[(size_t)k * n8 + b];
                if (v < 0 || v > n8) return fail(OAL_ERR_ARG, "neighbor_table entry out of range");
                const int nb16 = v == 0 ? -1 : rank[b16[v - 1]];
                if (found == -2) found = nb16;
                else if (found != nb16)
                    return fail(OAL_ERR_UNSUPPORTED, "8^3 blocks: the neighbour table does not regroup into 16^3 blocks");
            }
            R.nbr[(size_t)k * n16 + rank[B]] = found >= 0 ? found + 1 : 0;
        }
    }
    // cells
    R.cmap.resize((size_t)n8 * BS8_3);
    if (d.obstacle) R.obst.assign((size_t)n16 * BS3, 0);
    for (int b = 0; b < n8; ++b) {
        const int o = o8[b], i = rank[b16[b]];
        const int x0 = (o & 1) * BS8, y0 = ((o >> 1) & 1) * BS8, z0 = (o >> 2) * BS8;
        for (int z = 0; z < BS8; ++z)
            for (int y = 0; y < BS8; ++y)
                for (int x = 0; x < BS8; ++x) {
                    const size_t c8 = (size_t)b * BS8_3 + (size_t)(z * BS8 + y) * BS8 + x;
                    const uint32_t c16 = (uint32_t)i * BS3 + (uint32_t)((z0 + z) * BS + (y0 + y)) * BS + (uint32_t)(x0 + x);
                    R.cmap[c8] = c16;
                    if (d.obstacle) R.obst[c16] = d.obstacle[c8] ? 1 : 0;
                }
    }
    return OAL_OK;
}

// per-cell host arrays between the caller's 8^3 layout (comps x n8 cells) and the library's
// (comps x ncell_host cells; cells outside the caller's blocks take `fill`)
template <class T>
std::vector<T> cells_to_lib(const Level& L, const T* src, int comps, T fill) {
    const size_t n8 = L.c8map.size(), n = L.ncell_host;
    std::vector<T> out((size_t)comps * n, fill);
    for (int c = 0; c < comps; ++c)
        for (size_t i = 0; i < n8; ++i) out[(size_t)c * n + L.c8map[i]] = src[(size_t)c * n8 + i];
    return out;
}
template <class T>
void cells_from_lib(const Level& L, const T* src, T* dst, int comps) {
    const size_t n8 = L.c8map.size(), n = L.ncell_host;
    for (int c = 0; c < comps; ++c)
        for (size_t i = 0; i < n8; ++i) dst[(size_t)c * n8 + i] = src[(size_t)c * n + L.c8map[i]];
}

}  // namespace

extern "C" {

int oal_abi_version(void) { return OAL_ABI_VERSION; }

const char* oal_strerror(int status) {
    switch (status) {
        case OAL_OK: return "ok";
        case OAL_ERR_ARG: return "invalid argument";
        case OAL_ERR_HIP: return "HIP runtime error";
        case OAL_ERR_NOMEM: return "device allocation failed";
        case OAL_ERR_UNSUPPORTED: return "unsupported feature";
        case OAL_ERR_STATE: return "call out of order";
        case OAL_ERR_NONFINITE: return "stability guard tripped";
        default: return "unknown status";
    }
}

const char* oal_last_error(void) { return g_last_error.c_str(); }

int oal_plan_create(const OalPlanDesc* desc, OalPlan** out) {
    if (!desc || !out) return fail(OAL_ERR_ARG, "NULL descriptor or output pointer");
    *out = nullptr;
    if (desc->abi_version != OAL_ABI_VERSION) return fail(OAL_ERR_ARG, "ABI version mismatch");
    if (desc->storage < OAL_F32 || desc->storage > OAL_BF16S) return fail(OAL_ERR_ARG, "unknown storage");
    if (desc->block_size != BS && desc->block_size != BS8) return fail(OAL_ERR_UNSUPPORTED, "block_size must be 16 or 8");
    if (desc->block_size == BS8 && desc->n_levels != 1)
        return fail(OAL_ERR_UNSUPPORTED, "8^3 blocks: one level per plan (refined plans take 16^3 blocks)");
    if (desc->stream_mode != OAL_STREAM_AB && desc->stream_mode != OAL_STREAM_AA) return fail(OAL_ERR_ARG, "unknown stream_mode");
    if (desc->bb_mode != OAL_BB_REFERENCE && desc->bb_mode != OAL_BB_HALFWAY) return fail(OAL_ERR_ARG, "unknown bb_mode");
    if (desc->stream_mode == OAL_STREAM_AA && desc->bb_mode != OAL_BB_HALFWAY)
        return fail(OAL_ERR_UNSUPPORTED, "the AA pattern needs halfway bounce-back");
    if (desc->n_levels < 1 || desc->n_levels > 12) return fail(OAL_ERR_ARG, "n_levels must be 1 to 12");
    if (!desc->levels) return fail(OAL_ERR_ARG, "no level descriptors");
    const OalCollision& c = desc->collision;
    const bool ml = desc->n_levels > 1;
    if (ml && desc->stream_mode != OAL_STREAM_AB) return fail(OAL_ERR_UNSUPPORTED, "refined plans use the AB pattern");
    if (c.transverse_freestream != 0 && c.transverse_freestream != 1) return fail(OAL_ERR_ARG, "transverse_freestream must be 0 or 1");
    if (!(c.transverse_freestream_blend >= 0.0f && c.transverse_freestream_blend <= 1.0f))
        return fail(OAL_ERR_ARG, "transverse_freestream_blend must lie in [0, 1]");
    if (ml)
        for (int l = 0; l < desc->n_levels; ++l)
            for (int i = 0; i < desc->levels[l].n_domains && desc->levels[l].domains; ++i)
                if (desc->levels[l].domains[i].zwall_bc == 1)
                    return fail(OAL_ERR_UNSUPPORTED, "zwall_bc 1 on a refined plan (the image cell may lie on the parent level)");
    if (c.limiter < 0 || c.limiter > 2) return fail(OAL_ERR_ARG, "unknown limiter");
    if (desc->rounding != OAL_ROUND_STOCHASTIC && desc->rounding != OAL_ROUND_NEAREST)
        return fail(OAL_ERR_ARG, "rounding must be OAL_ROUND_STOCHASTIC or OAL_ROUND_NEAREST");

    OalPlan* p = new (std::nothrow) OalPlan();
    if (!p) return fail(OAL_ERR_NOMEM, "host allocation failed");
    p->device = desc->device;
    p->storage = desc->storage;
    p->pair_kernel = desc->storage == OAL_F32 ? 1 : 2;   // (the measured faster split per storage form)
    p->stream_mode = desc->stream_mode;
    p->bb_mode = desc->bb_mode;
    p->wale = c.c_wale != 0.0f;
    p->cp.omega_bulk = c.omega_bulk;
    p->cp.omega_3 = c.omega_3;
    p->cp.omega_4 = c.omega_4;
    p->cp.lambda = c.lambda;
    p->cp.adaptive_omega4 = c.adaptive_omega4;
    p->cp.limiter = c.limiter;
    p->cp.compressible = c.compressible;
    p->cp.bb_mode = desc->bb_mode;
    p->cp.c_wale = c.c_wale;
    p->cp.nu_sgs_bg = c.nu_sgs_background;
    p->cp.inlet_turb = c.inlet_turbulence;
    p->cp.dither = desc->rounding == OAL_ROUND_STOCHASTIC ? 1 : 0;
    p->cp.unit = (c.omega_bulk == 1.0f && c.omega_3 == 1.0f && c.omega_4 == 1.0f) ? 1 : 0;
    p->cp.tfs = c.transverse_freestream;
    p->cp.tfs_blend = c.transverse_freestream_blend;

    auto bail = [&](int rc) {
        oal_plan_destroy(p);
        return rc;
    };
    p->levels.resize(desc->n_levels);
    p->multilevel = ml;
    std::vector<HostLevel> host(desc->n_levels);
    std::vector<OctLevel> oct;
    std::vector<XferHost> fillX(desc->n_levels), restX(desc->n_levels), restFull(desc->n_levels);
    if (ml) {
        if (int rc = build_octree(desc, oct)) return bail(rc);
    }
    for (int l = 0; l < desc->n_levels; ++l) {
        OalLevelDesc d = desc->levels[l];
        std::vector<uint8_t> inactive;
        Regroup8 r8;
        if (desc->block_size == BS8) {
            if (int rc = regroup8(d, r8)) return bail(rc);
            d.n_blocks = (int32_t)r8.mx.size();
            d.map_x = r8.mx.data(); d.map_y = r8.my.data(); d.map_z = r8.mz.data();
            d.neighbor_table = r8.nbr.data();
            d.block_domain = d.block_domain ? r8.bdom.data() : nullptr;
            d.obstacle = d.obstacle ? r8.obst.data() : nullptr;
        }
        if (ml) {
            const OctLevel& o = oct[l];
            d.n_blocks = o.nb();
            d.map_x = o.mx.data(); d.map_y = o.my.data(); d.map_z = o.mz.data();
            d.neighbor_table = o.nbr.data();
            d.obstacle = o.solid.data();
            inactive.resize(o.nb());
            for (int b = 0; b < o.nb(); ++b) inactive[b] = o.status[b] != OCT_STEPPED;
        }
        // the pipelined kernels cover FP32-arithmetic storage
        const bool pipe_ok = desc->storage != OAL_F64;
        const int rc = prepare_level(d, desc->collision, desc->stream_mode == OAL_STREAM_AB, pipe_ok,
                                     storage_bytes(desc->storage), p->levels[l],
                                     host[l], ml ? &inactive : nullptr);
        if (rc) return bail(rc);
        if (desc->block_size == BS8) {
            Level& L = p->levels[l];
            const OalLevelDesc& d8 = desc->levels[l];
            L.c8map = std::move(r8.cmap);
            L.map8_h.resize(3 * (size_t)d8.n_blocks);
            for (int b = 0; b < d8.n_blocks; ++b) {
                L.map8_h[3 * b] = d8.map_x[b]; L.map8_h[3 * b + 1] = d8.map_y[b]; L.map8_h[3 * b + 2] = d8.map_z[b];
            }
            if (d8.obstacle) L.obst8_h.assign(d8.obstacle, d8.obstacle + (size_t)d8.n_blocks * BS8_3);
            for (auto& v : L.obst8_h) v = v ? 1 : 0;
        }
        if (ml) {
            Level& L = p->levels[l];
            const OctLevel& o = oct[l];
            L.n_caller_blocks = (uint32_t)o.n_caller;
            L.ncell_host = (uint32_t)o.n_caller * BS3;
            L.n_ghost_blocks = (uint32_t)(o.nb() - o.n_caller);
            L.n_covered_blocks = 0;
            for (int b = 0; b < o.n_caller; ++b) L.n_covered_blocks += o.status[b] == OCT_COVERED;
        }
    }
    if (ml) {
        const int nl = desc->n_levels;
        int64_t ncub = 0;
        for (int l = 1; l < nl; ++l) build_fill(oct[l], oct[l - 1], fillX[l], &ncub);
        for (int l = 0; l + 1 < nl; ++l) build_restrict(oct[l], oct[l + 1], restX[l], &ncub);
        // covered cells the stepping reads, coarsest level first: next to a stepped cell, donors of the
        // finer level's fill, donors of the coarser level's (reduced) restriction
        for (int l = 0; l + 1 < nl; ++l) {
            const OctLevel& o = oct[l];
            std::vector<char> need((size_t)o.nb() * BS3, 0);
            for (int b = 0; b < o.nb(); ++b) {
                if (o.status[b] != OCT_COVERED) continue;
                const int bx = (o.mx[b] - 1) * BS, by = (o.my[b] - 1) * BS, bz = (o.mz[b] - 1) * BS;
                for (int z = 0; z < BS; ++z)
                    for (int y = 0; y < BS; ++y)
                        for (int x = 0; x < BS; ++x) {
                            if (x > 0 && x < BS - 1 && y > 0 && y < BS - 1 && z > 0 && z < BS - 1) continue;
                            bool hit = false;
                            for (int k = 0; k < 27 && !hit; ++k) {
                                if (k == 13) continue;
                                int32_t nb = -1;
                                const int64_t c = o.cell(bx + x + k % 3 - 1, by + y + (k / 3) % 3 - 1, bz + z + k / 9 - 1, &nb);
                                hit = c >= 0 && o.status[nb] == OCT_STEPPED;
                            }
                            if (hit) need[(size_t)b * BS3 + (size_t)z * BS2 + (size_t)y * BS + x] = 1;
                        }
            }
            for (uint32_t e : fillX[l + 1].src)
                if (e >> 31) need[e & 0x7fffffffu] = 1;
            if (l > 0)
                for (uint32_t e : restX[l - 1].src) {
                    const uint32_t c = e & 0x7fffffffu;
                    if (o.status[c / BS3] == OCT_COVERED) need[c] = 1;
                }
            restFull[l] = restX[l];
            std::vector<char> keep(restX[l].dst.size());
            for (size_t i = 0; i < keep.size(); ++i) keep[i] = need[restX[l].dst[i]];
            restX[l] = xfer_subset(restFull[l], keep);
        }
        // every index the transfer kernels follow stays inside its buffer
        for (int l = 0; l < nl; ++l) {
            const uint64_t here = (uint64_t)oct[l].nb() * BS3;
            if (l > 0 && !xfer_in_bounds(fillX[l], here, (uint64_t)oct[l - 1].nb() * BS3))
                return bail(fail(OAL_ERR_ARG, "fill transfer list out of bounds"));
            if (l + 1 < nl && (!xfer_in_bounds(restX[l], here, (uint64_t)oct[l + 1].nb() * BS3) ||
                               !xfer_in_bounds(restFull[l], here, (uint64_t)oct[l + 1].nb() * BS3)))
                return bail(fail(OAL_ERR_ARG, "restriction transfer list out of bounds"));
        }
        // rescale of the non-equilibrium moments between levels (coupling.jl:134-178, 188-193): tau of
        // each level in its own lattice units, 1/(1 - omega) of the sending level (0 near omega = 1), with
        // each level's molecular rate 1/tau (the subgrid viscosity, background or WALE, left out: the
        // octree's rescale is between the levels' relaxation times)
        const bool dbl = p->storage == OAL_F64;
        std::vector<float> om(nl);
        std::vector<double> om_d(nl);
        for (int l = 0; l < nl; ++l) {
            const DevDomain& d = p->levels[l].doms_h[0];
            float o4;
            double o4d;
            cell_rates<float>(0.0f, d.tau, 1.0f, 0.0f, 0, 0.0f, 0.0f, om[l], o4);
            cell_rates<double>(0.0, (double)d.tau, 1.0, 0.0, 0, 0.0, 0.0, om_d[l], o4d);
        }
        for (int l = 0; l < nl; ++l) {
            const DevDomain& d = p->levels[l].doms_h[0];
            if (l > 0) {
                const DevDomain& dc = p->levels[l - 1].doms_h[0];
                if (dbl) {
                    const double pc = std::fabs(1.0 - om_d[l - 1]) > 0.05 ? 1.0 / (1.0 - om_d[l - 1]) : 0.0;
                    p->levels[l].fill_scale = ((double)d.tau / (2.0 * (double)dc.tau)) * pc;
                    p->levels[l].fill_keep = 1.0 - om_d[l];
                } else {
                    const float pc = std::fabs(1.0f - om[l - 1]) > 0.05f ? 1.0f / (1.0f - om[l - 1]) : 0.0f;
                    p->levels[l].fill_scale = (double)((d.tau / (2.0f * dc.tau)) * pc);
                    p->levels[l].fill_keep = (double)(1.0f - om[l]);
                }
            }
            if (l + 1 < nl) {
                const DevDomain& df = p->levels[l + 1].doms_h[0];
                if (dbl) {
                    const double pf = std::fabs(1.0 - om_d[l + 1]) > 0.05 ? 1.0 / (1.0 - om_d[l + 1]) : 0.0;
                    p->levels[l].restrict_scale = ((2.0 * (double)d.tau) / (double)df.tau) * pf;
                    p->levels[l].restrict_keep = 1.0 - om_d[l];
                } else {
                    const float pf = std::fabs(1.0f - om[l + 1]) > 0.05f ? 1.0f / (1.0f - om[l + 1]) : 0.0f;
                    p->levels[l].restrict_scale = (double)(((2.0f * d.tau) / df.tau) * pf);
                    p->levels[l].restrict_keep = (double)(1.0f - om[l]);
                }
            }
        }
    }
    hipError_t e = hipSetDevice(p->device);
    if (e != hipSuccess) return bail(fail(OAL_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)));
    if (int rc = scratch_init(p->device)) return bail(rc);
    if ((e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&p->ev0)) != hipSuccess || (e = hipEventCreate(&p->ev1)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&p->join, hipEventDisableTiming)) != hipSuccess ||
        (e = hipMalloc((void**)&p->ts_dev, 256)) != hipSuccess)
        return bail(fail(OAL_ERR_HIP, std::string("stream/event creation: ") + hipGetErrorString(e)));
    p->graph = ml;
    for (int l = 0; l < desc->n_levels; ++l) {
        const int rc = upload_level(p, p->levels[l], host[l]);
        if (rc) return bail(rc);
        if (ml) {
            Level& L = p->levels[l];
            // the rebuild kernel reads each written cell's 64 stencil nodes as one row: cell-major
            for (XferHost* X : {&fillX[l], &restX[l], &restFull[l]}) {
                X->slot = node_to_cell_major(X->slot, X->dst.size());
                X->w = node_to_cell_major(X->w, X->dst.size());
            }
            const XferHost& F = fillX[l];
            const XferHost& R = restX[l];
            const XferHost& RF = restFull[l];
            int r2;
            if ((r2 = upload(L.g_src, F.src)) || (r2 = upload(L.g_dst, F.dst)) || (r2 = upload(L.g_slot, F.slot)) ||
                (r2 = upload(L.g_w, F.w)) || (r2 = upload(L.r_src, R.src)) || (r2 = upload(L.r_dst, R.dst)) ||
                (r2 = upload(L.r_slot, R.slot)) || (r2 = upload(L.r_w, R.w)) || (r2 = upload(L.rf_src, RF.src)) ||
                (r2 = upload(L.rf_dst, RF.dst)) || (r2 = upload(L.rf_slot, RF.slot)) || (r2 = upload(L.rf_w, RF.w)))
                return bail(r2);
            L.n_gsrc = (uint32_t)F.src.size(); L.n_gdst = (uint32_t)F.dst.size();
            L.n_rsrc = (uint32_t)R.src.size(); L.n_rdst = (uint32_t)R.dst.size();
            L.n_rfsrc = (uint32_t)RF.src.size(); L.n_rfdst = (uint32_t)RF.dst.size();
            if (L.g_m.alloc((size_t)L.n_gsrc * XFER_STRIDE * work_bytes(p->storage)) != hipSuccess ||
                L.r_m.alloc((size_t)std::max(L.n_rsrc, L.n_rfsrc) * XFER_STRIDE * work_bytes(p->storage)) != hipSuccess)
                return bail(fail(OAL_ERR_NOMEM, "transfer buffer allocation failed"));
        }
    }
    *out = p;
    return OAL_OK;
}

int oal_plan_destroy(OalPlan* p) {
    if (!p) return OAL_OK;
    if (p->stream) {
        (void)hipSetDevice(p->device);
        (void)hipStreamSynchronize(p->stream);
    }
    drop_graphs(p);
    if (p->ts_dev) (void)hipFree(p->ts_dev);
    for (auto& L : p->levels) L.release();
    if (p->side) (void)hipStreamSynchronize(p->side);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    if (p->fork) (void)hipEventDestroy(p->fork);
    if (p->join) (void)hipEventDestroy(p->join);
    for (hipEvent_t e : p->bulk_ev) (void)hipEventDestroy(e);
    if (p->side) (void)hipStreamDestroy(p->side);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return OAL_OK;
}

int oal_level_cells(const OalPlan* p, int level, int64_t* n) {
    if (int rc = check_level(p, level)) return rc;
    if (!n) return fail(OAL_ERR_ARG, "NULL output");
    const Level& L = p->levels[level];
    *n = L.c8map.empty() ? (int64_t)L.ncell_host : (int64_t)L.c8map.size();
    return OAL_OK;
}

int oal_level_info(const OalPlan* p, int level, OalLevelInfo* info) {
    if (int rc = check_level(p, level)) return rc;
    if (!info) return fail(OAL_ERR_ARG, "NULL output");
    const Level& L = p->levels[level];
    *info = OalLevelInfo{};
    info->n_cells = L.c8map.empty() ? (int64_t)L.ncell_host : (int64_t)L.c8map.size();
    info->n_fluid_cells = L.n_fluid;
    info->n_blocks = L.c8map.empty() ? (int32_t)L.n_caller_blocks : (int32_t)(L.map8_h.size() / 3);
    info->n_ghost_blocks = (int32_t)L.n_ghost_blocks;
    info->n_covered_blocks = (int32_t)L.n_covered_blocks;
    info->n_regular = (int32_t)L.n_regular;
    info->n_partial = (int32_t)L.n_partial;
    info->n_sponge = (int32_t)L.n_sponge;
    info->n_solid_regular = L.pipe_ok ? (int32_t)L.n_br_solid : 0;
    info->n_interior = (int32_t)L.n_interior;
    info->n_boundary_regular = (int32_t)L.n_boundary_regular;
    info->n_boundary = (int32_t)L.n_boundary;
    info->n_face_cells = (int32_t)L.n_face;
    return OAL_OK;
}

static int upload_f_lib(OalPlan* p, int level, const void* host_f, int dt) {
    if (int rc = check_level(p, level)) return rc;
    if (!host_f || (dt != OAL_F32 && dt != OAL_F64)) return fail(OAL_ERR_ARG, "host_f must be OAL_F32 or OAL_F64 data");
    OAL_HIP(hipSetDevice(p->device));
    Level& L = p->levels[level];
    const size_t bytes = (size_t)L.ncell_host * 27 * (dt == OAL_F64 ? 8 : 4);
    void* stage = nullptr;
    OAL_HIP(hipMalloc(&stage, bytes));
    hipError_t e = hipMemcpyAsync(stage, host_f, bytes, hipMemcpyHostToDevice, p->stream);
    // ghost cells of a refined level start at rest (those no fill reaches keep that state)
    if (e == hipSuccess && L.ncell > L.ncell_host)
        e = launch_init_eq(p->storage, L.fA.p, L.ncell, 0, nullptr, nullptr, 1.0, 0.0, 0.0, 0.0, permute_of(p), nullptr,
                           p->stream);
    if (e == hipSuccess) e = launch_convert_in(p->storage, L.fA.p, stage, dt, L.ncell, L.ncell_host, permute_of(p), p->stream);
    if (e == hipSuccess && L.fB.p) e = hipMemcpyAsync(L.fB.p, L.fA.p, L.fA.n, hipMemcpyDeviceToDevice, p->stream);
    if (e == hipSuccess) e = aa_prep(p, L);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(OAL_ERR_HIP, std::string("oal_upload_f: ") + hipGetErrorString(e));
    L.have_f = true;
    L.aa_phase = 0;
    L.force_valid = false;
    p->coupled = false;
    return refresh_velocity(p, L);
}

static int download_f_lib(OalPlan* p, int level, void* host_f, int dt) {
    if (int rc = check_level(p, level)) return rc;
    if (!host_f || (dt != OAL_F32 && dt != OAL_F64)) return fail(OAL_ERR_ARG, "host_f must be OAL_F32 or OAL_F64 data");
    Level& L = p->levels[level];
    if (int rc = require_plain_state(p, L)) return rc;
    OAL_HIP(hipSetDevice(p->device));
    if (int rc = ml_sync_covered(p, p->stream)) return rc;
    const size_t bytes = (size_t)L.ncell_host * 27 * (dt == OAL_F64 ? 8 : 4);
    void* stage = nullptr;
    OAL_HIP(hipMalloc(&stage, bytes));
    hipError_t e = launch_convert_out(p->storage, L.fA.p, stage, dt, L.ncell, L.ncell_host, permute_of(p), p->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host_f, stage, bytes, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(OAL_ERR_HIP, std::string("oal_download_f: ") + hipGetErrorString(e));
    return OAL_OK;
}

static int init_equilibrium_lib(OalPlan* p, int level, const double* rho, const double* u, double rho0, const double u0[3]) {
    if (int rc = check_level(p, level)) return rc;
    OAL_HIP(hipSetDevice(p->device));
    Level& L = p->levels[level];
    double* drho = nullptr;
    double* du = nullptr;
    hipError_t e = hipSuccess;
    if (rho) {
        e = hipMalloc((void**)&drho, (size_t)L.ncell_host * 8);
        if (e == hipSuccess) e = hipMemcpyAsync(drho, rho, (size_t)L.ncell_host * 8, hipMemcpyHostToDevice, p->stream);
    }
    if (e == hipSuccess && u) {
        e = hipMalloc((void**)&du, (size_t)L.ncell_host * 24);
        if (e == hipSuccess) e = hipMemcpyAsync(du, u, (size_t)L.ncell_host * 24, hipMemcpyHostToDevice, p->stream);
    }
    const double ux0 = u0 ? u0[0] : 0.0, uy0 = u0 ? u0[1] : 0.0, uz0 = u0 ? u0[2] : 0.0;
    if (e == hipSuccess)
        e = launch_init_eq(p->storage, L.fA.p, L.ncell, L.ncell_host, drho, du, rho0, ux0, uy0, uz0, permute_of(p),
                           p->wale ? (void*)L.velA.p : nullptr, p->stream);
    if (e == hipSuccess && L.fB.p) e = hipMemcpyAsync(L.fB.p, L.fA.p, L.fA.n, hipMemcpyDeviceToDevice, p->stream);
    if (e == hipSuccess) e = aa_prep(p, L);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (drho) (void)hipFree(drho);
    if (du) (void)hipFree(du);
    if (e != hipSuccess) return fail(OAL_ERR_HIP, std::string("oal_init_equilibrium: ") + hipGetErrorString(e));
    L.have_f = true;
    L.aa_phase = 0;
    L.force_valid = false;
    p->coupled = false;
    return OAL_OK;
}

int oal_level_step(OalPlan* p, int level, int32_t time_step, int flags, void* stream) {
    if (int rc = check_level(p, level)) return rc;
    OAL_HIP(hipSetDevice(p->device));
    return level_step(p, level, time_step, flags, pick_stream(p, stream));
}

int oal_step(OalPlan* p, int n_steps, int32_t first_time_step, int flags, void* stream) {
    if (!p) return fail(OAL_ERR_ARG, "plan is NULL");
    if (n_steps < 0) return fail(OAL_ERR_ARG, "negative step count");
    OAL_HIP(hipSetDevice(p->device));
    hipStream_t s = pick_stream(p, stream);
    if (p->multilevel && !p->coupled)
        if (int rc = ml_couple(p, s)) return rc;
    for (int i = 0; i < n_steps; ++i) {
        const bool last = i == n_steps - 1;
        const int f = last ? flags : 0;
        if (p->multilevel) p->covered_stale = true;
        if (p->graph && s == p->stream && !(f & OAL_STEP_FORCE)) {
            for (const Level& L : p->levels)
                if (!L.have_f) return fail(OAL_ERR_STATE, "no populations: call oal_upload_f or oal_init_equilibrium first");
            if (int rc = replay_step(p, first_time_step + i)) return rc;
        } else if (p->multilevel) {
            if (int rc = ml_step_rec(p, 0, first_time_step + i, flags, last, s)) return rc;
        } else {
            if (int rc = level_step(p, 0, first_time_step + i, f, s)) return rc;
        }
    }
    return OAL_OK;
}

int oal_force(OalPlan* p, int level, double* force) {
    if (int rc = check_level(p, level)) return rc;
    if (!force) return fail(OAL_ERR_ARG, "NULL output");
    Level& L = p->levels[level];
    if (!L.force_valid) return fail(OAL_ERR_STATE, "the last step did not collect the force (OAL_STEP_FORCE)");
    OAL_HIP(hipSetDevice(p->device));
    std::vector<double> h((size_t)L.n_domains * 4);
    OAL_HIP(hipStreamSynchronize(p->stream));
    OAL_HIP(hipMemcpy(h.data(), L.red_out.p, h.size() * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < L.n_domains; ++i)
        for (int a = 0; a < 3; ++a) force[3 * i + a] = h[4 * (size_t)i + a];
    return OAL_OK;
}

int oal_mass(OalPlan* p, int level, double* out, void* stream) {
    if (int rc = check_level(p, level)) return rc;
    if (!out) return fail(OAL_ERR_ARG, "NULL output");
    Level& L = p->levels[level];
    if (int rc = require_plain_state(p, L)) return rc;
    OAL_HIP(hipSetDevice(p->device));
    hipStream_t s = pick_stream(p, stream);
    if (int rc = ml_sync_covered(p, s)) return rc;
    OAL_HIP(launch_mass(p->storage, L.fA.p, L.solid.p, L.borg.p, L.bdom.p, L.bflags.p, L.doms.p, L.nb, permute_of(p), L.mpart.p, s));
    // the reduction walks blocks here: one "item" per block
    ReduceArgs r{L.mpart.p, L.dom_block0.p, L.red_out.p, L.n_domains, 4};
    OAL_HIP(launch_reduce(r, s));
    OAL_HIP(hipStreamSynchronize(s));
    OAL_HIP(hipMemcpy(out, L.red_out.p, (size_t)L.n_domains * 32, hipMemcpyDeviceToHost));
    L.force_valid = false;   // red_out reused
    return OAL_OK;
}

int oal_mass_noalloc(OalPlan* p, int level, double* out_dev, void* stream) {
    if (int rc = check_level(p, level)) return rc;
    if (!out_dev) return fail(OAL_ERR_ARG, "NULL output");
    Level& L = p->levels[level];
    if (int rc = require_plain_state(p, L)) return rc;
    if (L.nb > NOALLOC_MAX_BLOCKS)
        return fail(OAL_ERR_UNSUPPORTED, "oal_mass_noalloc: the level has " + std::to_string(L.nb) +
                                             " blocks, its scratch holds the partials of at most " +
                                             std::to_string(NOALLOC_MAX_BLOCKS));
    OAL_HIP(hipSetDevice(p->device));
    double* part = scratch_addr(p->device);
    if (!part) return fail(OAL_ERR_STATE, "oal_mass_noalloc: no scratch address for this device");
    hipStream_t s = pick_stream(p, stream);
    if (int rc = ml_sync_covered(p, s)) return rc;
    // every block writes its own partials (zero for blocks not counted), so nothing is cleared
    OAL_HIP(launch_mass(p->storage, L.fA.p, L.solid.p, L.borg.p, L.bdom.p, L.bflags.p, L.doms.p, L.nb, permute_of(p), part, s));
    ReduceArgs r{part, L.dom_block0.p, out_dev, L.n_domains, 4};
    OAL_HIP(launch_reduce(r, s));
    return OAL_OK;
}

static int download_macro_lib(OalPlan* p, int level, double* rho, double* u) {
    if (int rc = check_level(p, level)) return rc;
    Level& L = p->levels[level];
    if (int rc = require_plain_state(p, L)) return rc;
    OAL_HIP(hipSetDevice(p->device));
    if (int rc = ml_sync_covered(p, p->stream)) return rc;
    double* d = nullptr;
    OAL_HIP(hipMalloc((void**)&d, (size_t)L.ncell_host * 32));
    hipError_t e = launch_macro(p->storage, L.fA.p, d, d + L.ncell_host, L.ncell, L.ncell_host, permute_of(p), nullptr, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e == hipSuccess && rho) e = hipMemcpy(rho, d, (size_t)L.ncell_host * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && u) e = hipMemcpy(u, d + L.ncell_host, (size_t)L.ncell_host * 24, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(OAL_ERR_HIP, std::string("oal_download_macro: ") + hipGetErrorString(e));
    return OAL_OK;
}

int oal_guard(OalPlan* p, int level, float rho_min, float u_max, int64_t* bad) {
    if (int rc = check_level(p, level)) return rc;
    Level& L = p->levels[level];
    if (!L.have_f) return fail(OAL_ERR_STATE, "no populations on this level");
    OAL_HIP(hipSetDevice(p->device));
    if (int rc = ml_sync_covered(p, p->stream)) return rc;
    unsigned long long h = 0;
    OAL_HIP(hipMemsetAsync(L.guard.p, 0, sizeof(unsigned long long), p->stream));
    OAL_HIP(launch_guard(p->storage, L.fA.p, L.solid.p, L.borg.p, L.bdom.p, L.bflags.p, L.doms.p, L.nb, rho_min, u_max, L.guard.p,
                         p->stream));
    OAL_HIP(hipStreamSynchronize(p->stream));
    OAL_HIP(hipMemcpy(&h, L.guard.p, sizeof(h), hipMemcpyDeviceToHost));
    if (bad) *bad = (int64_t)h;
    return h ? fail(OAL_ERR_NONFINITE, "stability guard: " + std::to_string(h) + " bad cells") : OAL_OK;
}

int oal_sync(OalPlan* p) {
    if (!p) return fail(OAL_ERR_ARG, "plan is NULL");
    OAL_HIP(hipSetDevice(p->device));
    OAL_HIP(hipStreamSynchronize(p->stream));
    return OAL_OK;
}

int oal_time_steps(OalPlan* p, int level, int n, int flags, double* ms_total, double* ms_kernel) {
    if (p && level == OAL_ALL_LEVELS) {
        if (n <= 0) return fail(OAL_ERR_ARG, "n must be positive");
        OAL_HIP(hipSetDevice(p->device));
        if (p->multilevel && !p->coupled)
            if (int rc = ml_couple(p, p->stream)) return rc;
        OAL_HIP(hipStreamSynchronize(p->stream));
        OAL_HIP(hipEventRecord(p->ev0, p->stream));
        if (int rc = oal_step(p, n, 0, flags, nullptr)) return rc;
        OAL_HIP(hipEventRecord(p->ev1, p->stream));
        OAL_HIP(hipEventSynchronize(p->ev1));
        float ms = 0.0f;
        OAL_HIP(hipEventElapsedTime(&ms, p->ev0, p->ev1));
        if (ms_total) *ms_total = ms;
        if (ms_kernel) *ms_kernel = (double)ms / n;
        return OAL_OK;
    }
    if (int rc = check_level(p, level)) return rc;
    if (n <= 0) return fail(OAL_ERR_ARG, "n must be positive");
    OAL_HIP(hipSetDevice(p->device));
    OAL_HIP(hipStreamSynchronize(p->stream));
    const Level& L = p->levels[level];
    const bool bulk = L.n_regular + L.n_interior > 0;
    if (bulk) {
        while (p->bulk_ev.size() < 2 * (size_t)n) {
            hipEvent_t e = nullptr;
            OAL_HIP(hipEventCreate(&e));
            p->bulk_ev.push_back(e);
        }
    }
    OAL_HIP(hipEventRecord(p->ev0, p->stream));
    for (int i = 0; i < n; ++i) {
        hipEvent_t b0 = bulk ? p->bulk_ev[2 * i] : nullptr, b1 = bulk ? p->bulk_ev[2 * i + 1] : nullptr;
        if (int rc = level_step(p, level, i, flags, p->stream, b0, b1)) return rc;
    }
    OAL_HIP(hipEventRecord(p->ev1, p->stream));
    OAL_HIP(hipEventSynchronize(p->ev1));
    float ms = 0.0f;
    OAL_HIP(hipEventElapsedTime(&ms, p->ev0, p->ev1));
    double kern = 0.0;
    for (int i = 0; bulk && i < n; ++i) {
        float d = 0.0f;
        OAL_HIP(hipEventElapsedTime(&d, p->bulk_ev[2 * i], p->bulk_ev[2 * i + 1]));
        kern += d;
    }
    if (ms_total) *ms_total = ms;
    if (ms_kernel) *ms_kernel = bulk ? kern / n : (double)ms / n;
    return OAL_OK;
}

int oal_copy_bandwidth(int device, int64_t bytes, int reps, double* gbs) {
    if (bytes < (1 << 20) || reps <= 0 || !gbs) return fail(OAL_ERR_ARG, "bytes >= 1 MiB, reps > 0 and an output are needed");
    OAL_HIP(hipSetDevice(device));
    const size_t nb = (size_t)bytes & ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, nb);
    if (e == hipSuccess) e = hipMalloc(&b, nb);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, nb, s);
    if (e == hipSuccess) e = launch_copy16(a, b, nb, s);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch_copy16(a, b, nb, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(OAL_ERR_HIP, std::string("oal_copy_bandwidth: ") + hipGetErrorString(e));
    *gbs = 2.0 * (double)nb * reps / (ms * 1e-3) / 1e9;
    return OAL_OK;
}

static int set_bouzidi_lib(OalPlan* p, int level, const uint16_t* q_map, int32_t ff_fallback, float q_min) {
    if (p && level >= 0 && level < (int)p->levels.size() && p->levels[level].has_sponge && q_map)
        return fail(OAL_ERR_UNSUPPORTED, "interpolated bounce-back on a level with a sponge is not supported");
    if (int rc = check_level(p, level)) return rc;
    if (p->stream_mode != OAL_STREAM_AB || p->bb_mode != OAL_BB_HALFWAY)
        return fail(OAL_ERR_ARG, "interpolated bounce-back needs the AB pattern and halfway bounce-back");
    if (ff_fallback < OAL_BFL_FALLBACK_HALFWAY || ff_fallback > OAL_BFL_FALLBACK_BOUNCE_BACK)
        return fail(OAL_ERR_ARG, "ff_fallback must be OAL_BFL_FALLBACK_HALFWAY, _WALL_EQ or _BOUNCE_BACK");
    if (!(q_min >= 0.0f && q_min < 1.0f)) return fail(OAL_ERR_ARG, "q_min must lie in [0, 1)");
    Level& L = p->levels[level];
    if (q_map && !L.has_solid) return fail(OAL_ERR_STATE, "the level has no obstacle cells");
    OAL_HIP(hipSetDevice(p->device));
    OAL_HIP(hipStreamSynchronize(p->stream));
    drop_graphs(p);
    L.qmap.release();
    L.bfl_fallback = ff_fallback;
    L.q_min = q_min;
    if (!q_map || L.n_lslots == 0) return OAL_OK;
    // the caller's q-map [x,y,z,b,k] regrouped per link-mask slot (blocks with solid cells near);
    // ghost blocks of a refined level (not stepped) keep q = 0
    std::vector<uint16_t> h((size_t)L.n_lslots * 27 * BS3, 0);
    for (uint32_t b = 0; b < L.n_caller_blocks; ++b) {
        const int32_t ls = L.lslot_h[b];
        if (ls < 0) continue;
        for (int k = 0; k < 27; ++k)
            std::memcpy(h.data() + ((size_t)ls * 27 + k) * BS3, q_map + (size_t)k * L.ncell_host + (size_t)b * BS3,
                        BS3 * sizeof(uint16_t));
    }
    OAL_HIP(L.qmap.alloc(h.size()));
    OAL_HIP(hipMemcpy(L.qmap.p, h.data(), h.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    return OAL_OK;
}

static int set_sponge_lib(OalPlan* p, int level, const float* sponge, int32_t blend_distributions) {
    if (int rc = check_level(p, level)) return rc;
    if (blend_distributions != 0 && blend_distributions != 1) return fail(OAL_ERR_ARG, "blend_distributions must be 0 or 1");
    Level& L = p->levels[level];
    if (sponge && p->stream_mode != OAL_STREAM_AB) return fail(OAL_ERR_UNSUPPORTED, "the sponge needs the AB pattern");
    if (sponge && L.qmap.p) return fail(OAL_ERR_UNSUPPORTED, "a sponge on a level with interpolated bounce-back is not supported");
    const uint32_t nhost = L.ncell_host;
    if (sponge) {
        for (uint32_t i = 0; i < nhost; ++i)
            if (!(sponge[i] >= 0.0f && sponge[i] <= 1.0f)) return fail(OAL_ERR_ARG, "sponge strengths must lie in [0, 1]");
    }
    // blocks with a nonzero strength: flagged, each with a slot of BS3 strengths
    std::vector<int32_t> slot(L.nb, -1);
    std::vector<float> vals;
    for (uint32_t b = 0; b < L.nb; ++b) {
        L.flags_h[b] &= ~BF_SPONGE;
        if (!sponge || b >= L.n_caller_blocks || (L.flags_h[b] & BF_INACTIVE)) continue;
        bool any = false;
        for (int c = 0; c < BS3 && !any; ++c) any = sponge[(size_t)b * BS3 + c] > 0.0f;
        if (!any) continue;
        slot[b] = (int32_t)(vals.size() / BS3);
        vals.insert(vals.end(), sponge + (size_t)b * BS3, sponge + (size_t)(b + 1) * BS3);
        L.flags_h[b] |= BF_SPONGE;
    }
    OAL_HIP(hipSetDevice(p->device));
    OAL_HIP(hipDeviceSynchronize());
    drop_graphs(p);
    HostLevel H;
    build_lists(L, H);
    auto re = [](auto& d, const auto& h) { d.release(); return upload(d, h); };
    int rc;
    if ((rc = re(L.bflags, L.flags_h)) || (rc = re(L.blist_rg, H.blist_rg)) || (rc = re(L.blist_in, H.blist_in)) ||
        (rc = re(L.blist_br, H.blist_br)) || (rc = re(L.blist_bd, H.blist_bd)) || (rc = re(L.blist_pr, H.blist_pr)) || (rc = re(L.blist_sp, H.blist_sp)) ||
        (rc = re(L.blist_rp, H.blist_rp)) || (rc = re(L.blist_rf, H.blist_rf)) || (rc = re(L.face_cells, H.face_cells)) ||
        (rc = re(L.sponge, vals)) || (rc = re(L.sslot, slot)))
        return rc;
    L.has_sponge = !vals.empty();
    L.sponge_blend = blend_distributions;
    return OAL_OK;
}

// ---- Surface-stress forces (src/forces/surface.jl) ---------------------------------------------------
int oal_set_surface(OalPlan* p, int level, int64_t n_tri, const float* centers, const float* normals,
                    const float* areas) {
    if (int rc = check_level(p, level)) return rc;
    Level& L = p->levels[level];
    if (n_tri < 0) return fail(OAL_ERR_ARG, "negative triangle count");
    OAL_HIP(hipSetDevice(p->device));
    if (n_tri == 0 || !centers) {
        OAL_HIP(hipDeviceSynchronize());
        L.release_surface();
        return OAL_OK;
    }
    if (!normals || !areas) return fail(OAL_ERR_ARG, "NULL normals or areas");
    if (n_tri > ((int64_t)1 << 31) - 1) return fail(OAL_ERR_UNSUPPORTED, "more than 2^31 - 1 triangles");
    if (L.n_domains != 1) return fail(OAL_ERR_UNSUPPORTED, "the surface forces need a level of one domain");
    for (int64_t i = 0; i < 3 * n_tri; ++i)
        if (!std::isfinite(centers[i]) || !std::isfinite(normals[i])) return fail(OAL_ERR_ARG, "non-finite triangle centre or normal");
    for (int64_t i = 0; i < n_tri; ++i)
        if (!std::isfinite(areas[i]) || areas[i] < 0.0f) return fail(OAL_ERR_ARG, "triangle areas must be finite and >= 0");
    // block of each block position of the domain (the reference's block_ptr), caller's blocks only
    const DevDomain& D = L.doms_h[0];
    int32_t nblk[3];
    for (int a = 0; a < 3; ++a) nblk[a] = (D.ext[a] + BS - 1) / BS;
    std::vector<int32_t> bmap((size_t)nblk[0] * nblk[1] * nblk[2], -1);
    for (uint32_t b = 0; b < L.n_caller_blocks; ++b) {
        const int bx = L.map_h[3 * b] - 1, by = L.map_h[3 * b + 1] - 1, bz = L.map_h[3 * b + 2] - 1;
        if (bx < 0 || by < 0 || bz < 0 || bx >= nblk[0] || by >= nblk[1] || bz >= nblk[2]) continue;
        bmap[((size_t)bz * nblk[1] + by) * nblk[0] + bx] = (int32_t)b;
    }
    OAL_HIP(hipDeviceSynchronize());
    L.release_surface();
    const size_t n = (size_t)n_tri, ngroups = (n + SURF_WG - 1) / SURF_WG;
    auto put = [](DevBuf<float>& d, const float* h, size_t count) -> hipError_t {
        hipError_t e = d.alloc(count);
        return e == hipSuccess ? hipMemcpy(d.p, h, count * sizeof(float), hipMemcpyHostToDevice) : e;
    };
    hipError_t e = put(L.s_cen, centers, 3 * n);
    if (e == hipSuccess) e = put(L.s_nrm, normals, 3 * n);
    if (e == hipSuccess) e = put(L.s_area, areas, n);
    if (e == hipSuccess) e = L.s_p.alloc(n);
    if (e == hipSuccess) e = L.s_shear.alloc(3 * n);
    if (e == hipSuccess) e = L.s_sign.alloc(n);
    if (e == hipSuccess) e = L.s_wd.alloc(n);
    if (e == hipSuccess) e = L.s_part.alloc(SURF_SUMS * ngroups);
    if (e == hipSuccess) e = L.s_out.alloc(SURF_SUMS);
    if (e != hipSuccess) {
        L.release_surface();
        return fail(OAL_ERR_HIP, std::string("oal_set_surface: ") + hipGetErrorString(e));
    }
    if (int rc = upload(L.s_bmap, bmap)) { L.release_surface(); return rc; }
    for (int a = 0; a < 3; ++a) L.s_nblk[a] = nblk[a];
    L.n_tri = n_tri;
    return OAL_OK;
}

int oal_surface_forces(OalPlan* p, int level, const OalSurfaceParams* sp, OalSurfaceForces* out, void* stream) {
    if (int rc = check_level(p, level)) return rc;
    if (!sp || !out) return fail(OAL_ERR_ARG, "NULL parameters or output");
    Level& L = p->levels[level];
    if (L.n_tri <= 0) return fail(OAL_ERR_STATE, "no surface mesh on this level (oal_set_surface)");
    if (int rc = require_plain_state(p, L)) return rc;
    if (!(sp->dx > 0.0 && std::isfinite(sp->dx))) return fail(OAL_ERR_ARG, "dx must be finite and > 0");
    if (sp->search_radius < 1 || sp->search_radius > 64) return fail(OAL_ERR_ARG, "search_radius must lie in [1, 64]");
    if (sp->recover_normal != 0 && sp->recover_normal != 1) return fail(OAL_ERR_ARG, "recover_normal must be 0 or 1");
    if (sp->symmetric != 0 && sp->symmetric != 1) return fail(OAL_ERR_ARG, "symmetric must be 0 or 1");
    OAL_HIP(hipSetDevice(p->device));
    hipStream_t s = pick_stream(p, stream);
    if (int rc = ml_sync_covered(p, s)) return rc;
    const DevDomain& D = L.doms_h[0];
    SurfaceArgs a{};
    a.f = L.fA.p;
    a.permute = permute_of(p);
    a.solid = L.solid.p;
    a.bmap = L.s_bmap.p;
    a.nbx = L.s_nblk[0]; a.nby = L.s_nblk[1]; a.nbz = L.s_nblk[2];
    for (int d = 0; d < 3; ++d) a.ext[d] = D.ext[d];
    a.n_tri = L.n_tri;
    a.cen = L.s_cen.p; a.nrm = L.s_nrm.p; a.area = L.s_area.p;
    a.dx = (float)sp->dx;
    for (int d = 0; d < 3; ++d) { a.off[d] = (float)sp->mesh_offset[d]; a.ref[d] = (double)(float)sp->moment_center[d]; }
    a.pressure_scale = (float)sp->pressure_scale; a.stress_scale = (float)sp->stress_scale;
    a.tau = (float)sp->tau; a.c_wale = (float)sp->c_wale; a.nu_sgs_bg = (float)sp->nu_sgs_bg; a.cs2 = (float)sp->cs2;
    a.search_radius = sp->search_radius; a.recover_normal = sp->recover_normal;
    a.p_map = L.s_p.p; a.s_map = L.s_shear.p; a.sign_map = L.s_sign.p; a.wd_map = L.s_wd.p;
    a.part = L.s_part.p; a.out = L.s_out.p;
    OAL_HIP(launch_surface(p->storage, a, s));
    double h[SURF_SUMS];
    OAL_HIP(hipMemcpyAsync(h, L.s_out.p, sizeof(h), hipMemcpyDeviceToHost, s));
    OAL_HIP(hipStreamSynchronize(s));
    L.surface_mapped = true;
    // integrate_surface_forces! (surface.jl:1485-1547): half model, totals and coefficients
    double Fp[3] = {h[0], h[1], h[2]}, Fv[3] = {h[3], h[4], h[5]}, M[3] = {h[6], h[7], h[8]};
    if (sp->symmetric) {
        Fp[0] *= 2.0; Fp[2] *= 2.0; Fv[0] *= 2.0; Fv[2] *= 2.0; M[1] *= 2.0;
        Fp[1] = 0.0; Fv[1] = 0.0; M[0] = 0.0; M[2] = 0.0;
    }
    *out = OalSurfaceForces{};
    for (int d = 0; d < 3; ++d) {
        out->F_pressure[d] = Fp[d]; out->F_viscous[d] = Fv[d]; out->F[d] = Fp[d] + Fv[d]; out->M[d] = M[d];
    }
    const double q_inf = 0.5 * sp->rho_ref * sp->u_ref * sp->u_ref;
    const double F_ref = q_inf * sp->area_ref, M_ref = F_ref * sp->chord_ref;
    if (F_ref > 1e-10) { out->Cd = out->F[0] / F_ref; out->Cl = out->F[2] / F_ref; out->Cs = out->F[1] / F_ref; }
    if (M_ref > 1e-10) for (int d = 0; d < 3; ++d) out->Cm[d] = M[d] / M_ref;
    out->n_mapped = (int64_t)h[9];
    return OAL_OK;
}

int oal_download_surface(OalPlan* p, int level, float* pressure, float* shear, float* normal_sign, float* wall_dist) {
    if (int rc = check_level(p, level)) return rc;
    Level& L = p->levels[level];
    if (!L.surface_mapped) return fail(OAL_ERR_STATE, "no surface map yet (oal_surface_forces)");
    OAL_HIP(hipSetDevice(p->device));
    const size_t n = (size_t)L.n_tri;
    if (pressure) OAL_HIP(hipMemcpy(pressure, L.s_p.p, n * 4, hipMemcpyDeviceToHost));
    if (shear) OAL_HIP(hipMemcpy(shear, L.s_shear.p, 3 * n * 4, hipMemcpyDeviceToHost));
    if (normal_sign) OAL_HIP(hipMemcpy(normal_sign, L.s_sign.p, n * 4, hipMemcpyDeviceToHost));
    if (wall_dist) OAL_HIP(hipMemcpy(wall_dist, L.s_wd.p, n * 4, hipMemcpyDeviceToHost));
    return OAL_OK;
}

// ---- Per-cell host arrays: the caller's layout (8^3 levels regrouped) ---------------------------------
int oal_upload_f(OalPlan* p, int level, const void* host_f, int dt) {
    if (int rc = check_level(p, level)) return rc;
    const Level& L = p->levels[level];
    if (L.c8map.empty() || !host_f) return upload_f_lib(p, level, host_f, dt);
    if (dt == OAL_F64) return upload_f_lib(p, level, cells_to_lib(L, (const double*)host_f, 27, 0.0).data(), dt);
    if (dt == OAL_F32) return upload_f_lib(p, level, cells_to_lib(L, (const float*)host_f, 27, 0.0f).data(), dt);
    return upload_f_lib(p, level, host_f, dt);
}

int oal_download_f(OalPlan* p, int level, void* host_f, int dt) {
    if (int rc = check_level(p, level)) return rc;
    const Level& L = p->levels[level];
    if (L.c8map.empty() || !host_f || (dt != OAL_F32 && dt != OAL_F64)) return download_f_lib(p, level, host_f, dt);
    if (dt == OAL_F64) {
        std::vector<double> t((size_t)27 * L.ncell_host);
        if (int rc = download_f_lib(p, level, t.data(), dt)) return rc;
        cells_from_lib(L, t.data(), (double*)host_f, 27);
    } else {
        std::vector<float> t((size_t)27 * L.ncell_host);
        if (int rc = download_f_lib(p, level, t.data(), dt)) return rc;
        cells_from_lib(L, t.data(), (float*)host_f, 27);
    }
    return OAL_OK;
}

int oal_init_equilibrium(OalPlan* p, int level, const double* rho, const double* u, double rho0, const double u0[3]) {
    if (int rc = check_level(p, level)) return rc;
    const Level& L = p->levels[level];
    if (L.c8map.empty()) return init_equilibrium_lib(p, level, rho, u, rho0, u0);
    std::vector<double> r, v;
    if (rho) r = cells_to_lib(L, rho, 1, rho0);
    if (u) v = cells_to_lib(L, u, 3, 0.0);
    return init_equilibrium_lib(p, level, rho ? r.data() : nullptr, u ? v.data() : nullptr, rho0, u0);
}

int oal_download_macro(OalPlan* p, int level, double* rho, double* u) {
    if (int rc = check_level(p, level)) return rc;
    const Level& L = p->levels[level];
    if (L.c8map.empty()) return download_macro_lib(p, level, rho, u);
    std::vector<double> r(rho ? L.ncell_host : 0), v(u ? 3 * (size_t)L.ncell_host : 0);
    if (int rc = download_macro_lib(p, level, rho ? r.data() : nullptr, u ? v.data() : nullptr)) return rc;
    if (rho) cells_from_lib(L, r.data(), rho, 1);
    if (u) cells_from_lib(L, v.data(), u, 3);
    return OAL_OK;
}

int oal_set_bouzidi(OalPlan* p, int level, const uint16_t* q_map, int32_t ff_fallback, float q_min) {
    if (int rc = check_level(p, level)) return rc;
    const Level& L = p->levels[level];
    if (L.c8map.empty() || !q_map) return set_bouzidi_lib(p, level, q_map, ff_fallback, q_min);
    return set_bouzidi_lib(p, level, cells_to_lib(L, q_map, 27, (uint16_t)0).data(), ff_fallback, q_min);
}

int oal_set_sponge(OalPlan* p, int level, const float* sponge, int32_t blend_distributions) {
    if (int rc = check_level(p, level)) return rc;
    const Level& L = p->levels[level];
    if (L.c8map.empty() || !sponge) return set_sponge_lib(p, level, sponge, blend_distributions);
    return set_sponge_lib(p, level, cells_to_lib(L, sponge, 1, 0.0f).data(), blend_distributions);
}

// ---- LBM1 raw level state (src/io/raw_state_dump.jl:10-47) ------------------------------------------
namespace {
constexpr int32_t LBM1_MAGIC = 0x4C424D31;   // "LBM1"

struct FileCloser {
    FILE* f;
    ~FileCloser() { if (f) std::fclose(f); }
};
}  // namespace

int oal_write_raw_state(OalPlan* p, int level, int32_t level_id, double dx, const char* path) {
    if (int rc = check_level(p, level)) return rc;
    if (!path) return fail(OAL_ERR_ARG, "NULL path");
    Level& L = p->levels[level];
    if (!L.have_f) return fail(OAL_ERR_STATE, "no populations on this level");
    // in the caller's blocks (8^3 levels: the caller's 8^3 blocks)
    const bool c8 = !L.c8map.empty();
    const size_t nc = c8 ? L.c8map.size() : L.ncell_host;
    const uint32_t nb = c8 ? (uint32_t)(L.map8_h.size() / 3) : L.n_caller_blocks;
    const std::vector<int32_t>& map = c8 ? L.map8_h : L.map_h;
    const std::vector<uint8_t>& obst = c8 ? L.obst8_h : L.obst_h;
    std::vector<double> rho(nc), u(3 * nc);
    if (int rc = oal_download_macro(p, level, rho.data(), u.data())) return rc;
    // the reference's obstacle cells carry rho = 1, u = 0 (kernel_cumulant.jl:405-410)
    std::vector<float> fr(nc), fv(3 * nc);
    std::vector<uint8_t> ob(nc, 0);
    for (size_t c = 0; c < nc; ++c) {
        const bool solid = !obst.empty() && obst[c];
        ob[c] = solid ? 1 : 0;
        fr[c] = solid ? 1.0f : (float)rho[c];
        for (int a = 0; a < 3; ++a) fv[a * nc + c] = solid ? 0.0f : (float)u[a * nc + c];
    }
    int32_t head[7] = {LBM1_MAGIC, level_id, (int32_t)nb, c8 ? BS8 : BS, 0, 0, 0};
    for (uint32_t b = 0; b < nb; ++b)
        for (int a = 0; a < 3; ++a) head[4 + a] = std::max(head[4 + a], map[3 * b + a]);
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) return fail(OAL_ERR_ARG, std::string("cannot open for writing: ") + path);
    bool ok = std::fwrite(head, sizeof(head), 1, fc.f) == 1 && std::fwrite(&dx, sizeof(dx), 1, fc.f) == 1 &&
              std::fwrite(map.data(), sizeof(int32_t), 3 * (size_t)nb, fc.f) == 3 * (size_t)nb &&
              std::fwrite(fr.data(), sizeof(float), nc, fc.f) == nc &&
              std::fwrite(fv.data(), sizeof(float), 3 * nc, fc.f) == 3 * nc &&
              std::fwrite(ob.data(), 1, nc, fc.f) == nc;
    ok = (std::fclose(fc.f) == 0) && ok;
    fc.f = nullptr;
    if (!ok) return fail(OAL_ERR_ARG, std::string("write failed: ") + path);
    return OAL_OK;
}

int oal_read_raw_state(OalPlan* p, int level, const char* path, int32_t* level_id, double* dx) {
    if (int rc = check_level(p, level)) return rc;
    if (!path) return fail(OAL_ERR_ARG, "NULL path");
    Level& L = p->levels[level];
    const bool c8 = !L.c8map.empty();
    const size_t nc = c8 ? L.c8map.size() : L.ncell_host;
    const uint32_t nb = c8 ? (uint32_t)(L.map8_h.size() / 3) : L.n_caller_blocks;
    const std::vector<int32_t>& map = c8 ? L.map8_h : L.map_h;
    const std::vector<uint8_t>& obst = c8 ? L.obst8_h : L.obst_h;
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) return fail(OAL_ERR_ARG, std::string("cannot open: ") + path);
    int32_t head[7];
    double h_dx = 0.0;
    if (std::fread(head, sizeof(head), 1, fc.f) != 1 || head[0] != LBM1_MAGIC)
        return fail(OAL_ERR_ARG, "not an LBM1 raw state file");
    if (std::fread(&h_dx, sizeof(h_dx), 1, fc.f) != 1) return fail(OAL_ERR_ARG, "truncated LBM1 header");
    if (head[2] != (int32_t)nb || head[3] != (c8 ? BS8 : BS))
        return fail(OAL_ERR_ARG, "LBM1 file: block count or block size differs from the level's");
    std::vector<int32_t> coords(3 * (size_t)nb);
    std::vector<float> fr(nc), fv(3 * nc);
    std::vector<uint8_t> ob(nc);
    if (std::fread(coords.data(), sizeof(int32_t), coords.size(), fc.f) != coords.size() ||
        std::fread(fr.data(), sizeof(float), nc, fc.f) != nc || std::fread(fv.data(), sizeof(float), 3 * nc, fc.f) != 3 * nc ||
        std::fread(ob.data(), 1, nc, fc.f) != nc)
        return fail(OAL_ERR_ARG, "truncated LBM1 file");
    unsigned char extra;
    if (std::fread(&extra, 1, 1, fc.f) != 0) return fail(OAL_ERR_ARG, "LBM1 file has trailing data");
    if (!std::equal(coords.begin(), coords.end(), map.begin()))
        return fail(OAL_ERR_ARG, "LBM1 file: block coordinates differ from the level's");
    for (size_t c = 0; c < nc; ++c) {
        const uint8_t want = !obst.empty() && obst[c] ? 1 : 0;
        if ((ob[c] != 0) != (want != 0)) return fail(OAL_ERR_ARG, "LBM1 file: obstacle mask differs from the level's");
    }
    std::vector<double> rho(nc), u(3 * nc);
    for (size_t c = 0; c < nc; ++c) rho[c] = fr[c];
    for (size_t i = 0; i < 3 * nc; ++i) u[i] = fv[i];
    if (int rc = oal_init_equilibrium(p, level, rho.data(), u.data(), 1.0, nullptr)) return rc;
    if (level_id) *level_id = head[1];
    if (dx) *dx = h_dx;
    return OAL_OK;
}

int oal_get_tuning(const OalPlan* p, OalTuning* t) {
    if (!p || !t) return fail(OAL_ERR_ARG, "NULL argument");
    *t = OalTuning{};
    t->pipe_planes = p->pipe_planes;
    t->overlap = p->overlap ? 1 : 0;
    t->face_kernel = p->face_kernel ? 1 : 0;
    t->graph = p->graph ? 1 : 0;
    t->pair_kernel = p->pair_kernel;
    return OAL_OK;
}

int oal_set_tuning(OalPlan* p, const OalTuning* t) {
    if (!p || !t) return fail(OAL_ERR_ARG, "NULL argument");
    const int np = t->pipe_planes;
    if (np != 1 && np != 2 && np != 4 && np != 8 && np != 16) return fail(OAL_ERR_ARG, "pipe_planes must be 1, 2, 4, 8 or 16");
    if (t->overlap != 0 && t->overlap != 1) return fail(OAL_ERR_ARG, "overlap must be 0 or 1");
    if (t->face_kernel != 0 && t->face_kernel != 1) return fail(OAL_ERR_ARG, "face_kernel must be 0 or 1");
    if (t->graph != 0 && t->graph != 1) return fail(OAL_ERR_ARG, "graph must be 0 or 1");
    if (t->pair_kernel < 0 || t->pair_kernel > 2) return fail(OAL_ERR_ARG, "pair_kernel must be 0, 1 or 2");
    OAL_HIP(hipSetDevice(p->device));
    OAL_HIP(hipStreamSynchronize(p->stream));
    drop_graphs(p);
    p->pipe_planes = np;
    p->overlap = t->overlap != 0;
    p->face_kernel = t->face_kernel != 0;
    p->graph = t->graph != 0;
    p->pair_kernel = t->pair_kernel;
    return OAL_OK;
}

int oal_autotune(OalPlan* p, int level, int n_steps, OalTuning* best, double* ms_per_step) {
    if (int rc = check_level(p, level)) return rc;
    if (n_steps <= 0 || (n_steps & 1)) return fail(OAL_ERR_ARG, "n_steps must be positive and even");
    Level& L = p->levels[level];
    if (!L.have_f) return fail(OAL_ERR_STATE, "no populations on this level");
    OAL_HIP(hipSetDevice(p->device));
    // save the state: the candidate runs advance it
    void* save = nullptr;
    OAL_HIP(hipMalloc(&save, L.fA.n));
    hipError_t e = hipMemcpyAsync(save, L.fA.p, L.fA.n, hipMemcpyDeviceToDevice, p->stream);
    void* vsave = nullptr;
    if (e == hipSuccess && p->wale) {
        e = hipMalloc(&vsave, L.velA.n);
        if (e == hipSuccess) e = hipMemcpyAsync(vsave, L.velA.p, L.velA.n, hipMemcpyDeviceToDevice, p->stream);
    }
    const int aa_phase = L.aa_phase;
    OAL_HIP(hipStreamSynchronize(p->stream));
    drop_graphs(p);
    const int cand[5] = {1, 2, 4, 8, 16};
    int bestc = p->pipe_planes;
    bool besto = p->overlap, bestf = p->face_kernel;
    double bestms = 1e300;
    int rc = OAL_OK;
    if (e != hipSuccess) rc = fail(OAL_ERR_HIP, std::string("oal_autotune: ") + hipGetErrorString(e));
    for (int i = 0; i < OAL_TUNE_SHAPES && rc == OAL_OK; ++i) {
        p->pipe_planes = cand[i % 5];
        p->overlap = (i / 5) % 2 == 0;
        p->face_kernel = i >= 10;
        double tot = 0.0, ker = 0.0;
        rc = oal_time_steps(p, level, 2, 0, &tot, &ker);   // warm-up of this shape
        if (rc == OAL_OK) rc = oal_time_steps(p, level, n_steps, 0, &tot, &ker);
        if (rc != OAL_OK) break;
        const double ms = tot / n_steps;   // whole steps, all launches
        if (ms_per_step) ms_per_step[i] = ms;
        if (ms < bestms) { bestms = ms; bestc = cand[i % 5]; besto = p->overlap; bestf = p->face_kernel; }
    }
    p->pipe_planes = bestc;
    p->overlap = besto;
    p->face_kernel = bestf;
    // restore
    if (e == hipSuccess) e = hipMemcpyAsync(L.fA.p, save, L.fA.n, hipMemcpyDeviceToDevice, p->stream);
    if (e == hipSuccess && L.fB.p) e = hipMemcpyAsync(L.fB.p, save, L.fA.n, hipMemcpyDeviceToDevice, p->stream);
    if (e == hipSuccess && vsave) e = hipMemcpyAsync(L.velA.p, vsave, L.velA.n, hipMemcpyDeviceToDevice, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    (void)hipFree(save);
    if (vsave) (void)hipFree(vsave);
    L.aa_phase = aa_phase;
    L.force_valid = false;
    if (rc == OAL_OK && e != hipSuccess) rc = fail(OAL_ERR_HIP, std::string("oal_autotune: ") + hipGetErrorString(e));
    if (rc == OAL_OK && best) {
        *best = OalTuning{};
        best->pipe_planes = bestc;
        best->overlap = besto ? 1 : 0;
        best->face_kernel = bestf ? 1 : 0;
        best->graph = p->graph ? 1 : 0;
        best->pair_kernel = p->pair_kernel;
    }
    return rc;
}

}  // extern "C"
