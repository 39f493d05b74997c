// liboal host side: the C ABI of include/oal.h.  A plan owns the device state of a solver run
// (every level's block tables, populations, reductions) and launches the kernels of
// oal_kernels.hip on its stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/oal.h"
#include "oal_kernels.h"

using namespace oal;

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
    DevBuf<uint32_t> face_cells;       // AB: face cells of the regular-list blocks
    uint32_t n_face = 0;
    bool regular_face = false;         // AB: the regular list holds blocks on a domain face
    DevBuf<int32_t> lslot;             // link-mask slot per block
    DevBuf<uint32_t> lmask;            // per-cell solid-link masks of blocks near solid cells
    DevBuf<unsigned char> obuf;        // AA: outlet-face populations with cx = -1
    DevBuf<unsigned long long> guard;
    bool has_solid = false;
    int64_t n_fluid = 0;
    bool have_f = false;
    bool force_valid = false;
    int aa_phase = 0;                // AA: 0 = populations in stored form (next step odd), 1 = after an odd step
    void release() {
        nbr.release(); borg.release(); bdom.release(); dom_item0.release(); dom_block0.release();
        bflags.release(); solid.release(); doms.release(); fA.release(); fB.release(); velA.release(); velB.release();
        fpart.release(); mpart.release(); red_out.release(); guard.release(); oslot.release(); obuf.release(); blist_rg.release(); blist_in.release(); blist_br.release(); blist_bd.release(); face_cells.release();
        lslot.release(); lmask.release();
    }
};

}  // namespace

struct OalPlan {
    int device = 0;
    int pipe_planes = 4;   // planes per workgroup of the pipelined interior kernel
    bool overlap = true;   // boundary and face launches on the side stream
    bool face_kernel = false;   // face cells by the face-cell kernel instead of rules in the pipelined one
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
    std::vector<Level> levels;
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
    std::vector<int32_t> borg, bdom, nbr, nbr_bm, oslot, blist_rg, blist_in, blist_br, blist_bd, lslot;
    std::vector<uint32_t> lmask;
    std::vector<uint32_t> face_cells;
    int32_t n_oslots = 0;
    std::vector<uint32_t> flags;
    std::vector<uint8_t> solid;
};

// Host-side construction of one level's tables from the reference-convention descriptor.
int prepare_level(const OalLevelDesc& d, const OalCollision& coll, bool ab, Level& L, HostLevel& H) {
    if (d.n_blocks <= 0) return fail(OAL_ERR_ARG, "level with no blocks");
    if (!d.map_x || !d.map_y || !d.map_z || !d.neighbor_table) return fail(OAL_ERR_ARG, "level block arrays missing");
    if (d.n_domains <= 0 || !d.domains) return fail(OAL_ERR_ARG, "level has no domain descriptors");
    const uint64_t ncell64 = (uint64_t)d.n_blocks * BS3;
    if (ncell64 * 27ull >= (1ull << 32)) return fail(OAL_ERR_UNSUPPORTED, "level too large for 32-bit cell indexing (n_blocks*4096*27 >= 2^32)");
    L.nb = (uint32_t)d.n_blocks;
    L.ncell = (uint32_t)ncell64;
    L.n_items = L.nb * PLANES;
    L.n_domains = d.n_domains;
    const uint32_t nb = L.nb;

    std::vector<int32_t>& borg = H.borg;
    std::vector<int32_t>& bdom = H.bdom;
    std::vector<int32_t>& nbr = H.nbr;
    borg.assign(3 * (size_t)nb, 0);
    bdom.assign(nb, 0);
    nbr.assign(27 * (size_t)nb, -1);
    for (uint32_t b = 0; b < nb; ++b) {
        if (d.map_x[b] < 1 || d.map_y[b] < 1 || d.map_z[b] < 1) return fail(OAL_ERR_ARG, "block coordinates are 1-based");
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
        if (s.zwall_bc != 0 && s.zwall_bc != 2) return fail(OAL_ERR_UNSUPPORTED, "zwall_bc 1 (exact specular slip) is not supported");
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
    // (a boundary block qualifies when no neighbour inside the extent is absent and every present one is
    // at its lexicographic offset: its out-of-extent sources take the face rules and are never read)
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
        if (flags[b] == 0) {
            (regular ? H.blist_rg : H.blist_in).push_back((int32_t)b);
        } else if (ab && regular && flags[b] == BF_EDGE) {
            // AB pattern: a whole block on a domain face, no solid near, joins the regular list; the
            // cells on its faces (sources outside the extent) go to the face-cell list
            H.blist_rg.push_back((int32_t)b);
            L.regular_face = true;
            const int o[3] = {borg[3 * b], borg[3 * b + 1], borg[3 * b + 2]};
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
        } else {
            (regular ? H.blist_br : H.blist_bd).push_back((int32_t)b);
        }
    }
    L.n_face = (uint32_t)H.face_cells.size();
    L.n_regular = (uint32_t)H.blist_rg.size();
    L.n_interior = (uint32_t)H.blist_in.size();
    L.n_boundary_regular = (uint32_t)H.blist_br.size();
    L.n_boundary = (uint32_t)H.blist_bd.size();
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
        (rc = upload(L.blist_br, H.blist_br)) || (rc = upload(L.blist_bd, H.blist_bd)) ||
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
    a.n_interior = L.n_interior;
    a.n_boundary_regular = L.n_boundary_regular;
    a.n_boundary = L.n_boundary;
    a.lslot = L.lslot.p;
    a.lmask = L.lmask.p;
    a.face_cells = L.face_cells.p;
    a.n_face = L.n_face;
    a.regular_face = L.regular_face ? 1 : 0;
    // the in-kernel face rules address by 32-bit byte offsets: levels of 4 GiB or more take the face-cell kernel
    const bool small = (uint64_t)L.ncell * 27u * (uint64_t)storage_bytes(p->storage) <= 0xffffffffull;
    a.face_kernel = (p->face_kernel || !small) ? 1 : 0;
    a.obuf = L.obuf.p;
    a.ncell = L.ncell;
    a.nb = L.nb;
    a.n_items = L.n_items;
    a.time_step = time_step;
    a.parity = p->stream_mode == OAL_STREAM_AA ? (L.aa_phase == 0 ? 1 : 0) : 0;
    a.pipe_planes = p->pipe_planes;
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
    const bool split = p->overlap && L.n_regular + L.n_interior > 0 && L.n_boundary_regular + L.n_boundary + L.n_face > 0;
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
    OAL_HIP(launch_macro(p->storage, L.fA.p, nullptr, nullptr, L.ncell, permute_of(p), L.velA.p, p->stream));
    return OAL_OK;
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
    if (desc->block_size != BS) return fail(OAL_ERR_UNSUPPORTED, "block_size must be 16");
    if (desc->stream_mode != OAL_STREAM_AB && desc->stream_mode != OAL_STREAM_AA) return fail(OAL_ERR_ARG, "unknown stream_mode");
    if (desc->bb_mode != OAL_BB_REFERENCE && desc->bb_mode != OAL_BB_HALFWAY) return fail(OAL_ERR_ARG, "unknown bb_mode");
    if (desc->stream_mode == OAL_STREAM_AA && desc->bb_mode != OAL_BB_HALFWAY)
        return fail(OAL_ERR_UNSUPPORTED, "the AA pattern needs halfway bounce-back");
    if (desc->n_levels != 1) return fail(OAL_ERR_UNSUPPORTED, "plans of more than one level are not supported in this version");
    if (!desc->levels) return fail(OAL_ERR_ARG, "no level descriptors");
    const OalCollision& c = desc->collision;
    if (c.limiter < 0 || c.limiter > 2) return fail(OAL_ERR_ARG, "unknown limiter");
    if (desc->rounding != OAL_ROUND_STOCHASTIC && desc->rounding != OAL_ROUND_NEAREST)
        return fail(OAL_ERR_ARG, "rounding must be OAL_ROUND_STOCHASTIC or OAL_ROUND_NEAREST");

    OalPlan* p = new (std::nothrow) OalPlan();
    if (!p) return fail(OAL_ERR_NOMEM, "host allocation failed");
    p->device = desc->device;
    p->storage = desc->storage;
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

    auto bail = [&](int rc) {
        oal_plan_destroy(p);
        return rc;
    };
    p->levels.resize(desc->n_levels);
    std::vector<HostLevel> host(desc->n_levels);
    for (int l = 0; l < desc->n_levels; ++l) {
        const int rc = prepare_level(desc->levels[l], desc->collision, desc->stream_mode == OAL_STREAM_AB, p->levels[l],
                                     host[l]);
        if (rc) return bail(rc);
    }
    hipError_t e = hipSetDevice(p->device);
    if (e != hipSuccess) return bail(fail(OAL_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)));
    if ((e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&p->ev0)) != hipSuccess || (e = hipEventCreate(&p->ev1)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&p->join, hipEventDisableTiming)) != hipSuccess)
        return bail(fail(OAL_ERR_HIP, std::string("stream/event creation: ") + hipGetErrorString(e)));
    for (int l = 0; l < desc->n_levels; ++l) {
        const int rc = upload_level(p, p->levels[l], host[l]);
        if (rc) return bail(rc);
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
    *n = p->levels[level].ncell;
    return OAL_OK;
}

int oal_level_info(const OalPlan* p, int level, OalLevelInfo* info) {
    if (int rc = check_level(p, level)) return rc;
    if (!info) return fail(OAL_ERR_ARG, "NULL output");
    const Level& L = p->levels[level];
    *info = OalLevelInfo{};
    info->n_cells = L.ncell;
    info->n_fluid_cells = L.n_fluid;
    info->n_blocks = (int32_t)L.nb;
    info->n_regular = (int32_t)L.n_regular;
    info->n_interior = (int32_t)L.n_interior;
    info->n_boundary_regular = (int32_t)L.n_boundary_regular;
    info->n_boundary = (int32_t)L.n_boundary;
    info->n_face_cells = (int32_t)L.n_face;
    return OAL_OK;
}

int oal_upload_f(OalPlan* p, int level, const void* host_f, int dt) {
    if (int rc = check_level(p, level)) return rc;
    if (!host_f || (dt != OAL_F32 && dt != OAL_F64)) return fail(OAL_ERR_ARG, "host_f must be OAL_F32 or OAL_F64 data");
    OAL_HIP(hipSetDevice(p->device));
    Level& L = p->levels[level];
    const size_t bytes = (size_t)L.ncell * 27 * (dt == OAL_F64 ? 8 : 4);
    void* stage = nullptr;
    OAL_HIP(hipMalloc(&stage, bytes));
    hipError_t e = hipMemcpyAsync(stage, host_f, bytes, hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) e = launch_convert_in(p->storage, L.fA.p, stage, dt, L.ncell, permute_of(p), p->stream);
    if (e == hipSuccess && L.fB.p) e = hipMemcpyAsync(L.fB.p, L.fA.p, L.fA.n, hipMemcpyDeviceToDevice, p->stream);
    if (e == hipSuccess) e = aa_prep(p, L);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(OAL_ERR_HIP, std::string("oal_upload_f: ") + hipGetErrorString(e));
    L.have_f = true;
    L.aa_phase = 0;
    L.force_valid = false;
    return refresh_velocity(p, L);
}

int oal_download_f(OalPlan* p, int level, void* host_f, int dt) {
    if (int rc = check_level(p, level)) return rc;
    if (!host_f || (dt != OAL_F32 && dt != OAL_F64)) return fail(OAL_ERR_ARG, "host_f must be OAL_F32 or OAL_F64 data");
    Level& L = p->levels[level];
    if (int rc = require_plain_state(p, L)) return rc;
    OAL_HIP(hipSetDevice(p->device));
    const size_t bytes = (size_t)L.ncell * 27 * (dt == OAL_F64 ? 8 : 4);
    void* stage = nullptr;
    OAL_HIP(hipMalloc(&stage, bytes));
    hipError_t e = launch_convert_out(p->storage, L.fA.p, stage, dt, L.ncell, permute_of(p), p->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host_f, stage, bytes, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(OAL_ERR_HIP, std::string("oal_download_f: ") + hipGetErrorString(e));
    return OAL_OK;
}

int oal_init_equilibrium(OalPlan* p, int level, const double* rho, const double* u, double rho0, const double u0[3]) {
    if (int rc = check_level(p, level)) return rc;
    OAL_HIP(hipSetDevice(p->device));
    Level& L = p->levels[level];
    double* drho = nullptr;
    double* du = nullptr;
    hipError_t e = hipSuccess;
    if (rho) {
        e = hipMalloc((void**)&drho, (size_t)L.ncell * 8);
        if (e == hipSuccess) e = hipMemcpyAsync(drho, rho, (size_t)L.ncell * 8, hipMemcpyHostToDevice, p->stream);
    }
    if (e == hipSuccess && u) {
        e = hipMalloc((void**)&du, (size_t)L.ncell * 24);
        if (e == hipSuccess) e = hipMemcpyAsync(du, u, (size_t)L.ncell * 24, hipMemcpyHostToDevice, p->stream);
    }
    const double ux0 = u0 ? u0[0] : 0.0, uy0 = u0 ? u0[1] : 0.0, uz0 = u0 ? u0[2] : 0.0;
    if (e == hipSuccess)
        e = launch_init_eq(p->storage, L.fA.p, L.ncell, drho, du, rho0, ux0, uy0, uz0, permute_of(p),
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
    for (int i = 0; i < n_steps; ++i) {
        const int f = (i == n_steps - 1) ? flags : 0;
        if (int rc = level_step(p, 0, first_time_step + i, f, s)) return rc;
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
    OAL_HIP(launch_mass(p->storage, L.fA.p, L.solid.p, L.borg.p, L.bdom.p, L.doms.p, L.nb, permute_of(p), L.mpart.p, s));
    // the reduction walks blocks here: one "item" per block
    ReduceArgs r{L.mpart.p, L.dom_block0.p, L.red_out.p, L.n_domains, 4};
    OAL_HIP(launch_reduce(r, s));
    OAL_HIP(hipStreamSynchronize(s));
    OAL_HIP(hipMemcpy(out, L.red_out.p, (size_t)L.n_domains * 32, hipMemcpyDeviceToHost));
    L.force_valid = false;   // red_out reused
    return OAL_OK;
}

int oal_download_macro(OalPlan* p, int level, double* rho, double* u) {
    if (int rc = check_level(p, level)) return rc;
    Level& L = p->levels[level];
    if (int rc = require_plain_state(p, L)) return rc;
    OAL_HIP(hipSetDevice(p->device));
    double* d = nullptr;
    OAL_HIP(hipMalloc((void**)&d, (size_t)L.ncell * 32));
    hipError_t e = launch_macro(p->storage, L.fA.p, d, d + L.ncell, L.ncell, permute_of(p), nullptr, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e == hipSuccess && rho) e = hipMemcpy(rho, d, (size_t)L.ncell * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && u) e = hipMemcpy(u, d + L.ncell, (size_t)L.ncell * 24, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(OAL_ERR_HIP, std::string("oal_download_macro: ") + hipGetErrorString(e));
    return OAL_OK;
}

int oal_guard(OalPlan* p, int level, float rho_min, float u_max, int64_t* bad) {
    if (int rc = check_level(p, level)) return rc;
    Level& L = p->levels[level];
    if (!L.have_f) return fail(OAL_ERR_STATE, "no populations on this level");
    OAL_HIP(hipSetDevice(p->device));
    unsigned long long h = 0;
    OAL_HIP(hipMemsetAsync(L.guard.p, 0, sizeof(unsigned long long), p->stream));
    OAL_HIP(launch_guard(p->storage, L.fA.p, L.solid.p, L.borg.p, L.bdom.p, L.doms.p, L.nb, rho_min, u_max, L.guard.p, p->stream));
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

int oal_get_tuning(const OalPlan* p, OalTuning* t) {
    if (!p || !t) return fail(OAL_ERR_ARG, "NULL argument");
    *t = OalTuning{};
    t->pipe_planes = p->pipe_planes;
    t->overlap = p->overlap ? 1 : 0;
    t->face_kernel = p->face_kernel ? 1 : 0;
    return OAL_OK;
}

int oal_set_tuning(OalPlan* p, const OalTuning* t) {
    if (!p || !t) return fail(OAL_ERR_ARG, "NULL argument");
    const int np = t->pipe_planes;
    if (np != 1 && np != 2 && np != 4 && np != 8 && np != 16) return fail(OAL_ERR_ARG, "pipe_planes must be 1, 2, 4, 8 or 16");
    if (t->overlap != 0 && t->overlap != 1) return fail(OAL_ERR_ARG, "overlap must be 0 or 1");
    if (t->face_kernel != 0 && t->face_kernel != 1) return fail(OAL_ERR_ARG, "face_kernel must be 0 or 1");
    p->pipe_planes = np;
    p->overlap = t->overlap != 0;
    p->face_kernel = t->face_kernel != 0;
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
    }
    return rc;
}

}  // extern "C"
