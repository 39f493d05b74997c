// Kernels of liboal for gfx950 other than the stream-collide step (oal_step.h): initialisation,
// layout conversion, macroscopic output, FP64 reductions, the stability guard and the bandwidth probe,
// and the storage dispatch of the step launches.
#include <hip/hip_runtime.h>

#include "oal_kernels.h"

namespace oal {

hipError_t launch_step(int storage, int stream_mode, bool wale, const StepArgs& a, hipStream_t stream, hipStream_t side,
                       hipEvent_t bulk0, hipEvent_t bulk1) {
    switch (storage) {
        case 0: return launch_step_f32(stream_mode, wale, a, stream, side, bulk0, bulk1);
        case 1: return launch_step_f64(stream_mode, wale, a, stream, side, bulk0, bulk1);
        case 2: return launch_step_f16s(stream_mode, wale, a, stream, side, bulk0, bulk1);
        case 3: return launch_step_bf16s(stream_mode, wale, a, stream, side, bulk0, bulk1);
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
__global__ void __launch_bounds__(WG) k_init_eq(typename SP::st* f, uint32_t ncell, uint32_t nhost, const double* rho,
                                                const double* u, double rho0, double ux0, double uy0, double uz0,
                                                int permute, typename SP::T* vel) {
    using T = typename SP::T;
    const uint32_t cidx = blockIdx.x * WG + threadIdx.x;
    if (cidx >= ncell) return;
    // host fields cover the first nhost cells; the rest (ghost blocks of a refined level) take (rho0, u0)
    const bool h = cidx < nhost;
    const T r = rho && h ? T(rho[cidx]) : T(rho0);
    const T ux = u && h ? T(u[cidx]) : T(ux0);
    const T uy = u && h ? T(u[nhost + cidx]) : T(uy0);
    const T uz = u && h ? T(u[2 * (size_t)nhost + cidx]) : T(uz0);
    const T cs2 = T(1) / T(3);
    const T cs4 = cs2 * cs2;
    const T usq = ux * ux + uy * uy + uz * uz;
    const T inv_cs2 = T(1) / cs2, inv_2cs4 = T(0.5) / cs4, inv_2cs2 = T(0.5) / cs2;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const T cu = T(dcx(k)) * ux + T(dcy(k)) * uy + T(dcz(k)) * uz;
        const T feq = r * wgt<T>(k) * (T(1) + inv_cs2 * cu + inv_2cs4 * cu * cu - inv_2cs2 * usq);
        SP::sto(f + pop_index(cidx, permute ? 26 - k : k), feq, wgt<T>(k));
    }
    if (vel) { vel[cidx] = ux; vel[ncell + cidx] = uy; vel[2 * (size_t)ncell + cidx] = uz; }
}

template <class SP>
static hipError_t init_sp(void* f, uint32_t ncell, uint32_t nhost, const double* rho, const double* u, double rho0,
                          double ux0, double uy0, double uz0, int permute, void* vel, hipStream_t s) {
    hipLaunchKernelGGL((k_init_eq<SP>), dim3((ncell + WG - 1) / WG), dim3(WG), 0, s, (typename SP::st*)f, ncell, nhost, rho,
                       u, rho0, ux0, uy0, uz0, permute, (typename SP::T*)vel);
    return hipGetLastError();
}

hipError_t launch_init_eq(int storage, void* f, uint32_t ncell, uint32_t nhost, const double* rho, const double* u,
                          double rho0, double ux0, double uy0, double uz0, int permute, void* vel, hipStream_t s) {
    switch (storage) {
        case 0: return init_sp<StF32>(f, ncell, nhost, rho, u, rho0, ux0, uy0, uz0, permute, vel, s);
        case 1: return init_sp<StF64>(f, ncell, nhost, rho, u, rho0, ux0, uy0, uz0, permute, vel, s);
        case 2: return init_sp<StF16S>(f, ncell, nhost, rho, u, rho0, ux0, uy0, uz0, permute, vel, s);
        case 3: return init_sp<StBF16S>(f, ncell, nhost, rho, u, rho0, ux0, uy0, uz0, permute, vel, s);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------------------------------------
// Layout conversion between host populations (unshifted f32/f64) and the storage form
// ------------------------------------------------------------------------------------------------
template <class SP, class H>
__global__ void __launch_bounds__(WG) k_convert_in(typename SP::st* f, const H* src, uint32_t ncell, uint32_t nhost,
                                                   int permute) {
    using T = typename SP::T;
    const uint32_t cidx = blockIdx.x * WG + threadIdx.x;
    const int k = blockIdx.y;
    if (cidx >= nhost) return;
    SP::sto(f + pop_index(cidx, permute ? 26 - k : k), T(src[(size_t)k * nhost + cidx]), wgt<T>(k));
}

template <class SP, class H>
__global__ void __launch_bounds__(WG) k_convert_out(const typename SP::st* f, H* dst, uint32_t ncell, uint32_t nhost,
                                                    int permute) {
    using T = typename SP::T;
    const uint32_t cidx = blockIdx.x * WG + threadIdx.x;
    const int k = blockIdx.y;
    if (cidx >= nhost) return;
    dst[(size_t)k * nhost + cidx] = H(SP::ld(f + pop_index(cidx, permute ? 26 - k : k), wgt<T>(k)));
}

template <class SP>
static hipError_t conv_in_sp(void* f, const void* src, int dt, uint32_t ncell, uint32_t nhost, int permute, hipStream_t s) {
    dim3 grid((nhost + WG - 1) / WG, 27);
    if (dt == 0) hipLaunchKernelGGL((k_convert_in<SP, float>), grid, dim3(WG), 0, s, (typename SP::st*)f, (const float*)src, ncell, nhost, permute);
    else hipLaunchKernelGGL((k_convert_in<SP, double>), grid, dim3(WG), 0, s, (typename SP::st*)f, (const double*)src, ncell, nhost, permute);
    return hipGetLastError();
}
template <class SP>
static hipError_t conv_out_sp(const void* f, void* dst, int dt, uint32_t ncell, uint32_t nhost, int permute, hipStream_t s) {
    dim3 grid((nhost + WG - 1) / WG, 27);
    if (dt == 0) hipLaunchKernelGGL((k_convert_out<SP, float>), grid, dim3(WG), 0, s, (const typename SP::st*)f, (float*)dst, ncell, nhost, permute);
    else hipLaunchKernelGGL((k_convert_out<SP, double>), grid, dim3(WG), 0, s, (const typename SP::st*)f, (double*)dst, ncell, nhost, permute);
    return hipGetLastError();
}

hipError_t launch_convert_in(int storage, void* f, const void* src, int dt, uint32_t ncell, uint32_t nhost, int permute,
                             hipStream_t s) {
    switch (storage) {
        case 0: return conv_in_sp<StF32>(f, src, dt, ncell, nhost, permute, s);
        case 1: return conv_in_sp<StF64>(f, src, dt, ncell, nhost, permute, s);
        case 2: return conv_in_sp<StF16S>(f, src, dt, ncell, nhost, permute, s);
        case 3: return conv_in_sp<StBF16S>(f, src, dt, ncell, nhost, permute, s);
        default: return hipErrorInvalidValue;
    }
}
hipError_t launch_convert_out(int storage, const void* f, void* dst, int dt, uint32_t ncell, uint32_t nhost, int permute,
                             hipStream_t s) {
    switch (storage) {
        case 0: return conv_out_sp<StF32>(f, dst, dt, ncell, nhost, permute, s);
        case 1: return conv_out_sp<StF64>(f, dst, dt, ncell, nhost, permute, s);
        case 2: return conv_out_sp<StF16S>(f, dst, dt, ncell, nhost, permute, s);
        case 3: return conv_out_sp<StBF16S>(f, dst, dt, ncell, nhost, permute, s);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------------------------------------
// Macroscopic output in FP64: rho = 1 + sum (f_k - w_k), u = sum c_k (f_k - w_k) / rho
// ------------------------------------------------------------------------------------------------
template <class SP>
__global__ void __launch_bounds__(WG) k_macro(const typename SP::st* f, double* rho, double* u, uint32_t ncell,
                                              uint32_t nhost, int permute, typename SP::T* vel) {
    using T = typename SP::T;
    const uint32_t cidx = blockIdx.x * WG + threadIdx.x;
    if (cidx >= nhost) return;
    double dr = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const double s = (double)SP::ld(f + pop_index(cidx, permute ? 26 - k : k), wgt<T>(k)) - (double)wgt<T>(k);
        dr += s; jx += s * dcx(k); jy += s * dcy(k); jz += s * dcz(k);
    }
    const double r = 1.0 + dr;
    if (rho) rho[cidx] = r;
    if (u) { u[cidx] = jx / r; u[nhost + cidx] = jy / r; u[2 * (size_t)nhost + cidx] = jz / r; }
    if (vel) { vel[cidx] = T(jx / r); vel[ncell + cidx] = T(jy / r); vel[2 * (size_t)ncell + cidx] = T(jz / r); }
}

template <class SP>
static hipError_t macro_sp(const void* f, double* rho, double* u, uint32_t ncell, uint32_t nhost, int permute, void* vel,
                          hipStream_t s) {
    hipLaunchKernelGGL((k_macro<SP>), dim3((nhost + WG - 1) / WG), dim3(WG), 0, s, (const typename SP::st*)f, rho, u, ncell,
                       nhost, permute, (typename SP::T*)vel);
    return hipGetLastError();
}

hipError_t launch_macro(int storage, const void* f, double* rho, double* u, uint32_t ncell, uint32_t nhost, int permute,
                        void* vel, hipStream_t s) {
    switch (storage) {
        case 0: return macro_sp<StF32>(f, rho, u, ncell, nhost, permute, vel, s);
        case 1: return macro_sp<StF64>(f, rho, u, ncell, nhost, permute, vel, s);
        case 2: return macro_sp<StF16S>(f, rho, u, ncell, nhost, permute, vel, s);
        case 3: return macro_sp<StBF16S>(f, rho, u, ncell, nhost, permute, vel, s);
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
    for (int q = 0; q < 9; ++q) o[((size_t)os * 9 + q) * BS2 + ly * BS + z] = f[pop_index(own, 26 - 3 * q)];
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
                                             const int32_t* bdom, const uint32_t* bflags, const DevDomain* doms,
                                             int permute, double* part) {
    using T = typename SP::T;
    const uint32_t b = blockIdx.x;
    if (bflags[b] & BF_INACTIVE) {   // covered or ghost block of a refined level: counted on another level
        if (threadIdx.x < 4) part[(size_t)b * 4 + threadIdx.x] = 0.0;
        return;
    }
    const DevDomain* dm = doms + bdom[b];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int z = 0; z < BS; ++z) {
        const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
        const uint32_t cidx = b * BS3 + z * BS2 + ly * BS + lx;
        const bool inside = borg[3 * b] + lx < dm->ext[0] && borg[3 * b + 1] + ly < dm->ext[1] && borg[3 * b + 2] + z < dm->ext[2];
        if (!inside || (solid && solid[cidx])) continue;
        double dr = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
        for (int k = 0; k < 27; ++k) {
            const double s = (double)SP::ld(f + pop_index(cidx, permute ? 26 - k : k), wgt<T>(k)) - (double)wgt<T>(k);
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
                       const uint32_t* bflags, const DevDomain* doms, uint32_t nb, int permute, double* part, hipStream_t s) {
    switch (storage) {
        case 0: hipLaunchKernelGGL((k_mass<StF32>), dim3(nb), dim3(WG), 0, s, (const float*)f, solid, borg, bdom, bflags, doms, permute, part); break;
        case 1: hipLaunchKernelGGL((k_mass<StF64>), dim3(nb), dim3(WG), 0, s, (const double*)f, solid, borg, bdom, bflags, doms, permute, part); break;
        case 2: hipLaunchKernelGGL((k_mass<StF16S>), dim3(nb), dim3(WG), 0, s, (const _Float16*)f, solid, borg, bdom, bflags, doms, permute, part); break;
        case 3: hipLaunchKernelGGL((k_mass<StBF16S>), dim3(nb), dim3(WG), 0, s, (const uint16_t*)f, solid, borg, bdom, bflags, doms, permute, part); break;
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
                                              const int32_t* bdom, const uint32_t* bflags, const DevDomain* doms,
                                              float rho_min, float u_max, unsigned long long* count) {
    using T = typename SP::T;
    const uint32_t b = blockIdx.x;
    if (bflags[b] & BF_INACTIVE) return;
    const DevDomain* dm = doms + bdom[b];
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
            const float v = (float)SP::ld(f + pop_index(cidx, k), wgt<T>(k));
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
                        const uint32_t* bflags, const DevDomain* doms, uint32_t nb, float rho_min, float u_max,
                        unsigned long long* count, hipStream_t s) {
    switch (storage) {
        case 0: hipLaunchKernelGGL((k_guard<StF32>), dim3(nb), dim3(WG), 0, s, (const float*)f, solid, borg, bdom, bflags, doms, rho_min, u_max, count); break;
        case 1: hipLaunchKernelGGL((k_guard<StF64>), dim3(nb), dim3(WG), 0, s, (const double*)f, solid, borg, bdom, bflags, doms, rho_min, u_max, count); break;
        case 2: hipLaunchKernelGGL((k_guard<StF16S>), dim3(nb), dim3(WG), 0, s, (const _Float16*)f, solid, borg, bdom, bflags, doms, rho_min, u_max, count); break;
        case 3: hipLaunchKernelGGL((k_guard<StBF16S>), dim3(nb), dim3(WG), 0, s, (const uint16_t*)f, solid, borg, bdom, bflags, doms, rho_min, u_max, count); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// Level transfers of a refined plan (the role of the octree's coupling, src/octree_q19/src/coupling.jl:
// 86-233): a written cell receives the stencil-weighted sum of its donors' post-collision populations,
// whose second-order non-equilibrium moments are then set to those of the receiving level: the donors'
// Pi_neq times `scale` (1/(1 - omega) of the sending level and the ratio of the levels' relaxation
// times) times `keep` (1 - omega of the receiving level), through the second-order Hermite term.  All
// other moments of the interpolated state are kept: the cumulant collision's higher-order state
// crosses the interface unchanged (a rebuild from rho, u and Pi_neq alone, as the octree's D3Q19
// regularised collision allows, would impose the polynomial equilibrium's higher moments and make a
// uniform flow change at the interfaces).  A cell whose corrected state has a population below 1e-6
// of its weight takes the equilibrium of its rho and u (kernel.jl:94-100).  Every operation is written
// out in a fixed order: the CPU oracle restates the same sequence and must agree bit for bit.
// ------------------------------------------------------------------------------------------------
// 16-byte vectors of the working type (records of the level transfers)
template <class T> struct Vec16;
template <> struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int n = 4;
};
template <> struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int n = 2;
};

// RAW (16-bit storage, no blend): records of the 27 stored codes padded to 32 (64 bytes), decoded by
// the rebuild; the decoded values are the ones a float record would hold.
constexpr int XFER_RAW_STRIDE = 32;

// gather: a few lanes per donor, each writing one 16-byte piece of its record (the pieces of a
// wavefront's donors are contiguous): RAW, 4 lanes of 8 codes; float records, 7 lanes of 4 values;
// double records, 14 lanes of 2
template <class SP, bool RAW>
constexpr int xfer_gather_lanes() {
    return RAW ? XFER_RAW_STRIDE / 8 : XFER_STRIDE / Vec16<typename SP::T>::n;
}
template <class SP, bool RAW>
__global__ void __launch_bounds__(WG) k_xfer_gather(const typename SP::st* fnew, const typename SP::st* fold,
                                                    const uint32_t* cells, uint32_t n, uint32_t ncell, typename SP::T w1,
                                                    typename SP::T* out) {
    using T = typename SP::T;
    constexpr int Q = xfer_gather_lanes<SP, RAW>();
    const uint32_t t = blockIdx.x * WG + threadIdx.x;
    if (t >= n * (uint32_t)Q) return;
    const uint32_t i = t / Q;
    const int q = (int)(t - i * Q);
    const uint32_t e = cells[i];
    const uint32_t c = e & 0x7fffffffu;
    if constexpr (RAW) {
        uint32_t h[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int k = 8 * q + r;
            h[r] = k < 27 ? (uint32_t)SP::ld_raw(fnew + pop_index(c, k)) : 0u;
        }
        reinterpret_cast<uint4*>(reinterpret_cast<uint16_t*>(out) + (size_t)i * XFER_RAW_STRIDE)[q] =
            make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16));
        return;
    }
    using V = typename Vec16<T>::type;
    constexpr int VN = Vec16<T>::n;
    const bool blend = fold && !(e >> 31);
    V v;
#pragma unroll
    for (int r = 0; r < VN; ++r) {
        const int k = VN * q + r;
        T x = T(0);
        if (k < 27) {
            const T w = wgt_rt<T>(k);
            x = SP::ld(fnew + pop_index(c, k), w);
            if (blend) x = SP::ld(fold + pop_index(c, k), w) * (T(1) - w1) + x * w1;
        }
        v[r] = x;
    }
    reinterpret_cast<V*>(out + (size_t)i * XFER_STRIDE)[q] = v;
}

// sum of a value over the 16 lanes of a DPP row, as a fixed pairwise tree: lanes l and l xor 1, then
// l xor 2 (quad permutations), then the two quads of each half-row (row_half_mirror: every lane of a
// quad holds its quad's sum, so the mirror pairs the quad sums) and the two half-rows (row_mirror).
// Every lane ends with ((p0 + p1) + (p2 + p3)) + ((p4 + p5) + (p6 + p7)) + ((p8 + ...)) in the same
// bits (addition commutes).
template <int CTRL>
__device__ __forceinline__ float dpp_get(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_get(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, 0xf, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <class T>
__device__ __forceinline__ T row_sum(T v) {
    v = v + dpp_get<0xb1>(v);    // quad_perm [1, 0, 3, 2]
    v = v + dpp_get<0x4e>(v);    // quad_perm [2, 3, 0, 1]
    v = v + dpp_get<0x141>(v);   // row_half_mirror
    v = v + dpp_get<0x140>(v);   // row_mirror
    return v;
}
// any lane of the lane's DPP row
__device__ __forceinline__ bool row_any(bool b) {
    const uint64_t m = __builtin_amdgcn_ballot_w64(b);
    return ((m >> (threadIdx.x & 48u)) & 0xffffull) != 0;
}

// rebuild: 16 lanes (one DPP row) per written cell, lane j summing stencil nodes 4j .. 4j + 3 in node
// order from zero, the 16 partial sums then added as a pairwise tree (row_sum); every lane forms the
// moments, lane j rebuilds and stores directions j and j + 16.  Stencils are cell-major: slot[i][64],
// wt[i][64].
template <class SP, bool RAW>
__global__ void __launch_bounds__(WG) k_xfer_rebuild(typename SP::st* f, uint32_t ncell, const uint32_t* tcells, uint32_t n,
                                                     const int32_t* slot, const float* wt, const typename SP::T* m,
                                                     typename SP::T scale, typename SP::T keep, typename SP::T* vel) {
    using T = typename SP::T;
    const uint32_t i = blockIdx.x * (WG / 16) + (threadIdx.x >> 4);
    const int j = (int)(threadIdx.x & 15);
    if (i >= n) return;   // whole rows: the lanes of a live cell stay together
    const int4 s4 = *reinterpret_cast<const int4*>(slot + (size_t)i * 64 + 4 * j);
    const float4 w4 = *reinterpret_cast<const float4*>(wt + (size_t)i * 64 + 4 * j);
    const int32_t sr[4] = {s4.x, s4.y, s4.z, s4.w};
    const T wn[4] = {T(w4.x), T(w4.y), T(w4.z), T(w4.w)};
    T acc[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) acc[k] = T(0);
    if constexpr (RAW) {
        uint4 rec[4][4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const uint4* row = reinterpret_cast<const uint4*>(reinterpret_cast<const uint16_t*>(m) +
                                                              (size_t)(sr[g] >= 0 ? sr[g] : 0) * XFER_RAW_STRIDE);
#pragma unroll
            for (int q = 0; q < 4; ++q) rec[g][q] = row[q];
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const bool use = sr[g] >= 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t wd[4] = {rec[g][q].x, rec[g][q].y, rec[g][q].z, rec[g][q].w};
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = 8 * q + e;
                    if (k < 27) {
                        const T add = acc[k] + wn[g] * SP::from_raw((wd[e >> 1] >> (16 * (e & 1))) & 0xffffu, wgt<T>(k));
                        acc[k] = use ? add : acc[k];
                    }
                }
            }
        }
    } else {
        using V = typename Vec16<T>::type;
        constexpr int VN = Vec16<T>::n;
        V rec[4][XFER_STRIDE / VN];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const V* row = reinterpret_cast<const V*>(m + (size_t)(sr[g] >= 0 ? sr[g] : 0) * XFER_STRIDE);
#pragma unroll
            for (int q = 0; q < XFER_STRIDE / VN; ++q) rec[g][q] = row[q];
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const bool use = sr[g] >= 0;
#pragma unroll
            for (int q = 0; q < XFER_STRIDE / VN; ++q)
#pragma unroll
                for (int e = 0; e < VN; ++e)
                    if (q * VN + e < 27) {
                        const T add = acc[q * VN + e] + wn[g] * rec[g][q][e];
                        acc[q * VN + e] = use ? add : acc[q * VN + e];
                    }
        }
    }
#pragma unroll
    for (int k = 0; k < 27; ++k) acc[k] = row_sum(acc[k]);
    // moments of the interpolated state (every lane)
    T rho = T(0), jx = T(0), jy = T(0), jz = T(0), pxx = T(0), pyy = T(0), pzz = T(0), pxy = T(0), pxz = T(0), pyz = T(0);
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int cx = dcx(K), cy = dcy(K), cz = dcz(K);
        const T v = acc[K];
        rho = rho + v;
        if constexpr (cx > 0) jx = jx + v;
        if constexpr (cx < 0) jx = jx - v;
        if constexpr (cy > 0) jy = jy + v;
        if constexpr (cy < 0) jy = jy - v;
        if constexpr (cz > 0) jz = jz + v;
        if constexpr (cz < 0) jz = jz - v;
        if constexpr (cx != 0) pxx = pxx + v;
        if constexpr (cy != 0) pyy = pyy + v;
        if constexpr (cz != 0) pzz = pzz + v;
        if constexpr (cx * cy > 0) pxy = pxy + v;
        if constexpr (cx * cy < 0) pxy = pxy - v;
        if constexpr (cx * cz > 0) pxz = pxz + v;
        if constexpr (cx * cz < 0) pxz = pxz - v;
        if constexpr (cy * cz > 0) pyz = pyz + v;
        if constexpr (cy * cz < 0) pyz = pyz - v;
    });
    const T ux = jx / rho, uy = jy / rho, uz = jz / rho;
    const T third = T(1.0 / 3.0);
    const T rx = rho * ux, ry = rho * uy;
    const T qxx = (pxx - rho * third) - rx * ux, qyy = (pyy - rho * third) - ry * uy, qzz = (pzz - rho * third) - (rho * uz) * uz;
    const T qxy = pxy - rx * uy, qxz = pxz - rx * uz, qyz = pyz - ry * uz;
    // change of the second-order non-equilibrium moments
    const T sk = scale * keep;
    const T dxx = sk * qxx - qxx, dyy = sk * qyy - qyy, dzz = sk * qzz - qzz;
    const T dxy = sk * qxy - qxy, dxz = sk * qxz - qxz, dyz = sk * qyz - qyz;
    const T tr = (dxx + dyy) + dzz;
    const T usq = (ux * ux + uy * uy) + uz * uz;
    // lane j: directions j and j + 16
    T out[2], eq[2];
    bool bad = false;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k = j + 16 * h;
        if (k < 27) {
            const T cx = T(k % 3 - 1), cy = T((k / 3) % 3 - 1), cz = T(k / 9 - 1);
            const T w = wgt_rt<T>(k);
            T a = acc[0];
#pragma unroll
            for (int kk = 1; kk < 27; ++kk) a = kk == k ? acc[kk] : a;
            const T cac = ((cx * cx * dxx + cy * cy * dyy) + cz * cz * dzz) + T(2) * ((cx * cy * dxy + cx * cz * dxz) + cy * cz * dyz);
            out[h] = a + (w * T(4.5)) * (cac - tr * third);
            const T cu = (cx * ux + cy * uy) + cz * uz;
            eq[h] = (w * rho) * (((T(1) + T(3) * cu) + T(4.5) * cu * cu) - T(1.5) * usq);
            bad = bad || out[h] < T(1e-6) * w;
        }
    }
    bad = row_any(bad);
    const uint32_t c = tcells[i];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k = j + 16 * h;
        if (k < 27) SP::sto(f + pop_index(c, k), bad ? eq[h] : out[h], wgt_rt<T>(k));
    }
    // WALE: the written cell's velocity (that of the interpolated state, which the rescale and the
    // equilibrium fallback keep) for the receiving level's gradient gather
    if (vel && j == 0) {
        vel[c] = ux;
        vel[ncell + c] = uy;
        vel[2 * (size_t)ncell + c] = uz;
    }
}

__global__ void k_set_i32(int32_t* p, int32_t v) {
    if (threadIdx.x == 0) p[0] = v;
}

hipError_t launch_set_i32(int32_t* p, int32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}


template <class SP, bool RAW>
static hipError_t xfer_sp(const void* fnew, const void* fold, const uint32_t* cells, uint32_t n, uint32_t ncell, double w1,
                          void* out, hipStream_t s) {
    using T = typename SP::T;
    if (!n) return hipSuccess;
    constexpr uint32_t Q = (uint32_t)xfer_gather_lanes<SP, RAW>();
    hipLaunchKernelGGL((k_xfer_gather<SP, RAW>), dim3((n * Q + WG - 1) / WG), dim3(WG), 0, s, (const typename SP::st*)fnew,
                       (const typename SP::st*)fold, cells, n, ncell, T(w1), (T*)out);
    return hipGetLastError();
}
template <class SP, bool RAW>
static hipError_t rebuild_sp(void* f, uint32_t ncell, const uint32_t* tcells, uint32_t n, const int32_t* slot, const float* w,
                             const void* m, double scale, double keep, void* vel, hipStream_t s) {
    using T = typename SP::T;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL((k_xfer_rebuild<SP, RAW>), dim3((n + WG / 16 - 1) / (WG / 16)), dim3(WG), 0, s, (typename SP::st*)f, ncell, tcells, n,
                       slot, w, (const T*)m, T(scale), T(keep), (T*)vel);
    return hipGetLastError();
}

// raw records when the storage is 16-bit and no time blend is asked for
hipError_t launch_xfer_gather(int storage, const void* fnew, const void* fold, const uint32_t* cells, uint32_t n,
                               uint32_t ncell, double w1, void* out, hipStream_t s) {
    const bool raw = fold == nullptr;
    switch (storage) {
        case 0: return xfer_sp<StF32, false>(fnew, fold, cells, n, ncell, w1, out, s);
        case 1: return xfer_sp<StF64, false>(fnew, fold, cells, n, ncell, w1, out, s);
        case 2: return raw ? xfer_sp<StF16S, true>(fnew, fold, cells, n, ncell, w1, out, s)
                           : xfer_sp<StF16S, false>(fnew, fold, cells, n, ncell, w1, out, s);
        case 3: return raw ? xfer_sp<StBF16S, true>(fnew, fold, cells, n, ncell, w1, out, s)
                           : xfer_sp<StBF16S, false>(fnew, fold, cells, n, ncell, w1, out, s);
        default: return hipErrorInvalidValue;
    }
}
hipError_t launch_xfer_rebuild(int storage, bool raw, void* f, uint32_t ncell, const uint32_t* tcells, uint32_t n,
                               const int32_t* slot, const float* w, const void* m, double scale, double keep, void* vel,
                               hipStream_t s) {
    switch (storage) {
        case 0: return rebuild_sp<StF32, false>(f, ncell, tcells, n, slot, w, m, scale, keep, vel, s);
        case 1: return rebuild_sp<StF64, false>(f, ncell, tcells, n, slot, w, m, scale, keep, vel, s);
        case 2: return raw ? rebuild_sp<StF16S, true>(f, ncell, tcells, n, slot, w, m, scale, keep, vel, s)
                           : rebuild_sp<StF16S, false>(f, ncell, tcells, n, slot, w, m, scale, keep, vel, s);
        case 3: return raw ? rebuild_sp<StBF16S, true>(f, ncell, tcells, n, slot, w, m, scale, keep, vel, s)
                           : rebuild_sp<StBF16S, false>(f, ncell, tcells, n, slot, w, m, scale, keep, vel, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace oal
