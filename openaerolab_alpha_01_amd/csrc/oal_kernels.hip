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

template <class SP, bool RAW>
__global__ void __launch_bounds__(WG) k_xfer_gather(const typename SP::st* fnew, const typename SP::st* fold,
                                                    const uint32_t* cells, uint32_t n, uint32_t ncell, typename SP::T w1,
                                                    typename SP::T* out) {
    using T = typename SP::T;
    const uint32_t i = blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    const uint32_t e = cells[i];
    const uint32_t c = e & 0x7fffffffu;
    if constexpr (RAW) {
        uint16_t codes[XFER_RAW_STRIDE];
#pragma unroll
        for (int k = 27; k < XFER_RAW_STRIDE; ++k) codes[k] = 0;
        for_dirs([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            codes[K] = (uint16_t)SP::ld_raw(fnew + pop_index(c, K));
        });
        uint4* o = reinterpret_cast<uint4*>(reinterpret_cast<uint16_t*>(out) + (size_t)i * XFER_RAW_STRIDE);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            o[j] = make_uint4(codes[8 * j] | ((uint32_t)codes[8 * j + 1] << 16), codes[8 * j + 2] | ((uint32_t)codes[8 * j + 3] << 16),
                              codes[8 * j + 4] | ((uint32_t)codes[8 * j + 5] << 16), codes[8 * j + 6] | ((uint32_t)codes[8 * j + 7] << 16));
        return;
    }
    const bool blend = fold && !(e >> 31);
    const T w0 = T(1) - w1;
    T vals[XFER_STRIDE];
    vals[XFER_STRIDE - 1] = T(0);
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        T v = SP::ld(fnew + pop_index(c, K), wgt<T>(K));
        if (blend) v = SP::ld(fold + pop_index(c, K), wgt<T>(K)) * w0 + v * w1;
        vals[K] = v;
    });
    // one 16-byte-aligned record per donor, written as whole 16-byte vectors
    using V = typename Vec16<T>::type;
    constexpr int VN = Vec16<T>::n;
    V* o = reinterpret_cast<V*>(out + (size_t)i * XFER_STRIDE);
#pragma unroll
    for (int j = 0; j < XFER_STRIDE / VN; ++j) {
        V v;
#pragma unroll
        for (int e = 0; e < VN; ++e) v[e] = vals[j * VN + e];
        o[j] = v;
    }
}

template <class SP, bool RAW>
__global__ void __launch_bounds__(WG) k_xfer_rebuild(typename SP::st* f, uint32_t ncell, const uint32_t* tcells, uint32_t n,
                                                     const int32_t* slot, const float* wt, const typename SP::T* m,
                                                     typename SP::T scale, typename SP::T keep, typename SP::T* vel) {
    using T = typename SP::T;
    const uint32_t i = blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    T acc[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) acc[k] = T(0);
    using V = typename Vec16<T>::type;
    constexpr int VN = Vec16<T>::n;
    // the 64 nodes in groups of G: the group's slots and weights, then all of its records, are in
    // flight at once (the compiler would otherwise wait for each node's loads in turn); an unused node
    // reads record 0 and leaves acc unchanged through a select; the sums are formed in node order
    constexpr int G = RAW ? 8 : (sizeof(T) == 8 ? 2 : 4);
    for (int nb = 0; nb < 64; nb += G) {
        int32_t sr[G];
        T wn[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            sr[g] = slot[(size_t)(nb + g) * n + i];
            wn[g] = T(wt[(size_t)(nb + g) * n + i]);
        }
        if constexpr (RAW) {
            uint4 rec[G][4];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const uint4* row = reinterpret_cast<const uint4*>(reinterpret_cast<const uint16_t*>(m) +
                                                                  (size_t)(sr[g] >= 0 ? sr[g] : 0) * XFER_RAW_STRIDE);
#pragma unroll
                for (int j = 0; j < 4; ++j) rec[g][j] = row[j];
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const bool use = sr[g] >= 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t wd[4] = {rec[g][j].x, rec[g][j].y, rec[g][j].z, rec[g][j].w};
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int k = 8 * j + e;
                        if (k < 27) {
                            const T add = acc[k] + wn[g] * SP::from_raw((wd[e >> 1] >> (16 * (e & 1))) & 0xffffu, wgt<T>(k));
                            acc[k] = use ? add : acc[k];
                        }
                    }
                }
            }
        } else {
            V rec[G][XFER_STRIDE / VN];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const V* row = reinterpret_cast<const V*>(m + (size_t)(sr[g] >= 0 ? sr[g] : 0) * XFER_STRIDE);
#pragma unroll
                for (int j = 0; j < XFER_STRIDE / VN; ++j) rec[g][j] = row[j];
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const bool use = sr[g] >= 0;
#pragma unroll
                for (int j = 0; j < XFER_STRIDE / VN; ++j)
#pragma unroll
                    for (int e = 0; e < VN; ++e)
                        if (j * VN + e < 27) {
                            const T add = acc[j * VN + e] + wn[g] * rec[g][j][e];
                            acc[j * VN + e] = use ? add : acc[j * VN + e];
                        }
            }
        }
    }
    // moments of the interpolated state
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
    T out[27], eq[27];
    bool bad = false;
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        const T cx = T(dcx(K)), cy = T(dcy(K)), cz = T(dcz(K));
        constexpr T w = wgt<T>(K);
        const T cac = ((cx * cx * dxx + cy * cy * dyy) + cz * cz * dzz) + T(2) * ((cx * cy * dxy + cx * cz * dxz) + cy * cz * dyz);
        out[K] = acc[K] + (w * T(4.5)) * (cac - tr * third);
        const T cu = (cx * ux + cy * uy) + cz * uz;
        eq[K] = (w * rho) * (((T(1) + T(3) * cu) + T(4.5) * cu * cu) - T(1.5) * usq);
        bad = bad || out[K] < T(1e-6) * w;
    });
    const uint32_t c = tcells[i];
    for_dirs([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        SP::sto(f + pop_index(c, K), bad ? eq[K] : out[K], wgt<T>(K));
    });
    // WALE: the written cell's velocity (that of the interpolated state, which the rescale and the
    // equilibrium fallback keep) for the receiving level's gradient gather
    if (vel) {
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
    hipLaunchKernelGGL((k_xfer_gather<SP, RAW>), dim3((n + WG - 1) / WG), dim3(WG), 0, s, (const typename SP::st*)fnew,
                       (const typename SP::st*)fold, cells, n, ncell, T(w1), (T*)out);
    return hipGetLastError();
}
template <class SP, bool RAW>
static hipError_t rebuild_sp(void* f, uint32_t ncell, const uint32_t* tcells, uint32_t n, const int32_t* slot, const float* w,
                             const void* m, double scale, double keep, void* vel, hipStream_t s) {
    using T = typename SP::T;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL((k_xfer_rebuild<SP, RAW>), dim3((n + WG - 1) / WG), dim3(WG), 0, s, (typename SP::st*)f, ncell, tcells, n,
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
