// Device-side building blocks of liboal: D3Q27 constants, population storage forms and the
// cumulant collision of one cell.  gfx950 only.
//
// The collision is the reference's (src/collision/kernel_cumulant.jl:420-908) with fused
// multiply-adds written out explicitly; the library is compiled with -ffp-contract=off so that no
// other operation is fused, which lets every storage form match a CPU restatement of the same
// arithmetic bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

namespace oal {

// ------------------------------------------------------------------------------------------------
// Lattice: direction k = (cx+1) + 3 (cy+1) + 9 (cz+1)   (src/collision/lattice.jl:16-19)
// ------------------------------------------------------------------------------------------------
__host__ __device__ constexpr int dcx(int k) { return k % 3 - 1; }
__host__ __device__ constexpr int dcy(int k) { return (k / 3) % 3 - 1; }
__host__ __device__ constexpr int dcz(int k) { return k / 9 - 1; }
__host__ __device__ constexpr int dopp(int k) { return 26 - k; }
__host__ __device__ constexpr int dmir_y(int k) { return k + 3 * (2 - 2 * ((k / 3) % 3)); }
__host__ __device__ constexpr int dmir_z(int k) { return k + 9 * (2 - 2 * (k / 9)); }
__host__ __device__ constexpr int dnorm2(int k) { return dcx(k) * dcx(k) + dcy(k) * dcy(k) + dcz(k) * dcz(k); }

// weights in the working precision, as build_lattice_arrays_gpu_d3q27 forms them (timestep.jl:2013)
template <class T>
__host__ __device__ constexpr T wgt(int k) {
    return dnorm2(k) == 0 ? T(8) / T(27) : dnorm2(k) == 1 ? T(2) / T(27) : dnorm2(k) == 2 ? T(1) / T(54) : T(1) / T(216);
}

// the same weights for a direction known at run time only (a select among the four constants)
template <class T>
__device__ __forceinline__ T wgt_rt(int k) {
    constexpr T w0 = wgt<T>(13), w1 = wgt<T>(12), w2 = wgt<T>(9), w3 = wgt<T>(0);
    const int n2 = (k % 3 - 1) * (k % 3 - 1) + ((k / 3) % 3 - 1) * ((k / 3) % 3 - 1) + (k / 9 - 1) * (k / 9 - 1);
    return n2 == 0 ? w0 : n2 == 1 ? w1 : n2 == 2 ? w2 : w3;
}

template <int... Ks, class F>
__device__ __forceinline__ void for_each_dir(std::integer_sequence<int, Ks...>, F&& f) {
    (f(std::integral_constant<int, Ks>{}), ...);
}
template <class F>
__device__ __forceinline__ void for_dirs(F&& f) { for_each_dir(std::make_integer_sequence<int, 27>{}, f); }

// max/min with the tie and NaN behaviour of `(a < b) ? b : a`
template <class T> __device__ __forceinline__ T mx(T a, T b) { return a < b ? b : a; }
template <class T> __device__ __forceinline__ T mn(T a, T b) { return b < a ? b : a; }

// ------------------------------------------------------------------------------------------------
// Storage forms.  ld() returns the working-precision population, sto() stores one rounded to
// nearest (initialisation, uploads), sto_r() stores one with the step's rounding: to nearest, or
// stochastic with the random bits r (the 16-bit forms).  rt_r() is the value a population stored by
// sto_r() reads back as.  sto_r() is the step kernels' store: non-temporal, since a step's output is
// read only by the next step, after the whole level has streamed through the caches (3.5 % faster on a
// 256^3 FP32 level).  The 16-bit forms hold f - w_k (shifted), which keeps the small deviations
// from rest that matter; FP16 holds it times 2^14, an exact power-of-two factor that keeps the
// shifted values of a low-Mach flow in FP16's normal range.
//
// Stochastic rounding: deterministic rounding of a population that barely changes from step to step
// makes the same error every step, a steady source of mass and momentum; unbiased random rounding
// does not.  The random bits come from a counter hash of (cell, direction, time step), so results are
// reproducible and independent of the launch shape and of the streaming pattern.
// ------------------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
// uniform per step
__host__ __device__ __forceinline__ uint32_t round_key(int32_t time_step) {
    return fmix32((uint32_t)time_step * 0x9e3779b9u + 0x632be5abu);
}
// per cell
__host__ __device__ __forceinline__ uint32_t round_cell(uint32_t key, uint32_t cell) { return fmix32(cell ^ key); }
// per population: 32 random bits, the high ones used
__host__ __device__ __forceinline__ uint32_t round_bits(uint32_t hc, int k) {
    return (hc ^ ((uint32_t)k * 0x9e3779b9u)) * 0x85ebca6bu;
}

// Population storage inside the library: block-major.  Each 16^3 block holds its 27 direction arrays
// of 4096 cells contiguously, so the loads and stores of a workgroup stay within one 432-element-wide
// region per block (the host layout f[x,y,z,b,k] is converted on upload and download).
__host__ __device__ __forceinline__ size_t pop_index(uint32_t cell, int k) {
    return ((size_t)(cell >> 12) * 27u + (size_t)k) * 4096u + (cell & 4095u);
}

struct StF32 {
    using st = float;
    using T = float;
    using raw = float;
    static constexpr int code = 0;
    static constexpr bool dithered = false;
    __device__ static __forceinline__ raw ld_raw(const st* p) { return *p; }
    __device__ static __forceinline__ T from_raw(raw r, T) { return r; }
    __device__ static __forceinline__ T ld(const st* p, T) { return *p; }
    __device__ static __forceinline__ void sto(st* p, T f, T) { *p = f; }
    __device__ static __forceinline__ void sto_r(st* p, T f, T, uint32_t, bool) { __builtin_nontemporal_store(f, p); }
    __device__ static __forceinline__ T rt_r(T f, T, uint32_t, bool) { return f; }
};
struct StF64 {
    using st = double;
    using T = double;
    using raw = double;
    static constexpr int code = 1;
    static constexpr bool dithered = false;
    __device__ static __forceinline__ raw ld_raw(const st* p) { return *p; }
    __device__ static __forceinline__ T from_raw(raw r, T) { return r; }
    __device__ static __forceinline__ T ld(const st* p, T) { return *p; }
    __device__ static __forceinline__ void sto(st* p, T f, T) { *p = f; }
    __device__ static __forceinline__ void sto_r(st* p, T f, T, uint32_t, bool) { __builtin_nontemporal_store(f, p); }
    __device__ static __forceinline__ T rt_r(T f, T, uint32_t, bool) { return f; }
};
struct StF16S {
    using st = _Float16;
    using T = float;
    using raw = uint32_t;   // the 16 stored bits, zero-extended
    static constexpr int code = 2;
    static constexpr bool dithered = true;
    static constexpr float scale = 16384.0f, inv_scale = 1.0f / 16384.0f;
    __device__ static __forceinline__ raw ld_raw(const st* p) { return (uint32_t)*(const uint16_t*)p; }
    // decode h * 2^-14 + w: the product is exact (a power-of-two scaling of an FP16 value stays in FP32's
    // normal range), so the fused form rounds once at the same place as the multiply-then-add form
    __device__ static __forceinline__ T from_raw(raw r, T w) {
        return __builtin_fmaf((float)__builtin_bit_cast(_Float16, (uint16_t)r), inv_scale, w);
    }
    __device__ static __forceinline__ T ld(const st* p, T w) { return __builtin_fmaf((float)(*p), inv_scale, w); }
    // encode (f - w) * 2^14 as f 2^14 - w 2^14, one rounding of the exact difference: the same value (scaling
    // by a power of two commutes with rounding away from overflow and underflow)
    __device__ static __forceinline__ float shifted(T f, T w) { return __builtin_fmaf(f, scale, -(w * scale)); }
    // FP32 -> FP16: to nearest, or stochastic with the hardware conversion: the random bits r >> 19
    // added below the result's last place, then truncated (normal and subnormal results alike)
    __device__ static __forceinline__ uint16_t enc(float x, uint32_t r, bool sr) {
        if (sr) {
            typedef _Float16 h2 __attribute__((ext_vector_type(2)));
            const h2 v = __builtin_amdgcn_cvt_sr_f16_f32(h2{0, 0}, x, r, false);
            return __builtin_bit_cast(uint16_t, v[0]);
        }
        return __builtin_bit_cast(uint16_t, (_Float16)x);
    }
    __device__ static __forceinline__ float dec(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }
    __device__ static __forceinline__ void sto(st* p, T f, T w) { *p = (_Float16)((f - w) * scale); }
    __device__ static __forceinline__ void sto_r(st* p, T f, T w, uint32_t r, bool sr) {
        __builtin_nontemporal_store(enc(shifted(f, w), r, sr), (uint16_t*)p);
    }
    __device__ static __forceinline__ T rt_r(T f, T w, uint32_t r, bool sr) {
        return __builtin_fmaf(dec(enc(shifted(f, w), r, sr)), inv_scale, w);
    }
    // packed pairs (two cells along x in one 32-bit word, the lower cell in the low half): decode one
    // half, or replace one half by a cell's rounded value (the same rounding as sto_r)
    __device__ static __forceinline__ T dec_lo(uint32_t p, T w) { return __builtin_fmaf(dec((uint16_t)p), inv_scale, w); }
    __device__ static __forceinline__ T dec_hi(uint32_t p, T w) { return __builtin_fmaf(dec((uint16_t)(p >> 16)), inv_scale, w); }
    template <bool HI>
    __device__ static __forceinline__ uint32_t put(uint32_t old, T f, T w, uint32_t r, bool sr) {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const float x = shifted(f, w);
        if (sr) return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(h2, old), x, r, HI));
        const uint32_t h = __builtin_bit_cast(uint16_t, (_Float16)x);
        return HI ? (old & 0xffffu) | (h << 16) : (old & 0xffff0000u) | h;
    }
};
// bfloat16 round to nearest even, done on integers so that host and device agree bit for bit
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
struct StBF16S {
    using st = uint16_t;
    using T = float;
    using raw = uint32_t;
    static constexpr int code = 3;
    static constexpr bool dithered = true;
    __device__ static __forceinline__ raw ld_raw(const st* p) { return (uint32_t)*p; }
    __device__ static __forceinline__ T from_raw(raw r, T w) { return __uint_as_float(r << 16) + w; }
    __device__ static __forceinline__ T ld(const st* p, T w) { return __uint_as_float((uint32_t)(*p) << 16) + w; }
    // FP32 -> BF16: to nearest even, or stochastic for finite values: the 16 dropped bits plus 16
    // random bits, truncated (subnormals included: their encoding is linear too)
    __device__ static __forceinline__ uint16_t enc(float x, uint32_t r, bool sr) {
        if (sr) {   // the hardware conversion: (u + (r >> 16)) >> 16 for finite values
            typedef __bf16 b2 __attribute__((ext_vector_type(2)));
            const b2 v = __builtin_amdgcn_cvt_sr_bf16_f32(b2{0, 0}, x, r, false);
            return __builtin_bit_cast(uint16_t, v[0]);
        }
        return f32_to_bf16_rne(x);
    }
    __device__ static __forceinline__ void sto(st* p, T f, T w) { *p = f32_to_bf16_rne(f - w); }
    __device__ static __forceinline__ void sto_r(st* p, T f, T w, uint32_t r, bool sr) {
        __builtin_nontemporal_store(enc(f - w, r, sr), p);
    }
    __device__ static __forceinline__ T rt_r(T f, T w, uint32_t r, bool sr) {
        return __uint_as_float((uint32_t)enc(f - w, r, sr) << 16) + w;
    }
    __device__ static __forceinline__ T dec_lo(uint32_t p, T w) { return __uint_as_float(p << 16) + w; }
    __device__ static __forceinline__ T dec_hi(uint32_t p, T w) { return __uint_as_float(p & 0xffff0000u) + w; }
    template <bool HI>
    __device__ static __forceinline__ uint32_t put(uint32_t old, T f, T w, uint32_t r, bool sr) {
        typedef __bf16 b2 __attribute__((ext_vector_type(2)));
        if (sr) return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(b2, old), f - w, r, HI));
        const uint32_t h = f32_to_bf16_rne(f - w);
        return HI ? (old & 0xffffu) | (h << 16) : (old & 0xffff0000u) | h;
    }
};

// ------------------------------------------------------------------------------------------------
// Equilibria (src/collision/physics_utils.jl:30-100) and inlet noise (physics_utils.jl:5-16)
// ------------------------------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T feq_auto(T rho, T ux, T uy, T uz, T w, T cx, T cy, T cz, int compressible) {
    T cu = cx * ux + cy * uy + cz * uz;
    T usq = ux * ux + uy * uy + uz * uz;
    if (compressible == 1) {
        T cu2 = cu * cu, cu3 = cu2 * cu, cu4 = cu2 * cu2, usq2 = usq * usq;
        T val = T(1) + T(3) * cu + T(4.5) * cu2 - T(1.5) * usq;
        T ho = (T(4.5) * cu3 - T(4.5) * cu * usq) + (T(3.375) * cu4 - T(6.75) * cu2 * usq + T(1.125) * usq2);
        val += ho;
        return rho * w * val;
    }
    return rho * w * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq);
}

__device__ __forceinline__ float inlet_noise(int32_t gy1, int32_t gz1, int32_t step, int32_t seed) {
    uint32_t h = (uint32_t)gy1 * 374761393u + (uint32_t)gz1 * 668265263u + (uint32_t)step * 1274126177u + (uint32_t)seed;
    h = (h ^ (h >> 16)) * 0x85ebca6bu;
    h = (h ^ (h >> 13)) * 0xc2b2ae35u;
    h = h ^ (h >> 16);
    return ((float)(h & 0xffffu) / 32768.0f) - 1.0f;
}

// ------------------------------------------------------------------------------------------------
// Collision parameters (the cumulant arguments of the reference kernel)
// ------------------------------------------------------------------------------------------------
struct CollParams {
    float omega_bulk, omega_3, omega_4, lambda;
    int32_t adaptive_omega4, limiter, compressible, bb_mode;
    float c_wale, nu_sgs_bg, inlet_turb;
    int32_t dither;   // 16-bit storage forms: stochastic rounding of the step's stores
    int32_t unit;     // omega_bulk = omega_3 = omega_4 = 1: the pipelined kernels use the UNIT collision
    int32_t tfs;      // transverse freestream at the y and z faces
    float tfs_blend;
};

// WALE eddy viscosity of the gradient g_ij = d u_i / d x_j (kernel_cumulant.jl:500-535)
template <class T>
__device__ __forceinline__ T wale_nu(T g11, T g12, T g13, T g21, T g22, T g23, T g31, T g32, T g33, T c_wale) {
    T gsq11 = g11 * g11 + g12 * g21 + g13 * g31;
    T gsq12 = g11 * g12 + g12 * g22 + g13 * g32;
    T gsq13 = g11 * g13 + g12 * g23 + g13 * g33;
    T gsq21 = g21 * g11 + g22 * g21 + g23 * g31;
    T gsq22 = g21 * g12 + g22 * g22 + g23 * g32;
    T gsq23 = g21 * g13 + g22 * g23 + g23 * g33;
    T gsq31 = g31 * g11 + g32 * g21 + g33 * g31;
    T gsq32 = g31 * g12 + g32 * g22 + g33 * g32;
    T gsq33 = g31 * g13 + g32 * g23 + g33 * g33;
    T tr_term = (gsq11 + gsq22 + gsq33) / T(3);
    T Sd11 = gsq11 - tr_term, Sd22 = gsq22 - tr_term, Sd33 = gsq33 - tr_term;
    T Sd12 = T(0.5) * (gsq12 + gsq21), Sd13 = T(0.5) * (gsq13 + gsq31), Sd23 = T(0.5) * (gsq23 + gsq32);
    T S12 = T(0.5) * (g12 + g21), S13 = T(0.5) * (g13 + g31), S23 = T(0.5) * (g23 + g32);
    T OP1 = Sd11 * Sd11 + Sd22 * Sd22 + Sd33 * Sd33 + T(2) * (Sd12 * Sd12 + Sd13 * Sd13 + Sd23 * Sd23);
    T OP2 = g11 * g11 + g22 * g22 + g33 * g33 + T(2) * (S12 * S12 + S13 * S13 + S23 * S23);
    T nu = T(0);
    if (OP1 > T(1e-12f)) {
        T OP1_32 = OP1 * sqrt(OP1);
        T OP2_52 = OP2 * OP2 * sqrt(mx(OP2, T(1e-12f)));
        T denom = OP2_52 + OP1 * sqrt(sqrt(mx(OP1, T(1e-12f))));
        if (denom > T(1e-12f)) nu = (c_wale * c_wale) * OP1_32 / denom;
    }
    return nu;
}

// One forward chimera triple about u (kernel_cumulant.jl:4-14): (f-, f0, f+) -> (k0, k1, k2)
template <class T>
__device__ __forceinline__ void chim_f(T& a, T& b, T& c, T u) {
    const T m0 = a + b + c;
    const T m1 = c - a;
    const T m2 = a + c;
    a = m0;
    b = m1 - u * m0;
    c = m2 - T(2) * u * m1 + u * u * m0;
}
// and its inverse (kernel_cumulant.jl:16-26)
template <class T>
__device__ __forceinline__ void chim_b(T& a, T& b, T& c, T u) {
    const T m0 = a;
    const T m1 = b + u * a;
    const T m2 = c + T(2) * u * b + u * u * a;
    const T fp = (m2 + m1) * T(0.5);
    const T fm = (m2 - m1) * T(0.5);
    a = fm;
    b = m0 - m2;
    c = fp;
}

template <class T> __device__ __forceinline__ T fmx(T a, T b, T c);
template <> __device__ __forceinline__ float fmx<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double fmx<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
// IEEE minNum: the smaller operand, the other one where one is NaN
__device__ __forceinline__ float fmn(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double fmn(double a, double b) { return __builtin_fmin(a, b); }

// omega1 and the diagonal 4th-order rate of a cell (kernel_cumulant.jl:592-610); computed per cell
// only with WALE, otherwise once per domain on the host with the same operations
template <class T>
__host__ __device__ inline void cell_rates(T nu_eddy, T tau, T amr_scale, T nu_bg, int adaptive, T lambda, T omega4,
                                           T& omega1, T& omega4_diag) {
    nu_eddy = nu_eddy * amr_scale;
    const T bg = nu_bg * amr_scale;
    nu_eddy = nu_eddy < bg ? bg : nu_eddy;
    const T nu_mol = (tau - T(0.5)) / T(3);
    const T nu_eff = nu_mol + nu_eddy;
    const T tau_eff = T(0.5) + T(3) * nu_eff;
    omega1 = T(1) / (tau_eff < T(0.500001f) ? T(0.500001f) : tau_eff);
    omega4_diag = omega4;
    if (adaptive == 1) {
        const T om1_m2 = omega1 - T(2);
        const T den = lambda * omega1 + T(1e-10f);
        const T a_val = -(om1_m2 * om1_m2) / den;
        const T lo = a_val < T(0.01f) ? T(0.01f) : a_val;
        omega4_diag = T(2) < lo ? T(2) : lo;
    }
}

// Per-cell constants of the collision
template <class T>
struct CellRates {
    T omega1, omega4_diag;
    T Fx, Fy, Fz;
    bool forced;
    // sponge (read by the SPONGE form only): strength sp of the cell, the inlet velocity it blends
    // toward, and whether the populations are blended too
    T sponge, sponge_u;
    int sponge_blend;
};

// The fluid-cell update after the gather (kernel_cumulant.jl:420-908): f holds the 27 arriving
// populations in direction order; on return the post-collision ones, with rho and u the macroscopic
// state the reference writes to rho_out / vel_out.
//
// Arithmetic: the reference's cumulant collision with fused multiply-adds, and with rho and j read
// off the raw partial sums of the transform passes: the x pass turns each row (f-, f0, f+) into
// (m0, d = f+ - f-, s = f- + f+); rho and j are sums of those over rows; the x, y and z passes then
// shift to central moments about u (k1 = d - u m0, k2 = s - 2 u d + u^2 m0).  The total of the last
// pass equals rho bit for bit, so 1/max(k000, 0.01) is the 1/max(rho, 0.01) already formed.
//
// UNIT: the reference's default rates omega_bulk = omega_3 = omega_4 = 1 (config_loader.jl:154-159), for
// which the relaxed trace is 1 and the 3rd-order, mixed 4th-order and 5th-order moments are zero: the
// terms they would contribute are left out, and the first backward pass uses the zeros it knows.  Each
// term left out adds an exact zero in the general form, so both give the same values (only the sign of
// a result that is exactly zero could differ).
//
// SPONGE: the sponge blend (kernel_cumulant.jl:428-446) of a cell with q.sponge > 0: rho and u blended
// toward (1, u_inlet) before the collision uses them, and with q.sponge_blend the populations toward the
// inlet equilibrium, the transform then summing the blended populations.
template <class T, bool UNIT = false, bool SPONGE = false>
__device__ __forceinline__ void collide_cell(T (&f)[27], const CellRates<T>& q, const CollParams& p,
                                             T& rho_o, T& ux_o, T& uy_o, T& uz_o) {
    T A[27];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const T s = f[3 * r] + f[3 * r + 2];
        const T d = f[3 * r + 2] - f[3 * r];
        A[3 * r] = s + f[3 * r + 1];
        A[3 * r + 1] = d;
        A[3 * r + 2] = s;
    }
    T y0[3], yd0[3], y1[3];
#pragma unroll
    for (int iz = 0; iz < 3; ++iz) {
        const T a = A[9 * iz], b = A[3 + 9 * iz], c = A[6 + 9 * iz];
        const T ys = a + c;
        yd0[iz] = c - a;
        y0[iz] = ys + b;
        y1[iz] = (A[1 + 9 * iz] + A[7 + 9 * iz]) + A[4 + 9 * iz];
    }
    const T rho = (y0[0] + y0[2]) + y0[1];
    const T jz = y0[2] - y0[0];
    const T jy = (yd0[0] + yd0[2]) + yd0[1];
    const T jx = (y1[0] + y1[2]) + y1[1];

    const T rc = mx(rho, T(0.01f));
    const T inv = T(1) / rc;
    T ux = jx * inv, uy = jy * inv, uz = jz * inv;
    // rb: the density the collision reports and its limiter uses; rr, invr: the mass of the populations
    // the transform sums and 1 / max(rr, 0.01) (the sponge's population blend changes them)
    T rb = rc, rr = rho, invr = inv;
    if constexpr (SPONGE) {
        const T sp = q.sponge;
        if (sp > T(0)) {
            const T g = T(1) - sp;
            rb = rc * g + T(1) * sp;
            ux = ux * g + q.sponge_u * sp;
            uy = uy * g;
            uz = uz * g;
            if (q.sponge_blend == 1) {
#pragma unroll
                for (int k = 0; k < 27; ++k)
                    f[k] = f[k] * g + feq_auto(T(1), q.sponge_u, T(0), T(0), wgt<T>(k), T(dcx(k)), T(dcy(k)), T(dcz(k)),
                                               p.compressible) * sp;
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    const T s = f[3 * r] + f[3 * r + 2];
                    const T d = f[3 * r + 2] - f[3 * r];
                    A[3 * r] = s + f[3 * r + 1];
                    A[3 * r + 1] = d;
                    A[3 * r + 2] = s;
                }
                T z0[3];
#pragma unroll
                for (int iz = 0; iz < 3; ++iz) z0[iz] = (A[9 * iz] + A[6 + 9 * iz]) + A[3 + 9 * iz];
                rr = (z0[0] + z0[2]) + z0[1];
                invr = T(1) / mx(rr, T(0.01f));
            }
        }
    }
    T uxe = ux, uye = uy, uze = uz;
    if (q.forced) {
        uxe = fmx(T(0.5) * q.Fx, inv, ux);
        uye = fmx(T(0.5) * q.Fy, inv, uy);
        uze = fmx(T(0.5) * q.Fz, inv, uz);
    }
    rho_o = rb; ux_o = ux; uy_o = uy; uz_o = uz;

    {
        const T u = uxe, uu = u * u, t = T(-2) * u;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const T m0 = A[3 * r], d = A[3 * r + 1], s = A[3 * r + 2];
            A[3 * r + 1] = fmx(-u, m0, d);
            A[3 * r + 2] = fmx(uu, m0, fmx(t, d, s));
        }
    }
    auto pass = [&](int i0, int i1, int i2, T u) {
        const T uu = u * u, t = T(-2) * u;
        const T a = A[i0], b = A[i1], c = A[i2];
        const T ys = a + c, yd = c - a, ym0 = ys + b;
        A[i0] = ym0;
        A[i1] = fmx(-u, ym0, yd);
        A[i2] = fmx(uu, ym0, fmx(t, yd, ys));
    };
#pragma unroll
    for (int iz = 0; iz < 3; ++iz)
#pragma unroll
        for (int ma = 0; ma < 3; ++ma) pass(ma + 9 * iz, ma + 3 + 9 * iz, ma + 6 + 9 * iz, uye);
#pragma unroll
    for (int r = 0; r < 9; ++r) pass(r, r + 9, r + 18, uze);

    // normalised central moments (index a + 3b + 9c for order (a,b,c)) and 4th-order cumulants
    const T c200 = A[2] * invr, c020 = A[6] * invr, c002 = A[18] * invr;
    const T c110 = A[4] * invr, c101 = A[10] * invr, c011 = A[12] * invr;
    const T c220n = A[8] * invr, c202n = A[20] * invr, c022n = A[24] * invr;
    const T x110 = T(2) * c110, x101 = T(2) * c101, x011 = T(2) * c011;
    const T C220 = fmx(-x110, c110, fmx(-c200, c020, c220n));
    const T C202 = fmx(-x101, c101, fmx(-c200, c002, c202n));
    const T C022 = fmx(-x011, c011, fmx(-c020, c002, c022n));

    // relaxation (kernel_cumulant.jl:705-754): trace toward 1 at omega_bulk, deviators at omega1,
    // realizability clamp, 3rd order at omega3, 4th-order cumulants at omega4 / omega4_diag
    const T third = T(1) / T(3);
    const T ob = T(p.omega_bulk);
    const T tr = (c200 + c020) + c002;
    const T tr_s = UNIT ? T(1) : fmx(T(1) - ob, tr, ob);
    const T g1 = T(1) - q.omega1;
    const T d1s = g1 * (c200 - c020), d2s = g1 * (c200 - c002);
    const T t3 = tr_s * third;
    const T s200 = fmx(d1s + d2s, third, t3);
    const T s020 = fmx(fmx(T(-2), d1s, d2s), third, t3);
    const T s002 = fmx(fmx(T(-2), d2s, d1s), third, t3);
    T s110 = g1 * c110, s101 = g1 * c101, s011 = g1 * c011;
    if (p.limiter == 1) {
        const T q110 = mx(s200 * s020, T(0)), q101 = mx(s200 * s002, T(0)), q011 = mx(s020 * s002, T(0));
        const bool l110 = s110 * s110 > q110, l101 = s101 * s101 > q101, l011 = s011 * s011 > q011;
        // the clamp is rare: a wave takes the square roots only when one of its lanes needs one
        if (__builtin_amdgcn_ballot_w64(l110 || l101 || l011) != 0) {
            if (l110) s110 = copysign(sqrt(q110), s110);
            if (l101) s101 = copysign(sqrt(q101), s101);
            if (l011) s011 = copysign(sqrt(q011), s011);
        }
    }
    const T g4d = T(1) - q.omega4_diag;
    const T S220 = g4d * C220, S202 = g4d * C202, S022 = g4d * C022;
    const T y110 = T(2) * s110, y101 = T(2) * s101, y011 = T(2) * s011;
    const T s220 = fmx(y110, s110, fmx(s200, s020, S220));
    const T s202 = fmx(y101, s101, fmx(s200, s002, S202));
    const T s022 = fmx(y011, s011, fmx(s020, s002, S022));
    const T r = rr;
    A[0] = r;
    A[1] = T(0); A[3] = T(0); A[9] = T(0);
    A[2] = r * s200; A[6] = r * s020; A[18] = r * s002;
    A[4] = r * s110; A[10] = r * s101; A[12] = r * s011;
    A[8] = r * s220; A[20] = r * s202; A[24] = r * s022;
    // central moments from the relaxed cumulants; 5th and 6th order cumulants vanish
    if constexpr (UNIT) {
        const T s211 = fmx(y110, s101, s200 * s011);
        const T s121 = fmx(y110, s011, s020 * s101);
        const T s112 = fmx(y101, s011, s002 * s110);
        T e = s200 * S022;
        e = fmx(s020, S202, e);
        e = fmx(s002, S220, e);
        e = fmx(s200 * s020, s002, e);
        e = fmx(s200 * y011, s011, e);
        e = fmx(s020 * y101, s101, e);
        e = fmx(s002 * y110, s110, e);
        e = fmx((T(4) * y110) * s101, s011, e);
        A[14] = r * s211; A[16] = r * s121; A[22] = r * s112;
        A[26] = r * e;
    } else {
        const T c210 = A[5] * invr, c120 = A[7] * invr, c201 = A[11] * invr;
        const T c021 = A[15] * invr, c102 = A[19] * invr, c012 = A[21] * invr;
        const T c111 = A[13] * invr;
        const T c211n = A[14] * invr, c121n = A[16] * invr, c112n = A[22] * invr;
        const T C211 = fmx(-x110, c101, fmx(-c200, c011, c211n));
        const T C121 = fmx(-x110, c011, fmx(-c020, c101, c121n));
        const T C112 = fmx(-x101, c011, fmx(-c002, c110, c112n));
        const T g3 = T(1) - T(p.omega_3), g4 = T(1) - T(p.omega_4);
        const T s210 = g3 * c210, s120 = g3 * c120, s201 = g3 * c201, s021 = g3 * c021;
        const T s102 = g3 * c102, s012 = g3 * c012, s111 = g3 * c111;
        const T S211 = g4 * C211, S121 = g4 * C121, S112 = g4 * C112;
        const T s211 = fmx(y110, s101, fmx(s200, s011, S211));
        const T s121 = fmx(y110, s011, fmx(s020, s101, S121));
        const T s112 = fmx(y101, s011, fmx(s002, s110, S112));
        const T s221 = fmx(y011, s210, fmx(y101, s120, fmx(T(2) * y110, s111, fmx(s020, s201, s200 * s021))));
        const T s212 = fmx(y011, s201, fmx(T(2) * y101, s111, fmx(y110, s102, fmx(s002, s210, s200 * s012))));
        const T s122 = fmx(y101, s021, fmx(T(2) * y011, s111, fmx(y110, s012, fmx(s002, s120, s020 * s102))));
        T e = s200 * S022;
        e = fmx(s020, S202, e);
        e = fmx(s002, S220, e);
        e = fmx(T(2) * y110, S112, e);
        e = fmx(T(2) * y101, S121, e);
        e = fmx(T(2) * y011, S211, e);
        e = fmx(T(2) * s210, s012, e);
        e = fmx(T(2) * s201, s021, e);
        e = fmx(T(2) * s120, s102, e);
        e = fmx(T(4) * s111, s111, e);
        e = fmx(s200 * s020, s002, e);
        e = fmx(s200 * y011, s011, e);
        e = fmx(s020 * y101, s101, e);
        e = fmx(s002 * y110, s110, e);
        e = fmx((T(4) * y110) * s101, s011, e);
        A[5] = r * s210; A[7] = r * s120; A[11] = r * s201;
        A[15] = r * s021; A[19] = r * s102; A[21] = r * s012;
        A[13] = r * s111;
        A[14] = r * s211; A[16] = r * s121; A[22] = r * s112;
        A[17] = r * s221; A[23] = r * s212; A[25] = r * s122;
        A[26] = r * e;
    }

    // backward passes z, y, x: (k0, k1, k2) -> (f-, f0, f+) about u
    auto back = [&](int i0, int i1, int i2, T u) {
        const T uu = u * u, t = T(2) * u;
        const T a = A[i0], b = A[i1], c = A[i2];
        const T m1 = fmx(u, a, b);
        const T m2 = fmx(uu, a, fmx(t, b, c));
        const T h = T(0.5) * m1;
        A[i0] = fmx(T(0.5), m2, -h);
        A[i1] = a - m2;
        A[i2] = fmx(T(0.5), m2, h);
    };
    if constexpr (UNIT) {
        // the z pass with its known zeros: rows (r, r + 9, r + 18) with r even have k1 = 0 (1st, 3rd and
        // 5th-order moments), rows with r odd have k0 = k2 = 0
        const T u = uze, uu = u * u, t = T(2) * u;
#pragma unroll
        for (int r2 = 0; r2 < 9; r2 += 2) {
            const T a = A[r2], c = A[r2 + 18];
            const T m1 = u * a;
            const T m2 = fmx(uu, a, c);
            const T h = T(0.5) * m1;
            A[r2] = fmx(T(0.5), m2, -h);
            A[r2 + 9] = a - m2;
            A[r2 + 18] = fmx(T(0.5), m2, h);
        }
#pragma unroll
        for (int r2 = 1; r2 < 9; r2 += 2) {
            const T b = A[r2 + 9];
            const T m2 = t * b;
            const T h = T(0.5) * b;
            A[r2] = fmx(T(0.5), m2, -h);
            A[r2 + 9] = -m2;
            A[r2 + 18] = fmx(T(0.5), m2, h);
        }
    } else {
#pragma unroll
        for (int r2 = 0; r2 < 9; ++r2) back(r2, r2 + 9, r2 + 18, uze);
    }
#pragma unroll
    for (int iz = 0; iz < 3; ++iz)
#pragma unroll
        for (int ma = 0; ma < 3; ++ma) back(ma + 9 * iz, ma + 3 + 9 * iz, ma + 6 + 9 * iz, uye);
#pragma unroll
    for (int r2 = 0; r2 < 9; ++r2) back(3 * r2, 3 * r2 + 1, 3 * r2 + 2, uxe);

    if (p.limiter == 2) {   // positivity (kernel_cumulant.jl:849-884)
        // The running minimum m = mn(m, A[k]) from m = A[0] is negative exactly when A[0] is not NaN
        // (a NaN first value stays) and some A[k] is negative (a NaN A[k] is passed over): the same
        // test as a min tree (fmin passes over NaN) with A[0] checked, 13 three-way minima.
        T m9[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) m9[i] = fmn(fmn(A[3 * i], A[3 * i + 1]), A[3 * i + 2]);
        const T fmin = fmn(fmn(fmn(fmn(m9[0], m9[1]), m9[2]), fmn(fmn(m9[3], m9[4]), m9[5])),
                           fmn(fmn(m9[6], m9[7]), m9[8]));
        if (fmin < T(0) && A[0] == A[0]) {
            T alpha = T(1);
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                const T fe = feq_auto(rb, uxe, uye, uze, wgt<T>(k), T(dcx(k)), T(dcy(k)), T(dcz(k)), p.compressible);
                const T fn = A[k] - fe;
                if (fn < T(0) && fe > T(1e-15f)) alpha = mn(alpha, fe / (fe - fn + T(1e-15f)));
            }
            alpha = mx(alpha, T(0));
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                const T fe = feq_auto(rb, uxe, uye, uze, wgt<T>(k), T(dcx(k)), T(dcy(k)), T(dcz(k)), p.compressible);
                A[k] = fe + alpha * (A[k] - fe);
            }
        }
    }
    if (q.forced) {   // Guo term of the force density (kernel_cumulant.jl:889-905)
        const T g = T(1) - T(0.5) * q.omega1;
#pragma unroll
        for (int k = 0; k < 27; ++k) {
            const T cxf = T(dcx(k)), cyf = T(dcy(k)), czf = T(dcz(k));
            const T cu = cxf * uxe + cyf * uye + czf * uze;
            const T ft = wgt<T>(k) * T(3) * ((cxf - ux + T(3) * cu * cxf) * q.Fx + (cyf - uy + T(3) * cu * cyf) * q.Fy +
                                             (czf - uz + T(3) * cu * czf) * q.Fz);
            A[k] = A[k] + g * ft;
        }
    }
#pragma unroll
    for (int k = 0; k < 27; ++k) f[k] = A[k];
}

}  // namespace oal
