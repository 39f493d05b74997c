// Device-side building blocks of liboal: D3Q27 constants, population storage forms and the
// cumulant collision of one cell.  gfx950 only.
//
// The collision arithmetic follows src/collision/kernel_cumulant.jl:420-908 operation for operation
// (the library is compiled with -ffp-contract=off, so every multiply and add rounds on its own as
// in the reference), which lets the FP32 path match a CPU restatement of the same arithmetic bit
// for bit.
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
// Storage forms.  ld() returns the working-precision population, sto() stores one.  The 16-bit forms
// hold f - w_k (shifted), which keeps the small deviations from rest that matter.  rt() is the value
// a stored population reads back as.
// ------------------------------------------------------------------------------------------------
struct StF32 {
    using st = float;
    using T = float;
    static constexpr int code = 0;
    __device__ static __forceinline__ T ld(const st* p, T) { return *p; }
    __device__ static __forceinline__ void sto(st* p, T f, T) { *p = f; }
    __device__ static __forceinline__ T rt(T f, T) { return f; }
};
struct StF64 {
    using st = double;
    using T = double;
    static constexpr int code = 1;
    __device__ static __forceinline__ T ld(const st* p, T) { return *p; }
    __device__ static __forceinline__ void sto(st* p, T f, T) { *p = f; }
    __device__ static __forceinline__ T rt(T f, T) { return f; }
};
struct StF16S {
    using st = _Float16;
    using T = float;
    static constexpr int code = 2;
    __device__ static __forceinline__ T ld(const st* p, T w) { return (float)(*p) + w; }
    __device__ static __forceinline__ void sto(st* p, T f, T w) { *p = (_Float16)(f - w); }
    __device__ static __forceinline__ T rt(T f, T w) { return (float)(_Float16)(f - w) + w; }
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
    static constexpr int code = 3;
    __device__ static __forceinline__ T ld(const st* p, T w) { return __uint_as_float((uint32_t)(*p) << 16) + w; }
    __device__ static __forceinline__ void sto(st* p, T f, T w) { *p = f32_to_bf16_rne(f - w); }
    __device__ static __forceinline__ T rt(T f, T w) { return __uint_as_float((uint32_t)f32_to_bf16_rne(f - w) << 16) + w; }
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
    int32_t pad;
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

// The fluid-cell update after the gather: f holds the 27 arriving populations, (rho, jx, jy, jz)
// their sums in direction order.  On return f holds the post-collision populations and rho, u the
// macroscopic state the reference writes to rho_out / vel_out.
template <class T>
__device__ __forceinline__ void collide_cell(T (&f)[27], T rho, T jx, T jy, T jz, T nu_eddy, T tau, T amr_scale,
                                             T Fx, T Fy, T Fz, bool forced, const CollParams& p,
                                             T& rho_o, T& ux_o, T& uy_o, T& uz_o) {
    rho = mx(rho, T(0.01f));
    const T inv_rho = T(1) / rho;
    const T ux = jx * inv_rho, uy = jy * inv_rho, uz = jz * inv_rho;
    const T ux_eq = ux + T(0.5) * Fx * inv_rho;
    const T uy_eq = uy + T(0.5) * Fy * inv_rho;
    const T uz_eq = uz + T(0.5) * Fz * inv_rho;
    rho_o = rho; ux_o = ux; uy_o = uy; uz_o = uz;

    nu_eddy = nu_eddy * amr_scale;
    nu_eddy = mx(nu_eddy, T(p.nu_sgs_bg) * amr_scale);
    const T nu_mol = (tau - T(0.5)) / T(3);
    const T nu_eff = nu_mol + nu_eddy;
    const T tau_eff = T(0.5) + T(3) * nu_eff;
    const T omega1 = T(1) / mx(tau_eff, T(0.500001f));
    const T omega_bulk = T(p.omega_bulk), omega3 = T(p.omega_3), omega4 = T(p.omega_4);
    T omega4_diag = omega4;
    if (p.adaptive_omega4 == 1) {
        const T om1_m2 = omega1 - T(2);
        const T den = T(p.lambda) * omega1 + T(1e-10f);
        const T a_val = -(om1_m2 * om1_m2) / den;
        omega4_diag = mn(mx(a_val, T(0.01f)), T(2));
    }

    // forward transform: x on (0,1,2)+3iy+9iz, y on ma+(0,3,6)+9iz, z on ma+3mb+(0,9,18)
#pragma unroll
    for (int r = 0; r < 9; ++r) chim_f(f[3 * r], f[3 * r + 1], f[3 * r + 2], ux_eq);
#pragma unroll
    for (int iz = 0; iz < 3; ++iz)
#pragma unroll
        for (int ma = 0; ma < 3; ++ma) chim_f(f[ma + 9 * iz], f[ma + 3 + 9 * iz], f[ma + 6 + 9 * iz], uy_eq);
#pragma unroll
    for (int r = 0; r < 9; ++r) chim_f(f[r], f[r + 9], f[r + 18], uz_eq);

    // normalised central moments, index a + 3b + 9c for order (a,b,c)
    const T k000 = f[0];
    const T inv_r = T(1) / mx(k000, T(0.01f));
    const T c200 = f[2] * inv_r, c020 = f[6] * inv_r, c002 = f[18] * inv_r;
    const T c110 = f[4] * inv_r, c101 = f[10] * inv_r, c011 = f[12] * inv_r;
    const T c210 = f[5] * inv_r, c120 = f[7] * inv_r, c201 = f[11] * inv_r;
    const T c021 = f[15] * inv_r, c102 = f[19] * inv_r, c012 = f[21] * inv_r;
    const T c111 = f[13] * inv_r;
    const T c220n = f[8] * inv_r, c202n = f[20] * inv_r, c022n = f[24] * inv_r;
    const T c211n = f[14] * inv_r, c121n = f[16] * inv_r, c112n = f[22] * inv_r;

    const T C220 = c220n - c200 * c020 - T(2) * c110 * c110;
    const T C202 = c202n - c200 * c002 - T(2) * c101 * c101;
    const T C022 = c022n - c020 * c002 - T(2) * c011 * c011;
    const T C211 = c211n - c200 * c011 - T(2) * c110 * c101;
    const T C121 = c121n - c020 * c101 - T(2) * c110 * c011;
    const T C112 = c112n - c002 * c110 - T(2) * c101 * c011;

    // second order: trace with omega_bulk, deviators with omega1 (kernel_cumulant.jl:708-723)
    const T tr = c200 + c020 + c002;
    const T tr_eq = T(3.0f * 0.333333333f);   // Float32 in the reference: rounds to 1
    const T tr_s = (T(1) - omega_bulk) * tr + omega_bulk * tr_eq;
    const T d1 = c200 - c020, d2 = c200 - c002;
    const T d1s = (T(1) - omega1) * d1, d2s = (T(1) - omega1) * d2;
    const T s200 = (tr_s + d1s + d2s) / T(3);
    const T s020 = (tr_s - T(2) * d1s + d2s) / T(3);
    const T s002 = (tr_s + d1s - T(2) * d2s) / T(3);
    T s110 = (T(1) - omega1) * c110;
    T s101 = (T(1) - omega1) * c101;
    T s011 = (T(1) - omega1) * c011;
    if (p.limiter == 1) {   // realizability clamp (kernel_cumulant.jl:726-735)
        const T b110 = sqrt(mx(s200 * s020, T(0)));
        const T b101 = sqrt(mx(s200 * s002, T(0)));
        const T b011 = sqrt(mx(s020 * s002, T(0)));
        s110 = mn(mx(s110, -b110), b110);
        s101 = mn(mx(s101, -b101), b101);
        s011 = mn(mx(s011, -b011), b011);
    }
    const T s210 = (T(1) - omega3) * c210, s120 = (T(1) - omega3) * c120;
    const T s201 = (T(1) - omega3) * c201, s021 = (T(1) - omega3) * c021;
    const T s102 = (T(1) - omega3) * c102, s012 = (T(1) - omega3) * c012;
    const T s111 = (T(1) - omega3) * c111;
    const T S220 = (T(1) - omega4_diag) * C220;
    const T S202 = (T(1) - omega4_diag) * C202;
    const T S022 = (T(1) - omega4_diag) * C022;
    const T S211 = (T(1) - omega4) * C211;
    const T S121 = (T(1) - omega4) * C121;
    const T S112 = (T(1) - omega4) * C112;

    // central moments from the relaxed cumulants; 5th and 6th order cumulants vanish
    const T s220 = S220 + s200 * s020 + T(2) * s110 * s110;
    const T s202 = S202 + s200 * s002 + T(2) * s101 * s101;
    const T s022 = S022 + s020 * s002 + T(2) * s011 * s011;
    const T s211 = S211 + s200 * s011 + T(2) * s110 * s101;
    const T s121 = S121 + s020 * s101 + T(2) * s110 * s011;
    const T s112 = S112 + s002 * s110 + T(2) * s101 * s011;
    const T s221 = s200 * s021 + s020 * s201 + T(4) * s110 * s111 + T(2) * s101 * s120 + T(2) * s011 * s210;
    const T s212 = s200 * s012 + s002 * s210 + T(2) * s110 * s102 + T(4) * s101 * s111 + T(2) * s011 * s201;
    const T s122 = s020 * s102 + s002 * s120 + T(2) * s110 * s012 + T(4) * s011 * s111 + T(2) * s101 * s021;
    T s222 = s200 * S022 + s020 * S202 + s002 * S220 + T(4) * (s110 * S112 + s101 * S121 + s011 * S211);
    s222 += T(2) * (s210 * s012 + s201 * s021 + s120 * s102) + T(4) * s111 * s111;
    s222 += s200 * s020 * s002 + T(2) * (s200 * s011 * s011 + s020 * s101 * s101 + s002 * s110 * s110) +
            T(8) * s110 * s101 * s011;

    const T r = k000;
    f[0] = r;
    f[1] = T(0); f[3] = T(0); f[9] = T(0);
    f[2] = r * s200; f[6] = r * s020; f[18] = r * s002;
    f[4] = r * s110; f[10] = r * s101; f[12] = r * s011;
    f[5] = r * s210; f[7] = r * s120; f[11] = r * s201;
    f[15] = r * s021; f[19] = r * s102; f[21] = r * s012;
    f[13] = r * s111;
    f[8] = r * s220; f[20] = r * s202; f[24] = r * s022;
    f[14] = r * s211; f[16] = r * s121; f[22] = r * s112;
    f[17] = r * s221; f[23] = r * s212; f[25] = r * s122;
    f[26] = r * s222;

    // backward transform z, y, x
#pragma unroll
    for (int q = 0; q < 9; ++q) chim_b(f[q], f[q + 9], f[q + 18], uz_eq);
#pragma unroll
    for (int iz = 0; iz < 3; ++iz)
#pragma unroll
        for (int ma = 0; ma < 3; ++ma) chim_b(f[ma + 9 * iz], f[ma + 3 + 9 * iz], f[ma + 6 + 9 * iz], uy_eq);
#pragma unroll
    for (int q = 0; q < 9; ++q) chim_b(f[3 * q], f[3 * q + 1], f[3 * q + 2], ux_eq);

    if (p.limiter == 2) {   // positivity (kernel_cumulant.jl:849-884)
        T fmin = f[0];
#pragma unroll
        for (int k = 1; k < 27; ++k) fmin = mn(fmin, f[k]);
        if (fmin < T(0)) {
            T alpha = T(1);
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                const T fe = feq_auto(rho, ux_eq, uy_eq, uz_eq, wgt<T>(k), T(dcx(k)), T(dcy(k)), T(dcz(k)), p.compressible);
                const T fn = f[k] - fe;
                if (fn < T(0) && fe > T(1e-15f)) alpha = mn(alpha, fe / (fe - fn + T(1e-15f)));
            }
            alpha = mx(alpha, T(0));
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                const T fe = feq_auto(rho, ux_eq, uy_eq, uz_eq, wgt<T>(k), T(dcx(k)), T(dcy(k)), T(dcz(k)), p.compressible);
                f[k] = fe + alpha * (f[k] - fe);
            }
        }
    }

    if (forced) {   // Guo term of the force density (kernel_cumulant.jl:889-905)
        const T g = T(1) - T(0.5) * omega1;
#pragma unroll
        for (int k = 0; k < 27; ++k) {
            const T cxf = T(dcx(k)), cyf = T(dcy(k)), czf = T(dcz(k));
            const T cu = cxf * ux_eq + cyf * uy_eq + czf * uz_eq;
            const T ft = wgt<T>(k) * T(3) * ((cxf - ux + T(3) * cu * cxf) * Fx + (cyf - uy + T(3) * cu * cyf) * Fy +
                                             (czf - uz + T(3) * cu * czf) * Fz);
            f[k] = f[k] + g * ft;
        }
    }
}

}  // namespace oal
