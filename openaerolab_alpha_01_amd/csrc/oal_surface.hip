// Surface-stress forces of liboal for gfx950 (src/forces/surface.jl, the default Prandtl method): one
// lane per triangle probes density and velocity along the triangle's normal, turns them into a
// pressure and a wall-law shear stress (FP32, in the reference's order of operations), and the
// workgroup sums the triangles' FP64 force and moment in a fixed tree; one more workgroup sums the
// workgroups' records in a fixed order.  The probe reads the stored populations of the few cells it
// touches (27 loads per cell), so no level-wide density or velocity field is kept.
#include <hip/hip_runtime.h>

#include "oal_kernels.h"

namespace oal {
namespace {

// The level's density and velocity at one cell: FP64 sums of f_k - w_k over the stored populations
// (the moments of k_macro), rho clamped to >= 0.01 as stream_collide_cumulant! clamps the density it
// writes (src/collision/kernel_cumulant.jl:421-425), then rounded to FP32.
template <class SP>
__device__ __forceinline__ void cell_rho_u(const typename SP::st* f, uint32_t cell, int permute, float& r, float& ux,
                                           float& uy, float& uz) {
    using T = typename SP::T;
    double dr = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const double s = (double)SP::ld(f + pop_index(cell, permute ? 26 - k : k), wgt<T>(k)) - (double)wgt<T>(k);
        dr += s; jx += s * dcx(k); jy += s * dcy(k); jz += s * dcz(k);
    }
    double rr = 1.0 + dr;
    if (rr < 0.01) rr = 0.01;
    r = (float)rr;
    ux = (float)(jx / rr); uy = (float)(jy / rr); uz = (float)(jz / rr);
}

// the fluid cell at 0-based cell coordinates of the domain, -1 when the position lies outside the
// extent, in an absent block or on a solid cell (get_cell_at_position and the obstacle test of the
// samplers, surface.jl:271-298, 530)
__device__ __forceinline__ int64_t fluid_cell(const SurfaceArgs& a, int gx, int gy, int gz) {
    if (gx < 0 || gy < 0 || gz < 0 || gx >= a.ext[0] || gy >= a.ext[1] || gz >= a.ext[2]) return -1;
    const int bx = gx >> 4, by = gy >> 4, bz = gz >> 4;
    const int32_t b = a.bmap[((int64_t)bz * a.nby + by) * a.nbx + bx];
    if (b < 0) return -1;
    const int64_t c = (int64_t)b * BS3 + (((gz & 15) * BS + (gy & 15)) * BS + (gx & 15));
    if (a.solid && a.solid[c]) return -1;
    return c;
}

__device__ __forceinline__ float fmax_j(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float fmin_j(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float clamp01(float x) { return x > 1.0f ? 1.0f : (x < 0.0f ? 0.0f : x); }
// floor of a sampling coordinate; false for coordinates no cell index reaches (and NaN)
__device__ __forceinline__ bool floor_index(float x, int& i) {
    if (!(x > -1.0e9f && x < 1.0e9f)) return false;
    i = (int)floorf(x);
    return true;
}

// trilinear_sample_flow (surface.jl:502-632): the eight cells around the probe point, solid or
// missing ones left out, renormalised by the weight found; ok when that weight exceeds 0.01 and the
// result is finite.  wq receives the weight (the probe's quality).
template <class SP>
__device__ bool sample_flow(const SurfaceArgs& a, float px, float py, float pz, float& r, float& ux, float& uy,
                            float& uz, float& wq) {
    r = 1.0f; ux = 0.0f; uy = 0.0f; uz = 0.0f; wq = 0.0f;
    const float inv_dx = 1.0f / a.dx;
    const float fx = px * inv_dx - 0.5f;
    const float fy = py * inv_dx - 0.5f;
    const float fz = pz * inv_dx - 0.5f;
    int ix0, iy0, iz0;
    if (!floor_index(fx, ix0) || !floor_index(fy, iy0) || !floor_index(fz, iz0)) return false;
    const float wx = clamp01(fx - (float)ix0), wy = clamp01(fy - (float)iy0), wz = clamp01(fz - (float)iz0);
    const float w0x = 1.0f - wx, w1x = wx, w0y = 1.0f - wy, w1y = wy, w0z = 1.0f - wz, w1z = wz;
    float racc = 0.0f, xacc = 0.0f, yacc = 0.0f, zacc = 0.0f, wacc = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int dx = n & 1, dy = (n >> 1) & 1, dz = n >> 2;
        const float wk = (dx ? w1x : w0x) * (dy ? w1y : w0y) * (dz ? w1z : w0z);
        const int64_t c = fluid_cell(a, ix0 + dx, iy0 + dy, iz0 + dz);
        if (c < 0) continue;
        float cr, cu, cv, cw;
        cell_rho_u<SP>((const typename SP::st*)a.f, (uint32_t)c, a.permute, cr, cu, cv, cw);
        racc += wk * cr;
        xacc += wk * cu;
        yacc += wk * cv;
        zacc += wk * cw;
        wacc += wk;
    }
    if (wacc > 0.01f) {
        const float inv_w = 1.0f / wacc;
        const float ro = racc * inv_w, uo = xacc * inv_w, vo = yacc * inv_w, wo = zacc * inv_w;
        if (!isfinite(ro) || !isfinite(uo) || !isfinite(vo) || !isfinite(wo)) return false;
        r = ro; ux = uo; uy = vo; uz = wo; wq = wacc;
        return true;
    }
    return false;
}

// compute_stress_from_cell (surface.jl:180-269), athermal (T = 1): gauge pressure (rho - 1) cs2 and the
// wall-law shear rho nu_eff |u_t| / max(d, 1) along the tangential velocity u_t, with the bounded SGS
// estimate nu_sgs = min(max(C^2 |u_t| / max(d, 1), nu_bg), 20 nu_mol); zeros for an unphysical input
__device__ void stress_from_cell(const SurfaceArgs& a, float rho, float ux, float uy, float uz, float nx, float ny,
                                 float nz, float dist, float& p, float& tx, float& ty, float& tz) {
    p = 0.0f; tx = 0.0f; ty = 0.0f; tz = 0.0f;
    if (!isfinite(rho) || rho < 0.01f || rho > 10.0f) return;
    if (!isfinite(ux) || !isfinite(uy) || !isfinite(uz)) return;
    const float umag2 = ux * ux + uy * uy + uz * uz;
    if (umag2 > 100.0f) return;
    const float pp = (rho - 1.0f) * a.cs2 * a.pressure_scale;
    const float udn = ux * nx + uy * ny + uz * nz;
    const float utx = ux - udn * nx, uty = uy - udn * ny, utz = uz - udn * nz;
    const float utan = sqrtf(utx * utx + uty * uty + utz * utz);
    const float nu_mol = (a.tau - 0.5f) * a.cs2;
    float nu_sgs = 0.0f;
    const float deff = fmax_j(dist, 1.0f);
    if (a.c_wale > 0.0f && utan > 1.0e-10f) nu_sgs = (a.c_wale * a.c_wale) * (utan / deff);
    nu_sgs = fmax_j(nu_sgs, a.nu_sgs_bg);
    nu_sgs = fmin_j(nu_sgs, 20.0f * nu_mol);
    const float nu_eff = nu_mol + nu_sgs;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    if (utan > 1.0e-10f && dist > 0.01f) {
        const float tmag = rho * nu_eff * utan / deff * a.stress_scale;
        sx = (utx / utan) * tmag;
        sy = (uty / utan) * tmag;
        sz = (utz / utan) * tmag;
    }
    if (!isfinite(pp) || !isfinite(sx) || !isfinite(sy) || !isfinite(sz)) return;
    p = pp; tx = sx; ty = sy; tz = sz;
}

// map_stresses_kernel! (surface.jl:636-1033, Prandtl path) for one triangle, then its FP64 force and
// moment (integrate_forces_kernel!, :1265-1331) into the workgroup's fixed-order sum
template <class SP>
__global__ void __launch_bounds__(SURF_WG) k_surface(SurfaceArgs a) {
    __shared__ double red[SURF_SUMS][SURF_WG];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * SURF_WG + tid;
    double contrib[SURF_SUMS];
#pragma unroll
    for (int j = 0; j < SURF_SUMS; ++j) contrib[j] = 0.0;
    if (i < a.n_tri) {
        const int64_t n = a.n_tri;
        const float tx = a.cen[i] + a.off[0], ty = a.cen[n + i] + a.off[1], tz = a.cen[2 * n + i] + a.off[2];
        float nx = a.nrm[i], ny = a.nrm[n + i], nz = a.nrm[2 * n + i];
        const float dx = a.dx;
        float best_rho = 1.0f, bux = 0.0f, buy = 0.0f, buz = 0.0f;
        float best_wd = (float)(a.search_radius + 1), best_q = 0.0f;
        bool found = false;
        float prho = 0.0f, pw = 0.0f;     // quality-weighted pressure probes (Shepard average)
        const float min_q = 0.25f;
        const int pmax = a.search_radius < 4 ? a.search_radius : 4;
        // pass 1: the closest probe of quality >= 0.25 along the normal locks the shear probe; the
        // first pmax probes of that quality average the density for the pressure
        for (int pi = 1; pi <= a.search_radius; ++pi) {
            const float d = (float)pi + 0.5f;
            float r, ux, uy, uz, wq;
            const bool ok = sample_flow<SP>(a, tx + nx * d * dx, ty + ny * d * dx, tz + nz * d * dx, r, ux, uy, uz, wq);
            if (ok && wq >= min_q && pi <= pmax) { prho += wq * r; pw += wq; }
            if (ok && wq >= min_q) {
                if (!found) { best_rho = r; bux = ux; buy = uy; buz = uz; best_wd = d; best_q = wq; found = true; }
                if (pi >= pmax) break;
            } else if (ok && wq > best_q && !found) {
                best_rho = r; bux = ux; buy = uy; buz = uz; best_wd = d; best_q = wq; found = true;
            }
        }
        // no fluid on the given side: the opposite side, the normal flipped when it has fluid
        float sign = 1.0f;
        if (!found && a.recover_normal) {
            for (int pi = 1; pi <= pmax; ++pi) {
                const float d = (float)pi + 0.5f;
                float r, ux, uy, uz, wq;
                const bool ok = sample_flow<SP>(a, tx - nx * d * dx, ty - ny * d * dx, tz - nz * d * dx, r, ux, uy, uz, wq);
                if (ok && wq >= min_q) {
                    prho += wq * r; pw += wq;
                    if (!found) { best_rho = r; bux = ux; buy = uy; buz = uz; best_wd = d; best_q = wq; found = true; sign = -1.0f; }
                }
            }
            if (sign < 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
        }
        // still none: the nearest fluid cell on the normal's side, shell by shell around the centroid
        int gx0, gy0, gz0;
        if (!found && floor_index(tx / dx, gx0) && floor_index(ty / dx, gy0) && floor_index(tz / dx, gz0)) {
            float best_d2 = 1.0e10f;
            for (int rad = 1; rad <= a.search_radius && !found; ++rad) {
                for (int dz = -rad; dz <= rad; ++dz)
                    for (int dy = -rad; dy <= rad; ++dy)
                        for (int ddx = -rad; ddx <= rad; ++ddx) {
                            if (abs(ddx) != rad && abs(dy) != rad && abs(dz) != rad) continue;
                            const int cx = gx0 + ddx, cy = gy0 + dy, cz = gz0 + dz;
                            const int64_t c = fluid_cell(a, cx, cy, cz);
                            if (c < 0) continue;
                            float cr, cu, cv, cw;
                            cell_rho_u<SP>((const typename SP::st*)a.f, (uint32_t)c, a.permute, cr, cu, cv, cw);
                            if (!isfinite(cr) || cr < 0.01f || cr > 100.0f) continue;
                            const float ddx_ = ((float)cx + 0.5f) * dx - tx;
                            const float ddy_ = ((float)cy + 0.5f) * dx - ty;
                            const float ddz_ = ((float)cz + 0.5f) * dx - tz;
                            const float dn = ddx_ * nx + ddy_ * ny + ddz_ * nz;
                            if (dn <= 0.0f) continue;
                            const float d2 = ddx_ * ddx_ + ddy_ * ddy_ + ddz_ * ddz_;
                            if (d2 < best_d2) {
                                best_d2 = d2; best_rho = cr; bux = cu; buy = cv; buz = cw; best_wd = dn / dx; found = true;
                            }
                        }
            }
        }
        const float pressure_rho = pw > 0.5f ? prho / pw : best_rho;
        float p = __builtin_nanf(""), sx = 0.0f, sy = 0.0f, sz = 0.0f;
        if (found) {
            stress_from_cell(a, best_rho, bux, buy, buz, nx, ny, nz, fmax_j(best_wd, 0.5f), p, sx, sy, sz);
            if (isfinite(pressure_rho) && pressure_rho >= 0.01f && pressure_rho <= 10.0f)
                p = (pressure_rho - 1.0f) * a.cs2 * a.pressure_scale;
        }
        a.p_map[i] = p;
        a.s_map[i] = sx; a.s_map[n + i] = sy; a.s_map[2 * n + i] = sz;
        a.sign_map[i] = sign;
        a.wd_map[i] = found ? best_wd * dx : 0.0f;
        if (isfinite(p) && isfinite(sx) && isfinite(sy) && isfinite(sz)) {
            const double A = (double)a.area[i], pd = (double)p;
            const double n3[3] = {(double)nx, (double)ny, (double)nz};
            const double fv[3] = {(double)sx * A, (double)sy * A, (double)sz * A};
            double fp[3], ft[3];
            for (int d = 0; d < 3; ++d) { fp[d] = -pd * n3[d] * A; ft[d] = fp[d] + fv[d]; }
            const double rx = (double)tx - a.ref[0], ry = (double)ty - a.ref[1], rz = (double)tz - a.ref[2];
            contrib[0] = fp[0]; contrib[1] = fp[1]; contrib[2] = fp[2];
            contrib[3] = fv[0]; contrib[4] = fv[1]; contrib[5] = fv[2];
            contrib[6] = ry * ft[2] - rz * ft[1];
            contrib[7] = rz * ft[0] - rx * ft[2];
            contrib[8] = rx * ft[1] - ry * ft[0];
            contrib[9] = 1.0;
        }
    }
#pragma unroll
    for (int j = 0; j < SURF_SUMS; ++j) red[j][tid] = contrib[j];
    __syncthreads();
    for (int s = SURF_WG / 2; s > 0; s >>= 1) {
        if (tid < s) {
#pragma unroll
            for (int j = 0; j < SURF_SUMS; ++j) red[j][tid] += red[j][tid + s];
        }
        __syncthreads();
    }
    if (tid < SURF_SUMS) a.part[(size_t)tid * gridDim.x + blockIdx.x] = red[tid][0];
}

// the workgroups' records summed in a fixed order: lane t takes records t, t + 256, ..., then a tree
__global__ void __launch_bounds__(SURF_WG) k_surface_final(const double* part, int ngroups, double* out) {
    __shared__ double red[SURF_SUMS][SURF_WG];
    const int tid = threadIdx.x;
    for (int j = 0; j < SURF_SUMS; ++j) {
        double s = 0.0;
        for (int g = tid; g < ngroups; g += SURF_WG) s += part[(size_t)j * ngroups + g];
        red[j][tid] = s;
    }
    __syncthreads();
    for (int s = SURF_WG / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int j = 0; j < SURF_SUMS; ++j) red[j][tid] += red[j][tid + s];
        __syncthreads();
    }
    if (tid < SURF_SUMS) out[tid] = red[tid][0];
}

template <class SP>
hipError_t surface_sp(const SurfaceArgs& a, hipStream_t s) {
    const int ngroups = (int)((a.n_tri + SURF_WG - 1) / SURF_WG);
    hipLaunchKernelGGL((k_surface<SP>), dim3(ngroups), dim3(SURF_WG), 0, s, a);
    hipLaunchKernelGGL(k_surface_final, dim3(1), dim3(SURF_WG), 0, s, (const double*)a.part, ngroups, a.out);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_surface(int storage, const SurfaceArgs& a, hipStream_t s) {
    if (a.n_tri <= 0) return hipErrorInvalidValue;
    switch (storage) {
        case 0: return surface_sp<StF32>(a, s);
        case 1: return surface_sp<StF64>(a, s);
        case 2: return surface_sp<StF16S>(a, s);
        case 3: return surface_sp<StBF16S>(a, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace oal
