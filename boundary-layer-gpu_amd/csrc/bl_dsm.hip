// Smagorinsky-type eddy viscosities on the device, the semantics of the
// reference's CPU build:
//   LES = 2  dynamic, z-averaged   sgs_dynamic_Smagorinsky (subgrid.f90:261-414)
//   LES = 3  dynamic, per cell     sgs_dynamic_Smagorinsky_noaverage (:436-583)
//   LES = 1  constant coefficient  sgs_Smagorinsky (:609-662)
// with filter_xzy (:140-187) and compute_Sij (:672-730).
//
// The dynamic models build work arrays, then fix their ghost entries in place
// with apply_Neumann_bc_x (LES = 2, boundary_conditions.f90:326-344) or
// apply_periodic_bc_x (LES = 3, :295-314) and apply_periodic_bc_z (:732-757).
// Here those fixes become index maps applied when an entry is read: an x index
// is clamped into the interior (Neumann) or redirected to its periodic source,
// and a z index on a ghost or duplicated plane is redirected to the plane it
// is copied from.
//   xmap_c periodic (x centres):               0 <- nxg-3, nxg-2 <- 1, nxg-1 <- 2
//   xmap_f periodic (x faces, U):              0 <- nx-2,  nx-1 <- 1
//   zmap_c (cell-centred in z, periodic id 1/2/4): 0 <- nzg-3, nzg-2 <- 1, nzg-1 <- 2
//   zmap_f (z faces, id 3):                      0 <- nz-2,  nz-1 <- 1
//
// Passes (0-based indices, every array in the padded field layout):
//   k_dsm_cells    cell-centred velocities (interpolation.f90:24-114 with the
//                  fixes of subgrid.f90:282-284), |S| and |S|*S_ij (:351-357)
//   k_dsm_vfilter  test-filtered velocities Uf, Vf, Wf (:329-344)
//   k_dsm_filter   per interior cell: the filtered products, velocities and
//                  |S|S_ij, the filtered strain, L_ij, M_ij, then L:M and M:M
//   k_dsm_column   z averages per (x, y) column, nu_t = max(-<L:M>/<M:M>/2*|S|, 0)
//                  and the y ghost rows (:371-412); LES = 3: the ratio per
//                  cell, periodic x and z copies, clipping, y ghost rows (:545-581)
//   k_smag         LES = 1: (Cs Delta_j f_j)^2 |S| on the interior cells
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/blgpu.h"
#include "bl_internal.hpp"

namespace blgpu {

namespace {

enum DsmSlot {
    D_CU = 0,  // term   = interpolate_x(U) with the x fix
    D_CV,      // term_1 = interpolate_y(V), rows 1..nyg-2
    D_CW,      // term_2 = interpolate_z(W) with the z fix
    D_S,       // |S| of the unfiltered field (interior cells)
    D_SS0,     // |S| S_ij, six components (interior cells)
    D_UF = D_SS0 + 6,
    D_VF,
    D_WF,
    D_LM,
    D_MM,
    D_END
};
static_assert(D_END == kDsmArrays, "work array count");

__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }
__device__ __forceinline__ int zmap_c(const Dims &d, int k) {
    return k == 0 ? d.nzg - 3 : (k == d.nzg - 2 ? 1 : (k == d.nzg - 1 ? 2 : k));
}
__device__ __forceinline__ int zmap_f(const Dims &d, int k) { return k == 0 ? d.nz - 2 : (k == d.nz - 1 ? 1 : k); }
// x fix of the cell-centred work arrays (extent nxg) and of the filtered U (extent nx)
template <bool PX>
__device__ __forceinline__ int xmap_c(const Dims &d, int i) {
    if constexpr (PX) return i == 0 ? d.nxg - 3 : (i == d.nxg - 2 ? 1 : (i == d.nxg - 1 ? 2 : i));
    return clampi(i, 1, d.nxg - 2);
}
template <bool PX>
__device__ __forceinline__ int xmap_f(const Dims &d, int i) {
    if constexpr (PX) return i == 0 ? d.nx - 2 : (i == d.nx - 1 ? 1 : i);
    return clampi(i, 1, d.nx - 2);
}

// compute_Sij at interior cell (i, j, k): diagonal terms at the centre, the
// off-diagonal terms averaged from the four surrounding edges; returns |S|.
template <class T>
__device__ double strain(const Dims &d, const Metrics &m, const T *__restrict__ U, const T *__restrict__ V,
                         const T *__restrict__ W, int i, int j, int k, double s[6]) {
    const long long sx = d.sx, sxy = d.sxy;
    const long long p = fidx(d, i, j, k);
    s[0] = ((double)U[p] - (double)U[p - 1]) * m.rdx[i];
    s[1] = ((double)V[p] - (double)V[p - sx]) * m.rdy[j];
    s[2] = ((double)W[p] - (double)W[p - sxy]) * m.rdz[k];
    auto e4 = [&](int a, int b, int c) {
        const long long q = fidx(d, a, b, c);
        return 0.5 * (((double)U[q - 1] - (double)U[q - 1 - sx]) * m.rdyg[b - 1] +
                      ((double)V[q - sx] - (double)V[q - 1 - sx]) * m.rdxg[a - 1]);
    };
    auto e5 = [&](int a, int b, int c) {
        const long long q = fidx(d, a, b, c);
        return 0.5 * (((double)U[q - 1] - (double)U[q - 1 - sxy]) * m.rdzg[c - 1] +
                      ((double)W[q - sxy] - (double)W[q - 1 - sxy]) * m.rdxg[a - 1]);
    };
    auto e6 = [&](int a, int b, int c) {
        const long long q = fidx(d, a, b, c);
        return 0.5 * (((double)V[q - sx] - (double)V[q - sx - sxy]) * m.rdzg[c - 1] +
                      ((double)W[q - sxy] - (double)W[q - sx - sxy]) * m.rdyg[b - 1]);
    };
    s[3] = 0.25 * (e4(i, j, k) + e4(i + 1, j, k) + e4(i, j + 1, k) + e4(i + 1, j + 1, k));
    s[4] = 0.25 * (e5(i, j, k) + e5(i + 1, j, k) + e5(i, j, k + 1) + e5(i + 1, j, k + 1));
    s[5] = 0.25 * (e6(i, j, k) + e6(i, j + 1, k) + e6(i, j, k + 1) + e6(i, j + 1, k + 1));
    double a = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) a = a + (c < 3 ? 2.0 : 4.0) * s[c] * s[c];
    return sqrt(a);
}

// xz weights of filter_xzy (subgrid.f90:163-173)
__device__ __forceinline__ double fil(int a, int c) {
    return (a == 0 && c == 0) ? 4.0 / 9.0 : ((a == 0 || c == 0) ? 1.0 / 9.0 : 1.0 / 36.0);
}

// y part of filter_xzy at row j of a section spanning rows jlo..jhi
// (one-sided at the section ends, subgrid.f90:179-181); q indexes row j.
template <class T>
__device__ __forceinline__ double yfilt(const T *__restrict__ F, long long q, long long sx, int j, int jlo, int jhi) {
    if (j == jlo) return 2.0 / 3.0 * (double)F[q] + 1.0 / 3.0 * (double)F[q + sx];
    if (j == jhi) return 1.0 / 3.0 * (double)F[q - sx] + 2.0 / 3.0 * (double)F[q];
    return 1.0 / 6.0 * (double)F[q - sx] + 2.0 / 3.0 * (double)F[q] + 1.0 / 6.0 * (double)F[q + sx];
}

// full filter_xzy value centred on (ic, j, kc) of a face array
template <class T>
__device__ double filt_face(const Dims &d, const T *__restrict__ F, int ic, int j, int kc, int jhi) {
    double acc = 0.0;
#pragma unroll
    for (int a = -1; a <= 1; ++a)
#pragma unroll
        for (int c = -1; c <= 1; ++c)
            acc = acc + yfilt(F, fidx(d, ic + a, j, kc + c), d.sx, j, 0, jhi) * fil(a, c);
    return acc;
}

template <bool PX, class T>
__global__ __launch_bounds__(256) void k_dsm_cells(Dims d, Metrics m, const T *__restrict__ U, const T *__restrict__ V,
                                                   const T *__restrict__ W, double *__restrict__ ws, long long nf) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= d.nxg || j >= d.nyg) return;
    const long long p = fidx(d, i, j, k);
    const int ic = xmap_c<PX>(d, i);
    ws[D_CU * nf + p] = 0.5 * ((double)U[fidx(d, ic - 1, j, k)] + (double)U[fidx(d, ic, j, k)]);
    ws[D_CV * nf + p] = (j >= 1 && j <= d.nyg - 2) ? 0.5 * ((double)V[p - d.sx] + (double)V[p]) : 0.0;
    const int kc = zmap_c(d, k);
    ws[D_CW * nf + p] = 0.5 * ((double)W[fidx(d, i, j, kc - 1)] + (double)W[fidx(d, i, j, kc)]);
    if (i >= 1 && i <= d.nxg - 2 && j >= 1 && j <= d.nyg - 2 && k >= 1 && k <= d.nzg - 2) {
        double s[6];
        const double mag = strain(d, m, U, V, W, i, j, k, s);
        ws[D_S * nf + p] = mag;
#pragma unroll
        for (int c = 0; c < 6; ++c) ws[(D_SS0 + c) * nf + p] = mag * s[c];
    }
}

template <bool PX, class T>
__global__ __launch_bounds__(256) void k_dsm_vfilter(Dims d, const T *__restrict__ U, const T *__restrict__ V,
                                                     const T *__restrict__ W, double *__restrict__ ws, long long nf) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= d.nxg || j >= d.nyg) return;
    const long long p = fidx(d, i, j, k);
    if (i < d.nx) ws[D_UF * nf + p] = filt_face(d, U, xmap_f<PX>(d, i), j, zmap_c(d, k), d.nyg - 1);
    if (j < d.ny) ws[D_VF * nf + p] = filt_face(d, V, xmap_c<PX>(d, i), j, zmap_c(d, k), d.ny - 1);
    if (k < d.nz) ws[D_WF * nf + p] = filt_face(d, W, xmap_c<PX>(d, i), j, zmap_f(d, k), d.nyg - 1);
}

template <bool PX>
__global__ __launch_bounds__(256) void k_dsm_filter(Dims d, Metrics m, double *__restrict__ ws, long long nf) {
    const int i = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = 1 + blockIdx.z;
    if (i > d.nxg - 2 || j > d.nyg - 2) return;
    const double *cu = ws + D_CU * nf, *cv = ws + D_CV * nf, *cw = ws + D_CW * nf;
    const long long sx = d.sx;
    const int jlo = 1, jhi = d.nyg - 2;
    const int r0 = (j == jlo) ? 0 : -1, r1 = (j == jhi) ? 0 : 1;  // rows used by the y filter
    double hu = 0.0, hv = 0.0, hw = 0.0, hp[6] = {0, 0, 0, 0, 0, 0}, hs[6] = {0, 0, 0, 0, 0, 0};
    for (int a = -1; a <= 1; ++a)
        for (int c = -1; c <= 1; ++c) {
            const double f = fil(a, c);
            const int in = i + a, kn = k + c;
            const long long q1 = fidx(d, in, j, kn);
            const long long q2 = fidx(d, xmap_c<PX>(d, in), j, zmap_c(d, kn));
            hu = hu + yfilt(cu, q1, sx, j, jlo, jhi) * f;
            hv = hv + yfilt(cv, q1, sx, j, jlo, jhi) * f;
            hw = hw + yfilt(cw, q1, sx, j, jlo, jhi) * f;
            // products u_i u_j at the mapped position, rows j-1..j+1
            double pr[3][6];
#pragma unroll
            for (int r = -1; r <= 1; ++r) {
                if (r < r0 || r > r1) continue;
                const long long q = q2 + r * sx;
                const double x = cu[q], y = cv[q], z = cw[q];
                pr[r + 1][0] = x * x;
                pr[r + 1][1] = y * y;
                pr[r + 1][2] = z * z;
                pr[r + 1][3] = x * y;
                pr[r + 1][4] = x * z;
                pr[r + 1][5] = y * z;
            }
#pragma unroll
            for (int e = 0; e < 6; ++e) {
                double yv;
                if (j == jlo)
                    yv = 2.0 / 3.0 * pr[1][e] + 1.0 / 3.0 * pr[2][e];
                else if (j == jhi)
                    yv = 1.0 / 3.0 * pr[0][e] + 2.0 / 3.0 * pr[1][e];
                else
                    yv = 1.0 / 6.0 * pr[0][e] + 2.0 / 3.0 * pr[1][e] + 1.0 / 6.0 * pr[2][e];
                hp[e] = hp[e] + yv * f;
                hs[e] = hs[e] + yfilt(ws + (D_SS0 + e) * nf, q2, sx, j, jlo, jhi) * f;
            }
        }
    double L[6];
    L[0] = hp[0] - hu * hu;
    L[1] = hp[1] - hv * hv;
    L[2] = hp[2] - hw * hw;
    L[3] = hp[3] - hu * hv;
    L[4] = hp[4] - hu * hw;
    L[5] = hp[5] - hv * hw;
    double sh[6];
    const double magh = strain(d, m, ws + D_UF * nf, ws + D_VF * nf, ws + D_WF * nf, i, j, k, sh);
    double M[6];
#pragma unroll
    for (int e = 0; e < 6; ++e) M[e] = 4.0 * magh * sh[e] - hs[e];
    double lm = M[0] * L[0];
    lm = lm + M[1] * L[1];
    lm = lm + M[2] * L[2];
    lm = lm + 2.0 * M[3] * L[3];
    lm = lm + 2.0 * M[4] * L[4];
    lm = lm + 2.0 * M[5] * L[5];
    const long long p = fidx(d, i, j, k);
    ws[D_LM * nf + p] = lm;
    ws[D_MM * nf + p] = M[0] * M[0] + M[1] * M[1] + M[2] * M[2] + 2.0 * (M[3] * M[3] + M[4] * M[4] + M[5] * M[5]);
}

template <class T>
__global__ __launch_bounds__(256) void k_dsm_column(Dims d, const double *__restrict__ ws, long long nf,
                                                    int dirichlet, T *__restrict__ nu_t) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= d.nxg || j >= d.nyg) return;
    const bool ghost_y = (j == 0 || j == d.nyg - 1);
    const int js = j == 0 ? 1 : (j == d.nyg - 1 ? d.nyg - 2 : j);
    const double sgn = (ghost_y && dirichlet == 1) ? -1.0 : 1.0;
    if (i == 0 || i == d.nxg - 1) {
        for (int k = 0; k < d.nzg; ++k) nu_t[fidx(d, i, j, k)] = (T)(sgn * 0.0);
        return;
    }
    const double *LM = ws + D_LM * nf, *MM = ws + D_MM * nf, *S = ws + D_S * nf;
    double slm = 0.0, smm = 0.0;
    for (int k = 1; k <= d.nzg - 3; ++k) {
        const long long q = fidx(d, i, js, k);
        slm += LM[q];
        smm += MM[q];
    }
    const double nz3 = (double)(d.nzg - 3);
    const double avg = (slm / nz3) / (smm / nz3);
    nu_t[fidx(d, i, j, 0)] = (T)(sgn * 0.0);
    nu_t[fidx(d, i, j, d.nzg - 1)] = (T)(sgn * 0.0);
    for (int k = 1; k <= d.nzg - 2; ++k)
        nu_t[fidx(d, i, j, k)] = (T)(sgn * fmax(-0.5 * avg * S[fidx(d, i, js, k)], 0.0));
}

// LES = 3: nu_t(i, j, k) = -0.5 (L:M)/(M:M) |S| per interior cell, then
// (subgrid.f90:571-581) the periodic x and z copies, clipping, and the y
// ghost rows from rows 1 and nyg-2; one thread per (i, j), all k
template <class T>
__global__ __launch_bounds__(256) void k_dsm_cellwise(Dims d, const double *__restrict__ ws, long long nf,
                                                      int dirichlet, T *__restrict__ nu_t) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= d.nxg || j >= d.nyg) return;
    const bool ghost_y = (j == 0 || j == d.nyg - 1);
    const int js = j == 0 ? 1 : (j == d.nyg - 1 ? d.nyg - 2 : j);
    const double sgn = (ghost_y && dirichlet == 1) ? -1.0 : 1.0;
    const int is = xmap_c<true>(d, i);
    const double *LM = ws + D_LM * nf, *MM = ws + D_MM * nf, *S = ws + D_S * nf;
    for (int k = 0; k < d.nzg; ++k) {
        const long long q = fidx(d, is, js, zmap_c(d, k));
        nu_t[fidx(d, i, j, k)] = (T)(sgn * fmax(-0.5 * (LM[q] / MM[q]) * S[q], 0.0));
    }
}

// LES = 1: nu_t = coef[j] |S| on the interior cells, coef[j] = (Cs Delta_j
// f_j)^2 from the host; the ghost entries are left as they are
template <class T>
__global__ __launch_bounds__(256) void k_smag(Dims d, Metrics m, const T *__restrict__ U, const T *__restrict__ V,
                                              const T *__restrict__ W, const double *__restrict__ coef,
                                              T *__restrict__ nu_t) {
    const int i = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = 1 + blockIdx.z;
    if (i > d.nxg - 2 || j > d.nyg - 2) return;
    double s[6];
    const double mag = strain(d, m, U, V, W, i, j, k, s);
    nu_t[fidx(d, i, j, k)] = (T)(coef[j] * mag);
}

void check() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        last_error() = std::string("kernel launch: ") + hipGetErrorString(e);
        throw DeviceError();
    }
}

template <bool PX, class T>
void dsm_t(Handle &h, const Fld3<T> &f, T *nu_t) {
    const Dims &d = h.d;
    const long long nf = d.sxy * d.nzg;
    const dim3 block(256);
    const dim3 gfull((d.nxg + 63) / 64, (d.nyg + 3) / 4, d.nzg);
    {
        ProfScope ps(h, KID_DSM_CELLS);
        hipLaunchKernelGGL((k_dsm_cells<PX, T>), gfull, block, 0, h.stream, d, h.m, f.u, f.v, f.w, h.dsm, nf);
        check();
    }
    {
        ProfScope ps(h, KID_DSM_VFILTER);
        hipLaunchKernelGGL((k_dsm_vfilter<PX, T>), gfull, block, 0, h.stream, d, f.u, f.v, f.w, h.dsm, nf);
        check();
    }
    {
        ProfScope ps(h, KID_DSM_FILTER);
        const dim3 g((d.nxg - 2 + 63) / 64, (d.nyg - 2 + 3) / 4, d.nzg - 2);
        hipLaunchKernelGGL(k_dsm_filter<PX>, g, block, 0, h.stream, d, h.m, h.dsm, nf);
        check();
    }
    {
        ProfScope ps(h, KID_DSM_COLUMN);
        const dim3 g((d.nxg + 63) / 64, (d.nyg + 3) / 4);
        if constexpr (PX)
            hipLaunchKernelGGL(k_dsm_cellwise<T>, g, block, 0, h.stream, d, h.dsm, nf, h.nut_bc, nu_t);
        else
            hipLaunchKernelGGL(k_dsm_column<T>, g, block, 0, h.stream, d, h.dsm, nf, h.nut_bc, nu_t);
        check();
    }
}

template <class T>
void smag_t(Handle &h, const Fld3<T> &f, T *nu_t) {
    const Dims &d = h.d;
    ProfScope ps(h, KID_DSM_CELLS);
    const dim3 g((d.nxg - 2 + 63) / 64, (d.nyg - 2 + 3) / 4, d.nzg - 2);
    hipLaunchKernelGGL(k_smag<T>, g, dim3(256), 0, h.stream, d, h.m, f.u, f.v, f.w, (const double *)h.smag_c, nu_t);
    check();
}

template <class T>
void eddy_t(Handle &h, const Fld3<T> &f, T *nu_t) {
    if (h.les == 1)
        smag_t(h, f, nu_t);
    else if (h.les == 3)
        dsm_t<true>(h, f, nu_t);
    else
        dsm_t<false>(h, f, nu_t);
}

}  // namespace

void launch_dsm(Handle &h, const Field3 &f) {
    if (h.precision == BL_PREC_MIXED)
        eddy_t(h, as<float>(f), reinterpret_cast<float *>(h.nu_t));
    else
        eddy_t(h, as<double>(f), h.nu_t);
}

}  // namespace blgpu
