// HIP kernels for the per-time-step path (gfx950, FP64 mode).
//
// Reference correspondence (0-based indices here, the reference is 1-based):
//   k_rk_march      equations.f90:24-415 fused with time_integration.f90:351-458
//   k_bc_*          boundary_conditions.f90:127-262
//   k_divergence    projection.f90:34-49
//   k_dct_*, k_zdft_*  the DCT-II/DFT stages of solve_poisson_gpu projection.f90:110-202
//   k_thomas        projection.f90:137-150
//   k_project       projection.f90:355-386 with apply_periodic_z_pressure :293-303,
//                   and the P record of the first sub-step :75-85
//   reductions      projection.f90:391-413, statistics.f90:137-160, monitor.f90:29-52,
//                   time_integration.f90:546-558
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "bl_internal.hpp"

namespace blgpu {

namespace {

constexpr int kWave = 64;

template <class T>
__device__ __forceinline__ T sq(T a) {
    return a * a;
}

// ------------------------------------------------------------------------
// Fused RHS + Runge-Kutta stage update, marching in z.
// MODE 0: new = old + c0*F              (first stage, and RK2 second stage;
//         FP64 RK3 stages 2 and 3 with old = G, below)
// MODE 1: new = old + dt*(c1*F1 + c0*F) (RK3 stages 2 and 3, FP32 fields)
// MODE 2: write F only (operator test)
// FOUT (first stage of RK3): 1 writes F1 = F; 2 writes G = old + g2*F
// instead (g2 = dt a21), so that stages 2 and 3 read one field per component
// instead of two (old and F1): a21 = a31 = 1/4 (initialization.f90:526,528)
// makes Uo + dt a21 F1 and Uo + dt a31 F1 the same field, and nothing
// writes G between stages 2 and 3.  A block owns a
// 64 (x) x 4 RPT (y) column of cells and a chunk of z planes; U and V of
// planes k, k+1 and W of planes k-1, k, k+1 live in an LDS ring with a
// one-cell halo in x and y (the stencil reads U and V of plane k-1 only at
// the cell itself, which each thread carries in registers from the previous
// plane), so every velocity value is read from HBM once per chunk (plus
// halo) instead of once per neighbour.  Loads are x-fastest and 512-byte
// aligned rows.  The x and y metrics of the tile are staged in LDS once:
// read from global memory inside each equation's branch they cost a load
// latency per branch and plane (2.69 -> 2.00 ms of RK stages per FP64 step,
// 1.71 -> 1.47 in FP32, at 814x257x193).  Thread (tx, ty) computes cells
// (i0+tx, j0+RPT ty+r), r < RPT.  FP64: 64 x 8 tiles (RPT = 2, 39.5 KB of
// LDS, 140 VGPRs: 3 blocks per CU; RPT = 1 took 16 % longer, and 4 blocks
// per CU forced by a 128-VGPR bound spilled and took 12 % longer); FP32:
// 64 x 12 tiles (RPT = 3, 27 KB; 64 x 16 tiles held 2 blocks per CU and
// took 3 % longer).  128-wide FP32 tiles (fewer x-halo lines) measured
// slower.
// ------------------------------------------------------------------------
template <class T>
struct MarchTile {
    static constexpr int TX = 64;                          // tile width (x)
    static constexpr int RPT = sizeof(T) == 4 ? 3 : 2;    // rows per thread
    static constexpr int TY = (256 / TX) * RPT, HX = TX + 2, HY = TY + 2, Plane = HX * HY;
};

// LES: the viscous terms use nu + nu_t (equations.f90:97-132, 231-270,
// 367-404); nu_t is staged in a fourth LDS ring.
template <int MODE, int FOUT, bool LES, class T>
__global__ __launch_bounds__(256) void k_rk_march(
    Dims d, MetricsT<T> m, T nu, T dPdx, T dPdz, int kchunk,
    const T *__restrict__ U, const T *__restrict__ V, const T *__restrict__ W,
    const T *__restrict__ NT,
    const T *__restrict__ Uo, const T *__restrict__ Vo, const T *__restrict__ Wo,
    const T *__restrict__ F1u, const T *__restrict__ F1v, const T *__restrict__ F1w,
    T *__restrict__ Un, T *__restrict__ Vn, T *__restrict__ Wn, T *__restrict__ Fou,
    T *__restrict__ Fov, T *__restrict__ Fow, T c0, T c1, T dt, T g2) {
    constexpr int RPT = MarchTile<T>::RPT, kTY = MarchTile<T>::TY, kPlane = MarchTile<T>::Plane;
    constexpr int kTX = MarchTile<T>::TX, kHX = MarchTile<T>::HX;
    // U and V: planes k and k+1 (their plane k-1 value at the thread's own
    // cells is carried in registers); W: planes k-1, k, k+1
    __shared__ T sU[2][kPlane], sV[2][kPlane], sW[3][kPlane];
    __shared__ T sN[LES ? 3 : 1][LES ? kPlane : 1];
    // the tile's x metrics (rdx(i), rdx(i+1), rdxg(i), rdxg(i-1)) and y metrics
    // (wy0(j), wy1(j), wy0(j-1), wy1(j-1), rdy(j), rdy(j+1), rdyg(j),
    // rdyg(j-1)): constant along the march, read from LDS with the velocities
    // instead of from global memory inside each equation's branch
    __shared__ T sMx[4][kTX], sMy[8][kTY];
    const int tid = threadIdx.x;
    const int tx = tid % kTX, ty = tid / kTX;
    const int nbx = (d.nxg - 1 + kTX - 1) / kTX, nby = (d.nyg - 2 + kTY - 1) / kTY;
    const int ntile = nbx * nby * ((d.nzg - 2 + kchunk - 1) / kchunk);
    const int t = xcd_tile(blockIdx.x, ntile);
    if (t >= ntile) return;
    // the tile's cells are i0 .. i0 + kTX - 1 with i0 a multiple of kTX, so
    // that the per-cell rows start on a 128-byte line (cell 0 is idle)
    const int i0 = (t % nbx) * kTX, j0 = ((t / nbx) % nby) * kTY + 1;
    const int kb = 1 + (t / (nbx * nby)) * kchunk;
    const int ke = min(kb + kchunk, d.nzg - 1);  // exclusive
    const long long sx = d.sx, sxy = d.sxy;
    // halo tile origin (i0-1, j0-1); loads clamped into the allocation (the
    // column i0-1 = -1 of the first tile feeds only the idle cell 0)
    // each thread moves up to kPer halo elements of a plane; they are fetched
    // into registers one iteration ahead and committed to LDS after the sync
    constexpr int kPer = (kPlane + 255) / 256;
    long long off[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int e = tid + 256 * q;
        const int ee = e < kPlane ? e : kPlane - 1;
        const int hy = ee / kHX, hx = ee - hy * kHX;
        const int ii = max(0, min(i0 - 1 + hx, d.sx - 1)), jj = min(j0 - 1 + hy, d.nyg - 1);
        off[q] = (long long)jj * sx + ii;
    }
    T ru[kPer], rv[kPer], rw[kPer], rn[LES ? kPer : 1];
    auto fetch = [&](int k) {
        const long long base = (long long)k * sxy;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            ru[q] = U[base + off[q]];
            rv[q] = V[base + off[q]];
            rw[q] = W[base + off[q]];
            if constexpr (LES) rn[q] = NT[base + off[q]];
        }
    };
    auto commit = [&](int plane) {
        const int s2 = plane & 1, s3 = plane % 3;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + 256 * q;
            if (e < kPlane) {
                sU[s2][e] = ru[q];
                sV[s2][e] = rv[q];
                sW[s3][e] = rw[q];
                if constexpr (LES) sN[s3][e] = rn[q];
            }
        }
    };
    if (tid < kTX) {
        const int ii = max(1, i0 + tid);  // entries of cells outside the grid are clamped and never used
        sMx[0][tid] = m.rdx[min(ii, d.nx - 1)];
        sMx[1][tid] = m.rdx[min(ii + 1, d.nx - 1)];
        sMx[2][tid] = m.rdxg[min(ii, d.nxg - 1)];
        sMx[3][tid] = m.rdxg[min(ii - 1, d.nxg - 1)];
    } else if (tid < kTX + kTY) {
        const int jt = tid - kTX, jj = j0 + jt;
        sMy[0][jt] = m.wy0[min(jj, d.ny - 1)];
        sMy[1][jt] = m.wy1[min(jj, d.ny - 1)];
        sMy[2][jt] = m.wy0[min(jj - 1, d.ny - 1)];
        sMy[3][jt] = m.wy1[min(jj - 1, d.ny - 1)];
        sMy[4][jt] = m.rdy[min(jj, d.ny - 1)];
        sMy[5][jt] = m.rdy[min(jj + 1, d.ny - 1)];
        sMy[6][jt] = m.rdyg[min(jj, d.nyg - 1)];
        sMy[7][jt] = m.rdyg[min(jj - 1, d.nyg - 1)];
    }
    fetch(kb - 1);
    commit(kb - 1);
    fetch(kb);
    commit(kb);
    fetch(kb + 1);
    const int i = i0 + tx;
    // U and V of plane kb-1 at the thread's cells, before plane kb+1 takes
    // their slot
    __syncthreads();
    T ubr[RPT], vbr[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int e = (RPT * ty + r + 1) * kHX + (tx + 1);
        ubr[r] = sU[(kb - 1) & 1][e];
        vbr[r] = sV[(kb - 1) & 1][e];
    }
    __syncthreads();  // before plane kb+1 overwrites that slot
    // a first stage starts from the field it differentiates: the old values
    // are then the LDS centre values and are not read again
    const bool old_is_cur = MODE == 0 && Uo == U && Vo == V && Wo == W;
    // MODE 0 with a separate old field (G, or the RK2 second stage) and
    // MODE 1: the old values per cell (and MODE 1's F1) are loaded one plane
    // ahead of their use; with the metrics in LDS MODE 1's six values fit
    // the registers at 3 blocks per CU too (FP32 RK stages 1.47 -> 1.39 ms
    // per step at 814x257x193)
    constexpr bool kAheadOld = (MODE == 0 && FOUT == 0) || MODE == 1;
    const bool ahead = kAheadOld && (MODE == 1 || !old_is_cur);
    T nuo[kAheadOld ? RPT : 1][3] = {}, nfo[MODE == 1 ? RPT : 1][3] = {};
    auto fetch_old = [&](int k) {
#pragma unroll
        for (int r = 0; r < (kAheadOld ? RPT : 1); ++r) {
            const long long p = fidx(d, max(1, min(i0 + tx, d.nxg - 2)), min(j0 + RPT * ty + r, d.nyg - 2), k);
            nuo[r][0] = Uo[p];
            nuo[r][1] = Vo[p];
            nuo[r][2] = Wo[p];
            if constexpr (MODE == 1) {
                nfo[r][0] = F1u[p];
                nfo[r][1] = F1v[p];
                nfo[r][2] = F1w[p];
            }
        }
    };
    if (ahead) fetch_old(kb);
    for (int k = kb; k < ke; ++k) {
        // this plane's step-start and first-stage values, issued before the
        // barrier so that their latency overlaps the stencil work
        T uo[RPT][3] = {}, fo[RPT][3] = {};
        if (ahead) {
#pragma unroll
            for (int r = 0; r < (kAheadOld ? RPT : 1); ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    uo[r][c] = nuo[r][c];
                    if constexpr (MODE == 1) fo[r][c] = nfo[r][c];
                }
        } else if (MODE == 1 || (MODE == 0 && !old_is_cur)) {
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int j = j0 + RPT * ty + r;
                if (i < 1 || i > d.nxg - 2 || j > d.nyg - 2) continue;
                const long long p = fidx(d, i, j, k);
                uo[r][0] = Uo[p];
                uo[r][1] = Vo[p];
                uo[r][2] = Wo[p];
                if constexpr (MODE == 1) {
                    fo[r][0] = F1u[p];
                    fo[r][1] = F1v[p];
                    fo[r][2] = F1w[p];
                }
            }
        }
        commit(k + 1);
        __syncthreads();
        if (k + 2 <= ke) fetch(k + 2);
        if (ahead && k + 1 < ke) fetch_old(k + 1);
        // W (and nu_t): three slots; U, V: two
        const int sm = (k - 1) % 3, sc = k % 3, sp = (k + 1) % 3;
        const int uc = k & 1, up1 = (k + 1) & 1;
        const T rdzk = m.rdz[k], rdzgk = m.rdzg[k], rdzgm = m.rdzg[k - 1];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int jt = RPT * ty + r;     // row in the tile
            const int j = j0 + jt;
            const int e = (jt + 1) * kHX + (tx + 1);  // LDS index of (i, j)
            // U, V of plane k-1 at this cell, and plane k's for the next plane
            const T ubc = ubr[r], vbc = vbr[r];
            ubr[r] = sU[uc][e];
            vbr[r] = sV[uc][e];
            if (i < 1 || i > d.nxg - 2 || j > d.nyg - 2) continue;
            const T x_rdx = sMx[0][tx], x_rdx1 = sMx[1][tx], x_rdxg = sMx[2][tx], x_rdxgm = sMx[3][tx];
            const T y_w0 = sMy[0][jt], y_w1 = sMy[1][jt], y_w0m = sMy[2][jt], y_w1m = sMy[3][jt];
            const T y_rdy = sMy[4][jt], y_rdy1 = sMy[5][jt], y_rdyg = sMy[6][jt], y_rdygm = sMy[7][jt];
            const long long p = fidx(d, i, j, k);
            // ---- u (equations.f90:47-94, 149-154)
            if (i <= d.nx - 2) {
                T rr = 0.0;
                if (i < d.nx - 2) {
                    const T up = sU[uc][e], ue = sU[uc][e + 1], uw = sU[uc][e - 1];
                    const T un = sU[uc][e + kHX], us = sU[uc][e - kHX];
                    const T ut = sU[up1][e], ub = ubc;
                    rr = -(sq(0.5 * (up + ue)) - sq(0.5 * (uw + up))) * x_rdxg;
                    rr = rr - ((y_w0 * up + y_w1 * un) * 0.5 * (sV[uc][e] + sV[uc][e + 1]) -
                               (y_w0m * us + y_w1m * up) * 0.5 *
                                   (sV[uc][e - kHX] + sV[uc][e + 1 - kHX])) *
                                  y_rdy;
                    rr = rr - (0.5 * (up + ut) * 0.5 * (sW[sc][e] + sW[sc][e + 1]) -
                               0.5 * (ub + up) * 0.5 * (sW[sm][e] + sW[sm][e + 1])) *
                                  rdzk;
                    if constexpr (!LES) {
                        rr = rr + dPdx +
                             nu * ((x_rdx1 * (ue - up) - x_rdx * (up - uw)) * x_rdxg +
                                   ((un - up) * y_rdyg - (up - us) * y_rdygm) * y_rdy +
                                   (rdzgk * (ut - up) - rdzgm * (up - ub)) * rdzk);
                    } else {
                        const T *N = sN[sc], *Nm = sN[sm], *Np = sN[sp];
                        const T a0 = y_w0m, a1 = y_w1m, b0 = y_w0, b1 = y_w1;
                        const T nx1 = nu + N[e];
                        const T nx2 = nu + N[e + 1];
                        const T ny1 = nu + 0.5 * (a0 * N[e - kHX] + a1 * N[e] + a0 * N[e + 1 - kHX] + a1 * N[e + 1]);
                        const T ny2 = nu + 0.5 * (b0 * N[e] + b1 * N[e + kHX] + b0 * N[e + 1] + b1 * N[e + 1 + kHX]);
                        const T nz1 = nu + 0.25 * (N[e] + Nm[e] + N[e + 1] + Nm[e + 1]);
                        const T nz2 = nu + 0.25 * (N[e] + Np[e] + N[e + 1] + Np[e + 1]);
                        rr = rr + dPdx +
                             (nx2 * x_rdx1 * (ue - up) - nx1 * x_rdx * (up - uw)) * x_rdxg +
                             (ny2 * ((un - up) * y_rdyg) - ny1 * ((up - us) * y_rdygm)) * y_rdy +
                             (nz2 * rdzgk * (ut - up) - nz1 * rdzgm * (up - ub)) * rdzk;
                    }
                }
                if (MODE == 2) {
                    Fou[p] = rr;
                } else {
                    const T base = old_is_cur ? sU[uc][e] : uo[r][0];
                    if (FOUT == 1) Fou[p] = rr;
                    if (FOUT == 2) Fou[p] = base + g2 * rr;
                    Un[p] = (MODE == 0) ? base + c0 * rr
                                        : uo[r][0] + dt * (c1 * fo[r][0] + c0 * rr);
                }
            }
            // ---- v (equations.f90:183-230, 287-292)
            if (j <= d.ny - 2) {
                T rr = 0.0;
                if (i < d.nxg - 2) {
                    const T vp = sV[uc][e], ve = sV[uc][e + 1], vw = sV[uc][e - 1];
                    const T vn = sV[uc][e + kHX], vs = sV[uc][e - kHX];
                    const T vt = sV[up1][e], vb = vbc;
                    rr = -(sq(0.5 * (vp + vn)) - sq(0.5 * (vs + vp))) * y_rdyg;
                    rr = rr - ((y_w0 * sU[uc][e] + y_w1 * sU[uc][e + kHX]) * 0.5 * (vp + ve) -
                               (y_w0 * sU[uc][e - 1] + y_w1 * sU[uc][e - 1 + kHX]) * 0.5 * (vw + vp)) *
                                  x_rdx;
                    rr = rr - (0.5 * (vp + vt) * (y_w0 * sW[sc][e] + y_w1 * sW[sc][e + kHX]) -
                               0.5 * (vb + vp) * (y_w0 * sW[sm][e] + y_w1 * sW[sm][e + kHX])) *
                                  rdzk;
                    if constexpr (!LES) {
                        rr = rr + nu * ((x_rdxg * (ve - vp) - x_rdxgm * (vp - vw)) * x_rdx +
                                        (y_rdy1 * (vn - vp) - y_rdy * (vp - vs)) * y_rdyg +
                                        (rdzgk * (vt - vp) - rdzgm * (vp - vb)) * rdzk);
                    } else {
                        const T *N = sN[sc], *Nm = sN[sm], *Np = sN[sp];
                        const T b0 = y_w0, b1 = y_w1;
                        const T nx1 = nu + 0.5 * (b0 * N[e - 1] + b1 * N[e - 1 + kHX] + b0 * N[e] + b1 * N[e + kHX]);
                        const T nx2 = nu + 0.5 * (b0 * N[e] + b1 * N[e + kHX] + b0 * N[e + 1] + b1 * N[e + 1 + kHX]);
                        const T ny1 = nu + N[e];
                        const T ny2 = nu + N[e + kHX];
                        const T nz1 = nu + 0.5 * (b0 * Nm[e] + b1 * Nm[e + kHX] + b0 * N[e] + b1 * N[e + kHX]);
                        const T nz2 = nu + 0.5 * (b0 * N[e] + b1 * N[e + kHX] + b0 * Np[e] + b1 * Np[e + kHX]);
                        rr = rr + x_rdx * (nx2 * x_rdxg * (ve - vp) - nx1 * x_rdxgm * (vp - vw)) +
                             y_rdyg * (ny2 * y_rdy1 * (vn - vp) - ny1 * y_rdy * (vp - vs)) +
                             rdzk * (nz2 * rdzgk * (vt - vp) - nz1 * rdzgm * (vp - vb));
                    }
                }
                if (MODE == 2) {
                    Fov[p] = rr;
                } else {
                    const T base = old_is_cur ? sV[uc][e] : uo[r][1];
                    if (FOUT == 1) Fov[p] = rr;
                    if (FOUT == 2) Fov[p] = base + g2 * rr;
                    Vn[p] = (MODE == 0) ? base + c0 * rr
                                        : uo[r][1] + dt * (c1 * fo[r][1] + c0 * rr);
                }
            }
            // ---- w (equations.f90:323-366, 408-413)
            if (k <= d.nz - 2) {
                T rr = 0.0;
                if (i < d.nxg - 2) {
                    const T wp = sW[sc][e], we = sW[sc][e + 1], ww = sW[sc][e - 1];
                    const T wn = sW[sc][e + kHX], ws = sW[sc][e - kHX];
                    const T wt = sW[sp][e], wb = sW[sm][e];
                    rr = -(sq(0.5 * (wp + wt)) - sq(0.5 * (wb + wp))) * rdzgk;
                    rr = rr - (0.5 * (sU[uc][e] + sU[up1][e]) * 0.5 * (wp + we) -
                               0.5 * (sU[uc][e - 1] + sU[up1][e - 1]) * 0.5 * (ww + wp)) *
                                  x_rdx;
                    rr = rr - (0.5 * (sV[uc][e] + sV[up1][e]) * (y_w0 * wp + y_w1 * wn) -
                               0.5 * (sV[uc][e - kHX] + sV[up1][e - kHX]) * (y_w0m * ws + y_w1m * wp)) *
                                  y_rdy;
                    if constexpr (!LES) {
                        rr = rr + dPdz +
                             nu * ((x_rdxg * (we - wp) - x_rdxgm * (wp - ww)) * x_rdx +
                                   ((wn - wp) * y_rdyg - (wp - ws) * y_rdygm) * y_rdy +
                                   (m.rdz[k + 1] * (wt - wp) - rdzk * (wp - wb)) * rdzgk);
                    } else {
                        const T *N = sN[sc], *Np = sN[sp];
                        const T a0 = y_w0m, a1 = y_w1m, b0 = y_w0, b1 = y_w1;
                        const T nx1 = nu + 0.25 * (N[e] + Np[e] + N[e - 1] + Np[e - 1]);
                        const T nx2 = nu + 0.25 * (N[e] + Np[e] + N[e + 1] + Np[e + 1]);
                        const T ny1 = nu + 0.5 * (a0 * N[e - kHX] + a1 * N[e] + a0 * Np[e - kHX] + a1 * Np[e]);
                        const T ny2 = nu + 0.5 * (b0 * N[e] + b1 * N[e + kHX] + b0 * Np[e] + b1 * Np[e + kHX]);
                        const T nz1 = nu + N[e];
                        const T nz2 = nu + Np[e];
                        rr = rr + dPdz +
                             (nx2 * x_rdxg * (we - wp) - nx1 * x_rdxgm * (wp - ww)) * x_rdx +
                             (ny2 * ((wn - wp) * y_rdyg) - ny1 * ((wp - ws) * y_rdygm)) * y_rdy +
                             (nz2 * m.rdz[k + 1] * (wt - wp) - nz1 * rdzk * (wp - wb)) * rdzgk;
                    }
                }
                if (MODE == 2) {
                    Fow[p] = rr;
                } else {
                    const T base = old_is_cur ? sW[sc][e] : uo[r][2];
                    if (FOUT == 1) Fow[p] = rr;
                    if (FOUT == 2) Fow[p] = base + g2 * rr;
                    Wn[p] = (MODE == 0) ? base + c0 * rr
                                        : uo[r][2] + dt * (c1 * fo[r][2] + c0 * rr);
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// Boundary conditions, in the reference's order (boundary_conditions.f90:
// 135-260).  Each kernel covers one phase over the planes it writes.
// ------------------------------------------------------------------------

// inflow (:137-190) and outflow (:193-202); one thread per (j,k)
// deterministic block sum of 256 values; returns the sum in thread 0
__device__ __forceinline__ double block_sum_256(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    return sh[0];
}

// temporal inflow modes Re(q e^{i(beta z - omega t)}) of row j, plane k at
// time t (boundary_conditions.f90:153-170)
__device__ __forceinline__ void inflow_modes(const Dims &d, const Metrics &m, const BCTables &b, int j, int k,
                                             double t, double &ut, double &vt, double &wt) {
    ut = vt = wt = 0.0;
    const double zc = m.zg[k];
    const double zf = (k < d.nz) ? m.z[k] : 0.0;
    for (int n = 0; n < b.n_modes; ++n)
        for (int mm = 0; mm < b.m_modes; ++mm) {
            const double pc = b.zmode[n] * zc - b.tmode[mm] * t;
            const double pf = b.zmode[n] * zf - b.tmode[mm] * t;
            const long long qi = (long long)n + (long long)b.n_modes * mm;
            const long long qu = j + (long long)d.nyg * qi;
            const double cc = cos(pc), sc = sin(pc);
            ut = ut + (b.qu_re[qu] * cc - b.qu_im[qu] * sc);
            if (j < d.ny) {
                const long long qv = j + (long long)d.ny * qi;
                vt = vt + (b.qv_re[qv] * cc - b.qv_im[qv] * sc);
            }
            const double cf = cos(pf), sf = sin(pf);
            wt = wt + (b.qw_re[qu] * cf - b.qw_im[qu] * sf);
        }
}

// inflow_flag 4: uniform [0, 1) value of (step, component, j, k) from a
// counter hash (splitmix64 finaliser) of the stream's seed
__device__ __forceinline__ unsigned long long bc_mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double inflow_rand(unsigned long long seed, long long step, int comp, int j, int k) {
    const unsigned long long key = (((unsigned long long)step * 4ULL + (unsigned long long)comp) << 40) ^
                                   ((unsigned long long)k << 20) ^ (unsigned long long)j;
    return (double)(bc_mix64(seed ^ bc_mix64(key)) >> 11) * 0x1.0p-53;
}

// the inflow values of row j, plane k (FP64, before storage rounding):
// Blasius + temporal modes at time t (boundary_conditions.f90:183-190), or
// Blasius + the perturbations of step `step` (inflow_flag 4, :423-445)
__device__ __forceinline__ void inflow_values(const Dims &d, const Metrics &m, const BCTables &b, int j, int k,
                                              double t, long long step, double &u, double &v, double &w) {
    if (b.inflow_flag == 4) {
        const double a = b.amp * (1.0 - b.U_inlet[j]);
        u = b.U_inlet[j] + a * (inflow_rand(b.seed, step, 0, j, k) - 0.5);
        w = b.W_inlet[j] + a * (inflow_rand(b.seed, step, 2, j, k) - 0.5);
        v = (j < d.ny) ? b.V_inlet[j] + a * (inflow_rand(b.seed, step, 1, j, k) - 0.5) : 0.0;
        return;
    }
    double ut = 0.0, vt = 0.0, wt = 0.0;
    if (b.n_modes > 0) inflow_modes(d, m, b, j, k, t, ut, vt, wt);
    u = b.U_inlet[j] + ut;
    w = b.W_inlet[j] + wt;
    v = (j < d.ny) ? b.V_inlet[j] + vt : 0.0;
}

template <class T>
__global__ __launch_bounds__(256) void k_bc_inout(Dims d, Metrics m, BCTables b, T *U, T *V, T *W,
                                                  const T *Uo, const T *Vo, const T *Wo, double coef,
                                                  double t, double *qpart, double *inc1, double t_old) {
    __shared__ double sh[256];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    // mass flux through the x boundaries (boundary_conditions.f90:246-249):
    // (U(1,j,k) - U(nx-1,j,k))*(y(j)-y(j-1)) for j in [1,nyg-2], k in [1,nzg-3]
    double q = 0.0;
    if (j < d.nyg) {
    const int nx = d.nx, nxg = d.nxg;
    double ui, vi, wi;
    const bool lund = b.inflow_flag == 3 || b.inflow_flag == 5;
    if (lund) {
        // the step's rescaled plane (:569-597); flag 5 adds it to the inlet
        // means, with W's planes shifted by 16 (:512-559)
        const bool f5 = b.inflow_flag == 5;
        ui = (f5 ? b.U_inlet[j] : 0.0) + b.lund_u[j + (long long)d.nyg * k];
        wi = 0.0;
        if (k < d.nz) {
            const int ks = f5 ? (k + 17) % d.nz : k;
            wi = (f5 ? b.W_inlet[j] : 0.0) + b.lund_w[j + (long long)d.nyg * ks];
        }
        vi = (j < d.ny) ? (f5 ? b.V_inlet[j] : 0.0) + b.lund_v[j + (long long)d.ny * k] : 0.0;
    } else {
        inflow_values(d, m, b, j, k, t, b.istep, ui, vi, wi);
    }
    const long long p0 = fidx(d, 0, j, k);
    // U (all j < nyg, k < nzg)
    const double uin = (double)(T)ui;  // the value stored
    const long long pl = j + (long long)d.nyg * k;
    if (inc1) {
        // FP64 increment of the inflow face over the step-start value: that
        // value is taken unrounded when it is the stored boundary value of
        // t_old (the previous step), as it is after any step
        double uo = 0.0, vo, wo;
        if (!lund) inflow_values(d, m, b, j, k, t_old, b.istep - 1, uo, vo, wo);
        const double so = (double)Uo[p0];
        // (the previous step's rescaled plane is gone: the stored value)
        inc1[pl] = ui - (!lund && (double)(T)uo == so ? uo : so);
    }
    U[p0] = (T)uin;
    {
        const long long pa = fidx(d, nx - 2, j, k);
        const double o1 = Uo[pa], o0 = Uo[pa - 1];
        const double du = coef * (o1 - o0) / (m.x[nx - 2] - m.x[nx - 3]);
        const double v = (double)(T)(o1 - du);
        U[pa] = (T)v;
        U[pa + 1] = (T)v;
        // FP64 increment of the outflow face before the mass-flux shift
        if (inc1) inc1[(long long)d.nyg * d.nzg + pl] = -du;
        if (j >= 1 && j <= d.nyg - 2 && k >= 1 && k <= d.nzg - 3) q = (uin - v) * (m.y[j] - m.y[j - 1]);
    }
    if (k < d.nz) {
        W[p0] = (T)wi;
        const long long pa = fidx(d, nxg - 2, j, k);
        const double o1 = Wo[pa], o0 = Wo[pa - 1];
        const double v = o1 - coef * (o1 - o0) / (m.xg[nxg - 2] - m.xg[nxg - 3]);
        W[pa] = (T)v;
        W[pa + 1] = (T)v;
    }
    if (j < d.ny) {
        V[p0] = (T)vi;
        const long long pa = fidx(d, nxg - 2, j, k);
        const double o1 = Vo[pa], o0 = Vo[pa - 1];
        const double v = o1 - coef * (o1 - o0) / (m.xg[nxg - 2] - m.xg[nxg - 3]);
        V[pa] = (T)v;
        V[pa + 1] = (T)v;
    }
    }
    q = block_sum_256(q, sh);
    if (threadIdx.x == 0) qpart[blockIdx.x + gridDim.x * blockIdx.y] = q;
}

// periodic z (:205-217); one thread per (i,j)
template <class T>
__global__ __launch_bounds__(256) void k_bc_periodic(Dims d, T *U, T *V, T *W) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= d.nxg) return;
    const long long base = fidx(d, i, j, 0), sxy = d.sxy;
    const int nzg = d.nzg, nz = d.nz;
    if (i < d.nx) {
        U[base] = U[base + sxy * (nzg - 3)];
        U[base + sxy * (nzg - 2)] = U[base + sxy * 1];
        U[base + sxy * (nzg - 1)] = U[base + sxy * 2];
    }
    if (j < d.ny) {
        V[base] = V[base + sxy * (nzg - 3)];
        V[base + sxy * (nzg - 2)] = V[base + sxy * 1];
        V[base + sxy * (nzg - 1)] = V[base + sxy * 2];
    }
    W[base] = W[base + sxy * (nz - 2)];
    W[base + sxy * (nz - 1)] = W[base + sxy * 1];
}

// top Dirichlet (:220-228) then wall (:231-238); one thread per (i,k)
template <class T>
__global__ __launch_bounds__(256) void k_bc_topwall(Dims d, Metrics m, BCTables b, T *U, T *V, T *W,
                                                    const T *__restrict__ Vo, double *qpart) {
    __shared__ double sh[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    // mass flux through the y boundaries (boundary_conditions.f90:251-253) from
    // the values written here: (V(i,1,k) - V(i,ny,k))*(x(i)-x(i-1))
    // V(i, ny) as stored at the top: V_top (top_flag 0), blowing/suction
    // (1, 2: :648-673; columns below i_rescale keep it, because the
    // reference's V_top copy of zero_wz_top, :279, writes V(i, nyg), one row
    // past the top face, which the wall condition then overwrites), or 0
    // (Falkner-Skan, :1001)
    const int ii = min(i, d.nxg - 1);
    const double vtop = b.top_flag == 0 ? (double)(T)b.V_top[ii]
                                        : (b.top_flag == 3 ? 0.0 : (double)(T)b.vtop_bs[ii]);
    // V(i, 1) as the wall condition below stores it: 0, or with a Robin wall
    // model the pseudo-Robin alpha_y dVo/dy less its wall mean beta_y
    // (apply_pseudo_Robin_bc_y_bottom :1152-1180, id 2)
    const bool robin = b.iwall != 0 && b.iwall != 11 && b.iwall != 15;
    double vwall = 0.0;
    if (robin) {
        const long long pw = fidx(d, ii, 0, k);
        vwall = (double)(T)(b.wm_ay[ii] * ((double)Vo[pw + d.sx] - (double)Vo[pw]) / (m.y[1] - m.y[0]) - b.wm_beta[0]);
    }
    double q = 0.0;
    if (i >= 1 && i <= d.nxg - 3 && k >= 1 && k <= d.nzg - 3) q = (vwall - vtop) * (m.x[i] - m.x[i - 1]);
    q = block_sum_256(q, sh);
    if (threadIdx.x == 0) qpart[blockIdx.x + gridDim.x * blockIdx.y] = q;
    if (i >= d.nxg) return;
    const long long sx = d.sx;
    const long long p0 = fidx(d, i, 0, k);
    const long long ut_ = p0 + sx * (d.nyg - 1), vt_ = p0 + sx * (d.ny - 1), wt_ = p0 + sx * (d.nyg - 1);
    if (b.top_flag == 0) {
        // Dirichlet (:220-228, :947-964)
        if (i < d.nx) U[ut_] = (T)b.U_top[i];
        V[vt_] = (T)b.V_top[i];
        if (k < d.nz) W[wt_] = (T)b.W_top[i];
    } else if (b.top_flag == 3) {
        // Falkner-Skan (:979-1007)
        if (i < d.nx) U[ut_] = (T)b.utop_fs[i];
        V[vt_] = (T)0;
        if (k < d.nz) W[wt_] = W[wt_ - sx];
    } else {
        // blowing/suction, then zero_wz_top (:264-283)
        V[vt_] = (T)b.vtop_bs[i];
        if (i < d.nx) {
            const double vp = (double)(T)b.vtop_bs[i + 1], vm = (double)(T)b.vtop_bs[i];
            U[ut_] = (T)(b.dy_dx * (vp - vm) + (double)U[ut_ - sx]);
        }
        if (k < d.nz) W[wt_] = W[wt_ - sx];
        if (i < b.i_rescale) {
            if (i < d.nx) U[ut_] = (T)b.U_top[i];
            if (k < d.nz) W[wt_] = (T)b.W_top[i];
        }
    }
    // walls: no-slip (:231-238), or the wall model's conditions (:73-108):
    // Neumann U (iwall_model 11, 15: apply_Nozero_Neumann_bc_y :359-372)
    // with V = W = 0, else Robin U, W (apply_Robin_bc_y_bottom :1107-1132,
    // id 2) and the pseudo-Robin V
    if (b.iwall == 0) {
        if (i < d.nx) U[p0] = -U[p0 + sx];
        V[p0] = (T)0;
        if (k < d.nz) W[p0] = -W[p0 + sx];
    } else {
        const double dyg = m.yg[1] - m.yg[0];
        if (!robin) {
            if (i < d.nx) U[p0] = (T)((double)U[p0 + sx] - b.wm_ax[i] * dyg);
            V[p0] = (T)0;
            if (k < d.nz) W[p0] = -W[p0 + sx];
        } else {
            if (i < d.nx) {
                const double r = 2.0 * b.wm_ax[i] / dyg;
                U[p0] = (T)((r - 1.0) * (double)U[p0 + sx] / (r + 1.0));
            }
            V[p0] = (T)vwall;
            if (k < d.nz) {
                const double r = 2.0 * b.wm_az[i] / dyg;
                W[p0] = (T)((r - 1.0) * (double)W[p0 + sx] / (r + 1.0));
            }
        }
    }
}

// global mass conservation (:241-260): every block re-sums the flux partials
// in the same order (so all agree bit for bit) and corrects its part of the
// outflow planes.  One thread per (j,k).
template <class T>
__global__ __launch_bounds__(256) void k_bc_massflux(Dims d, double length_y, const double *qpart,
                                                     int nq, T *U, double *du_out) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int t = threadIdx.x; t < nq; t += 256) acc += qpart[t];
    const double Q = block_sum_256(acc, sh);
    const double du = Q / length_y / (double)(d.nzg - 3);
    if (du_out && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) du_out[0] = du;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= d.nyg) return;
    const long long pa = fidx(d, d.nx - 2, j, k);
    const double v = (double)U[pa] + du;
    U[pa] = (T)v;
    U[pa + 1] = (T)v;
}

// ------------------------------------------------------------------------
// Projection pieces
// ------------------------------------------------------------------------
// phi with the x-Neumann and z-periodic copies of apply_periodic_z_pressure
__device__ __forceinline__ long long phi_at(const Dims &d, int i, int j, int k) {
    if (i > d.nxg - 3) i = d.nxg - 3;
    if (k == d.nzg - 2) k = 1;
    return fidx(d, i, j, k);
}

// a thread walks kProjZ planes of one (i, j) column, so phi(k+1) of a plane
// is phi(k) of the next; all loads of its planes are issued before the first
// update (0.81 -> 0.71 ms/step in FP32 at 814x257x193).  Tiles cover cells
// 64 bx .. 64 bx + 63 (cell 0 idle), so that a wave's row segment starts on
// a 128-byte line (-5 % FP64, -7 % FP32).  Tiles in XCD-contiguous order in
// FP32 only: at that size it is 3 % faster there and 4 % slower in FP64 than
// plain order.
constexpr int kProjZ = 4;

// With P set (the first sub-step), also P = phi / den on the same cells
// (projection.f90:75-85), and on the ghost cells next to them the same value
// (the copies of apply_periodic_z_pressure and the P ghost faces), without a
// second read of phi.
template <class T>
__global__ __launch_bounds__(256) void k_project(Dims d, MetricsT<T> m, T *U, T *V, T *W, const T *phi, T *P,
                                                 double den) {
    const int nbx = (d.nxg - 1 + kWave - 1) / kWave, nby = (d.nyg - 2 + 3) / 4;
    const int ntile = nbx * nby * ((d.nzg - 2 + kProjZ - 1) / kProjZ);
    const int t = sizeof(T) == 4 ? xcd_tile(blockIdx.x, ntile) : blockIdx.x;
    if (t >= ntile) return;
    const int i = (t % nbx) * kWave + threadIdx.x;
    const int j = ((t / nbx) % nby) * 4 + threadIdx.y + 1;
    const int k0 = (t / (nbx * nby)) * kProjZ + 1;
    if (i < 1 || i > d.nxg - 2 || j > d.nyg - 2) return;
    const int k1 = min(k0 + kProjZ, d.nzg - 1);  // exclusive
    const T rdxg = m.rdxg[i], rdyg = m.rdyg[j];
    // every load of the kProjZ planes first (planes past the last one clamped
    // and not stored), then the updates
    T pz[kProjZ + 1], px[kProjZ], py[kProjZ], u[kProjZ], v[kProjZ], w[kProjZ], rz[kProjZ];
    pz[0] = phi[phi_at(d, i, j, k0)];
#pragma unroll
    for (int r = 0; r < kProjZ; ++r) {
        const int k = min(k0 + r, k1 - 1);
        const long long p = fidx(d, i, j, k);
        pz[r + 1] = phi[phi_at(d, i, j, k + 1)];
        px[r] = phi[phi_at(d, i + 1, j, k)];
        py[r] = phi[phi_at(d, i, j + 1, k)];
        u[r] = U[p];
        v[r] = V[p];
        w[r] = W[p];
        rz[r] = m.rdzg[k];
    }
#pragma unroll
    for (int r = 0; r < kProjZ; ++r) {
        const int k = k0 + r;
        if (k >= k1) break;
        const long long p = fidx(d, i, j, k);
        if (i <= d.nx - 2) U[p] = u[r] - (px[r] - pz[r]) * rdxg;
        if (j <= d.ny - 2) V[p] = v[r] - (py[r] - pz[r]) * rdyg;
        if (k <= d.nz - 2) W[p] = w[r] - (pz[r + 1] - pz[r]) * rz[r];
    }
    if (P) {
        const int gi = i == 1 ? 0 : (i == d.nxg - 2 ? d.nxg - 1 : -1);
        const int gj = j == 1 ? 0 : (j == d.nyg - 2 ? d.nyg - 1 : -1);
#pragma unroll
        for (int r = 0; r < kProjZ; ++r) {
            const int k = k0 + r;
            if (k >= k1) break;
            const int gk = k == 1 ? 0 : (k == d.nzg - 2 ? d.nzg - 1 : -1);
            const T pv = (T)((double)pz[r] / den);
            for (int a = 0; a < 2; ++a) {
                const int ii = a ? gi : i;
                if (ii < 0) continue;
                for (int b = 0; b < 2; ++b) {
                    const int jj = b ? gj : j;
                    if (jj < 0) continue;
                    P[fidx(d, ii, jj, k)] = pv;
                    if (gk >= 0) P[fidx(d, ii, jj, gk)] = pv;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// Real DFT along z (M = nzp points, periodic) of many columns, folded by the
// symmetry m <-> M-m so that each (cos, sin) pair of outputs costs M FMAs.
// Spectrum rows: q = 0 (k = 0), q = 2k-1 Re X_k, q = 2k Im X_k with
// X_k = sum_m T[m] e^{-2 pi i k m/M} (the z part of the reference's 2-D FFT,
// projection.f90:123); the inverse is the unnormalised sum over all k.
// Tables (transposed for wave-uniform scalar loads): ctab[m][k] = cos(2 pi k m/M),
// stab[m][k] = sin(2 pi k m/M), m, k < lda rows/cols (zero padded).
// Columns are c = kx + N*j, contiguous; stride between z planes is `plane`.
// ------------------------------------------------------------------------
// A block is 256 columns (four waves of 64) times one group of RW output
// rows, so that its four waves read the same table rows.  Tiles run in
// XCD-contiguous order with the row group fastest: the row groups of one
// column block follow each other on one XCD and share its L2 for the input.
// The next PF inputs of a column are loaded ahead of their use.
__device__ __forceinline__ void zdft_tile(int M, int C, int RW, int &k0, int &c0, bool &live) {
    const int ng = (M / 2 + 1 + RW - 1) / RW, nt = ng * ((C + 255) / 256);
    const int t = xcd_tile(blockIdx.x, nt);
    live = t < nt;
    k0 = (t % ng) * RW;
    c0 = (t / ng) * 256;
}

template <int RW, int PF>
__global__ __launch_bounds__(256) void k_zdft_fwd(int M, int C, long long plane, int lda,
                                                  const double *__restrict__ ctab,
                                                  const double *__restrict__ stab,
                                                  const double *__restrict__ in,
                                                  double *__restrict__ out) {
    int k0, c0;
    bool live;
    zdft_tile(M, C, RW, k0, c0, live);
    const int c = c0 + threadIdx.x;
    if (!live || c >= C) return;
    const int Kc = M / 2 + 1;  // cos rows k = 0..M/2
    const double *src = in + c;
    const double t0 = src[0];
    double ac[RW], as[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        ac[r] = t0;
        as[r] = 0.0;
    }
    const int hm = M / 2;
    if (hm >= 1) {
        double rp[PF], rm[PF];
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int m = min(1 + u, hm);
            rp[u] = src[(long long)m * plane];
            rm[u] = src[(long long)(M - m) * plane];
        }
        for (int m0 = 1; m0 <= hm; m0 += PF) {
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int m = m0 + u;
                const double tp = rp[u], tm = rm[u];
                const int mn = min(m + PF, hm);
                rp[u] = src[(long long)mn * plane];
                rm[u] = src[(long long)(M - mn) * plane];
                if (m <= hm) {
                    const double e = (2 * m == M) ? tp : tp + tm;
                    const double o = tp - tm;  // zero for the middle element of even M
                    const double *cr = ctab + (long long)m * lda + k0;
                    const double *sr = stab + (long long)m * lda + k0;
#pragma unroll
                    for (int r = 0; r < RW; ++r) {
                        ac[r] = fma(e, cr[r], ac[r]);
                        as[r] = fma(o, sr[r], as[r]);
                    }
                }
            }
        }
    }
    double *dst = out + c;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int k = k0 + r;
        if (k >= Kc) break;
        dst[(long long)(k == 0 ? 0 : 2 * k - 1) * plane] = ac[r];
        if (k >= 1 && 2 * k <= M - 1) dst[(long long)(2 * k) * plane] = -as[r];
    }
}

// inverse: g[m] = S0 + sum_k c_k (Re_k cos - Im_k sin)(2 pi k m/M), c_k = 2
// (1 for the Nyquist row of even M); outputs m and M-m come from the same
// cos and sin sums.  Row groups cover the outputs m = 0..M/2.
template <int RW, int PF>
__global__ __launch_bounds__(256) void k_zdft_inv(int M, int C, long long plane, int lda,
                                                  const double *__restrict__ ctab,
                                                  const double *__restrict__ stab,
                                                  const double *__restrict__ in,
                                                  double *__restrict__ out) {
    int m0, c0;
    bool live;
    zdft_tile(M, C, RW, m0, c0, live);
    const int c = c0 + threadIdx.x;
    if (!live || c >= C) return;
    const int hm = M / 2;  // outputs m = 0..M/2 (paired with M-m)
    const double *src = in + c;
    const double s0 = src[0];
    double cc[RW], ss[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        cc[r] = s0;
        ss[r] = 0.0;
    }
    const int Kc = M / 2 + 1;
    if (Kc >= 2) {
        // Re_k = row 2k-1, Im_k = row 2k (absent for the Nyquist row of even M)
        double rr[PF], ri[PF];
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int k = min(1 + u, Kc - 1);
            rr[u] = src[(long long)(2 * k - 1) * plane];
            ri[u] = src[(long long)min(2 * k, M - 1) * plane];
        }
        for (int kb = 1; kb < Kc; kb += PF) {
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int k = kb + u;
                const double re = rr[u], imr = ri[u];
                const int kn = min(k + PF, Kc - 1);
                rr[u] = src[(long long)(2 * kn - 1) * plane];
                ri[u] = src[(long long)min(2 * kn, M - 1) * plane];
                if (k < Kc) {
                    const bool nyq = (2 * k == M);
                    const double im = nyq ? 0.0 : imr;
                    const double cw = nyq ? re : 2.0 * re;
                    const double sw = 2.0 * im;
                    // tables indexed [k][m] here: ctab is symmetric in (k, m)
                    const double *cr = ctab + (long long)k * lda + m0;
                    const double *sr = stab + (long long)k * lda + m0;
#pragma unroll
                    for (int r = 0; r < RW; ++r) {
                        cc[r] = fma(cw, cr[r], cc[r]);
                        ss[r] = fma(sw, sr[r], ss[r]);
                    }
                }
            }
        }
    }
    double *dst = out + c;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int m = m0 + r;
        if (m > hm) break;
        dst[(long long)m * plane] = cc[r] - ss[r];
        if (m >= 1 && M - m != m) dst[(long long)(M - m) * plane] = cc[r] + ss[r];
    }
}

// ------------------------------------------------------------------------
// DCT-II x2 along x (forward) and its normalised inverse, per line, by
// Makhoul's reordering and a two-factor (N = N1*N2) DFT of the real
// reordered sequence, all in LDS.  Only k1 <= N1/2 is formed; the other half
// follows from Hermitian symmetry.  Forward: X[k] = sum_n 2 f[n]
// cos(pi k (2n+1)/(2N)) (the reference's 2N-point symmetric extension plus
// twiddle, projection.f90:110-134).  Inverse: f[n] = (1/(N M)) (g[0]/2 +
// sum_k g[k] cos(pi k (2n+1)/(2N))) (projection.f90:160-202 with the
// 1/(2 nxp nzp) factor).
// Tables (cos, sin interleaved): w1[q] = e^{-2 pi i q/N1}, w2[q] = e^{-2 pi i q/N2},
// wN[q] = e^{-2 pi i q/N}, rot[k] = e^{-i pi k/(2N)}.
// ------------------------------------------------------------------------
struct DctArgs {
    int N, N1, N2, K1h;
    int C1, C;  // lines c = c1 + C1*c2
    long long in_base, in_c1, in_c2;
    long long out_base, out_c1, out_c2;
    double scale;  // inverse only: 1/(N*M)
    const double2 *w1, *w2, *wN, *rot;
    // FP64 MFMA operand tables of the two short DFTs, in lane order (see
    // k_dct_fwd_m); null when the shape does not fit those kernels
    const double *fA, *fC, *iC, *iA;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// DIV: the line is the pseudo-pressure rhs (projection.f90:38-47) evaluated
// on the fly from U, V, W (in = U, in2 = V, in3 = W, lines over cells
// i = 1..N, j = 1..NM, k = 1..M) instead of being read from memory.
template <int LB, bool DIV>
__global__ __launch_bounds__(256) void k_dct_fwd(DctArgs a, Dims d, Metrics mt,
                                                 const double *__restrict__ in,
                                                 const double *__restrict__ in2,
                                                 const double *__restrict__ in3,
                                                 double *__restrict__ out) {
    extern __shared__ double smem[];
    const int N = a.N, N1 = a.N1, N2 = a.N2, K1 = a.K1h + 1;
    double *v = smem;                                    // LB*N
    double2 *B = reinterpret_cast<double2 *>(v + LB * N);  // LB*K1*N2
    double2 *w1 = B + LB * K1 * N2;                      // N1
    double2 *w2 = w1 + N1;                               // N2
    const int tid = threadIdx.x;
    for (int q = tid; q < N1; q += 256) w1[q] = a.w1[q];
    for (int q = tid; q < N2; q += 256) w2[q] = a.w2[q];
    const int line0 = blockIdx.x * LB;
    // load and reorder: v[n/2] = f[n] (n even), v[N-1-(n-1)/2] = f[n] (n odd)
    for (int idx = tid; idx < LB * N; idx += 256) {
        const int l = idx / N, n = idx - l * N;
        const int c = min(line0 + l, a.C - 1);
        const int c1 = c % a.C1, c2 = c / a.C1;
        const long long p = a.in_base + c1 * a.in_c1 + c2 * a.in_c2 + n;
        double f;
        if (DIV) {
            f = (in[p] - in[p - 1]) * mt.rdx[n + 1] + (in2[p] - in2[p - d.sx]) * mt.rdy[c1 + 1] +
                (in3[p] - in3[p - d.sxy]) * mt.rdz[c2 + 1];
        } else {
            f = in[p];
        }
        v[l * N + ((n & 1) ? N - 1 - (n >> 1) : (n >> 1))] = f;
    }
    __syncthreads();
    // step A: B[k1][n2] = W_N^{n2 k1} sum_n1 v[N2 n1 + n2] W_N1^{n1 k1}
    for (int idx = tid; idx < LB * K1 * N2; idx += 256) {
        const int l = idx / (K1 * N2), r = idx - l * K1 * N2;
        const int k1 = r / N2, n2 = r - k1 * N2;
        const double *vl = v + l * N + n2;
        double re = 0.0, im = 0.0;
        int q = 0;
        for (int n1 = 0; n1 < N1; ++n1) {
            const double x = vl[n1 * N2];
            const double2 w = w1[q];
            re = fma(x, w.x, re);
            im = fma(x, w.y, im);
            q += k1;
            if (q >= N1) q -= N1;
        }
        B[idx] = cmul(make_double2(re, im), a.wN[(n2 * k1) % N]);
    }
    __syncthreads();
    // step C: V[k1 + N1 k2] = sum_n2 B[k1][n2] W_N2^{n2 k2}; X[k] = 2 Re(rot_k V),
    // X[N-k] = -2 Im(rot_k V)
    for (int idx = tid; idx < LB * K1 * N2; idx += 256) {
        const int l = idx / (K1 * N2), r = idx - l * K1 * N2;
        const int k1 = r / N2, k2 = r - k1 * N2;
        const double2 *Bl = B + l * K1 * N2 + k1 * N2;
        double re = 0.0, im = 0.0;
        int q = 0;
        for (int n2 = 0; n2 < N2; ++n2) {
            const double2 b = Bl[n2], w = w2[q];
            re = fma(b.x, w.x, fma(-b.y, w.y, re));
            im = fma(b.x, w.y, fma(b.y, w.x, im));
            q += k2;
            if (q >= N2) q -= N2;
        }
        const int k = k1 + N1 * k2;
        const double2 wv = cmul(a.rot[k], make_double2(re, im));
        v[l * N + k] = 2.0 * wv.x;
        // the mirror N - k of k1 = 0 and k1 = N1/2 is written directly (k1' = k1)
        if (k1 != 0 && 2 * k1 != N1) v[l * N + N - k] = -2.0 * wv.y;
    }
    __syncthreads();
    for (int idx = tid; idx < LB * N; idx += 256) {
        const int l = idx / N, kx = idx - l * N;
        const int c = line0 + l;
        if (c >= a.C) continue;
        const int c1 = c % a.C1, c2 = c / a.C1;
        out[a.out_base + c1 * a.out_c1 + c2 * a.out_c2 + kx] = v[l * N + kx];
    }
}

template <int LB>
__global__ __launch_bounds__(256) void k_dct_inv(DctArgs a, const double *__restrict__ in,
                                                 double *__restrict__ out) {
    extern __shared__ double smem[];
    const int N = a.N, N1 = a.N1, N2 = a.N2, K1h = a.K1h, K1 = K1h + 1;
    double2 *Vs = reinterpret_cast<double2 *>(smem);  // LB*N, later the output lines (as doubles)
    double2 *C = Vs + LB * N;                         // LB*K1*N2
    double2 *w1 = C + LB * K1 * N2;
    double2 *w2 = w1 + N1;
    const int tid = threadIdx.x;
    for (int q = tid; q < N1; q += 256) w1[q] = a.w1[q];
    for (int q = tid; q < N2; q += 256) w2[q] = a.w2[q];
    const int line0 = blockIdx.x * LB;
    // V[k] = conj(rot_k) (g[k] - i g[N-k]) / 2, g read straight from the spectrum line
    for (int idx = tid; idx < LB * N; idx += 256) {
        const int l = idx / N, k = idx - l * N;
        const int c = min(line0 + l, a.C - 1);
        const int c1 = c % a.C1, c2 = c / a.C1;
        const double *gl = in + a.in_base + c1 * a.in_c1 + c2 * a.in_c2;
        const double gk = gl[k], gm = (k == 0) ? 0.0 : gl[N - k];
        Vs[idx] = cmulc(make_double2(0.5 * gk, -0.5 * gm), a.rot[k]);
    }
    __syncthreads();
    // step C': C~[k1][n2] = W_N^{-n2 k1} sum_k2 V[k1 + N1 k2] W_N2^{-n2 k2}
    for (int idx = tid; idx < LB * K1 * N2; idx += 256) {
        const int l = idx / (K1 * N2), r = idx - l * K1 * N2;
        const int k1 = r / N2, n2 = r - k1 * N2;
        const double2 *Vl = Vs + l * N + k1;
        double re = 0.0, im = 0.0;
        int q = 0;
        for (int k2 = 0; k2 < N2; ++k2) {
            const double2 V = Vl[N1 * k2];
            const double2 w = w2[q];  // use conj(w)
            re = fma(V.x, w.x, fma(V.y, w.y, re));
            im = fma(V.y, w.x, fma(-V.x, w.y, im));
            q += n2;
            if (q >= N2) q -= N2;
        }
        C[idx] = cmulc(make_double2(re, im), a.wN[(n2 * k1) % N]);
    }
    __syncthreads();
    // step A': v[N2 n1 + n2] = (1/N) (Re C~[0] + sum_{k1>=1} c_k1 Re(C~[k1] W_N1^{-n1 k1}))
    double *o = smem;  // LB*N doubles inside the V area, which is no longer read
    for (int idx = tid; idx < LB * N; idx += 256) {
        const int l = idx / N, n = idx - l * N;
        const int n1 = n / N2, n2 = n - n1 * N2;
        const double2 *Cl = C + l * K1 * N2 + n2;
        double acc = Cl[0].x;
        int q = 0;
        for (int k1 = 1; k1 <= K1h; ++k1) {
            q += n1;
            if (q >= N1) q -= N1;
            const double2 c = Cl[k1 * N2], w = w1[q];
            const double t = c.x * w.x + c.y * w.y;  // Re(c * conj(w))
            acc += ((N1 % 2 == 0) && k1 == K1h) ? t : 2.0 * t;
        }
        const int pos = (n < (N + 1) / 2) ? 2 * n : 2 * (N - 1 - n) + 1;
        o[l * N + pos] = acc * a.scale;
    }
    __syncthreads();
    for (int idx = tid; idx < LB * N; idx += 256) {
        const int l = idx / N, n = idx - l * N;
        const int c = line0 + l;
        if (c >= a.C) continue;
        const int c1 = c % a.C1, c2 = c / a.C1;
        out[a.out_base + c1 * a.out_c1 + c2 * a.out_c2 + n] = o[idx];
    }
}

// ------------------------------------------------------------------------
// The same transforms with both short DFTs on the FP64 matrix cores
// (v_mfma_f64_16x16x4_f64; lane l holds A[l%16][l/16], B[l/16][l%16] and
// D[l/16 + 4r][l%16], r < 4).  kMfL lines per block, four waves.
// Step A / A' use 2 kMfL column tiles (kMfL / 2 per wave), step C / C' kMfL.
// Forward, step A: [Re, Im of B[k1][n2]] = WA [N1 x (line, n2)] with the
//   rows of WA ordered so that a lane's D rows r = 0, 1 are Re, Im of one k1
//   and r = 2, 3 those of k1 + 4; the twiddle W_N^{n2 k1} is applied in
//   registers and B' goes to LDS as complex.
// Forward, step C: [Re, Im of V[k1 + N1 k2]] = WC [(n2, re/im) x (line, k1)],
//   rows paired the same way over k2; the post-rotation and the Makhoul
//   placement write the output line into LDS.
// Inverse, step C': the conjugate DFT over k2, rows paired over n2, then the
//   conjugate twiddle; step A': the real output v[N2 n1 + n2] as one
//   real matrix WA' [n1 x (k1, re/im)] with the Hermitian weights (1, 2, .., 1).
// The operand tables are built on the host in lane order (one double per
// lane per (row tile, k step)); shapes: N1 <= 32, N1/2 + 1 <= 16, N2 <= 32.
// ------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr int kMfL = 2;                   // lines per block
constexpr int kLd = 4;                    // loads per thread in flight in the line load phases
constexpr int kMfKA = 8, kMfKC = 16;      // k steps: step A (N1 <= 32), step C (2 N2 <= 64)

__device__ __forceinline__ dbl4 mfma64(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------------
// The folded real z DFT above on the FP64 matrix cores, for M/2 + 1 <= 96
// rows.  Both directions are two real products D = A B over the folded
// index kk = 0..M/2 with A[row][kk] = cos or sin(2 pi row kk / M), the entry
// (row kk) mod M of a single M-point table:
//   forward  B = e[kk], o[kk] (the pairs T[kk] +- T[M-kk]); rows k give
//            Re X_k and -Im X_k;
//   inverse  B = R[k], I[k] (the Hermitian-weighted Re and Im rows); rows m
//            give c[m], s[m], and g[m] = c - s, g[M-m] = c + s.
// The products run in kk order, so the sums are those of the VALU kernels
// bit for bit.  A block holds every A fragment (cos, sin) in LDS in lane
// order (RT x KS x 64 pairs, 144 KB at M = 191; the table zf of
// z_dft_tables) and its waves loop over
// 16-column tiles, each owning every row tile of both products for its
// columns; the raw inputs of the next kZmPF k steps (into the next tile at
// the end of one) are loaded ahead.
// ------------------------------------------------------------------------
constexpr int kZmRT = 6, kZmKS = 24, kZmPF = 4, kZmW = 8;

// RT = (M/2 + 1 + 15) / 16 row tiles
template <bool FWD, int RT>
__global__ __launch_bounds__(64 * kZmW) void k_zdft_mf(int M, int C, long long plane,
                                                       const double2 *__restrict__ zf,
                                                       const double *__restrict__ in,
                                                       double *__restrict__ out) {
    __shared__ double2 sA[RT * kZmKS * 64];
    const int hm = M / 2, Kc = hm + 1, KS = (Kc + 3) / 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, q4 = lane >> 4;
    // fragment (t, ks), lane l: A[16 t + l % 16][4 ks + l / 16] as (cos, sin),
    // built on the host in this order (zf)
    for (int e = threadIdx.x; e < RT * KS * 64; e += 64 * kZmW) sA[e] = zf[e];
    __syncthreads();
    const int ntile = (C + 15) / 16, stride = gridDim.x * kZmW;
    int tile = blockIdx.x * kZmW + wave;
    if (tile >= ntile) return;
    // k steps past M/2 multiply zeros
    const int KSr = (KS + kZmPF - 1) / kZmPF * kZmPF;
    // raw inputs of folded index kk = 4 ks + q4 of tile tl (rows clamped to
    // M/2, columns to C - 1); the loads are unconditional and masked only
    // where they are used, so that those issued ahead stay in flight
    auto raw = [&](int tl, int ks, double &a, double &b) {
        const int kk = min(4 * ks + q4, hm);
        const bool single = kk == 0 || 2 * kk == M;
        int ra, rb;
        if (FWD) {
            ra = kk;
            rb = single ? kk : M - kk;
        } else {
            ra = kk == 0 ? 0 : (2 * kk == M ? M - 1 : 2 * kk - 1);
            rb = single ? ra : 2 * kk;
        }
        const double *src = in + min(16 * tl + r16, C - 1);
        a = src[(long long)ra * plane];
        b = src[(long long)rb * plane];
    };
    double pa[kZmPF], pb[kZmPF];
#pragma unroll
    for (int u = 0; u < kZmPF; ++u) raw(tile, u, pa[u], pb[u]);

    for (; tile < ntile; tile += stride) {
        dbl4 ac[RT], as[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            ac[t] = dbl4{0.0, 0.0, 0.0, 0.0};
            as[t] = dbl4{0.0, 0.0, 0.0, 0.0};
        }
        for (int ks0 = 0; ks0 < KSr; ks0 += kZmPF) {
#pragma unroll
            for (int u = 0; u < kZmPF; ++u) {
                const int ks = ks0 + u;
                const int kk = 4 * ks + q4;
                const bool live = kk <= hm, single = kk == 0 || 2 * kk == M;
                const double a = live ? pa[u] : 0.0, b = live && !single ? pb[u] : 0.0;
                // ahead: k step ks + kZmPF of this tile, or of the next one
                const int kn = ks + kZmPF;
                const bool here = kn < KSr;
                raw(here ? tile : min(tile + stride, ntile - 1), here ? kn : kn - KSr, pa[u], pb[u]);
                double bc, bs;
                if (FWD) {
                    bc = a + b;                  // b = 0 for the single elements
                    bs = single ? 0.0 : a - b;
                } else {
                    bc = single ? a : 2.0 * a;   // R: weight 1 for k = 0 and the Nyquist row
                    bs = 2.0 * b;                // I: b = 0 there
                }
                // A fragments of this step (clamped to the table: steps past
                // M/2 multiply zero operands)
                const double2 *fa = sA + min(ks, KS - 1) * 64 + lane;
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    const double2 w = fa[t * KS * 64];
                    ac[t] = mfma64(w.x, bc, ac[t]);
                    as[t] = mfma64(w.y, bs, as[t]);
                }
            }
        }
        const int col = 16 * tile + r16;
        if (col < C) {
            double *dst = out + col;
#pragma unroll
            for (int t = 0; t < RT; ++t) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * t + q4 + 4 * r;
                    if (row > hm) continue;
                    const double c = ac[t][r], s = as[t][r];
                    if (FWD) {
                        dst[(long long)(row == 0 ? 0 : 2 * row - 1) * plane] = c;
                        if (row >= 1 && 2 * row <= M - 1) dst[(long long)(2 * row) * plane] = -s;
                    } else {
                        dst[(long long)row * plane] = c - s;
                        if (row >= 1 && M - row != row) dst[(long long)(M - row) * plane] = c + s;
                    }
                }
            }
        }
    }
}

template <bool DIV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_dct_fwd_m(DctArgs a, Dims d, Metrics mt, const double *__restrict__ in,
                                                   const double *__restrict__ in2, const double *__restrict__ in3,
                                                   double *__restrict__ out) {
    extern __shared__ double smem[];
    const int N = a.N, N1 = a.N1, N2 = a.N2, K1 = a.K1h + 1;
    // twiddle indices n2 k1 stay below N (n2 < N2, k1 <= N1/2): no reduction
    const float rN2 = 1.0f / (float)N2, rK1 = 1.0f / (float)K1;
    // v (kMfL*N doubles) and Bc (kMfL*K1*N2 complex) share one LDS area: each
    // step finishes reading its input before a barrier and then writes its output
    double *v = smem;
    double2 *Bc = reinterpret_cast<double2 *>(smem);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, q = lane >> 4, c16 = lane & 15;
    const int line0 = blockIdx.x * kMfL;
    // load and reorder: v[n/2] = f[n] (n even), v[N-1-(n-1)/2] = f[n] (n odd);
    // kLd elements per thread are loaded before the first one is used
    static_assert(kMfL == 2, "line index below assumes two lines per block");
    long long lb[kMfL];
    double ry[kMfL], rz[kMfL];
#pragma unroll
    for (int l = 0; l < kMfL; ++l) {
        const int c = min(line0 + l, a.C - 1);
        const int c1 = c % a.C1, c2 = c / a.C1;
        lb[l] = a.in_base + c1 * a.in_c1 + c2 * a.in_c2;
        ry[l] = DIV ? mt.rdy[c1 + 1] : 0.0;
        rz[l] = DIV ? mt.rdz[c2 + 1] : 0.0;
    }
    for (int i0 = tid; i0 < kMfL * N; i0 += 256 * kLd) {
        double f[kLd];
#pragma unroll
        for (int u = 0; u < kLd; ++u) {
            const int idx = min(i0 + 256 * u, kMfL * N - 1);
            const int l = idx >= N ? 1 : 0, n = idx - l * N;
            const long long p = lb[l] + n;
            if (DIV) {
                f[u] = (in[p] - in[p - 1]) * mt.rdx[n + 1] + (in2[p] - in2[p - d.sx]) * ry[l] +
                       (in3[p] - in3[p - d.sxy]) * rz[l];
            } else {
                f[u] = in[p];
            }
        }
#pragma unroll
        for (int u = 0; u < kLd; ++u) {
            const int idx = i0 + 256 * u;
            if (idx < kMfL * N) {
                const int l = idx >= N ? 1 : 0, n = idx - l * N;
                v[l * N + ((n & 1) ? N - 1 - (n >> 1) : (n >> 1))] = f[u];
            }
        }
    }
    const int KSA = (N1 + 3) / 4, KSC = (2 * N2 + 3) / 4;
    double fa[2][kMfKA];
#pragma unroll
    for (int ks = 0; ks < kMfKA; ++ks) {
        fa[0][ks] = ks < KSA ? a.fA[(0 * KSA + ks) * 64 + lane] : 0.0;
        fa[1][ks] = ks < KSA ? a.fA[(1 * KSA + ks) * 64 + lane] : 0.0;
    }
    __syncthreads();
    // ---- step A: wave w owns kMfL / 2 column tiles (columns (line, n2))
    for (int ct = wave * (kMfL / 2); ct < (wave + 1) * (kMfL / 2); ++ct) {
        const int col = ct * 16 + c16;
        const int ln = small_div(col, rN2), n2 = col - ln * N2;
        const bool cok = col < kMfL * N2;
        dbl4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < kMfKA; ++ks) {
            if (ks < KSA) {
                const int n1 = 4 * ks + q;
                const double b = (cok && n1 < N1) ? v[ln * N + N2 * n1 + n2] : 0.0;
                acc0 = mfma64(fa[0][ks], b, acc0);
                acc1 = mfma64(fa[1][ks], b, acc1);
            }
        }
        static_assert(kMfL == 2, "one step-A column tile per wave");
        __syncthreads();  // v is read; Bc overwrites it
        if (cok) {
#pragma unroll
            for (int h = 0; h < 4; ++h) {  // (row tile, pair): k1 = 8 (h/2) + q + 4 (h%2)
                const dbl4 &acc = (h < 2) ? acc0 : acc1;
                const int k1 = 8 * (h / 2) + q + 4 * (h % 2);
                if (k1 < K1) {
                    const double re = acc[2 * (h % 2)], im = acc[2 * (h % 2) + 1];
                    Bc[(ln * K1 + k1) * N2 + n2] = cmul(make_double2(re, im), a.wN[n2 * k1]);
                }
            }
        }
    }
    __syncthreads();
    // ---- step C: wave w owns row tile w (k2 = 8w + q, 8w + q + 4), all
    // kMfL column tiles (columns (line, k1))
    {
        double fc[kMfKC];
#pragma unroll
        for (int ks = 0; ks < kMfKC; ++ks) fc[ks] = ks < KSC ? a.fC[(wave * KSC + ks) * 64 + lane] : 0.0;
        dbl4 accs[kMfL];
#pragma unroll
        for (int ct = 0; ct < kMfL; ++ct) {
            const int col = ct * 16 + c16;
            const int ln = small_div(col, rK1), k1 = col - ln * K1;
            const bool cok = col < kMfL * K1;
            const double *Bl = reinterpret_cast<const double *>(Bc + (ln * K1 + k1) * N2);
            dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < kMfKC; ++ks) {
                if (ks < KSC) {
                    const int kk = 4 * ks + q;  // (n2, re/im) = (kk / 2, kk % 2)
                    const double b = (cok && (kk >> 1) < N2) ? Bl[kk] : 0.0;
                    acc = mfma64(fc[ks], b, acc);
                }
            }
            accs[ct] = acc;
        }
        __syncthreads();  // Bc is read; v overwrites it
#pragma unroll
        for (int ct = 0; ct < kMfL; ++ct) {
            const int col = ct * 16 + c16;
            const int ln = small_div(col, rK1), k1 = col - ln * K1;
            const bool cok = col < kMfL * K1;
            const dbl4 acc = accs[ct];
            if (cok) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int k2 = 8 * wave + q + 4 * h;
                    if (k2 < N2) {
                        const int k = k1 + N1 * k2;
                        const double2 wv = cmul(a.rot[k], make_double2(acc[2 * h], acc[2 * h + 1]));
                        v[ln * N + k] = 2.0 * wv.x;
                        // the mirror N - k of k1 = 0 and k1 = N1/2 is written
                        // directly by another lane (k1' = k1): one writer each
                        if (k1 != 0 && 2 * k1 != N1) v[ln * N + N - k] = -2.0 * wv.y;
                    }
                }
            }
        }
    }
    __syncthreads();
    // per-line output bases, then one element per thread and step (kMfL = 2)
    long long ob[kMfL];
#pragma unroll
    for (int l = 0; l < kMfL; ++l) {
        const int c = min(line0 + l, a.C - 1);
        ob[l] = a.out_base + (long long)(c % a.C1) * a.out_c1 + (long long)(c / a.C1) * a.out_c2;
    }
    for (int idx = tid; idx < kMfL * N; idx += 256) {
        const int l = idx >= N ? 1 : 0, kx = idx - l * N;
        if (line0 + l >= a.C) continue;
        out[ob[l] + kx] = v[idx];
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_dct_inv_m(DctArgs a, const double *__restrict__ in, double *__restrict__ out) {
    extern __shared__ double smem[];
    const int N = a.N, N1 = a.N1, N2 = a.N2, K1h = a.K1h, K1 = K1h + 1;
    // twiddle indices n2 k1 stay below N (n2 < N2, k1 <= N1/2): no reduction
    const float rN2 = 1.0f / (float)N2, rK1 = 1.0f / (float)K1;
    // Vs, Cs (kMfL*K1*N2 complex each) and the output lines o (kMfL*N doubles)
    // share one LDS area: each step finishes reading its input before a
    // barrier and then writes its output
    double2 *Vs = reinterpret_cast<double2 *>(smem);
    double2 *Cs = Vs;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, q = lane >> 4, c16 = lane & 15;
    const int line0 = blockIdx.x * kMfL;
    // V[k1 + N1 k2] = conj(rot_k) (g[k] - i g[N-k]) / 2 for k1 <= N1/2: Vs[l][k1][k2];
    // kLd elements per thread are loaded before the first one is used
    static_assert(kMfL == 2, "line index below assumes two lines per block");
    const double *gls[kMfL];
#pragma unroll
    for (int l = 0; l < kMfL; ++l) {
        const int c = min(line0 + l, a.C - 1);
        const int c1 = c % a.C1, c2 = c / a.C1;
        gls[l] = in + a.in_base + c1 * a.in_c1 + c2 * a.in_c2;
    }
    const int nv = kMfL * K1 * N2;
    for (int i0 = tid; i0 < nv; i0 += 256 * kLd) {
        double gk[kLd], gm[kLd];
        double2 rt[kLd];
#pragma unroll
        for (int u = 0; u < kLd; ++u) {
            const int idx = min(i0 + 256 * u, nv - 1);
            const int l = idx >= K1 * N2 ? 1 : 0, r = idx - l * K1 * N2;
            const int k1 = small_div(r, rN2), k2 = r - k1 * N2;
            const int k = k1 + N1 * k2;
            gk[u] = gls[l][k];
            gm[u] = gls[l][k == 0 ? 0 : N - k];
            rt[u] = a.rot[k];
            if (k == 0) gm[u] = 0.0;
        }
#pragma unroll
        for (int u = 0; u < kLd; ++u) {
            const int idx = i0 + 256 * u;
            if (idx < nv) Vs[idx] = cmulc(make_double2(0.5 * gk[u], -0.5 * gm[u]), rt[u]);
        }
    }
    const int KSC = (2 * N2 + 3) / 4, KSA = (2 * K1 + 3) / 4;
    __syncthreads();
    // ---- step C': wave w owns row tile w (n2 = 8w + q, 8w + q + 4), all
    // kMfL column tiles (columns (line, k1))
    {
        double fc[kMfKC];
#pragma unroll
        for (int ks = 0; ks < kMfKC; ++ks) fc[ks] = ks < KSC ? a.iC[(wave * KSC + ks) * 64 + lane] : 0.0;
        dbl4 accs[kMfL];
#pragma unroll
        for (int ct = 0; ct < kMfL; ++ct) {
            const int col = ct * 16 + c16;
            const int ln = small_div(col, rK1), k1 = col - ln * K1;
            const bool cok = col < kMfL * K1;
            const double *Vl = reinterpret_cast<const double *>(Vs + (ln * K1 + k1) * N2);
            dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < kMfKC; ++ks) {
                if (ks < KSC) {
                    const int kk = 4 * ks + q;  // (k2, re/im) = (kk / 2, kk % 2)
                    const double b = (cok && (kk >> 1) < N2) ? Vl[kk] : 0.0;
                    acc = mfma64(fc[ks], b, acc);
                }
            }
            accs[ct] = acc;
        }
        __syncthreads();  // Vs is read; Cs overwrites it
#pragma unroll
        for (int ct = 0; ct < kMfL; ++ct) {
            const int col = ct * 16 + c16;
            const int ln = small_div(col, rK1), k1 = col - ln * K1;
            const bool cok = col < kMfL * K1;
            const dbl4 acc = accs[ct];
            if (cok) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int n2 = 8 * wave + q + 4 * h;
                    if (n2 < N2)
                        Cs[(ln * K1 + k1) * N2 + n2] =
                            cmulc(make_double2(acc[2 * h], acc[2 * h + 1]), a.wN[n2 * k1]);
                }
            }
        }
    }
    __syncthreads();
    // ---- step A': v[N2 n1 + n2] = sum over (k1, re/im) of WA'[n1][.] C~;
    // wave w owns kMfL / 2 column tiles (columns (line, n2)), both row tiles
    double *o = smem;  // kMfL*N doubles over Cs, written after Cs is read
    {
        double fa[2][kMfKA];
#pragma unroll
        for (int ks = 0; ks < kMfKA; ++ks) {
            fa[0][ks] = ks < KSA ? a.iA[(0 * KSA + ks) * 64 + lane] : 0.0;
            fa[1][ks] = ks < KSA ? a.iA[(1 * KSA + ks) * 64 + lane] : 0.0;
        }
        for (int ct = wave * (kMfL / 2); ct < (wave + 1) * (kMfL / 2); ++ct) {
            const int col = ct * 16 + c16;
            const int ln = small_div(col, rN2), n2 = col - ln * N2;
            const bool cok = col < kMfL * N2;
            dbl4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < kMfKA; ++ks) {
                if (ks < KSA) {
                    const int kk = 4 * ks + q;  // (k1, re/im)
                    const int k1 = kk >> 1;
                    const double b =
                        (cok && k1 < K1) ? reinterpret_cast<const double *>(Cs + (ln * K1 + k1) * N2 + n2)[kk & 1]
                                         : 0.0;
                    acc0 = mfma64(fa[0][ks], b, acc0);
                    acc1 = mfma64(fa[1][ks], b, acc1);
                }
            }
            static_assert(kMfL == 2, "one step-A' column tile per wave");
            __syncthreads();  // Cs is read; o overwrites it
            if (cok) {
#pragma unroll
                for (int r = 0; r < 8; ++r) {  // n1 = 16 (r / 4) + q + 4 (r % 4)
                    const int n1 = 16 * (r / 4) + q + 4 * (r % 4);
                    if (n1 < N1) {
                        const int n = N2 * n1 + n2;
                        const int pos = (n < (N + 1) / 2) ? 2 * n : 2 * (N - 1 - n) + 1;
                        o[ln * N + pos] = (r < 4 ? acc0[r % 4] : acc1[r % 4]) * a.scale;
                    }
                }
            }
        }
    }
    __syncthreads();
    // per-line output bases, then one element per thread and step (kMfL = 2)
    long long ob[kMfL];
#pragma unroll
    for (int l = 0; l < kMfL; ++l) {
        const int c = min(line0 + l, a.C - 1);
        ob[l] = a.out_base + (long long)(c % a.C1) * a.out_c1 + (long long)(c / a.C1) * a.out_c2;
    }
    for (int idx = tid; idx < kMfL * N; idx += 256) {
        const int l = idx >= N ? 1 : 0, n = idx - l * N;
        if (line0 + l >= a.C) continue;
        out[ob[l] + n] = o[idx];
    }
}

// Thomas sweep with precomputed factors (projection.f90:137-150).  One real
// system (kx, q) per lane: q = 0 is the kz = 0 row, q = 2k-1 / 2k the real /
// imaginary rows of wavenumber k, which share the factors of k.  The sweep is
// unrolled by TJ rows so that the loads of a chunk are issued together.
template <int TJ>
__global__ __launch_bounds__(256) void k_thomas(Dims d, const double *__restrict__ dl,
                                                const double *__restrict__ dp,
                                                const double *__restrict__ du, double *S) {
    const int N = d.N, NM = d.NM;
    int kxb, q;
    thomas_block(blockIdx.x, (N + 255) / 256, d.M, kxb, q);
    const int kx = kxb * 256 + threadIdx.x;
    if (kx >= N) return;
    const long long plane = (long long)N * NM;
    const int kk = (q + 1) / 2;
    double *a = S + q * plane + kx;
    const double *L = dl + kk * plane + kx;
    const double *D = dp + kk * plane + kx;
    // forward elimination with the stored multipliers
    double r = a[0];
    int j = 1;
    for (; j + TJ <= NM; j += TJ) {
        double x[TJ], l[TJ];
#pragma unroll
        for (int t = 0; t < TJ; ++t) {
            x[t] = a[(long long)(j + t) * N];
            l[t] = L[(long long)(j + t - 1) * N];
        }
#pragma unroll
        for (int t = 0; t < TJ; ++t) {
            r = x[t] - l[t] * r;
            x[t] = r;
        }
#pragma unroll
        for (int t = 0; t < TJ; ++t) a[(long long)(j + t) * N] = x[t];
    }
    for (; j < NM; ++j) {
        r = a[(long long)j * N] - L[(long long)(j - 1) * N] * r;
        a[(long long)j * N] = r;
    }
    // back substitution with the stored reciprocal pivots
    r = r * D[(long long)(NM - 1) * N];
    a[(long long)(NM - 1) * N] = r;
    j = NM - 2;
    for (; j - TJ + 1 >= 0; j -= TJ) {
        double x[TJ], p[TJ], u[TJ];
#pragma unroll
        for (int t = 0; t < TJ; ++t) {
            x[t] = a[(long long)(j - t) * N];
            p[t] = D[(long long)(j - t) * N];
            u[t] = du[j - t];
        }
#pragma unroll
        for (int t = 0; t < TJ; ++t) {
            r = (x[t] - u[t] * r) * p[t];
            x[t] = r;
        }
#pragma unroll
        for (int t = 0; t < TJ; ++t) a[(long long)(j - t) * N] = x[t];
    }
    for (; j >= 0; --j) {
        r = (a[(long long)j * N] - du[j] * r) * D[(long long)j * N];
        a[(long long)j * N] = r;
    }
}

// a copy of a uniform pointer that the compiler cannot prove equal to the
// original (loads through it are not merged with loads through the original)
// 1/x for the normal, finite pivots of the Thomas sweeps: the hardware
// reciprocal refined by two Newton steps (within an ulp or so of the rounded
// quotient, without the special-case scaling of a full division)
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// in the constant address space, so that uniform loads through it are scalar
using cdptr = const double __attribute__((address_space(4))) *;
__device__ __forceinline__ cdptr opaque(const double *p) {
    asm volatile("" : "+s"(p));
    return (cdptr)p;
}

template <int NMAX>
__device__ void solve00_block(const Solve00 &z, int NM);

// Partitioned Thomas solve held in registers (the systems of
// projection.f90:137-150, factors formed on the fly instead of read from
// tables).  A block is P waves; wave p holds rows [32 p, 32 p + 32) of the 64
// systems (q, kx0 + lane), so S is read and written once.  Per segment:
//   LU sweep (top-down): reciprocal pivots ip (kept), y_last, and the last
//     row's responses to the left and right neighbour values xL, xR;
//   UL sweep (bottom-up, nothing kept): y_first and the first row's responses;
// with x = y + alpha xL + beta xR on the segment ends, the 2P interface
// values follow from a two-term recurrence over the segments (every lane
// runs it for its own system), and the segment is solved again with the
// neighbour values moved to its right-hand side.  Rows past NM are identity
// rows (a = c = 0, b = 1).  co = {a[L], b[L], c[L], kxx[N], kzz[Kz]},
// L = 32 P: a[j] couples x(j-1), c[j] couples x(j+1), b[j] + kxx + kzz is
// the diagonal.  The (0,0) system (kx = 0, q = 0, pinned first pivot) is
// skipped (zero00 = 0) or zeroed (zero00 = 1); block 0 solves it (z) while
// the other blocks run.  The other blocks (one resident per CU: 8 waves of
// more than 128 VGPRs each) loop over the nq * ceil(N/64) tiles of 64 systems.
template <class T>
__global__ __launch_bounds__(64 * kThMaxP) void k_thomas_seg(int N, int NM, int P, int nq, int ldj, long long ldq,
                                                             const double *__restrict__ co, int zero00,
                                                             T *__restrict__ S, Solve00 z) {
    __shared__ double xs[6][kThMaxP][64];
    if (blockIdx.x == 0) {  // the (0,0) system, alongside the other blocks
        solve00_block<kThSeg * kThMaxP>(z, NM);
        return;
    }
    const int lane = threadIdx.x & 63;
    const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nkb = (N + 63) / 64, ntiles = nkb * nq, nblk = gridDim.x - 1;
    const int L = P * kThSeg;
    const double *ca = co, *cb = co + L, *cc = co + 2 * L;
    const int r0 = p * kThSeg;
    // tiles (q, 64 kx) bid, bid + nblk, ... (holding the next tile's rows
    // in registers during a solve spills them)
    for (int tile = blockIdx.x - 1; tile < ntiles; tile += nblk) {
        const int q = tile / nkb;
        const int kx = (tile - q * nkb) * 64 + lane;
        const bool live = kx < N;
        const int kxc = live ? kx : N - 1;
        const double lam = co[3 * L + kxc] + co[3 * L + N + (q + 1) / 2];
        T *col = S + q * ldq + kxc;
        double dv[kThSeg], ip[kThSeg];
        {
            // unconditional loads (rows past NM clamped, then zeroed), all
            // issued before the first conversion waits on one
            T raw[kThSeg];
#pragma unroll
            for (int i = 0; i < kThSeg; ++i) raw[i] = col[(long long)min(r0 + i, NM - 1) * ldj];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < kThSeg; ++i) dv[i] = (r0 + i < NM) ? (double)raw[i] : 0.0;
        }
        // Each sweep reads the coefficients through its own opaque copy of the
        // pointer, so that a value loaded by one sweep is not kept in registers
        // for the other sweeps' use of the same row.
        const cdptr la = opaque(ca), lb = opaque(cb), lc = opaque(cc);
        const cdptr ua = opaque(ca), ub = opaque(cb), uc = opaque(cc);
        // LU (index i) and UL (index kThSeg-1-i) sweeps, interleaved
        double dp = dv[0], s = 1.0;
        ip[0] = rcp64(lb[r0] + (r0 < NM ? lam : 0.0));
        const int rl = r0 + kThSeg - 1;
        double iu = rcp64(ub[rl] + (rl < NM ? lam : 0.0)), du = dv[kThSeg - 1], t = 1.0;
#pragma unroll
        for (int i = 1; i < kThSeg; ++i) {
            const int j = r0 + i;
            const double m = la[j] * ip[i - 1];
            ip[i] = rcp64(lb[j] + (j < NM ? lam : 0.0) - m * lc[j - 1]);
            dp = dv[i] - m * dp;
            s = -m * s;
            const int iv = kThSeg - 1 - i, jv = r0 + iv;
            const double mu = uc[jv] * iu;
            iu = rcp64(ub[jv] + (jv < NM ? lam : 0.0) - mu * ua[jv + 1]);
            du = dv[iv] - mu * du;
            t = -mu * t;
            if (i % 4 == 0) __builtin_amdgcn_sched_barrier(0);
        }
        const double aL = ca[r0], cR = cc[r0 + kThSeg - 1];
        xs[0][p][lane] = du * iu;                        // y first
        xs[1][p][lane] = -aL * iu;                       // d x_first / d xL
        xs[2][p][lane] = -cR * t * iu;                   // d x_first / d xR
        xs[3][p][lane] = dp * ip[kThSeg - 1];            // y last
        xs[4][p][lane] = -aL * s * ip[kThSeg - 1];       // d x_last / d xL
        xs[5][p][lane] = -cR * ip[kThSeg - 1];           // d x_last / d xR
        __syncthreads();
        // segment ends: A_p (first row), B_p (last row); B_{p-1} = u + v A_p
        double g[kThMaxP], hh[kThMaxP];
        double u = 0.0, v = 0.0, ul = 0.0, vl = 0.0;  // (ul, vl): B_{p-1} = ul + vl A_p
#pragma unroll
        for (int pp = 0; pp < kThMaxP; ++pp) {
            if (pp < P) {
                if (pp == p) {
                    ul = u;
                    vl = v;
                }
                const double a0 = xs[1][pp][lane], an = xs[4][pp][lane];
                const double den = 1.0 / (1.0 - a0 * v);
                g[pp] = (xs[0][pp][lane] + a0 * u) * den;
                hh[pp] = xs[2][pp][lane] * den;
                const double un = xs[3][pp][lane] + an * (u + v * g[pp]);
                v = an * v * hh[pp] + xs[5][pp][lane];
                u = un;
            }
        }
        __syncthreads();  // xs is read; the next tile writes it
        double Anext = 0.0, xL = 0.0, xR = 0.0;
#pragma unroll
        for (int pp = kThMaxP - 1; pp >= 0; --pp) {
            if (pp < P) {
                const double Ap = g[pp] + hh[pp] * Anext;
                if (pp == p + 1) xR = Ap;
                if (pp == p) xL = ul + vl * Ap;
                Anext = Ap;
            }
        }
        // the segment with its neighbour values on the right-hand side
        const cdptr fa = opaque(ca), fc = opaque(cc);
        dv[0] -= aL * xL;
        dv[kThSeg - 1] -= cR * xR;
#pragma unroll
        for (int i = 1; i < kThSeg; ++i) dv[i] -= fa[r0 + i] * ip[i - 1] * dv[i - 1];
        dv[kThSeg - 1] *= ip[kThSeg - 1];
#pragma unroll
        for (int i = kThSeg - 2; i >= 0; --i) dv[i] = (dv[i] - fc[r0 + i] * dv[i + 1]) * ip[i];
        const bool is00 = q == 0 && kx == 0;
        if (live && !(is00 && !zero00)) {
#pragma unroll
            for (int i = 0; i < kThSeg; ++i)
                if (r0 + i < NM) col[(long long)(r0 + i) * ldj] = is00 ? (T)0 : (T)dv[i];
        }
    }
}

// The pinned (0,0) system (kx = 0, kz = 0) in place in z.x (row stride
// z.ldx), scaled by z.scale: one lane, in the elimination order of the
// table kernel, from LDS copies of the stored multipliers L00 and reciprocal
// pivots D00.  With z.part set, the right-hand side is first formed as
// r(j) = 2 sum_q part[j][q] (the FP64 plane sums of the mixed path's
// divergence), one wave per row.  The sweeps read one array and write
// another, so that the LDS reads of the unrolled iterations are issued ahead
// of the dependent chain.  Any block size that is a multiple of 64.
template <int NMAX>
__device__ void solve00_block(const Solve00 &z, int NM) {
    __shared__ double r[NMAX], o[NMAX], L[NMAX], D[NMAX], U[NMAX];
    const int tid = threadIdx.x, nt = blockDim.x;
    if (z.part) {
        const int lane = tid & 63;
        for (int j = tid >> 6; j < NM; j += nt >> 6) {
            double v = 0.0;
            for (int q = lane; q < z.npart; q += 64) v += z.part[(long long)j * z.npart + q];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
            if (lane == 0) r[j] = 2.0 * v;
        }
    } else {
        for (int j = tid; j < NM; j += nt) r[j] = z.x[(long long)j * z.ldx];
    }
    for (int j = tid; j < NM; j += nt) {
        L[j] = j > 0 ? z.L00[j - 1] : 0.0;
        D[j] = z.D00[j];
        U[j] = z.du[j];
    }
    __syncthreads();
    if (tid == 0) {
        double v = r[0];
        o[0] = v;
#pragma unroll 8
        for (int j = 1; j < NM; ++j) {
            v = r[j] - L[j] * v;
            o[j] = v;
        }
        v *= D[NM - 1];
        r[NM - 1] = v;
#pragma unroll 8
        for (int j = NM - 2; j >= 0; --j) {
            v = (o[j] - U[j] * v) * D[j];
            r[j] = v;
        }
    }
    __syncthreads();
    for (int j = tid; j < NM; j += nt) z.x[(long long)j * z.ldx] = r[j] * z.scale;
}

__global__ __launch_bounds__(256) void k_solve00(int NM, Solve00 z) { solve00_block<kMax00>(z, NM); }

// ------------------------------------------------------------------------
// Reductions
// ------------------------------------------------------------------------
// block max of non-negative values over 4 waves; the caller stores one
// partial per block and k_reduce_max reduces the partials
__device__ __forceinline__ double block_max_256(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o));
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    if ((t & 63) == 0) sh[t >> 6] = v;
    __syncthreads();
    return fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// max |div u| partials (check_divergence projection.f90:391-413); each field
// has its own row and plane strides (the padded solver layout, or the
// reference's unpadded extents for bl_max_divergence_noalloc)
template <class T>
__global__ __launch_bounds__(256) void k_maxdiv(DivArgs a, const T *U, const T *V, const T *W, double *out) {
    __shared__ double sh[4];
    const int i = blockIdx.x * kWave + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y + 1;
    double v = 0.0;
    if (i >= 1 && i <= a.nxg - 3 && j <= a.nyg - 2)
        for (int k = 1 + blockIdx.z; k <= a.nzg - 2; k += gridDim.z) {
            const long long pu = i + a.u1 * j + a.u2 * k, pv = i + a.v1 * j + a.v2 * k;
            const long long pw = i + a.w1 * j + a.w2 * k;
            v = fmax(v, fabs(((double)U[pu] - (double)U[pu - 1]) * a.rdx[i] +
                             ((double)V[pv] - (double)V[pv - a.v1]) * a.rdy[j] +
                             ((double)W[pw] - (double)W[pw - a.w2]) * a.rdz[k]));
        }
    v = block_max_256(v, sh);
    const long long bid = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    if (threadIdx.x == 0 && threadIdx.y == 0) out[bid] = v;
}

template <class T>
__global__ __launch_bounds__(256) void k_dtconv(Dims d, Metrics m, const T *U, const T *V, const T *W,
                                                double *out) {
    __shared__ double sh[3][4];
    const int i = blockIdx.x * kWave + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y + 1;
    double a = 0.0, b = 0.0, c = 0.0;
    if (i >= 1 && i <= d.nxg - 2 && j <= d.nyg - 2) {
        const double dxi = m.xg[i + 1] - m.xg[i], dyj = m.yg[j + 1] - m.yg[j];
        const int gz = gridDim.z;
        // four of the thread's planes loaded before the first is used
        for (int k0 = 1 + blockIdx.z; k0 <= d.nzg - 2; k0 += 4 * gz) {
            T u[4], v[4], w[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long long p = fidx(d, i, j, min(k0 + r * gz, d.nzg - 2));
                u[r] = U[p];
                v[r] = V[p];
                w[r] = W[p];
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = k0 + r * gz;
                if (k > d.nzg - 2) break;
                a = fmax(a, dxi / fabs((double)u[r]));
                b = fmax(b, dyj / fabs((double)v[r]));
                c = fmax(c, (m.zg[k + 1] - m.zg[k]) / fabs((double)w[r]));
            }
        }
    }
    a = block_max_256(a, sh[0]);
    b = block_max_256(b, sh[1]);
    c = block_max_256(c, sh[2]);
    const long long bid = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        out[3 * bid + 0] = a;
        out[3 * bid + 1] = b;
        out[3 * bid + 2] = c;
    }
}

// max over n partial values part[c + stride*q] for c < ncomp, one block
__global__ __launch_bounds__(1024) void k_reduce_max(const double *part, long long n, int ncomp,
                                                     double *out) {
    __shared__ double sh[1024];
    for (int c = 0; c < ncomp; ++c) {
        double v = 0.0;
        for (long long q = threadIdx.x; q < n; q += 1024) v = fmax(v, part[c + ncomp * q]);
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int st = 512; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + st]);
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = sh[0];
        __syncthreads();
    }
}

// Cf(x) (statistics.f90:137-160): one thread per x face
template <class T>
__global__ __launch_bounds__(256) void k_wallstats(Dims d, Metrics m, double nu, const T *U, const T *V,
                                                   const T *nu_t, double *cf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nx) return;
    const double nz2 = (double)(d.nzg - 2);
    // the x neighbour of the last station is column nxg of nu_t_mean, which
    // compute_statistics never fills (statistics.f90:77 runs ii = 1..nx): 0
    const bool last = i == d.nx - 1;
    double u1 = 0.0, u2 = 0.0, uv = 0.0, n1a = 0.0, n2a = 0.0, n1b = 0.0, n2b = 0.0;
    for (int k = 1; k <= d.nzg - 2; ++k) {
        const long long p = fidx(d, i, 0, k);
        const double a = (double)U[p], b = (double)U[p + d.sx];
        u1 += a;
        u2 += b;
        uv += (m.wy0[0] * a + m.wy1[0] * b) * (0.5 * ((double)V[p] + (double)V[p + 1]));
        n1a += (double)nu_t[p];
        n2a += (double)nu_t[p + d.sx];
        if (!last) {
            n1b += (double)nu_t[p + 1];
            n2b += (double)nu_t[p + 1 + d.sx];
        }
    }
    const double dudy = (u2 / nz2 - u1 / nz2) / (m.yg[1] - m.yg[0]);
    const double tau = 0.25 * (n1a / nz2 + n2a / nz2 + n1b / nz2 + n2b / nz2) * dudy;
    cf[i] = 2.0 * (-(uv / nz2) + nu * dudy + tau) / 1.0;
}

// per-plane sums and maxima of the four monitored interiors (monitor.f90:29-52)
template <class T>
__global__ __launch_bounds__(256) void k_monitor(Dims d, const T *U, const T *V, const T *W, const T *nu_t,
                                                 double *part) {
    const int k = blockIdx.x + 1;  // plane
    __shared__ double sh[8][256];
    double s[4] = {0, 0, 0, 0}, mx[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int j = 1; j <= d.nyg - 2; ++j)
        for (int i = threadIdx.x + 1; i <= d.nxg - 2; i += 256) {
            const long long p = fidx(d, i, j, k);
            if (i <= d.nx - 2) {
                s[0] += (double)U[p];
                mx[0] = fmax(mx[0], (double)U[p]);
            }
            if (j <= d.ny - 2) {
                s[1] += (double)V[p];
                mx[1] = fmax(mx[1], (double)V[p]);
            }
            if (k <= d.nz - 2) {
                s[2] += (double)W[p];
                mx[2] = fmax(mx[2], (double)W[p]);
            }
            s[3] += (double)nu_t[p];
            mx[3] = fmax(mx[3], (double)nu_t[p]);
        }
    for (int q = 0; q < 4; ++q) {
        sh[q][threadIdx.x] = s[q];
        sh[4 + q][threadIdx.x] = mx[q];
    }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int q = 0; q < 4; ++q) {
                sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
                sh[4 + q][threadIdx.x] = fmax(sh[4 + q][threadIdx.x], sh[4 + q][threadIdx.x + st]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 8) part[(long long)(k - 1) * 8 + threadIdx.x] = sh[threadIdx.x][0];
}

// reductions: each thread walks z planes with stride kRedZ, so the partials
// per launch are kRedZ times fewer than the interior blocks; x tiles cover
// cells 64 bx .. 64 bx + 63 (cell 0 idle) so that rows start on a 128-byte line
constexpr int kRedZ = 12;
dim3 reduce_grid(const Dims &d) {
    return dim3((d.nxg - 1 + kWave - 1) / kWave, (d.nyg - 2 + 3) / 4, std::min(d.nzg - 2, kRedZ));
}

void check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        last_error() = std::string("kernel launch: ") + hipGetErrorString(e);
        throw DeviceError();
    }
}

constexpr int kDctLines = 2;
constexpr int kZRows = 16, kZPre = 4;

// the matrix-core z DFT with RT row tiles (RT = rt, one instantiation per
// count); persistent blocks, one per compute unit
template <int RT>
void zdft_mf_launch(hipStream_t s, int rt, const ZTables &z, int C, long long plane, bool forward,
                    const double *in, double *out) {
    if constexpr (RT > 1) {
        if (rt < RT) {
            zdft_mf_launch<RT - 1>(s, rt, z, C, plane, forward, in, out);
            return;
        }
    }
    const double2 *zf = reinterpret_cast<const double2 *>(z.zf);
    int dev = 0, ncu = 0;
    BL_HIP(hipGetDevice(&dev));
    BL_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int tiles = (C + 15) / 16;
    const dim3 grid((unsigned)std::max(1, std::min(ncu, (tiles + kZmW - 1) / kZmW)));
    if (forward)
        hipLaunchKernelGGL((k_zdft_mf<true, RT>), grid, dim3(64 * kZmW), 0, s, z.M, C, plane, zf, in, out);
    else
        hipLaunchKernelGGL((k_zdft_mf<false, RT>), grid, dim3(64 * kZmW), 0, s, z.M, C, plane, zf, in, out);
    check_launch();
}

// the real z DFT (M points) of C columns with plane stride `plane`
void zdft_raw(hipStream_t s, const ZTables &z, int C, long long plane, bool forward, const double *in,
              double *out) {
    const int M = z.M, lda = z.lda, rows = M / 2 + 1;
    const double *zc = z.zc, *zs = z.zs;
    if (rows <= 16 * kZmRT) {
        if (!z.zf) {
            last_error() = "z transform: matrix-core tables missing";
            throw DeviceError();
        }
        zdft_mf_launch<kZmRT>(s, (rows + 15) / 16, z, C, plane, forward, in, out);
        return;
    }
    const dim3 grid(xcd_blocks((rows + kZRows - 1) / kZRows * ((C + 255) / 256)));
    if (forward)
        hipLaunchKernelGGL((k_zdft_fwd<kZRows, kZPre>), grid, dim3(256), 0, s, M, C, plane, lda, zc, zs, in, out);
    else
        hipLaunchKernelGGL((k_zdft_inv<kZRows, kZPre>), grid, dim3(256), 0, s, M, C, plane, lda, zc, zs, in, out);
    check_launch();
}

void zdft_launch(Handle &h, int kid, bool forward, const double *in, double *out) {
    ProfScope ps(h, kid);
    const Dims &d = h.d;
    const int C = d.N * d.NM;
    ZTables z;
    z.M = d.M;
    z.lda = h.lda_zt;
    z.zc = h.zc;
    z.zs = h.zs;
    z.zf = h.zf;
    zdft_raw(h.stream, z, C, (long long)C, forward, in, out);
}

void dct_launch(Handle &h, int kid, const DctArgs &a, bool forward, const double *in, double *out,
                const Field3 *div = nullptr) {
    ProfScope ps(h, kid);
    const int K1 = a.K1h + 1;
    if (a.fA) {
        // matrix-core kernels (the shape was checked with the tables)
        // one shared area per block: the lines (kMfL*N doubles) and the
        // intermediate spectra (kMfL*K1*N2 complex) in turn
        const size_t lds = std::max(sizeof(double) * (size_t)kMfL * a.N, sizeof(double2) * (size_t)kMfL * K1 * a.N2);
        if (lds > 163840) {
            last_error() = "x transform length too large for the LDS-resident DCT (N = " + std::to_string(a.N) + ")";
            throw ArgError();
        }
        const dim3 grid((a.C + kMfL - 1) / kMfL);
        if (forward && div)
            hipLaunchKernelGGL(k_dct_fwd_m<true>, grid, dim3(256), lds, h.stream, a, h.d, h.m, div->u, div->v,
                               div->w, out);
        else if (forward)
            hipLaunchKernelGGL(k_dct_fwd_m<false>, grid, dim3(256), lds, h.stream, a, h.d, h.m, in, nullptr,
                               nullptr, out);
        else
            hipLaunchKernelGGL(k_dct_inv_m, grid, dim3(256), lds, h.stream, a, in, out);
        check_launch();
        return;
    }
    const size_t lds = (forward ? sizeof(double) : sizeof(double2)) * (size_t)kDctLines * a.N +
                       sizeof(double2) * ((size_t)kDctLines * K1 * a.N2 + a.N1 + a.N2);
    if (lds > 65536) {
        last_error() = "x transform length too large for the LDS-resident DCT (N = " + std::to_string(a.N) + ")";
        throw ArgError();
    }
    const dim3 grid((a.C + kDctLines - 1) / kDctLines);
    if (forward && div)
        hipLaunchKernelGGL((k_dct_fwd<kDctLines, true>), grid, dim3(256), lds, h.stream, a, h.d, h.m,
                           div->u, div->v, div->w, out);
    else if (forward)
        hipLaunchKernelGGL((k_dct_fwd<kDctLines, false>), grid, dim3(256), lds, h.stream, a, h.d, h.m, in,
                           nullptr, nullptr, out);
    else
        hipLaunchKernelGGL(k_dct_inv<kDctLines>, grid, dim3(256), lds, h.stream, a, in, out);
    check_launch();
}

}  // namespace

void launch_zdft(hipStream_t s, const ZTables &z, int C, long long plane, bool forward, const double *in,
                 double *out) {
    zdft_raw(s, z, C, plane, forward, in, out);
}

const char *kernel_name(int id) {
    static const char *names[KID_COUNT] = {
        "rk_stage", "bc_inout", "bc_periodic", "bc_topwall", "bc_massflux", "divergence",
        "xform_x_fwd", "xform_z_fwd", "thomas", "xform_z_inv", "xform_x_inv", "project",
        "reduce", "dsm_cells", "dsm_vfilter", "dsm_filter", "dsm_column", "closure"};
    return (id >= 0 && id < KID_COUNT) ? names[id] : "";
}

// ------------------------------------------------------------------------
// launch wrappers
// ------------------------------------------------------------------------
template <>
const MetricsT<double> &metrics<double>(const Handle &h) {
    return h.m;
}
template <>
const MetricsT<float> &metrics<float>(const Handle &h) {
    return h.mf;
}


template <class T>
dim3 march_grid(const Dims &d, int chunk) {
    constexpr int kTX = MarchTile<T>::TX, kTY = MarchTile<T>::TY;
    return dim3(xcd_blocks((d.nxg - 1 + kTX - 1) / kTX * ((d.nyg - 2 + kTY - 1) / kTY) *
                           ((d.nzg - 2 + chunk - 1) / chunk)));
}

// z-chunk length of the march.  The tiles run in XCD-contiguous order with
// the z chunk slowest, so with a multiple of 8 chunks each XCD holds whole
// chunk layers and a tile's x and y neighbours (whose halos it shares) sit in
// the same L2; the count is the multiple of 8 that gives chunks of about 24
// planes.  At 814x257x193 (8 chunks of 24 planes) this measured 1 % (FP64)
// and 2.4 % (FP32) faster than 12 chunks of 16, and 20-plane chunks (10
// chunks) 4-5 % slower.  Below 128 planes, chunks of up to 16 planes (the
// two re-read halo planes stay a small share).
inline int march_chunk(const Dims &d) {
    const int nz = d.nzg - 2;
    const int nch = nz >= 8 * 16 ? 8 * std::max(1, (nz + 8 * 12) / (8 * 24))  // multiple of 8 nearest nz / 24
                                 : (nz + 15) / 16;
    return std::max(1, (nz + nch - 1) / nch);
}

template <int MODE, int FOUT, class T>
void march(Handle &h, const Fld3<T> &cur, const Fld3<T> *old, const Fld3<T> *f1, Fld3<T> *nxt, Fld3<T> *fo,
           double c0, double c1, double dt, double g2 = 0.0) {
    const Dims &d = h.d;
    const int chunk = march_chunk(d);
    const dim3 grid = march_grid<T>(d, chunk), block(256);
    const T *o[3] = {old ? old->u : nullptr, old ? old->v : nullptr, old ? old->w : nullptr};
    const T *f[3] = {f1 ? f1->u : nullptr, f1 ? f1->v : nullptr, f1 ? f1->w : nullptr};
    T *n[3] = {nxt ? nxt->u : nullptr, nxt ? nxt->v : nullptr, nxt ? nxt->w : nullptr};
    T *r[3] = {fo ? fo->u : nullptr, fo ? fo->v : nullptr, fo ? fo->w : nullptr};
    const MetricsT<T> &m = metrics<T>(h);
    if (h.les != 0)
        hipLaunchKernelGGL((k_rk_march<MODE, FOUT, true, T>), grid, block, 0, h.stream, d, m, (T)h.nu,
                           (T)h.dPdx, (T)h.dPdz, chunk, cur.u, cur.v, cur.w, reinterpret_cast<const T *>(h.nu_t),
                           o[0], o[1], o[2], f[0], f[1], f[2], n[0], n[1], n[2], r[0], r[1], r[2], (T)c0, (T)c1,
                           (T)dt, (T)g2);
    else
        hipLaunchKernelGGL((k_rk_march<MODE, FOUT, false, T>), grid, block, 0, h.stream, d, m, (T)h.nu,
                           (T)h.dPdx, (T)h.dPdz, chunk, cur.u, cur.v, cur.w, nullptr, o[0], o[1], o[2], f[0],
                           f[1], f[2], n[0], n[1], n[2], r[0], r[1], r[2], (T)c0, (T)c1, (T)dt, (T)g2);
    check_launch();
}

template <class T>
void launch_rk_stage(Handle &h, const Fld3<T> &cur, const Fld3<T> &old, Fld3<T> &nxt, int stage, double dt) {
    const bool rk3 = h.scheme == 3;
    Fld3<T> F1 = as<T>(h.F1);
    ProfScope ps(h, KID_RK_STAGE);
    // FP64 RK3 carries G = Uo + dt a21 F1 (= Uo + dt a31 F1) from the first
    // stage in the F1 buffers; FP32 fields carry Uo and F1, since storing G in
    // FP32 would round a velocity-sized value once more per stage
    const bool gform = rk3 && sizeof(T) == 8;
    if (stage == 1) {
        const double c0 = dt * (rk3 ? h.rk_coef[1][1] : h.rk2_coef[1][1]);
        if (gform)
            march<0, 2>(h, cur, &old, (const Fld3<T> *)nullptr, &nxt, &F1, c0, 0.0, dt, dt * h.rk_coef[2][1]);
        else if (rk3 || sizeof(T) == 4)
            // FP32 RK2 keeps F1 too: it is the first stage's pressure source
            march<0, 1>(h, cur, &old, (const Fld3<T> *)nullptr, &nxt, &F1, c0, 0.0, dt);
        else
            march<0, 0>(h, cur, &old, (const Fld3<T> *)nullptr, &nxt, (Fld3<T> *)nullptr, c0, 0.0, dt);
    } else if (!rk3) {
        // RK2 second stage: Uo + dt*(0*F1 + 1*F2) == Uo + dt*F2 exactly
        march<0, 0>(h, cur, &old, (const Fld3<T> *)nullptr, &nxt, (Fld3<T> *)nullptr, dt * h.rk2_coef[2][2],
                    0.0, dt);
    } else if (gform) {
        // RK3 stages 2, 3: new = G + dt a_ss F (a_32 = 0 so F2 is never needed)
        march<0, 0>(h, cur, &F1, (const Fld3<T> *)nullptr, &nxt, (Fld3<T> *)nullptr, dt * h.rk_coef[stage][stage],
                    0.0, dt);
    } else {
        // RK3 stages 2, 3: a_32 = 0 so F2 is never needed
        const double c1 = h.rk_coef[stage][1], c0 = h.rk_coef[stage][stage];
        march<1, 0>(h, cur, &old, &F1, &nxt, (Fld3<T> *)nullptr, c0, c1, dt);
    }
}

template <class T>
void launch_rhs_only(Handle &h, const Fld3<T> &cur, Fld3<T> &rhs) {
    march<2, 0>(h, cur, (const Fld3<T> *)nullptr, (const Fld3<T> *)nullptr, (Fld3<T> *)nullptr, &rhs, 0.0, 0.0,
                0.0);
}

template <class T>
void launch_bc(Handle &h, Fld3<T> &f, const Fld3<T> &old, int rk_step, double dt, double t, double *inc1,
               double t_old) {
    const Dims &d = h.d;
    const double coef = h.rk_t[rk_step] * dt * h.bc_uc;
    prof_begin(h, KID_BC_INOUT);
    const dim3 g_in((d.nyg + 255) / 256, d.nzg), g_tw((d.nxg + 255) / 256, d.nzg);
    const int nq_in = g_in.x * g_in.y, nq_tw = g_tw.x * g_tw.y;
    double *qpart = h.red;
    hipLaunchKernelGGL(k_bc_inout<T>, g_in, dim3(256), 0, h.stream, d, h.m, h.bc, f.u, f.v, f.w, old.u, old.v,
                       old.w, coef, t, qpart, inc1, t_old);
    check_launch();
    prof_end(h, KID_BC_INOUT);
    prof_begin(h, KID_BC_PERIODIC);
    hipLaunchKernelGGL(k_bc_periodic<T>, dim3((d.nxg + 255) / 256, d.nyg), dim3(256), 0, h.stream, d, f.u, f.v,
                       f.w);
    check_launch();
    prof_end(h, KID_BC_PERIODIC);
    prof_begin(h, KID_BC_TOPWALL);
    hipLaunchKernelGGL(k_bc_topwall<T>, g_tw, dim3(256), 0, h.stream, d, h.m, h.bc, f.u, f.v, f.w, old.v,
                       qpart + nq_in);
    check_launch();
    prof_end(h, KID_BC_TOPWALL);
    prof_begin(h, KID_BC_MASSFLUX);
    hipLaunchKernelGGL(k_bc_massflux<T>, g_in, dim3(256), 0, h.stream, d, h.length_y, qpart, nq_in + nq_tw, f.u,
                       inc1 ? inc1 + 2 * (long long)d.nyg * d.nzg : nullptr);
    check_launch();
    prof_end(h, KID_BC_MASSFLUX);
    // Falkner-Skan top: U_top is the Falkner-Skan profile from the first call
    // on, and with it the outflow's convective velocity U_top(nx)
    if (h.fs_uc_pending) {
        h.bc_uc = h.fs_uc;
        h.fs_uc_pending = false;
    }
}

void launch_poisson(Handle &h, double *phi, const Field3 *div) {
    const Dims &d = h.d;
    const int N = d.N, M = d.M, NM = d.NM;
    const long long plane = (long long)N * NM;
    // forward x: lines (j, m) of the padded field -> T1[m][j][kx]
    DctArgs da{};
    da.N = N;
    da.N1 = h.dct_n1;
    da.N2 = N / h.dct_n1;
    da.K1h = h.dct_n1 / 2;
    da.C1 = NM;
    da.C = NM * M;
    da.w1 = reinterpret_cast<const double2 *>(h.dct_w1);
    da.w2 = reinterpret_cast<const double2 *>(h.dct_w2);
    da.wN = reinterpret_cast<const double2 *>(h.dct_wN);
    da.rot = reinterpret_cast<const double2 *>(h.dct_rot);
    if (h.dct_mfma) {
        da.fA = h.dct_mf + h.dct_mf_off[0];
        da.fC = h.dct_mf + h.dct_mf_off[1];
        da.iC = h.dct_mf + h.dct_mf_off[2];
        da.iA = h.dct_mf + h.dct_mf_off[3];
    }
    da.in_base = fidx(d, 1, 1, 1);
    da.in_c1 = d.sx;
    da.in_c2 = d.sxy;
    da.out_base = 0;
    da.out_c1 = N;
    da.out_c2 = plane;
    da.scale = 1.0;
    dct_launch(h, KID_XFORM_X_FWD, da, true, phi, h.T1, div);
    // forward z: columns (kx, j) of T1 -> S[q][j][kx]
    zdft_launch(h, KID_XFORM_Z_FWD, true, h.T1, h.S);
    {
    ProfScope ps(h, KID_THOMAS);
    // the partitioned kernel leaves the (0,0) column of S to its block 0,
    // which solves it in place; the table kernel (no length limit) solves it
    // with the other systems
    if (h.th_p != 0) {
        launch_thomas_seg<double>(h, h.S, N, (long long)plane, solve00_args(h, h.S, N, 1.0, nullptr, 0));
    } else {
        hipLaunchKernelGGL(k_thomas<8>, dim3((N + 255) / 256 * M), dim3(256), 0, h.stream, d, h.lu_dl,
                           h.lu_dp, h.lu_du, h.S);
        check_launch();
    }
    }
    // inverse z: S -> T1
    zdft_launch(h, KID_XFORM_Z_INV, false, h.S, h.T1);
    // inverse x: T1[m][j][kx] -> phi cells (1..N, 1..NM, 1..M)
    da.in_base = 0;
    da.in_c1 = N;
    da.in_c2 = plane;
    da.out_base = fidx(d, 1, 1, 1);
    da.out_c1 = d.sx;
    da.out_c2 = d.sxy;
    da.scale = 1.0 / ((double)N * (double)M);
    dct_launch(h, KID_XFORM_X_INV, da, false, h.T1, phi);
}

template <class T>
bool launch_thomas_seg(Handle &h, T *S, int ldj, long long ldq, const Solve00 &z) {
    const Dims &d = h.d;
    if (h.th_p == 0) return false;
    // one resident block per CU (the registers of 8 waves), each looping over
    // its tiles, plus block 0 for the (0,0) system
    const int ntiles = (d.N + 63) / 64 * d.M;
    const int nblk = std::max(1, std::min(ntiles, h.n_cu > 1 ? h.n_cu - 1 : ntiles));
    hipLaunchKernelGGL(k_thomas_seg<T>, dim3(1 + nblk), dim3(64 * h.th_p), 0, h.stream, d.N, d.NM, h.th_p, d.M,
                       ldj, ldq, (const double *)h.th_co, z.part ? 1 : 0, S, z);
    check_launch();
    return true;
}

Solve00 solve00_args(const Handle &h, double *x, int ldx, double scale, const double *part, int npart) {
    if (h.d.NM > kMax00) {
        last_error() = "internal: the (0,0) system is longer than the solve kernel holds";
        throw ArgError();
    }
    Solve00 z;
    z.L00 = h.lu00;
    z.D00 = h.lu00 + h.d.NM;
    z.du = h.lu_du;
    z.part = part;
    z.npart = npart;
    z.x = x;
    z.ldx = ldx;
    z.scale = scale;
    return z;
}

void launch_solve00(Handle &h, const Solve00 &z) {
    hipLaunchKernelGGL(k_solve00, dim3(1), dim3(256), 0, h.stream, h.d.NM, z);
    check_launch();
}

template <class T>
void launch_project(Handle &h, Fld3<T> &f, const T *phi, T *P, double dt) {
    const Dims &d = h.d;
    ProfScope ps(h, KID_PROJECT);
    const dim3 grid(xcd_blocks((d.nxg - 1 + kWave - 1) / kWave * ((d.nyg - 2 + 3) / 4) *
                               ((d.nzg - 2 + kProjZ - 1) / kProjZ)));
    const double den = dt * h.rk2_coef[1][1];
    hipLaunchKernelGGL(k_project<T>, grid, dim3(kWave, 4), 0, h.stream, d, metrics<T>(h), f.u, f.v, f.w, phi, P,
                       den);
    check_launch();
}

long long maxdiv_partials(int nxg, int nyg, int nzg) {
    Dims d{};
    d.nxg = nxg;
    d.nyg = nyg;
    d.nzg = nzg;
    const dim3 g = reduce_grid(d);
    return (long long)g.x * g.y * g.z;
}

template <class T>
void launch_maxdiv(hipStream_t stream, const DivArgs &a, const T *U, const T *V, const T *W, double *part,
                   double *out) {
    Dims d{};
    d.nxg = a.nxg;
    d.nyg = a.nyg;
    d.nzg = a.nzg;
    const dim3 g = reduce_grid(d);
    const long long nb = (long long)g.x * g.y * g.z;
    hipLaunchKernelGGL(k_maxdiv<T>, g, dim3(kWave, 4), 0, stream, a, U, V, W, part);
    check_launch();
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, stream, part, nb, 1, out);
    check_launch();
}
template void launch_maxdiv<double>(hipStream_t, const DivArgs &, const double *, const double *, const double *,
                                    double *, double *);

template <class T>
double max_divergence(Handle &h, const Fld3<T> &f) {
    const Dims &d = h.d;
    DivArgs a{};
    a.nxg = d.nxg;
    a.nyg = d.nyg;
    a.nzg = d.nzg;
    a.u1 = a.v1 = a.w1 = d.sx;
    a.u2 = a.v2 = a.w2 = d.sxy;
    a.rdx = h.m.rdx;
    a.rdy = h.m.rdy;
    a.rdz = h.m.rdz;
    launch_maxdiv(h.stream, a, f.u, f.v, f.w, h.red, h.scal);
    double v = 0.0;
    BL_HIP(hipMemcpyAsync(&v, h.scal, sizeof(double), hipMemcpyDeviceToHost, h.stream));
    BL_HIP(hipStreamSynchronize(h.stream));
    return v;
}

template <class T>
double dt_convective(Handle &h, const Fld3<T> &f) {
    const Dims &d = h.d;
    ProfScope ps(h, KID_REDUCE);
    const dim3 g = reduce_grid(d);
    const long long nb = (long long)g.x * g.y * g.z;
    hipLaunchKernelGGL(k_dtconv<T>, g, dim3(kWave, 4), 0, h.stream, d, h.m, f.u, f.v, f.w, h.red);
    check_launch();
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, h.stream, h.red, nb, 3, h.scal);
    check_launch();
    double v[3];
    BL_HIP(hipMemcpyAsync(v, h.scal, 3 * sizeof(double), hipMemcpyDeviceToHost, h.stream));
    BL_HIP(hipStreamSynchronize(h.stream));
    const double CFLa = std::fabs(h.cfl);
    const double dcu = CFLa * v[0], dcv = CFLa * v[1], dcw = CFLa * v[2];
    return std::fmin(dcu, std::fmin(dcv, dcw));
}

template <class T>
void wall_stats(Handle &h, const Fld3<T> &f, double *cf_host) {
    const Dims &d = h.d;
    hipLaunchKernelGGL(k_wallstats<T>, dim3((d.nx + 255) / 256), dim3(256), 0, h.stream, d, h.m, h.nu, f.u, f.v,
                       reinterpret_cast<const T *>(h.nu_t), h.red);
    check_launch();
    BL_HIP(hipMemcpyAsync(cf_host, h.red, sizeof(double) * d.nx, hipMemcpyDeviceToHost, h.stream));
    BL_HIP(hipStreamSynchronize(h.stream));
}

template <class T>
void monitor(Handle &h, const Fld3<T> &f, double *out) {
    const Dims &d = h.d;
    const int np = d.nzg - 2;
    hipLaunchKernelGGL(k_monitor<T>, dim3(np), dim3(256), 0, h.stream, d, f.u, f.v, f.w,
                       reinterpret_cast<const T *>(h.nu_t), h.red);
    check_launch();
    std::vector<double> part((size_t)np * 8);
    BL_HIP(hipMemcpyAsync(part.data(), h.red, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h.stream));
    BL_HIP(hipStreamSynchronize(h.stream));
    for (int q = 0; q < 4; ++q) {
        out[q] = 0.0;
        out[4 + q] = -INFINITY;
    }
    for (int k = 0; k < np; ++k)
        for (int q = 0; q < 4; ++q) {
            out[q] += part[(size_t)k * 8 + q];
            out[4 + q] = std::fmax(out[4 + q], part[(size_t)k * 8 + 4 + q]);
        }
}

#define BL_INSTANTIATE(T)                                                                                   \
    template void launch_rk_stage<T>(Handle &, const Fld3<T> &, const Fld3<T> &, Fld3<T> &, int, double); \
    template void launch_rhs_only<T>(Handle &, const Fld3<T> &, Fld3<T> &);                                \
    template void launch_bc<T>(Handle &, Fld3<T> &, const Fld3<T> &, int, double, double, double *, double); \
    template void launch_project<T>(Handle &, Fld3<T> &, const T *, T *, double);                         \
    template bool launch_thomas_seg<T>(Handle &, T *, int, long long, const Solve00 &);                   \
    template double max_divergence<T>(Handle &, const Fld3<T> &);                                         \
    template double dt_convective<T>(Handle &, const Fld3<T> &);                                          \
    template void wall_stats<T>(Handle &, const Fld3<T> &, double *);                                     \
    template void monitor<T>(Handle &, const Fld3<T> &, double *);
BL_INSTANTIATE(double)
BL_INSTANTIATE(float)
#undef BL_INSTANTIATE

}  // namespace blgpu
