// Lund rescaled inflow (inflow_flag 3 and 5) on the device:
// compute_rescaled_inflow (rescaled_inlet_bc.f90:60-300) once per step from
// the step-start field, as apply_inflow_bc_x_rescaling
// (boundary_conditions.f90:569-597) and apply_inflow_bc_x_Turbulent_rescaled_flu
// (:512-559) call it at the first boundary call of a step.
//
//   k_lund_sums     z means over planes 1..nzg-2 (0-based) of U and V at the
//                   rescaling plane and at the inlet, and <U V> at the wall
//                   row of the rescaling plane (one block per y row, fixed
//                   summation order)
//   k_lund_profile  one block: the time-averaged means (state), U_inf,
//                   delta99 and theta of both planes, u_tau, gamma, the inner
//                   (y+) and outer (y/delta) coordinates of every row, the
//                   interpolation rows and weights of interp_vel (:305-345),
//                   the interpolated mean profiles and the blending weights
//   k_lund_plane    the inflow plane: rescaled fluctuations (and means, for
//                   flag 3) blended with the weights, into FP64 arrays that
//                   every boundary call of the step copies into its buffer
//
// Row 0 of each plane (the wall ghost, which interp_vel leaves unset and the
// wall condition overwrites) is 0.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/blgpu.h"
#include "bl_internal.hpp"

namespace blgpu {
namespace {

void check() { BL_HIP(hipGetLastError()); }

// deterministic block sum of 256 values (thread 0 gets the total)
__device__ __forceinline__ double lund_sum_256(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

template <class T>
__global__ __launch_bounds__(256) void k_lund_sums(Dims d, int ir, const T *__restrict__ U, const T *__restrict__ V,
                                                  LundArrays L) {
    __shared__ double sh[256];
    const int j = blockIdx.x;  // 0 .. nyg-1
    double a = 0.0, b = 0.0, c = 0.0, e = 0.0;
    for (int k = 1 + (int)threadIdx.x; k <= d.nzg - 2; k += 256) {
        a += (double)U[fidx(d, ir, j, k)];
        b += (double)U[fidx(d, 0, j, k)];
        if (j < d.ny) {
            c += (double)V[fidx(d, ir, j, k)];
            e += (double)V[fidx(d, 0, j, k)];
        }
    }
    const double n = (double)(d.nzg - 2);
    a = lund_sum_256(a, sh);
    b = lund_sum_256(b, sh);
    c = lund_sum_256(c, sh);
    e = lund_sum_256(e, sh);
    if (threadIdx.x == 0) {
        L.umr[j] = a / n;
        L.umi[j] = b / n;
        if (j < d.ny) {
            L.vmr[j] = c / n;
            L.vmi[j] = e / n;
        }
    }
    if (j == 0) {
        // <0.5 (U(ir, 0) + U(ir, 1)) V(ir, 0)> over the same planes (:196-199)
        double w = 0.0;
        for (int k = 1 + (int)threadIdx.x; k <= d.nzg - 2; k += 256)
            w += 0.5 * ((double)U[fidx(d, ir, 0, k)] + (double)U[fidx(d, ir, 1, k)]) * (double)V[fidx(d, ir, 0, k)];
        w = lund_sum_256(w, sh);
        if (threadIdx.x == 0) L.scal[0] = w / n;
    }
}

// first row (0-based) at or above `level`, nyg/2 - 1 if it is row 0 or none,
// and the interpolated height (rescaled_inlet_bc.f90:144-175)
__device__ double lund_height99(int n, const double *m, const double *yg, double level) {
    int jref = -1;
    for (int j = 0; j < n; ++j)
        if (m[j] >= level) {
            jref = j;
            break;
        }
    if (jref < 1) jref = n / 2 - 1;
    const double w1 = (level - m[jref - 1]) / (m[jref] - m[jref - 1]);
    const double w0 = 1.0 - w1;
    const double h = w1 * yg[jref] + w0 * yg[jref - 1];
    return h < 0 ? fabs(h) : h;
}

__device__ double lund_theta(int n, const double *m, const double *yg, double uinf) {
    double th = 0.0;
    for (int j = 1; j < n; ++j) {
        const double a = m[j] / uinf * (1.0 - m[j] / uinf), b = m[j - 1] / uinf * (1.0 - m[j - 1] / uinf);
        th = th + 0.5 * (a + b) * (yg[j] - yg[j - 1]);
    }
    return th < 0 ? fabs(th) : th;
}

__device__ __forceinline__ double lund_interp(const double *f, int n, int j, double coef) {
    return j >= 1 ? coef * f[j - 1] + (1.0 - coef) * f[j] : f[n - 1];
}

__global__ __launch_bounds__(256) void k_lund_profile(Dims d, Metrics m, double nu, double dt, double t_resc,
                                                     double delta_inlet, LundArrays L) {
    __shared__ double s_utau[3];
    const int nyg = d.nyg, ny = d.ny;
    const double a = dt / t_resc, b1 = 1.0 - dt / t_resc;
    for (int j = threadIdx.x; j < nyg; j += 256) {
        L.umrT[j] = a * L.umr[j] + b1 * L.umrT[j];
        L.umiT[j] = a * L.umi[j] + b1 * L.umiT[j];
    }
    for (int j = threadIdx.x; j < ny; j += 256) {
        L.vmrT[j] = a * L.vmr[j] + b1 * L.vmrT[j];
        L.vmiT[j] = a * L.vmi[j] + b1 * L.vmiT[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double uinf = L.umr[nyg - 1];
        const double delta_resc = lund_height99(nyg, L.umrT, m.yg, 0.99 * uinf);
        const double th_r = lund_theta(nyg, L.umrT, m.yg, uinf), th_i = lund_theta(nyg, L.umiT, m.yg, uinf);
        const double dU = (L.umrT[1] - L.umrT[0]) / (m.yg[1] - m.yg[0]);
        const double utau_r = pow(fabs(nu * dU - L.scal[0]), 0.5);
        const double utau_i = utau_r * pow(th_r / th_i, 1.0 / 8.0);
        s_utau[0] = utau_r;
        s_utau[1] = utau_i;
        s_utau[2] = delta_resc;
        L.scal[1] = utau_i / utau_r;  // gamma
        L.scal[2] = uinf;
    }
    __syncthreads();
    const double utau_r = s_utau[0], utau_i = s_utau[1], delta_resc = s_utau[2], gamma = L.scal[1],
                 uinf = L.scal[2];
    const double alpha = 4.0, bb = 0.2;
    // cell rows (U, W): inner y+ and outer yg/delta coordinates
    for (int j = threadIdx.x; j < nyg; j += 256) {
        const double ti = utau_i * m.yg[j] / nu, to = m.yg[j] / delta_inlet;
        L.wg[j] = 0.5 * (1.0 + tanh(alpha * (to - bb) / ((1.0 - 2.0 * bb) * to + bb)) / tanh(alpha));
        // interpolation points of the rescaling plane: utau_r yg / nu and yg / delta_resc
        int ji, jo;
        double ci, co;
        {
            // interp_vel's row (rescaled_inlet_bc.f90:305-345): the first k
            // whose coordinate exceeds the target, its lower neighbour's weight
            ji = -1;
            for (int k = 0; k < nyg; ++k)
                if (utau_r * m.yg[k] / nu > ti) {
                    ji = k;
                    break;
                }
            ci = ji >= 1 ? (utau_r * m.yg[ji] / nu - ti) / (utau_r * m.yg[ji] / nu - utau_r * m.yg[ji - 1] / nu) : 0.0;
            if (ji == 0) ji = -1;
            jo = -1;
            for (int k = 0; k < nyg; ++k)
                if (m.yg[k] / delta_resc > to) {
                    jo = k;
                    break;
                }
            co = jo >= 1 ? (m.yg[jo] / delta_resc - to) / (m.yg[jo] / delta_resc - m.yg[jo - 1] / delta_resc) : 0.0;
            if (jo == 0) jo = -1;
        }
        L.jgi[j] = ji;
        L.cgi[j] = ci;
        L.jgo[j] = jo;
        L.cgo[j] = co;
        const double mi = j == 0 ? 0.0 : lund_interp(L.umrT, nyg, ji, ci);
        const double mo = j == 0 ? 0.0 : lund_interp(L.umrT, nyg, jo, co);
        L.ugi[j] = gamma * mi;
        L.ugo[j] = gamma * mo + (1.0 - gamma) * uinf;
    }
    // face rows (V)
    for (int j = threadIdx.x; j < ny; j += 256) {
        const double ti = utau_i * m.y[j] / nu, to = m.y[j] / delta_inlet;
        L.wf[j] = 0.5 * (1.0 + tanh(alpha * (to - bb) / ((1.0 - 2.0 * bb) * to + bb)) / tanh(alpha));
        int ji = -1, jo = -1;
        for (int k = 0; k < ny; ++k)
            if (utau_r * m.y[k] / nu > ti) {
                ji = k;
                break;
            }
        const double ci =
            ji >= 1 ? (utau_r * m.y[ji] / nu - ti) / (utau_r * m.y[ji] / nu - utau_r * m.y[ji - 1] / nu) : 0.0;
        if (ji == 0) ji = -1;
        for (int k = 0; k < ny; ++k)
            if (m.y[k] / delta_resc > to) {
                jo = k;
                break;
            }
        const double co = jo >= 1 ? (m.y[jo] / delta_resc - to) / (m.y[jo] / delta_resc - m.y[jo - 1] / delta_resc)
                                  : 0.0;
        if (jo == 0) jo = -1;
        L.jfi[j] = ji;
        L.cfi[j] = ci;
        L.jfo[j] = jo;
        L.cfo[j] = co;
        L.vfi[j] = j == 0 ? 0.0 : lund_interp(L.vmrT, ny, ji, ci);
        L.vfo[j] = j == 0 ? 0.0 : lund_interp(L.vmrT, ny, jo, co);
    }
}

// the inflow plane, one thread per (j, k) of the largest extent
template <class T>
__global__ __launch_bounds__(256) void k_lund_plane(Dims d, int ir, int fluc_only, const T *__restrict__ U,
                                                   const T *__restrict__ V, const T *__restrict__ W, LundArrays L) {
    const int j = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
    const double gamma = L.scal[1];
    const int nyg = d.nyg, ny = d.ny;
    // fluctuation of the rescaling plane at row r: U - <U>_T, V - <V>_T, W
    auto fu = [&](int r) { return (double)U[fidx(d, ir, r, k)] - L.umrT[r]; };
    auto fv = [&](int r) { return (double)V[fidx(d, ir, r, k)] - L.vmrT[r]; };
    auto fw = [&](int r) { return (double)W[fidx(d, ir, r, k)]; };
    auto itp = [&](auto f, int n, int jj, double c) { return jj >= 1 ? c * f(jj - 1) + (1.0 - c) * f(jj) : f(n - 1); };
    if (j < nyg && k < d.nzg) {
        double v = 0.0;
        if (j > 0) {
            const double ui = gamma * itp(fu, nyg, L.jgi[j], L.cgi[j]), uo = gamma * itp(fu, nyg, L.jgo[j], L.cgo[j]);
            const double mi = fluc_only ? 0.0 : L.ugi[j], mo = fluc_only ? 0.0 : L.ugo[j];
            v = (mi + ui) * (1.0 - L.wg[j]) + (mo + uo) * L.wg[j];
        }
        L.pu[j + (long long)nyg * k] = v;
    }
    if (j < nyg && k < d.nz) {
        double v = 0.0;
        if (j > 0) {
            const double wi = gamma * itp(fw, nyg, L.jgi[j], L.cgi[j]), wo = gamma * itp(fw, nyg, L.jgo[j], L.cgo[j]);
            v = wi * (1.0 - L.wg[j]) + wo * L.wg[j];
        }
        L.pw[j + (long long)nyg * k] = v;
    }
    if (j < ny && k < d.nzg) {
        double v = 0.0;
        if (j > 0) {
            const double vi = gamma * itp(fv, ny, L.jfi[j], L.cfi[j]), vo = gamma * itp(fv, ny, L.jfo[j], L.cfo[j]);
            const double mi = fluc_only ? 0.0 : L.vfi[j], mo = fluc_only ? 0.0 : L.vfo[j];
            v = (mi + vi) * (1.0 - L.wf[j]) + (mo + vo) * L.wf[j];
        }
        L.pv[j + (long long)ny * k] = v;
    }
}

}  // namespace

size_t lund_doubles(const Dims &d) {
    const size_t g = d.nyg, f = d.ny;
    // sums (4), T means (4), scalars, per-row tables (cell: 2 coefs + 2 means + weight;
    // face: the same) and rows as doubles (4 per grid), planes
    return 2 * g + 2 * f + 2 * g + 2 * f + 8 + 5 * g + 5 * f + 2 * g + 2 * f + g * d.nzg + f * d.nzg + g * d.nz;
}

LundArrays lund_layout(const Dims &d, double *base) {
    LundArrays L{};
    double *p = base;
    auto take = [&](size_t n) {
        double *q = p;
        p += n;
        return q;
    };
    const size_t g = d.nyg, f = d.ny;
    L.umr = take(g);
    L.umi = take(g);
    L.vmr = take(f);
    L.vmi = take(f);
    L.umrT = take(g);
    L.umiT = take(g);
    L.vmrT = take(f);
    L.vmiT = take(f);
    L.scal = take(8);
    L.cgi = take(g);
    L.cgo = take(g);
    L.ugi = take(g);
    L.ugo = take(g);
    L.wg = take(g);
    L.cfi = take(f);
    L.cfo = take(f);
    L.vfi = take(f);
    L.vfo = take(f);
    L.wf = take(f);
    L.jgi = reinterpret_cast<int *>(take(g));
    L.jgo = reinterpret_cast<int *>(take(g));
    L.jfi = reinterpret_cast<int *>(take(f));
    L.jfo = reinterpret_cast<int *>(take(f));
    L.pu = take(g * d.nzg);
    L.pv = take(f * d.nzg);
    L.pw = take(g * d.nz);
    return L;
}

template <class T>
void launch_lund_sums(Handle &h, const Fld3<T> &f) {
    hipLaunchKernelGGL(k_lund_sums<T>, dim3(h.d.nyg), dim3(256), 0, h.stream, h.d, h.lund_ir, f.u, f.v, h.lund);
    check();
}

template <class T>
void launch_lund(Handle &h, const Fld3<T> &old, double dt) {
    const Dims &d = h.d;
    launch_lund_sums<T>(h, old);
    hipLaunchKernelGGL(k_lund_profile, dim3(1), dim3(256), 0, h.stream, d, h.m, h.nu, dt, h.lund_t_resc,
                       h.lund_delta_inlet, h.lund);
    check();
    const dim3 g((d.nyg + 255) / 256, d.nzg);
    hipLaunchKernelGGL(k_lund_plane<T>, g, dim3(256), 0, h.stream, d, h.lund_ir, h.bc.inflow_flag == 5 ? 1 : 0,
                       old.u, old.v, old.w, h.lund);
    check();
}

template <class T>
void lund_init_means(Handle &h, const Fld3<T> &f) {
    // the first-start means of the field (initialization.f90:640-660)
    launch_lund_sums<T>(h, f);
    const Dims &d = h.d;
    BL_HIP(hipMemcpyAsync(h.lund.umrT, h.lund.umr, sizeof(double) * d.nyg, hipMemcpyDeviceToDevice, h.stream));
    BL_HIP(hipMemcpyAsync(h.lund.umiT, h.lund.umi, sizeof(double) * d.nyg, hipMemcpyDeviceToDevice, h.stream));
    BL_HIP(hipMemcpyAsync(h.lund.vmrT, h.lund.vmr, sizeof(double) * d.ny, hipMemcpyDeviceToDevice, h.stream));
    BL_HIP(hipMemcpyAsync(h.lund.vmiT, h.lund.vmi, sizeof(double) * d.ny, hipMemcpyDeviceToDevice, h.stream));
}

template void launch_lund<double>(Handle &, const Fld3<double> &, double);
template void launch_lund<float>(Handle &, const Fld3<float> &, double);
template void lund_init_means<double>(Handle &, const Fld3<double> &);
template void lund_init_means<float>(Handle &, const Fld3<float> &);

}  // namespace blgpu
