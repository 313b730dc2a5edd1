// Wall models on the device: compute_wall_model (wallmodel.f90:25-144) once
// per RK stage after the eddy viscosity (time_integration.f90:341-346), and
// the wall mean of the pseudo-Robin V (apply_pseudo_Robin_bc_y_bottom,
// boundary_conditions.f90:1152-1180) that the boundary kernel subtracts.
//
//   k_wm_cols   one block per x column: the z sums each model needs (the U
//               row 3 mean of the law of the wall, the resolved tau_uv and
//               (nu + nu_t)/tau_w at the wall edge, the wall-normal gradients
//               and their products at both walls, the rows of nu_t that the
//               constant-coefficient Smagorinsky averages, the step-start dV/dy
//               at the wall), fixed summation order
//   k_wm_final  one block: utau_model (Newton per column), the slip lengths
//               alpha_x(nx), alpha_y(nxg), alpha_z(nxg), and beta_y
//
// Every model fills alpha(i, :, :) with one value per x column, so the slip
// lengths are x vectors.  Values that depend only on t and constants (model 5
// and the log-layer factor of models 6 and 8) and the x tables (utau_ref,
// the Cf correlation's tau_w) are evaluated on the host with the same libm
// calls the reference makes, then used by the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "../../include/blgpu.h"
#include "bl_internal.hpp"

namespace blgpu {
namespace {

void check() { BL_HIP(hipGetLastError()); }

// per-column sums, rows of cols[q * nxg + i]
enum WmQ { Q_U3 = 0, Q_TAU, Q_NUTW, Q_DVO, Q_SB, Q_ST, Q_DUV, Q_DU, Q_N2, Q_NN, Q_UW, Q_UR, Q_DUVB, Q_DUB, Q_H, kWmQ };

struct WmArgs {
    int model, istress, les, nut_dirichlet;
    int nut_flag;       // models 9, 12: iwall_model_nut (0 fixed, 2 dynamic viscous fraction)
    double frac_vis;    // models 9, 12: frac_vis_wall_model
    int beta_only;      // only the pseudo-Robin wall mean (slip lengths kept)
    double nu, dPdx, Ly;
    double a5x, a5y, a5z;  // model 5 at the current t
    double logf, utau_p;   // models 6, 8: log(Delta utau/nu e^(B kappa)) and dPdx^0.5
    double *ax, *ay, *az, *utau, *beta;
    const double *utau_ref, *tauw;
    double *cols;
};

__device__ __forceinline__ double wm_sum_256(double v, double *sh) {
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

// the models that use the Robin / pseudo-Robin wall (boundary_conditions.f90:80-108)
__host__ __device__ __forceinline__ bool wm_robin(int model) { return model != 0 && model != 11 && model != 15; }

// nu_t at the wall edge (x(i), y(1)) of plane k: interpolate_y (weighted,
// centres to faces) then interpolate_x (interpolation.f90:24-85)
template <class T>
__device__ __forceinline__ double nut_edge(const Dims &d, const Metrics &m, const T *nt, int i, int k) {
    const double a = m.wy0[0] * (double)nt[fidx(d, i, 0, k)] + m.wy1[0] * (double)nt[fidx(d, i, 1, k)];
    const double b = m.wy0[0] * (double)nt[fidx(d, i + 1, 0, k)] + m.wy1[0] * (double)nt[fidx(d, i + 1, 1, k)];
    return 0.5 * (a + b);
}

// (dU/dy + dV/dx)/2 at the cell edge (x(i), y(j)) of plane k
template <class T>
__device__ __forceinline__ double suv(const Dims &d, const Metrics &m, const T *U, const T *V, int i, int j, int k) {
    return 0.5 * (((double)U[fidx(d, i, j + 1, k)] - (double)U[fidx(d, i, j, k)]) / (m.yg[j + 1] - m.yg[j]) +
                  ((double)V[fidx(d, i + 1, j, k)] - (double)V[fidx(d, i, j, k)]) / (m.xg[i + 1] - m.xg[i]));
}

template <class T>
__global__ __launch_bounds__(256) void k_wm_cols(Dims d, Metrics m, WmArgs a, const T *__restrict__ U,
                                                 const T *__restrict__ V, const T *__restrict__ Vo,
                                                 const T *__restrict__ nt) {
    __shared__ double sh[256];
    const int i = blockIdx.x;  // 0 .. nxg-1
    const int nx = d.nx, nxg = d.nxg, ny = d.ny, nyg = d.nyg, nzg = d.nzg;
    double s[kWmQ];
#pragma unroll
    for (int q = 0; q < kWmQ; ++q) s[q] = 0.0;
    const bool robin = wm_robin(a.model);
    if (robin && i >= 1 && i <= nxg - 3) {
        // step-start dV at the wall over planes 1 .. nzg-3 (reference 2 .. nzg-2)
        for (int k = 1 + (int)threadIdx.x; k <= nzg - 3; k += 256)
            s[Q_DVO] += (double)Vo[fidx(d, i, 1, k)] - (double)Vo[fidx(d, i, 0, k)];
    }
    if (!a.beta_only) {
        for (int k = 1 + (int)threadIdx.x; k <= nzg - 2; k += 256) {
            if (a.istress == 1 && i >= 1 && i <= nx - 2) s[Q_U3] += (double)U[fidx(d, i, 2, k)];
            if (a.model == 11 && i <= nx - 1) {
                const double ne = nt ? nut_edge(d, m, nt, i, k) : 0.0;
                s[Q_TAU] += -2.0 * ne * suv(d, m, U, V, i, 0, k);
            }
            if ((a.model == 9 || a.model == 12) && i <= nx - 1) {
                // compute_alpha_v_from_Cf (:1532-1697) / compute_alpha_u_from_Cf
                // (:1263-1424): U over the two wall rows (and rows 3:4 for the
                // dynamic viscous fraction), tau_uv at the wall edge (model 9
                // with dU/dy alone), <dU/dy dV/dy> (12; model 9: <U dVo/dy>
                // of the step-start V) and <dU/dy>
                const double u0 = (double)U[fidx(d, i, 0, k)], u1 = (double)U[fidx(d, i, 1, k)];
                s[Q_UW] += u0 + u1;
                if (a.nut_flag == 2) s[Q_UR] += (double)U[fidx(d, i, 2, k)] + (double)U[fidx(d, i, 3, k)];
                const double ne = nt ? nut_edge(d, m, nt, i, k) : 0.0;
                const double sv = a.model == 12 ? suv(d, m, U, V, i, 0, k) : 0.5 * ((u1 - u0) / (m.yg[1] - m.yg[0]));
                s[Q_TAU] += -2.0 * ne * sv;
                const T *VV = a.model == 12 ? V : Vo;
                const double v1 = 0.5 * ((double)VV[fidx(d, i, 0, k)] + (double)VV[fidx(d, i + 1, 0, k)]);
                const double v2 = 0.5 * ((double)VV[fidx(d, i, 1, k)] + (double)VV[fidx(d, i + 1, 1, k)]);
                const double g = (u1 - u0) / (m.yg[1] - m.yg[0]);
                s[Q_DUVB] += a.model == 12 ? g * (v2 - v1) / (m.y[1] - m.y[0]) : 0.5 * (u1 + u0) * (v2 - v1) / (m.y[1] - m.y[0]);
                s[Q_DUB] += g;
            }
            if (a.model == 15 && i <= nx - 1) {
                const double ne = nt ? nut_edge(d, m, nt, i, k) : 0.0;
                s[Q_NUTW] += (a.nu + ne) / a.tauw[i];
            }
            if (a.model == 4 && i >= 1 && i <= nx - 2) {
                s[Q_SB] += suv(d, m, U, V, i, 0, k);
                s[Q_ST] += suv(d, m, U, V, i, ny - 1, k);
                // V on the cell edges x(i) (interpolate_x, faces to centres)
                const double v1 = 0.5 * ((double)V[fidx(d, i, 0, k)] + (double)V[fidx(d, i + 1, 0, k)]);
                const double v2 = 0.5 * ((double)V[fidx(d, i, 1, k)] + (double)V[fidx(d, i + 1, 1, k)]);
                const double vn = 0.5 * ((double)V[fidx(d, i, ny - 1, k)] + (double)V[fidx(d, i + 1, ny - 1, k)]);
                const double vn1 = 0.5 * ((double)V[fidx(d, i, ny - 2, k)] + (double)V[fidx(d, i + 1, ny - 2, k)]);
                const double gb = ((double)U[fidx(d, i, 1, k)] - (double)U[fidx(d, i, 0, k)]) / (m.yg[1] - m.yg[0]);
                const double gt =
                    ((double)U[fidx(d, i, nyg - 1, k)] - (double)U[fidx(d, i, nyg - 2, k)]) / (m.yg[nyg - 1] - m.yg[nyg - 2]);
                const double bot = gb * (v2 - v1) / (m.y[1] - m.y[0]);
                const double top = -(gt * (vn - vn1) / (m.y[ny - 1] - m.y[ny - 2]));
                s[Q_DUV] += bot + top;
                s[Q_DU] += gb - gt;
            }
            // the interior mean of nu_t rows 2 and nyg-1 (reference numbering)
            // that sgs_Smagorinsky keeps in avg_nu_t(1, j) (subgrid.f90:648-650)
            if ((a.model == 4 || a.model == 6) && a.les == 1 && nt && i >= 1 && i <= nxg - 2) {
                s[Q_N2] += (double)nt[fidx(d, i, 1, k)];
                s[Q_NN] += (double)nt[fidx(d, i, nyg - 2, k)];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kWmQ; ++q) {
        const double r = wm_sum_256(s[q], sh);
        if (threadIdx.x == 0) a.cols[q * nxg + i] = r;
    }
    // models 9, 12: U_inf = U(2, nyg, 2) (reference numbering) in the last
    // column of the scratch row, which the per-column h (i < nx) leaves free
    if ((a.model == 9 || a.model == 12) && i == nxg - 1 && threadIdx.x == 0)
        a.cols[Q_H * nxg + i] = (double)U[fidx(d, 1, nyg - 1, 1)];
}

// f_law_of_wall, df_law_of_wall (Newton_solver.f90:98-118), kappa 0.38, B 5.2
__device__ __forceinline__ double law_f(double u, double ygm, double nu, double Um) {
    return u / 0.38 * log(ygm * u / nu) + 5.20 * u - Um;
}
__device__ __forceinline__ double law_df(double u, double ygm, double nu) {
    return 1.0 / 0.38 * log(ygm * u / nu) + 1.0 / 0.38 + 5.20;
}

__global__ __launch_bounds__(256) void k_wm_final(Dims d, Metrics m, WmArgs a) {
    __shared__ double sh[256];
    const int nx = d.nx, nxg = d.nxg, ny = d.ny, nzg = d.nzg;
    const int tid = threadIdx.x;
    const double *cols = a.cols;
    // model 9 reads beta_y of the last boundary call before this kernel
    // replaces it
    const double beta_prev = a.beta[0];
    __syncthreads();
    if (!a.beta_only) {
        // utau_model (compute_wall_model :40-46)
        if (a.istress == 1) {
            // compute_law_of_the_wall_model (:1951-1985): Newton_iter_solver
            // (Newton_solver.f90:13-85), at most 30 steps, |f| < 1e-10
            const double ygm = m.yg[2];
            for (int i = 1 + tid; i <= nx - 2; i += 256) {
                const double Um = cols[Q_U3 * nxg + i] / (double)(nzg - 2);
                double x = a.utau[i];
                for (int it = 1; it <= 30; ++it) {
                    const double fx = law_f(x, ygm, a.nu, Um);
                    const double fp = law_df(x, ygm, a.nu);
                    if (fabs(fx) < 1e-10) break;
                    x = x - fx / fp;
                }
                a.utau[i] = isnan(x) ? 0.0 : x;
            }
            __syncthreads();
            if (tid == 0) {
                a.utau[0] = a.utau[1];
                a.utau[nx - 1] = a.utau[nx - 2];
            }
        } else {
            for (int i = tid; i < nx; i += 256) a.utau[i] = a.utau_ref[i];
        }
        __syncthreads();
        // avg_nu_t(1, j) at the wall faces y(1) and y(ny): the constant-
        // coefficient Smagorinsky's clipped interior means with their ghost
        // rows (subgrid.f90:648-660); 0 for the other eddy viscosities
        double nb = 0.0, nt = 0.0;
        if ((a.model == 4 || a.model == 6) && a.les == 1) {
            double p2 = 0.0, pn = 0.0;
            for (int i = 1 + tid; i <= nxg - 2; i += 256) {
                p2 += cols[Q_N2 * nxg + i];
                pn += cols[Q_NN * nxg + i];
            }
            const double cnt = (double)((nxg - 2) * (nzg - 2));
            const double a2 = fmax(wm_sum_256(p2, sh) / cnt, 0.0);
            const double an = fmax(wm_sum_256(pn, sh) / cnt, 0.0);
            const double g2 = a.nut_dirichlet == 1 ? -a2 : a2;
            const double gn = a.nut_dirichlet == 1 ? -an : an;
            nb = m.wy0[0] * g2 + m.wy1[0] * a2;
            nt = m.wy0[ny - 1] * an + m.wy1[ny - 1] * gn;
        }
        if (a.model == 4) {
            // compute_alpha_momentum_wall_model (:674-764)
            double sb = 0.0, st = 0.0, duv = 0.0, du = 0.0;
            for (int i = 1 + tid; i <= nx - 2; i += 256) {
                sb += cols[Q_SB * nxg + i];
                st += cols[Q_ST * nxg + i];
                duv += cols[Q_DUV * nxg + i];
                du += cols[Q_DU * nxg + i];
            }
            sb = wm_sum_256(sb, sh);
            st = wm_sum_256(st, sh);
            duv = wm_sum_256(duv, sh);
            du = wm_sum_256(du, sh);
            const double cnt = (double)(2 * (nx - 2) * (nzg - 2));
            const double mtau = (-2.0 * nb * sb + 2.0 * nt * st) / cnt;
            double al = (a.nu * (du / cnt) - a.dPdx * a.Ly / 2.0 - mtau) / (duv / cnt);
            al = sqrt(fmax(al, 0.0));
            for (int i = tid; i < nx; i += 256) a.ax[i] = al;
            for (int i = tid; i < nxg; i += 256) a.ay[i] = al, a.az[i] = al;
        } else if (a.model == 5) {
            for (int i = tid; i < nx; i += 256) a.ax[i] = a.a5x;
            for (int i = tid; i < nxg; i += 256) a.ay[i] = a.a5y, a.az[i] = a.a5z;
        } else if (a.model == 6 || a.model == 8) {
            // compute_log_layer_wall_model (:987-1026)
            const double al = a.model == 6 ? (nb + a.nu) / (0.38 * a.utau_p) * a.logf : (m.y[1] - m.y[0]) * a.logf;
            for (int i = tid; i < nx; i += 256) a.ax[i] = al;
            for (int i = tid; i < nxg; i += 256) a.ay[i] = al, a.az[i] = al;
        } else if (a.model == 9 || a.model == 12) {
            // compute_alpha_v_from_Cf (:1532-1697), compute_alpha_u_from_Cf
            // (:1263-1424): tau_w from Cf = 0.027 Re_x^(-1/7) with U_inf the
            // field's U(2, nyg, 2); alpha_x from the viscous fraction; per
            // column h = (nu <dU/dy> - tau_w - <tau_uv> [+ beta_y <U>_wall])
            // / ([alpha_x] <...>), alpha_y the mean of h over the two
            // columns beside each V face (model 12 clipped to 2e-3)
            const double Uinf = cols[Q_H * nxg + nxg - 1];
            const double c2 = (double)(2 * (nzg - 2)), c1 = (double)(nzg - 2);
            for (int i = tid; i < nx; i += 256) {
                const double tw = 0.5 * (0.027 * pow(Uinf * m.x[i] / a.nu, -1.0 / 7.0)) * (Uinf * Uinf);
                const double um = cols[Q_UW * nxg + i] / c2;
                double al;
                if (a.nut_flag == 2) {
                    const double ur = cols[Q_UR * nxg + i] / c2;
                    const double fd = a.nu * ur / (tw * m.y[2]);
                    al = a.nu * um / (fd * tw);
                } else {
                    al = a.nu * um / (a.frac_vis * tw);
                }
                if (al < 0.0) al = 1e-3;
                if (fabs(al) < 1e-8) al = 1e-3;
                a.ax[i] = al;
                const double mt = cols[Q_TAU * nxg + i] / c1, duv = cols[Q_DUVB * nxg + i] / c1,
                             du = cols[Q_DUB * nxg + i] / c1;
                a.cols[Q_H * nxg + i] = a.model == 12 ? (a.nu * du - tw - mt) / (al * duv)
                                                      : (a.nu * du - tw - mt + um * beta_prev) / duv;
            }
            __syncthreads();
            for (int i = 1 + tid; i <= nxg - 2; i += 256) {
                double al = 0.5 * (cols[Q_H * nxg + i - 1] + cols[Q_H * nxg + i]);
                if (a.model == 12 && fabs(al) > 2e-3) al = 2e-3 * al / fabs(al);
                a.ay[i] = al;
                a.az[i] = al;
            }
            __syncthreads();
            if (tid == 0) {
                a.ay[0] = a.ay[1], a.az[0] = a.az[1];
                a.ay[nxg - 1] = a.ay[nxg - 2], a.az[nxg - 1] = a.az[nxg - 2];
            }
        } else if (a.model == 11 || a.model == 15) {
            // compute_alpha_u_momentum_wall_model (:1444-1500),
            // compute_exact_Neumann_from_Cf (:2003-2066)
            for (int i = tid; i < nx; i += 256) {
                double al;
                if (a.model == 11) {
                    const double u = a.utau[i];
                    const double mt = cols[Q_TAU * nxg + i] / (double)(nzg - 2);
                    al = fabs(u) < 1e-10 ? 0.0 : (u * u + mt) / a.nu;
                    al = fmax(al, 0.0);
                } else {
                    al = 1.0 / (cols[Q_NUTW * nxg + i] / (double)(nzg - 2));
                }
                a.ax[i] = al;
            }
            for (int i = tid; i < nxg; i += 256) a.ay[i] = 0.0, a.az[i] = 0.0;
        }
        __syncthreads();
    }
    // beta_y: the mean of alpha_y dVo/dy over the interior wall
    // (boundary_conditions.f90:1168-1174, single rank: planes 2 .. nzg-2)
    if (wm_robin(a.model)) {
        double p = 0.0;
        for (int i = 1 + tid; i <= nxg - 3; i += 256) p += a.ay[i] * cols[Q_DVO * nxg + i] / (m.y[1] - m.y[0]);
        const double b = wm_sum_256(p, sh) / (double)((nxg - 3) * (nzg - 3));
        if (tid == 0) a.beta[0] = b;
    }
}

// device layout of the wall-model memory
struct WmLayout {
    double *ax, *ay, *az, *utau, *utau_ref, *tauw, *cols, *beta;
};
WmLayout wm_layout(const Dims &d, double *p) {
    WmLayout L;
    L.ax = p;
    L.ay = L.ax + d.nx;
    L.az = L.ay + d.nxg;
    L.utau = L.az + d.nxg;
    L.utau_ref = L.utau + d.nx;
    L.tauw = L.utau_ref + d.nx;
    L.beta = L.tauw + d.nx;
    L.cols = L.beta + 2;
    return L;
}
size_t wm_doubles(const Dims &d) { return (size_t)4 * d.nx + 2 * d.nxg + 2 + (size_t)kWmQ * d.nxg; }

double *wm_alloc(size_t n) {
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, n * sizeof(double));
    if (e != hipSuccess) {
        last_error() = "hipMalloc of " + std::to_string(n * sizeof(double)) + " bytes: " + hipGetErrorString(e);
        throw std::bad_alloc();
    }
    return static_cast<double *>(p);
}

}  // namespace

bool wall_model_robin(int model) { return wm_robin(model); }

void wall_model_set(Handle &h, int model, int istress, const double am[3], double astd, double freq, int nut_flag,
                    double frac_vis) {
    const Dims &d = h.d;
    const bool ok = model == 0 || model == 4 || model == 5 || model == 6 || model == 8 || model == 9 || model == 11 ||
                    model == 12 || model == 15;
    if (!ok) {
        last_error() = "bl_set_wall_model: iwall_model " + std::to_string(model) +
                       " is not built (0, 4, 5, 6, 8, 9, 11, 12 and 15 are)";
        throw ArgError();
    }
    if ((model == 9 || model == 12) && nut_flag != 0 && nut_flag != 2) {
        last_error() = "bl_set_wall_model: iwall_model_nut must be 0 (fixed viscous fraction) or 2 (dynamic)";
        throw ArgError();
    }
    if ((model == 9 || model == 12) && nut_flag == 0 && frac_vis == 0.0) {
        last_error() = "bl_set_wall_model: models 9 and 12 with iwall_model_nut 0 need frac_vis != 0";
        throw ArgError();
    }
    if (istress != 0 && istress != 1) {
        last_error() = "bl_set_wall_model: istress_model must be 0 (utau_ref) or 1 (law of the wall)";
        throw ArgError();
    }
    if ((model == 6 || model == 8) && !(h.dPdx > 0.0)) {
        last_error() = "bl_set_wall_model: the log-layer models take utau = dPdx^0.5 and need dPdx > 0";
        throw ArgError();
    }
    if (d.nx < 4 || d.ny < 3 || d.nzg < 5) {
        last_error() = "bl_set_wall_model: grid too small for the wall sums";
        throw ArgError();
    }
    if (!h.wm_mem) h.wm_mem = wm_alloc(wm_doubles(d));
    const WmLayout L = wm_layout(d, h.wm_mem);
    // initialization.f90:745-746 and the Cf correlation of
    // compute_exact_Neumann_from_Cf (wallmodel.f90:2024-2026), U_inf = 1
    std::vector<double> ur(d.nx), tw(d.nx);
    for (int i = 0; i < d.nx; ++i) {
        ur[i] = std::pow(0.5 * 0.027 * std::pow(h.hx[i] / h.nu, -1.0 / 7.0), 0.5);
        const double Cf = 0.027 * std::pow(1.0 * h.hx[i] / h.nu, -1.0 / 7.0);
        tw[i] = 0.5 * Cf * std::pow(1.0, 2.0);
    }
    // kernels: the slip lengths and beta start at 0; utau_model = utau_ref
    std::vector<double> zeros(2 * (size_t)d.nxg + d.nx + 2, 0.0);
    BL_HIP(hipMemcpyAsync(L.ax, zeros.data(), sizeof(double) * (d.nx + 2 * (size_t)d.nxg), hipMemcpyHostToDevice,
                          h.stream));
    BL_HIP(hipMemcpyAsync(L.beta, zeros.data(), sizeof(double) * 2, hipMemcpyHostToDevice, h.stream));
    BL_HIP(hipMemcpyAsync(L.utau, ur.data(), sizeof(double) * d.nx, hipMemcpyHostToDevice, h.stream));
    BL_HIP(hipMemcpyAsync(L.utau_ref, ur.data(), sizeof(double) * d.nx, hipMemcpyHostToDevice, h.stream));
    BL_HIP(hipMemcpyAsync(L.tauw, tw.data(), sizeof(double) * d.nx, hipMemcpyHostToDevice, h.stream));
    BL_HIP(hipStreamSynchronize(h.stream));
    h.wm_model = model;
    h.wm_istress = istress;
    for (int c = 0; c < 3; ++c) h.wm_am[c] = am[c];
    h.wm_std = astd;
    h.wm_freq = freq;
    h.wm_nut_flag = nut_flag;
    h.wm_frac_vis = frac_vis;
    h.bc.iwall = model;
    h.bc.wm_ax = L.ax;
    h.bc.wm_ay = L.ay;
    h.bc.wm_az = L.az;
    h.bc.wm_beta = L.beta;
}

template <class T>
void launch_wall_model(Handle &h, const Fld3<T> &cur, const Fld3<T> &old, bool beta_only) {
    if (!h.wm_mem) return;
    if (beta_only && !wm_robin(h.wm_model)) return;
    const Dims &d = h.d;
    const WmLayout L = wm_layout(d, h.wm_mem);
    WmArgs a{};
    a.model = h.wm_model;
    a.istress = h.wm_istress;
    a.les = h.les;
    a.nut_dirichlet = h.nut_bc;
    a.nut_flag = h.wm_nut_flag;
    a.frac_vis = h.wm_frac_vis;
    a.beta_only = beta_only ? 1 : 0;
    a.nu = h.nu;
    a.dPdx = h.dPdx;
    a.Ly = h.hy[d.ny - 1] - h.hy[0];
    // compute_constant_alpha (wallmodel.f90:952-961) at the current t
    const double pi = 4.0 * std::atan(1.0);
    const double omega = 2.0 * pi * h.wm_freq / (h.hy[1] - h.hy[0]) * std::pow(h.dPdx, 0.5);
    const double f5 = 1.0 + h.wm_std * std::sin(omega * h.t);
    a.a5x = h.wm_am[0] * f5;
    a.a5y = h.wm_am[1] * f5;
    a.a5z = h.wm_am[2] * f5;
    // compute_log_layer_wall_model (:1010-1015): kappa 0.38, B 5.2
    if (h.wm_model == 6 || h.wm_model == 8) {
        a.utau_p = std::pow(h.dPdx, 0.5);
        a.logf = std::log((h.hy[1] - h.hy[0]) * a.utau_p / h.nu * std::pow(std::exp(1.0), 5.20 * 0.38));
    }
    a.ax = L.ax;
    a.ay = L.ay;
    a.az = L.az;
    a.utau = L.utau;
    a.beta = L.beta;
    a.utau_ref = L.utau_ref;
    a.tauw = L.tauw;
    a.cols = L.cols;
    // LES = 0: compute_eddy_viscosity sets nu_t = 0 (subgrid.f90:41-45)
    const T *nt = h.les != 0 ? reinterpret_cast<const T *>(h.nu_t) : nullptr;
    hipLaunchKernelGGL(k_wm_cols<T>, dim3(d.nxg), dim3(256), 0, h.stream, d, h.m, a, cur.u, cur.v, old.v, nt);
    check();
    hipLaunchKernelGGL(k_wm_final, dim3(1), dim3(256), 0, h.stream, d, h.m, a);
    check();
}
template void launch_wall_model<double>(Handle &, const Fld3<double> &, const Fld3<double> &, bool);
template void launch_wall_model<float>(Handle &, const Fld3<float> &, const Fld3<float> &, bool);

void wall_model_get(Handle &h, double *ax, double *ay, double *az, double *utau) {
    const Dims &d = h.d;
    if (!h.wm_mem) {
        if (ax) std::fill(ax, ax + d.nx, 0.0);
        if (ay) std::fill(ay, ay + d.nxg, 0.0);
        if (az) std::fill(az, az + d.nxg, 0.0);
        if (utau) std::fill(utau, utau + d.nx, 0.0);
        return;
    }
    const WmLayout L = wm_layout(d, h.wm_mem);
    BL_HIP(hipStreamSynchronize(h.stream));
    if (ax) BL_HIP(hipMemcpy(ax, L.ax, sizeof(double) * d.nx, hipMemcpyDeviceToHost));
    if (ay) BL_HIP(hipMemcpy(ay, L.ay, sizeof(double) * d.nxg, hipMemcpyDeviceToHost));
    if (az) BL_HIP(hipMemcpy(az, L.az, sizeof(double) * d.nxg, hipMemcpyDeviceToHost));
    if (utau) BL_HIP(hipMemcpy(utau, L.utau, sizeof(double) * d.nx, hipMemcpyDeviceToHost));
}

}  // namespace blgpu
