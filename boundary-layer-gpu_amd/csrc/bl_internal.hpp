// Internal state of a blgpu handle and the kernel launch interface.
//
// Device layout: every 3-D array (U, V, W, P, phi, F1, nu_t) uses one padded
// x-fastest layout idx = i + sx*(j + sy*k) with 0-based (i, j, k),
// sx = nxg rounded up to 16 doubles (128-byte rows), sy = nyg, nzg planes.
// A staggered field simply leaves the unused tail of its shorter dimension
// untouched, so one offset addresses U(i,j,k), V(i,j,k) and W(i,j,k) alike.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

namespace blgpu {

struct Dims {
    int nx, ny, nz;     // face counts
    int nxg, nyg, nzg;  // cells + ghosts
    int sx, sy;         // padded strides
    long long sxy;      // sx*sy
    int N, M, NM;       // transform sizes: nxp = nx-2, nzp = nz-2, nm = ny-1
    int Kz;             // distinct z wavenumbers in the half spectrum: M/2 + 1
};

__host__ __device__ inline long long fidx(const Dims &d, int i, int j, int k) {
    return (long long)i + (long long)d.sx * ((long long)j + (long long)d.sy * k);
}

// Metric arrays (0-based, reciprocal spacings) living on the device.
struct Metrics {
    const double *x, *y, *z, *xg, *yg, *zg;
    const double *rdx;   // 1/(x[i]-x[i-1]), i>=1
    const double *rdxg;  // 1/(xg[i+1]-xg[i])
    const double *rdy;   // 1/(y[j]-y[j-1]), j>=1
    const double *rdyg;  // 1/(yg[j+1]-yg[j])
    const double *rdz;   // 1/(z[k]-z[k-1]), k>=1
    const double *rdzg;  // 1/(zg[k+1]-zg[k])
    const double *wy0, *wy1;
};

struct BCTables {
    const double *U_inlet, *V_inlet, *W_inlet, *U_top, *V_top, *W_top;
    int n_modes, m_modes;
    const double *zmode, *tmode, *qu_re, *qu_im, *qv_re, *qv_im, *qw_re, *qw_im;
};

struct Field3 {
    double *u, *v, *w;
};

// Kernel classes timed by the optional HIP-event profiler.
enum KernelId {
    KID_RK_STAGE = 0,
    KID_BC_INOUT,
    KID_BC_PERIODIC,
    KID_BC_TOPWALL,
    KID_BC_MASSFLUX,
    KID_DIVERGENCE,
    KID_XFORM_X_FWD,
    KID_XFORM_Z_FWD,
    KID_THOMAS,
    KID_XFORM_Z_INV,
    KID_XFORM_X_INV,
    KID_PROJECT,
    KID_PRESSURE,
    KID_REDUCE,
    KID_DSM_CELLS,
    KID_DSM_VFILTER,
    KID_DSM_FILTER,
    KID_DSM_COLUMN,
    KID_COUNT
};
const char *kernel_name(int id);

struct ProfState;  // bl_api.hip

struct Handle {
    Dims d;
    double nu, dPdx, dPdz, cfl;
    int scheme, les, precision;
    int nut_bc;  // nutBC: 1 = Dirichlet (antisymmetric) nu_t ghost rows in y, else Neumann
    double t, dt;
    double length_y;
    double bc_uc;      // U_top(nx): outflow convection speed
    double dxmin, dymin, dzmin;  // smallest centre spacings (initialization.f90:220-222)
    double rk_t[4];
    double rk_coef[4][4];
    double rk2_t[3];
    double rk2_coef[3][3];
    hipStream_t stream;

    // device memory
    Field3 buf[3];  // rotating state buffers
    int iold;       // index of the step-start buffer (Uo)
    int icur;       // index of the current (most recent) buffer
    Field3 F1;      // first-stage RHS (RK3)
    double *P, *phi, *nu_t;
    double *dsm;  // dynamic Smagorinsky work arrays (LES = 2 only), kDsmArrays field-sized slots
    double *T1, *S;  // transform intermediates: [m][j][kx] and [q][j][kx]
    int dct_n1;                      // x-DCT factorisation N = dct_n1 * (N / dct_n1)
    int lda_zt;                      // leading dimension of the z cos/sin tables
    double *zc, *zs;                 // z tables cos/sin(2 pi k m / M), [m][k]
    double *dct_w1, *dct_w2, *dct_wN, *dct_rot;  // x-DCT twiddle tables (complex)
    double *lu_dl, *lu_dp, *lu_du;  // Thomas factors [k][j][kx], [k][j][kx], [j]
    double *scal;                   // device scalars (mass-flux correction, reductions)
    double *red;                    // reduction scratch
    void *mem_grid;                 // one allocation for metrics and tables
    Metrics m;
    BCTables bc;
    ProfState *prof;                // null unless profiling is enabled
};

// Records HIP events around one launch on the handle's stream when
// profiling is enabled.
void prof_begin(Handle &h, int id);
void prof_end(Handle &h, int id);
struct ProfScope {
    Handle &h;
    int id;
    ProfScope(Handle &hh, int i) : h(hh), id(i) { prof_begin(h, id); }
    ~ProfScope() { prof_end(h, id); }
};

// ---- kernels (bl_kernels.hip) ----
void launch_rk_stage(Handle &h, const Field3 &cur, const Field3 &old, Field3 &nxt, int stage,
                     double dt);
void launch_bc(Handle &h, Field3 &f, const Field3 &old, int rk_step, double dt, double t);
void launch_divergence(Handle &h, const Field3 &f, double *out);
// div != null: the rhs is the divergence of *div (computed inside the x transform)
void launch_poisson(Handle &h, double *phi_inout, const Field3 *div = nullptr);
void launch_project(Handle &h, Field3 &f, const double *phi);
void launch_pressure(Handle &h, const double *phi, double *P, double dt);
void launch_rhs_only(Handle &h, const Field3 &cur, Field3 &rhs);
double max_divergence(Handle &h, const Field3 &f);
void wall_stats(Handle &h, const Field3 &f, double *cf_host);
void monitor(Handle &h, const Field3 &f, double *out_host);
double dt_convective(Handle &h, const Field3 &f);
// dynamic Smagorinsky eddy viscosity of f into h.nu_t (bl_dsm.hip)
constexpr int kDsmArrays = 15;
void launch_dsm(Handle &h, const Field3 &f);

std::string &last_error();

}  // namespace blgpu

struct bl_handle : blgpu::Handle {};

#define BL_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) {                                                            \
            blgpu::last_error() = std::string(#call) + ": " + hipGetErrorString(_e);       \
            throw blgpu::DeviceError();                                                    \
        }                                                                                  \
    } while (0)

namespace blgpu {
struct DeviceError {};
struct ArgError {};
}  // namespace blgpu
