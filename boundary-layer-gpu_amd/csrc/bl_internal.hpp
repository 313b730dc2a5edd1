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
#include <vector>

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

// a / b for 0 <= a < 16384 and 1 <= b <= 128, with rb = 1.0f / b: exact on
// that range (checked exhaustively), and a few instructions instead of the
// integer division sequence
__device__ __forceinline__ int small_div(int a, float rb) { return (int)(((float)a + 0.5f) * rb); }

// XCD-aware tile order: the T tiles of a launch are cut into 8 contiguous
// runs, one per XCD (the hardware deals consecutive block ids round-robin
// over the 8 XCDs), so that tiles sharing halo lines share an L2.  The launch
// has xcd_blocks(T) blocks; a block whose tile is >= T returns at once.
__host__ __device__ inline int xcd_blocks(int T) { return 8 * ((T + 7) / 8); }
__host__ __device__ inline int xcd_tile(int b, int T) { return (b % 8) * ((T + 7) / 8) + b / 8; }

// Partitioned Thomas solve: rows per segment (held in registers) and
// segments per system, so systems up to NM = 256 rows take that kernel.
constexpr int kThSeg = 32, kThMaxP = 8;

// Block order of the Thomas sweeps (nkb kx blocks times M spectral rows, one
// block id b each).  Rows q = 2k-1 and 2k (real and imaginary parts of z
// wavenumber k) read the same factors, so their blocks for one kx range go
// to one XCD back to back: the hardware deals consecutive block ids
// round-robin over 8 XCDs, and b = 16 g + s pairs with b + 8 (s < 8).  The
// unpaired rows (q = 0, and q = M-1 when M is even) come last.
__host__ __device__ inline void thomas_block(int b, int nkb, int M, int &kxb, int &q) {
    const int P = nkb * ((M - 1) / 2);  // (kx block, k) pairs
    if (b < 2 * P) {
        const int G = P / 8;
        int p, member;
        if (b < 16 * G) {
            const int s = b % 16;
            p = 8 * (b / 16) + s % 8;
            member = s / 8;
        } else {  // the last, partial group: its pairs one after the other
            const int s = b - 16 * G;
            p = 8 * G + s / 2;
            member = s % 2;
        }
        kxb = p % nkb;
        q = 2 * (p / nkb + 1) - 1 + member;
    } else {
        const int s = b - 2 * P;
        kxb = s % nkb;
        q = s < nkb ? 0 : M - 1;
    }
}

// Metric arrays (0-based, reciprocal spacings) living on the device, in the
// arithmetic type of the kernels that read them.
template <class T>
struct MetricsT {
    const T *x, *y, *z, *xg, *yg, *zg;
    const T *rdx;   // 1/(x[i]-x[i-1]), i>=1
    const T *rdxg;  // 1/(xg[i+1]-xg[i])
    const T *rdy;   // 1/(y[j]-y[j-1]), j>=1
    const T *rdyg;  // 1/(yg[j+1]-yg[j])
    const T *rdz;   // 1/(z[k]-z[k-1]), k>=1
    const T *rdzg;  // 1/(zg[k+1]-zg[k])
    const T *wy0, *wy1;
};
using Metrics = MetricsT<double>;

// Lund rescaling state and work arrays (bl_lund.hip), FP64 on the device;
// rows per grid: cell rows (nyg) and face rows (ny)
struct LundArrays {
    double *umr, *umi, *vmr, *vmi;      // z means of the step-start field (rescaling plane, inlet)
    double *umrT, *umiT, *vmrT, *vmiT;  // their time averages (Umean_resc_T, ..., global.f90)
    double *scal;                       // <UV> at the wall, gamma, U_inf
    double *cgi, *cgo, *ugi, *ugo, *wg; // cell rows: interpolation weights, mean profiles, blend weight
    double *cfi, *cfo, *vfi, *vfo, *wf; // face rows: the same for V
    int *jgi, *jgo, *jfi, *jfo;         // interpolation rows (inner, outer) per target row
    double *pu, *pv, *pw;               // the inflow plane: U(nyg, nzg), V(ny, nzg), W(nyg, nz)
};

struct BCTables {
    const double *U_inlet, *V_inlet, *W_inlet, *U_top, *V_top, *W_top;
    int n_modes, m_modes;
    const double *zmode, *tmode, *qu_re, *qu_im, *qv_re, *qv_im, *qw_re, *qw_im;
    // the general dispatcher's flags (bl_set_boundary); defaults: the GPU set
    int inflow_flag = 1, top_flag = 0;
    double amp = 0.0;               // inflow_flag 4 amplitude
    unsigned long long seed = 0;    // inflow_flag 4 stream
    long long istep = 0;            // inflow_flag 4: the step whose values are drawn
    const double *vtop_bs = nullptr;  // top_flag 1, 2: V(i, ny) of blowing/suction (nxg)
    const double *utop_fs = nullptr;  // top_flag 3: Falkner-Skan U(i, nyg) (nx)
    double dy_dx = 0.0;             // top_flag 1, 2: (yg(nyg) - yg(nyg-1)) / dxmin
    int i_rescale = 0;              // top_flag 1, 2: columns kept at U/V/W_top
    const double *lund_u = nullptr, *lund_v = nullptr, *lund_w = nullptr;  // inflow_flag 3, 5: the step's plane
    // the wall model's bottom-wall conditions (bl_set_wall_model): slip
    // lengths alpha_x (nx), alpha_y, alpha_z (nxg) and the pseudo-Robin V's
    // wall mean beta_y
    int iwall = 0;
    const double *wm_ax = nullptr, *wm_ay = nullptr, *wm_az = nullptr, *wm_beta = nullptr;
};

template <class T>
struct Fld3 {
    T *u, *v, *w;
};
// Storage of a velocity triple.  In mixed precision the three arrays hold
// float; as<float>() gives the typed view.
using Field3 = Fld3<double>;
template <class T>
inline Fld3<T> as(const Field3 &f) {
    return {reinterpret_cast<T *>(f.u), reinterpret_cast<T *>(f.v), reinterpret_cast<T *>(f.w)};
}

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
    KID_REDUCE,
    KID_DSM_CELLS,
    KID_DSM_VFILTER,
    KID_DSM_FILTER,
    KID_DSM_COLUMN,
    KID_CLOSURE,
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
    // the general dispatcher (bl_set_boundary)
    long long bc_step = 0;          // steps since the flags were set (inflow_flag 4 stream)
    double *bc_extra = nullptr;     // device V(i, ny) blowing/suction (nxg) + Falkner-Skan U(i, nyg) (nx)
    bool fs_uc_pending = false;     // top_flag 3: U_top(nx) changes after the next BC call
    double fs_uc = 0.0;
    std::vector<double> hx, hxg, hyg, hy;  // host copies of x, xg, yg, y
    // wall model (bl_wall.hip): iwall_model, istress_model, alphas(5)
    int wm_model = 0, wm_istress = 0;
    double wm_am[3] = {0.0, 0.0, 0.0}, wm_std = 0.0, wm_freq = 0.0;
    int wm_nut_flag = 0;         // iwall_model_nut (models 9, 12)
    double wm_frac_vis = 0.0;    // frac_vis_wall_model (models 9, 12)
    double *wm_mem = nullptr;
    // Lund rescaling (inflow_flag 3, 5)
    double *lund_mem = nullptr;
    LundArrays lund{};
    int lund_ir = 0;                   // rescaling plane, 0-based
    double lund_t_resc = 0.0, lund_delta_inlet = 0.0;
    bool lund_pending = false;         // the next boundary call starts a step (step_beginning)
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
    Field3 F1;      // first-stage RHS (RK3); FP64 RK3: G = Uo + dt a21 F1 instead (a21 = a31)
    double *P, *phi, *nu_t;  // float in mixed precision
    double *dsm;  // dynamic Smagorinsky work arrays (LES = 2, 3 and 4), kDsmArrays field-sized slots
    double *smag_c;  // LES = 1: (Cs Delta_j f_j)^2 per y row (nyg)
    // learned closure (LES = 4): BF16 weights, FP32 biases, input range
    unsigned short *cl_w;
    float *cl_b, *cl_lo, *cl_hi;
    double cl_inv_d99;
    double *cl_cnt;  // fallback cells per block of the last evaluation
    int cl_nblk;
    int cl_ready;
    double *T1, *S;  // transform intermediates: [m][j][kx] and [q][j][kx]
    int dct_n1;                      // x-DCT factorisation N = dct_n1 * (N / dct_n1)
    int lda_zt;                      // leading dimension of the z cos/sin tables
    double *zc, *zs;                 // z tables cos/sin(2 pi k m / M), [m][k]
    double *zf = nullptr;            // their matrix-core fragments (M/2 + 1 <= 96)
    double *dct_w1, *dct_w2, *dct_wN, *dct_rot;  // x-DCT twiddle tables (complex)
    bool dct_mfma = false;                       // x-DCT on the FP64 matrix cores (shape fits)
    double *dct_mf = nullptr;                    // their operand tables: fA, fC, iC, iA
    size_t dct_mf_off[4] = {0, 0, 0, 0};
    double *lu_dl, *lu_dp, *lu_du;  // Thomas factors [k][j][kx], [k][j][kx], [j] (table kernel only)
    double *lu00;                   // the pinned (0,0) system: multipliers [NM], reciprocal pivots [NM]
    int th_p = 0;                   // segments of the partitioned Thomas kernel (0: table kernel)
    int n_cu = 0;                   // compute units of the device (persistent launches)
    double *th_co = nullptr;        // its coefficients: a, b, c [32 th_p] each, kxx [N], kzz [Kz]
    double *scal;                   // device scalars (mass-flux correction, reductions)
    double *red;                    // reduction scratch
    void *mem_grid;                 // one allocation for metrics and tables
    void *mem_gridf;                // the float metric copies
    Metrics m;
    MetricsT<float> mf;             // float copies for the FP32 stencils
    BCTables bc;
    // mixed precision: BF16 hi/lo transform matrices, FP64 (0,0) mode
    int NP;                         // kx stride of the FP32 spectral arrays (N rounded up to 32)
    int kpx, kpz;                   // GEMM depths N and M rounded up to 32
    bool mx_two = false;            // x transform as two short DFTs (k_xfwd_mx / k_xinv_mx)
    void *mx2 = nullptr;            // their operand and twiddle tables (one allocation)
    const void *x2_fA = nullptr, *x2_fC = nullptr, *x2_iC = nullptr, *x2_iA = nullptr;
    const float *x2_tw = nullptr;   // wN[N], rot[N] as float2
    unsigned short *mx_tab;         // one allocation for the eight matrices below
    unsigned short *zmf = nullptr;  // folded z transform fragments (k_zdft_mx), M/2 + 1 <= 96
    int zmf_rt = 0;
    const unsigned short *dxf_hi, *dxf_lo, *dxi_hi, *dxi_lo, *zf_hi, *zf_lo, *zi_hi, *zi_lo;
    double *p00;                    // FP64 plane partials of the divergence
    double *c00;                    // phi00(j) / (2 N M)
    double *inc1 = nullptr;         // first-stage FP64 boundary increments: inflow and outflow U
                                    // planes [k][j] (nyg nzg each), then the mass-flux shift
    bool inc_ok = false;            // the current fields are the output of a step (projected: their
                                    // divergence is rounding only), not uploaded or set otherwise
    double *stats = nullptr;        // device statistics: 9 (nx, ny) means, then nu_t_mean (nxg, nyg)
    // asynchronous snapshots: device staging and pinned host copy of U, V, W,
    // P, nu_t in the reference's unpadded FP64 layout (offsets snap_off), the
    // copy stream and its events
    double *snap_dev = nullptr, *snap_host = nullptr;
    size_t snap_off[5] = {0, 0, 0, 0, 0};
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_packed = nullptr, snap_copied = nullptr;
    bool snap_pending = false;
    int np00;                       // partials per y row
    ProfState *prof;                // null unless profiling is enabled
    unsigned long long prof_mask = ~0ull;  // kernel classes the profile brackets (bit = class id)
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

// ---- kernels (bl_kernels.hip); T = double (FP64 mode) or float (mixed) ----
template <class T>
const MetricsT<T> &metrics(const Handle &h);
template <class T>
void launch_rk_stage(Handle &h, const Fld3<T> &cur, const Fld3<T> &old, Fld3<T> &nxt, int stage, double dt);
// inc1 != null: also write the FP64 increments of the inflow and outflow U
// planes over the step-start field (t_old its time) and the mass-flux shift
// into inc1 (layout of Handle::inc1)
template <class T>
void launch_bc(Handle &h, Fld3<T> &f, const Fld3<T> &old, int rk_step, double dt, double t,
               double *inc1 = nullptr, double t_old = 0.0);
// div != null: the rhs is the divergence of *div (computed inside the x transform)
void launch_poisson(Handle &h, double *phi_inout, const Field3 *div = nullptr);
// u -= grad phi (project_velocity, projection.f90:355-386); with P set, also
// P = phi / (dt a11) on every cell (projection.f90:75-85)
template <class T>
void launch_project(Handle &h, Fld3<T> &f, const T *phi, T *P = nullptr, double dt = 0.0);
// The pinned (0,0) system (NM <= kMax00 rows) solved in place in x (row
// stride ldx) and scaled; with part set, its right-hand side is first formed
// from the FP64 plane sums part[j][0..npart) (mixed path).
constexpr int kMax00 = 1024;
struct Solve00 {
    const double *L00, *D00, *du;  // multipliers, reciprocal pivots, super-diagonal
    const double *part;            // null: the rhs is x itself
    int npart;
    double *x;
    int ldx;
    double scale;
};
Solve00 solve00_args(const Handle &h, double *x, int ldx, double scale, const double *part, int npart);
void launch_solve00(Handle &h, const Solve00 &z);
// the y solves of all spectral systems S[q][j][kx] (row stride ldj, q stride
// ldq) by the partitioned kernel, with the (0,0) system z solved by one more
// block of the same launch; false if the handle uses the table kernel.  The
// (0,0) column of S is left alone (z.part null: z solves it in place) or
// zeroed (z.part set: z writes its own output).
template <class T>
bool launch_thomas_seg(Handle &h, T *S, int ldj, long long ldq, const Solve00 &z);
template <class T>
void launch_rhs_only(Handle &h, const Fld3<T> &cur, Fld3<T> &rhs);
template <class T>
double max_divergence(Handle &h, const Fld3<T> &f);
// the max-divergence reduction on fields with their own row (x1) and plane
// (x2) strides and reciprocal face spacings rdx[i] = 1/(x[i]-x[i-1]) (i >= 1)
struct DivArgs {
    int nxg, nyg, nzg;
    long long u1, u2, v1, v2, w1, w2;
    const double *rdx, *rdy, *rdz;
};
// partials the reduction writes (maxdiv_partials of the same extents) into
// part, then the maximum into out[0]; both device memory, stream-ordered
template <class T>
void launch_maxdiv(hipStream_t stream, const DivArgs &a, const T *U, const T *V, const T *W, double *part,
                   double *out);
long long maxdiv_partials(int nxg, int nyg, int nzg);
template <class T>
void wall_stats(Handle &h, const Fld3<T> &f, double *cf_host);
template <class T>
void monitor(Handle &h, const Fld3<T> &f, double *out_host);
template <class T>
double dt_convective(Handle &h, const Fld3<T> &f);
// the z DFT tables cos/sin(2 pi k m / M), [m][k] with leading dimension lda,
// and for M/2 + 1 <= 96 the matrix-core operand fragments zf: for row tile
// t < (M/2 + 16) / 16, k step ks < (M/2 + 4) / 4 and lane l, the pair
// (cos, sin)(2 pi r kk / M) with r = 16 t + l % 16, kk = 4 ks + l / 16, at
// zf[2 ((t KS + ks) 64 + l)] (empty otherwise)
void z_dft_tables(int M, int &lda, std::vector<double> &zc, std::vector<double> &zs, std::vector<double> &zf);
// device z tables of one length M
struct ZTables {
    int M = 0, lda = 0;
    const double *zc = nullptr, *zs = nullptr, *zf = nullptr;
};
// the real z DFT of C columns (plane stride `plane`) on stream s: forward to
// the packed half spectrum (rows 0, 2k-1, 2k), or its unnormalised inverse
void launch_zdft(hipStream_t s, const ZTables &z, int C, long long plane, bool forward, const double *in,
                 double *out);
// ---- Lund rescaling (bl_lund.hip) ----
size_t lund_doubles(const Dims &d);
LundArrays lund_layout(const Dims &d, double *base);
// the step's inflow plane from the step-start field `old` (updates the time averages)
template <class T>
void launch_lund(Handle &h, const Fld3<T> &old, double dt);
// the time averages start as the means of f
template <class T>
void lund_init_means(Handle &h, const Fld3<T> &f);
// ---- wall models (bl_wall.hip) ----
// validate (ArgError with a message), allocate, and reset: slip lengths 0,
// utau_model = utau_ref
void wall_model_set(Handle &h, int model, int istress, const double am[3], double astd, double freq, int nut_flag,
                    double frac_vis);
// compute_wall_model on `cur` at h.t (old: the step-start field of the
// pseudo-Robin V); beta_only: only the pseudo-Robin wall mean, slip lengths kept
template <class T>
void launch_wall_model(Handle &h, const Fld3<T> &cur, const Fld3<T> &old, bool beta_only);
void wall_model_get(Handle &h, double *ax, double *ay, double *az, double *utau);
// the models whose wall is Robin (U, W) and pseudo-Robin (V)
bool wall_model_robin(int model);
// ---- statistics and snapshots (bl_stats.hip) ----
long long stats_size(const Dims &d);
// compute_statistics' z means of f (and the handle's P, nu_t) into out
// (device, stats_size doubles): Umean, Vmean, Wmean, U2mean, V2mean, W2mean,
// UVmean, Pmean, P2mean as (nx, ny), then nu_t_mean (nxg, nyg)
void launch_statistics(Handle &h, const Field3 &f, double *out);
// pack the current U, V, W, P, nu_t on the solver's stream and copy them to
// pinned host memory on the copy stream
void snapshot_begin(Handle &h);
// wait for that copy; ptr[0..4] point into the pinned copy
void snapshot_end(Handle &h, const double **ptr);
void snapshot_free(Handle &h);
// ---- mixed precision (bl_mixed.hip) ----
void build_mixed(Handle &h);
// The first stage's pressure source in the mixed path: the divergence of the
// stage's increment U* - Un, taken as c0 F1 on the faces the stage update
// moves and as the FP64 boundary increments (Handle::inc1) or the difference
// of the stored values on the boundary faces, instead of div U*.  In exact
// arithmetic the two agree (div Un is zero after the previous projection);
// in FP32 storage div Un is rounding noise of order eps |u| / dx, which the
// projection would divide by dt a11 into the recorded pressure.
struct IncSrc {
    const float *fu, *fv, *fw;     // F1 of the stage
    const float *vs, *vn;          // V after the stage's boundary update, and at the step start
    const double *inc1;            // Handle::inc1
    double c0;                     // dt a11
};
void launch_poisson_mixed(Handle &h, float *phi, const Fld3<float> &f, const IncSrc *inc = nullptr);
// eddy viscosity of f into h.nu_t (bl_dsm.hip): LES = 1 constant-coefficient
// Smagorinsky, 2 and 4 dynamic Smagorinsky z-averaged, 3 per cell
constexpr int kDsmArrays = 15;
void launch_dsm(Handle &h, const Field3 &f);
// learned closure over the DSM values in h.nu_t (bl_closure.hip)
void launch_closure(Handle &h, const Field3 &f);

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
struct StateError {};  // the call is not valid in the handle's current state
}  // namespace blgpu
