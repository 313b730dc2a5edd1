/*
 * blgpu — MI355X (gfx950) implementation of the boundary-layer solver's
 * per-time-step path behind a plain C ABI.
 *
 * Conventions shared by every entry point:
 *  - 3-D fields cross the boundary exactly as the reference Fortran stores
 *    them: column-major, x fastest, full allocated extent
 *      U(nx,nyg,nzg)  V(nxg,ny,nzg)  W(nxg,nyg,nz)  P(nxg,nyg,nzg)  nu_t(nxg,nyg,nzg)
 *    with nxg = nx+1, nyg = ny+1, nzg = nz+1 (initialization.f90:84-104, 126-129).
 *  - Every function returns 0 on success or a nonzero BL_E* code; the message
 *    of the last failure on the calling thread is bl_last_error().  The
 *    reference Stops with a message on errors (cufft_solver.f90:39); a Fortran
 *    caller does `if (ierr /= 0) stop` with that text.
 *  - One handle = one device-resident solver state on the current HIP device,
 *    driven from one host thread on one stream.  Handles are not thread-safe.
 *  - No host pointer passed in is retained after the call returns.
 */
#ifndef BLGPU_H
#define BLGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bl_handle bl_handle;

enum {
    BL_OK = 0,
    BL_EINVAL = 1,   /* bad argument or unsupported configuration   */
    BL_EDEVICE = 2,  /* HIP runtime / device error                  */
    BL_ENOMEM = 3,   /* device allocation failed                    */
    BL_ESTATE = 4    /* call not valid in the handle's current state */
};

enum { BL_FIELD_U = 0, BL_FIELD_V = 1, BL_FIELD_W = 2, BL_FIELD_P = 3, BL_FIELD_NUT = 4 };

enum { BL_PREC_FP64 = 0, BL_PREC_MIXED = 1 };

/*
 * Solver configuration.  Replaces the `!$acc data` region set-up
 * (main.f90:61-81), cufft_init_plans (cufft_solver.f90:26-62) and the Thomas
 * LU / twiddle precompute (initialization.f90:313-320, 427-469): the library
 * derives metrics, transform matrices and LU factors itself from the grids.
 *  x(nx), y(ny), z(nz): face coordinates (x_global, y_global, z_global).
 *  scheme: RKscheme (2 or 3), les: LES flag as compute_eddy_viscosity
 *  dispatches it (subgrid.f90:32-45): 0 DNS, 1 constant-coefficient
 *  Smagorinsky with van Driest damping (sgs_Smagorinsky :609-662), 2 dynamic
 *  Smagorinsky averaged in z (:261-414), 3 dynamic Smagorinsky per cell
 *  (:436-583); and 4, this library's learned closure with dynamic Smagorinsky
 *  fallback,
 *  cfl: .turbb CFL (negative = fixed dt = -cfl).
 *  Boundary tables as built by compute_blasius_solution_for_bc
 *  (boundary_conditions.f90:1452-1797): U_inlet(nyg) V_inlet(ny) W_inlet(nyg)
 *  U_top(nx) V_top(nxg) W_top(nxg).
 *  Temporal inflow modes (optional, n_modes = 0 disables): zmode(n_modes),
 *  tmode(m_modes) already multiplied by beta/omega, and q*_re/q*_im laid out
 *  (j, n, m) with j fastest, already interpolated to yg (u, w) and y (v).
 *  nut_bc: .turbb nutBC (params.f90:156), the y ghost rows of nu_t after the
 *  dynamic Smagorinsky models: 1 = antisymmetric (apply_Dirichlet_bc_y),
 *  otherwise symmetric (apply_Neumann_bc_y), subgrid.f90:406-411; LES = 1
 *  applies its van Driest damping unless nut_bc = 0 and leaves the ghost
 *  entries of nu_t untouched (subgrid.f90:635-646).
 */
typedef struct bl_config {
    int nx, ny, nz;
    const double *x, *y, *z;
    double nu, dPdx, dPdz, cfl;
    int scheme;
    int les;
    int precision;
    const double *U_inlet, *V_inlet, *W_inlet;
    const double *U_top, *V_top, *W_top;
    int n_modes, m_modes;
    const double *zmode, *tmode;
    const double *qu_re, *qu_im, *qv_re, *qv_im, *qw_re, *qw_im;
    int nut_bc;
} bl_config;

/* Create a solver on the current HIP device.  Replaces main.f90:61-81 +
 * cufft_init_plans + initialization.f90:427-469. */
int bl_create(const bl_config *cfg, bl_handle **out);
/* Free all device memory (cufft_destroy_plans cufft_solver.f90:64-72,
 * finalize finalization.f90:20-36). */
void bl_destroy(bl_handle *h);
/* Message of the last failure on this thread ("" if none). */
const char *bl_last_error(void);

/* `!$acc update device(field)` / `!$acc update self(field)` (main.f90:100,
 * time_integration.f90:342-345).  Synchronous; full reference extent. */
int bl_upload(bl_handle *h, int field, const double *host);
int bl_download(bl_handle *h, int field, double *host);

int bl_set_time(bl_handle *h, double t);
int bl_get_time(bl_handle *h, double *t);

/* Boundary-condition flags of the general dispatcher apply_boundary_conditions
 * (boundary_conditions.f90:22-121), read from the .turbb file in
 * params.f90:122-162.  The default (bl_create) is the set the reference's
 * GPU path hard-wires (boundary_conditions.f90:125-262): inflow_flag 1,
 * top_flag 0, no-slip walls (iwall_model 0).
 *   inflow_flag 1, 2: Blasius + temporal modes (:382-402; 2 reads the
 *     profile from a file, which arrives here as the same U/V/W_inlet tables)
 *   inflow_flag 4: Blasius + uniform random perturbations of amplitude
 *     `amplitude` weighted by (1 - U_inlet(j)), drawn once per time step
 *     (:412-451).  The random stream is a counter hash of (seed, step,
 *     component, j, k), not gfortran's random_number: the values differ from
 *     a reference run, their distribution and the once-per-step update do not.
 *   top_flag 0: Dirichlet U_top, V_top, W_top (:947-964)
 *   top_flag 1, 2: blowing/suction V at the top (Coleman 2018 / Abe 2017,
 *     :648-673, parameters vbs_max, x_bs, sigma_bs, phi_bs = .turbb Vmax,
 *     x0, sigma, phi), then zero spanwise vorticity and dW/dy = 0 at the top
 *     with columns 1..i_rescale (.turbb Lund_ix) kept at U/V/W_top (:264-283)
 *   top_flag 3: Falkner-Skan U_top(x) = C (x + 1)^(beta/(2 - beta)) matched
 *     to 1 at x(1), V = 0, dW/dy = 0 (:979-1007); beta = beta_hartree.  As in
 *     the reference, U_top takes the Falkner-Skan values at the first
 *     boundary-condition call, so the outflow's convective velocity U_top(nx)
 *     changes from the next call on.
 *   inflow_flag 3: Lund rescaled inflow (apply_inflow_bc_x_rescaling :569-597,
 *     compute_rescaled_inflow rescaled_inlet_bc.f90:60-300): once per step,
 *     from the step-start field, the plane i_rescale (.turbb Lund_ix,
 *     1-based) rescaled in inner (u_tau) and outer (delta99 against
 *     delta_inlet = Lund_deltai) coordinates, with the means averaged in time
 *     over t_resc (Lund_T) from the field's means at this call
 *     (initialization.f90:640-660); inflow_flag 5: its fluctuations only, on
 *     the U/V/W_inlet means (:512-559).  The first boundary call after this
 *     one (bl_apply_bc) or the next bl_step computes the plane.
 * Wall models: bl_set_wall_model.  Resets the step counter of the
 * inflow_flag 4 stream. */
typedef struct bl_boundary {
    int inflow_flag, top_flag;
    double amplitude;
    unsigned long long seed;
    double vbs_max, x_bs, sigma_bs, phi_bs;
    int i_rescale;
    double beta_hartree;
    double delta_inlet, t_resc;
} bl_boundary;
int bl_set_boundary(bl_handle *h, const bl_boundary *b);

/* Wall model (.turbb WM, TauwModel, alphas; params.f90:154-155, 198-206) and
 * the bottom-wall conditions it feeds (apply_boundary_conditions
 * boundary_conditions.f90:73-108).  compute_wall_model
 * (wallmodel.f90:25-144) runs on the device at every RK stage after the eddy
 * viscosity, as time_integration.f90:341-346 calls it when LES > 0 or
 * WM > 0, and sets the slip lengths alpha_x(nx), alpha_y(nxg), alpha_z(nxg)
 * (functions of x, as every model builds them):
 *   iwall_model 0: no-slip (the default);
 *   4: momentum balance at both walls with the wall value of avg_nu_t
 *      (compute_alpha_momentum_wall_model :674-764; a channel model: its
 *      averages include the top wall);
 *   5: alpha_mean_[xyz] (1 + alpha_std sin(omega t)),
 *      omega = 2 pi freq_mult/(y(2) - y(1)) dPdx^0.5 (compute_constant_alpha
 *      :952-961);
 *   6, 8: log-layer slip length with utau = dPdx^0.5, kappa 0.38, B 5.2,
 *      with the wall value of avg_nu_t (6) or nu_t = kappa utau y (8)
 *      (compute_log_layer_wall_model :987-1026); dPdx > 0;
 *   9, 12: slip lengths from the Cf = 0.027 Re_x^(-1/7) correlation with
 *      U_inf = U(2, nyg, 2): alpha_x = nu <U>_wall/(frac tau_w) with the fixed
 *      viscous fraction frac_vis (iwall_model_nut 0) or the dynamic one
 *      nu <U>_rows 3:4/(tau_w y(3)) (2); alpha_y = alpha_z from the wall
 *      stress balance, averaged over the two columns beside each face
 *      (compute_alpha_v_from_Cf :1532-1697 with the step-start V and the
 *      last boundary call's beta_y; compute_alpha_u_from_Cf :1263-1424,
 *      |alpha_y| <= 2e-3);
 *   11: Neumann dU/dy = (utau_model^2 + <tau_uv>_z)/nu, V = W = 0
 *      (compute_alpha_u_momentum_wall_model :1444-1500);
 *   15: Neumann dU/dy = 1/<(nu + nu_t)/tau_w>_z with tau_w from
 *      Cf = 0.027 Re_x^(-1/7), V = W = 0 (compute_exact_Neumann_from_Cf
 *      :2003-2066).
 * With 4, 5, 6, 8, 9, 12 the wall is Robin for U and W (apply_Robin_bc_y_bottom
 * :1107-1132) and pseudo-Robin for V from the step-start field with its wall
 * mean removed (apply_pseudo_Robin_bc_y_bottom :1152-1180).  istress_model 1
 * updates utau_model by Newton's method on the law of the wall at yg(3)
 * (compute_law_of_the_wall_model :1951-1985); otherwise utau_model =
 * utau_ref (initialization.f90:745).  avg_nu_t(1, :) of models 4 and 6 is the
 * constant-coefficient Smagorinsky's wall-parallel mean (LES 1,
 * subgrid.f90:648-660) and 0 for the other eddy viscosities, whose routines
 * do not write it.  Models 1, 2, 3, 7 (test-filtered slip lengths) and 13, 14
 * (Cf-fitted wall blowing, whose alpha_z = alpha_x/10 assigns arrays of
 * different shapes, wallmodel.f90:1933) are not built and give BL_EINVAL.  Resets the slip lengths to 0 and utau_model to utau_ref. */
typedef struct bl_wall_params {
    int iwall_model, istress_model;
    double alpha_mean_x, alpha_mean_y, alpha_mean_z, alpha_std, freq_mult;
    int iwall_model_nut;  /* WMnutflag (models 9, 12): 0 fixed, 2 dynamic viscous fraction */
    double frac_vis;      /* WMnut (models 9, 12): the fixed viscous fraction */
} bl_wall_params;
int bl_set_wall_model(bl_handle *h, const bl_wall_params *p);
/* compute_wall_model (wallmodel.f90:25-144) once on the current fields at
 * the current time (the step does it itself at every stage). */
int bl_compute_wall_model(bl_handle *h);
/* The slip lengths alpha_x(nx), alpha_y(nxg), alpha_z(nxg) and utau_model(nx)
 * (host arrays; any may be null). */
int bl_get_wall_model(bl_handle *h, double *alpha_x, double *alpha_y, double *alpha_z, double *utau_model);

/* compute_dt (time_integration.f90:528-591): fixed dt for cfl < 0, else the
 * reference's max-based convective bound (a device reduction) and the
 * viscous bound.  The result is also stored as the handle's dt. */
int bl_compute_dt(bl_handle *h, double *dt);

/* compute_time_step_RK2 / RK3 (time_integration.f90:145-300, 305-465):
 * one full step with the handle's dt, asynchronous on the handle's stream
 * (no host synchronisation inside).  t advances to t + dt. */
int bl_step(bl_handle *h);
/* Run nsteps steps back to back (the time loop of main.f90:84-113 without
 * the I/O). */
int bl_run(bl_handle *h, int nsteps);
/* Block until queued work is finished. */
int bl_synchronize(bl_handle *h);

/* compute_eddy_viscosity (subgrid.f90:23-47) on the current fields:
 * recomputes nu_t on the device for the handle's LES model (1, 2, 3, 4 as in
 * bl_config; LES = 0: nu_t = 0).  bl_step
 * does this itself at the start of every RK stage, as the reference does
 * (time_integration.f90:341-346, 383-388, 425-430). */
int bl_eddy_viscosity(bl_handle *h);

/* check_divergence (projection.f90:391-413), FP64 device max-reduction. */
int bl_max_divergence(bl_handle *h, double *div);
/* Cf part of compute_statistics (statistics.f90:137-160): cf(nx). */
int bl_wall_stats(bl_handle *h, double *cf);
/* output_monitor reductions (monitor.f90:29-52): out[8] = sum U, V, W,
 * nu_t over the interiors, then max U, V, W, nu_t. */
int bl_monitor(bl_handle *h, double *out);
/* The z means of compute_statistics (statistics.f90:44-80, 125-135) on the
 * device, FP64 accumulation: means[9*nx*ny] receives Umean, Vmean, Wmean,
 * U2mean, V2mean, W2mean, UVmean, Pmean, P2mean, each (nx, ny) x fastest,
 * from term_1 = interpolate_y(U), term_2 = interpolate_x(V) over z planes
 * 2..nzg-1 and term = interpolate_y(interpolate_x(W)) over planes 1..nz-2;
 * nu_t_mean[nxg*nyg] (may be NULL) the nu_t mean, with column nxg and row
 * nyg 0 as compute_statistics leaves them; cf[nx] (may be NULL) as
 * bl_wall_stats.  Replaces the field download of statistics.f90:58. */
int bl_statistics(bl_handle *h, double *means, double *nu_t_mean, double *cf);
/* Snapshot fields for output_data (input_output.f90:496-680) without
 * stalling the time loop: bl_snapshot_begin queues a copy of the current U,
 * V, W, P, nu_t (FP64, unpadded, the extents above) into library-owned
 * pinned host memory on a second stream and returns at once; later steps may
 * be queued meanwhile.  bl_snapshot_end waits for that copy and returns
 * pointers into it, valid until the next bl_snapshot_begin or bl_destroy;
 * BL_ESTATE if no snapshot was begun.  Replaces the `!$acc update self` of
 * the fields before output_data (main.f90:100). */
int bl_snapshot_begin(bl_handle *h);
int bl_snapshot_end(bl_handle *h, const double **U, const double **V, const double **W, const double **P,
                    const double **nu_t);

/* ---- single operators, for parity checks against the reference ---- */

/* compute_rhs_u/v/w (equations.f90:24-415) of the current fields, written to
 * host arrays of the full U/V/W extents (interior ranges only). */
int bl_rhs(bl_handle *h, double *ru, double *rv, double *rw);
/* Set the step-start fields Uo, Vo, Wo used by the outflow condition
 * (time_integration.f90:314-337 copy). */
int bl_set_old(bl_handle *h, const double *Uo, const double *Vo, const double *Wo);
/* apply_boundary_conditions_gpu (boundary_conditions.f90:127-262) on the
 * current fields with rk_step and dt at the handle's time. */
int bl_apply_bc(bl_handle *h, int rk_step, double dt);
/* compute_projection_step (projection.f90:21-87, 293-386) on the current
 * fields; P is updated when rk_step == 1. */
int bl_projection(bl_handle *h, int rk_step, double dt);
/* Poisson solve alone (solve_poisson_gpu projection.f90:93-204): rhs and
 * result in a host array of the P extent; cells i=2..nxg-2, j=2..nyg-1,
 * k=2..nzg-2 are read and written. */
int bl_poisson(bl_handle *h, double *rhs_inout);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the maximum
 * divergence of check_divergence (projection.f90:391-413) of caller-owned
 * device arrays U(nx,nyg,nzg), V(nxg,ny,nzg), W(nxg,nyg,nz) in the layout
 * above, with the face coordinates x(nx), y(ny), z(nz) also on the device,
 * written to the device double div_out on `stream` (a hipStream_t; NULL is
 * the null stream), asynchronously.  No allocation: the reduction partials
 * and reciprocal spacings live in a file-scope device array of 1 MiB per
 * device; a grid that needs more (about 1.3e5 reduction blocks plus
 * nx + ny + nz) is rejected with BL_EINVAL.  Every call on a device shares
 * that array: calls that overlap on one device race. */
int bl_max_divergence_noalloc(const double *U, const double *V, const double *W, int nx, int ny, int nz,
                              const double *x, const double *y, const double *z, double *div_out, void *stream);

/* ---- the fftz module (fftz.f90:32-127) ----
 * Real DFT along z of xz planes with the solver's z-DFT kernels.  The
 * reference plans it at start-up (initialize_fftz, initialization.f90:720)
 * and never executes it on the time step.
 *  bl_fftz_create     initialize_fftz(N): N points in z
 *  bl_fft_z           compute_fft_z: U(n1, N) -> U_hat(n1, N/2+1), complex
 *                     as interleaved (re, im) doubles, unnormalised (r2c)
 *  bl_ifft_z          compute_ifft_z: U_hat -> U, divided by N (c2r; the
 *                     imaginary parts of the zero and, for even N, Nyquist
 *                     terms are ignored)
 *  bl_fftz_destroy    finalize_fftz
 * Host arrays, x (the n1 index) fastest, as the reference's. */
typedef struct bl_fftz bl_fftz;
int bl_fftz_create(int N, bl_fftz **out);
int bl_fft_z(bl_fftz *p, int n1, const double *U, double *U_hat);
int bl_ifft_z(bl_fftz *p, int n1, const double *U_hat, double *U);
void bl_fftz_destroy(bl_fftz *p);

/* ---- per-kernel timing (no counterpart in the reference, which prints a
 * system_clock sub-step profile, time_integration.f90:283-298) ---- */

/* When enabled, every kernel launch of the handle is bracketed by HIP events
 * on the handle's stream. */
int bl_profile_enable(bl_handle *h, int on);
/* The kernel classes that an enabled profile brackets: bit id of mask for
 * class id (default: all).  Fewer classes, fewer events in the stream: the
 * events cost about 0.3 ms per 814x257x193 RK3 step when all are on. */
int bl_profile_select(bl_handle *h, unsigned long long mask);
/* Accumulated milliseconds and launch counts per kernel class id 0..n-1
 * since profiling was enabled. */
int bl_profile_read(bl_handle *h, double *ms, long long *launches, int n);
/* Number of kernel classes and the name of class id. */
int bl_profile_count(void);
const char *bl_profile_name(int id);

/* ---- learned subgrid closure: training (no counterpart in the reference,
 * whose only dynamic model is subgrid.f90:261-414) ----
 * A 10 -> 64 -> 64 -> 64 -> 1 SiLU network for c = nu_t / (Delta^2 |S|),
 * FP32 master parameters in the layout W1[64][10], b1[64], W2[64][64], b2,
 * W3[64][64], b3, W4[64], b4 (BL_MLP_PARAMS floats), trained with Adam on a
 * device-resident dataset of BL_MLP_ROW-float rows: 10 features, target c,
 * weight, Delta^2 |S|, S_ij (6: 11 22 33 12 13 23), tau_ij (6). */
typedef struct bl_mlp bl_mlp;
enum { BL_MLP_IN = 10, BL_MLP_HIDDEN = 64, BL_MLP_PARAMS = 9089, BL_MLP_ROW = 28 };
int bl_mlp_create(long long capacity, int batch_max, bl_mlp **out);
void bl_mlp_destroy(bl_mlp *m);
/* set the FP32 master parameters (and reset the Adam moments) / read them */
int bl_mlp_set_params(bl_mlp *m, const float *theta);
int bl_mlp_get_params(bl_mlp *m, float *theta);
/* FP32 forward of n <= batch_max samples x[n][10] -> y[n] (training path) */
int bl_mlp_forward(bl_mlp *m, const float *x, int n, float *y);
/* gradient of mean((y - t)^2) over n samples, and that loss */
int bl_mlp_gradient(bl_mlp *m, const float *x, const float *t, int n, float *g, float *loss);
/* dataset: append host rows, read rows back, count, clear */
int bl_mlp_add_rows(bl_mlp *m, const float *rows, long long n);
int bl_mlp_get_rows(bl_mlp *m, long long first, long long n, float *rows);
int bl_mlp_count(bl_mlp *m, long long *n);
int bl_mlp_clear(bl_mlp *m);
/* steps of Adam (lr; betas 0.9, 0.999; eps 1e-8) on batches of `batch` rows
 * drawn by a counter hash of (seed, step, slot), minimising the weighted
 * loss (1/B) sum w (y - t)^2 with the row weights rescaled to mean 1 over the
 * dataset's positive-weight rows; loss_hist[steps] optional: that loss before
 * each update */
int bl_mlp_train(bl_mlp *m, int steps, int batch, float lr, unsigned long long seed, float *loss_hist);

/* ---- learned closure in the LES (les = 4; no counterpart in the reference) ----
 * Training rows from a DNS (fine, FP64) filtered onto an LES grid (coarse,
 * FP64) refined from it by (fx, fy, fz) cells with coinciding faces: box
 * (top-hat) averages of the DNS faces give the filtered staggered field,
 * which is left in the coarse solver's state; per interior LES cell the row
 * holds the features the LES computes at run time and the least-squares eddy
 * viscosity of the exact SGS stress.  Row weight: (Delta^2 |S|)^2 S:S, the
 * factor that turns the squared error in c into the squared deviatoric SGS
 * stress error; rows with |S| <= smin or next to the x boundaries get weight 0. */
int bl_closure_samples(bl_handle *fine, bl_handle *coarse, bl_mlp *m, int fx, int fy, int fz, double delta99,
                       double smin);
/* install the trained network in an les = 4 solver (BF16 weights) with the
 * input range [min - nsigma sd, max + nsigma sd] of the weighted rows; cells
 * outside it keep the dynamic Smagorinsky value */
int bl_set_closure(bl_handle *h, bl_mlp *m, double delta99, double nsigma);
/* cells evaluated and cells that fell back to DSM in the last evaluation */
int bl_closure_fallback(bl_handle *h, long long *cells, long long *fallback);
/* the BF16-MFMA inference kernel on feature rows x[n][10] (parity checks) */
int bl_closure_infer(bl_mlp *m, const float *x, long long n, float *y);

/* ---- host-side set-up helpers (not on the per-step path) ---- */

/* Blasius similarity profile f''' + f f'' = 0, f(0)=f'(0)=0, f''(0)=fpp0,
 * integrated by classical RK4 with `sub` sub-steps per output interval on
 * eta = 0..eta_max (n points).  The reference ships the table it makes with
 * fpp0 = 0.46960, eta_max = 20, n = 1000 (laminar_test/generate_blasius.py). */
int bl_blasius_profile(int n, double eta_max, double fpp0, int sub, double *eta, double *f,
                       double *df);
/* Boundary tables from a Blasius table, inflow_flag 1/3/4
 * (boundary_conditions.f90:1463-1627). */
int bl_blasius_tables(int nx, int ny, const double *x, const double *y, double nu, int n,
                      const double *eta, const double *f, const double *df, double *U_inlet,
                      double *V_inlet, double *W_inlet, double *U_top, double *V_top,
                      double *W_top);
/* Blasius initial field on any grid (init_flow input_output.f90:206-259). */
int bl_blasius_field(int nx, int ny, int nz, const double *x, const double *y, double nu, int n,
                     const double *eta, const double *f, const double *df, double *U, double *V,
                     double *W);
/* Uniform noise amp*(2r-1) on the interior of U, V, W, r from a counter-based
 * generator keyed by (seed, field, linear index): identical on any host. */
int bl_add_noise(int nx, int ny, int nz, double amp, unsigned long long seed, double *U, double *V,
                 double *W);
/* tanh-stretched grid: x = x0 + Lx*i/(nx-1), y = Ly*(1 - tanh(a(1-j/(ny-1)))/tanh a),
 * z = Lz*k/(nz-2). */
int bl_grid_tanh(int nx, int ny, int nz, double Lx, double Ly, double Lz, double alpha, double x0,
                 double *x, double *y, double *z);

#ifdef __cplusplus
}
#endif
#endif
