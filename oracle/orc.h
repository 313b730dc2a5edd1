/*
 * CPU-ORACLE TEST INFRASTRUCTURE — not part of the product.
 *
 * FP64 C restatement of the reference solver's per-time-step path, used
 * only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker.  Every function cites the reference file:line it follows.
 *
 * Array convention: every 3-D field is stored exactly as the reference
 * stores it — Fortran column-major, x fastest — at its full allocated
 * extent: U(nx,nyg,nzg), V(nxg,ny,nzg), W(nxg,nyg,nz), P(nxg,nyg,nzg).
 * 1-D grids are 0-based C arrays holding the Fortran 1-based entries
 * (x[i-1] is the reference's x(i)).
 */
#ifndef ORC_H
#define ORC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

/* Build the state from face grids x(nx), y(ny), z(nz) (initialization.f90:
 * 174-227, 254-263, 328-344, 394-425, 479-484, 503-530).  scheme = 2 or 3
 * (RKscheme), les = LES flag (0 DNS, 1 constant-coefficient Smagorinsky,
 * 2 dynamic Smagorinsky z-averaged, 3 dynamic Smagorinsky per cell;
 * subgrid.f90:32-45).  cfl is the .turbb CFL (negative = fixed dt).
 * Returns NULL for any other les. */
orc_state *orc_create(int nx, int ny, int nz, const double *x, const double *y, const double *z,
                      double nu, double dPdx, double dPdz, int scheme, int les, double cfl);
void orc_destroy(orc_state *s);

/* Boundary tables (compute_blasius_solution_for_bc, boundary_conditions.f90:
 * 1452-1797): U_inlet(nyg), V_inlet(ny), W_inlet(nyg), U_top(nx), V_top(nxg),
 * W_top(nxg). */
void orc_set_bc_tables(orc_state *s, const double *U_inlet, const double *V_inlet,
                       const double *W_inlet, const double *U_top, const double *V_top,
                       const double *W_top);
/* Temporal inflow modes already interpolated to the grid
 * (boundary_conditions.f90:1734-1776): q*_re/q*_im laid out
 * (j, n, m) Fortran order with j fastest; qu/qw over nyg, qv over ny. */
void orc_set_temporal_modes(orc_state *s, int n_modes, int m_modes, const double *zmode,
                            const double *tmode, const double *qu_re, const double *qu_im,
                            const double *qv_re, const double *qv_im, const double *qw_re,
                            const double *qw_im);

/* field: 0=U 1=V 2=W 3=P 4=nu_t ; copies the full allocated extent */
int orc_field_size(const orc_state *s, int field);
void orc_set_field(orc_state *s, int field, const double *src);
void orc_get_field(const orc_state *s, int field, double *dst);
void orc_set_time(orc_state *s, double t);
double orc_get_time(const orc_state *s);

/* force periodicity of the initial field (initialization.f90:244-246) */
void orc_apply_periodic_z(orc_state *s);

/* compute_dt (time_integration.f90:528-591); returns dt */
double orc_compute_dt(orc_state *s);
/* compute_time_step_RK2 / RK3 (time_integration.f90:145-300, 305-465) */
void orc_step(orc_state *s);
/* The general dispatcher apply_boundary_conditions (boundary_conditions.f90:
 * 22-121) for inflow_flag 1, 2 (Blasius + temporal modes), 4 (Blasius +
 * random perturbations once per step, :412-451) and top_flag 0 (Dirichlet,
 * :947-964), 1, 2 (blowing/suction :648-673 with zero_wz_top :264-283), 3
 * (Falkner-Skan :979-1007).  The random values are a counter hash of (seed,
 * step, component, j, k) in place of gfortran's random_number.  Returns 0, or
 * -1 for flags outside those.  Resets the step counter. */
int orc_set_boundary(orc_state *s, int inflow_flag, int top_flag, double amplitude, unsigned long long seed,
                     double vbs_max, double x_bs, double sigma_bs, double phi_bs, int i_rescale, double beta_hartree,
                     double delta_inlet, double t_resc);
/* inflow_flag 3 (Lund rescaling, apply_inflow_bc_x_rescaling
 * boundary_conditions.f90:569-597) and 5 (rescaled fluctuations on the
 * U/V/W_inlet mean, :512-559) via compute_rescaled_inflow
 * (rescaled_inlet_bc.f90:60-300), once per step from the step-start field;
 * the time-averaged means start from the field at orc_set_boundary
 * (initialization.f90:640-660).  Lund means (out): Umean_resc_T(nyg),
 * Vmean_resc_T(ny), Umean_inlet_T(nyg), Vmean_inlet_T(ny). */
void orc_lund_means(orc_state *s, double *umr, double *vmr, double *umi, double *vmi);
/* Wall model (wallmodel.f90:25-144) and its bottom-wall conditions
 * (boundary_conditions.f90:73-108): iwall_model 0, 4, 5, 6, 8, 9, 11, 12, 15
 * (9, 12: nut_flag = WMnutflag 0 or 2, frac_vis = WMnut);
 * istress_model 1 = law of the wall for utau_model.  alphas(5) of the .turbb
 * as amx, amy, amz, astd, freq_mult.  Resets the slip lengths to 0 and
 * utau_model to utau_ref (initialization.f90:745-746).  Returns 0, or -1
 * for other models.  orc_step computes the model once per RK stage after the
 * eddy viscosity (time_integration.f90:341-346) when LES > 0 or
 * iwall_model > 0; orc_wall_model does it once on the current fields at the
 * current t. */
int orc_set_wall_model(orc_state *s, int iwall, int istress, double amx, double amy, double amz, double astd,
                       double freq_mult, int nut_flag, double frac_vis);
void orc_wall_model(orc_state *s);
/* alpha_x(nx), alpha_y(nxg), alpha_z(nxg), utau_model(nx) */
void orc_get_wall_model(const orc_state *s, double *ax, double *ay, double *az, double *utau);
/* beta_y of the last pseudo-Robin wall condition (model 9 reads it) */
double orc_wall_beta(const orc_state *s);

/* compute_rhs_u/v/w (equations.f90:24-415) on the current U,V,W into
 * full-extent arrays ru(nx,nyg,nzg), rv(nxg,ny,nzg), rw(nxg,nyg,nz)
 * (only the interior ranges are written). */
void orc_rhs(orc_state *s, double *ru, double *rv, double *rw);
/* apply_boundary_conditions_gpu (boundary_conditions.f90:127-262) with the
 * given rk_step (selects rk_t) and current t, dt; Uo/Vo/Wo are taken from the
 * state's snapshot buffers. */
void orc_apply_bc(orc_state *s, int rk_step, double dt);
void orc_set_old(orc_state *s, const double *Uo, const double *Vo, const double *Wo);
/* compute_projection_step (projection.f90:21-87, 210-386) with rk_step */
void orc_projection(orc_state *s, int rk_step, double dt);
/* Poisson solve alone: rhs(nxg,nyg,nzg) interior ranges in, phi out
 * (solve_poisson_cpu projection.f90:210-288; only i=2..nxg-2, j=2..nyg-1,
 * k=2..nzg-2 are read and written). */
void orc_poisson(orc_state *s, double *rhs);
/* check_divergence (projection.f90:391-413) */
double orc_max_divergence(const orc_state *s);
/* Cf (statistics.f90:137-160) into cf(nx) */
void orc_wall_stats(const orc_state *s, double *cf);
/* the z-averaged statistics of compute_statistics (statistics.f90:44-80,
 * 125-135) on the current U, V, W, P, nu_t: means = Umean, Vmean, Wmean,
 * U2mean, V2mean, W2mean, UVmean, Pmean, P2mean, each (nx, ny), x fastest;
 * nu_t_mean (nxg, nyg), whose column nxg and row nyg compute_statistics never
 * fills (they stay 0) */
void orc_statistics(const orc_state *s, double *means, double *nu_t_mean);
/* monitor sums/maxima (monitor.f90:29-52): out[8] = sumU,sumV,sumW,sumNut,
 * maxU,maxV,maxW,maxNut */
void orc_monitor(const orc_state *s, double *out);
/* eddy viscosity (subgrid.f90:23-47) on the current U,V,W into nu_t */
void orc_eddy_viscosity(orc_state *s);
/* nutBC (params.f90:156): 1 = Dirichlet (antisymmetric) nu_t ghost in y, else Neumann */
void orc_set_nut_bc(orc_state *s, int dirichlet);

/* accessors for precomputed data used by tests */
void orc_get_kxx_kzz(const orc_state *s, double *kxx, double *kzz);
void orc_get_dyy(const orc_state *s, double *diag, double *lower, double *upper);
double orc_get_dt(const orc_state *s);

/* ---- host initialisation restatements (input_output.f90, boundary_conditions.f90) ---- */

/* compute_blasius_solution_for_bc part 1 (boundary_conditions.f90:1463-1627)
 * for inflow_flag 1/3/4 from the table (eta,f,df)[n].  Outputs as in
 * orc_set_bc_tables plus U_outlet(nyg), V_outlet(ny). */
void orc_blasius_tables(int nx, int ny, const double *x, const double *y, const double *xg,
                        const double *yg, double nu, int n, const double *eta, const double *f,
                        const double *df, double *U_inlet, double *V_inlet, double *W_inlet,
                        double *U_outlet, double *V_outlet, double *U_top, double *V_top,
                        double *W_top);
/* temporal-mode interpolation (boundary_conditions.f90:1734-1776): source
 * q_re/q_im (ny_inlet, n, m) onto yg (nyg) or y (ny) */
void orc_interp_modes(int ny_inlet, const double *ymesh, int nmodes_total, const double *q_re,
                      const double *q_im, int nout, const double *ygrid, double *out_re,
                      double *out_im);
/* init_flow grid + Blasius field (input_output.f90:130-279) */
void orc_init_flow(int nx, int ny, int nz, double Lx, double Ly, double Lz, double alpha,
                   double nu, int n, const double *eta, const double *f, const double *df,
                   double *x, double *y, double *z, double *U, double *V, double *W);
/* grid helpers: xg from x (initialization.f90:181-207) */
void orc_centers(int n, const double *x, double *xm, double *xg);
/* OpenMP threads of the oracle's parallel loops (the CPU-baseline timing) */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
