/*
 * CPU-ORACLE TEST INFRASTRUCTURE — not part of the product.
 * Shared state layout and 1-based index helpers of the oracle sources.
 */
#pragma once

#include <stddef.h>

#include "orc.h"

#define PI_D 3.141592653589793238462643383279502884

struct orc_state {
    int nx, ny, nz, nxg, nyg, nzg;
    int nxp, nzp, mx, mz, nm;
    double nu, dPdx, dPdz, cfl;
    int scheme, les;
    int nut_dirichlet; /* nutBC (params.f90:156): 1 = Dirichlet nu_t ghost in y */
    double *x, *y, *z, *xm, *ym, *zm, *xg, *yg, *zg;
    double *wy0, *wy1;
    double dxmin, dymin, dzmin, dx, dz;
    double *kxx, *kzz;
    double *dyy_d, *dyy_l, *dyy_u; /* Dyy(j,j), Dyy(j+1,j), Dyy(j,j+1), j=2..nyg-1 */
    double rk_t[4], rk_coef[4][4], rk2_t[3], rk2_coef[3][3];
    double *U_inlet, *V_inlet, *W_inlet, *U_top, *V_top, *W_top;
    int n_modes, m_modes;
    double *zmode, *tmode, *qu_re, *qu_im, *qv_re, *qv_im, *qw_re, *qw_im;
    double *Ut, *Vt, *Wt;
    double *U, *V, *W, *P, *Uo, *Vo, *Wo, *nu_t, *avg_nu_t;
    double *Fu[4], *Fv[4], *Fw[4];
    double *rhs_p;
    double t, dt;
    /* the general dispatcher's flags (orc_set_boundary) */
    int inflow_flag, top_flag, i_rescale, step_beginning;
    double amp, vbs_max, x_bs, sigma_bs, phi_bs, beta_hartree;
    unsigned long long seed;
    long long bc_step;
    /* Lund rescaling (inflow_flag 3, 5): parameters, the time-averaged mean
     * profiles (global.f90) and the rescaled inflow plane of the step */
    double delta_inlet, t_resc;
    double *Umean_resc_T, *Vmean_resc_T, *Umean_inlet_T, *Vmean_inlet_T;
    double *Ut_in, *Vt_in, *Wt_in;
    /* wall model (orc_wall.c) */
    int iwall, istress, nut_flag;
    double frac_vis;
    double alpha_mean_x, alpha_mean_y, alpha_mean_z, alpha_std, freq_mult, beta_y;
    double *alpha_x, *alpha_y, *alpha_z, *utau_model, *utau_ref;
};

/* bottom-wall conditions of apply_boundary_conditions (orc_wall.c) */
void orc__wall_bc(orc_state *s);
/* calloc of n doubles (n = 0: one) */
double *orc__dalloc(size_t n);

/* ---- 1-based index helpers (Fortran allocation shapes, x fastest) ---- */
#define X(i) s->x[(i)-1]
#define Y(j) s->y[(j)-1]
#define Z(k) s->z[(k)-1]
#define XG(i) s->xg[(i)-1]
#define YG(j) s->yg[(j)-1]
#define ZG(k) s->zg[(k)-1]
#define WY0(j) s->wy0[(j)-1]
#define WY1(j) s->wy1[(j)-1]

#define IU(i, j, k) ((size_t)((i)-1) + (size_t)s->nx * ((size_t)((j)-1) + (size_t)s->nyg * (size_t)((k)-1)))
#define IV(i, j, k) ((size_t)((i)-1) + (size_t)s->nxg * ((size_t)((j)-1) + (size_t)s->ny * (size_t)((k)-1)))
#define IW(i, j, k) ((size_t)((i)-1) + (size_t)s->nxg * ((size_t)((j)-1) + (size_t)s->nyg * (size_t)((k)-1)))
#define IP(i, j, k) IW(i, j, k)

