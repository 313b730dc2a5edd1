/*
 * CPU-ORACLE TEST INFRASTRUCTURE — not part of the product.
 *
 * FP64 restatement of the reference per-time-step path.  Loops, index ranges,
 * expression order and divisions follow the reference Fortran line by line so
 * that the file can be read side by side with it; each routine cites its
 * source.  1-based index macros mirror the Fortran arrays exactly.
 *
 * Parity status: see the header of each routine and tests/test_oracle_*.py
 * for the checks that pin it (closed forms, library routines, brute force,
 * reference fixtures).
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc_fft.h"
#include "orc_internal.h"

void orc_dsm(orc_state *s);          /* orc_dsm.c */
void orc_dsm_noavg(orc_state *s);    /* orc_dsm.c */
void orc_smagorinsky(orc_state *s);  /* orc_dsm.c */

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
double *orc__dalloc(size_t n) { return dalloc(n); }
static double sq(double a) { return a * a; }

static size_t nU(const orc_state *s) { return (size_t)s->nx * s->nyg * s->nzg; }
static size_t nV(const orc_state *s) { return (size_t)s->nxg * s->ny * s->nzg; }
static size_t nW(const orc_state *s) { return (size_t)s->nxg * s->nyg * s->nz; }
static size_t nP(const orc_state *s) { return (size_t)s->nxg * s->nyg * s->nzg; }

/* initialization.f90:181-207: interior centres and ghost-extended centres */
void orc_centers(int n, const double *x, double *xm, double *xg) {
    for (int i = 0; i < n - 1; i++) xm[i] = 0.5 * (x[i] + x[i + 1]);
    for (int i = 0; i < n - 1; i++) xg[i + 1] = xm[i];
    xg[0] = xm[0] - 2.0 * (xm[0] - x[0]);
    xg[n] = xm[n - 2] + 2.0 * (x[n - 1] - xm[n - 2]);
}

orc_state *orc_create(int nx, int ny, int nz, const double *x, const double *y, const double *z,
                      double nu, double dPdx, double dPdz, int scheme, int les, double cfl) {
    /* LES_model 0..3 (subgrid.f90:32-45); anything else has no model */
    if (les < 0 || les > 3) return NULL;
    orc_state *s = (orc_state *)calloc(1, sizeof(*s));
    s->inflow_flag = 1; /* the GPU path's set; orc_set_boundary changes it */
    s->step_beginning = 1;
    s->nx = nx;
    s->ny = ny;
    s->nz = nz;
    /* initialization.f90:84-104 (single rank) */
    s->nxg = nx + 1;
    s->nyg = ny + 1;
    s->nzg = nz + 1;
    s->nu = nu;
    s->dPdx = dPdx;
    s->dPdz = dPdz;
    s->scheme = scheme;
    s->les = les;
    s->cfl = cfl;
    s->x = dalloc(nx);
    s->y = dalloc(ny);
    s->z = dalloc(nz);
    memcpy(s->x, x, sizeof(double) * nx);
    memcpy(s->y, y, sizeof(double) * ny);
    memcpy(s->z, z, sizeof(double) * nz);
    s->xm = dalloc(nx - 1);
    s->ym = dalloc(ny - 1);
    s->zm = dalloc(nz - 1);
    s->xg = dalloc(nx + 1);
    s->yg = dalloc(ny + 1);
    s->zg = dalloc(nz + 1);
    orc_centers(nx, x, s->xm, s->xg);
    orc_centers(ny, y, s->ym, s->yg);
    orc_centers(nz, z, s->zm, s->zg);
    /* initialization.f90:220-222 */
    s->dxmin = 1e300;
    s->dymin = 1e300;
    s->dzmin = 1e300;
    for (int i = 0; i < s->nxg - 1; i++) s->dxmin = fmin(s->dxmin, s->xg[i + 1] - s->xg[i]);
    for (int i = 0; i < s->nyg - 1; i++) s->dymin = fmin(s->dymin, s->yg[i + 1] - s->yg[i]);
    for (int i = 0; i < s->nzg - 1; i++) s->dzmin = fmin(s->dzmin, s->zg[i + 1] - s->zg[i]);
    /* initialization.f90:254-278 */
    s->dx = s->dxmin;
    s->dz = s->dzmin;
    s->nxp = (nx - 1) - 1;
    s->nzp = (nz - 1) - 1;
    s->mx = s->nxp - 1;
    s->mz = s->nzp - 1;
    s->nm = s->nyg - 2;
    /* initialization.f90:329-344 */
    s->kxx = dalloc(s->mx + 1);
    s->kzz = dalloc(s->mz + 1);
    for (int i = 0; i <= s->mx; i++)
        s->kxx[i] = 2.0 * (cos(PI_D * (double)i / (double)s->nxp) - 1.0) / pow(s->dx, 2.0);
    int half = (s->nzp + 1) / 2; /* Ceiling(nzp/2) */
    for (int k = 0; k <= half && k <= s->mz; k++)
        s->kzz[k] = 2.0 * (cos(2.0 * PI_D * (double)k / (double)s->nzp) - 1.0) / pow(s->dz, 2.0);
    for (int k = half + 1; k <= s->mz; k++)
        s->kzz[k] = 2.0 * (cos(2.0 * PI_D * (double)(-s->nzp + k) / (double)s->nzp) - 1.0) / pow(s->dz, 2.0);
    /* Dyy, initialization.f90:394-425; stored for j = 2..nyg-1 at [j-2] */
    int nyg = s->nyg;
    s->dyy_d = dalloc(s->nm);
    s->dyy_l = dalloc(s->nm);
    s->dyy_u = dalloc(s->nm);
    for (int j = 2; j <= nyg - 1; j++) {
        double a = 1.0 / (Y(j) - Y(j - 1)) / (YG(j + 1) - YG(j));
        double b = 1.0 / (Y(j) - Y(j - 1)) * (-1.0 / (YG(j + 1) - YG(j)) - 1.0 / (YG(j) - YG(j - 1)));
        double c = 1.0 / (Y(j) - Y(j - 1)) / (YG(j) - YG(j - 1));
        if (j == 2) {
            s->dyy_d[j - 2] = b + c; /* Neumann p(1)=p(2) */
            s->dyy_u[j - 2] = a;
        } else if (j == nyg - 1) {
            s->dyy_d[j - 2] = a + b; /* Neumann p(nyg)=p(nyg-1) */
            s->dyy_l[j - 3] = c;     /* Dyy(nyg-1,nyg-2) */
        } else {
            s->dyy_d[j - 2] = b;
            s->dyy_u[j - 2] = a; /* Dyy(j,j+1) */
            s->dyy_l[j - 3] = c; /* Dyy(j,j-1) -> stored as lower of column j-1 */
        }
    }
    /* initialization.f90:482-484 */
    s->wy0 = dalloc(ny);
    s->wy1 = dalloc(ny);
    for (int j = 1; j <= ny; j++) {
        s->wy0[j - 1] = (YG(j + 1) - Y(j)) / (YG(j + 1) - YG(j));
        s->wy1[j - 1] = 1.0 - s->wy0[j - 1];
    }
    /* initialization.f90:503-530 */
    s->rk2_t[0] = 0.0;
    s->rk2_t[1] = 1.0 / 2.0;
    s->rk2_t[2] = 1.0;
    s->rk2_coef[1][1] = 1.0 / 2.0;
    s->rk2_coef[2][1] = 0.0;
    s->rk2_coef[2][2] = 1.0;
    s->rk_t[0] = 0.0;
    s->rk_t[1] = 8.0 / 15.0;
    s->rk_t[2] = 2.0 / 3.0;
    s->rk_t[3] = 1.0;
    s->rk_coef[1][1] = 8.0 / 15.0;
    s->rk_coef[2][1] = 1.0 / 4.0;
    s->rk_coef[2][2] = 5.0 / 12.0;
    s->rk_coef[3][1] = 1.0 / 4.0;
    s->rk_coef[3][2] = 0.0;
    s->rk_coef[3][3] = 3.0 / 4.0;
    /* fields and BC tables */
    s->U = dalloc(nU(s));
    s->V = dalloc(nV(s));
    s->W = dalloc(nW(s));
    s->P = dalloc(nP(s));
    s->Uo = dalloc(nU(s));
    s->Vo = dalloc(nV(s));
    s->Wo = dalloc(nW(s));
    s->nu_t = dalloc(nP(s));
    s->avg_nu_t = dalloc((size_t)s->nxg * s->nyg);
    s->rhs_p = dalloc(nP(s));
    for (int st = 1; st <= 3; st++) {
        s->Fu[st] = dalloc(nU(s));
        s->Fv[st] = dalloc(nV(s));
        s->Fw[st] = dalloc(nW(s));
    }
    s->U_inlet = dalloc(nyg);
    s->V_inlet = dalloc(ny);
    s->W_inlet = dalloc(nyg);
    s->U_top = dalloc(nx);
    s->V_top = dalloc(s->nxg);
    s->W_top = dalloc(s->nxg);
    s->Ut = dalloc((size_t)nyg * s->nzg);
    s->Vt = dalloc((size_t)ny * s->nzg);
    s->Wt = dalloc((size_t)nyg * nz);
    return s;
}

void orc_destroy(orc_state *s) {
    if (!s) return;
    double *ptrs[] = {s->x, s->y, s->z, s->xm, s->ym, s->zm, s->xg, s->yg, s->zg, s->wy0,
                      s->wy1, s->kxx, s->kzz, s->dyy_d, s->dyy_l, s->dyy_u, s->U_inlet,
                      s->V_inlet, s->W_inlet, s->U_top, s->V_top, s->W_top, s->zmode,
                      s->tmode, s->qu_re, s->qu_im, s->qv_re, s->qv_im, s->qw_re, s->qw_im,
                      s->Ut, s->Vt, s->Wt, s->U, s->V, s->W, s->P, s->Uo, s->Vo, s->Wo,
                      s->nu_t, s->avg_nu_t, s->rhs_p, s->Umean_resc_T, s->Vmean_resc_T, s->Umean_inlet_T,
                      s->Vmean_inlet_T, s->Ut_in, s->Vt_in, s->Wt_in, s->alpha_x, s->alpha_y,
                      s->alpha_z, s->utau_model, s->utau_ref};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    for (int st = 1; st <= 3; st++) {
        free(s->Fu[st]);
        free(s->Fv[st]);
        free(s->Fw[st]);
    }
    free(s);
}

void orc_set_bc_tables(orc_state *s, const double *U_inlet, const double *V_inlet,
                       const double *W_inlet, const double *U_top, const double *V_top,
                       const double *W_top) {
    memcpy(s->U_inlet, U_inlet, sizeof(double) * s->nyg);
    memcpy(s->V_inlet, V_inlet, sizeof(double) * s->ny);
    memcpy(s->W_inlet, W_inlet, sizeof(double) * s->nyg);
    memcpy(s->U_top, U_top, sizeof(double) * s->nx);
    memcpy(s->V_top, V_top, sizeof(double) * s->nxg);
    memcpy(s->W_top, W_top, sizeof(double) * s->nxg);
}

void orc_set_temporal_modes(orc_state *s, int n_modes, int m_modes, const double *zmode,
                            const double *tmode, const double *qu_re, const double *qu_im,
                            const double *qv_re, const double *qv_im, const double *qw_re,
                            const double *qw_im) {
    s->n_modes = n_modes;
    s->m_modes = m_modes;
    size_t nm = (size_t)n_modes * m_modes;
    s->zmode = dalloc(n_modes);
    s->tmode = dalloc(m_modes);
    memcpy(s->zmode, zmode, sizeof(double) * n_modes);
    memcpy(s->tmode, tmode, sizeof(double) * m_modes);
    double **dst[6] = {&s->qu_re, &s->qu_im, &s->qv_re, &s->qv_im, &s->qw_re, &s->qw_im};
    const double *src[6] = {qu_re, qu_im, qv_re, qv_im, qw_re, qw_im};
    for (int c = 0; c < 6; c++) {
        size_t n = nm * (size_t)((c == 2 || c == 3) ? s->ny : s->nyg);
        *dst[c] = dalloc(n);
        memcpy(*dst[c], src[c], sizeof(double) * n);
    }
}

int orc_field_size(const orc_state *s, int field) {
    switch (field) {
    case 0: return (int)nU(s);
    case 1: return (int)nV(s);
    case 2: return (int)nW(s);
    default: return (int)nP(s);
    }
}

static double *field_ptr(const orc_state *s, int field) {
    switch (field) {
    case 0: return s->U;
    case 1: return s->V;
    case 2: return s->W;
    case 3: return s->P;
    default: return s->nu_t;
    }
}

void orc_set_field(orc_state *s, int field, const double *src) {
    memcpy(field_ptr(s, field), src, sizeof(double) * (size_t)orc_field_size(s, field));
}
void orc_get_field(const orc_state *s, int field, double *dst) {
    memcpy(dst, field_ptr(s, field), sizeof(double) * (size_t)orc_field_size(s, field));
}
void orc_set_time(orc_state *s, double t) { s->t = t; }
double orc_get_time(const orc_state *s) { return s->t; }
double orc_get_dt(const orc_state *s) { return s->dt; }

void orc_set_old(orc_state *s, const double *Uo, const double *Vo, const double *Wo) {
    memcpy(s->Uo, Uo, sizeof(double) * nU(s));
    memcpy(s->Vo, Vo, sizeof(double) * nV(s));
    memcpy(s->Wo, Wo, sizeof(double) * nW(s));
}

void orc_get_kxx_kzz(const orc_state *s, double *kxx, double *kzz) {
    memcpy(kxx, s->kxx, sizeof(double) * (s->mx + 1));
    memcpy(kzz, s->kzz, sizeof(double) * (s->mz + 1));
}
void orc_get_dyy(const orc_state *s, double *diag, double *lower, double *upper) {
    memcpy(diag, s->dyy_d, sizeof(double) * s->nm);
    memcpy(lower, s->dyy_l, sizeof(double) * s->nm);
    memcpy(upper, s->dyy_u, sizeof(double) * s->nm);
}

/* apply_periodic_bc_z single-rank branch (boundary_conditions.f90:739-757) */
static void periodic_z_cc(orc_state *s, double *F, int n1, int n2) {
    int nzg = s->nzg;
    size_t pl = (size_t)n1 * n2;
    memcpy(F + 0 * pl, F + (size_t)(nzg - 3) * pl, sizeof(double) * pl);
    memcpy(F + (size_t)(nzg - 2) * pl, F + 1 * pl, sizeof(double) * pl);
    memcpy(F + (size_t)(nzg - 1) * pl, F + 2 * pl, sizeof(double) * pl);
}
static void periodic_z_w(orc_state *s, double *F) {
    int nz = s->nz;
    size_t pl = (size_t)s->nxg * s->nyg;
    memcpy(F + 0 * pl, F + (size_t)(nz - 2) * pl, sizeof(double) * pl);
    memcpy(F + (size_t)(nz - 1) * pl, F + 1 * pl, sizeof(double) * pl);
}

void orc_apply_periodic_z(orc_state *s) {
    periodic_z_cc(s, s->U, s->nx, s->nyg);
    periodic_z_cc(s, s->V, s->nxg, s->ny);
    periodic_z_w(s, s->W);
}

/* ------------------------------------------------------------------ */
/* compute_rhs_u (equations.f90:24-156)                                */
/* ------------------------------------------------------------------ */
static void rhs_u(orc_state *s, const double *U, const double *V, const double *W, double *ru) {
    int nx = s->nx, nyg = s->nyg, nzg = s->nzg;
    const double nu = s->nu, dPdx = s->dPdx;
    const double *nu_t = s->nu_t;
#define U_(i, j, k) U[IU(i, j, k)]
#define V_(i, j, k) V[IV(i, j, k)]
#define W_(i, j, k) W[IW(i, j, k)]
#define NT(i, j, k) nu_t[IP(i, j, k)]
#define R(i, j, k) ru[IU(i, j, k)]
    /* convective terms, equations.f90:47-67 */
#pragma omp parallel for
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                double r = -(sq(0.5 * (U_(i, j, k) + U_(i + 1, j, k))) -
                             sq(0.5 * (U_(i - 1, j, k) + U_(i, j, k)))) /
                           (XG(i + 1) - XG(i));
                r = r - ((WY0(j) * U_(i, j, k) + WY1(j) * U_(i, j + 1, k)) * 0.5 *
                             (V_(i, j, k) + V_(i + 1, j, k)) -
                         (WY0(j - 1) * U_(i, j - 1, k) + WY1(j - 1) * U_(i, j, k)) * 0.5 *
                             (V_(i, j - 1, k) + V_(i + 1, j - 1, k))) /
                            (Y(j) - Y(j - 1));
                r = r - (0.5 * (U_(i, j, k) + U_(i, j, k + 1)) * 0.5 * (W_(i, j, k) + W_(i + 1, j, k)) -
                         0.5 * (U_(i, j, k - 1) + U_(i, j, k)) * 0.5 *
                             (W_(i, j, k - 1) + W_(i + 1, j, k - 1))) /
                            (Z(k) - Z(k - 1));
                R(i, j, k) = r;
            }
    if (s->les == 0) {
        /* DNS viscous + dPdx, equations.f90:76-94 */
#pragma omp parallel for
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nx - 1; i++) {
                    double dx_1 = X(i) - X(i - 1);
                    double dx_2 = X(i + 1) - X(i);
                    double dx_3 = XG(i + 1) - XG(i);
                    double dy_3 = Y(j) - Y(j - 1);
                    double dz_1 = ZG(k) - ZG(k - 1);
                    double dz_2 = ZG(k + 1) - ZG(k);
                    double dz_3 = Z(k) - Z(k - 1);
                    R(i, j, k) = R(i, j, k) + dPdx +
                                 nu * ((1.0 / dx_2 * (U_(i + 1, j, k) - U_(i, j, k)) -
                                        1.0 / dx_1 * (U_(i, j, k) - U_(i - 1, j, k))) /
                                           dx_3 +
                                       ((U_(i, j + 1, k) - U_(i, j, k)) / (YG(j + 1) - YG(j)) -
                                        (U_(i, j, k) - U_(i, j - 1, k)) / (YG(j) - YG(j - 1))) /
                                           dy_3 +
                                       (1.0 / dz_2 * (U_(i, j, k + 1) - U_(i, j, k)) -
                                        1.0 / dz_1 * (U_(i, j, k) - U_(i, j, k - 1))) /
                                           dz_3);
                }
    } else {
        /* LES viscous, equations.f90:97-132 (term_2 inlined) */
#define T2(i, j, k) ((U_(i, (j) + 1, k) - U_(i, j, k)) / (YG((j) + 1) - YG(j)))
#pragma omp parallel for
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nx - 1; i++) {
                    double dz_1 = ZG(k) - ZG(k - 1);
                    double dz_2 = ZG(k + 1) - ZG(k);
                    double dz_3 = Z(k) - Z(k - 1);
                    double dy_3 = Y(j) - Y(j - 1);
                    double dx_1 = X(i) - X(i - 1);
                    double dx_2 = X(i + 1) - X(i);
                    double dx_3 = XG(i + 1) - XG(i);
                    double nu_x1 = nu + NT(i, j, k);
                    double nu_x2 = nu + NT(i + 1, j, k);
                    double nu_y1 = nu + 0.5 * (WY0(j - 1) * NT(i, j - 1, k) + WY1(j - 1) * NT(i, j, k) +
                                               WY0(j - 1) * NT(i + 1, j - 1, k) + WY1(j - 1) * NT(i + 1, j, k));
                    double nu_y2 = nu + 0.5 * (WY0(j) * NT(i, j, k) + WY1(j) * NT(i, j + 1, k) +
                                               WY0(j) * NT(i + 1, j, k) + WY1(j) * NT(i + 1, j + 1, k));
                    double nu_z1 = nu + 0.25 * (NT(i, j, k) + NT(i, j, k - 1) + NT(i + 1, j, k) + NT(i + 1, j, k - 1));
                    double nu_z2 = nu + 0.25 * (NT(i, j, k) + NT(i, j, k + 1) + NT(i + 1, j, k) + NT(i + 1, j, k + 1));
                    R(i, j, k) = R(i, j, k) + dPdx +
                                 (nu_x2 / dx_2 * (U_(i + 1, j, k) - U_(i, j, k)) -
                                  nu_x1 / dx_1 * (U_(i, j, k) - U_(i - 1, j, k))) /
                                     dx_3 +
                                 (nu_y2 * T2(i, j, k) - nu_y1 * T2(i, j - 1, k)) / dy_3 +
                                 (nu_z2 / dz_2 * (U_(i, j, k + 1) - U_(i, j, k)) -
                                  nu_z1 / dz_1 * (U_(i, j, k) - U_(i, j, k - 1))) /
                                     dz_3;
                }
#undef T2
    }
    /* last plane is dummy, equations.f90:149-154 */
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++) R(nx - 1, j, k) = 0.0;
#undef R
}

/* compute_rhs_v (equations.f90:163-294) */
static void rhs_v(orc_state *s, const double *U, const double *V, const double *W, double *rv) {
    int nxg = s->nxg, ny = s->ny, nzg = s->nzg;
    const double nu = s->nu;
    const double *nu_t = s->nu_t;
#define R(i, j, k) rv[IV(i, j, k)]
#pragma omp parallel for
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                double r = -(sq(0.5 * (V_(i, j, k) + V_(i, j + 1, k))) -
                             sq(0.5 * (V_(i, j - 1, k) + V_(i, j, k)))) /
                           (YG(j + 1) - YG(j));
                r = r - ((WY0(j) * U_(i, j, k) + WY1(j) * U_(i, j + 1, k)) * 0.5 *
                             (V_(i, j, k) + V_(i + 1, j, k)) -
                         (WY0(j) * U_(i - 1, j, k) + WY1(j) * U_(i - 1, j + 1, k)) * 0.5 *
                             (V_(i - 1, j, k) + V_(i, j, k))) /
                            (X(i) - X(i - 1));
                r = r - (0.5 * (V_(i, j, k) + V_(i, j, k + 1)) *
                             (WY0(j) * W_(i, j, k) + WY1(j) * W_(i, j + 1, k)) -
                         0.5 * (V_(i, j, k - 1) + V_(i, j, k)) *
                             (WY0(j) * W_(i, j, k - 1) + WY1(j) * W_(i, j + 1, k - 1))) /
                            (Z(k) - Z(k - 1));
                R(i, j, k) = r;
            }
    if (s->les == 0) {
        /* equations.f90:208-230 */
#pragma omp parallel for
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    double dx_1 = XG(i) - XG(i - 1);
                    double dx_2 = XG(i + 1) - XG(i);
                    double dx_3 = X(i) - X(i - 1);
                    double dy_1 = Y(j) - Y(j - 1);
                    double dy_2 = Y(j + 1) - Y(j);
                    double dy_3 = YG(j + 1) - YG(j);
                    double dz_1 = ZG(k) - ZG(k - 1);
                    double dz_2 = ZG(k + 1) - ZG(k);
                    double dz_3 = Z(k) - Z(k - 1);
                    R(i, j, k) = R(i, j, k) +
                                 nu * ((1.0 / dx_2 * (V_(i + 1, j, k) - V_(i, j, k)) -
                                        1.0 / dx_1 * (V_(i, j, k) - V_(i - 1, j, k))) /
                                           dx_3 +
                                       (1.0 / dy_2 * (V_(i, j + 1, k) - V_(i, j, k)) -
                                        1.0 / dy_1 * (V_(i, j, k) - V_(i, j - 1, k))) /
                                           dy_3 +
                                       (1.0 / dz_2 * (V_(i, j, k + 1) - V_(i, j, k)) -
                                        1.0 / dz_1 * (V_(i, j, k) - V_(i, j, k - 1))) /
                                           dz_3);
                }
    } else {
        /* equations.f90:231-270 */
#pragma omp parallel for
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    double dz_1 = ZG(k) - ZG(k - 1);
                    double dz_2 = ZG(k + 1) - ZG(k);
                    double dz_3 = Z(k) - Z(k - 1);
                    double dy_1 = Y(j) - Y(j - 1);
                    double dy_2 = Y(j + 1) - Y(j);
                    double dy_3 = YG(j + 1) - YG(j);
                    double dx_1 = XG(i) - XG(i - 1);
                    double dx_2 = XG(i + 1) - XG(i);
                    double dx_3 = X(i) - X(i - 1);
                    double nu_x1 = nu + 0.5 * (WY0(j) * NT(i - 1, j, k) + WY1(j) * NT(i - 1, j + 1, k) +
                                               WY0(j) * NT(i, j, k) + WY1(j) * NT(i, j + 1, k));
                    double nu_x2 = nu + 0.5 * (WY0(j) * NT(i, j, k) + WY1(j) * NT(i, j + 1, k) +
                                               WY0(j) * NT(i + 1, j, k) + WY1(j) * NT(i + 1, j + 1, k));
                    double nu_y1 = nu + NT(i, j, k);
                    double nu_y2 = nu + NT(i, j + 1, k);
                    double nu_z1 = nu + 0.5 * (WY0(j) * NT(i, j, k - 1) + WY1(j) * NT(i, j + 1, k - 1) +
                                               WY0(j) * NT(i, j, k) + WY1(j) * NT(i, j + 1, k));
                    double nu_z2 = nu + 0.5 * (WY0(j) * NT(i, j, k) + WY1(j) * NT(i, j + 1, k) +
                                               WY0(j) * NT(i, j, k + 1) + WY1(j) * NT(i, j + 1, k + 1));
                    R(i, j, k) = R(i, j, k) +
                                 1.0 / dx_3 *
                                     (nu_x2 * 1.0 / dx_2 * (V_(i + 1, j, k) - V_(i, j, k)) -
                                      nu_x1 * 1.0 / dx_1 * (V_(i, j, k) - V_(i - 1, j, k))) +
                                 1.0 / dy_3 *
                                     (nu_y2 * 1.0 / dy_2 * (V_(i, j + 1, k) - V_(i, j, k)) -
                                      nu_y1 * 1.0 / dy_1 * (V_(i, j, k) - V_(i, j - 1, k))) +
                                 1.0 / dz_3 *
                                     (nu_z2 * 1.0 / dz_2 * (V_(i, j, k + 1) - V_(i, j, k)) -
                                      nu_z1 * 1.0 / dz_1 * (V_(i, j, k) - V_(i, j, k - 1)));
                }
    }
    /* equations.f90:287-292 */
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= ny - 1; j++) R(nxg - 1, j, k) = 0.0;
#undef R
}

/* compute_rhs_w (equations.f90:301-415) */
static void rhs_w(orc_state *s, const double *U, const double *V, const double *W, double *rw) {
    int nxg = s->nxg, nyg = s->nyg, nz = s->nz;
    const double nu = s->nu, dPdz = s->dPdz;
    const double *nu_t = s->nu_t;
#define R(i, j, k) rw[IW(i, j, k)]
#pragma omp parallel for
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                double r = -(sq(0.5 * (W_(i, j, k) + W_(i, j, k + 1))) -
                             sq(0.5 * (W_(i, j, k - 1) + W_(i, j, k)))) /
                           (ZG(k + 1) - ZG(k));
                r = r - (0.5 * (U_(i, j, k) + U_(i, j, k + 1)) * 0.5 * (W_(i, j, k) + W_(i + 1, j, k)) -
                         0.5 * (U_(i - 1, j, k) + U_(i - 1, j, k + 1)) * 0.5 *
                             (W_(i - 1, j, k) + W_(i, j, k))) /
                            (X(i) - X(i - 1));
                r = r - (0.5 * (V_(i, j, k) + V_(i, j, k + 1)) *
                             (WY0(j) * W_(i, j, k) + WY1(j) * W_(i, j + 1, k)) -
                         0.5 * (V_(i, j - 1, k) + V_(i, j - 1, k + 1)) *
                             (WY0(j - 1) * W_(i, j - 1, k) + WY1(j - 1) * W_(i, j, k))) /
                            (Y(j) - Y(j - 1));
                R(i, j, k) = r;
            }
    if (s->les == 0) {
        /* equations.f90:346-366 */
#pragma omp parallel for
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    double dx_1 = XG(i) - XG(i - 1);
                    double dx_2 = XG(i + 1) - XG(i);
                    double dx_3 = X(i) - X(i - 1);
                    double dy_3 = Y(j) - Y(j - 1);
                    double dz_1 = Z(k) - Z(k - 1);
                    double dz_2 = Z(k + 1) - Z(k);
                    double dz_3 = ZG(k + 1) - ZG(k);
                    R(i, j, k) = R(i, j, k) + dPdz +
                                 nu * ((1.0 / dx_2 * (W_(i + 1, j, k) - W_(i, j, k)) -
                                        1.0 / dx_1 * (W_(i, j, k) - W_(i - 1, j, k))) /
                                           dx_3 +
                                       ((W_(i, j + 1, k) - W_(i, j, k)) / (YG(j + 1) - YG(j)) -
                                        (W_(i, j, k) - W_(i, j - 1, k)) / (YG(j) - YG(j - 1))) /
                                           dy_3 +
                                       (1.0 / dz_2 * (W_(i, j, k + 1) - W_(i, j, k)) -
                                        1.0 / dz_1 * (W_(i, j, k) - W_(i, j, k - 1))) /
                                           dz_3);
                }
    } else {
        /* equations.f90:367-404 (the interpolate_y(nu_t) at :377 is dead) */
#define T2(i, j, k) ((W_(i, (j) + 1, k) - W_(i, j, k)) / (YG((j) + 1) - YG(j)))
#pragma omp parallel for
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    double dz_1 = Z(k) - Z(k - 1);
                    double dz_2 = Z(k + 1) - Z(k);
                    double dz_3 = ZG(k + 1) - ZG(k);
                    double dy_3 = Y(j) - Y(j - 1);
                    double dx_1 = XG(i) - XG(i - 1);
                    double dx_2 = XG(i + 1) - XG(i);
                    double dx_3 = X(i) - X(i - 1);
                    double nu_x1 = nu + 0.25 * (NT(i, j, k) + NT(i, j, k + 1) + NT(i - 1, j, k) + NT(i - 1, j, k + 1));
                    double nu_x2 = nu + 0.25 * (NT(i, j, k) + NT(i, j, k + 1) + NT(i + 1, j, k) + NT(i + 1, j, k + 1));
                    double nu_y1 = nu + 0.5 * (WY0(j - 1) * NT(i, j - 1, k) + WY1(j - 1) * NT(i, j, k) +
                                               WY0(j - 1) * NT(i, j - 1, k + 1) + WY1(j - 1) * NT(i, j, k + 1));
                    double nu_y2 = nu + 0.5 * (WY0(j) * NT(i, j, k) + WY1(j) * NT(i, j + 1, k) +
                                               WY0(j) * NT(i, j, k + 1) + WY1(j) * NT(i, j + 1, k + 1));
                    double nu_z1 = nu + NT(i, j, k);
                    double nu_z2 = nu + NT(i, j, k + 1);
                    R(i, j, k) = R(i, j, k) + dPdz +
                                 (nu_x2 / dx_2 * (W_(i + 1, j, k) - W_(i, j, k)) -
                                  nu_x1 / dx_1 * (W_(i, j, k) - W_(i - 1, j, k))) /
                                     dx_3 +
                                 (nu_y2 * T2(i, j, k) - nu_y1 * T2(i, j - 1, k)) / dy_3 +
                                 (nu_z2 / dz_2 * (W_(i, j, k + 1) - W_(i, j, k)) -
                                  nu_z1 / dz_1 * (W_(i, j, k) - W_(i, j, k - 1))) /
                                     dz_3;
                }
#undef T2
    }
    /* equations.f90:408-413 */
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= nyg - 1; j++) R(nxg - 1, j, k) = 0.0;
#undef R
}

void orc_rhs(orc_state *s, double *ru, double *rv, double *rw) {
    rhs_u(s, s->U, s->V, s->W, ru);
    rhs_v(s, s->U, s->V, s->W, rv);
    rhs_w(s, s->U, s->V, s->W, rw);
}

int orc_set_boundary(orc_state *s, int inflow_flag, int top_flag, double amplitude, unsigned long long seed,
                     double vbs_max, double x_bs, double sigma_bs, double phi_bs, int i_rescale, double beta_hartree,
                     double delta_inlet, double t_resc) {
    if (inflow_flag < 1 || inflow_flag > 5) return -1;
    if (top_flag < 0 || top_flag > 3) return -1;
    if ((inflow_flag == 3 || inflow_flag == 5) &&
        (i_rescale < 2 || i_rescale > s->nx - 1 || !(delta_inlet > 0.0) || !(t_resc > 0.0)))
        return -1;
    s->inflow_flag = inflow_flag;
    s->top_flag = top_flag;
    s->amp = amplitude;
    s->seed = seed;
    s->vbs_max = vbs_max;
    s->x_bs = x_bs;
    s->sigma_bs = sigma_bs;
    s->phi_bs = phi_bs;
    s->i_rescale = i_rescale;
    s->beta_hartree = beta_hartree;
    s->delta_inlet = delta_inlet;
    s->t_resc = t_resc;
    s->bc_step = 0;
    s->step_beginning = 1;
    if (inflow_flag == 3 || inflow_flag == 5) {
        int nyg = s->nyg, ny = s->ny, nzg = s->nzg;
        if (!s->Umean_resc_T) {
            s->Umean_resc_T = dalloc(nyg);
            s->Vmean_resc_T = dalloc(ny);
            s->Umean_inlet_T = dalloc(nyg);
            s->Vmean_inlet_T = dalloc(ny);
            s->Ut_in = dalloc((size_t)nyg * nzg);
            s->Vt_in = dalloc((size_t)ny * nzg);
            s->Wt_in = dalloc((size_t)nyg * s->nz);
        }
        /* first-time means from the field, initialization.f90:640-660 */
        const double *U = s->U, *V = s->V;
        for (int j = 1; j <= nyg; j++) {
            double a = 0.0, b = 0.0;
            for (int k = 2; k <= nzg - 1; k++) {
                a += U[IU(i_rescale, j, k)];
                b += U[IU(1, j, k)];
            }
            s->Umean_resc_T[j - 1] = a / (double)(nzg - 2);
            s->Umean_inlet_T[j - 1] = b / (double)(nzg - 2);
        }
        for (int j = 1; j <= ny; j++) {
            double a = 0.0, b = 0.0;
            for (int k = 2; k <= nzg - 1; k++) {
                a += V[IV(i_rescale, j, k)];
                b += V[IV(1, j, k)];
            }
            s->Vmean_resc_T[j - 1] = a / (double)(nzg - 2);
            s->Vmean_inlet_T[j - 1] = b / (double)(nzg - 2);
        }
    }
    return 0;
}

void orc_lund_means(orc_state *s, double *umr, double *vmr, double *umi, double *vmi) {
    if (!s->Umean_resc_T) return;
    memcpy(umr, s->Umean_resc_T, sizeof(double) * s->nyg);
    memcpy(vmr, s->Vmean_resc_T, sizeof(double) * s->ny);
    memcpy(umi, s->Umean_inlet_T, sizeof(double) * s->nyg);
    memcpy(vmi, s->Vmean_inlet_T, sizeof(double) * s->ny);
}

/* interp_vel (rescaled_inlet_bc.f90:305-345) of `cols` columns of a profile
 * f(n, cols) given at y_ to the points y_i (both n long), rows 2..n; row 1
 * (the ghost, left unset there and overwritten by the wall condition) is 0 */
static void interp_vel(int n, int cols, const double *f, const double *y_, const double *y_i, double *out) {
    for (int c = 0; c < cols; c++) out[(size_t)n * c] = 0.0;
    for (int i = 2; i <= n; i++) {
        int j = 0;
        for (int k = 1; k <= n; k++)
            if (y_[k - 1] > y_i[i - 1]) {
                j = k;
                break;
            }
        for (int c = 0; c < cols; c++) {
            double *o = out + (size_t)n * c;
            const double *g = f + (size_t)n * c;
            if (j > 1) {
                double coef = (y_[j - 1] - y_i[i - 1]) / (y_[j - 1] - y_[j - 2]);
                o[i - 1] = coef * g[j - 2] + (1.0 - coef) * g[j - 1];
            } else {
                o[i - 1] = g[n - 1]; /* default: the top value (j = 1 stops the reference) */
            }
        }
    }
}

/* the first j (1-based) with m(j) >= level, nyg/2 if below 2, and the
 * interpolated height (rescaled_inlet_bc.f90:144-175) */
static double height99(int nyg, const double *m, const double *yg, double level) {
    int jref = 0;
    for (int j = 1; j <= nyg; j++)
        if (m[j - 1] >= level) {
            jref = j;
            break;
        }
    if (jref < 2) jref = nyg / 2;
    double w1 = (level - m[jref - 2]) / (m[jref - 1] - m[jref - 2]);
    double w0 = 1.0 - w1;
    double d = w1 * yg[jref - 1] + w0 * yg[jref - 2];
    return d < 0 ? fabs(d) : d;
}

static double theta_of(int nyg, const double *m, const double *yg, double uinf) {
    double th = 0.0;
    for (int j = 2; j <= nyg; j++) {
        double a = m[j - 1] / uinf * (1.0 - m[j - 1] / uinf), b = m[j - 2] / uinf * (1.0 - m[j - 2] / uinf);
        th = th + 0.5 * (a + b) * (yg[j - 1] - yg[j - 2]);
    }
    return th < 0 ? fabs(th) : th;
}

/* compute_rescaled_inflow (rescaled_inlet_bc.f90:60-300) from the
 * step-start field into Ut_in(nyg,nzg), Vt_in(ny,nzg), Wt_in(nyg,nz) */
static void lund_inflow(orc_state *s, int iflag_fluc) {
    int ny = s->ny, nz = s->nz, nyg = s->nyg, nzg = s->nzg, ir = s->i_rescale;
    const double *U = s->Uo, *V = s->Vo, *W = s->Wo;
    double nu = s->nu, dt = s->dt, tr = s->t_resc;
    double *umr = dalloc(nyg), *vmr = dalloc(ny), *umi = dalloc(nyg), *vmi = dalloc(ny);
    for (int j = 1; j <= nyg; j++) {
        double a = 0.0, b = 0.0;
        for (int k = 2; k <= nzg - 1; k++) {
            a += U[IU(ir, j, k)];
            b += U[IU(1, j, k)];
        }
        umr[j - 1] = a / (double)(nzg - 2);
        umi[j - 1] = b / (double)(nzg - 2);
    }
    for (int j = 1; j <= ny; j++) {
        double a = 0.0, b = 0.0;
        for (int k = 2; k <= nzg - 1; k++) {
            a += V[IV(ir, j, k)];
            b += V[IV(1, j, k)];
        }
        vmr[j - 1] = a / (double)(nzg - 2);
        vmi[j - 1] = b / (double)(nzg - 2);
    }
    double Uinf = umr[nyg - 1];
    for (int j = 0; j < nyg; j++) {
        s->Umean_resc_T[j] = dt / tr * umr[j] + (1.0 - dt / tr) * s->Umean_resc_T[j];
        s->Umean_inlet_T[j] = dt / tr * umi[j] + (1.0 - dt / tr) * s->Umean_inlet_T[j];
    }
    for (int j = 0; j < ny; j++) {
        s->Vmean_resc_T[j] = dt / tr * vmr[j] + (1.0 - dt / tr) * s->Vmean_resc_T[j];
        s->Vmean_inlet_T[j] = dt / tr * vmi[j] + (1.0 - dt / tr) * s->Vmean_inlet_T[j];
    }
    /* fluctuations of the rescaling plane (W has no mean removed) */
    double *uf = dalloc((size_t)nyg * nzg), *vf = dalloc((size_t)ny * nzg), *wf = dalloc((size_t)nyg * nz);
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++) uf[(j - 1) + (size_t)nyg * (k - 1)] = U[IU(ir, j, k)] - s->Umean_resc_T[j - 1];
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= ny; j++) vf[(j - 1) + (size_t)ny * (k - 1)] = V[IV(ir, j, k)] - s->Vmean_resc_T[j - 1];
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= nyg; j++) wf[(j - 1) + (size_t)nyg * (k - 1)] = W[IW(ir, j, k)];
    double delta_resc = height99(nyg, s->Umean_resc_T, s->yg, 0.99 * Uinf);
    double theta_resc = theta_of(nyg, s->Umean_resc_T, s->yg, Uinf);
    double theta_inlet = theta_of(nyg, s->Umean_inlet_T, s->yg, Uinf);
    double uvw = 0.0;
    for (int k = 2; k <= nzg - 1; k++) uvw += 0.5 * (U[IU(ir, 1, k)] + U[IU(ir, 2, k)]) * V[IV(ir, 1, k)];
    uvw = uvw / (double)(nzg - 2);
    double dU = (s->Umean_resc_T[1] - s->Umean_resc_T[0]) / (YG(2) - YG(1));
    double utau_resc = pow(fabs(nu * dU - uvw), 0.5);
    double utau_inlet = utau_resc * pow(theta_resc / theta_inlet, 1.0 / 8.0);
    double gamma = utau_inlet / utau_resc;
    double *ygpi = dalloc(nyg), *ygpr = dalloc(nyg), *etgi = dalloc(nyg), *etgr = dalloc(nyg), *weig = dalloc(nyg);
    double *ypi = dalloc(ny), *ypr = dalloc(ny), *eti = dalloc(ny), *etr = dalloc(ny), *wei = dalloc(ny);
    double alpha = 4.0, b = 0.2;
    for (int j = 0; j < nyg; j++) {
        ygpi[j] = utau_inlet * s->yg[j] / nu;
        ygpr[j] = utau_resc * s->yg[j] / nu;
        etgi[j] = s->yg[j] / s->delta_inlet;
        etgr[j] = s->yg[j] / delta_resc;
        weig[j] = 0.5 * (1.0 + tanh(alpha * (etgi[j] - b) / ((1.0 - 2.0 * b) * etgi[j] + b)) / tanh(alpha));
    }
    for (int j = 0; j < ny; j++) {
        ypi[j] = utau_inlet * s->y[j] / nu;
        ypr[j] = utau_resc * s->y[j] / nu;
        eti[j] = s->y[j] / s->delta_inlet;
        etr[j] = s->y[j] / delta_resc;
        wei[j] = 0.5 * (1.0 + tanh(alpha * (eti[j] - b) / ((1.0 - 2.0 * b) * eti[j] + b)) / tanh(alpha));
    }
    double *umi_in = dalloc(nyg), *umi_out = dalloc(nyg), *vmi_in = dalloc(ny), *vmi_out = dalloc(ny);
    interp_vel(nyg, 1, s->Umean_resc_T, ygpr, ygpi, umi_in);
    interp_vel(nyg, 1, s->Umean_resc_T, etgr, etgi, umi_out);
    for (int j = 0; j < nyg; j++) {
        umi_in[j] = gamma * umi_in[j];
        umi_out[j] = gamma * umi_out[j] + (1.0 - gamma) * Uinf;
    }
    interp_vel(ny, 1, s->Vmean_resc_T, ypr, ypi, vmi_in);
    interp_vel(ny, 1, s->Vmean_resc_T, etr, eti, vmi_out);
    double *ufi = dalloc((size_t)nyg * nzg), *ufo = dalloc((size_t)nyg * nzg);
    double *vfi = dalloc((size_t)ny * nzg), *vfo = dalloc((size_t)ny * nzg);
    double *wfi = dalloc((size_t)nyg * nz), *wfo = dalloc((size_t)nyg * nz);
    interp_vel(nyg, nzg, uf, ygpr, ygpi, ufi);
    interp_vel(nyg, nzg, uf, etgr, etgi, ufo);
    interp_vel(ny, nzg, vf, ypr, ypi, vfi);
    interp_vel(ny, nzg, vf, etr, eti, vfo);
    interp_vel(nyg, nz, wf, ygpr, ygpi, wfi);
    interp_vel(nyg, nz, wf, etgr, etgi, wfo);
    if (iflag_fluc == 1) {
        for (int j = 0; j < nyg; j++) umi_in[j] = umi_out[j] = 0.0;
        for (int j = 0; j < ny; j++) vmi_in[j] = vmi_out[j] = 0.0;
    }
    for (int k = 0; k < nzg; k++)
        for (int j = 0; j < nyg; j++) {
            size_t q = j + (size_t)nyg * k;
            s->Ut_in[q] = (umi_in[j] + gamma * ufi[q]) * (1.0 - weig[j]) + (umi_out[j] + gamma * ufo[q]) * weig[j];
        }
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < nyg; j++) {
            size_t q = j + (size_t)nyg * k;
            s->Wt_in[q] = (gamma * wfi[q]) * (1.0 - weig[j]) + (gamma * wfo[q]) * weig[j];
        }
    for (int k = 0; k < nzg; k++)
        for (int j = 0; j < ny; j++) {
            size_t q = j + (size_t)ny * k;
            s->Vt_in[q] = (vmi_in[j] + gamma * vfi[q]) * (1.0 - wei[j]) + (vmi_out[j] + gamma * vfo[q]) * wei[j];
        }
    /* row 1 (the wall ghost, unset by interp_vel and overwritten by the wall
     * condition) is 0 */
    for (int k = 0; k < nzg; k++) s->Ut_in[(size_t)nyg * k] = s->Vt_in[(size_t)ny * k] = 0.0;
    for (int k = 0; k < nz; k++) s->Wt_in[(size_t)nyg * k] = 0.0;
    double *fr[] = {umr, vmr, umi, vmi, uf, vf, wf, ygpi, ygpr, etgi, etgr, weig, ypi, ypr, eti, etr, wei,
                    umi_in, umi_out, vmi_in, vmi_out, ufi, ufo, vfi, vfo, wfi, wfo};
    for (size_t q = 0; q < sizeof(fr) / sizeof(fr[0]); q++) free(fr[q]);
}

/* the random_number of apply_inflow_bc_x_Blasius_random: a uniform [0, 1)
 * value of (seed, step, component, j, k) from the splitmix64 finaliser */
static unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double inflow_rand(const orc_state *s, int comp, int j0, int k0) {
    unsigned long long key = (((unsigned long long)s->bc_step * 4ULL + (unsigned long long)comp) << 40) ^
                             ((unsigned long long)k0 << 20) ^ (unsigned long long)j0;
    return (double)(mix64(s->seed ^ mix64(key)) >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------------ */
/* apply_boundary_conditions_gpu (boundary_conditions.f90:127-262), and */
/* the general dispatcher's inflow and top flags (:22-121)              */
/* ------------------------------------------------------------------ */
void orc_apply_bc(orc_state *s, int rk_step, double dt) {
    int nx = s->nx, ny = s->ny, nz = s->nz, nxg = s->nxg, nyg = s->nyg, nzg = s->nzg;
    double *U = s->U, *V = s->V, *W = s->W;
    const double *Uo = s->Uo, *Vo = s->Vo, *Wo = s->Wo;
    const double t = s->t;
    /* temporal inflow modes, :137-180 */
    if (s->inflow_flag != 4 && s->n_modes > 0) {
        int nm = s->n_modes, mm = s->m_modes;
        for (int j = 1; j <= nyg; j++)
            for (int k = 1; k <= nzg; k++) {
                double acc = 0.0;
                for (int n = 1; n <= nm; n++)
                    for (int m = 1; m <= mm; m++) {
                        size_t q = (size_t)(j - 1) + (size_t)nyg * ((size_t)(n - 1) + (size_t)nm * (size_t)(m - 1));
                        double ph = s->zmode[n - 1] * ZG(k) - s->tmode[m - 1] * t;
                        acc = acc + (s->qu_re[q] * cos(ph) - s->qu_im[q] * sin(ph));
                    }
                s->Ut[(size_t)(j - 1) + (size_t)nyg * (k - 1)] = acc;
            }
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nzg; k++) {
                double acc = 0.0;
                for (int n = 1; n <= nm; n++)
                    for (int m = 1; m <= mm; m++) {
                        size_t q = (size_t)(j - 1) + (size_t)ny * ((size_t)(n - 1) + (size_t)nm * (size_t)(m - 1));
                        double ph = s->zmode[n - 1] * ZG(k) - s->tmode[m - 1] * t;
                        acc = acc + (s->qv_re[q] * cos(ph) - s->qv_im[q] * sin(ph));
                    }
                s->Vt[(size_t)(j - 1) + (size_t)ny * (k - 1)] = acc;
            }
        for (int j = 1; j <= nyg; j++)
            for (int k = 1; k <= nz; k++) {
                double acc = 0.0;
                for (int n = 1; n <= nm; n++)
                    for (int m = 1; m <= mm; m++) {
                        size_t q = (size_t)(j - 1) + (size_t)nyg * ((size_t)(n - 1) + (size_t)nm * (size_t)(m - 1));
                        double ph = s->zmode[n - 1] * Z(k) - s->tmode[m - 1] * t;
                        acc = acc + (s->qw_re[q] * cos(ph) - s->qw_im[q] * sin(ph));
                    }
                s->Wt[(size_t)(j - 1) + (size_t)nyg * (k - 1)] = acc;
            }
    }
    /* inflow, :183-190 */
    if (s->inflow_flag == 3 || s->inflow_flag == 5) {
        /* Lund rescaling once per step from the step-start field, :569-597
         * (flag 3) and :512-559 (flag 5, on the inlet means, W shifted by 16
         * planes in z) */
        if (s->step_beginning == 1) {
            lund_inflow(s, s->inflow_flag == 5 ? 1 : 0);
            int fl = s->inflow_flag == 5;
            for (int k = 1; k <= nzg; k++)
                for (int j = 1; j <= nyg; j++)
                    U[IU(1, j, k)] = (fl ? s->U_inlet[j - 1] : 0.0) + s->Ut_in[(j - 1) + (size_t)nyg * (k - 1)];
            for (int k = 1; k <= nz; k++) {
                int ks = fl ? (k + 16) % nz + 1 : k;
                for (int j = 1; j <= nyg; j++)
                    W[IW(1, j, k)] = (fl ? s->W_inlet[j - 1] : 0.0) + s->Wt_in[(j - 1) + (size_t)nyg * (ks - 1)];
            }
            for (int k = 1; k <= nzg; k++)
                for (int j = 1; j <= ny; j++)
                    V[IV(1, j, k)] = (fl ? s->V_inlet[j - 1] : 0.0) + s->Vt_in[(j - 1) + (size_t)ny * (k - 1)];
            s->step_beginning = 0;
        }
    } else if (s->inflow_flag == 4) {
        /* Blasius + random perturbations once per step, :423-447 */
        if (s->step_beginning == 1) {
            for (int j = 1; j <= nyg; j++) {
                double amp_loc = s->amp * (1.0 - s->U_inlet[j - 1]);
                for (int k = 1; k <= nzg; k++)
                    U[IU(1, j, k)] = s->U_inlet[j - 1] + amp_loc * (inflow_rand(s, 0, j - 1, k - 1) - 0.5);
            }
            for (int j = 1; j <= nyg; j++) {
                double amp_loc = s->amp * (1.0 - s->U_inlet[j - 1]);
                for (int k = 1; k <= nz; k++)
                    W[IW(1, j, k)] = s->W_inlet[j - 1] + amp_loc * (inflow_rand(s, 2, j - 1, k - 1) - 0.5);
            }
            for (int j = 1; j <= ny; j++) {
                double amp_loc = s->amp * (1.0 - s->U_inlet[(j < nyg ? j : nyg) - 1]);
                for (int k = 1; k <= nzg; k++)
                    V[IV(1, j, k)] = s->V_inlet[j - 1] + amp_loc * (inflow_rand(s, 1, j - 1, k - 1) - 0.5);
            }
            s->step_beginning = 0;
        }
    } else {
        for (int k = 1; k <= nzg; k++)
            for (int j = 1; j <= nyg; j++) U[IU(1, j, k)] = s->U_inlet[j - 1] + s->Ut[(j - 1) + (size_t)nyg * (k - 1)];
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= nyg; j++) W[IW(1, j, k)] = s->W_inlet[j - 1] + s->Wt[(j - 1) + (size_t)nyg * (k - 1)];
        for (int k = 1; k <= nzg; k++)
            for (int j = 1; j <= ny; j++) V[IV(1, j, k)] = s->V_inlet[j - 1] + s->Vt[(j - 1) + (size_t)ny * (k - 1)];
    }
    /* outflow, :193-202 (rk_t of RK3 for every scheme) */
    double Uc = s->U_top[nx - 1];
    double c = s->rk_t[rk_step];
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++)
            U[IU(nx - 1, j, k)] = Uo[IU(nx - 1, j, k)] - c * dt * Uc * (Uo[IU(nx - 1, j, k)] - Uo[IU(nx - 2, j, k)]) / (X(nx - 1) - X(nx - 2));
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= ny; j++)
            V[IV(nxg - 1, j, k)] = Vo[IV(nxg - 1, j, k)] - c * dt * Uc * (Vo[IV(nxg - 1, j, k)] - Vo[IV(nxg - 2, j, k)]) / (XG(nxg - 1) - XG(nxg - 2));
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= nyg; j++)
            W[IW(nxg - 1, j, k)] = Wo[IW(nxg - 1, j, k)] - c * dt * Uc * (Wo[IW(nxg - 1, j, k)] - Wo[IW(nxg - 2, j, k)]) / (XG(nxg - 1) - XG(nxg - 2));
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++) U[IU(nx, j, k)] = U[IU(nx - 1, j, k)];
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= ny; j++) V[IV(nxg, j, k)] = V[IV(nxg - 1, j, k)];
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= nyg; j++) W[IW(nxg, j, k)] = W[IW(nxg - 1, j, k)];
    /* periodic z, :205-217 */
    periodic_z_cc(s, U, nx, nyg);
    periodic_z_cc(s, V, nxg, ny);
    periodic_z_w(s, W);
    /* top, :220-228 (top_flag 0, :947-964) */
    if (s->top_flag == 1 || s->top_flag == 2) {
        /* apply_Dirichlet_bc_y_top_BlowingSuction(V, top_flag), :648-673 */
        for (int i = 1; i <= nxg; i++) {
            double r = (s->x_bs - XG(i)) / s->sigma_bs;
            double f = s->top_flag == 1
                           ? sqrt(2.0) * s->vbs_max * (s->x_bs - XG(i)) / s->sigma_bs * exp(0.5 - r * r) + s->phi_bs
                           : sqrt(2.0) * s->vbs_max * (s->x_bs - XG(i)) / s->sigma_bs * exp(s->phi_bs - r * r);
            for (int k = 1; k <= nzg; k++) V[IV(i, ny, k)] = f;
        }
        /* zero_wz_top, :264-283 (dx = dxmin, initialization.f90:254) */
        double dy = YG(nyg) - YG(nyg - 1);
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nxg - 1; i++)
                U[IU(i, nyg, k)] = (dy / s->dx) * (V[IV(i + 1, ny, k)] - V[IV(i, ny, k)]) + U[IU(i, nyg - 1, k)];
        for (int k = 1; k <= nz; k++)
            for (int i = 1; i <= nxg; i++) W[IW(i, nyg, k)] = W[IW(i, nyg - 1, k)];
        /* columns 1..i_rescale: U and W back to U_top, W_top.  The V_top copy
         * of :279 stores V(i, nyg, :), one row past V's top face; with V's
         * column-major shape that is V(i, 1, k+1), which the wall condition
         * below sets to 0 again, so it leaves no trace and is not restated */
        for (int i = 1; i <= s->i_rescale; i++) {
            if (i <= nx)
                for (int k = 1; k <= nzg; k++) U[IU(i, nyg, k)] = s->U_top[i - 1];
            if (i <= nxg)
                for (int k = 1; k <= nz; k++) W[IW(i, nyg, k)] = s->W_top[i - 1];
        }
    } else if (s->top_flag == 3) {
        /* apply_top_bc_y_Falkner_Skan, :979-1007 */
        double x_falkner = -1.0, Uinf = 1.0, m = s->beta_hartree / (2.0 - s->beta_hartree);
        double C_falkner = Uinf / pow(X(1) - x_falkner, m);
        for (int i = 1; i <= nx; i++) {
            double u = C_falkner * pow(XG(i) - x_falkner, m);
            for (int k = 1; k <= nzg; k++) U[IU(i, nyg, k)] = u;
            s->U_top[i - 1] = u;
        }
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nxg; i++) V[IV(i, ny, k)] = 0.0;
        for (int k = 1; k <= nz; k++)
            for (int i = 1; i <= nxg; i++) W[IW(i, nyg, k)] = W[IW(i, nyg - 1, k)];
    } else {
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nx; i++) U[IU(i, nyg, k)] = s->U_top[i - 1];
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nxg; i++) V[IV(i, ny, k)] = s->V_top[i - 1];
        for (int k = 1; k <= nz; k++)
            for (int i = 1; i <= nxg; i++) W[IW(i, nyg, k)] = s->W_top[i - 1];
    }
    /* wall, :231-238 (no-slip), or the wall model's conditions, :73-108 */
    orc__wall_bc(s);
    /* global mass conservation, :241-260 (single rank: kk = 2) */
    double Qx = 0.0, length_y = 0.0;
    int kk = 2;
    for (int j = 2; j <= nyg - 1; j++) {
        double sum = 0.0;
        for (int k = 2; k <= nzg - kk; k++) sum += (U[IU(1, j, k)] - U[IU(nx - 1, j, k)]) * (Y(j) - Y(j - 1));
        Qx = Qx + sum;
        length_y = length_y + Y(j) - Y(j - 1);
    }
    double Qy = 0.0;
    for (int i = 2; i <= nxg - 2; i++) {
        double sum = 0.0;
        for (int k = 2; k <= nzg - kk; k++) sum += (V[IV(i, 1, k)] - V[IV(i, ny, k)]) * (X(i) - X(i - 1));
        Qy = Qy + sum;
    }
    double Q_total = Qx + Qy;
    double Delta_U = Q_total / length_y / (double)(nzg - 3);
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++) {
            U[IU(nx - 1, j, k)] = U[IU(nx - 1, j, k)] + Delta_U;
            U[IU(nx, j, k)] = U[IU(nx - 1, j, k)];
        }
}

/* ------------------------------------------------------------------ */
/* solve_poisson_cpu (projection.f90:210-288)                          */
/* ------------------------------------------------------------------ */

/* LAPACK zgtsv restated for a real tridiagonal matrix and one complex RHS:
 * Gaussian elimination with partial pivoting (row interchange when the
 * sub-diagonal entry exceeds the pivot in magnitude), then back substitution
 * with the second superdiagonal fill-in.  dl, d, du are overwritten. */
static int gtsv(int n, double *dl, double *d, double *du, orc_cplx *b) {
    for (int k = 0; k < n - 1; k++) {
        if (dl[k] == 0.0) {
            if (d[k] == 0.0) return k + 1;
        } else if (fabs(d[k]) >= fabs(dl[k])) {
            double mult = dl[k] / d[k];
            d[k + 1] = d[k + 1] - mult * du[k];
            b[k + 1].re = b[k + 1].re - mult * b[k].re;
            b[k + 1].im = b[k + 1].im - mult * b[k].im;
            if (k < n - 2) dl[k] = 0.0;
        } else {
            double mult = d[k] / dl[k];
            d[k] = dl[k];
            double temp = d[k + 1];
            d[k + 1] = du[k] - mult * temp;
            if (k < n - 2) {
                dl[k] = du[k + 1];
                du[k + 1] = -mult * dl[k];
            }
            du[k] = temp;
            orc_cplx tb = b[k];
            b[k] = b[k + 1];
            b[k + 1].re = tb.re - mult * b[k + 1].re;
            b[k + 1].im = tb.im - mult * b[k + 1].im;
        }
    }
    if (d[n - 1] == 0.0) return n;
    b[n - 1].re = b[n - 1].re / d[n - 1];
    b[n - 1].im = b[n - 1].im / d[n - 1];
    if (n > 1) {
        b[n - 2].re = (b[n - 2].re - du[n - 2] * b[n - 1].re) / d[n - 2];
        b[n - 2].im = (b[n - 2].im - du[n - 2] * b[n - 1].im) / d[n - 2];
    }
    for (int k = n - 3; k >= 0; k--) {
        b[k].re = (b[k].re - du[k] * b[k + 1].re - dl[k] * b[k + 2].re) / d[k];
        b[k].im = (b[k].im - du[k] * b[k + 1].im - dl[k] * b[k + 2].im) / d[k];
    }
    return 0;
}

/* 2-D DFT of a (n1 x n2) complex array, n1 fastest: rows then columns */
static void fft2(orc_fft_plan *p1, orc_fft_plan *p2, orc_cplx *a, int n1, int n2) {
    for (int k = 0; k < n2; k++) orc_fft_exec(p1, a + (size_t)k * n1, a + (size_t)k * n1, 1, 1);
    for (int i = 0; i < n1; i++) orc_fft_exec(p2, a + i, a + i, n1, n1);
}

void orc_poisson(orc_state *s, double *rhs_p) {
    int nxp = s->nxp, nzp = s->nzp, mx = s->mx, mz = s->mz, nyg = s->nyg;
    int nxpe = 4 * nxp;
    int nm = s->nm;
    size_t nhat = (size_t)(mx + 1) * nm * (mz + 1);
    orc_cplx *hat = (orc_cplx *)calloc(nhat, sizeof(orc_cplx));
#define HAT(i, j, k) hat[(size_t)(i) + (size_t)(mx + 1) * ((size_t)((j)-2) + (size_t)nm * (size_t)(k))]
#define RP(i, j, k) rhs_p[IP(i, j, k)]
    /* forward cosine+Fourier transform, :219-235 */
#pragma omp parallel
    {
        orc_fft_plan *pf1 = orc_fft_plan_create(nxpe, -1);
        orc_fft_plan *pf2 = orc_fft_plan_create(nzp, -1);
        orc_cplx *plane = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)nxpe * nzp);
#pragma omp for
        for (int j = 2; j <= nyg - 1; j++) {
            memset(plane, 0, sizeof(orc_cplx) * (size_t)nxpe * nzp);
            for (int kk = 0; kk < nzp; kk++) {
                /* plane(i,:) for i = 2,4,...,2*nxp (1-based) */
                int ii = 1;
                for (int i = 2; i <= 2 * nxp; i += 2) {
                    plane[(size_t)(i - 1) + (size_t)nxpe * kk].re = RP(ii + 1, j, kk + 2);
                    ii++;
                }
                /* plane(i,:) for i = 4*nxp, 4*nxp-2, ..., 2*nxp+2 */
                ii = 1;
                for (int i = 4 * nxp; i >= 2 * nxp + 2; i -= 2) {
                    plane[(size_t)(i - 1) + (size_t)nxpe * kk].re = RP(ii + 1, j, kk + 2);
                    ii++;
                }
            }
            fft2(pf1, pf2, plane, nxpe, nzp);
            for (int k = 0; k <= mz; k++)
                for (int i = 0; i <= mx; i++) HAT(i, j, k) = plane[(size_t)i + (size_t)nxpe * k];
        }
        orc_fft_plan_destroy(pf1);
        orc_fft_plan_destroy(pf2);
        free(plane);
    }
    /* tridiagonal solve, :238-256 */
#pragma omp parallel
    {
        double *D = dalloc(nm), *DL = dalloc(nm), *DU = dalloc(nm);
        orc_cplx *aux = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)nm);
#pragma omp for collapse(2)
        for (int k = 0; k <= mz; k++)
            for (int i = 0; i <= mx; i++) {
                for (int j = 2; j <= nyg - 1; j++) D[j - 2] = s->dyy_d[j - 2] + s->kxx[i] + s->kzz[k];
                for (int j = 2; j <= nyg - 2; j++) DL[j - 2] = s->dyy_l[j - 2];
                for (int j = 2; j <= nyg - 2; j++) DU[j - 2] = s->dyy_u[j - 2];
                if (i == 0 && k == 0) D[0] = 3.0 / 2.0 * D[0];
                for (int j = 2; j <= nyg - 1; j++) aux[j - 2] = HAT(i, j, k);
                gtsv(nm, DL, D, DU, aux);
                for (int j = 2; j <= nyg - 1; j++) HAT(i, j, k) = aux[j - 2];
            }
        free(D);
        free(DL);
        free(DU);
        free(aux);
    }
    /* inverse cosine+Fourier transform, :259-283 */
    double norm = (double)((long)nxpe * nzp);
#pragma omp parallel
    {
        orc_fft_plan *pb1 = orc_fft_plan_create(nxpe, +1);
        orc_fft_plan *pb2 = orc_fft_plan_create(nzp, +1);
        orc_cplx *plane = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)nxpe * nzp);
#pragma omp for
        for (int j = 2; j <= nyg - 1; j++) {
            memset(plane, 0, sizeof(orc_cplx) * (size_t)nxpe * nzp);
            for (int kk = 0; kk <= mz; kk++) {
                orc_cplx *pl = plane + (size_t)nxpe * kk;
                for (int i = 0; i <= nxp - 1; i++) pl[i] = HAT(i, j, kk);
                int ii = nxp + 2 - 1;
                for (int i = nxp - 1; i >= 0; i--) {
                    pl[ii].re = -HAT(i, j, kk).re; /* plane(ii+1) 1-based */
                    pl[ii].im = -HAT(i, j, kk).im;
                    ii++;
                }
                ii = 2 - 1;
                for (int i = 2 * nxp + 2 - 1; i <= 3 * nxp - 1; i++) {
                    pl[i].re = -HAT(ii, j, kk).re;
                    pl[i].im = -HAT(ii, j, kk).im;
                    ii++;
                }
                ii = 4 * nxp - nxp + 2 - 1;
                for (int i = nxp - 1; i >= 2 - 1; i--) {
                    pl[ii] = HAT(i, j, kk);
                    ii++;
                }
            }
            fft2(pb1, pb2, plane, nxpe, nzp);
            for (int kk = 0; kk < nzp; kk++) {
                int ii = 2;
                for (int i = 2; i <= 2 * nxp; i += 2) {
                    RP(ii, j, kk + 2) = plane[(size_t)(i - 1) + (size_t)nxpe * kk].re / norm;
                    ii++;
                }
            }
        }
        orc_fft_plan_destroy(pb1);
        orc_fft_plan_destroy(pb2);
        free(plane);
    }
    free(hat);
#undef HAT
}

/* compute_projection_step (projection.f90:21-87, 293-303, 355-386) */
void orc_projection(orc_state *s, int rk_step, double dt) {
    int nx = s->nx, ny = s->ny, nz = s->nz, nxg = s->nxg, nyg = s->nyg, nzg = s->nzg;
    double *U = s->U, *V = s->V, *W = s->W, *rhs_p = s->rhs_p;
    /* compute_pseudo_pressure_rhs, :38-47 */
#pragma omp parallel for
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                RP(i, j, k) = (U[IU(i, j, k)] - U[IU(i - 1, j, k)]) / (X(i) - X(i - 1)) +
                              (V[IV(i, j, k)] - V[IV(i, j - 1, k)]) / (Y(j) - Y(j - 1)) +
                              (W[IW(i, j, k)] - W[IW(i, j, k - 1)]) / (Z(k) - Z(k - 1));
    orc_poisson(s, rhs_p);
    /* apply_periodic_z_pressure, :295-302 (rhs_p spans k = 2..nzg) */
    for (int k = 2; k <= nzg; k++)
        for (int j = 2; j <= nyg - 1; j++) RP(nxg - 1, j, k) = RP(nxg - 2, j, k);
    for (int j = 2; j <= nyg - 1; j++)
        for (int i = 2; i <= nxg - 1; i++) RP(i, j, s->nzp + 1 + 1) = RP(i, j, 2);
    /* P for statistics at the first sub-step, :75-85 */
    if (rk_step == 1) {
        double *P = s->P;
        double den = dt * s->rk2_coef[1][1];
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) P[IP(i, j, k)] = RP(i, j, k) / den;
        for (int k = 1; k <= nzg; k++)
            for (int j = 1; j <= nyg; j++) {
                P[IP(1, j, k)] = P[IP(2, j, k)];
                P[IP(nxg, j, k)] = P[IP(nxg - 1, j, k)];
            }
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nxg; i++) {
                P[IP(i, 1, k)] = P[IP(i, 2, k)];
                P[IP(i, nyg, k)] = P[IP(i, nyg - 1, k)];
            }
        for (int j = 1; j <= nyg; j++)
            for (int i = 1; i <= nxg; i++) {
                P[IP(i, j, 1)] = P[IP(i, j, 2)];
                P[IP(i, j, nzg)] = P[IP(i, j, nzg - 1)];
            }
    }
    /* project_velocity, :359-384 */
#pragma omp parallel for
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nx - 1; i++)
                U[IU(i, j, k)] = U[IU(i, j, k)] - (RP(i + 1, j, k) - RP(i, j, k)) / (XG(i + 1) - XG(i));
#pragma omp parallel for
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                V[IV(i, j, k)] = V[IV(i, j, k)] - (RP(i, j + 1, k) - RP(i, j, k)) / (YG(j + 1) - YG(j));
#pragma omp parallel for
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                W[IW(i, j, k)] = W[IW(i, j, k)] - (RP(i, j, k + 1) - RP(i, j, k)) / (ZG(k + 1) - ZG(k));
}

/* check_divergence (projection.f90:391-413) */
double orc_max_divergence(const orc_state *s) {
    int nxg = s->nxg, nyg = s->nyg, nzg = s->nzg;
    const double *U = s->U, *V = s->V, *W = s->W;
    double div = 0.0;
#pragma omp parallel for reduction(max : div)
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 2; i++) {
                double d = fabs((U[IU(i, j, k)] - U[IU(i - 1, j, k)]) / (X(i) - X(i - 1)) +
                                (V[IV(i, j, k)] - V[IV(i, j - 1, k)]) / (Y(j) - Y(j - 1)) +
                                (W[IW(i, j, k)] - W[IW(i, j, k - 1)]) / (Z(k) - Z(k - 1)));
                div = fmax(div, d);
            }
    return div;
}

/* compute_dt (time_integration.f90:528-591) */
double orc_compute_dt(orc_state *s) {
    double CFLa = fabs(s->cfl);
    if (s->cfl < 0) {
        s->dt = -s->cfl;
        return s->dt;
    }
    int nxg = s->nxg, nyg = s->nyg, nzg = s->nzg;
    const double *U = s->U, *V = s->V, *W = s->W;
    double lU = 0.0, lV = 0.0, lW = 0.0;
    for (int i = 2; i <= nxg - 1; i++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int k = 2; k <= nzg - 1; k++) {
                lU = fmax(lU, (XG(i + 1) - XG(i)) / fabs(U[IU(i, j, k)]));
                lV = fmax(lV, (YG(j + 1) - YG(j)) / fabs(V[IV(i, j, k)]));
                lW = fmax(lW, (ZG(k + 1) - ZG(k)) / fabs(W[IW(i, j, k)]));
            }
    double dcu = CFLa * lU, dcv = CFLa * lV, dcw = CFLa * lW;
    double dt_conv = fmin(dcu, fmin(dcv, dcw));
    double dvu = CFLa * pow(s->dxmin, 2.0) / s->nu;
    double dvv = CFLa * pow(s->dymin, 2.0) / s->nu;
    double dvw = CFLa * pow(s->dzmin, 2.0) / s->nu;
    double dt_vis = fmin(dvu, fmin(dvv, dvw));
    s->dt = fmin(dt_conv, dt_vis);
    return s->dt;
}

/* compute_time_step_RK3 / RK2 (time_integration.f90:305-465, 145-300) */
void orc_step(orc_state *s) {
    int nx = s->nx, ny = s->ny, nz = s->nz, nxg = s->nxg, nyg = s->nyg, nzg = s->nzg;
    double dt = s->dt;
    double to = s->t;
    s->step_beginning = 1; /* time_integration.f90:29 */
    memcpy(s->Uo, s->U, sizeof(double) * nU(s));
    memcpy(s->Vo, s->V, sizeof(double) * nV(s));
    memcpy(s->Wo, s->W, sizeof(double) * nW(s));
    int nst = s->scheme == 2 ? 2 : 3;
    for (int st = 1; st <= nst; st++) {
        int rk_step = st;
        /* time_integration.f90:341-346: eddy viscosity, then the wall model */
        if (s->les > 0 || s->iwall > 0) {
            orc_eddy_viscosity(s);
            orc_wall_model(s);
        }
        rhs_u(s, s->U, s->V, s->W, s->Fu[st]);
        rhs_v(s, s->U, s->V, s->W, s->Fv[st]);
        rhs_w(s, s->U, s->V, s->W, s->Fw[st]);
        double a[4] = {0, 0, 0, 0};
        for (int q = 1; q <= st; q++) a[q] = (nst == 2) ? s->rk2_coef[st][q] : s->rk_coef[st][q];
        double *Fu1 = s->Fu[1], *Fu2 = s->Fu[2], *Fu3 = s->Fu[3];
        double *Fv1 = s->Fv[1], *Fv2 = s->Fv[2], *Fv3 = s->Fv[3];
        double *Fw1 = s->Fw[1], *Fw2 = s->Fw[2], *Fw3 = s->Fw[3];
#define ADV(F1, F2, F3, idx)                                                                       \
    (st == 1 ? dt * a[1] * F1[idx]                                                                 \
             : (st == 2 ? dt * (a[1] * F1[idx] + a[2] * F2[idx])                                   \
                        : dt * (a[1] * F1[idx] + a[2] * F2[idx] + a[3] * F3[idx])))
#pragma omp parallel for
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nx - 1; i++) {
                    size_t id = IU(i, j, k);
                    s->U[id] = s->Uo[id] + ADV(Fu1, Fu2, Fu3, id);
                }
#pragma omp parallel for
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    size_t id = IV(i, j, k);
                    s->V[id] = s->Vo[id] + ADV(Fv1, Fv2, Fv3, id);
                }
#pragma omp parallel for
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    size_t id = IW(i, j, k);
                    s->W[id] = s->Wo[id] + ADV(Fw1, Fw2, Fw3, id);
                }
#undef ADV
        s->t = to + ((nst == 2) ? s->rk2_t[rk_step] : s->rk_t[rk_step]) * dt;
        orc_apply_bc(s, rk_step, dt);
        orc_projection(s, rk_step, dt);
        orc_apply_bc(s, rk_step, dt);
    }
    s->bc_step++;
}

/* Cf part of compute_statistics (statistics.f90:44-52, 61-80, 137-160) */
void orc_wall_stats(const orc_state *s, double *cf) {
    int nx = s->nx, nzg = s->nzg;
    const double *U = s->U, *V = s->V, *nu_t = s->nu_t;
    double nz2 = (double)(nzg - 2);
    double *dUdy = dalloc(nx), *ntm1 = dalloc(s->nxg), *ntm2 = dalloc(s->nxg);
    /* nu_t_mean(ii, 1:2) for ii = 1..nx (statistics.f90:77, 135); column nxg
     * is never filled and keeps its initial 0 (initialization.f90:501), so the
     * SGS term of the last station takes 0 for its ii+1 neighbours */
    for (int ii = 1; ii <= nx; ii++) {
        double a1 = 0.0, a2 = 0.0;
        for (int k = 2; k <= nzg - 1; k++) {
            a1 += nu_t[IP(ii, 1, k)];
            a2 += nu_t[IP(ii, 2, k)];
        }
        ntm1[ii - 1] = a1 / nz2;
        ntm2[ii - 1] = a2 / nz2;
    }
    for (int ii = 1; ii <= nx; ii++) {
        double u1 = 0.0, u2 = 0.0, uv = 0.0;
        for (int k = 2; k <= nzg - 1; k++) {
            u1 += U[IU(ii, 1, k)];
            u2 += U[IU(ii, 2, k)];
            /* term_1 = interpolate_y(U, di=2) at j=1, term_2 = interpolate_x(V, di=1) at j=1 */
            double t1 = WY0(1) * U[IU(ii, 1, k)] + WY1(1) * U[IU(ii, 2, k)];
            double t2 = 0.5 * (V[IV(ii, 1, k)] + V[IV(ii + 1, 1, k)]);
            uv += t1 * t2;
        }
        u1 = u1 / nz2;
        u2 = u2 / nz2;
        double uvw = uv / nz2;
        double dUdy_wall = (u2 - u1) / (YG(2) - YG(1));
        double tau = 0.25 * (ntm1[ii - 1] + ntm2[ii - 1] + ntm1[ii] + ntm2[ii]) * dUdy_wall;
        dUdy[ii - 1] = dUdy_wall;
        cf[ii - 1] = 2.0 * (-uvw + s->nu * dUdy_wall + tau) / 1.0;
    }
    free(dUdy);
    free(ntm1);
    free(ntm2);
}

/* compute_statistics (statistics.f90:44-80, 125-135), single rank:
 * term_1 = interpolate_y(U, 2), term_2 = interpolate_x(V, 1) and
 * term = interpolate_y(interpolate_x(W, 1), 2) (interpolation.f90:24-85),
 * summed over z planes 2..nzg-1 (W: 1..nz-2) and divided by nzg-2 (nz-2) */
void orc_statistics(const orc_state *s, double *means, double *nu_t_mean) {
    int nx = s->nx, ny = s->ny, nz = s->nz, nxg = s->nxg, nyg = s->nyg, nzg = s->nzg;
    const double *U = s->U, *V = s->V, *W = s->W, *P = s->P, *nu_t = s->nu_t;
    size_t n2 = (size_t)nx * ny;
    memset(nu_t_mean, 0, sizeof(double) * (size_t)nxg * nyg);
    for (int jj = 1; jj <= ny; jj++)
        for (int ii = 1; ii <= nx; ii++) {
            double su = 0, sv = 0, sw = 0, su2 = 0, sv2 = 0, sw2 = 0, suv = 0, sp = 0, sp2 = 0, sn = 0;
            for (int k = 2; k <= nzg - 1; k++) {
                double t1 = WY0(jj) * U[IU(ii, jj, k)] + WY1(jj) * U[IU(ii, jj + 1, k)];
                double t2 = 0.5 * (V[IV(ii, jj, k)] + V[IV(ii + 1, jj, k)]);
                double p = P[IP(ii, jj, k)];
                su += t1;
                sv += t2;
                su2 += pow(t1, 2.0);
                sv2 += pow(t2, 2.0);
                suv += t1 * t2;
                sp += p;
                sp2 += pow(p, 2.0);
                sn += nu_t[IP(ii, jj, k)];
            }
            for (int k = 1; k <= nz - 2; k++) {
                double a = 0.5 * (W[IW(ii, jj, k)] + W[IW(ii + 1, jj, k)]);
                double b = 0.5 * (W[IW(ii, jj + 1, k)] + W[IW(ii + 1, jj + 1, k)]);
                double t = WY0(jj) * a + WY1(jj) * b;
                sw += t;
                sw2 += pow(t, 2.0);
            }
            size_t q = (size_t)(ii - 1) + (size_t)nx * (jj - 1);
            double dz = (double)(nzg - 2), dw = (double)(nz - 2);
            means[0 * n2 + q] = su / dz;
            means[1 * n2 + q] = sv / dz;
            means[2 * n2 + q] = sw / dw;
            means[3 * n2 + q] = su2 / dz;
            means[4 * n2 + q] = sv2 / dz;
            means[5 * n2 + q] = sw2 / dw;
            means[6 * n2 + q] = suv / dz;
            means[7 * n2 + q] = sp / dz;
            means[8 * n2 + q] = sp2 / dz;
            nu_t_mean[(size_t)(ii - 1) + (size_t)nxg * (jj - 1)] = sn / dz;
        }
}

/* output_monitor reductions (monitor.f90:29-52) */
void orc_monitor(const orc_state *s, double *out) {
    int nx = s->nx, ny = s->ny, nz = s->nz, nxg = s->nxg, nyg = s->nyg, nzg = s->nzg;
    double su = 0, sv = 0, sw = 0, sn = 0, mu = -INFINITY, mv = -INFINITY, mw = -INFINITY, mn = -INFINITY;
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
                su += s->U[IU(i, j, k)];
                mu = fmax(mu, s->U[IU(i, j, k)]);
            }
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                sv += s->V[IV(i, j, k)];
                mv = fmax(mv, s->V[IV(i, j, k)]);
            }
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                sw += s->W[IW(i, j, k)];
                mw = fmax(mw, s->W[IW(i, j, k)]);
            }
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                sn += s->nu_t[IP(i, j, k)];
                mn = fmax(mn, s->nu_t[IP(i, j, k)]);
            }
    out[0] = su;
    out[1] = sv;
    out[2] = sw;
    out[3] = sn;
    out[4] = mu;
    out[5] = mv;
    out[6] = mw;
    out[7] = mn;
}

void orc_eddy_viscosity(orc_state *s) {
    /* subgrid.f90:32-45 */
    if (s->les == 1) {
        orc_smagorinsky(s);
    } else if (s->les == 2) {
        orc_dsm(s);
    } else if (s->les == 3) {
        orc_dsm_noavg(s);
    } else {
        memset(s->nu_t, 0, sizeof(double) * nP(s));
        memset(s->avg_nu_t, 0, sizeof(double) * (size_t)s->nxg * s->nyg);
    }
}

void orc_set_nut_bc(orc_state *s, int dirichlet) { s->nut_dirichlet = dirichlet; }

/* accessors for orc_dsm.c */
double *orc__nu_t(orc_state *s) { return s->nu_t; }

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }
