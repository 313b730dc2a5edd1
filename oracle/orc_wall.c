/*
 * CPU-ORACLE TEST INFRASTRUCTURE — not part of the product.
 *
 * Wall models (wallmodel.f90) and the bottom-wall conditions that use their
 * slip lengths (boundary_conditions.f90:73-108), restated in plain C.
 *
 * compute_wall_model (wallmodel.f90:25-144), per RK stage after the eddy
 * viscosity:
 *   - istress_model 1: utau_model from the law of the wall at yg(3) by
 *     Newton's method (compute_law_of_the_wall_model :1951-1985,
 *     Newton_solver.f90:13-85, 98-118), otherwise utau_model = utau_ref
 *     (initialization.f90:745);
 *   - iwall_model 4: momentum balance at both walls
 *     (compute_alpha_momentum_wall_model :674-764);
 *     5: alpha_mean_[xyz]*(1 + alpha_std*sin(omega t))
 *     (compute_constant_alpha :952-961);
 *     6, 8: log-layer slip length with the wall value of avg_nu_t or with
 *     nu_t = kappa utau y (compute_log_layer_wall_model :987-1026);
 *     9, 12: slip lengths from the Cf correlation and the wall stress
 *     balance (compute_alpha_v_from_Cf :1532-1697, compute_alpha_u_from_Cf
 *     :1263-1424);
 *     11: dU/dy at the wall from utau_model and the resolved tau_uv
 *     (compute_alpha_u_momentum_wall_model :1444-1500);
 *     15: dU/dy at the wall from the turbulent Cf correlation
 *     (compute_exact_Neumann_from_Cf :2003-2066).
 * The alphas are functions of x only (every model fills alpha(i, :, :) with
 * one value), so they are kept as alpha_x(nx), alpha_y(nxg), alpha_z(nxg).
 *
 * Bottom-wall conditions (apply_boundary_conditions :73-108): iwall_model 0
 * no-slip; 11 and 15 Neumann for U (apply_Nozero_Neumann_bc_y :359-372) with
 * V = W = 0; every other model Robin for U and W (apply_Robin_bc_y_bottom
 * :1107-1132, id 2) and the pseudo-Robin V from the step-start field with its
 * mean over the wall removed (apply_pseudo_Robin_bc_y_bottom :1152-1180,
 * id 2).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

#define KAPPA_LAW 0.38 /* kappa_model, wallmodel.f90:1965 */
#define B_LAW 5.20     /* B_model, :1966 */

static int model_ok(int iwall) {
    return iwall == 0 || iwall == 4 || iwall == 5 || iwall == 6 || iwall == 8 || iwall == 9 || iwall == 11 ||
           iwall == 12 || iwall == 15;
}

int orc_set_wall_model(orc_state *s, int iwall, int istress, double amx, double amy, double amz, double astd,
                       double freq_mult, int nut_flag, double frac_vis) {
    if (!model_ok(iwall)) return -1;
    if ((iwall == 9 || iwall == 12) && nut_flag != 0 && nut_flag != 2) return -1;
    s->nut_flag = nut_flag;
    s->frac_vis = frac_vis;
    s->beta_y = 0.0;
    s->iwall = iwall;
    s->istress = istress;
    s->alpha_mean_x = amx;
    s->alpha_mean_y = amy;
    s->alpha_mean_z = amz;
    s->alpha_std = astd;
    s->freq_mult = freq_mult;
    if (!s->alpha_x) {
        s->alpha_x = orc__dalloc(s->nx);
        s->alpha_y = orc__dalloc(s->nxg);
        s->alpha_z = orc__dalloc(s->nxg);
        s->utau_model = orc__dalloc(s->nx);
        s->utau_ref = orc__dalloc(s->nx);
    }
    memset(s->alpha_x, 0, sizeof(double) * s->nx);
    memset(s->alpha_y, 0, sizeof(double) * s->nxg);
    memset(s->alpha_z, 0, sizeof(double) * s->nxg);
    /* initialization.f90:745-746 */
    for (int i = 1; i <= s->nx; i++) {
        s->utau_ref[i - 1] = pow(0.5 * 0.027 * pow(X(i) / s->nu, -1.0 / 7.0), 0.5);
        s->utau_model[i - 1] = s->utau_ref[i - 1];
    }
    return 0;
}

void orc_get_wall_model(const orc_state *s, double *ax, double *ay, double *az, double *utau) {
    if (!s->alpha_x) {
        memset(ax, 0, sizeof(double) * s->nx);
        memset(ay, 0, sizeof(double) * s->nxg);
        memset(az, 0, sizeof(double) * s->nxg);
        memset(utau, 0, sizeof(double) * s->nx);
        return;
    }
    memcpy(ax, s->alpha_x, sizeof(double) * s->nx);
    memcpy(ay, s->alpha_y, sizeof(double) * s->nxg);
    memcpy(az, s->alpha_z, sizeof(double) * s->nxg);
    memcpy(utau, s->utau_model, sizeof(double) * s->nx);
}

double orc_wall_beta(const orc_state *s) { return s->beta_y; }

/* f_law_of_wall, df_law_of_wall (Newton_solver.f90:98-118) */
static double law_f(double u, double ygm, double nu, double Um) {
    return u / KAPPA_LAW * log(ygm * u / nu) + B_LAW * u - Um;
}
static double law_df(double u, double ygm, double nu) { return 1.0 / KAPPA_LAW * log(ygm * u / nu) + 1.0 / KAPPA_LAW + B_LAW; }

/* Newton_iter_solver (Newton_solver.f90:13-85): at most 30 steps, stop when
 * |f| < 1e-10 */
static double newton_law(double x0, double ygm, double nu, double Um) {
    double x = x0;
    for (int k = 1; k <= 30; k++) {
        double fx = law_f(x, ygm, nu, Um);
        double fp = law_df(x, ygm, nu);
        if (fabs(fx) < 1e-10) break;
        x = x - fx / fp;
    }
    return x;
}

/* compute_law_of_the_wall_model (wallmodel.f90:1951-1985), jref = 3 */
static void law_of_the_wall(orc_state *s) {
    const int nx = s->nx, nzg = s->nzg;
    const double ygm = YG(3);
    for (int i = 2; i <= nx - 1; i++) {
        double u0 = s->utau_model[i - 1];
        double Um = 0.0;
        for (int k = 2; k <= nzg - 1; k++) Um += s->U[IU(i, 3, k)];
        Um = Um / (double)(nzg - 2);
        double u1 = newton_law(u0, ygm, s->nu, Um);
        s->utau_model[i - 1] = isnan(u1) ? 0.0 : u1;
    }
    s->utau_model[0] = s->utau_model[1];
    s->utau_model[nx - 1] = s->utau_model[nx - 2];
}

/* nu_t at the wall edge (i, y(1), k): interpolate_y (weighted, centres to
 * faces) then interpolate_x (interpolation.f90:24-85) */
static double nut_edge(const orc_state *s, int i, int k) {
    const double *nt = s->nu_t;
    double a = WY0(1) * nt[IP(i, 1, k)] + WY1(1) * nt[IP(i, 2, k)];
    double b = WY0(1) * nt[IP(i + 1, 1, k)] + WY1(1) * nt[IP(i + 1, 2, k)];
    return 0.5 * (a + b);
}

/* Suv = (dU/dy + dV/dx)/2 at the cell edge (i, y(j), k) */
static double suv(const orc_state *s, int i, int j, int k) {
    return 0.5 * ((s->U[IU(i, j + 1, k)] - s->U[IU(i, j, k)]) / (YG(j + 1) - YG(j)) +
                  (s->V[IV(i + 1, j, k)] - s->V[IV(i, j, k)]) / (XG(i + 1) - XG(i)));
}

/* avg_nu_t(1, j) interpolated to the face y(j) (interpolate_y, di = 2) */
static double avg_face(const orc_state *s, int j) {
    const double *avg = s->avg_nu_t;
    return WY0(j) * avg[(size_t)s->nxg * (j - 1)] + WY1(j) * avg[(size_t)s->nxg * j];
}

static void fill3(orc_state *s, double a) {
    for (int i = 0; i < s->nx; i++) s->alpha_x[i] = a;
    for (int i = 0; i < s->nxg; i++) s->alpha_y[i] = a;
    for (int i = 0; i < s->nxg; i++) s->alpha_z[i] = a;
}

/* compute_alpha_momentum_wall_model (wallmodel.f90:674-764) */
static void model_momentum_wall(orc_state *s) {
    const int nx = s->nx, ny = s->ny, nyg = s->nyg, nzg = s->nzg;
    const double nb = avg_face(s, 1), nt = avg_face(s, ny);
    const double cnt = (double)(2 * (nx - 2) * (nzg - 2));
    double mtau = 0.0, dUV = 0.0, dU = 0.0;
    for (int k = 2; k <= nzg - 1; k++)
        for (int i = 2; i <= nx - 1; i++) {
            double tb = -2.0 * nb * suv(s, i, 1, k);
            double tt = -2.0 * nt * suv(s, i, ny, k);
            mtau += tb - tt;
        }
    mtau = mtau / cnt;
    const double *U = s->U, *V = s->V;
    for (int k = 2; k <= nzg - 1; k++)
        for (int i = 2; i <= nx - 1; i++) {
            /* interpolate_x(V, term_1, 1): V at the cell edges x(i) */
            double v1 = 0.5 * (V[IV(i, 1, k)] + V[IV(i + 1, 1, k)]);
            double v2 = 0.5 * (V[IV(i, 2, k)] + V[IV(i + 1, 2, k)]);
            double vn = 0.5 * (V[IV(i, ny, k)] + V[IV(i + 1, ny, k)]);
            double vn1 = 0.5 * (V[IV(i, ny - 1, k)] + V[IV(i + 1, ny - 1, k)]);
            double bot = (U[IU(i, 2, k)] - U[IU(i, 1, k)]) / (YG(2) - YG(1)) * (v2 - v1) / (Y(2) - Y(1));
            double top = -((U[IU(i, nyg, k)] - U[IU(i, nyg - 1, k)]) / (YG(nyg) - YG(nyg - 1)) * (vn - vn1) /
                           (Y(ny) - Y(ny - 1)));
            dUV += bot + top;
            dU += (U[IU(i, 2, k)] - U[IU(i, 1, k)]) / (YG(2) - YG(1)) -
                  (U[IU(i, nyg, k)] - U[IU(i, nyg - 1, k)]) / (YG(nyg) - YG(nyg - 1));
        }
    dUV = dUV / cnt;
    dU = dU / cnt;
    const double Ly = Y(ny) - Y(1);
    double a = (s->nu * dU - s->dPdx * Ly / 2.0 - mtau) / dUV;
    a = fmax(a, 0.0);
    fill3(s, pow(a, 0.5));
}

/* compute_alpha_u_momentum_wall_model (wallmodel.f90:1444-1500) */
static void model_u_momentum(orc_state *s) {
    const int nx = s->nx, nzg = s->nzg;
    for (int i = 1; i <= nx; i++) {
        double m = 0.0;
        for (int k = 2; k <= nzg - 1; k++) m += -2.0 * nut_edge(s, i, k) * suv(s, i, 1, k);
        m = m / (double)(nzg - 2);
        double u = s->utau_model[i - 1];
        double a = fabs(u) < 1e-10 ? 0.0 : (pow(u, 2.0) + m) / s->nu;
        s->alpha_x[i - 1] = fmax(a, 0.0);
    }
    memset(s->alpha_y, 0, sizeof(double) * s->nxg);
    memset(s->alpha_z, 0, sizeof(double) * s->nxg);
}

/* compute_exact_Neumann_from_Cf (wallmodel.f90:2003-2066), Uinf = 1 */
static void model_exact_neumann(orc_state *s) {
    const int nx = s->nx, nzg = s->nzg;
    for (int i = 1; i <= nx; i++) {
        double Cf = 0.027 * pow(1.0 * X(i) / s->nu, -1.0 / 7.0);
        double tw = 0.5 * Cf * pow(1.0, 2.0);
        double m = 0.0;
        for (int k = 2; k <= nzg - 1; k++) m += (s->nu + nut_edge(s, i, k)) / tw;
        s->alpha_x[i - 1] = 1.0 / (m / (double)(nzg - 2));
    }
    memset(s->alpha_y, 0, sizeof(double) * s->nxg);
    memset(s->alpha_z, 0, sizeof(double) * s->nxg);
}

/* compute_alpha_v_from_Cf (wallmodel.f90:1532-1697, model 9) and
 * compute_alpha_u_from_Cf (:1263-1424, model 12) */
static void model_cf_slip(orc_state *s) {
    const int nx = s->nx, nxg = s->nxg, nzg = s->nzg, m12 = s->iwall == 12;
    const double *U = s->U, *V = m12 ? s->V : s->Vo;
    const double Uinf = U[IU(2, s->nyg, 2)];
    const double c2 = (double)(2 * (nzg - 2)), c1 = (double)(nzg - 2);
    double *h = orc__dalloc(nx);
    for (int i = 1; i <= nx; i++) {
        double Cf = 0.027 * pow(Uinf * X(i) / s->nu, -1.0 / 7.0);
        double tw = 0.5 * Cf * pow(Uinf, 2.0);
        /* Part 0: alpha_x from the viscous fraction */
        double um = 0.0, ur = 0.0;
        for (int k = 2; k <= nzg - 1; k++) {
            um += U[IU(i, 1, k)] + U[IU(i, 2, k)];
            ur += U[IU(i, 3, k)] + U[IU(i, 4, k)];
        }
        um = um / c2;
        ur = ur / c2;
        double a;
        if (s->nut_flag == 2) {
            double fd = s->nu * ur / (tw * Y(3));
            a = s->nu * um / (fd * tw);
        } else {
            a = s->nu * um / (s->frac_vis * tw);
        }
        if (a < 0.0) a = 1e-3;
        if (fabs(a) < 1e-8) a = 1e-3;
        s->alpha_x[i - 1] = a;
        /* Parts 1-3: <tau_uv>, <dU/dy dV/dy> (9: <U dVo/dy>), <dU/dy> at the wall */
        double mt = 0.0, duv = 0.0, du = 0.0;
        for (int k = 2; k <= nzg - 1; k++) {
            double g = (U[IU(i, 2, k)] - U[IU(i, 1, k)]) / (YG(2) - YG(1));
            double sv = m12 ? suv(s, i, 1, k) : 0.5 * ((U[IU(i, 2, k)] - U[IU(i, 1, k)]) / (YG(2) - YG(1)));
            mt += -2.0 * nut_edge(s, i, k) * sv;
            double v1 = 0.5 * (V[IV(i, 1, k)] + V[IV(i + 1, 1, k)]);
            double v2 = 0.5 * (V[IV(i, 2, k)] + V[IV(i + 1, 2, k)]);
            duv += m12 ? (U[IU(i, 2, k)] - U[IU(i, 1, k)]) / (YG(2) - YG(1)) * (v2 - v1) / (Y(2) - Y(1))
                       : 0.5 * (U[IU(i, 2, k)] + U[IU(i, 1, k)]) * (v2 - v1) / (Y(2) - Y(1));
            du += g;
        }
        mt = mt / c1;
        duv = duv / c1;
        du = du / c1;
        /* Part 4's per-column quotient */
        h[i - 1] = m12 ? (s->nu * du - tw - mt) / (a * duv) : (s->nu * du - tw - mt + um * s->beta_y) / duv;
    }
    for (int i = 2; i <= nxg - 1; i++) s->alpha_y[i - 1] = 0.5 * (h[i - 2] + h[i - 1]);
    s->alpha_y[0] = s->alpha_y[1];
    s->alpha_y[nxg - 1] = s->alpha_y[nxg - 2];
    if (m12)
        for (int i = 1; i <= nxg; i++) {
            double a = s->alpha_y[i - 1];
            if (fabs(a) > 2e-3) s->alpha_y[i - 1] = 2e-3 * a / fabs(a);
        }
    memcpy(s->alpha_z, s->alpha_y, sizeof(double) * nxg);
    free(h);
}

/* the factor log(Delta utau/nu e^(B kappa)) of compute_log_layer_wall_model,
 * kappa = 0.38, B = 5.2, Delta = y(2) - y(1), utau = dPdx^0.5 */
static double log_factor(const orc_state *s, double *utau) {
    const double kappa = 0.38, B = 5.20, e = exp(1.0);
    *utau = pow(s->dPdx, 0.5);
    return log((Y(2) - Y(1)) * *utau / s->nu * pow(e, B * kappa));
}

void orc_wall_model(orc_state *s) {
    if (!s->alpha_x) return;
    if (s->istress == 1)
        law_of_the_wall(s);
    else
        memcpy(s->utau_model, s->utau_ref, sizeof(double) * s->nx);
    if (s->iwall == 4) {
        model_momentum_wall(s);
    } else if (s->iwall == 5) {
        const double omega = 2.0 * PI_D * s->freq_mult / (Y(2) - Y(1)) * pow(s->dPdx, 0.5);
        const double f = 1.0 + s->alpha_std * sin(omega * s->t);
        for (int i = 0; i < s->nx; i++) s->alpha_x[i] = s->alpha_mean_x * f;
        for (int i = 0; i < s->nxg; i++) s->alpha_y[i] = s->alpha_mean_y * f;
        for (int i = 0; i < s->nxg; i++) s->alpha_z[i] = s->alpha_mean_z * f;
    } else if (s->iwall == 6 || s->iwall == 8) {
        double utau;
        const double L = log_factor(s, &utau);
        const double a = s->iwall == 6 ? (avg_face(s, 1) + s->nu) / (0.38 * utau) * L : (Y(2) - Y(1)) * L;
        fill3(s, a);
    } else if (s->iwall == 9 || s->iwall == 12) {
        model_cf_slip(s);
    } else if (s->iwall == 11) {
        model_u_momentum(s);
    } else if (s->iwall == 15) {
        model_exact_neumann(s);
    }
}

/* bottom-wall conditions, apply_boundary_conditions :73-108 */
void orc__wall_bc(orc_state *s) {
    const int nx = s->nx, nxg = s->nxg, nz = s->nz, nzg = s->nzg;
    double *U = s->U, *V = s->V, *W = s->W;
    const int iw = s->alpha_x ? s->iwall : 0;
    const double dyg = YG(2) - YG(1);
    if (iw == 0) {
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nx; i++) U[IU(i, 1, k)] = -U[IU(i, 2, k)];
    } else if (iw == 11 || iw == 15) {
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nx; i++) U[IU(i, 1, k)] = U[IU(i, 2, k)] - s->alpha_x[i - 1] * dyg;
    } else {
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nx; i++) {
                double a = s->alpha_x[i - 1];
                U[IU(i, 1, k)] = (2.0 * a / dyg - 1.0) * U[IU(i, 2, k)] / (2.0 * a / dyg + 1.0);
            }
    }
    if (iw == 0 || iw == 11 || iw == 15) {
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nxg; i++) V[IV(i, 1, k)] = 0.0;
    } else {
        const double *Vo = s->Vo;
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nxg; i++)
                V[IV(i, 1, k)] = s->alpha_y[i - 1] * (Vo[IV(i, 2, k)] - Vo[IV(i, 1, k)]) / (Y(2) - Y(1));
        /* beta_y: mean over i = 2..nxg-2, k = 2..nzg-2 (single rank, kk = 2) */
        double b = 0.0;
        for (int k = 2; k <= nzg - 2; k++)
            for (int i = 2; i <= nxg - 2; i++) b += V[IV(i, 1, k)];
        b = b / (double)((nxg - 3) * (nzg - 3));
        s->beta_y = b;
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nxg; i++) V[IV(i, 1, k)] = V[IV(i, 1, k)] - b;
    }
    if (iw == 0 || iw == 11 || iw == 15) {
        for (int k = 1; k <= nz; k++)
            for (int i = 1; i <= nxg; i++) W[IW(i, 1, k)] = -W[IW(i, 2, k)];
    } else {
        for (int k = 1; k <= nz; k++)
            for (int i = 1; i <= nxg; i++) {
                double a = s->alpha_z[i - 1];
                W[IW(i, 1, k)] = (2.0 * a / dyg - 1.0) * W[IW(i, 2, k)] / (2.0 * a / dyg + 1.0);
            }
    }
}
