/*
 * CPU-ORACLE TEST INFRASTRUCTURE — not part of the product.
 *
 * Restatement of the reference's host-side initialisation that feeds the
 * per-step path: Blasius boundary tables, temporal inflow-mode
 * interpolation, and init_flow's grid and Blasius initial field.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

/* Search used by every Blasius lookup in the reference: first jj in 2..n
 * with eta_source(jj) > eta gives j0 = jj-1, j1 = jj (1-based); returns j0
 * (0 when none) and the weights. */
static int blasius_bracket(int n, const double *eta_s, double eta, double *w0, double *w1) {
    for (int jj = 2; jj <= n; jj++) {
        if (eta_s[jj - 1] > eta) {
            int j0 = jj - 1, j1 = jj;
            *w1 = (eta - eta_s[j0 - 1]) / (eta_s[j1 - 1] - eta_s[j0 - 1]);
            *w0 = 1.0 - *w1;
            return j0;
        }
    }
    return 0;
}

/* compute_blasius_solution_for_bc part 1, inflow_flag 1/3/4
 * (boundary_conditions.f90:1463-1627) */
void orc_blasius_tables(int nx, int ny, const double *x, const double *y, const double *xg,
                        const double *yg, double nu, int n, const double *eta, const double *f,
                        const double *df, double *U_inlet, double *V_inlet, double *W_inlet,
                        double *U_outlet, double *V_outlet, double *U_top, double *V_top,
                        double *W_top) {
    int nyg = ny + 1, nxg = nx + 1;
    double U_inf = 1.0, w0, w1;
    /* U inlet :1495-1515 */
    for (int j = 1; j <= nyg; j++) U_inlet[j - 1] = 1.0;
    for (int j = 1; j <= nyg; j++) {
        double e = yg[j - 1] * pow(U_inf / (2.0 * nu * x[0]), 0.5);
        int j0 = blasius_bracket(n, eta, e, &w0, &w1);
        U_inlet[j - 1] = j0 > 0 ? w0 * df[j0 - 1] + w1 * df[j0] : 1.0;
    }
    /* V inlet :1517-1539 */
    for (int j = 1; j <= ny; j++) V_inlet[j - 1] = 0.0;
    double Rex_ref = 1.0 * xg[0] / nu;
    for (int j = 1; j <= ny; j++) {
        double e = y[j - 1] * pow(U_inf / (2.0 * nu * xg[0]), 0.5);
        int j0 = blasius_bracket(n, eta, e, &w0, &w1);
        if (j0 > 0) {
            V_inlet[j - 1] = w0 * 1.0 / pow(2.0 * Rex_ref, 0.5) * (eta[j0 - 1] * df[j0 - 1] - f[j0 - 1]) +
                             w1 * 1.0 / pow(2.0 * Rex_ref, 0.5) * (eta[j0] * df[j0] - f[j0]);
        } else {
            double mx = V_inlet[0];
            for (int q = 1; q < ny; q++) mx = fmax(mx, V_inlet[q]);
            V_inlet[j - 1] = mx;
        }
    }
    for (int j = 1; j <= nyg; j++) W_inlet[j - 1] = 0.0;
    /* outlet :1566-1613 */
    for (int j = 1; j <= nyg; j++) {
        double e = yg[j - 1] * pow(U_inf / (2.0 * nu * x[nx - 2]), 0.5);
        int j0 = blasius_bracket(n, eta, e, &w0, &w1);
        U_outlet[j - 1] = j0 > 0 ? w0 * df[j0 - 1] + w1 * df[j0] : 1.0;
    }
    for (int j = 1; j <= ny; j++) V_outlet[j - 1] = 0.0;
    Rex_ref = 1.0 * xg[nxg - 2] / nu;
    for (int j = 1; j <= ny; j++) {
        double e = y[j - 1] * pow(U_inf / (2.0 * nu * xg[nxg - 2]), 0.5);
        int j0 = blasius_bracket(n, eta, e, &w0, &w1);
        if (j0 > 0) {
            V_outlet[j - 1] = w0 * 1.0 / pow(2.0 * Rex_ref, 0.5) * (eta[j0 - 1] * df[j0 - 1] - f[j0 - 1]) +
                              w1 * 1.0 / pow(2.0 * Rex_ref, 0.5) * (eta[j0] * df[j0] - f[j0]);
        } else {
            double mx = V_outlet[0];
            for (int q = 1; q < ny; q++) mx = fmax(mx, V_outlet[q]);
            V_outlet[j - 1] = mx;
        }
    }
    /* top :1617-1627 */
    for (int i = 1; i <= nx; i++) U_top[i - 1] = 1.0;
    for (int i = 1; i <= nxg; i++) {
        double R = 1.0 * xg[i - 1] / nu;
        V_top[i - 1] = 1.0 / pow(2.0 * R, 0.5) * (eta[n - 1] * df[n - 1] - f[n - 1]);
    }
    for (int i = 1; i <= nxg; i++) W_top[i - 1] = 0.0;
}

/* temporal-mode interpolation (boundary_conditions.f90:1734-1776).
 * Source arrays (ny_inlet, nmodes_total) with j fastest. */
void orc_interp_modes(int ny_inlet, const double *ymesh, int nmodes_total, const double *q_re,
                      const double *q_im, int nout, const double *ygrid, double *out_re,
                      double *out_im) {
    for (int j = 1; j <= nout; j++) {
        int j1 = 0;
        for (int jj = 2; jj <= ny_inlet; jj++) {
            if (ymesh[jj - 1] >= ygrid[j - 1]) {
                j1 = jj;
                break;
            }
            if (jj == ny_inlet) j1 = ny_inlet;
        }
        for (int q = 0; q < nmodes_total; q++) {
            size_t o = (size_t)(j - 1) + (size_t)nout * q;
            if (j1 == ny_inlet) {
                out_re[o] = 0.0;
                out_im[o] = 0.0;
            } else {
                int j0 = j1 - 1;
                size_t s0 = (size_t)(j0 - 1) + (size_t)ny_inlet * q;
                size_t s1 = (size_t)(j1 - 1) + (size_t)ny_inlet * q;
                double dy = ymesh[j1 - 1] - ymesh[j0 - 1];
                double fy = ygrid[j - 1] - ymesh[j0 - 1];
                out_re[o] = q_re[s0] + fy * (q_re[s1] - q_re[s0]) / dy;
                out_im[o] = q_im[s0] + fy * (q_im[s1] - q_im[s0]) / dy;
            }
        }
    }
}

/* init_flow (input_output.f90:130-279): sinh grid and Blasius field. */
void orc_init_flow(int nx, int ny, int nz, double Lx, double Ly, double Lz, double alpha,
                   double nu, int n, const double *eta, const double *f, const double *df,
                   double *x, double *y, double *z, double *U, double *V, double *W) {
    int nxg = nx + 1, nyg = ny + 1, nzg = nz + 1;
    /* x mesh :150-154 */
    for (int i = 1; i <= nx; i++) x[i - 1] = (double)(i - 1);
    double x1 = x[0];
    for (int i = 0; i < nx; i++) x[i] = x[i] - x1;
    double mx = x[0];
    for (int i = 1; i < nx; i++) mx = fmax(mx, x[i]);
    for (int i = 0; i < nx; i++) x[i] = 1.0 + x[i] / mx * Lx;
    /* z mesh :157-160 */
    for (int k = 1; k <= nz; k++) z[k - 1] = (double)(k - 1);
    double zl = z[nz - 2];
    for (int k = 0; k < nz; k++) z[k] = z[k] / zl * Lz;
    /* y mesh :164-174 */
    for (int i = 1; i <= ny; i++) y[i - 1] = (double)(i - 1);
    double y1 = y[0];
    for (int i = 0; i < ny; i++) y[i] = y[i] - y1;
    double my = y[0];
    for (int i = 1; i < ny; i++) my = fmax(my, y[i]);
    for (int i = 0; i < ny; i++) y[i] = y[i] / my;
    for (int i = 0; i < ny; i++) y[i] = sinh(alpha * y[i]) / sinh(alpha);
    y1 = y[0];
    for (int i = 0; i < ny; i++) y[i] = y[i] - y1;
    my = y[0];
    for (int i = 1; i < ny; i++) my = fmax(my, y[i]);
    for (int i = 0; i < ny; i++) y[i] = y[i] / my * Ly;
    /* centre grids :179-191 */
    double *ygi = (double *)calloc((size_t)nyg, sizeof(double));
    double *xgi = (double *)calloc((size_t)nxg, sizeof(double));
    for (int j = 1; j <= ny - 1; j++) ygi[j] = 0.5 * (y[j - 1] + y[j]);
    ygi[0] = ygi[1] - 2.0 * (ygi[1] - y[0]);
    ygi[nyg - 1] = ygi[ny - 1] + 2.0 * (y[ny - 1] - ygi[ny - 1]);
    for (int i = 1; i <= nx - 1; i++) xgi[i] = 0.5 * (x[i - 1] + x[i]);
    xgi[0] = xgi[1] - 2.0 * (xgi[1] - x[0]);
    xgi[nxg - 1] = xgi[nx - 1] + 2.0 * (x[nx - 1] - xgi[nx - 1]);
    double U_inf = 1.0, w0, w1;
    /* U :208-228 */
    for (int ii = 1; ii <= nx; ii++)
        for (int jj = 1; jj <= nyg; jj++) {
            double e = ygi[jj - 1] * sqrt(U_inf / (2.0 * nu * x[ii - 1]));
            int j0 = blasius_bracket(n, eta, e, &w0, &w1);
            double val = j0 > 0 ? U_inf * (w0 * df[j0 - 1] + w1 * df[j0]) : U_inf;
            for (int k = 0; k < nzg; k++) U[(size_t)(ii - 1) + (size_t)nx * ((jj - 1) + (size_t)nyg * k)] = val;
        }
    /* V :232-256 */
    for (int ii = 1; ii <= nxg; ii++) {
        double Rex_ref = U_inf * xgi[ii - 1] / nu;
        for (int jj = 1; jj <= ny; jj++) {
            double e = y[jj - 1] * sqrt(U_inf / (2.0 * nu * xgi[ii - 1]));
            int j0 = blasius_bracket(n, eta, e, &w0, &w1);
            double val;
            if (j0 > 0)
                val = U_inf / sqrt(2.0 * Rex_ref) *
                      (w0 * (eta[j0 - 1] * df[j0 - 1] - f[j0 - 1]) + w1 * (eta[j0] * df[j0] - f[j0]));
            else
                val = U_inf / sqrt(2.0 * Rex_ref) * (eta[n - 1] * df[n - 1] - f[n - 1]);
            for (int k = 0; k < nzg; k++) V[(size_t)(ii - 1) + (size_t)nxg * ((jj - 1) + (size_t)ny * k)] = val;
        }
    }
    /* W = 0 :259 */
    memset(W, 0, sizeof(double) * (size_t)nxg * nyg * nz);
    free(ygi);
    free(xgi);
}
