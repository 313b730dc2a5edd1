/*
 * CPU-ORACLE TEST INFRASTRUCTURE — not part of the product.
 *
 * Smagorinsky-type eddy viscosities: FP64 restatements of
 * sgs_dynamic_Smagorinsky (LES = 2, z-averaged, subgrid.f90:261-414),
 * sgs_dynamic_Smagorinsky_noaverage (LES = 3, per cell, :436-583) and
 * sgs_Smagorinsky (LES = 1, constant coefficient with van Driest damping,
 * :609-662), with filter_xzy (:140-187), compute_Sij (:672-730), and the
 * single-rank behaviour of the helpers they call: apply_Neumann_bc_x
 * (boundary_conditions.f90:326-344), apply_periodic_bc_x (:295-314),
 * apply_periodic_bc_z (:732-757),
 * update_ghost_interior_planes (:1185-1354, a no-op on one rank),
 * apply_Neumann_bc_y / apply_Dirichlet_bc_y (:1072-1087, :1020-1035), and the
 * face-to-centre interpolations (interpolation.f90:24-114).
 *
 * The CPU build's semantics are followed: the interpolations write the same
 * host arrays that the filters then read.  Work arrays carry the reference's
 * full allocation extents; entries the reference leaves stale are never read
 * (each read below is of an entry written earlier in the same call).
 *
 * Pinned by tests/test_oracle_dsm.py: uniform-strain and linear fields give
 * nu_t = 0 (L_ij = 0), a brute-force numpy restatement on tiny grids written
 * independently from the formulas, the filter's moments (it reproduces
 * linear fields in x and z exactly and has unit sum), and nu_t >= 0; LES = 3
 * against LES = 2 on fields uniform in z and x-periodic (where the z average
 * and the x fixes change nothing), and LES = 1 against the closed form
 * (Cs Delta f)^2 |S| of uniform-shear fields.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

/* Array of the P allocation extent (nxg, nyg, nzg), 1-based */
#define A3(a, i, j, k) (a)[IP(i, j, k)]

/*
 * filter_xzy (subgrid.f90:140-187) applied to the section
 * F(i0:i0+n1-1, j0:j0+n2-1, k0:k0+n3-1) of an array with leading extents
 * (ld1, ld2), writing O(oi0:oi0+n1-3, oj0:oj0+n2-1, ok0:ok0+n3-3) of an array
 * with leading extents (lo1, lo2).  Loop nest and accumulation order as in
 * the reference: ii, i, kk, k, then j.
 */
static void filter_xzy(const double *F, int ld1, int ld2, int i0, int n1, int j0, int n2, int k0, int n3,
                       double *O, int lo1, int lo2, int oi0, int oj0, int ok0) {
#define FF(a, b, c) \
    F[(size_t)(i0 + (a)-2) + (size_t)ld1 * ((size_t)(j0 + (b)-2) + (size_t)ld2 * (size_t)(k0 + (c)-2))]
#define OO(a, b, c) \
    O[(size_t)(oi0 + (a)-2) + (size_t)lo1 * ((size_t)(oj0 + (b)-2) + (size_t)lo2 * (size_t)(ok0 + (c)-2))]
    double fil[3][3];
    fil[0][0] = 1.0 / 36.0;
    fil[2][2] = 1.0 / 36.0;
    fil[0][2] = 1.0 / 36.0;
    fil[2][0] = 1.0 / 36.0;
    fil[1][2] = 1.0 / 9.0;
    fil[2][1] = 1.0 / 9.0;
    fil[1][0] = 1.0 / 9.0;
    fil[0][1] = 1.0 / 9.0;
    fil[1][1] = 4.0 / 9.0;
    for (int c = 1; c <= n3 - 2; c++)
        for (int b = 1; b <= n2; b++)
            for (int a = 1; a <= n1 - 2; a++) OO(a, b, c) = 0.0;
    for (int ii = 2; ii <= n1 - 1; ii++)
        for (int i = -1; i <= 1; i++)
            for (int kk = 2; kk <= n3 - 1; kk++)
                for (int k = -1; k <= 1; k++) {
                    const double f = fil[i + 1][k + 1];
                    for (int j = 2; j <= n2 - 1; j++)
                        OO(ii - 1, j, kk - 1) =
                            OO(ii - 1, j, kk - 1) + (1.0 / 6.0 * FF(ii + i, j - 1, kk + k) +
                                                     2.0 / 3.0 * FF(ii + i, j, kk + k) +
                                                     1.0 / 6.0 * FF(ii + i, j + 1, kk + k)) *
                                                        f;
                    OO(ii - 1, 1, kk - 1) =
                        OO(ii - 1, 1, kk - 1) + (2.0 / 3.0 * FF(ii + i, 1, kk + k) + 1.0 / 3.0 * FF(ii + i, 2, kk + k)) * f;
                    OO(ii - 1, n2, kk - 1) =
                        OO(ii - 1, n2, kk - 1) +
                        (1.0 / 3.0 * FF(ii + i, n2 - 1, kk + k) + 2.0 / 3.0 * FF(ii + i, n2, kk + k)) * f;
                }
#undef FF
#undef OO
}

/* apply_Neumann_bc_x on an array with first extent n1, id 1 (x faces, uses
 * nx) or 2 (x centres, uses nxg) */
static void neumann_x(double *F, int n1, int n2, int n3, int last) {
    for (int k = 0; k < n3; k++)
        for (int j = 0; j < n2; j++) {
            double *row = F + (size_t)n1 * ((size_t)j + (size_t)n2 * k);
            row[0] = row[1];
            row[last - 1] = row[last - 2];
        }
}

/* apply_periodic_bc_x: id 1 (x faces, extent nx): F(1)=F(nx-1), F(nx)=F(2);
 * otherwise (x centres, extent nxg): F(1)=F(nxg-2), F(nxg-1)=F(2),
 * F(nxg)=F(3) -- the last interior centre is overwritten too */
static void periodic_x(double *F, int n1, int n2, int n3, int faces) {
    for (int k = 0; k < n3; k++)
        for (int j = 0; j < n2; j++) {
            double *row = F + (size_t)n1 * ((size_t)j + (size_t)n2 * k);
            if (faces) {
                row[0] = row[n1 - 2];
                row[n1 - 1] = row[1];
            } else {
                row[0] = row[n1 - 3];
                row[n1 - 2] = row[1];
                row[n1 - 1] = row[2];
            }
        }
}

/* the x fix of the dynamic models: Neumann (LES = 2) or periodic (LES = 3) */
static void fix_x(int per, double *F, int n1, int n2, int n3, int faces) {
    if (per)
        periodic_x(F, n1, n2, n3, faces);
    else
        neumann_x(F, n1, n2, n3, n1);
}

/* apply_periodic_bc_z, single rank; id 1/2/4 (z centres) or 3 (z faces) */
static void periodic_z(orc_state *s, double *F, int n1, int n2, int id) {
    size_t pl = (size_t)n1 * n2;
    if (id == 3) {
        memcpy(F, F + pl * (s->nz - 2), pl * sizeof(double));
        memcpy(F + pl * (s->nz - 1), F + pl * 1, pl * sizeof(double));
    } else {
        memcpy(F, F + pl * (s->nzg - 3), pl * sizeof(double));
        memcpy(F + pl * (s->nzg - 2), F + pl * 1, pl * sizeof(double));
        memcpy(F + pl * (s->nzg - 1), F + pl * 2, pl * sizeof(double));
    }
}

/* compute_Sij (subgrid.f90:672-730): Sij(2:nxg,2:nyg,2:nzg,6) held in six
 * arrays of the P extent, S(2:nxg-1,2:nyg-1,2:nzg-1) likewise */
static void compute_sij(orc_state *s, const double *U, const double *V, const double *W, double *Sij[6],
                        double *S) {
    int nxg = s->nxg, nyg = s->nyg, nzg = s->nzg;
#define U_(i, j, k) U[IU(i, j, k)]
#define V_(i, j, k) V[IV(i, j, k)]
#define W_(i, j, k) W[IW(i, j, k)]
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                A3(Sij[0], i, j, k) = (U_(i, j, k) - U_(i - 1, j, k)) / (X(i) - X(i - 1));
                A3(Sij[1], i, j, k) = (V_(i, j, k) - V_(i, j - 1, k)) / (Y(j) - Y(j - 1));
                A3(Sij[2], i, j, k) = (W_(i, j, k) - W_(i, j, k - 1)) / (Z(k) - Z(k - 1));
            }
    for (int k = 2; k <= nzg; k++)
        for (int j = 2; j <= nyg; j++)
            for (int i = 2; i <= nxg; i++) {
                A3(Sij[3], i, j, k) = 0.5 * ((U_(i - 1, j, k) - U_(i - 1, j - 1, k)) / (YG(j) - YG(j - 1)) +
                                             (V_(i, j - 1, k) - V_(i - 1, j - 1, k)) / (XG(i) - XG(i - 1)));
                A3(Sij[4], i, j, k) = 0.5 * ((U_(i - 1, j, k) - U_(i - 1, j, k - 1)) / (ZG(k) - ZG(k - 1)) +
                                             (W_(i, j, k - 1) - W_(i - 1, j, k - 1)) / (XG(i) - XG(i - 1)));
                A3(Sij[5], i, j, k) = 0.5 * ((V_(i, j - 1, k) - V_(i, j - 1, k - 1)) / (ZG(k) - ZG(k - 1)) +
                                             (W_(i, j, k - 1) - W_(i, j - 1, k - 1)) / (YG(j) - YG(j - 1)));
            }
    /* edges to centres, in place; every neighbour read is not yet updated */
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                A3(Sij[3], i, j, k) = 0.25 * (A3(Sij[3], i, j, k) + A3(Sij[3], i + 1, j, k) +
                                              A3(Sij[3], i, j + 1, k) + A3(Sij[3], i + 1, j + 1, k));
                A3(Sij[4], i, j, k) = 0.25 * (A3(Sij[4], i, j, k) + A3(Sij[4], i + 1, j, k) +
                                              A3(Sij[4], i, j, k + 1) + A3(Sij[4], i + 1, j, k + 1));
                A3(Sij[5], i, j, k) = 0.25 * (A3(Sij[5], i, j, k) + A3(Sij[5], i, j + 1, k) +
                                              A3(Sij[5], i, j, k + 1) + A3(Sij[5], i, j + 1, k + 1));
            }
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                double a = 0.0;
                for (int c = 0; c < 6; c++) {
                    double v = A3(Sij[c], i, j, k);
                    a = a + (c < 3 ? 2.0 : 4.0) * v * v;
                }
                A3(S, i, j, k) = sqrt(a);
            }
#undef U_
#undef V_
#undef W_
}

/* LES = 2 (noavg = 0): sgs_dynamic_Smagorinsky; LES = 3 (noavg = 1):
 * sgs_dynamic_Smagorinsky_noaverage, which differs in its x fixes
 * (apply_periodic_bc_x where LES = 2 has apply_Neumann_bc_x), in the ratio
 * <L:M>/<M:M> taken per cell instead of over z, in the periodic x and z
 * fixes of nu_t before clipping, and in avg_nu_t = 0 */
static void dsm_common(orc_state *s, int noavg) {
    const int nx = s->nx, ny = s->ny, nz = s->nz, nxg = s->nxg, nyg = s->nyg, nzg = s->nzg;
    const size_t nP = (size_t)nxg * nyg * nzg;
    const double *U = s->U, *V = s->V, *W = s->W;
    double *work = (double *)calloc(nP * 28 + (size_t)nx * nyg * nzg + (size_t)nxg * ny * nzg +
                                        (size_t)nxg * nyg * nz,
                                    sizeof(double));
    double *term = work, *term_1 = work + nP, *term_2 = work + 2 * nP, *S = work + 3 * nP;
    double *Sij[6], *Lij[6], *Mij[6], *ten[6];
    for (int c = 0; c < 6; c++) {
        Sij[c] = work + (4 + c) * nP;
        Lij[c] = work + (10 + c) * nP;
        Mij[c] = work + (16 + c) * nP;
        ten[c] = work + (22 + c) * nP;
    }
    double *Uf = work + 28 * nP;
    double *Vf = Uf + (size_t)nx * nyg * nzg;
    double *Wf = Vf + (size_t)nxg * ny * nzg;
    double *nu_t = s->nu_t, *avg = s->avg_nu_t;

    /* ---- Part 1: Leonard term (subgrid.f90:275-323) ---- */
    /* interpolate_x(U, term(2:nxg-1,:,:)), faces to centres */
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++)
            for (int i = 2; i <= nxg - 1; i++) A3(term, i, j, k) = 0.5 * (U[IU(i - 1, j, k)] + U[IU(i, j, k)]);
    /* interpolate_y(V, term_1(:,2:nyg-1,:)) */
    for (int k = 1; k <= nzg; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 1; i <= nxg; i++) A3(term_1, i, j, k) = 0.5 * (V[IV(i, j - 1, k)] + V[IV(i, j, k)]);
    /* interpolate_z(W, term_2(:,:,2:nzg)); its last plane is zeroed */
    for (int k = 2; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++)
            for (int i = 1; i <= nxg; i++)
                A3(term_2, i, j, k) = (k <= nz) ? 0.5 * (W[IW(i, j, k - 1)] + W[IW(i, j, k)]) : 0.0;
    fix_x(noavg, term, nxg, nyg, nzg, 0);
    periodic_z(s, term_2, nxg, nyg, 4);

    /* products at cell centres (Sij as placeholder), copied to ten_buf */
    for (int k = 2; k <= nzg; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg; i++) {
                double a = A3(term, i, j, k), b = A3(term_1, i, j, k), c = A3(term_2, i, j, k);
                A3(ten[0], i, j, k) = a * a;
                A3(ten[1], i, j, k) = b * b;
                A3(ten[2], i, j, k) = c * c;
                A3(ten[3], i, j, k) = a * b;
                A3(ten[4], i, j, k) = a * c;
                A3(ten[5], i, j, k) = b * c;
            }
    for (int c = 0; c < 6; c++) {
        fix_x(noavg, ten[c], nxg, nyg, nzg, 0);
        periodic_z(s, ten[c], nxg, nyg, 4);
        /* hat(ui uj): filter_xzy(ten_buf(1:nxg,2:nyg-1,1:nzg), Lij(2:nxg-1,2:nyg-1,2:nzg-1)) */
        filter_xzy(ten[c], nxg, nyg, 1, nxg, 2, nyg - 2, 1, nzg, Lij[c], nxg, nyg, 2, 2, 2);
    }
    /* hat(ui) */
    filter_xzy(term, nxg, nyg, 1, nxg, 2, nyg - 2, 1, nzg, Sij[0], nxg, nyg, 2, 2, 2);
    filter_xzy(term_1, nxg, nyg, 1, nxg, 2, nyg - 2, 1, nzg, Sij[1], nxg, nyg, 2, 2, 2);
    filter_xzy(term_2, nxg, nyg, 1, nxg, 2, nyg - 2, 1, nzg, Sij[2], nxg, nyg, 2, 2, 2);
    static const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    for (int c = 0; c < 6; c++)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++)
                    A3(Lij[c], i, j, k) = A3(Lij[c], i, j, k) - A3(Sij[pa[c]], i, j, k) * A3(Sij[pb[c]], i, j, k);

    /* ---- Part 2: Mij (subgrid.f90:325-367) ---- */
    filter_xzy(U, nx, nyg, 1, nx, 1, nyg, 1, nzg, Uf, nx, nyg, 2, 1, 2);
    filter_xzy(V, nxg, ny, 1, nxg, 1, ny, 1, nzg, Vf, nxg, ny, 2, 1, 2);
    filter_xzy(W, nxg, nyg, 1, nxg, 1, nyg, 1, nz, Wf, nxg, nyg, 2, 1, 2);
    fix_x(noavg, Uf, nx, nyg, nzg, 1);
    periodic_z(s, Uf, nx, nyg, 1);
    fix_x(noavg, Vf, nxg, ny, nzg, 0);
    periodic_z(s, Vf, nxg, ny, 2);
    fix_x(noavg, Wf, nxg, nyg, nz, 0);
    periodic_z(s, Wf, nxg, nyg, 3);

    compute_sij(s, Uf, Vf, Wf, Sij, S);
    /* fil_size**2 with fil_size = 2 (set by filter_xzy, subgrid.f90:160) */
    const double fil2 = pow(2.0, 2.0);
    for (int c = 0; c < 6; c++)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++)
                    A3(Mij[c], i, j, k) = fil2 * A3(S, i, j, k) * A3(Sij[c], i, j, k);

    compute_sij(s, U, V, W, Sij, S);
    for (int c = 0; c < 6; c++) {
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) A3(term, i, j, k) = A3(S, i, j, k) * A3(Sij[c], i, j, k);
        fix_x(noavg, term, nxg, nyg, nzg, 0);
        periodic_z(s, term, nxg, nyg, 4);
        filter_xzy(term, nxg, nyg, 1, nxg, 2, nyg - 2, 1, nzg, Sij[c], nxg, nyg, 2, 2, 2);
    }
    for (int c = 0; c < 6; c++)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) A3(Mij[c], i, j, k) = A3(Mij[c], i, j, k) - A3(Sij[c], i, j, k);

    /* ---- Part 3: nu_t (subgrid.f90:369-412) ---- */
    memset(nu_t, 0, nP * sizeof(double));
    for (int c = 0; c < 6; c++)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    double ml = (c < 3) ? A3(Mij[c], i, j, k) * A3(Lij[c], i, j, k)
                                        : 2.0 * A3(Mij[c], i, j, k) * A3(Lij[c], i, j, k);
                    A3(nu_t, i, j, k) = A3(nu_t, i, j, k) + ml;
                }
    if (noavg) {
        /* subgrid.f90:545-581 */
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    double m1 = A3(Mij[0], i, j, k), m2 = A3(Mij[1], i, j, k), m3 = A3(Mij[2], i, j, k);
                    double m4 = A3(Mij[3], i, j, k), m5 = A3(Mij[4], i, j, k), m6 = A3(Mij[5], i, j, k);
                    A3(nu_t, i, j, k) = A3(nu_t, i, j, k) /
                                        (m1 * m1 + m2 * m2 + m3 * m3 + 2.0 * (m4 * m4 + m5 * m5 + m6 * m6));
                }
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) A3(nu_t, i, j, k) = -0.5 * A3(nu_t, i, j, k) * A3(S, i, j, k);
        periodic_x(nu_t, nxg, nyg, nzg, 0);
        periodic_z(s, nu_t, nxg, nyg, 4);
        memset(avg, 0, sizeof(double) * (size_t)nxg * nyg);
    } else {
    const double nz3 = (double)(nzg - 3);
#define AVG(i, j) avg[(size_t)((i)-1) + (size_t)nxg * (size_t)((j)-1)]
    for (int i = 2; i <= nxg - 1; i++)
        for (int j = 2; j <= nyg - 1; j++) {
            double a = 0.0;
            for (int k = 2; k <= nzg - 2; k++) a += A3(nu_t, i, j, k);
            AVG(i, j) = a / nz3;
        }
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                double m1 = A3(Mij[0], i, j, k), m2 = A3(Mij[1], i, j, k), m3 = A3(Mij[2], i, j, k);
                double m4 = A3(Mij[3], i, j, k), m5 = A3(Mij[4], i, j, k), m6 = A3(Mij[5], i, j, k);
                A3(nu_t, i, j, k) = (m1 * m1 + m2 * m2 + m3 * m3 + 2.0 * (m4 * m4 + m5 * m5 + m6 * m6));
            }
    for (int i = 2; i <= nxg - 1; i++)
        for (int j = 2; j <= nyg - 1; j++) {
            double a = 0.0;
            for (int k = 2; k <= nzg - 2; k++) a += A3(nu_t, i, j, k);
            AVG(i, j) = AVG(i, j) / (a / nz3);
        }
    for (int i = 2; i <= nxg - 1; i++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int k = 2; k <= nzg - 1; k++) A3(nu_t, i, j, k) = -0.5 * AVG(i, j) * A3(S, i, j, k);
#undef AVG
    }
    /* clipping (Max(nu_t, 0): a NaN ratio becomes 0), then the y ghosts */
    for (size_t q = 0; q < nP; q++) nu_t[q] = fmax(nu_t[q], 0.0);
    for (int k = 1; k <= nzg; k++)
        for (int i = 1; i <= nxg; i++) {
            if (s->nut_dirichlet == 1) {
                A3(nu_t, i, 1, k) = -A3(nu_t, i, 2, k);
                A3(nu_t, i, nyg, k) = -A3(nu_t, i, nyg - 1, k);
            } else {
                A3(nu_t, i, 1, k) = A3(nu_t, i, 2, k);
                A3(nu_t, i, nyg, k) = A3(nu_t, i, nyg - 1, k);
            }
        }
    free(work);
}

void orc_dsm(orc_state *s) { dsm_common(s, 0); }

void orc_dsm_noavg(orc_state *s) { dsm_common(s, 1); }

/* sgs_Smagorinsky (subgrid.f90:609-662): nu_t = (Cs Delta f)^2 |S| on the
 * interior cells, Cs = 0.11, Delta = (dx dz dy_j)^(1/3) with dx, dz the
 * smallest spacings, van Driest f = 1 - exp(-y_wall/25/(nu/utau)) with
 * utau = dPdx^0.5 and y_wall the distance to the nearer wall (j < nyg/2:
 * bottom), f = 1 unless nutBC = 1.  The ghost entries of nu_t are not
 * written (they keep their previous values); the wall-parallel average
 * avg_nu_t(j) is clipped and gets the y ghost condition. */
void orc_smagorinsky(orc_state *s) {
    const int nxg = s->nxg, nyg = s->nyg, nzg = s->nzg;
    double *Sij[6], *S;
    const size_t nP = (size_t)nxg * nyg * nzg;
    double *work = (double *)calloc(nP * 7, sizeof(double));
    for (int c = 0; c < 6; c++) Sij[c] = work + (size_t)c * nP;
    S = work + 6 * nP;
    compute_sij(s, s->U, s->V, s->W, Sij, S);
    const double Cs = 0.11;
    const double utau_t = pow(s->dPdx, 0.5);
    double *nu_t = s->nu_t, *avg = s->avg_nu_t;
    for (int j = 2; j <= nyg - 1; j++) {
        double f;
        if (j < nyg / 2)
            f = 1.0 - exp(-(YG(j) - Y(1)) / 25.0 / (s->nu / utau_t));
        else
            f = 1.0 - exp(-(Y(s->ny) - YG(j)) / 25.0 / (s->nu / utau_t));
        if (s->nut_dirichlet == 0) f = 1.0;
        const double Delta = pow(s->dx * s->dz * (Y(j) - Y(j - 1)), 1.0 / 3.0);
        const double c = pow(Cs * Delta * f, 2.0);
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) A3(nu_t, i, j, k) = c * A3(S, i, j, k);
    }
    /* avg_nu_t(1, j, 1) in the first column of the (nxg, nyg) array */
    for (int j = 2; j <= nyg - 1; j++) {
        double a = 0.0;
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) a += A3(nu_t, i, j, k);
        avg[(size_t)nxg * (j - 1)] = fmax(a / (double)((nxg - 2) * (nzg - 2)), 0.0);
    }
    avg[0] = s->nut_dirichlet == 1 ? -avg[(size_t)nxg] : avg[(size_t)nxg];
    avg[(size_t)nxg * (nyg - 1)] = s->nut_dirichlet == 1 ? -avg[(size_t)nxg * (nyg - 2)] : avg[(size_t)nxg * (nyg - 2)];
    free(work);
}
