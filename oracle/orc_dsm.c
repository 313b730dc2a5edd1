/*
 * CPU-ORACLE TEST INFRASTRUCTURE — not part of the product.
 *
 * Dynamic Smagorinsky eddy viscosity, z-averaged (LES = 2): FP64
 * restatement of sgs_dynamic_Smagorinsky (subgrid.f90:261-414) with
 * filter_xzy (:140-187), compute_Sij (:672-730), and the single-rank
 * behaviour of the helpers it calls: apply_Neumann_bc_x
 * (boundary_conditions.f90:326-344), apply_periodic_bc_z (:732-757),
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
 * linear fields in x and z exactly and has unit sum), and nu_t >= 0.
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

void orc_dsm(orc_state *s) {
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
    neumann_x(term, nxg, nyg, nzg, nxg);
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
        neumann_x(ten[c], nxg, nyg, nzg, nxg);
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
    neumann_x(Uf, nx, nyg, nzg, nx);
    periodic_z(s, Uf, nx, nyg, 1);
    neumann_x(Vf, nxg, ny, nzg, nxg);
    periodic_z(s, Vf, nxg, ny, 2);
    neumann_x(Wf, nxg, nyg, nz, nxg);
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
        neumann_x(term, nxg, nyg, nzg, nxg);
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
