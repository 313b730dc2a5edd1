// Host-side set-up helpers of blgpu (not on the per-step path): Blasius
// profile and boundary tables, initial fields, grids and seeded noise.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/blgpu.h"

namespace {

// first jj (1-based, 2..n) with eta(jj) > e: returns j0 = jj-1 (0 if none)
int bracket(int n, const double *eta, double e, double *w0, double *w1) {
    for (int jj = 2; jj <= n; ++jj)
        if (eta[jj - 1] > e) {
            const int j0 = jj - 1;
            *w1 = (e - eta[j0 - 1]) / (eta[jj - 1] - eta[j0 - 1]);
            *w0 = 1.0 - *w1;
            return j0;
        }
    return 0;
}

void centers(int n, const double *x, std::vector<double> &xg) {
    std::vector<double> xm(n - 1);
    for (int i = 0; i < n - 1; ++i) xm[i] = 0.5 * (x[i] + x[i + 1]);
    xg.assign(n + 1, 0.0);
    for (int i = 0; i < n - 1; ++i) xg[i + 1] = xm[i];
    xg[0] = xm[0] - 2.0 * (xm[0] - x[0]);
    xg[n] = xm[n - 2] + 2.0 * (x[n - 1] - xm[n - 2]);
}

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

}  // namespace

extern "C" {

int bl_blasius_profile(int n, double eta_max, double fpp0, int sub, double *eta, double *f, double *df) {
    if (n < 2 || sub < 1 || !eta || !f || !df || !(eta_max > 0)) return BL_EINVAL;
    // y = (f, f', f''), y' = (f', f'', -f f'')
    double y0 = 0.0, y1 = 0.0, y2 = fpp0;
    const double H = eta_max / (n - 1), h = H / sub;
    auto rhs = [](double a, double b, double c, double &da, double &db, double &dc) {
        da = b;
        db = c;
        dc = -a * c;
    };
    eta[0] = 0.0;
    f[0] = y0;
    df[0] = y1;
    for (int i = 1; i < n; ++i) {
        for (int s = 0; s < sub; ++s) {
            double k1a, k1b, k1c, k2a, k2b, k2c, k3a, k3b, k3c, k4a, k4b, k4c;
            rhs(y0, y1, y2, k1a, k1b, k1c);
            rhs(y0 + 0.5 * h * k1a, y1 + 0.5 * h * k1b, y2 + 0.5 * h * k1c, k2a, k2b, k2c);
            rhs(y0 + 0.5 * h * k2a, y1 + 0.5 * h * k2b, y2 + 0.5 * h * k2c, k3a, k3b, k3c);
            rhs(y0 + h * k3a, y1 + h * k3b, y2 + h * k3c, k4a, k4b, k4c);
            y0 += h / 6.0 * (k1a + 2 * k2a + 2 * k3a + k4a);
            y1 += h / 6.0 * (k1b + 2 * k2b + 2 * k3b + k4b);
            y2 += h / 6.0 * (k1c + 2 * k2c + 2 * k3c + k4c);
        }
        eta[i] = H * i;
        f[i] = y0;
        df[i] = y1;
    }
    return BL_OK;
}

// boundary_conditions.f90:1463-1627 (inflow_flag 1/3/4; outlet tables are not
// used by the time step and are not produced)
int bl_blasius_tables(int nx, int ny, const double *x, const double *y, double nu, int n, const double *eta,
                      const double *f, const double *df, double *U_inlet, double *V_inlet, double *W_inlet,
                      double *U_top, double *V_top, double *W_top) {
    if (nx < 3 || ny < 3 || n < 2 || !x || !y || !eta || !f || !df || !U_inlet || !V_inlet || !W_inlet ||
        !U_top || !V_top || !W_top)
        return BL_EINVAL;
    const int nyg = ny + 1, nxg = nx + 1;
    std::vector<double> xg, yg;
    centers(nx, x, xg);
    centers(ny, y, yg);
    double w0, w1;
    for (int j = 0; j < nyg; ++j) {
        const double e = yg[j] * std::pow(1.0 / (2.0 * nu * x[0]), 0.5);
        const int j0 = bracket(n, eta, e, &w0, &w1);
        U_inlet[j] = j0 > 0 ? w0 * df[j0 - 1] + w1 * df[j0] : 1.0;
    }
    for (int j = 0; j < ny; ++j) V_inlet[j] = 0.0;
    const double rex = 1.0 * xg[0] / nu;
    for (int j = 0; j < ny; ++j) {
        const double e = y[j] * std::pow(1.0 / (2.0 * nu * xg[0]), 0.5);
        const int j0 = bracket(n, eta, e, &w0, &w1);
        if (j0 > 0) {
            V_inlet[j] = w0 * 1.0 / std::pow(2.0 * rex, 0.5) * (eta[j0 - 1] * df[j0 - 1] - f[j0 - 1]) +
                         w1 * 1.0 / std::pow(2.0 * rex, 0.5) * (eta[j0] * df[j0] - f[j0]);
        } else {
            double mx = V_inlet[0];
            for (int q = 1; q < ny; ++q) mx = std::fmax(mx, V_inlet[q]);
            V_inlet[j] = mx;
        }
    }
    for (int j = 0; j < nyg; ++j) W_inlet[j] = 0.0;
    for (int i = 0; i < nx; ++i) U_top[i] = 1.0;
    for (int i = 0; i < nxg; ++i) {
        const double r = 1.0 * xg[i] / nu;
        V_top[i] = 1.0 / std::pow(2.0 * r, 0.5) * (eta[n - 1] * df[n - 1] - f[n - 1]);
        W_top[i] = 0.0;
    }
    return BL_OK;
}

// input_output.f90:206-259 on the given grid
int bl_blasius_field(int nx, int ny, int nz, const double *x, const double *y, double nu, int n,
                     const double *eta, const double *f, const double *df, double *U, double *V, double *W) {
    if (nx < 3 || ny < 3 || nz < 3 || !x || !y || !U || !V || !W) return BL_EINVAL;
    const int nxg = nx + 1, nyg = ny + 1, nzg = nz + 1;
    std::vector<double> xg, yg;
    centers(nx, x, xg);
    centers(ny, y, yg);
    double w0, w1;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nyg; ++j) {
            const double e = yg[j] * std::sqrt(1.0 / (2.0 * nu * x[i]));
            const int j0 = bracket(n, eta, e, &w0, &w1);
            const double v = j0 > 0 ? 1.0 * (w0 * df[j0 - 1] + w1 * df[j0]) : 1.0;
            for (int k = 0; k < nzg; ++k) U[i + (size_t)nx * (j + (size_t)nyg * k)] = v;
        }
    for (int i = 0; i < nxg; ++i) {
        const double rex = 1.0 * xg[i] / nu;
        for (int j = 0; j < ny; ++j) {
            const double e = y[j] * std::sqrt(1.0 / (2.0 * nu * xg[i]));
            const int j0 = bracket(n, eta, e, &w0, &w1);
            double v;
            if (j0 > 0)
                v = 1.0 / std::sqrt(2.0 * rex) *
                    (w0 * (eta[j0 - 1] * df[j0 - 1] - f[j0 - 1]) + w1 * (eta[j0] * df[j0] - f[j0]));
            else
                v = 1.0 / std::sqrt(2.0 * rex) * (eta[n - 1] * df[n - 1] - f[n - 1]);
            for (int k = 0; k < nzg; ++k) V[i + (size_t)nxg * (j + (size_t)ny * k)] = v;
        }
    }
    for (size_t q = 0; q < (size_t)nxg * nyg * nz; ++q) W[q] = 0.0;
    return BL_OK;
}

int bl_add_noise(int nx, int ny, int nz, double amp, unsigned long long seed, double *U, double *V, double *W) {
    if (nx < 3 || ny < 3 || nz < 3 || !U || !V || !W) return BL_EINVAL;
    const int ext[3][3] = {{nx, ny + 1, nz + 1}, {nx + 1, ny, nz + 1}, {nx + 1, ny + 1, nz}};
    double *fld[3] = {U, V, W};
    const uint64_t key = splitmix64(seed);
    for (int c = 0; c < 3; ++c) {
        const int n1 = ext[c][0], n2 = ext[c][1], n3 = ext[c][2];
        for (int k = 1; k < n3 - 1; ++k)
            for (int j = 1; j < n2 - 1; ++j)
                for (int i = 1; i < n1 - 1; ++i) {
                    const uint64_t idx = (uint64_t)i + (uint64_t)n1 * ((uint64_t)j + (uint64_t)n2 * k);
                    const uint64_t r = splitmix64(key ^ splitmix64(((uint64_t)c << 56) ^ idx));
                    const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
                    fld[c][idx] += amp * (2.0 * u - 1.0);
                }
    }
    return BL_OK;
}

int bl_grid_tanh(int nx, int ny, int nz, double Lx, double Ly, double Lz, double alpha, double x0, double *x,
                 double *y, double *z) {
    if (nx < 2 || ny < 2 || nz < 3 || !x || !y || !z || !(alpha > 0)) return BL_EINVAL;
    for (int i = 0; i < nx; ++i) x[i] = x0 + Lx * (double)i / (double)(nx - 1);
    for (int j = 0; j < ny; ++j) {
        const double e = (double)j / (double)(ny - 1);
        y[j] = Ly * (1.0 - std::tanh(alpha * (1.0 - e)) / std::tanh(alpha));
    }
    y[0] = 0.0;
    for (int k = 0; k < nz; ++k) z[k] = Lz * (double)k / (double)(nz - 2);
    return BL_OK;
}

}  // extern "C"
