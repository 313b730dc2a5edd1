// C ABI of blgpu: handle life cycle, transfers and step orchestration.
//
// The per-step sequence mirrors compute_time_step_RK3 / RK2
// (time_integration.f90:145-465): for each stage the fused RHS+update kernel
// writes the next state buffer, then boundary conditions, projection and
// boundary conditions again.  The T1 copy U->Uo of the reference is replaced
// by rotating three state buffers: the step-start buffer is read as Uo for the
// whole step and never written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/blgpu.h"
#include "bl_internal.hpp"
#include "bl_mlp.hpp"

namespace blgpu {

std::string &last_error() {
    static thread_local std::string msg;
    return msg;
}

namespace {

constexpr long double kPi = 3.141592653589793238462643383279502884L;
constexpr double kPiD = 3.141592653589793238462643383279502884;  // = 4*atan(1) in double

int round_up(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace

// cos/sin(2 pi k m / M) in long double with the exact integer reduction of
// k m mod M, rounded once; symmetric in (k, m), lda covers padded row blocks
void z_dft_tables(int M, int &lda, std::vector<double> &zc, std::vector<double> &zs, std::vector<double> &zf) {
    lda = round_up(M + 1, 64);
    zc.assign((size_t)lda * lda, 0.0);
    zs.assign((size_t)lda * lda, 0.0);
    for (int m = 0; m < M; ++m)
        for (int k = 0; k < M; ++k) {
            const long long r = ((long long)k * m) % M;
            const long double ang = 2.0L * kPi * (long double)r / (long double)M;
            zc[(size_t)m * lda + k] = (double)cosl(ang);
            zs[(size_t)m * lda + k] = (double)sinl(ang);
        }
    zf.clear();
    const int Kc = M / 2 + 1;
    if (Kc > 96) return;
    const int RT = (Kc + 15) / 16, KS = (Kc + 3) / 4;
    zf.assign(2 * (size_t)RT * KS * 64, 0.0);
    for (int t = 0; t < RT; ++t)
        for (int ks = 0; ks < KS; ++ks)
            for (int l = 0; l < 64; ++l) {
                // the same entries as row 1 of zc, zs
                const int j = (int)(((long long)(16 * t + l % 16) * (4 * ks + l / 16)) % M);
                const size_t e = 2 * (((size_t)t * KS + ks) * 64 + l);
                zf[e] = zc[(size_t)lda + j];
                zf[e + 1] = zs[(size_t)lda + j];
            }
}

namespace {

// initialization.f90:181-207
void centers(int n, const double *x, std::vector<double> &xg) {
    std::vector<double> xm(n - 1);
    for (int i = 0; i < n - 1; ++i) xm[i] = 0.5 * (x[i] + x[i + 1]);
    xg.assign(n + 1, 0.0);
    for (int i = 0; i < n - 1; ++i) xg[i + 1] = xm[i];
    xg[0] = xm[0] - 2.0 * (xm[0] - x[0]);
    xg[n] = xm[n - 2] + 2.0 * (x[n - 1] - xm[n - 2]);
}

struct HostSetup {
    std::vector<double> x, y, z, xg, yg, zg, rdx, rdxg, rdy, rdyg, rdz, rdzg, wy0, wy1;
    double dxmin, dymin, dzmin;
};

double *dmalloc(size_t n) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double));
    if (e != hipSuccess) {
        last_error() = "hipMalloc of " + std::to_string(n * sizeof(double)) + " bytes: " + hipGetErrorString(e);
        throw std::bad_alloc();
    }
    return static_cast<double *>(p);
}

void upload_vec(double *dst, const std::vector<double> &v, hipStream_t s) {
    BL_HIP(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, s));
}

void field_extent(const Dims &d, int field, int &n1, int &n2, int &n3) {
    switch (field) {
    case BL_FIELD_U: n1 = d.nx; n2 = d.nyg; n3 = d.nzg; break;
    case BL_FIELD_V: n1 = d.nxg; n2 = d.ny; n3 = d.nzg; break;
    case BL_FIELD_W: n1 = d.nxg; n2 = d.nyg; n3 = d.nz; break;
    default: n1 = d.nxg; n2 = d.nyg; n3 = d.nzg; break;
    }
}

double *field_dev(Handle &h, int field) {
    Field3 &f = h.buf[h.icur];
    switch (field) {
    case BL_FIELD_U: return f.u;
    case BL_FIELD_V: return f.v;
    case BL_FIELD_W: return f.w;
    case BL_FIELD_P: return h.P;
    case BL_FIELD_NUT: return h.nu_t;
    default: last_error() = "unknown field id"; throw ArgError();
    }
}

// strided copies between a host array of extent (n1, n2, n3) and the padded
// device layout; element type E on the device
template <class E>
void copy_in(Handle &h, E *dev, const E *host, int n1, int n2, int n3) {
    const Dims &d = h.d;
    for (int k = 0; k < n3; ++k)
        BL_HIP(hipMemcpy2DAsync(dev + (size_t)k * d.sxy, (size_t)d.sx * sizeof(E), host + (size_t)k * n1 * n2,
                                (size_t)n1 * sizeof(E), (size_t)n1 * sizeof(E), n2, hipMemcpyHostToDevice,
                                h.stream));
    BL_HIP(hipStreamSynchronize(h.stream));
}

template <class E>
void copy_out(Handle &h, E *host, const E *dev, int n1, int n2, int n3) {
    const Dims &d = h.d;
    BL_HIP(hipStreamSynchronize(h.stream));
    for (int k = 0; k < n3; ++k)
        BL_HIP(hipMemcpy2DAsync(host + (size_t)k * n1 * n2, (size_t)n1 * sizeof(E), dev + (size_t)k * d.sxy,
                                (size_t)d.sx * sizeof(E), (size_t)n1 * sizeof(E), n2, hipMemcpyDeviceToHost,
                                h.stream));
    BL_HIP(hipStreamSynchronize(h.stream));
}

// host FP64 arrays <-> device fields (rounded to float in mixed precision)
size_t elem_size(const Handle &h) { return h.precision == BL_PREC_MIXED ? sizeof(float) : sizeof(double); }

void copy_to_device(Handle &h, double *dev, const double *host, int n1, int n2, int n3) {
    if (h.precision != BL_PREC_MIXED) {
        copy_in(h, dev, host, n1, n2, n3);
        return;
    }
    std::vector<float> tmp((size_t)n1 * n2 * n3);
    for (size_t q = 0; q < tmp.size(); ++q) tmp[q] = (float)host[q];
    copy_in(h, reinterpret_cast<float *>(dev), tmp.data(), n1, n2, n3);
}

void copy_to_host(Handle &h, double *host, const double *dev, int n1, int n2, int n3) {
    if (h.precision != BL_PREC_MIXED) {
        copy_out(h, host, dev, n1, n2, n3);
        return;
    }
    std::vector<float> tmp((size_t)n1 * n2 * n3);
    copy_out(h, tmp.data(), reinterpret_cast<const float *>(dev), n1, n2, n3);
    for (size_t q = 0; q < tmp.size(); ++q) host[q] = (double)tmp[q];
}

void check_config(const bl_config *c) {
    auto bad = [](const char *m) {
        last_error() = m;
        throw ArgError();
    };
    if (!c) bad("bl_create: null config");
    if (c->nx < 5 || c->ny < 4 || c->nz < 5) bad("bl_create: need nx>=5, ny>=4, nz>=5");
    if (!c->x || !c->y || !c->z) bad("bl_create: missing grid");
    if (!c->U_inlet || !c->V_inlet || !c->W_inlet || !c->U_top || !c->V_top || !c->W_top)
        bad("bl_create: missing boundary tables");
    if (c->scheme != 2 && c->scheme != 3) bad("bl_create: scheme must be 2 (RK2) or 3 (RK3)");
    if (c->les < 0 || c->les > 4)
        bad("bl_create: les must be 0 (DNS), 1 (Smagorinsky), 2 (dynamic Smagorinsky), 3 (dynamic Smagorinsky "
            "without the z average) or 4 (learned closure with dynamic Smagorinsky fallback)");
    if (c->precision != BL_PREC_FP64 && c->precision != BL_PREC_MIXED)
        bad("bl_create: precision must be BL_PREC_FP64 or BL_PREC_MIXED");

    if (c->n_modes < 0 || c->m_modes < 0) bad("bl_create: negative mode count");
    if (c->n_modes > 0 && (c->m_modes <= 0 || !c->zmode || !c->tmode || !c->qu_re || !c->qu_im ||
                           !c->qv_re || !c->qv_im || !c->qw_re || !c->qw_im))
        bad("bl_create: incomplete temporal-mode tables");
    for (int i = 1; i < c->nx; ++i)
        if (!(c->x[i] > c->x[i - 1])) bad("bl_create: x must increase");
    for (int i = 1; i < c->ny; ++i)
        if (!(c->y[i] > c->y[i - 1])) bad("bl_create: y must increase");
    for (int i = 1; i < c->nz; ++i)
        if (!(c->z[i] > c->z[i - 1])) bad("bl_create: z must increase");
}

// initialization.f90:394-425 (Dyy rows for j = 2..nyg-1, 0-based row r = j-2)
void build_dyy(const HostSetup &g, int nyg, std::vector<double> &dd, std::vector<double> &dl,
               std::vector<double> &du) {
    const int nm = nyg - 2;
    dd.assign(nm, 0.0);
    dl.assign(nm, 0.0);
    du.assign(nm, 0.0);
    auto Y = [&](int j) { return g.y[j - 1]; };
    auto YG = [&](int j) { return g.yg[j - 1]; };
    for (int j = 2; j <= nyg - 1; ++j) {
        const double a = 1.0 / (Y(j) - Y(j - 1)) / (YG(j + 1) - YG(j));
        const double b = 1.0 / (Y(j) - Y(j - 1)) * (-1.0 / (YG(j + 1) - YG(j)) - 1.0 / (YG(j) - YG(j - 1)));
        const double c = 1.0 / (Y(j) - Y(j - 1)) / (YG(j) - YG(j - 1));
        const int r = j - 2;
        if (j == 2) {
            dd[r] = b + c;
            du[r] = a;
        } else if (j == nyg - 1) {
            dd[r] = a + b;
            dl[r - 1] = c;
        } else {
            dd[r] = b;
            du[r] = a;
            dl[r - 1] = c;
        }
    }
}

void build(Handle &h, const bl_config *c) {
    {
        int dev = 0;
        BL_HIP(hipGetDevice(&dev));
        BL_HIP(hipDeviceGetAttribute(&h.n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    Dims &d = h.d;
    d.nx = c->nx;
    d.ny = c->ny;
    d.nz = c->nz;
    d.nxg = d.nx + 1;
    d.nyg = d.ny + 1;
    d.nzg = d.nz + 1;
    d.sx = round_up(d.nxg, 16);
    d.sy = d.nyg;
    d.sxy = (long long)d.sx * d.sy;
    d.N = d.nx - 2;
    d.M = d.nz - 2;
    d.NM = d.ny - 1;
    d.Kz = d.M / 2 + 1;
    h.nu = c->nu;
    h.dPdx = c->dPdx;
    h.dPdz = c->dPdz;
    h.cfl = c->cfl;
    h.scheme = c->scheme;
    h.les = c->les;
    h.nut_bc = c->nut_bc;
    h.precision = c->precision;
    h.t = 0.0;
    h.dt = c->cfl < 0 ? -c->cfl : 0.0;
    // RK tables (initialization.f90:503-530)
    std::memset(h.rk_t, 0, sizeof(h.rk_t));
    std::memset(h.rk_coef, 0, sizeof(h.rk_coef));
    std::memset(h.rk2_t, 0, sizeof(h.rk2_t));
    std::memset(h.rk2_coef, 0, sizeof(h.rk2_coef));
    h.rk2_t[1] = 1.0 / 2.0;
    h.rk2_t[2] = 1.0;
    h.rk2_coef[1][1] = 1.0 / 2.0;
    h.rk2_coef[2][1] = 0.0;
    h.rk2_coef[2][2] = 1.0;
    h.rk_t[1] = 8.0 / 15.0;
    h.rk_t[2] = 2.0 / 3.0;
    h.rk_t[3] = 1.0;
    h.rk_coef[1][1] = 8.0 / 15.0;
    h.rk_coef[2][1] = 1.0 / 4.0;
    h.rk_coef[2][2] = 5.0 / 12.0;
    h.rk_coef[3][1] = 1.0 / 4.0;
    h.rk_coef[3][2] = 0.0;
    h.rk_coef[3][3] = 3.0 / 4.0;
    // the FP64 RK3 stages 2 and 3 share one carried field Uo + dt a_s1 F1
    if (h.rk_coef[2][1] != h.rk_coef[3][1]) {
        last_error() = "bl_create: RK3 tableau with a21 != a31";
        throw ArgError();
    }

    HostSetup g;
    g.x.assign(c->x, c->x + d.nx);
    g.y.assign(c->y, c->y + d.ny);
    g.z.assign(c->z, c->z + d.nz);
    centers(d.nx, c->x, g.xg);
    centers(d.ny, c->y, g.yg);
    centers(d.nz, c->z, g.zg);
    auto recip_diff = [](const std::vector<double> &a, bool backward) {
        std::vector<double> r(a.size(), 0.0);
        for (size_t i = 0; i + 1 < a.size(); ++i) {
            if (backward)
                r[i + 1] = 1.0 / (a[i + 1] - a[i]);
            else
                r[i] = 1.0 / (a[i + 1] - a[i]);
        }
        return r;
    };
    g.rdx = recip_diff(g.x, true);
    g.rdy = recip_diff(g.y, true);
    g.rdz = recip_diff(g.z, true);
    g.rdxg = recip_diff(g.xg, false);
    g.rdyg = recip_diff(g.yg, false);
    g.rdzg = recip_diff(g.zg, false);
    g.wy0.resize(d.ny);
    g.wy1.resize(d.ny);
    for (int j = 0; j < d.ny; ++j) {
        g.wy0[j] = (g.yg[j + 1] - g.y[j]) / (g.yg[j + 1] - g.yg[j]);
        g.wy1[j] = 1.0 - g.wy0[j];
    }
    auto mindiff = [](const std::vector<double> &a) {
        double m = 1e300;
        for (size_t i = 0; i + 1 < a.size(); ++i) m = std::min(m, a[i + 1] - a[i]);
        return m;
    };
    g.dxmin = mindiff(g.xg);
    g.dymin = mindiff(g.yg);
    g.dzmin = mindiff(g.zg);
    h.length_y = 0.0;
    for (int j = 1; j <= d.nyg - 2; ++j) h.length_y = h.length_y + (g.y[j] - g.y[j - 1]);
    h.bc_uc = c->U_top[d.nx - 1];
    h.dxmin = g.dxmin;
    h.hx = g.x;
    h.hxg = g.xg;
    h.hyg = g.yg;
    h.hy = g.y;
    h.dymin = g.dymin;
    h.dzmin = g.dzmin;

    BL_HIP(hipStreamCreate(&h.stream));

    // ---- grid, metric and table storage in one allocation
    std::vector<std::pair<const double **, std::vector<double>>> tabs;
    auto add = [&](const double **slot, std::vector<double> v) { tabs.emplace_back(slot, std::move(v)); };
    add(&h.m.x, g.x);
    add(&h.m.y, g.y);
    add(&h.m.z, g.z);
    add(&h.m.xg, g.xg);
    add(&h.m.yg, g.yg);
    add(&h.m.zg, g.zg);
    add(&h.m.rdx, g.rdx);
    add(&h.m.rdxg, g.rdxg);
    add(&h.m.rdy, g.rdy);
    add(&h.m.rdyg, g.rdyg);
    add(&h.m.rdz, g.rdz);
    add(&h.m.rdzg, g.rdzg);
    add(&h.m.wy0, g.wy0);
    add(&h.m.wy1, g.wy1);
    add(&h.bc.U_inlet, std::vector<double>(c->U_inlet, c->U_inlet + d.nyg));
    add(&h.bc.V_inlet, std::vector<double>(c->V_inlet, c->V_inlet + d.ny));
    add(&h.bc.W_inlet, std::vector<double>(c->W_inlet, c->W_inlet + d.nyg));
    add(&h.bc.U_top, std::vector<double>(c->U_top, c->U_top + d.nx));
    add(&h.bc.V_top, std::vector<double>(c->V_top, c->V_top + d.nxg));
    add(&h.bc.W_top, std::vector<double>(c->W_top, c->W_top + d.nxg));
    h.bc.n_modes = c->n_modes;
    h.bc.m_modes = c->m_modes;
    if (c->n_modes > 0) {
        const size_t nm = (size_t)c->n_modes * c->m_modes;
        add(&h.bc.zmode, std::vector<double>(c->zmode, c->zmode + c->n_modes));
        add(&h.bc.tmode, std::vector<double>(c->tmode, c->tmode + c->m_modes));
        add(&h.bc.qu_re, std::vector<double>(c->qu_re, c->qu_re + nm * d.nyg));
        add(&h.bc.qu_im, std::vector<double>(c->qu_im, c->qu_im + nm * d.nyg));
        add(&h.bc.qv_re, std::vector<double>(c->qv_re, c->qv_re + nm * d.ny));
        add(&h.bc.qv_im, std::vector<double>(c->qv_im, c->qv_im + nm * d.ny));
        add(&h.bc.qw_re, std::vector<double>(c->qw_re, c->qw_re + nm * d.nyg));
        add(&h.bc.qw_im, std::vector<double>(c->qw_im, c->qw_im + nm * d.nyg));
    }
    size_t total = 0;
    for (auto &t : tabs) total += round_up((int)t.second.size(), 32);
    double *blk = dmalloc(total);
    h.mem_grid = blk;
    size_t off = 0;
    for (auto &t : tabs) {
        upload_vec(blk + off, t.second, h.stream);
        *t.first = blk + off;
        off += round_up((int)t.second.size(), 32);
    }
    // float copies of the metrics for the FP32 stencils
    {
        const std::vector<double> *src[14] = {&g.x,   &g.y,    &g.z,   &g.xg,   &g.yg,   &g.zg,  &g.rdx,
                                              &g.rdxg, &g.rdy, &g.rdyg, &g.rdz, &g.rdzg, &g.wy0, &g.wy1};
        const float **dst[14] = {&h.mf.x,   &h.mf.y,   &h.mf.z,    &h.mf.xg,  &h.mf.yg,  &h.mf.zg,  &h.mf.rdx,
                                 &h.mf.rdxg, &h.mf.rdy, &h.mf.rdyg, &h.mf.rdz, &h.mf.rdzg, &h.mf.wy0, &h.mf.wy1};
        size_t tot = 0;
        for (auto *v : src) tot += round_up((int)v->size(), 64);
        std::vector<float> host(tot, 0.f);
        size_t o = 0;
        std::vector<size_t> offs;
        for (auto *v : src) {
            offs.push_back(o);
            for (size_t q = 0; q < v->size(); ++q) host[o + q] = (float)(*v)[q];
            o += round_up((int)v->size(), 64);
        }
        float *fb = nullptr;
        BL_HIP(hipMalloc(&fb, tot * sizeof(float)));
        h.mem_gridf = fb;
        BL_HIP(hipMemcpyAsync(fb, host.data(), tot * sizeof(float), hipMemcpyHostToDevice, h.stream));
        BL_HIP(hipStreamSynchronize(h.stream));
        for (int q = 0; q < 14; ++q) *dst[q] = fb + offs[q];
    }

    // ---- state buffers (float in mixed precision)
    const size_t nf = (size_t)d.sxy * d.nzg;
    const size_t es = h.precision == BL_PREC_MIXED ? sizeof(float) : sizeof(double);
    auto falloc = [&]() {
        double *p = dmalloc((nf * es + sizeof(double) - 1) / sizeof(double));
        BL_HIP(hipMemsetAsync(p, 0, nf * es, h.stream));
        return p;
    };
    for (int b = 0; b < 3; ++b) {
        h.buf[b].u = falloc();
        h.buf[b].v = falloc();
        h.buf[b].w = falloc();
    }
    h.icur = h.iold = 0;
    h.F1.u = falloc();
    h.F1.v = falloc();
    h.F1.w = falloc();
    h.P = falloc();
    h.phi = falloc();
    h.nu_t = falloc();
    if (h.les >= 2) {
        h.dsm = dmalloc(nf * kDsmArrays);
        BL_HIP(hipMemsetAsync(h.dsm, 0, nf * kDsmArrays * sizeof(double), h.stream));
    }
    if (h.les == 1) {
        // sgs_Smagorinsky (subgrid.f90:609-662): Cs = 0.11, Delta_j =
        // (dx dz dy_j)^(1/3) with dx, dz the smallest spacings, van Driest
        // damping of the distance to the nearer wall (rows j < nyg/2 count
        // from the bottom) in units nu/utau, utau = dPdx^0.5; f = 1 when
        // nutBC = 0
        std::vector<double> cj(d.nyg, 0.0);
        const double utau = std::pow(c->dPdx, 0.5);
        for (int j = 2; j <= d.nyg - 1; ++j) {
            const double yw = j < d.nyg / 2 ? g.yg[j - 1] - g.y[0] : g.y[d.ny - 1] - g.yg[j - 1];
            double f = 1.0 - std::exp(-yw / 25.0 / (c->nu / utau));
            if (c->nut_bc == 0) f = 1.0;
            const double delta = std::pow(g.dxmin * g.dzmin * (g.y[j - 1] - g.y[j - 2]), 1.0 / 3.0);
            cj[j - 1] = std::pow(0.11 * delta * f, 2.0);
        }
        h.smag_c = dmalloc(cj.size());
        upload_vec(h.smag_c, cj, h.stream);
    }
    const size_t nspec = (size_t)d.N * d.NM * d.M;
    if (h.precision == BL_PREC_MIXED) {
        // FP32 spectra with kx stride NP; T1 doubles as the FP64 Thomas scratch
        build_mixed(h);
        const size_t nsf = (size_t)h.NP * d.NM * d.M;
        const size_t t1 = std::max(nspec, (nsf + 1) / 2), s1 = (nsf + 1) / 2;
        h.T1 = dmalloc(t1);
        h.S = dmalloc(s1);
        BL_HIP(hipMemsetAsync(h.T1, 0, t1 * sizeof(double), h.stream));
        BL_HIP(hipMemsetAsync(h.S, 0, s1 * sizeof(double), h.stream));
    } else {
        h.T1 = dmalloc(nspec);
        h.S = dmalloc(nspec);
    }

    const int N = d.N, M = d.M;
    // x-DCT tables: N = N1*N2 with N1 the largest divisor <= sqrt(N)
    {
        int n1 = 1;
        for (int a = 1; (long long)a * a <= N; ++a)
            if (N % a == 0) n1 = a;
        h.dct_n1 = n1;
        const int n2 = N / n1;
        auto cis = [](long long num, long long den, long double frac) {
            // e^{-2 pi i num/den * frac}, argument reduced exactly
            const long long r = num % den;
            const long double ang = -2.0L * kPi * (long double)r / (long double)den * frac;
            return std::pair<double, double>((double)cosl(ang), (double)sinl(ang));
        };
        std::vector<double> w1(2 * n1), w2(2 * n2), wN(2 * N), rot(2 * N);
        for (int q = 0; q < n1; ++q) std::tie(w1[2 * q], w1[2 * q + 1]) = cis(q, n1, 1.0L);
        for (int q = 0; q < n2; ++q) std::tie(w2[2 * q], w2[2 * q + 1]) = cis(q, n2, 1.0L);
        for (int q = 0; q < N; ++q) std::tie(wN[2 * q], wN[2 * q + 1]) = cis(q, N, 1.0L);
        for (int k = 0; k < N; ++k) std::tie(rot[2 * k], rot[2 * k + 1]) = cis(k, 4LL * N, 1.0L);
        h.dct_w1 = dmalloc(w1.size());
        h.dct_w2 = dmalloc(w2.size());
        h.dct_wN = dmalloc(wN.size());
        h.dct_rot = dmalloc(rot.size());
        upload_vec(h.dct_w1, w1, h.stream);
        upload_vec(h.dct_w2, w2, h.stream);
        upload_vec(h.dct_wN, wN, h.stream);
        upload_vec(h.dct_rot, rot, h.stream);
        // FP64 MFMA operand tables of the two short DFTs (k_dct_fwd_m /
        // k_dct_inv_m), one double per lane l per (row tile, k step):
        // entry A[16 rt + l % 16][4 ks + l / 16].  A tile row R = q + 4 r
        // (q = R % 4) of a paired table is (Re, Im) = (r % 2) of index
        // 8 rt + q + 4 (r / 2).
        const int k1n = n1 / 2 + 1;
        h.dct_mfma = n1 <= 32 && k1n <= 16 && n2 <= 32;
        if (!h.dct_mfma && h.precision == BL_PREC_FP64) {
            // the VALU kernels keep two lines and their spectra in 64 KB of LDS
            const size_t lds = sizeof(double2) * (2 * (size_t)N + 2 * (size_t)k1n * n2 + n1 + n2);
            if (lds > 65536) {
                last_error() = "bl_create: the x transform of nx - 2 = " + std::to_string(N) + " points (" +
                               std::to_string(n1) + " x " + std::to_string(n2) +
                               ") does not fit the LDS-resident DCT; choose nx - 2 with a divisor near its square root";
                throw ArgError();
            }
        }
        if (h.dct_mfma) {
            auto frag = [](int tiles, int ks_n, auto f) {
                std::vector<double> t((size_t)tiles * ks_n * 64, 0.0);
                for (int rt = 0; rt < tiles; ++rt)
                    for (int ks = 0; ks < ks_n; ++ks)
                        for (int l = 0; l < 64; ++l) t[((size_t)rt * ks_n + ks) * 64 + l] = f(rt, l % 16, 4 * ks + l / 16);
                return t;
            };
            auto pair_row = [](int rt, int R, int &idx, int &part) {
                idx = 8 * rt + R % 4 + 4 * ((R / 4) / 2);
                part = (R / 4) % 2;
            };
            const int ksa = (n1 + 3) / 4, ksc = (2 * n2 + 3) / 4, ksi = (2 * k1n + 3) / 4;
            // forward step A: B[k1][n2] = sum_n1 v W1^{n1 k1}, W1^q = (w1[2q], w1[2q+1])
            std::vector<double> fA = frag(2, ksa, [&](int rt, int R, int k) {
                int k1, part;
                pair_row(rt, R, k1, part);
                if (k1 >= k1n || k >= n1) return 0.0;
                const int w = (int)(((long long)k * k1) % n1);
                return w1[2 * w + part];
            });
            // forward step C: V = sum_n2 B' W2^{n2 k2}; column k = (n2, re/im)
            std::vector<double> fC = frag(4, ksc, [&](int rt, int R, int k) {
                int k2, part;
                pair_row(rt, R, k2, part);
                const int nn = k / 2, pin = k % 2;
                if (k2 >= n2 || nn >= n2) return 0.0;
                const int w = (int)(((long long)nn * k2) % n2);
                const double wr = w2[2 * w], wi = w2[2 * w + 1];
                return part == 0 ? (pin == 0 ? wr : -wi) : (pin == 0 ? wi : wr);
            });
            // inverse step C': sum_k2 V conj(W2^{n2 k2}); column k = (k2, re/im)
            std::vector<double> iC = frag(4, ksc, [&](int rt, int R, int k) {
                int nn, part;
                pair_row(rt, R, nn, part);
                const int k2 = k / 2, pin = k % 2;
                if (nn >= n2 || k2 >= n2) return 0.0;
                const int w = (int)(((long long)nn * k2) % n2);
                const double wr = w2[2 * w], wi = w2[2 * w + 1];
                return part == 0 ? (pin == 0 ? wr : wi) : (pin == 0 ? -wi : wr);
            });
            // inverse step A': v = sum_k1 c_k1 Re(C~ conj(W1^{n1 k1})), rows n1
            std::vector<double> iA = frag(2, ksi, [&](int rt, int R, int k) {
                const int nn1 = 16 * rt + R, k1 = k / 2, pin = k % 2;
                if (nn1 >= n1 || k1 >= k1n) return 0.0;
                const double c = (k1 == 0 || (n1 % 2 == 0 && k1 == n1 / 2)) ? 1.0 : 2.0;
                const int w = (int)(((long long)nn1 * k1) % n1);
                return c * w1[2 * w + pin];
            });
            h.dct_mf = dmalloc(fA.size() + fC.size() + iC.size() + iA.size());
            std::vector<double> all;
            for (auto *t : {&fA, &fC, &iC, &iA}) all.insert(all.end(), t->begin(), t->end());
            upload_vec(h.dct_mf, all, h.stream);
            h.dct_mf_off[0] = 0;
            h.dct_mf_off[1] = fA.size();
            h.dct_mf_off[2] = fA.size() + fC.size();
            h.dct_mf_off[3] = fA.size() + fC.size() + iC.size();
        }
    }
    // z cos/sin tables, symmetric in (k, m): lda covers the padded row blocks
    {
        std::vector<double> zc, zs, zf;
        z_dft_tables(M, h.lda_zt, zc, zs, zf);
        h.zc = dmalloc(zc.size());
        h.zs = dmalloc(zs.size());
        upload_vec(h.zc, zc, h.stream);
        upload_vec(h.zs, zs, h.stream);
        if (!zf.empty()) {
            h.zf = dmalloc(zf.size());
            upload_vec(h.zf, zf, h.stream);
        }
    }
    // ---- Thomas factors on the half spectrum (initialization.f90:328-344, 427-469)
    const int NM = d.NM, Kz = d.Kz;
    std::vector<double> kxx(N), kzz(Kz);
    for (int i = 0; i < N; ++i) kxx[i] = 2.0 * (std::cos(kPiD * (double)i / (double)N) - 1.0) / std::pow(g.dxmin, 2.0);
    for (int k = 0; k < Kz; ++k)
        kzz[k] = 2.0 * (std::cos(2.0 * kPiD * (double)k / (double)M) - 1.0) / std::pow(g.dzmin, 2.0);
    std::vector<double> dd, dlo, dup;
    build_dyy(g, d.nyg, dd, dlo, dup);
    // partitioned kernel for NM <= 32 * kThMaxP rows: row coefficients only
    h.th_p = (NM + kThSeg - 1) / kThSeg;
    if (h.th_p > kThMaxP) h.th_p = 0;
    const size_t plane = (size_t)N * NM;
    const bool tables = h.th_p == 0;
    std::vector<double> fdl(tables ? plane * Kz : 0, 0.0), fdp(tables ? plane * Kz : 0, 0.0);
    std::vector<double> lu00(2 * (size_t)NM, 0.0);
    std::vector<double> piv(NM), low(NM);
    for (int kk = 0; kk < (tables ? Kz : 1); ++kk)
        for (int kx = 0; kx < (tables ? N : 1); ++kx) {
            for (int j = 0; j < NM; ++j) piv[j] = dd[j] + kxx[kx] + kzz[kk];
            for (int j = 0; j < NM - 1; ++j) low[j] = dlo[j];
            if (kx == 0 && kk == 0) piv[0] = 3.0 / 2.0 * piv[0];
            for (int j = 1; j < NM; ++j) {
                const double m = low[j - 1] / piv[j - 1];
                low[j - 1] = m;
                piv[j] = piv[j] - m * dup[j - 1];
            }
            if (kx == 0 && kk == 0)
                for (int j = 0; j < NM; ++j) {
                    lu00[NM + j] = 1.0 / piv[j];
                    if (j < NM - 1) lu00[j] = low[j];
                }
            if (tables)
                for (int j = 0; j < NM; ++j) {
                    fdp[(size_t)kk * plane + (size_t)j * N + kx] = 1.0 / piv[j];
                    if (j < NM - 1) fdl[(size_t)kk * plane + (size_t)j * N + kx] = low[j];
                }
        }
    if (tables) {
        h.lu_dl = dmalloc(fdl.size());
        h.lu_dp = dmalloc(fdp.size());
        upload_vec(h.lu_dl, fdl, h.stream);
        upload_vec(h.lu_dp, fdp, h.stream);
    } else {
        const int L = kThSeg * h.th_p;
        std::vector<double> co(3 * (size_t)L + N + Kz, 0.0);
        for (int j = 0; j < L; ++j) {
            co[j] = (j >= 1 && j < NM) ? dlo[j - 1] : 0.0;
            co[L + j] = j < NM ? dd[j] : 1.0;
            co[2 * L + j] = j < NM - 1 ? dup[j] : 0.0;
        }
        for (int i = 0; i < N; ++i) co[3 * L + i] = kxx[i];
        for (int k = 0; k < Kz; ++k) co[3 * L + N + k] = kzz[k];
        h.th_co = dmalloc(co.size());
        upload_vec(h.th_co, co, h.stream);
    }
    h.lu00 = dmalloc(lu00.size());
    upload_vec(h.lu00, lu00, h.stream);
    h.lu_du = dmalloc(NM);
    upload_vec(h.lu_du, dup, h.stream);

    h.scal = dmalloc(64);
    {
        // reduction scratch: 3 partials per interior block, flux partials, Cf, monitor planes
        const size_t nb = (size_t)((d.nxg - 2 + 63) / 64) * ((d.nyg - 2 + 3) / 4) * (d.nzg - 2);
        const size_t nq = (size_t)(((d.nyg + 255) / 256) + ((d.nxg + 255) / 256)) * d.nzg;
        h.red = dmalloc(std::max({3 * nb, nq, (size_t)d.nxg, (size_t)8 * d.nzg, (size_t)1024}));
    }
    BL_HIP(hipStreamSynchronize(h.stream));
}

void prof_free(Handle &h);

void free_all(Handle &h) {
    snapshot_free(h);
    auto f = [](void *p) {
        if (p) (void)hipFree(p);
    };
    for (int b = 0; b < 3; ++b) {
        f(h.buf[b].u);
        f(h.buf[b].v);
        f(h.buf[b].w);
    }
    for (void *p : {(void *)h.F1.u, (void *)h.F1.v, (void *)h.F1.w, (void *)h.P, (void *)h.phi,
                    (void *)h.nu_t, (void *)h.T1, (void *)h.S, (void *)h.dct_w1, (void *)h.dct_w2,
                    (void *)h.dct_wN, (void *)h.dct_rot, (void *)h.dct_mf, h.mx2, (void *)h.zc, (void *)h.zs, (void *)h.zf, (void *)h.bc_extra, (void *)h.lund_mem, (void *)h.wm_mem, (void *)h.lu_dl, (void *)h.lu_dp, (void *)h.lu_du,
                    (void *)h.lu00, (void *)h.th_co,
                    (void *)h.scal, (void *)h.red, (void *)h.dsm, (void *)h.smag_c, h.mem_grid, h.mem_gridf, (void *)h.mx_tab, (void *)h.zmf,
                    (void *)h.p00, (void *)h.inc1, (void *)h.stats, (void *)h.cl_w, (void *)h.cl_b, (void *)h.cl_cnt})
        f(p);
    if (h.stream) (void)hipStreamSynchronize(h.stream);
    prof_free(h);
    if (h.stream) (void)hipStreamDestroy(h.stream);
}

// one projection sub-step on the current buffer (projection.f90:21-87);
// inc: the mixed path's first-stage source (see IncSrc)
void projection_on(Handle &h, Field3 &f, int rk_step, double dt, const IncSrc *inc = nullptr) {
    if (h.precision == BL_PREC_MIXED) {
        Fld3<float> ff = as<float>(f);
        float *phi = reinterpret_cast<float *>(h.phi);
        launch_poisson_mixed(h, phi, ff, inc);
        launch_project(h, ff, (const float *)phi, rk_step == 1 ? reinterpret_cast<float *>(h.P) : nullptr, dt);
        return;
    }
    launch_poisson(h, h.phi, &f);
    launch_project(h, f, (const double *)h.phi, rk_step == 1 ? h.P : nullptr, dt);
}

template <class T>
void bc_on(Handle &h, Field3 &f, const Field3 &old, int rk_step, double dt, double t) {
    Fld3<T> ff = as<T>(f);
    launch_bc(h, ff, as<T>(old), rk_step, dt, t);
}

template <class T>
void do_step_t(Handle &h) {
    const double dt = h.dt;
    const double to = h.t;
    const int X = h.icur;
    const int Y = (X + 1) % 3, Z = (X + 2) % 3;
    const Fld3<T> old = as<T>(h.buf[X]);
    const int nst = h.scheme == 3 ? 3 : 2;
    const int order3[3] = {Y, Z, Y};
    int cur = X;
    h.bc.istep = h.bc_step;  // the inflow_flag 4 draw of this step
    // Lund rescaling: the step's inflow plane from the step-start field
    if (h.bc.inflow_flag == 3 || h.bc.inflow_flag == 5) launch_lund<T>(h, old, dt);
    h.lund_pending = false;
    for (int s = 1; s <= nst; ++s) {
        const int nxt = order3[s - 1];
        // eddy viscosity of the stage's input fields (time_integration.f90:341-346)
        if (h.les != 0) launch_dsm(h, h.buf[cur]);
        if (h.les == 4) launch_closure(h, h.buf[cur]);
        // then the wall model (time_integration.f90:344): slip lengths of
        // the stage, and the pseudo-Robin V's wall mean
        if (h.wm_mem && (h.wm_model != 0 || (h.les != 0 && h.wm_istress == 1)))
            launch_wall_model<T>(h, as<T>(h.buf[cur]), old, false);
        Fld3<T> next = as<T>(h.buf[nxt]);
        launch_rk_stage(h, as<T>(h.buf[cur]), old, next, s, dt);
        const double ts = to + (nst == 3 ? h.rk_t[s] : h.rk2_t[s]) * dt;
        // FP32 fields that a previous step projected: the first stage's
        // pressure comes from the divergence of the stage increment, not of
        // U* (see IncSrc); fields set from outside may carry a divergence of
        // their own, which the first stage projects as the reference does
        const bool inc = sizeof(T) == 4 && s == 1 && h.inc_ok;
        launch_bc(h, next, old, s, dt, ts, inc ? h.inc1 : nullptr, to);
        if (inc) {
            const Fld3<float> f1 = as<float>(h.F1), vn = as<float>(h.buf[X]), vs = as<float>(h.buf[nxt]);
            const IncSrc src{f1.u, f1.v, f1.w, vs.v, vn.v, h.inc1,
                             dt * (nst == 3 ? h.rk_coef[1][1] : h.rk2_coef[1][1])};
            projection_on(h, h.buf[nxt], s, dt, &src);
        } else {
            projection_on(h, h.buf[nxt], s, dt);
        }
        launch_bc(h, next, old, s, dt, ts);
        h.t = ts;
        cur = nxt;
    }
    h.icur = h.iold = cur;
    h.inc_ok = true;
    ++h.bc_step;
}

void do_step(Handle &h) {
    if (h.precision == BL_PREC_MIXED)
        do_step_t<float>(h);
    else
        do_step_t<double>(h);
}

}  // namespace

struct ProfState {
    struct Rec {
        int id;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double ms[KID_COUNT] = {};
    long long n[KID_COUNT] = {};
    hipEvent_t open_ev[KID_COUNT] = {};
    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        BL_HIP(hipEventCreate(&e));
        return e;
    }
    void drain() {
        for (auto &r : pending) {
            BL_HIP(hipEventSynchronize(r.b));
            float t = 0.f;
            BL_HIP(hipEventElapsedTime(&t, r.a, r.b));
            ms[r.id] += t;
            n[r.id] += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    ~ProfState() {
        for (auto &r : pending) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

namespace {
void prof_free(Handle &h) {
    delete h.prof;
    h.prof = nullptr;
}
}  // namespace

void prof_begin(Handle &h, int id) {
    if (!h.prof || !((h.prof_mask >> id) & 1ull)) return;
    hipEvent_t e = h.prof->take();
    BL_HIP(hipEventRecord(e, h.stream));
    h.prof->open_ev[id] = e;
}

void prof_end(Handle &h, int id) {
    if (!h.prof || !h.prof->open_ev[id]) return;
    hipEvent_t e = h.prof->take();
    BL_HIP(hipEventRecord(e, h.stream));
    h.prof->pending.push_back({id, h.prof->open_ev[id], e});
    h.prof->open_ev[id] = nullptr;
    if (h.prof->pending.size() > 4096) h.prof->drain();
}

namespace {

template <class F>
int guarded(F &&fn) {
    try {
        fn();
        return BL_OK;
    } catch (const ArgError &) {
        return BL_EINVAL;
    } catch (const DeviceError &) {
        return BL_EDEVICE;
    } catch (const StateError &) {
        return BL_ESTATE;
    } catch (const std::bad_alloc &) {
        if (last_error().empty()) last_error() = "out of memory";
        return BL_ENOMEM;
    } catch (...) {
        last_error() = "unexpected exception";
        return BL_EDEVICE;
    }
}

Handle &H(bl_handle *h) {
    if (!h) {
        last_error() = "null handle";
        throw ArgError();
    }
    return *h;
}

}  // namespace
}  // namespace blgpu

using namespace blgpu;

extern "C" {

const char *bl_last_error(void) { return last_error().c_str(); }

int bl_create(const bl_config *cfg, bl_handle **out) {
    last_error().clear();
    if (!out) {
        last_error() = "bl_create: null output pointer";
        return BL_EINVAL;
    }
    *out = nullptr;
    // value-initialised: members without an initialiser start at zero
    bl_handle *h = new (std::nothrow) bl_handle();
    if (!h) return BL_ENOMEM;
    int rc = guarded([&] {
        check_config(cfg);
        int ndev = 0;
        BL_HIP(hipGetDeviceCount(&ndev));
        if (ndev < 1) {
            last_error() = "bl_create: no HIP device";
            throw DeviceError();
        }
        build(*h, cfg);
    });
    if (rc != BL_OK) {
        free_all(*h);
        delete h;
        return rc;
    }
    *out = h;
    return BL_OK;
}

void bl_destroy(bl_handle *h) {
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    free_all(*h);
    delete h;
}

int bl_upload(bl_handle *hh, int field, const double *host) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!host) {
            last_error() = "bl_upload: null host pointer";
            throw ArgError();
        }
        int n1, n2, n3;
        field_extent(h.d, field, n1, n2, n3);
        copy_to_device(h, field_dev(h, field), host, n1, n2, n3);
        if (field <= BL_FIELD_W) h.inc_ok = false;
    });
}

int bl_download(bl_handle *hh, int field, double *host) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!host) {
            last_error() = "bl_download: null host pointer";
            throw ArgError();
        }
        int n1, n2, n3;
        field_extent(h.d, field, n1, n2, n3);
        copy_to_host(h, host, field_dev(h, field), n1, n2, n3);
    });
}

int bl_set_boundary(bl_handle *hh, const bl_boundary *b) {
    return guarded([&] {
        Handle &h = H(hh);
        const Dims &d = h.d;
        auto bad = [](const char *m) {
            last_error() = m;
            throw ArgError();
        };
        if (!b) bad("bl_set_boundary: null flags");
        if (b->inflow_flag < 1 || b->inflow_flag > 5) bad("bl_set_boundary: inflow_flag must be 1 to 5");
        const bool lund = b->inflow_flag == 3 || b->inflow_flag == 5;
        if (lund && (b->i_rescale < 2 || b->i_rescale > d.nx - 1))
            bad("bl_set_boundary: Lund rescaling needs i_rescale in [2, nx - 1]");
        if (lund && !(b->delta_inlet > 0.0 && std::isfinite(b->delta_inlet) && b->t_resc > 0.0 &&
                      std::isfinite(b->t_resc)))
            bad("bl_set_boundary: Lund rescaling needs delta_inlet > 0 and t_resc > 0");
        if (b->top_flag < 0 || b->top_flag > 3) bad("bl_set_boundary: top_flag must be 0, 1, 2 or 3");
        if (b->inflow_flag == 4 && !(std::isfinite(b->amplitude) && b->amplitude >= 0.0))
            bad("bl_set_boundary: amplitude must be finite and >= 0");
        if ((b->top_flag == 1 || b->top_flag == 2) &&
            !(std::isfinite(b->vbs_max) && std::isfinite(b->x_bs) && b->sigma_bs > 0.0 && std::isfinite(b->sigma_bs) &&
              std::isfinite(b->phi_bs)))
            bad("bl_set_boundary: blowing/suction needs finite vbs_max, x_bs, phi_bs and sigma_bs > 0");
        if ((b->top_flag == 1 || b->top_flag == 2) && (b->i_rescale < 0 || b->i_rescale > d.nxg))
            bad("bl_set_boundary: i_rescale must be in [0, nx + 1]");
        if (b->top_flag == 3 && !(std::isfinite(b->beta_hartree) && b->beta_hartree < 2.0))
            bad("bl_set_boundary: beta_hartree must be finite and below 2");
        // host tables: V(i, ny) of blowing/suction (boundary_conditions.f90:
        // 657-667) at xg(i); Falkner-Skan U(i, nyg) at xg(i) (:994-1000)
        std::vector<double> ex((size_t)d.nxg + d.nx, 0.0);
        if (b->top_flag == 1 || b->top_flag == 2) {
            for (int i = 0; i < d.nxg; ++i) {
                const double r = (b->x_bs - h.hxg[i]) / b->sigma_bs;
                const double e = b->top_flag == 1 ? std::exp(0.5 - r * r) : std::exp(b->phi_bs - r * r);
                ex[i] = std::sqrt(2.0) * b->vbs_max * (b->x_bs - h.hxg[i]) / b->sigma_bs * e +
                        (b->top_flag == 1 ? b->phi_bs : 0.0);
            }
        }
        double uc_fs = 0.0;
        if (b->top_flag == 3) {
            const double xf = -1.0, ex_m = b->beta_hartree / (2.0 - b->beta_hartree);
            const double c = 1.0 / std::pow(h.hx[0] - xf, ex_m);
            for (int i = 0; i < d.nx; ++i) ex[(size_t)d.nxg + i] = c * std::pow(h.hxg[i] - xf, ex_m);
            uc_fs = ex[(size_t)d.nxg + d.nx - 1];
        }
        BL_HIP(hipStreamSynchronize(h.stream));
        if (!h.bc_extra) h.bc_extra = dmalloc(ex.size());
        BL_HIP(hipMemcpy(h.bc_extra, ex.data(), ex.size() * sizeof(double), hipMemcpyHostToDevice));
        h.bc.inflow_flag = b->inflow_flag;
        h.bc.top_flag = b->top_flag;
        h.bc.amp = b->amplitude;
        h.bc.seed = b->seed;
        h.bc.vtop_bs = h.bc_extra;
        h.bc.utop_fs = h.bc_extra + d.nxg;
        h.bc.dy_dx = (h.hyg[d.nyg - 1] - h.hyg[d.nyg - 2]) / h.dxmin;
        h.bc.i_rescale = b->i_rescale;
        h.bc_step = 0;
        h.fs_uc_pending = b->top_flag == 3;
        h.fs_uc = uc_fs;
        h.inc_ok = false;
        if (lund) {
            if (!h.lund_mem) {
                h.lund_mem = dmalloc(lund_doubles(d));
                BL_HIP(hipMemsetAsync(h.lund_mem, 0, lund_doubles(d) * sizeof(double), h.stream));
                h.lund = lund_layout(d, h.lund_mem);
            }
            h.lund_ir = b->i_rescale - 1;
            h.lund_t_resc = b->t_resc;
            h.lund_delta_inlet = b->delta_inlet;
            h.bc.lund_u = h.lund.pu;
            h.bc.lund_v = h.lund.pv;
            h.bc.lund_w = h.lund.pw;
            if (h.precision == BL_PREC_MIXED)
                lund_init_means<float>(h, as<float>(h.buf[h.icur]));
            else
                lund_init_means<double>(h, as<double>(h.buf[h.icur]));
            BL_HIP(hipStreamSynchronize(h.stream));
        }
        h.lund_pending = lund;
    });
}

int bl_set_time(bl_handle *hh, double t) {
    return guarded([&] { H(hh).t = t; });
}

int bl_get_time(bl_handle *hh, double *t) {
    return guarded([&] {
        if (!t) throw ArgError();
        *t = H(hh).t;
    });
}

int bl_compute_dt(bl_handle *hh, double *dt) {
    return guarded([&] {
        Handle &h = H(hh);
        // time_integration.f90:536-574
        const double CFLa = std::fabs(h.cfl);
        if (h.cfl < 0) {
            h.dt = -h.cfl;
        } else {
            const double dt_conv = h.precision == BL_PREC_MIXED ? dt_convective(h, as<float>(h.buf[h.icur]))
                                                                : dt_convective(h, as<double>(h.buf[h.icur]));
            const double dvu = CFLa * std::pow(h.dxmin, 2.0) / h.nu;
            const double dvv = CFLa * std::pow(h.dymin, 2.0) / h.nu;
            const double dvw = CFLa * std::pow(h.dzmin, 2.0) / h.nu;
            const double dt_vis = std::fmin(dvu, std::fmin(dvv, dvw));
            h.dt = std::fmin(dt_conv, dt_vis);
        }
        if (dt) *dt = h.dt;
    });
}

int bl_step(bl_handle *hh) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!(h.dt > 0.0)) {
            last_error() = "bl_step: dt is not positive (call bl_compute_dt)";
            throw ArgError();
        }
        do_step(h);
    });
}

int bl_run(bl_handle *hh, int nsteps) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!(h.dt > 0.0)) {
            last_error() = "bl_run: dt is not positive (call bl_compute_dt)";
            throw ArgError();
        }
        for (int n = 0; n < nsteps; ++n) do_step(h);
    });
}

int bl_synchronize(bl_handle *hh) {
    return guarded([&] { BL_HIP(hipStreamSynchronize(H(hh).stream)); });
}

int bl_eddy_viscosity(bl_handle *hh) {
    return guarded([&] {
        Handle &h = H(hh);
        if (h.les != 0) {
            launch_dsm(h, h.buf[h.icur]);
            if (h.les == 4) launch_closure(h, h.buf[h.icur]);
        } else
            BL_HIP(hipMemsetAsync(h.nu_t, 0, (size_t)h.d.sxy * h.d.nzg * elem_size(h), h.stream));
    });
}

int bl_max_divergence(bl_handle *hh, double *div) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!div) throw ArgError();
        *div = h.precision == BL_PREC_MIXED ? max_divergence(h, as<float>(h.buf[h.icur]))
                                            : max_divergence(h, as<double>(h.buf[h.icur]));
    });
}

int bl_wall_stats(bl_handle *hh, double *cf) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!cf) throw ArgError();
        if (h.precision == BL_PREC_MIXED)
            wall_stats(h, as<float>(h.buf[h.icur]), cf);
        else
            wall_stats(h, as<double>(h.buf[h.icur]), cf);
    });
}

int bl_statistics(bl_handle *hh, double *means, double *nu_t_mean, double *cf) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!means) {
            last_error() = "bl_statistics: null means array";
            throw ArgError();
        }
        const Dims &d = h.d;
        const long long n2 = (long long)d.nx * d.ny;
        if (!h.stats) BL_HIP(hipMalloc(&h.stats, sizeof(double) * stats_size(d)));
        launch_statistics(h, h.buf[h.icur], h.stats);
        BL_HIP(hipMemcpyAsync(means, h.stats, sizeof(double) * 9 * n2, hipMemcpyDeviceToHost, h.stream));
        if (nu_t_mean)
            BL_HIP(hipMemcpyAsync(nu_t_mean, h.stats + 9 * n2, sizeof(double) * (size_t)d.nxg * d.nyg,
                                  hipMemcpyDeviceToHost, h.stream));
        BL_HIP(hipStreamSynchronize(h.stream));
        if (cf) {
            if (h.precision == BL_PREC_MIXED)
                wall_stats(h, as<float>(h.buf[h.icur]), cf);
            else
                wall_stats(h, as<double>(h.buf[h.icur]), cf);
        }
    });
}

int bl_snapshot_begin(bl_handle *hh) {
    return guarded([&] { snapshot_begin(H(hh)); });
}

int bl_snapshot_end(bl_handle *hh, const double **U, const double **V, const double **W, const double **P,
                    const double **nu_t) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!U || !V || !W || !P || !nu_t) {
            last_error() = "bl_snapshot_end: null output pointer";
            throw ArgError();
        }
        const double *ptr[5];
        snapshot_end(h, ptr);
        *U = ptr[0];
        *V = ptr[1];
        *W = ptr[2];
        *P = ptr[3];
        *nu_t = ptr[4];
    });
}

int bl_monitor(bl_handle *hh, double *out) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!out) throw ArgError();
        if (h.precision == BL_PREC_MIXED)
            monitor(h, as<float>(h.buf[h.icur]), out);
        else
            monitor(h, as<double>(h.buf[h.icur]), out);
    });
}

int bl_rhs(bl_handle *hh, double *ru, double *rv, double *rw) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!ru || !rv || !rw) throw ArgError();
        const size_t nf = (size_t)h.d.sxy * h.d.nzg;
        for (double *p : {h.F1.u, h.F1.v, h.F1.w}) BL_HIP(hipMemsetAsync(p, 0, nf * elem_size(h), h.stream));
        if (h.precision == BL_PREC_MIXED) {
            Fld3<float> r = as<float>(h.F1);
            launch_rhs_only(h, as<float>(h.buf[h.icur]), r);
        } else {
            launch_rhs_only(h, as<double>(h.buf[h.icur]), h.F1);
        }
        int n1, n2, n3;
        field_extent(h.d, BL_FIELD_U, n1, n2, n3);
        copy_to_host(h, ru, h.F1.u, n1, n2, n3);
        field_extent(h.d, BL_FIELD_V, n1, n2, n3);
        copy_to_host(h, rv, h.F1.v, n1, n2, n3);
        field_extent(h.d, BL_FIELD_W, n1, n2, n3);
        copy_to_host(h, rw, h.F1.w, n1, n2, n3);
    });
}

int bl_set_old(bl_handle *hh, const double *Uo, const double *Vo, const double *Wo) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!Uo || !Vo || !Wo) throw ArgError();
        const int io = (h.icur + 1) % 3;
        int n1, n2, n3;
        field_extent(h.d, BL_FIELD_U, n1, n2, n3);
        copy_to_device(h, h.buf[io].u, Uo, n1, n2, n3);
        field_extent(h.d, BL_FIELD_V, n1, n2, n3);
        copy_to_device(h, h.buf[io].v, Vo, n1, n2, n3);
        field_extent(h.d, BL_FIELD_W, n1, n2, n3);
        copy_to_device(h, h.buf[io].w, Wo, n1, n2, n3);
        h.iold = io;
    });
}

int bl_apply_bc(bl_handle *hh, int rk_step, double dt) {
    return guarded([&] {
        Handle &h = H(hh);
        if (rk_step < 1 || rk_step > 3) throw ArgError();
        h.inc_ok = false;
        h.bc.istep = h.bc_step;
        if (h.lund_pending) {
            // the first boundary call after set-up computes the rescaled plane
            // from the old fields with the handle's dt (step_beginning)
            if (h.precision == BL_PREC_MIXED)
                launch_lund<float>(h, as<float>(h.buf[h.iold]), h.dt);
            else
                launch_lund<double>(h, as<double>(h.buf[h.iold]), h.dt);
            h.lund_pending = false;
        }
        // the pseudo-Robin V's wall mean from the old fields and the current
        // slip lengths, as apply_pseudo_Robin_bc_y_bottom forms it at every call
        if (h.precision == BL_PREC_MIXED) {
            launch_wall_model<float>(h, as<float>(h.buf[h.icur]), as<float>(h.buf[h.iold]), true);
            bc_on<float>(h, h.buf[h.icur], h.buf[h.iold], rk_step, dt, h.t);
        } else {
            launch_wall_model<double>(h, as<double>(h.buf[h.icur]), as<double>(h.buf[h.iold]), true);
            bc_on<double>(h, h.buf[h.icur], h.buf[h.iold], rk_step, dt, h.t);
        }
        BL_HIP(hipStreamSynchronize(h.stream));
    });
}

int bl_set_wall_model(bl_handle *hh, const bl_wall_params *p) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!p) {
            last_error() = "bl_set_wall_model: null parameters";
            throw ArgError();
        }
        const double am[3] = {p->alpha_mean_x, p->alpha_mean_y, p->alpha_mean_z};
        wall_model_set(h, p->iwall_model, p->istress_model, am, p->alpha_std, p->freq_mult, p->iwall_model_nut,
                       p->frac_vis);
    });
}

int bl_compute_wall_model(bl_handle *hh) {
    return guarded([&] {
        Handle &h = H(hh);
        if (h.precision == BL_PREC_MIXED)
            launch_wall_model<float>(h, as<float>(h.buf[h.icur]), as<float>(h.buf[h.iold]), false);
        else
            launch_wall_model<double>(h, as<double>(h.buf[h.icur]), as<double>(h.buf[h.iold]), false);
        BL_HIP(hipStreamSynchronize(h.stream));
    });
}

int bl_get_wall_model(bl_handle *hh, double *alpha_x, double *alpha_y, double *alpha_z, double *utau_model) {
    return guarded([&] {
        Handle &h = H(hh);
        wall_model_get(h, alpha_x, alpha_y, alpha_z, utau_model);
    });
}

int bl_projection(bl_handle *hh, int rk_step, double dt) {
    return guarded([&] {
        Handle &h = H(hh);
        h.inc_ok = false;
        projection_on(h, h.buf[h.icur], rk_step, dt);
        BL_HIP(hipStreamSynchronize(h.stream));
    });
}

int bl_profile_enable(bl_handle *hh, int on) {
    return guarded([&] {
        Handle &h = H(hh);
        if (on && !h.prof) h.prof = new ProfState();
        if (!on && h.prof) {
            BL_HIP(hipStreamSynchronize(h.stream));
            delete h.prof;
            h.prof = nullptr;
        }
    });
}

int bl_profile_select(bl_handle *hh, unsigned long long mask) {
    return guarded([&] { H(hh).prof_mask = mask; });
}

int bl_profile_read(bl_handle *hh, double *ms, long long *launches, int n) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!ms || !launches || n < 1) throw ArgError();
        for (int i = 0; i < n; ++i) {
            ms[i] = 0.0;
            launches[i] = 0;
        }
        if (!h.prof) return;
        h.prof->drain();
        for (int i = 0; i < n && i < KID_COUNT; ++i) {
            ms[i] = h.prof->ms[i];
            launches[i] = h.prof->n[i];
        }
    });
}

const char *bl_profile_name(int id) { return kernel_name(id); }

int bl_profile_count(void) { return KID_COUNT; }

int bl_poisson(bl_handle *hh, double *rhs) {
    return guarded([&] {
        Handle &h = H(hh);
        if (!rhs) throw ArgError();
        if (h.precision == BL_PREC_MIXED) {
            last_error() = "bl_poisson: the stand-alone Poisson solve is the FP64 path (the mixed path forms its "
                           "rhs from the fields, see bl_projection)";
            throw ArgError();
        }
        int n1, n2, n3;
        field_extent(h.d, BL_FIELD_P, n1, n2, n3);
        copy_to_device(h, h.phi, rhs, n1, n2, n3);
        launch_poisson(h, h.phi);
        copy_to_host(h, rhs, h.phi, n1, n2, n3);
    });
}

int bl_closure_samples(bl_handle *fine, bl_handle *coarse, bl_mlp *m, int fx, int fy, int fz, double delta99,
                       double smin) {
    return guarded([&] {
        if (!m || !(delta99 > 0.0)) {
            last_error() = "bl_closure_samples: need a closure handle and delta99 > 0";
            throw ArgError();
        }
        closure_samples(H(fine), H(coarse), *m, fx, fy, fz, delta99, smin);
    });
}

int bl_set_closure(bl_handle *hh, bl_mlp *m, double delta99, double nsigma) {
    return guarded([&] {
        if (!m || !(delta99 > 0.0) || !(nsigma >= 0.0)) {
            last_error() = "bl_set_closure: need a closure handle, delta99 > 0 and nsigma >= 0";
            throw ArgError();
        }
        closure_install(H(hh), *m, delta99, nsigma);
    });
}

int bl_closure_fallback(bl_handle *hh, long long *cells, long long *fallback) {
    return guarded([&] {
        if (!cells || !fallback) throw ArgError();
        closure_fallback(H(hh), cells, fallback);
    });
}

int bl_closure_infer(bl_mlp *m, const float *x, long long n, float *y) {
    return guarded([&] {
        if (!m || !x || !y || n < 1) {
            last_error() = "bl_closure_infer: need a closure handle, n >= 1 and buffers";
            throw ArgError();
        }
        closure_infer_rows(*m, x, n, y);
    });
}

}  // extern "C"
