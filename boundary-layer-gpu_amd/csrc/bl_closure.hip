// Learned subgrid closure: training samples from a filtered DNS, BF16-MFMA
// inference, and the LES eddy viscosity with dynamic-Smagorinsky fallback
// (LES = 4).  None of this has a counterpart in the reference solver; the
// fallback and the a-priori judge are its dynamic Smagorinsky model
// (subgrid.f90:261-414, bl_dsm.hip).
//
// Features of an LES cell (0-based interior cell i, j, k of a staggered
// field): the velocity-gradient tensor g_ab = d u_a / d x_b at the cell
// centre, with the diagonal from the cell's own faces and each off-diagonal
// derivative averaged from the four cell edges around the centre (the
// edge-to-centre averaging of compute_Sij, subgrid.f90:695-724), normalised
// by |g| = sqrt(g:g); and the wall distance yg / delta99.  S_ab = (g_ab +
// g_ba)/2, |S| = sqrt(2 S:S), Delta = (dx dy dz)^(1/3).  The network predicts
// c = nu_t / (Delta^2 |S|); the LES uses nu_t = max(c, 0) Delta^2 |S| unless a
// feature leaves the training range by more than nsigma standard deviations,
// in which case the cell keeps the dynamic Smagorinsky value.
//
// Samples: a DNS on a grid that refines the LES grid by (fx, fy, fz) cells
// (x: nx_f - 1 = fx (nx_c - 1), y: ny_f - 1 = fy (ny_c - 1), z: nz_f - 2 =
// fz (nz_c - 2), coinciding faces) is top-hat filtered onto the LES grid: each
// LES face value is the area average of the DNS face values inside it (dy
// weighted), each LES cell gets tau_ab = <u_a u_b> - <u_a><u_b> from the DNS
// cell-centred velocities of its block (volume weighted).  The target is the
// least-squares eddy viscosity nu* = -tau^d:S / (2 S:S) in units of
// Delta^2 |S|.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/blgpu.h"
#include "bl_mlp.hpp"

namespace blgpu {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int kH = kMlpHidden;

// velocity-gradient tensor at interior cell (i, j, k): g[3a + b] = d u_a / d x_b
template <class T>
__device__ void grad_tensor(const Dims &d, const Metrics &m, const T *__restrict__ U, const T *__restrict__ V,
                            const T *__restrict__ W, int i, int j, int k, double g[9]) {
    const long long sx = d.sx, sxy = d.sxy;
    const long long p = fidx(d, i, j, k);
    g[0] = ((double)U[p] - (double)U[p - 1]) * m.rdx[i];
    g[4] = ((double)V[p] - (double)V[p - sx]) * m.rdy[j];
    g[8] = ((double)W[p] - (double)W[p - sxy]) * m.rdz[k];
    double uy = 0, vx = 0, uz = 0, wx = 0, vz = 0, wy = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int da = e & 1, db = e >> 1;
        {  // xy edge (i+da, j+db, k)
            const int a = i + da, b = j + db;
            const long long q = fidx(d, a, b, k);
            uy += ((double)U[q - 1] - (double)U[q - 1 - sx]) * m.rdyg[b - 1];
            vx += ((double)V[q - sx] - (double)V[q - 1 - sx]) * m.rdxg[a - 1];
        }
        {  // xz edge (i+da, j, k+db)
            const int a = i + da, c = k + db;
            const long long q = fidx(d, a, j, c);
            uz += ((double)U[q - 1] - (double)U[q - 1 - sxy]) * m.rdzg[c - 1];
            wx += ((double)W[q - sxy] - (double)W[q - 1 - sxy]) * m.rdxg[a - 1];
        }
        {  // yz edge (i, j+da, k+db)
            const int b = j + da, c = k + db;
            const long long q = fidx(d, i, b, c);
            vz += ((double)V[q - sx] - (double)V[q - sx - sxy]) * m.rdzg[c - 1];
            wy += ((double)W[q - sxy] - (double)W[q - sx - sxy]) * m.rdyg[b - 1];
        }
    }
    g[1] = 0.25 * uy;
    g[3] = 0.25 * vx;
    g[2] = 0.25 * uz;
    g[6] = 0.25 * wx;
    g[5] = 0.25 * vz;
    g[7] = 0.25 * wy;
}

struct CellFeat {
    float f[kMlpIn];
    double s[6];   // S11 S22 S33 S12 S13 S23
    double smag;   // |S|
    double scale;  // Delta^2 |S|
};

template <class T>
__device__ void cell_features(const Dims &d, const Metrics &m, const T *U, const T *V, const T *W, int i, int j,
                              int k, double inv_d99, CellFeat &cf) {
    double g[9];
    grad_tensor(d, m, U, V, W, i, j, k, g);
    double gn = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) gn += g[q] * g[q];
    gn = sqrt(gn);
    const double ig = gn > 0.0 ? 1.0 / gn : 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) cf.f[q] = (float)(g[q] * ig);
    cf.f[9] = (float)(m.yg[j] * inv_d99);
    cf.s[0] = g[0];
    cf.s[1] = g[4];
    cf.s[2] = g[8];
    cf.s[3] = 0.5 * (g[1] + g[3]);
    cf.s[4] = 0.5 * (g[2] + g[6]);
    cf.s[5] = 0.5 * (g[5] + g[7]);
    cf.smag = sqrt(2.0 * (cf.s[0] * cf.s[0] + cf.s[1] * cf.s[1] + cf.s[2] * cf.s[2]) +
                   4.0 * (cf.s[3] * cf.s[3] + cf.s[4] * cf.s[4] + cf.s[5] * cf.s[5]));
    const double vol = (m.x[i] - m.x[i - 1]) * (m.y[j] - m.y[j - 1]) * (m.z[k] - m.z[k - 1]);
    const double dl = cbrt(vol);
    cf.scale = dl * dl * cf.smag;
}

// ------------------------------------------------------------------------
// Filtering a DNS onto the LES grid
// ------------------------------------------------------------------------
struct Coarsen {
    Dims f, c;       // fine and coarse dimensions
    int fx, fy, fz;  // refinement per coarse cell
    const double *fy_dy;  // fine y(j) - y(j-1), j >= 1
};

__device__ __forceinline__ int zmap_c(const Dims &d, int k) {
    return k == 0 ? d.nzg - 3 : (k == d.nzg - 2 ? 1 : (k == d.nzg - 1 ? 2 : k));
}
__device__ __forceinline__ int zmap_f(const Dims &d, int k) { return k == 0 ? d.nz - 2 : (k == d.nz - 1 ? 1 : k); }

// dy-weighted average over the fine rows of coarse cell row J and the fine
// planes of coarse cell plane K of fine array F at fine x index i
__device__ double block_avg_yz(const Coarsen &cz, const double *F, int i, int J, int K) {
    const int j0 = cz.fy * (J - 1) + 1, k0 = cz.fz * (K - 1) + 1;
    double s = 0.0, wsum = 0.0;
    for (int jj = 0; jj < cz.fy; ++jj) {
        const double w = cz.fy_dy[j0 + jj];
        double a = 0.0;
        for (int kk = 0; kk < cz.fz; ++kk) a += F[fidx(cz.f, i, j0 + jj, k0 + kk)];
        s += w * a;
        wsum += w;
    }
    return s / (wsum * cz.fz);
}

// U at coarse faces (I, J, K), all J, K of the coarse extent
__global__ __launch_bounds__(256) void k_coarsen_u(Coarsen cz, const double *__restrict__ Uf, double *__restrict__ Uc) {
    const Dims &c = cz.c;
    const int I = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int K = blockIdx.z;
    if (I >= c.nx || J >= c.nyg) return;
    const int Kc = zmap_c(c, K);
    const int i = cz.fx * I;
    double v;
    if (J == 0) {
        v = -block_avg_yz(cz, Uf, i, 1, Kc);
    } else if (J == c.nyg - 1) {
        const int k0 = cz.fz * (Kc - 1) + 1;
        double a = 0.0;
        for (int kk = 0; kk < cz.fz; ++kk) a += Uf[fidx(cz.f, i, cz.f.nyg - 1, k0 + kk)];
        v = a / cz.fz;
    } else {
        v = block_avg_yz(cz, Uf, i, J, Kc);
    }
    Uc[fidx(c, I, J, K)] = v;
}

// V at coarse (I, y face J, K): average over the fine x cells of I and fine z cells of K
__global__ __launch_bounds__(256) void k_coarsen_v(Coarsen cz, const double *__restrict__ Vf, double *__restrict__ Vc) {
    const Dims &c = cz.c;
    const int I = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int K = blockIdx.z;
    if (I >= c.nxg || J >= c.ny) return;
    const int Kc = zmap_c(c, K);
    const int j = cz.fy * J, k0 = cz.fz * (Kc - 1) + 1;
    double a = 0.0;
    if (I == 0 || I == c.nxg - 1) {
        const int i = I == 0 ? 0 : cz.f.nxg - 1;
        for (int kk = 0; kk < cz.fz; ++kk) a += Vf[fidx(cz.f, i, j, k0 + kk)];
        a /= cz.fz;
    } else {
        const int i0 = cz.fx * (I - 1) + 1;
        for (int ii = 0; ii < cz.fx; ++ii)
            for (int kk = 0; kk < cz.fz; ++kk) a += Vf[fidx(cz.f, i0 + ii, j, k0 + kk)];
        a /= (double)(cz.fx * cz.fz);
    }
    Vc[fidx(c, I, J, K)] = a;
}

// W at coarse (I, J, z face K)
__global__ __launch_bounds__(256) void k_coarsen_w(Coarsen cz, const double *__restrict__ Wf, double *__restrict__ Wc) {
    const Dims &c = cz.c;
    const int I = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int K = blockIdx.z;
    if (I >= c.nxg || J >= c.nyg || K >= c.nz) return;
    const int k = cz.fz * zmap_f(c, K);
    const int Ja = J == 0 ? 1 : J;
    double v;
    if (J == c.nyg - 1) {
        double a = 0.0;
        const int i0 = (I == 0 || I == c.nxg - 1) ? (I == 0 ? 0 : cz.f.nxg - 1) : cz.fx * (I - 1) + 1;
        const int ni = (I == 0 || I == c.nxg - 1) ? 1 : cz.fx;
        for (int ii = 0; ii < ni; ++ii) a += Wf[fidx(cz.f, i0 + ii, cz.f.nyg - 1, k)];
        v = a / ni;
    } else {
        const int j0 = cz.fy * (Ja - 1) + 1;
        const int i0 = (I == 0 || I == c.nxg - 1) ? (I == 0 ? 0 : cz.f.nxg - 1) : cz.fx * (I - 1) + 1;
        const int ni = (I == 0 || I == c.nxg - 1) ? 1 : cz.fx;
        double s = 0.0, wsum = 0.0;
        for (int jj = 0; jj < cz.fy; ++jj) {
            const double w = cz.fy_dy[j0 + jj];
            double a = 0.0;
            for (int ii = 0; ii < ni; ++ii) a += Wf[fidx(cz.f, i0 + ii, j0 + jj, k)];
            s += w * a;
            wsum += w;
        }
        v = s / (wsum * ni);
        if (J == 0) v = -v;
    }
    Wc[fidx(c, I, J, K)] = v;
}

// one dataset row per interior coarse cell (I, J, K), I in [1, nxg-2],
// J in [1, nyg-2], K in [1, nzg-3]
__global__ __launch_bounds__(256) void k_samples(Coarsen cz, Metrics mc, const double *__restrict__ Uf,
                                                 const double *__restrict__ Vf, const double *__restrict__ Wf,
                                                 const double *__restrict__ Uc, const double *__restrict__ Vc,
                                                 const double *__restrict__ Wc, double inv_d99, double smin,
                                                 float *__restrict__ rows) {
    const Dims &c = cz.c, &f = cz.f;
    const int I = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int K = 1 + blockIdx.z;
    if (I > c.nxg - 2 || J > c.nyg - 2) return;
    // exact SGS stress of the block of DNS cells (volume weighted)
    const int i0 = cz.fx * (I - 1) + 1, j0 = cz.fy * (J - 1) + 1, k0 = cz.fz * (K - 1) + 1;
    double su = 0, sv = 0, sw = 0, suu = 0, svv = 0, sww = 0, suv = 0, suw = 0, svw = 0, wsum = 0;
    for (int jj = 0; jj < cz.fy; ++jj) {
        const int j = j0 + jj;
        const double w = cz.fy_dy[j];
        for (int kk = 0; kk < cz.fz; ++kk)
            for (int ii = 0; ii < cz.fx; ++ii) {
                const int i = i0 + ii, k = k0 + kk;
                const long long p = fidx(f, i, j, k);
                const double u = 0.5 * (Uf[p - 1] + Uf[p]);
                const double v = 0.5 * (Vf[p - f.sx] + Vf[p]);
                const double ww = 0.5 * (Wf[p - f.sxy] + Wf[p]);
                su += w * u;
                sv += w * v;
                sw += w * ww;
                suu += w * u * u;
                svv += w * v * v;
                sww += w * ww * ww;
                suv += w * u * v;
                suw += w * u * ww;
                svw += w * v * ww;
                wsum += w;
            }
    }
    const double r = 1.0 / wsum;
    su *= r;
    sv *= r;
    sw *= r;
    const double tau[6] = {suu * r - su * su, svv * r - sv * sv, sww * r - sw * sw,
                           suv * r - su * sv, suw * r - su * sw, svw * r - sv * sw};
    CellFeat cf;
    cell_features(c, mc, Uc, Vc, Wc, I, J, K, inv_d99, cf);
    const double ss = cf.s[0] * cf.s[0] + cf.s[1] * cf.s[1] + cf.s[2] * cf.s[2] +
                      2.0 * (cf.s[3] * cf.s[3] + cf.s[4] * cf.s[4] + cf.s[5] * cf.s[5]);
    const double ts = tau[0] * cf.s[0] + tau[1] * cf.s[1] + tau[2] * cf.s[2] +
                      2.0 * (tau[3] * cf.s[3] + tau[4] * cf.s[4] + tau[5] * cf.s[5]);
    const bool ok = cf.smag > smin && I >= 2 && I <= c.nxg - 3 && ss > 0.0;
    const double nus = ok ? -ts / (2.0 * ss) : 0.0;
    const long long slot = (long long)(I - 1) + (long long)(c.nxg - 2) * ((J - 1) + (long long)(c.nyg - 2) * (K - 1));
    float *row = rows + slot * kMlpRow;
#pragma unroll
    for (int q = 0; q < kMlpIn; ++q) row[q] = cf.f[q];
    row[kRowTarget] = ok ? (float)(nus / cf.scale) : 0.f;
    // stress-space weight: the deviatoric stress error of a model c at this
    // cell is 4 (Delta^2 |S|)^2 (S:S) (c - target)^2 + const, so weighting the
    // squared error in c by (Delta^2 |S|)^2 S:S makes the regression minimise
    // the SGS stress error itself (training rescales weights to mean 1)
    row[kRowWeight] = ok ? (float)(cf.scale * cf.scale * ss) : 0.f;
    row[kRowScale] = (float)cf.scale;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        row[kRowS + q] = (float)cf.s[q];
        row[kRowTau + q] = (float)tau[q];
    }
#pragma unroll
    for (int q = kRowTau + 6; q < kMlpRow; ++q) row[q] = 0.f;
}

// ------------------------------------------------------------------------
// BF16 MFMA inference of the network for 64 cells per wave
// ------------------------------------------------------------------------
// LDS weights: BF16 W1 [64][32] (inputs padded to 32), W2, W3 [64][64],
// W4 as a [16][64] block with row 0 = W4; FP32 biases.
struct InferLds {
    unsigned short w1[kH * 32], w2[kH * kH], w3[kH * kH], w4[16 * kH];
    float b1[kH], b2[kH], b3[kH], b4;
};

__device__ void load_infer_weights(InferLds &L, const unsigned short *__restrict__ wbf, const float *__restrict__ bias) {
    constexpr int n1 = kH * 32, n2 = kH * kH, n4 = 16 * kH;
    for (int q = threadIdx.x; q < n1 + 2 * n2 + n4; q += blockDim.x) {
        const unsigned short v = wbf[q];
        if (q < n1)
            L.w1[q] = v;
        else if (q < n1 + n2)
            L.w2[q - n1] = v;
        else if (q < n1 + 2 * n2)
            L.w3[q - n1 - n2] = v;
        else
            L.w4[q - n1 - 2 * n2] = v;
    }
    for (int q = threadIdx.x; q < 3 * kH + 1; q += blockDim.x) {
        const float v = bias[q];
        if (q < kH)
            L.b1[q] = v;
        else if (q < 2 * kH)
            L.b2[q - kH] = v;
        else if (q < 3 * kH)
            L.b3[q - 2 * kH] = v;
        else
            L.b4 = v;
    }
}

__device__ __forceinline__ unsigned short to_bf16(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
// SiLU from the hardware exp and reciprocal (~1 ulp each; the result is
// rounded to BF16 next); finite for all finite z: exp(-z) = inf gives z * 0
__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.f + __expf(-z)); }

// one hidden layer of a wave's 64 cells: act[cell][0..K) (BF16) times W^T
// (W stored [64 outputs][K]), bias, SiLU, written back to act[cell][0..64)
template <int K>
__device__ __forceinline__ void hidden_layer(const unsigned short *Wt, const float *bias, unsigned short *act) {
    const int lane = threadIdx.x & 63, l16 = lane & 15, lq = lane >> 4;
    constexpr int KS = K / 32;
    bf16x8 a[4][KS];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            a[cb][ks] = __builtin_bit_cast(
                bf16x8, *reinterpret_cast<const u16x8 *>(act + (cb * 16 + l16) * kH + ks * 32 + 8 * lq));
    __builtin_amdgcn_wave_barrier();
    // one 16-output block at a time keeps the register count low enough for
    // two waves per SIMD (the LDS use allows two blocks per CU)
#pragma unroll 1
    for (int ob = 0; ob < 4; ++ob) {
        f32x4 acc[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const bf16x8 b =
                __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(Wt + (ob * 16 + l16) * K + ks * 32 + 8 * lq));
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[cb][ks], b, acc[cb], 0, 0, 0);
        }
        const int feat = ob * 16 + l16;
        const float bv = bias[feat];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) act[(cb * 16 + 4 * lq + r) * kH + feat] = to_bf16(silu(acc[cb][r] + bv));
    }
    __builtin_amdgcn_wave_barrier();
}

// act: this wave's [64 cells][64] BF16 activation buffer; on entry row
// `lane` holds the lane's 10 features in columns 0..9 (others zero).
// Returns the network output for the lane's cell.
__device__ float infer_wave(const InferLds &L, unsigned short *act, float *outb) {
    const int lane = threadIdx.x & 63, l16 = lane & 15, lq = lane >> 4;
    hidden_layer<32>(L.w1, L.b1, act);
    hidden_layer<kH>(L.w2, L.b2, act);
    hidden_layer<kH>(L.w3, L.b3, act);
    // output layer: W4 as the first of 16 output columns
    bf16x8 a[4][2];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
            a[cb][ks] =
                __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(act + (cb * 16 + l16) * kH + ks * 32 + 8 * lq));
    f32x4 o[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) o[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        const bf16x8 b = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(L.w4 + l16 * kH + ks * 32 + 8 * lq));
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) o[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[cb][ks], b, o[cb], 0, 0, 0);
    }
    if (l16 == 0)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) outb[cb * 16 + 4 * lq + r] = o[cb][r] + L.b4;
    __builtin_amdgcn_wave_barrier();
    return outb[lane];
}

// inference on feature rows (parity checks): x[n][10] -> y[n]
__global__ __launch_bounds__(256) void k_infer_rows(const unsigned short *__restrict__ wbf,
                                                    const float *__restrict__ bias, const float *__restrict__ x,
                                                    long long n, float *__restrict__ y) {
    __shared__ InferLds L;
    __shared__ unsigned short act[4][64 * kH];
    __shared__ float outb[4][64];
    load_infer_weights(L, wbf, bias);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long cell = ((long long)blockIdx.x * 4 + wave) * 64 + lane;
    unsigned short *A = act[wave];
    for (int q = 0; q < kH; ++q) A[lane * kH + q] = 0;
    if (cell < n)
        for (int q = 0; q < kMlpIn; ++q) A[lane * kH + q] = to_bf16(x[cell * kMlpIn + q]);
    __builtin_amdgcn_wave_barrier();
    const float c = infer_wave(L, A, outb[wave]);
    if (cell < n) y[cell] = c;
}

// LES = 4: per interior cell, the network's eddy viscosity or the DSM value
// already in nu_t when a feature is out of range; fallback counts per block.
template <class T>
__global__ __launch_bounds__(256) void k_closure(Dims d, Metrics m, const T *__restrict__ U, const T *__restrict__ V,
                                                 const T *__restrict__ W,
                                                 const unsigned short *__restrict__ wbf, const float *__restrict__ bias,
                                                 const float *__restrict__ lo, const float *__restrict__ hi,
                                                 double inv_d99, int dirichlet, T *__restrict__ nu_t,
                                                 double *__restrict__ cnt) {
    __shared__ InferLds L;
    __shared__ unsigned short act[4][64 * kH];
    __shared__ float outb[4][64];
    __shared__ int nfb[4];
    load_infer_weights(L, wbf, bias);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long ni = d.nxg - 2, nj = d.nyg - 2, nk = d.nzg - 2;
    const long long cell = ((long long)blockIdx.x * 4 + wave) * 64 + lane;
    const bool live = cell < ni * nj * nk;
    int i = 1, j = 1, k = 1;
    if (live) {
        i = 1 + (int)(cell % ni);
        j = 1 + (int)((cell / ni) % nj);
        k = 1 + (int)(cell / (ni * nj));
    }
    CellFeat cf;
    cell_features(d, m, U, V, W, i, j, k, inv_d99, cf);
    bool fb = false;
#pragma unroll
    for (int q = 0; q < kMlpIn; ++q) fb = fb || (cf.f[q] < lo[q]) || (cf.f[q] > hi[q]);
    unsigned short *A = act[wave];
    for (int q = 0; q < kH; ++q) A[lane * kH + q] = 0;
    for (int q = 0; q < kMlpIn; ++q) A[lane * kH + q] = to_bf16(cf.f[q]);
    __builtin_amdgcn_wave_barrier();
    const float c = infer_wave(L, A, outb[wave]);
    if (live && !fb) {
        // the cell and, next to a y boundary, its ghost (apply_Neumann_bc_y /
        // apply_Dirichlet_bc_y on nu_t, subgrid.f90:406-411); DSM cells keep
        // the ghosts the DSM pass wrote
        const T v = (T)(fmax((double)c, 0.0) * cf.scale);
        const T sg = dirichlet == 1 ? (T)-1 : (T)1;
        nu_t[fidx(d, i, j, k)] = v;
        if (j == 1) nu_t[fidx(d, i, 0, k)] = sg * v;
        if (j == d.nyg - 2) nu_t[fidx(d, i, d.nyg - 1, k)] = sg * v;
    }
    // fallback count of the block
    const unsigned long long bal = __ballot(live && fb);
    if (lane == 0) nfb[wave] = __popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = (double)(nfb[0] + nfb[1] + nfb[2] + nfb[3]);
}

// per-feature min / max / sum / sum of squares over weighted rows, per block
__global__ __launch_bounds__(256) void k_row_stats(const float *__restrict__ rows, long long n, double *__restrict__ part) {
    __shared__ double sh[256];
    double mn[kMlpIn], mx[kMlpIn], s[kMlpIn], s2[kMlpIn], cnt = 0.0;
#pragma unroll
    for (int q = 0; q < kMlpIn; ++q) {
        mn[q] = 1e300;
        mx[q] = -1e300;
        s[q] = s2[q] = 0.0;
    }
    for (long long r = (long long)blockIdx.x * 256 + threadIdx.x; r < n; r += (long long)gridDim.x * 256) {
        const float *row = rows + r * kMlpRow;
        if (row[kRowWeight] <= 0.f) continue;
        cnt += 1.0;
#pragma unroll
        for (int q = 0; q < kMlpIn; ++q) {
            const double v = row[q];
            mn[q] = fmin(mn[q], v);
            mx[q] = fmax(mx[q], v);
            s[q] += v;
            s2[q] += v * v;
        }
    }
    auto red = [&](double v, int mode) {
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) {
                const double a = sh[threadIdx.x], b = sh[threadIdx.x + st];
                sh[threadIdx.x] = mode == 0 ? a + b : (mode == 1 ? fmin(a, b) : fmax(a, b));
            }
            __syncthreads();
        }
        const double out = sh[0];
        __syncthreads();
        return out;
    };
    double *P = part + (long long)blockIdx.x * (4 * kMlpIn + 1);
    for (int q = 0; q < kMlpIn; ++q) {
        const double a = red(mn[q], 1), b = red(mx[q], 2), c = red(s[q], 0), e = red(s2[q], 0);
        if (threadIdx.x == 0) {
            P[4 * q + 0] = a;
            P[4 * q + 1] = b;
            P[4 * q + 2] = c;
            P[4 * q + 3] = e;
        }
    }
    const double cc = red(cnt, 0);
    if (threadIdx.x == 0) P[4 * kMlpIn] = cc;
}

void check() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        last_error() = std::string("kernel launch: ") + hipGetErrorString(e);
        throw DeviceError();
    }
}

}  // namespace

// host: BF16 inference weights from the FP32 master parameters, in the LDS
// order of load_infer_weights, and the FP32 biases
void closure_weights(const float *theta, std::vector<unsigned short> &wbf, std::vector<float> &bias) {
    auto bf = [](float x) {
        uint32_t u;
        std::memcpy(&u, &x, 4);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (unsigned short)(u >> 16);
    };
    const float *W1 = theta, *b1 = W1 + kH * kMlpIn, *W2 = b1 + kH, *b2 = W2 + kH * kH, *W3 = b2 + kH,
                *b3 = W3 + kH * kH, *W4 = b3 + kH, *b4 = W4 + kH;
    wbf.assign(kH * 32 + 2 * kH * kH + 16 * kH, 0);
    for (int o = 0; o < kH; ++o)
        for (int i = 0; i < kMlpIn; ++i) wbf[o * 32 + i] = bf(W1[o * kMlpIn + i]);
    for (int q = 0; q < kH * kH; ++q) {
        wbf[kH * 32 + q] = bf(W2[q]);
        wbf[kH * 32 + kH * kH + q] = bf(W3[q]);
    }
    for (int i = 0; i < kH; ++i) wbf[kH * 32 + 2 * kH * kH + i] = bf(W4[i]);
    bias.assign(3 * kH + 1, 0.f);
    for (int q = 0; q < kH; ++q) {
        bias[q] = b1[q];
        bias[kH + q] = b2[q];
        bias[2 * kH + q] = b3[q];
    }
    bias[3 * kH] = b4[0];
}

void closure_infer_rows(Mlp &m, const float *x_host, long long n, float *y_host) {
    std::vector<float> th(kMlpParams);
    BL_HIP(hipMemcpyAsync(th.data(), m.theta, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m.stream));
    BL_HIP(hipStreamSynchronize(m.stream));
    std::vector<unsigned short> wbf;
    std::vector<float> bias;
    closure_weights(th.data(), wbf, bias);
    unsigned short *dw = nullptr;
    float *db = nullptr, *dx = nullptr, *dy = nullptr;
    BL_HIP(hipMalloc(&dw, wbf.size() * 2));
    BL_HIP(hipMalloc(&db, bias.size() * 4));
    BL_HIP(hipMalloc(&dx, sizeof(float) * kMlpIn * n));
    BL_HIP(hipMalloc(&dy, sizeof(float) * n));
    BL_HIP(hipMemcpyAsync(dw, wbf.data(), wbf.size() * 2, hipMemcpyHostToDevice, m.stream));
    BL_HIP(hipMemcpyAsync(db, bias.data(), bias.size() * 4, hipMemcpyHostToDevice, m.stream));
    BL_HIP(hipMemcpyAsync(dx, x_host, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice, m.stream));
    hipLaunchKernelGGL(k_infer_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, dw, db, dx, n, dy);
    check();
    BL_HIP(hipMemcpyAsync(y_host, dy, sizeof(float) * n, hipMemcpyDeviceToHost, m.stream));
    BL_HIP(hipStreamSynchronize(m.stream));
    (void)hipFree(dw);
    (void)hipFree(db);
    (void)hipFree(dx);
    (void)hipFree(dy);
}

void launch_closure(Handle &h, const Field3 &f) {
    if (!h.cl_ready) {
        last_error() = "LES = 4 needs a trained closure (bl_set_closure) before the first step";
        throw ArgError();
    }
    const Dims &d = h.d;
    ProfScope ps(h, KID_CLOSURE);
    if (h.precision == BL_PREC_MIXED) {
        const Fld3<float> ff = as<float>(f);
        float *nut = reinterpret_cast<float *>(h.nu_t);
        hipLaunchKernelGGL(k_closure<float>, dim3(h.cl_nblk), dim3(256), 0, h.stream, d, h.m, ff.u, ff.v, ff.w, h.cl_w,
                           h.cl_b, h.cl_lo, h.cl_hi, h.cl_inv_d99, h.nut_bc, nut, h.cl_cnt);
    } else {
        hipLaunchKernelGGL(k_closure<double>, dim3(h.cl_nblk), dim3(256), 0, h.stream, d, h.m, f.u, f.v, f.w, h.cl_w,
                           h.cl_b, h.cl_lo, h.cl_hi, h.cl_inv_d99, h.nut_bc, h.nu_t, h.cl_cnt);
    }
    check();
}

int closure_blocks(const Dims &d) {
    const long long n = (long long)(d.nxg - 2) * (d.nyg - 2) * (d.nzg - 2);
    return (int)((n + 255) / 256);
}

// training rows from a DNS filtered onto the LES grid; the filtered field
// is left in the coarse solver's current state
void closure_samples(Handle &fine, Handle &coarse, Mlp &m, int fx, int fy, int fz, double delta99, double smin) {
    const Dims &f = fine.d, &c = coarse.d;
    auto bad = [](const std::string &msg) {
        last_error() = "bl_closure_samples: " + msg;
        throw ArgError();
    };
    if (fine.precision != BL_PREC_FP64 || coarse.precision != BL_PREC_FP64) bad("both solvers must be FP64");
    if (fx < 1 || fy < 1 || fz < 1) bad("refinement factors must be >= 1");
    if (f.nx - 1 != fx * (c.nx - 1) || f.ny - 1 != fy * (c.ny - 1) || f.nz - 2 != fz * (c.nz - 2))
        bad("fine grid is not the coarse grid refined by (fx, fy, fz)");
    std::vector<double> xf(f.nx), yf(f.ny), zf(f.nz), xc(c.nx), yc(c.ny), zc(c.nz);
    BL_HIP(hipMemcpy(xf.data(), fine.m.x, sizeof(double) * f.nx, hipMemcpyDeviceToHost));
    BL_HIP(hipMemcpy(yf.data(), fine.m.y, sizeof(double) * f.ny, hipMemcpyDeviceToHost));
    BL_HIP(hipMemcpy(zf.data(), fine.m.z, sizeof(double) * f.nz, hipMemcpyDeviceToHost));
    BL_HIP(hipMemcpy(xc.data(), coarse.m.x, sizeof(double) * c.nx, hipMemcpyDeviceToHost));
    BL_HIP(hipMemcpy(yc.data(), coarse.m.y, sizeof(double) * c.ny, hipMemcpyDeviceToHost));
    BL_HIP(hipMemcpy(zc.data(), coarse.m.z, sizeof(double) * c.nz, hipMemcpyDeviceToHost));
    auto near = [](double a, double b) { return std::fabs(a - b) <= 1e-12 * (1.0 + std::fabs(a)); };
    for (int I = 0; I < c.nx; ++I)
        if (!near(xc[I], xf[(size_t)fx * I])) bad("x faces do not coincide");
    for (int J = 0; J < c.ny; ++J)
        if (!near(yc[J], yf[(size_t)fy * J])) bad("y faces do not coincide");
    for (int K = 0; K < c.nz - 1; ++K)
        if (!near(zc[K], zf[(size_t)fz * K])) bad("z faces do not coincide");
    const long long ncell = (long long)(c.nxg - 2) * (c.nyg - 2) * (c.nzg - 3);
    if (m.n + ncell > m.cap) bad("dataset capacity exceeded");
    std::vector<double> dy(f.ny, 0.0);
    for (int j = 1; j < f.ny; ++j) dy[j] = yf[j] - yf[j - 1];
    double *ddy = nullptr;
    BL_HIP(hipMalloc(&ddy, sizeof(double) * f.ny));
    BL_HIP(hipMemcpy(ddy, dy.data(), sizeof(double) * f.ny, hipMemcpyHostToDevice));
    BL_HIP(hipStreamSynchronize(coarse.stream));
    BL_HIP(hipStreamSynchronize(m.stream));
    Coarsen cz{f, c, fx, fy, fz, ddy};
    const Field3 &F = fine.buf[fine.icur], &C = coarse.buf[coarse.icur];
    coarse.inc_ok = false;
    const dim3 blk(256);
    hipLaunchKernelGGL(k_coarsen_u, dim3((c.nx + 63) / 64, (c.nyg + 3) / 4, c.nzg), blk, 0, fine.stream, cz, F.u, C.u);
    check();
    hipLaunchKernelGGL(k_coarsen_v, dim3((c.nxg + 63) / 64, (c.ny + 3) / 4, c.nzg), blk, 0, fine.stream, cz, F.v, C.v);
    check();
    hipLaunchKernelGGL(k_coarsen_w, dim3((c.nxg + 63) / 64, (c.nyg + 3) / 4, c.nz), blk, 0, fine.stream, cz, F.w, C.w);
    check();
    hipLaunchKernelGGL(k_samples, dim3((c.nxg - 2 + 63) / 64, (c.nyg - 2 + 3) / 4, c.nzg - 3), blk, 0, fine.stream, cz,
                       coarse.m, F.u, F.v, F.w, C.u, C.v, C.w, 1.0 / delta99, smin, m.rows + m.n * kMlpRow);
    check();
    BL_HIP(hipStreamSynchronize(fine.stream));
    (void)hipFree(ddy);
    m.n += ncell;
    coarse.t = fine.t;
}

// install the network of m in an LES = 4 solver: BF16 weights and the input
// range [min - nsigma sd, max + nsigma sd] of the weighted training rows
void closure_install(Handle &h, Mlp &m, double delta99, double nsigma) {
    if (h.les != 4) {
        last_error() = "bl_set_closure: the solver was not created with les = 4";
        throw ArgError();
    }
    if (m.n < 1) {
        last_error() = "bl_set_closure: the closure has no training rows (needed for the input range)";
        throw ArgError();
    }
    const int nb = 256;
    double *part = nullptr;
    BL_HIP(hipMalloc(&part, sizeof(double) * nb * (4 * kMlpIn + 1)));
    hipLaunchKernelGGL(k_row_stats, dim3(nb), dim3(256), 0, m.stream, m.rows, m.n, part);
    check();
    std::vector<double> hp((size_t)nb * (4 * kMlpIn + 1));
    BL_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, m.stream));
    std::vector<float> th(kMlpParams);
    BL_HIP(hipMemcpyAsync(th.data(), m.theta, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m.stream));
    BL_HIP(hipStreamSynchronize(m.stream));
    (void)hipFree(part);
    std::vector<float> lo(kMlpIn), hi(kMlpIn);
    double cnt = 0.0;
    for (int b = 0; b < nb; ++b) cnt += hp[(size_t)b * (4 * kMlpIn + 1) + 4 * kMlpIn];
    if (cnt < 2.0) {
        last_error() = "bl_set_closure: fewer than two weighted training rows";
        throw ArgError();
    }
    for (int q = 0; q < kMlpIn; ++q) {
        double mn = 1e300, mx = -1e300, s = 0.0, s2 = 0.0;
        for (int b = 0; b < nb; ++b) {
            const double *P = &hp[(size_t)b * (4 * kMlpIn + 1)];
            mn = std::fmin(mn, P[4 * q]);
            mx = std::fmax(mx, P[4 * q + 1]);
            s += P[4 * q + 2];
            s2 += P[4 * q + 3];
        }
        const double mean = s / cnt, var = std::fmax(s2 / cnt - mean * mean, 0.0), sd = std::sqrt(var);
        lo[q] = (float)(mn - nsigma * sd);
        hi[q] = (float)(mx + nsigma * sd);
    }
    std::vector<unsigned short> wbf;
    std::vector<float> bias;
    closure_weights(th.data(), wbf, bias);
    if (!h.cl_w) {
        BL_HIP(hipMalloc(&h.cl_w, wbf.size() * sizeof(unsigned short)));
        BL_HIP(hipMalloc(&h.cl_b, (bias.size() + 2 * kMlpIn) * sizeof(float)));
        h.cl_lo = h.cl_b + bias.size();
        h.cl_hi = h.cl_lo + kMlpIn;
        h.cl_nblk = closure_blocks(h.d);
        BL_HIP(hipMalloc(&h.cl_cnt, sizeof(double) * h.cl_nblk));
    }
    BL_HIP(hipMemcpyAsync(h.cl_w, wbf.data(), wbf.size() * sizeof(unsigned short), hipMemcpyHostToDevice, h.stream));
    BL_HIP(hipMemcpyAsync(h.cl_b, bias.data(), bias.size() * sizeof(float), hipMemcpyHostToDevice, h.stream));
    BL_HIP(hipMemcpyAsync(h.cl_lo, lo.data(), kMlpIn * sizeof(float), hipMemcpyHostToDevice, h.stream));
    BL_HIP(hipMemcpyAsync(h.cl_hi, hi.data(), kMlpIn * sizeof(float), hipMemcpyHostToDevice, h.stream));
    BL_HIP(hipMemsetAsync(h.cl_cnt, 0, sizeof(double) * h.cl_nblk, h.stream));
    BL_HIP(hipStreamSynchronize(h.stream));
    h.cl_inv_d99 = 1.0 / delta99;
    h.cl_ready = 1;
}

void closure_fallback(Handle &h, long long *cells, long long *fallback) {
    *cells = (long long)(h.d.nxg - 2) * (h.d.nyg - 2) * (h.d.nzg - 2);
    *fallback = 0;
    if (!h.cl_ready) return;
    std::vector<double> c(h.cl_nblk);
    BL_HIP(hipMemcpyAsync(c.data(), h.cl_cnt, sizeof(double) * h.cl_nblk, hipMemcpyDeviceToHost, h.stream));
    BL_HIP(hipStreamSynchronize(h.stream));
    double s = 0.0;
    for (double v : c) s += v;
    *fallback = (long long)s;
}

}  // namespace blgpu
