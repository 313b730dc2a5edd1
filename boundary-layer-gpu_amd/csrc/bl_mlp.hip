// Learned subgrid closure: training on the device (no counterpart in the
// reference solver, which only has the dynamic Smagorinsky model).
//
// Network: 10 -> 64 -> 64 -> 64 -> 1 with SiLU, predicting the dimensionless
// eddy viscosity c = nu_t / (Delta^2 |S|) from the normalised velocity-
// gradient tensor and the wall distance.  FP32 master parameters, Adam; the
// training forward/backward passes run in FP32 (one sample per lane, weights
// broadcast from LDS), the weight gradients are per-block partial sums over a
// batch slice reduced in a fixed order (deterministic).  A BF16 copy of the
// weights feeds the MFMA inference kernel (bl_closure.hip).
//
// Dataset rows (kRow floats): 10 features, target c, weight, Delta^2 |S|,
// then S_ij and tau_ij (6 + 6, the a-priori checks).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/blgpu.h"
#include "bl_mlp.hpp"

namespace blgpu {

namespace {

constexpr int kH = kMlpHidden, kIn = kMlpIn;
// parameter offsets (W1[64][10], b1, W2[64][64], b2, W3[64][64], b3, W4[64], b4)
constexpr int oW1 = 0, ob1 = oW1 + kH * kIn, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH,
              ob3 = oW3 + kH * kH, oW4 = ob3 + kH, ob4 = oW4 + kH;
static_assert(ob4 + 1 == kMlpParams, "parameter layout");

__device__ __forceinline__ float silu(float z) { return z / (1.f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.f / (1.f + expf(-z));
    return s * (1.f + z * (1.f - s));
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// batch gather: sample b of step s is row hash(seed, s, b) mod n
__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ rows, long long n, int B,
                                                unsigned long long seed, long long step, float wscale,
                                                float *__restrict__ Xb, float *__restrict__ tb,
                                                float *__restrict__ wb) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    const unsigned long long h = mix64(seed ^ mix64((unsigned long long)step * 0x100000001B3ull + (unsigned long long)b));
    const long long r = (long long)(h % (unsigned long long)n);
    const float *row = rows + r * kMlpRow;
#pragma unroll
    for (int f = 0; f < kIn; ++f) Xb[(long long)f * B + b] = row[f];
    tb[b] = row[kIn];
    wb[b] = row[kIn + 1] * wscale;
}

// sum and count of the positive row weights, per block (fixed order)
__global__ __launch_bounds__(256) void k_wsum(const float *__restrict__ rows, long long n, double *__restrict__ part) {
    __shared__ double sh[2][4];
    double s = 0.0, c = 0.0;
    for (long long r = (long long)blockIdx.x * 256 + threadIdx.x; r < n; r += (long long)gridDim.x * 256) {
        const float w = rows[r * kMlpRow + kIn + 1];
        if (w > 0.f) {
            s += w;
            c += 1.0;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_down(s, o);
        c += __shfl_down(c, o);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = s;
        sh[1][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
        part[2 * blockIdx.x + 1] = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    }
}

// stage the parameters in LDS
__device__ __forceinline__ void load_params(float *sp, const float *__restrict__ theta) {
    for (int q = threadIdx.x; q < kMlpParams; q += blockDim.x) sp[q] = theta[q];
    __syncthreads();
}

// forward, one sample per lane: Z[l][f][b] pre-activations, y[b]
__global__ __launch_bounds__(256) void k_fwd(const float *__restrict__ theta, const float *__restrict__ Xb, int B,
                                             float *__restrict__ Z, float *__restrict__ y) {
    __shared__ float sp[kMlpParams];
    load_params(sp, theta);
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    float x[kIn], h[kH], z[kH];
#pragma unroll
    for (int f = 0; f < kIn; ++f) x[f] = Xb[(long long)f * B + b];
#pragma unroll
    for (int o = 0; o < kH; ++o) {
        float a = sp[ob1 + o];
#pragma unroll
        for (int i = 0; i < kIn; ++i) a += sp[oW1 + o * kIn + i] * x[i];
        z[o] = a;
    }
#pragma unroll
    for (int o = 0; o < kH; ++o) {
        Z[(long long)o * B + b] = z[o];
        h[o] = silu(z[o]);
    }
    for (int l = 1; l < 3; ++l) {
        const int oW = l == 1 ? oW2 : oW3, ob = l == 1 ? ob2 : ob3;
#pragma unroll
        for (int o = 0; o < kH; ++o) {
            float a = sp[ob + o];
#pragma unroll
            for (int i = 0; i < kH; ++i) a += sp[oW + o * kH + i] * h[i];
            z[o] = a;
        }
#pragma unroll
        for (int o = 0; o < kH; ++o) {
            Z[((long long)l * kH + o) * B + b] = z[o];
            h[o] = silu(z[o]);
        }
    }
    float a = sp[ob4];
#pragma unroll
    for (int i = 0; i < kH; ++i) a += sp[oW4 + i] * h[i];
    y[b] = a;
}

// backward, one sample per lane: loss = (1/B) sum w (y - t)^2;
// D[l][f][b] = dloss/dz, dy[b] = dloss/dy, sq[b] = w (y - t)^2
__global__ __launch_bounds__(256) void k_bwd(const float *__restrict__ theta, const float *__restrict__ Z,
                                             const float *__restrict__ y, const float *__restrict__ tb,
                                             const float *__restrict__ wb, int B, float *__restrict__ D,
                                             float *__restrict__ dyo, float *__restrict__ sq) {
    __shared__ float sp[kMlpParams];
    load_params(sp, theta);
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    const float e = y[b] - tb[b];
    const float w = wb[b];
    sq[b] = w * e * e;
    const float dy = 2.f * w * e / (float)B;
    dyo[b] = dy;
    float d[kH], a[kH];
#pragma unroll
    for (int o = 0; o < kH; ++o) d[o] = dy * sp[oW4 + o] * dsilu(Z[(2LL * kH + o) * B + b]);
#pragma unroll
    for (int o = 0; o < kH; ++o) D[(2LL * kH + o) * B + b] = d[o];
    for (int l = 2; l >= 1; --l) {
        const int oW = l == 2 ? oW3 : oW2;
#pragma unroll
        for (int i = 0; i < kH; ++i) a[i] = 0.f;
#pragma unroll
        for (int o = 0; o < kH; ++o)
#pragma unroll
            for (int i = 0; i < kH; ++i) a[i] += sp[oW + o * kH + i] * d[o];
#pragma unroll
        for (int i = 0; i < kH; ++i) {
            d[i] = a[i] * dsilu(Z[((long long)(l - 1) * kH + i) * B + b]);
            D[((long long)(l - 1) * kH + i) * B + b] = d[i];
        }
    }
}

// per-block partial parameter gradients over a slice of kSlice samples
constexpr int kSub = 16;  // samples staged in LDS per sub-step
constexpr int kSlice = 512;
__global__ __launch_bounds__(256) void k_wgrad(const float *__restrict__ Xb, const float *__restrict__ Z,
                                               const float *__restrict__ D, const float *__restrict__ dy, int B,
                                               float *__restrict__ part) {
    __shared__ float sx[kIn][kSub], sh[3][kH][kSub], sd[3][kH][kSub], sdy[kSub];
    const int t = threadIdx.x;
    const int s0 = blockIdx.x * kSlice, s1 = min(s0 + kSlice, B);
    // thread t owns: W2/W3 rows o = t/4, columns 16*(t%4) .. +16 (32 accumulators)
    const int o = t >> 2, i0 = (t & 3) * 16;
    float a2[16], a3[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) a2[q] = a3[q] = 0.f;
    // and W1 entries t, t+256, t+512 (< 640), one bias or W4 entry each, b4 on thread 0
    float a1[3] = {0.f, 0.f, 0.f}, ab = 0.f, ab4 = 0.f;
    for (int s = s0; s < s1; s += kSub) {
        const int ns = min(kSub, s1 - s);
        for (int q = t; q < kSub * (kIn + 6 * kH + 1); q += 256) {
            const int c = q % kSub, r = q / kSub;
            const int b = s + c;
            const bool ok = c < ns;
            if (r < kIn) {
                sx[r][c] = ok ? Xb[(long long)r * B + b] : 0.f;
            } else if (r < kIn + 3 * kH) {
                const int l = (r - kIn) / kH, f = (r - kIn) % kH;
                sh[l][f][c] = ok ? silu(Z[((long long)l * kH + f) * B + b]) : 0.f;
            } else if (r < kIn + 6 * kH) {
                const int l = (r - kIn - 3 * kH) / kH, f = (r - kIn - 3 * kH) % kH;
                sd[l][f][c] = ok ? D[((long long)l * kH + f) * B + b] : 0.f;
            } else {
                sdy[c] = ok ? dy[b] : 0.f;
            }
        }
        __syncthreads();
        for (int c = 0; c < kSub; ++c) {
            const float d2 = sd[1][o][c], d3 = sd[2][o][c];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                a2[q] += d2 * sh[0][i0 + q][c];
                a3[q] += d3 * sh[1][i0 + q][c];
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const int p = t + 256 * r;
                if (p < kH * kIn) a1[r] += sd[0][p / kIn][c] * sx[p % kIn][c];
            }
            if (t < kH)
                ab += sd[0][t][c];  // b1
            else if (t < 2 * kH)
                ab += sd[1][t - kH][c];  // b2
            else if (t < 3 * kH)
                ab += sd[2][t - 2 * kH][c];  // b3
            else
                ab += sdy[c] * sh[2][t - 3 * kH][c];  // W4
            if (t == 0) ab4 += sdy[c];            // b4
        }
        __syncthreads();
    }
    float *P = part + (long long)blockIdx.x * kMlpParams;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        P[oW2 + o * kH + i0 + q] = a2[q];
        P[oW3 + o * kH + i0 + q] = a3[q];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int p = t + 256 * r;
        if (p < kH * kIn) P[oW1 + p] = a1[r];
    }
    if (t < kH)
        P[ob1 + t] = ab;
    else if (t < 2 * kH)
        P[ob2 + t - kH] = ab;
    else if (t < 3 * kH)
        P[ob3 + t - 2 * kH] = ab;
    else
        P[oW4 + t - 3 * kH] = ab;
    if (t == 0) P[ob4] = ab4;
}

// g[p] = sum over blocks of part[blk][p], fixed order
__global__ __launch_bounds__(256) void k_gsum(const float *__restrict__ part, int nblk, float *__restrict__ g) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kMlpParams) return;
    float a = 0.f;
    for (int q = 0; q < nblk; ++q) a += part[(long long)q * kMlpParams + p];
    g[p] = a;
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ theta, float *__restrict__ m, float *__restrict__ v,
                                              const float *__restrict__ g, float lr, float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kMlpParams) return;
    const float gp = g[p];
    const float mp = 0.9f * m[p] + 0.1f * gp;
    const float vp = 0.999f * v[p] + 0.001f * gp * gp;
    m[p] = mp;
    v[p] = vp;
    theta[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + 1e-8f);
}

// sum of n values into out[0], one block, fixed order
__global__ __launch_bounds__(256) void k_sum(const float *__restrict__ a, int n, float *__restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += 256) s += a[q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)sh[0];
}

void check() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        last_error() = std::string("kernel launch: ") + hipGetErrorString(e);
        throw DeviceError();
    }
}

template <class T>
T *alloc(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
        last_error() = "out of device memory (learned closure, " + std::to_string(n * sizeof(T)) + " bytes)";
        throw std::bad_alloc();
    }
    return static_cast<T *>(p);
}

}  // namespace

void mlp_init(Mlp &m, long long capacity, int batch_max) {
    m.cap = capacity;
    m.n = 0;
    m.bmax = batch_max;
    BL_HIP(hipStreamCreate(&m.stream));
    m.rows = alloc<float>((size_t)capacity * kMlpRow);
    m.theta = alloc<float>(kMlpParams);
    m.adam_m = alloc<float>(kMlpParams);
    m.adam_v = alloc<float>(kMlpParams);
    m.g = alloc<float>(kMlpParams);
    m.nslice = (batch_max + kSlice - 1) / kSlice;
    m.part = alloc<float>((size_t)m.nslice * kMlpParams);
    m.Xb = alloc<float>((size_t)kIn * batch_max);
    m.tb = alloc<float>(batch_max);
    m.wb = alloc<float>(batch_max);
    m.Z = alloc<float>((size_t)3 * kH * batch_max);
    m.D = alloc<float>((size_t)3 * kH * batch_max);
    m.y = alloc<float>(batch_max);
    m.dy = alloc<float>(batch_max);
    m.sq = alloc<float>(batch_max);
    m.scal = alloc<float>(16);
    BL_HIP(hipMemsetAsync(m.theta, 0, kMlpParams * sizeof(float), m.stream));
    BL_HIP(hipMemsetAsync(m.adam_m, 0, kMlpParams * sizeof(float), m.stream));
    BL_HIP(hipMemsetAsync(m.adam_v, 0, kMlpParams * sizeof(float), m.stream));
    m.step = 0;
}

void mlp_free(Mlp &m) {
    for (void *p : {(void *)m.rows, (void *)m.theta, (void *)m.adam_m, (void *)m.adam_v, (void *)m.g, (void *)m.part,
                    (void *)m.Xb, (void *)m.tb, (void *)m.wb, (void *)m.Z, (void *)m.D, (void *)m.y, (void *)m.dy,
                    (void *)m.sq, (void *)m.scal})
        if (p) (void)hipFree(p);
    if (m.stream) {
        (void)hipStreamSynchronize(m.stream);
        (void)hipStreamDestroy(m.stream);
    }
}

// forward + backward + weight gradient of the batch already in Xb/tb/wb;
// leaves g and the batch loss sum in scal[0]
void mlp_fwd_bwd(Mlp &m, int B, bool backward) {
    const dim3 gb((B + 255) / 256), blk(256);
    hipLaunchKernelGGL(k_fwd, gb, blk, 0, m.stream, m.theta, m.Xb, B, m.Z, m.y);
    check();
    if (!backward) return;
    hipLaunchKernelGGL(k_bwd, gb, blk, 0, m.stream, m.theta, m.Z, m.y, m.tb, m.wb, B, m.D, m.dy, m.sq);
    check();
    const int ns = (B + kSlice - 1) / kSlice;
    hipLaunchKernelGGL(k_wgrad, dim3(ns), blk, 0, m.stream, m.Xb, m.Z, m.D, m.dy, B, m.part);
    check();
    hipLaunchKernelGGL(k_gsum, dim3((kMlpParams + 255) / 256), blk, 0, m.stream, m.part, ns, m.g);
    check();
    hipLaunchKernelGGL(k_sum, dim3(1), blk, 0, m.stream, m.sq, B, m.scal);
    check();
}

void mlp_load_batch(Mlp &m, const float *x, const float *t, int n) {
    // host rows x[n][10] -> feature-major Xb; targets; unit weights
    std::vector<float> xb((size_t)kIn * n), w((size_t)n, 1.f), tt((size_t)n, 0.f);
    for (int b = 0; b < n; ++b)
        for (int f = 0; f < kIn; ++f) xb[(size_t)f * n + b] = x[(size_t)b * kIn + f];
    if (t) std::memcpy(tt.data(), t, sizeof(float) * n);
    BL_HIP(hipMemcpyAsync(m.Xb, xb.data(), xb.size() * sizeof(float), hipMemcpyHostToDevice, m.stream));
    BL_HIP(hipMemcpyAsync(m.tb, tt.data(), tt.size() * sizeof(float), hipMemcpyHostToDevice, m.stream));
    BL_HIP(hipMemcpyAsync(m.wb, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, m.stream));
    BL_HIP(hipStreamSynchronize(m.stream));
}

void mlp_train(Mlp &m, int steps, int B, float lr, unsigned long long seed, float *loss_hist) {
    if (m.n < 1) {
        last_error() = "bl_mlp_train: the dataset is empty";
        throw ArgError();
    }
    if (B < 1 || B > m.bmax) {
        last_error() = "bl_mlp_train: batch must be in [1, batch_max]";
        throw ArgError();
    }
    const dim3 blk(256);
    // weights rescaled to mean 1 over the weighted rows of the dataset (the
    // partial sums go through the gradient buffer, which each step rewrites)
    constexpr int kWBlk = 128;
    static_assert(2 * kWBlk * sizeof(double) <= kMlpParams * sizeof(float), "partials fit in g");
    double *wpart = reinterpret_cast<double *>(m.g);
    hipLaunchKernelGGL(k_wsum, dim3(kWBlk), blk, 0, m.stream, m.rows, m.n, wpart);
    check();
    double hp[2 * kWBlk];
    BL_HIP(hipMemcpyAsync(hp, wpart, sizeof(hp), hipMemcpyDeviceToHost, m.stream));
    BL_HIP(hipStreamSynchronize(m.stream));
    double wsum = 0.0, wcnt = 0.0;
    for (int q = 0; q < kWBlk; ++q) {
        wsum += hp[2 * q];
        wcnt += hp[2 * q + 1];
    }
    if (!(wsum > 0.0)) {
        last_error() = "bl_mlp_train: the dataset has no row with a positive weight";
        throw ArgError();
    }
    const float wscale = (float)(wcnt / wsum);
    for (int s = 0; s < steps; ++s) {
        hipLaunchKernelGGL(k_gather, dim3((B + 255) / 256), blk, 0, m.stream, m.rows, m.n, B, seed, m.step, wscale,
                           m.Xb, m.tb, m.wb);
        check();
        mlp_fwd_bwd(m, B, true);
        m.step += 1;
        const float c1 = 1.f - powf(0.9f, (float)m.step), c2 = 1.f - powf(0.999f, (float)m.step);
        hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), blk, 0, m.stream, m.theta, m.adam_m, m.adam_v,
                           m.g, lr, c1, c2);
        check();
        if (loss_hist) {
            float l = 0.f;
            BL_HIP(hipMemcpyAsync(&l, m.scal, sizeof(float), hipMemcpyDeviceToHost, m.stream));
            BL_HIP(hipStreamSynchronize(m.stream));
            loss_hist[s] = l / (float)B;
        }
    }
    BL_HIP(hipStreamSynchronize(m.stream));
}

}  // namespace blgpu

// ------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------
using namespace blgpu;

namespace {
template <class F>
int mlp_guarded(F &&fn) {
    try {
        fn();
        return BL_OK;
    } catch (const ArgError &) {
        return BL_EINVAL;
    } catch (const DeviceError &) {
        return BL_EDEVICE;
    } catch (const std::bad_alloc &) {
        if (last_error().empty()) last_error() = "out of memory";
        return BL_ENOMEM;
    } catch (...) {
        last_error() = "unexpected exception";
        return BL_EDEVICE;
    }
}
Mlp &MM(bl_mlp *m) {
    if (!m) {
        last_error() = "null closure handle";
        throw ArgError();
    }
    return *m;
}
void need(bool ok, const char *msg) {
    if (!ok) {
        last_error() = msg;
        throw ArgError();
    }
}
}  // namespace

extern "C" {

int bl_mlp_create(long long capacity, int batch_max, bl_mlp **out) {
    last_error().clear();
    if (!out || capacity < 1 || batch_max < 1) {
        last_error() = "bl_mlp_create: need capacity >= 1, batch_max >= 1 and an output pointer";
        return BL_EINVAL;
    }
    *out = nullptr;
    bl_mlp *m = new (std::nothrow) bl_mlp();
    if (!m) return BL_ENOMEM;
    std::memset(static_cast<Mlp *>(m), 0, sizeof(Mlp));
    const int rc = mlp_guarded([&] { mlp_init(*m, capacity, batch_max); });
    if (rc != BL_OK) {
        mlp_free(*m);
        delete m;
        return rc;
    }
    *out = m;
    return BL_OK;
}

void bl_mlp_destroy(bl_mlp *m) {
    if (!m) return;
    mlp_free(*m);
    delete m;
}

int bl_mlp_set_params(bl_mlp *mm, const float *theta) {
    return mlp_guarded([&] {
        Mlp &m = MM(mm);
        need(theta != nullptr, "bl_mlp_set_params: null parameters");
        BL_HIP(hipMemcpyAsync(m.theta, theta, kMlpParams * sizeof(float), hipMemcpyHostToDevice, m.stream));
        BL_HIP(hipMemsetAsync(m.adam_m, 0, kMlpParams * sizeof(float), m.stream));
        BL_HIP(hipMemsetAsync(m.adam_v, 0, kMlpParams * sizeof(float), m.stream));
        BL_HIP(hipStreamSynchronize(m.stream));
        m.step = 0;
    });
}

int bl_mlp_get_params(bl_mlp *mm, float *theta) {
    return mlp_guarded([&] {
        Mlp &m = MM(mm);
        need(theta != nullptr, "bl_mlp_get_params: null output");
        BL_HIP(hipMemcpyAsync(theta, m.theta, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, m.stream));
        BL_HIP(hipStreamSynchronize(m.stream));
    });
}

int bl_mlp_forward(bl_mlp *mm, const float *x, int n, float *y) {
    return mlp_guarded([&] {
        Mlp &m = MM(mm);
        need(x && y && n >= 1 && n <= m.bmax, "bl_mlp_forward: need 1 <= n <= batch_max and buffers");
        mlp_load_batch(m, x, nullptr, n);
        mlp_fwd_bwd(m, n, false);
        BL_HIP(hipMemcpyAsync(y, m.y, sizeof(float) * n, hipMemcpyDeviceToHost, m.stream));
        BL_HIP(hipStreamSynchronize(m.stream));
    });
}

int bl_mlp_gradient(bl_mlp *mm, const float *x, const float *t, int n, float *g, float *loss) {
    return mlp_guarded([&] {
        Mlp &m = MM(mm);
        need(x && t && g && n >= 1 && n <= m.bmax, "bl_mlp_gradient: need 1 <= n <= batch_max and buffers");
        mlp_load_batch(m, x, t, n);
        mlp_fwd_bwd(m, n, true);
        BL_HIP(hipMemcpyAsync(g, m.g, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m.stream));
        float l = 0.f;
        BL_HIP(hipMemcpyAsync(&l, m.scal, sizeof(float), hipMemcpyDeviceToHost, m.stream));
        BL_HIP(hipStreamSynchronize(m.stream));
        if (loss) *loss = l / (float)n;
    });
}

int bl_mlp_add_rows(bl_mlp *mm, const float *rows, long long n) {
    return mlp_guarded([&] {
        Mlp &m = MM(mm);
        need(rows && n >= 0, "bl_mlp_add_rows: null rows");
        need(m.n + n <= m.cap, "bl_mlp_add_rows: dataset capacity exceeded");
        BL_HIP(hipMemcpyAsync(m.rows + m.n * kMlpRow, rows, sizeof(float) * kMlpRow * n, hipMemcpyHostToDevice,
                              m.stream));
        BL_HIP(hipStreamSynchronize(m.stream));
        m.n += n;
    });
}

int bl_mlp_get_rows(bl_mlp *mm, long long first, long long n, float *rows) {
    return mlp_guarded([&] {
        Mlp &m = MM(mm);
        need(rows && first >= 0 && n >= 0 && first + n <= m.n, "bl_mlp_get_rows: range outside the dataset");
        BL_HIP(hipMemcpyAsync(rows, m.rows + first * kMlpRow, sizeof(float) * kMlpRow * n, hipMemcpyDeviceToHost,
                              m.stream));
        BL_HIP(hipStreamSynchronize(m.stream));
    });
}

int bl_mlp_count(bl_mlp *mm, long long *n) {
    return mlp_guarded([&] {
        need(n != nullptr, "bl_mlp_count: null output");
        *n = MM(mm).n;
    });
}

int bl_mlp_clear(bl_mlp *mm) {
    return mlp_guarded([&] { MM(mm).n = 0; });
}

int bl_mlp_train(bl_mlp *mm, int steps, int batch, float lr, unsigned long long seed, float *loss_hist) {
    return mlp_guarded([&] {
        Mlp &m = MM(mm);
        need(steps >= 0 && lr > 0.f, "bl_mlp_train: need steps >= 0 and lr > 0");
        mlp_train(m, steps, batch, lr, seed, loss_hist);
    });
}

}  // extern "C"
