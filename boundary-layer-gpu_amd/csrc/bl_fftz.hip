// The fftz module (fftz.f90:32-127): real DFT along z of xz planes,
// U(1:n1, 1:N) -> U_hat(1:n1, 1:N/2+1) (r2c, unnormalised) and back (c2r,
// divided by N).  The reference plans it once and never executes it on the
// time step (initialization.f90:720); this plan runs the solver's z-DFT
// kernels (the same as the Poisson solve's z transform) on host arrays, which
// also pins those kernels against any r2c/c2r implementation.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/blgpu.h"
#include "bl_internal.hpp"

struct bl_fftz {
    int N = 0;
    int lda = 0;
    double *zc = nullptr, *zs = nullptr, *zf = nullptr;
    hipStream_t stream = nullptr;
};

namespace blgpu {
namespace {

template <class F>
int guard(F &&fn) {
    try {
        fn();
        return BL_OK;
    } catch (const ArgError &) {
        return BL_EINVAL;
    } catch (const DeviceError &) {
        return BL_EDEVICE;
    } catch (const std::bad_alloc &) {
        last_error() = "out of memory";
        return BL_ENOMEM;
    } catch (...) {
        last_error() = "unexpected exception";
        return BL_EDEVICE;
    }
}

// packed half spectrum S[q][i] (q = 0: Re X_0; 2k-1, 2k: Re, Im X_k) <->
// complex U_hat[k][i] interleaved (re, im)
__global__ __launch_bounds__(256) void k_packed_to_complex(int N, int n1, const double *__restrict__ S,
                                                           double *__restrict__ H) {
    const int i = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
    if (i >= n1) return;
    double re, im;
    if (k == 0) {
        re = S[i];
        im = 0.0;
    } else if (2 * k == N) {
        re = S[(long long)(N - 1) * n1 + i];
        im = 0.0;
    } else {
        re = S[(long long)(2 * k - 1) * n1 + i];
        im = S[(long long)(2 * k) * n1 + i];
    }
    const long long o = 2 * (i + (long long)n1 * k);
    H[o] = re;
    H[o + 1] = im;
}

// c2r reads only the real part of the zero and (even N) Nyquist terms; every
// row q = 0..N-1 of S is written
__global__ __launch_bounds__(256) void k_complex_to_packed(int N, int n1, const double *__restrict__ H,
                                                           double *__restrict__ S) {
    const int i = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
    if (i >= n1) return;
    const long long o = 2 * (i + (long long)n1 * k);
    if (k == 0) {
        S[i] = H[o];
    } else if (2 * k == N) {
        S[(long long)(N - 1) * n1 + i] = H[o];
    } else {
        S[(long long)(2 * k - 1) * n1 + i] = H[o];
        S[(long long)(2 * k) * n1 + i] = H[o + 1];
    }
}

__global__ __launch_bounds__(256) void k_scale(long long n, double s, double *__restrict__ a) {
    const long long q = blockIdx.x * 256LL + threadIdx.x;
    if (q < n) a[q] *= s;
}

bl_fftz &P(bl_fftz *p) {
    if (!p) {
        last_error() = "null fftz plan";
        throw ArgError();
    }
    return *p;
}

struct DevBuf {
    double *p = nullptr;
    explicit DevBuf(size_t n) { BL_HIP(hipMalloc(&p, n * sizeof(double))); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

ZTables tables(const bl_fftz &f) {
    ZTables z;
    z.M = f.N;
    z.lda = f.lda;
    z.zc = f.zc;
    z.zs = f.zs;
    z.zf = f.zf;
    return z;
}

void run(bl_fftz &f, int n1, const double *in, double *out, bool forward) {
    if (n1 < 1) {
        last_error() = "fftz: n1 must be positive";
        throw ArgError();
    }
    const int N = f.N, K = N / 2 + 1;
    const size_t nreal = (size_t)n1 * N, ncplx = 2 * (size_t)n1 * K;
    DevBuf a(nreal), b(nreal), c(ncplx);
    const dim3 gk((n1 + 255) / 256, K);
    if (forward) {
        BL_HIP(hipMemcpyAsync(a.p, in, nreal * sizeof(double), hipMemcpyHostToDevice, f.stream));
        launch_zdft(f.stream, tables(f), n1, n1, true, a.p, b.p);
        hipLaunchKernelGGL(k_packed_to_complex, gk, dim3(256), 0, f.stream, N, n1, b.p, c.p);
        BL_HIP(hipGetLastError());
        BL_HIP(hipMemcpyAsync(out, c.p, ncplx * sizeof(double), hipMemcpyDeviceToHost, f.stream));
    } else {
        BL_HIP(hipMemcpyAsync(c.p, in, ncplx * sizeof(double), hipMemcpyHostToDevice, f.stream));
        hipLaunchKernelGGL(k_complex_to_packed, gk, dim3(256), 0, f.stream, N, n1, c.p, b.p);
        BL_HIP(hipGetLastError());
        launch_zdft(f.stream, tables(f), n1, n1, false, b.p, a.p);
        hipLaunchKernelGGL(k_scale, dim3((unsigned)((nreal + 255) / 256)), dim3(256), 0, f.stream, (long long)nreal,
                           1.0 / (double)N, a.p);
        BL_HIP(hipGetLastError());
        BL_HIP(hipMemcpyAsync(out, a.p, nreal * sizeof(double), hipMemcpyDeviceToHost, f.stream));
    }
    BL_HIP(hipStreamSynchronize(f.stream));
}

}  // namespace
}  // namespace blgpu

using namespace blgpu;

extern "C" {

int bl_fftz_create(int N, bl_fftz **out) {
    if (!out) return BL_EINVAL;
    *out = nullptr;
    bl_fftz *f = new bl_fftz();
    const int rc = guard([&] {
        if (N < 2) {
            last_error() = "bl_fftz_create: N must be at least 2";
            throw ArgError();
        }
        f->N = N;
        std::vector<double> zc, zs, zf;
        z_dft_tables(N, f->lda, zc, zs, zf);
        BL_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
        BL_HIP(hipMalloc(&f->zc, zc.size() * sizeof(double)));
        BL_HIP(hipMalloc(&f->zs, zs.size() * sizeof(double)));
        BL_HIP(hipMemcpy(f->zc, zc.data(), zc.size() * sizeof(double), hipMemcpyHostToDevice));
        BL_HIP(hipMemcpy(f->zs, zs.data(), zs.size() * sizeof(double), hipMemcpyHostToDevice));
        if (!zf.empty()) {
            BL_HIP(hipMalloc(&f->zf, zf.size() * sizeof(double)));
            BL_HIP(hipMemcpy(f->zf, zf.data(), zf.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    });
    if (rc != BL_OK) {
        bl_fftz_destroy(f);
        return rc;
    }
    *out = f;
    return BL_OK;
}

void bl_fftz_destroy(bl_fftz *f) {
    if (!f) return;
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    if (f->zc) (void)hipFree(f->zc);
    if (f->zs) (void)hipFree(f->zs);
    if (f->zf) (void)hipFree(f->zf);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
}

int bl_fft_z(bl_fftz *f, int n1, const double *U, double *U_hat) {
    return guard([&] {
        if (!U || !U_hat) throw ArgError();
        run(P(f), n1, U, U_hat, true);
    });
}

int bl_ifft_z(bl_fftz *f, int n1, const double *U_hat, double *U) {
    return guard([&] {
        if (!U || !U_hat) throw ArgError();
        run(P(f), n1, U_hat, U, false);
    });
}

}  // extern "C"
