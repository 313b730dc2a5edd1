// Statistics on the device and asynchronous snapshot copies.
//
// compute_statistics (statistics.f90:44-80, 125-135) reduces the fields to
// z means on the (x faces, y faces) grid at every nstats-th step; the
// reference downloads three interpolated fields for it
// (statistics.f90:56).  Here one thread per (ii, jj) column sums its z planes
// in FP64 in plane order and only the (nx, ny) means cross to the host.
//
// output_data (input_output.f90:496-680) needs the full fields every
// nsave-th step.  The snapshot path packs U, V, W, P and nu_t into one
// device staging area in the reference's unpadded FP64 layout on the
// solver's stream, then copies it to pinned host memory on a second stream,
// so that the copy overlaps the steps that follow.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/blgpu.h"
#include "bl_internal.hpp"

namespace blgpu {
namespace {

void check_launch() { BL_HIP(hipGetLastError()); }

// one thread per (ii, jj) = (i, j) 0-based, i < nx, j < ny:
// term_1 = interpolate_y(U, 2), term_2 = interpolate_x(V, 1) over planes
// k = 1..nzg-2 (1-based 2..nzg-1, the periodic duplicate nzg-1 included);
// term = interpolate_y(interpolate_x(W, 1), 2) over k = 0..nz-3 (1-based
// 1..nz-2) (interpolation.f90:24-85)
template <class T>
__global__ __launch_bounds__(256) void k_statistics(Dims d, Metrics m, const T *__restrict__ U,
                                                    const T *__restrict__ V, const T *__restrict__ W,
                                                    const T *__restrict__ P, const T *__restrict__ NT,
                                                    double *__restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= d.nx) return;
    const double w0 = m.wy0[j], w1 = m.wy1[j];
    double su = 0, sv = 0, sw = 0, su2 = 0, sv2 = 0, sw2 = 0, suv = 0, sp = 0, sp2 = 0, sn = 0;
    for (int k = 1; k <= d.nzg - 2; ++k) {
        const long long p = fidx(d, i, j, k);
        const double t1 = w0 * (double)U[p] + w1 * (double)U[p + d.sx];
        const double t2 = 0.5 * ((double)V[p] + (double)V[p + 1]);
        const double pp = (double)P[p];
        su += t1;
        sv += t2;
        su2 += t1 * t1;
        sv2 += t2 * t2;
        suv += t1 * t2;
        sp += pp;
        sp2 += pp * pp;
        sn += (double)NT[p];
    }
    for (int k = 0; k <= d.nz - 3; ++k) {
        const long long p = fidx(d, i, j, k);
        const double a = 0.5 * ((double)W[p] + (double)W[p + 1]);
        const double b = 0.5 * ((double)W[p + d.sx] + (double)W[p + d.sx + 1]);
        const double t = w0 * a + w1 * b;
        sw += t;
        sw2 += t * t;
    }
    const long long n2 = (long long)d.nx * d.ny, q = i + (long long)d.nx * j;
    const double dz = (double)(d.nzg - 2), dw = (double)(d.nz - 2);
    out[0 * n2 + q] = su / dz;
    out[1 * n2 + q] = sv / dz;
    out[2 * n2 + q] = sw / dw;
    out[3 * n2 + q] = su2 / dz;
    out[4 * n2 + q] = sv2 / dz;
    out[5 * n2 + q] = sw2 / dw;
    out[6 * n2 + q] = suv / dz;
    out[7 * n2 + q] = sp / dz;
    out[8 * n2 + q] = sp2 / dz;
    // nu_t_mean (nxg, nyg): column nxg and row nyg stay 0 (written below)
    out[9 * n2 + i + (long long)d.nxg * j] = sn / dz;
}

// the parts of nu_t_mean that compute_statistics never writes (column nxg,
// row nyg): zero, as initialised (initialization.f90:501)
__global__ __launch_bounds__(256) void k_statistics_edges(Dims d, double *__restrict__ ntm) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < d.nyg) ntm[d.nx + (long long)d.nxg * t] = 0.0;
    if (t < d.nxg) ntm[t + (long long)d.nxg * d.ny] = 0.0;
}

// padded device field (element T) -> unpadded FP64 array (n1, n2, n3)
template <class T>
__global__ __launch_bounds__(256) void k_pack(Dims d, const T *__restrict__ src, double *__restrict__ dst, int n1,
                                              int n2, int n3) {
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= n1 || j >= n2 || k >= n3) return;
    dst[i + (long long)n1 * (j + (long long)n2 * k)] = (double)src[fidx(d, i, j, k)];
}

template <class T>
void pack(Handle &h, const double *src, double *dst, int n1, int n2, int n3) {
    hipLaunchKernelGGL(k_pack<T>, dim3((n1 + 255) / 256, n2, n3), dim3(256), 0, h.stream, h.d,
                       reinterpret_cast<const T *>(src), dst, n1, n2, n3);
    check_launch();
}

}  // namespace

long long stats_size(const Dims &d) { return 9LL * d.nx * d.ny + (long long)d.nxg * d.nyg; }

void launch_statistics(Handle &h, const Field3 &f, double *out) {
    const Dims &d = h.d;
    const dim3 grid((d.nx + 255) / 256, d.ny);
    if (h.precision == BL_PREC_MIXED) {
        const Fld3<float> g = as<float>(f);
        hipLaunchKernelGGL(k_statistics<float>, grid, dim3(256), 0, h.stream, d, h.m, g.u, g.v, g.w,
                           reinterpret_cast<const float *>(h.P), reinterpret_cast<const float *>(h.nu_t), out);
    } else {
        hipLaunchKernelGGL(k_statistics<double>, grid, dim3(256), 0, h.stream, d, h.m, f.u, f.v, f.w, h.P,
                           h.nu_t, out);
    }
    check_launch();
    hipLaunchKernelGGL(k_statistics_edges, dim3((std::max(d.nxg, d.nyg) + 255) / 256), dim3(256), 0, h.stream, d,
                       out + 9LL * d.nx * d.ny);
    check_launch();
}

void snapshot_extents(const Dims &d, int ext[5][3]) {
    const int e[5][3] = {{d.nx, d.nyg, d.nzg}, {d.nxg, d.ny, d.nzg}, {d.nxg, d.nyg, d.nz},
                         {d.nxg, d.nyg, d.nzg}, {d.nxg, d.nyg, d.nzg}};
    for (int a = 0; a < 5; ++a)
        for (int b = 0; b < 3; ++b) ext[a][b] = e[a][b];
}

void snapshot_begin(Handle &h) {
    const Dims &d = h.d;
    int ext[5][3];
    snapshot_extents(d, ext);
    size_t total = 0;
    for (int a = 0; a < 5; ++a) {
        h.snap_off[a] = total;
        total += (size_t)ext[a][0] * ext[a][1] * ext[a][2];
    }
    if (!h.snap_dev) {
        BL_HIP(hipMalloc(&h.snap_dev, total * sizeof(double)));
        BL_HIP(hipHostMalloc((void **)&h.snap_host, total * sizeof(double), hipHostMallocDefault));
        BL_HIP(hipStreamCreateWithFlags(&h.copy_stream, hipStreamNonBlocking));
        BL_HIP(hipEventCreateWithFlags(&h.snap_packed, hipEventDisableTiming));
        BL_HIP(hipEventCreateWithFlags(&h.snap_copied, hipEventDisableTiming));
    } else if (h.snap_pending) {
        // the previous copy still reads the staging area
        BL_HIP(hipStreamWaitEvent(h.stream, h.snap_copied, 0));
    }
    const Field3 &f = h.buf[h.icur];
    const double *src[5] = {f.u, f.v, f.w, h.P, h.nu_t};
    for (int a = 0; a < 5; ++a) {
        if (h.precision == BL_PREC_MIXED)
            pack<float>(h, src[a], h.snap_dev + h.snap_off[a], ext[a][0], ext[a][1], ext[a][2]);
        else
            pack<double>(h, src[a], h.snap_dev + h.snap_off[a], ext[a][0], ext[a][1], ext[a][2]);
    }
    BL_HIP(hipEventRecord(h.snap_packed, h.stream));
    BL_HIP(hipStreamWaitEvent(h.copy_stream, h.snap_packed, 0));
    BL_HIP(hipMemcpyAsync(h.snap_host, h.snap_dev, total * sizeof(double), hipMemcpyDeviceToHost, h.copy_stream));
    BL_HIP(hipEventRecord(h.snap_copied, h.copy_stream));
    h.snap_pending = true;
}

void snapshot_end(Handle &h, const double **ptr) {
    if (!h.snap_dev) {
        last_error() = "bl_snapshot_end: no snapshot was begun";
        throw StateError();
    }
    BL_HIP(hipEventSynchronize(h.snap_copied));
    h.snap_pending = false;
    for (int a = 0; a < 5; ++a) ptr[a] = h.snap_host + h.snap_off[a];
}

void snapshot_free(Handle &h) {
    if (h.snap_pending && h.snap_copied) (void)hipEventSynchronize(h.snap_copied);
    if (h.snap_dev) (void)hipFree(h.snap_dev);
    if (h.snap_host) (void)hipHostFree(h.snap_host);
    if (h.copy_stream) (void)hipStreamDestroy(h.copy_stream);
    if (h.snap_packed) (void)hipEventDestroy(h.snap_packed);
    if (h.snap_copied) (void)hipEventDestroy(h.snap_copied);
    h.snap_dev = h.snap_host = nullptr;
    h.copy_stream = nullptr;
    h.snap_packed = h.snap_copied = nullptr;
    h.snap_pending = false;
}

}  // namespace blgpu
