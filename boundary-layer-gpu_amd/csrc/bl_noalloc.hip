// bl_max_divergence_noalloc: check_divergence (projection.f90:391-413) of
// caller-owned device fields without any device allocation, for callers that
// cannot allocate (HIP graph capture, serving loops).  Its working memory is
// the file-scope array `scratch` below: every call on a device shares that
// device's copy, so calls that overlap on one device race.  Nothing in it has
// to start at zero: each call writes every word it reads.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/blgpu.h"
#include "bl_internal.hpp"

namespace blgpu {
namespace {

// partials of the max reduction, then the reciprocal face spacings
constexpr long long kScratchDoubles = 1 << 17;  // 1 MiB per device
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// the device address of this device's copy of `scratch`, looked up once per
// device with that device current (the array's name is not a device address
// in host code)
double *scratch_on_current_device() {
    int dev = 0;
    BL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        last_error() = "bl_max_divergence_noalloc: device ordinal out of range";
        throw ArgError();
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        BL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    return g_scratch_ptr[dev];
}

template <class F>
int guard(F &&fn) {
    try {
        fn();
        return BL_OK;
    } catch (const ArgError &) {
        return BL_EINVAL;
    } catch (const DeviceError &) {
        return BL_EDEVICE;
    } catch (...) {
        last_error() = "unexpected exception";
        return BL_EDEVICE;
    }
}

// r[i] = 1 / (c[i] - c[i-1]) for i >= 1, r[0] = 0 (the handle's metric form)
__global__ __launch_bounds__(256) void k_recip_faces(const double *__restrict__ c, int n, double *__restrict__ r) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) r[i] = i == 0 ? 0.0 : 1.0 / (c[i] - c[i - 1]);
}

}  // namespace
}  // namespace blgpu

using namespace blgpu;

extern "C" int bl_max_divergence_noalloc(const double *U, const double *V, const double *W, int nx, int ny, int nz,
                                         const double *x, const double *y, const double *z, double *div_out,
                                         void *stream) {
    return guard([&] {
        if (!U || !V || !W || !x || !y || !z || !div_out) {
            last_error() = "bl_max_divergence_noalloc: null pointer";
            throw ArgError();
        }
        if (nx < 3 || ny < 3 || nz < 3) {
            last_error() = "bl_max_divergence_noalloc: nx, ny, nz must be at least 3";
            throw ArgError();
        }
        const long long nparts = maxdiv_partials(nx + 1, ny + 1, nz + 1);
        const long long need = nparts + (long long)nx + ny + nz;
        if (need > kScratchDoubles) {
            last_error() = "bl_max_divergence_noalloc: grid needs " + std::to_string(need) +
                           " scratch doubles, the maximum is " + std::to_string(kScratchDoubles);
            throw ArgError();
        }
        double *base = scratch_on_current_device();
        double *part = base, *rdx = base + nparts, *rdy = rdx + nx, *rdz = rdy + ny;
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_recip_faces, dim3((nx + 255) / 256), dim3(256), 0, s, x, nx, rdx);
        hipLaunchKernelGGL(k_recip_faces, dim3((ny + 255) / 256), dim3(256), 0, s, y, ny, rdy);
        hipLaunchKernelGGL(k_recip_faces, dim3((nz + 255) / 256), dim3(256), 0, s, z, nz, rdz);
        BL_HIP(hipGetLastError());
        // the reference's unpadded layouts U(nx,nyg,nzg), V(nxg,ny,nzg), W(nxg,nyg,nz)
        DivArgs a{};
        a.nxg = nx + 1;
        a.nyg = ny + 1;
        a.nzg = nz + 1;
        a.u1 = nx;
        a.u2 = (long long)nx * (ny + 1);
        a.v1 = nx + 1;
        a.v2 = (long long)(nx + 1) * ny;
        a.w1 = nx + 1;
        a.w2 = (long long)(nx + 1) * (ny + 1);
        a.rdx = rdx;
        a.rdy = rdy;
        a.rdz = rdz;
        launch_maxdiv<double>(s, a, U, V, W, part, div_out);
    });
}
