// calib.hip -- the achievable HBM streaming rate on this device, measured with
// the access pattern of a stencil step at its minimum: three FP32 streams
// read, one written, each element touched once (d = a + b * c), 16-byte
// vector accesses, one launch over the whole array.  This is the ceiling a
// bandwidth-bound time-step kernel can approach; the propagators' roofline
// fractions are quoted against the published 8 TB/s and read beside it.
#include <vector>

#include "common.h"

namespace sssi {

__global__ void __launch_bounds__(256) stream_triad(const float4 *__restrict__ a,
                                                    const float4 *__restrict__ b,
                                                    const float4 *__restrict__ c,
                                                    float4 *__restrict__ d, int64_t n4) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n4) return;
    const float4 x = a[g], y = b[g], z = c[g];
    float4 r;
    r.x = x.x + y.x * z.x;
    r.y = x.y + y.y * z.y;
    r.z = x.z + y.z * z.z;
    r.w = x.w + y.w * z.w;
    d[g] = r;
}

__global__ void fill_ones(float4 *p, int64_t n4) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n4) p[g] = make_float4(1.f, 1.f, 1.f, 1.f);
}

}  // namespace sssi

using namespace sssi;

extern "C" int sssi_calib_stream(int64_t ncells, int reps, double *us_per_launch,
                                 double *gb_per_s) {
    try {
        if (ncells < 1024 || ncells % 4 || reps < 1 || ncells > ((int64_t)1 << 32))
            fail(SSSI_EINVAL, "calib: ncells must be a multiple of 4 in [1024, 2^32]");
        const int64_t n4 = ncells / 4;
        DevBuf<float4> a(n4), b(n4), c(n4), d(n4);
        OwnedStream s;
        s.create();
        const unsigned nb = (unsigned)((n4 + 255) / 256);
        for (float4 *p : {a.p, b.p, c.p, d.p})
            hipLaunchKernelGGL(fill_ones, dim3(nb), dim3(256), 0, s, p, n4);
        hipLaunchKernelGGL(stream_triad, dim3(nb), dim3(256), 0, s, a.p, b.p, c.p, d.p, n4);
        hipEvent_t e0, e1;
        SSSI_HIP(hipEventCreate(&e0));
        SSSI_HIP(hipEventCreate(&e1));
        SSSI_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(stream_triad, dim3(nb), dim3(256), 0, s, a.p, b.p, c.p, d.p, n4);
        SSSI_HIP(hipGetLastError());
        SSSI_HIP(hipEventRecord(e1, s));
        SSSI_HIP(hipStreamSynchronize(s));
        float ms = 0;
        SSSI_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const double us = ms * 1e3 / reps;
        if (us_per_launch) *us_per_launch = us;
        if (gb_per_s) *gb_per_s = 16.0 * (double)ncells / (us * 1e-6) / 1e9;
        set_last_error("");
        return SSSI_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return SSSI_EDEVICE;
    }
}
