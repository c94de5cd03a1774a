// health.h -- non-finite (NaN/Inf) watchdog shared by the propagators.
//
// A device word holds kHealthy until a kernel finds a NaN or Inf in the
// wavefield; it then holds the smallest launch step that found one.  The
// step kernels check the top (receiver) row every step; a scan of the whole
// level runs at the end of a one-shot call.  The host reads the word at its
// synchronisation points and fails with SSSI_EUNSTABLE.
#pragma once

#include <hip/hip_runtime.h>

namespace sssi {

constexpr int kHealthy = 0x7f7f7f7f;  // what a byte fill of 0x7f leaves

__device__ __forceinline__ bool nonfinite(float x) { return !(fabsf(x) <= 3.402823466e38f); }

// Called in wave-uniform control flow: one atomic per wave that has a bad lane.
__device__ __forceinline__ void report_nonfinite(int *health, bool bad, int step) {
    const unsigned long long m = __ballot(bad);
    if (m && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicMin(health, step);
}

// Report a NaN/Inf anywhere in n floats of p into *health as step `tag`.
void launch_nonfinite_scan(const float *p, size_t n, int *health, int tag, hipStream_t s);
// The watchdog word after the stream's work (synchronises the stream).
int read_health(const int *health, hipStream_t s);
// Arm a watchdog word (kHealthy).
void arm_health(int *health, hipStream_t s);

}  // namespace sssi
