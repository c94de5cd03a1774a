// Device-side helpers shared by the 2-D kernels: raw buffer access with a
// wave-uniform scalar offset, scalar loads of uniform tables, and the
// wave index as a scalar.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sssi {

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000);
}
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                       uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, 0);
}
// Byte offset marking a lane with no value in a buffer of fewer than 2^31
// bytes: the range check covers the vector offset, so a load there returns 0
// and a store is dropped.  Offsets of the form kOOB + k (k < 2^31) stay out
// of range.
constexpr uint32_t kOOB = 0x80000000u;

// Load of wave-uniform data that no launch of this kernel writes (source
// tables, checkpoint scales of earlier steps): through the constant address
// space the compiler emits a scalar load, which is waited for with lgkmcnt and
// leaves the in-order vmcnt of the prefetched field loads untouched.
template <class T>
__device__ __forceinline__ T ld_uniform(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}

// Index of the wavefront in its block, as a wave-uniform (scalar) value: the
// compiler cannot prove threadIdx.x >> 6 uniform, and a row offset derived
// from it would turn every buffer load into a waterfall loop.
__device__ __forceinline__ int wave_id() {
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

}  // namespace sssi
