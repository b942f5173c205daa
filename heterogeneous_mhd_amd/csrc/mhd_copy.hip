// mhd_copy.hip -- copy_matrix on the device (C/copy_matrix.c:9-38: update <-
// main for the 5 u and 3 b components) as one streaming kernel for both
// arrays, and the library's stream-bandwidth probe.
//
// The two arrays form one index space of 16-byte chunks (u's, then b's).
// Each thread moves K chunks a workgroup-width apart, so every load and store
// instruction of a wavefront covers 1 KiB of contiguous memory; all K loads
// are issued before the first store.  Stores are nontemporal: the copy's
// destination is not read again soon, so it should not displace the L2.
// Measured on MI355X, 2 x 16 GiB: K = 1 6.38 TB/s, K = 2 6.03, K = 4 5.98,
// K = 8 4.49; K = 4 with ordinary stores 5.82; hipMemcpyAsync 4.94.
#include "mhd_internal.h"

namespace mhd {
namespace {

constexpr int kCopyBlock = 256;

template <int K>
__global__ __launch_bounds__(kCopyBlock) void stream_copy_kernel(v4f *__restrict__ d0, const v4f *__restrict__ s0,
                                                                 size_t n0, v4f *__restrict__ d1,
                                                                 const v4f *__restrict__ s1, size_t n1)
{
    const size_t base = (size_t)blockIdx.x * (kCopyBlock * K) + threadIdx.x;
    const size_t n = n0 + n1;
    v4f r[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const size_t i = base + (size_t)k * kCopyBlock;
        if (i < n0) r[k] = s0[i];
        else if (i < n) r[k] = s1[i - n0];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const size_t i = base + (size_t)k * kCopyBlock;
        if (i < n0) __builtin_nontemporal_store(r[k], d0 + i);
        else if (i < n) __builtin_nontemporal_store(r[k], d1 + i - n0);
    }
}

// the bytes past the last whole chunk of each array (fewer than 16 each)
__global__ void copy_tail_kernel(float *d0, const float *s0, int t0, float *d1, const float *s1, int t1)
{
    const int i = threadIdx.x;
    if (i < t0) d0[i] = s0[i];
    else if (i < t0 + t1) d1[i - t0] = s1[i - t0];
}

}  // namespace

// dst0 <- src0 (n0 floats) and dst1 <- src1 (n1 floats); 16-byte aligned
// arrays use the streaming kernel, others a runtime copy.
int stream_copy2(float *dst0, const float *src0, size_t n0, float *dst1, const float *src1, size_t n1,
                 hipStream_t s)
{
    constexpr int K = 1;
    const bool aligned = ((uintptr_t)dst0 | (uintptr_t)src0 | (uintptr_t)dst1 | (uintptr_t)src1) % 16 == 0;
    if (!aligned) {
        hipError_t e = n0 ? hipMemcpyAsync(dst0, src0, 4 * n0, hipMemcpyDeviceToDevice, s) : hipSuccess;
        if (e == hipSuccess && n1) e = hipMemcpyAsync(dst1, src1, 4 * n1, hipMemcpyDeviceToDevice, s);
        return e == hipSuccess ? MHD_OK : MHD_ERR_HIP;
    }
    const size_t c0 = n0 / 4, c1 = n1 / 4;
    const size_t per = (size_t)kCopyBlock * K;
    const size_t blocks = (c0 + c1 + per - 1) / per;
    (void)hipGetLastError();
    if (blocks > 0x7fffffffu) return MHD_ERR_ARG;
    if (blocks)
        hipLaunchKernelGGL(stream_copy_kernel<K>, dim3((unsigned)blocks), dim3(kCopyBlock), 0, s, (v4f *)dst0,
                           (const v4f *)src0, c0, (v4f *)dst1, (const v4f *)src1, c1);
    const int t0 = (int)(n0 - 4 * c0), t1 = (int)(n1 - 4 * c1);
    if (t0 + t1)
        hipLaunchKernelGGL(copy_tail_kernel, dim3(1), dim3(64), 0, s, dst0 + 4 * c0, src0 + 4 * c0, t0,
                           dst1 + 4 * c1, src1 + 4 * c1, t1);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

}  // namespace mhd
