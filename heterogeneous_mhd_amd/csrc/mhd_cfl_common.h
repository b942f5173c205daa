// mhd_cfl_common.h -- the CFL bound of one cell and the order-aware global
// maximum shared by the stand-alone calcfl kernel and the fluid sweep that
// fuses it (device code, internal to libmhd_hip.so).
//
// Reference: calcfl C/calcfl.c:13-63 (max_3num :65-89, max_2num :91-101).
// The reference keeps a running max_2num over the cells in memory order,
// starting from 0.  For non-NaN values that is an order-free maximum.  A NaN
// cell makes the running value NaN and the next cell then replaces it, so the
// result is the maximum of the cells after the last NaN (NaN if the last cell
// is NaN).
//
// A run of consecutive cells is summarised by (has_nan, after): whether it
// holds a NaN, and the maximum of its cells after the last NaN (of all its
// cells when it has none; -inf when empty).  Two adjacent runs A, B combine as
//   B.has_nan ? B : (A.has_nan, max(A.after, B.after)),
// which is associative, so the summary of the whole array is an ordered tree
// reduction of per-workgroup records: producers (the calcfl kernel, or the
// fused fluid sweep) write one record per workgroup, cfl_fold_kernel folds
// them in two levels and writes 1/c and dt.
#pragma once
#include "mhd_internal.h"

namespace mhd {

struct CflRec {
    int has_nan;
    float after;
};

// Workspace: [0, kCflHeader) ticket counter; then kFoldGroups level-2
// records; then one level-1 record per producer workgroup.
constexpr size_t kCflHeader = 256;
constexpr int kFoldGroups = 256;
constexpr int kFoldBlock = 256;

__host__ __device__ inline CflRec *cfl_level2(void *ws) { return (CflRec *)((char *)ws + kCflHeader); }
__host__ __device__ inline CflRec *cfl_level1(void *ws) { return cfl_level2(ws) + kFoldGroups; }

__device__ __forceinline__ CflRec cfl_combine(CflRec a, CflRec b)
{
    if (b.has_nan) return b;
    CflRec r;
    r.has_nan = a.has_nan;
    r.after = b.after > a.after ? b.after : a.after;
    return r;
}

__device__ __forceinline__ float cfl_max3(float a, float b, float c)
{
    if (a > b) return a > c ? a : c;
    return b > c ? b : c;
}

// Per-cell bound, C/calcfl.c:45-54 (face B, not averaged).
__device__ __forceinline__ float cfl_cell(const float *q, const float *bb)
{
    const float gam = (float)(5.0 / 3.0);
    const float rho = q[0];
    const float t1 = fabsf(q[1] / rho), t2 = fabsf(q[2] / rho), t3 = fabsf(q[3] / rho);
    const float v = cfl_max3(t1, t2, t3);
    const float msq = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const float bsq = bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2];
    const float ps = (float)(((double)q[4] - (double)(msq / rho) / 2.0) * ((double)gam - 1.0)
                             + (2.0 - (double)gam) * (double)bsq / 2.0);
    const float p = (float)((double)ps - (double)bsq / 2.0);
    return (float)((double)v + sqrt(fabs(((double)bsq * 2.0 + (double)(gam * p)) / (double)rho)));
}

// cfl_cell in FP32 only (the tolerance mode's fused CFL): hardware
// reciprocal and square root, FMA; max_i |m_i/rho| as max_i |m_i| * |1/rho|.
// NaN inputs still give NaN (through msq, rho or the energy term).
__device__ __forceinline__ float cfl_cell_fast(const float *q, const float *bb)
{
    const float gam = 5.0f / 3.0f;
    const float rr = __builtin_amdgcn_rcpf(q[0]);
    const float v = fmaxf(fmaxf(fabsf(q[1]), fabsf(q[2])), fabsf(q[3])) * fabsf(rr);
    const float msq = fmaf(q[1], q[1], fmaf(q[2], q[2], q[3] * q[3]));
    const float bsq = fmaf(bb[0], bb[0], fmaf(bb[1], bb[1], bb[2] * bb[2]));
    const float ps = fmaf(fmaf(-0.5f * msq, rr, q[4]), gam - 1.0f, (1.0f - 0.5f * gam) * bsq);
    const float p = fmaf(-0.5f, bsq, ps);
    return v + __builtin_amdgcn_sqrtf(fabsf(fmaf(2.0f, bsq, gam * p) * rr));
}

template <int BLOCK>
__device__ __forceinline__ int cfl_block_max_int(int x, int *sh)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        int o = __shfl_xor(x, off, 64);
        x = o > x ? o : x;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = x;
    __syncthreads();
    x = sh[0];
    for (int w = 1; w < BLOCK / 64; w++) x = sh[w] > x ? sh[w] : x;
    __syncthreads();
    return x;
}

// max of non-NaN values with -inf as the empty value
template <int BLOCK>
__device__ __forceinline__ float cfl_block_max_f(float x, float *sh)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float o = __shfl_xor(x, off, 64);
        x = o > x ? o : x;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = x;
    __syncthreads();
    x = sh[0];
    for (int w = 1; w < BLOCK / 64; w++) x = sh[w] > x ? sh[w] : x;
    __syncthreads();
    return x;
}

// Called by every thread of a producer workgroup: tval[] are this thread's
// cell values with their offsets toff[] in the workgroup's run (cells that
// do not exist hold -inf), lnan the offset of its last NaN cell (-1 if none).
// Writes the workgroup's record at index rec, the position of its run in
// memory order.  Uses ~48 bytes of static LDS.
template <int BLOCK, int K>
__device__ __forceinline__ void cfl_block_record(const float (&tval)[K], const int (&toff)[K],
                                                 int lnan, void *ws, int rec)
{
    __shared__ int shi[BLOCK / 64];
    __shared__ float shf[BLOCK / 64];
    const int L = cfl_block_max_int<BLOCK>(lnan, shi);
    float m = -INFINITY;
#pragma unroll
    for (int r = 0; r < K; r++)
        if (toff[r] > L && tval[r] > m) m = tval[r];
    m = cfl_block_max_f<BLOCK>(m, shf);
    if (threadIdx.x == 0) {
        CflRec r;
        r.has_nan = L >= 0;
        r.after = m;
        cfl_level1(ws)[rec] = r;
    }
}

// c of C/calcfl.c from the summary of all cells (running max_2num from 0).
__device__ __forceinline__ float cfl_finish(CflRec r)
{
    if (!r.has_nan) return r.after > 0.0f ? r.after : 0.0f;
    return r.after == -INFINITY ? __builtin_nanf("") : r.after;  // NaN was the last cell
}

struct CflOut {
    void *ws;         // workspace, see above
    float *inv_c;     // 1/c
    float *dt;        // (float)(cfl * 1/c), or null
    float *t;         // when non-null: clamp dt to (tf - t)/2 if tf > 0, t += 2 dt
    double cfl;
    float tf;
};

// Ordered tree over the workgroup: thread t's summary precedes thread t+s's.
// (kFoldBlock threads; sh holds kFoldBlock records.)
__device__ __forceinline__ CflRec ordered_block_fold(CflRec acc, CflRec *sh)
{
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 1; s < kFoldBlock; s <<= 1) {
        if ((threadIdx.x & (2 * s - 1)) == 0) sh[threadIdx.x] = cfl_combine(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    return sh[0];
}

// Folds the nrec level-1 records a producer launch wrote and writes the
// result (defined in mhd_cfl.hip).
int launch_cfl_fold(int nrec, const CflOut &o, hipStream_t s);

}  // namespace mhd
