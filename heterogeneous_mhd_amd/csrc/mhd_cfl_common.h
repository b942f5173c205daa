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

// Screen for the exact maximum (the tolerance mode's fused CFL): an FP32
// estimate f of cfl_cell with hardware reciprocal and square root and FMA,
// or +inf for a cell outside the screened domain D below.
//
// Claim: for a cell in D, |f - e| <= 2^-12 e, with e = cfl_cell(q, bb).
// Both f and e approximate the real-valued V + sqrt|X|, with
//   V = max|m_i|/rho, K = msq/rho, P = (E - K/2)(g-1) + (1-g/2)B^2 - B^2/2,
//   X = (2B^2 + g P)/rho, g = (float)(5/3) (g-1 and 1-g/2 are exact floats).
// With u = 2^-24 and W = (2/3)(|E| + K/2 + B^2) >= |P|, S = (2B^2 + g W)/rho
// (every rounding error of X is a few u times a term bounded by S):
//   fast path (v_rcp_f32 within 2u, FMA chain): |Xf - X| <= 24 u S;
//   exact path (float divisions, double sums, float stores of ps, p and
//   g*p): |Xe - X| <= 12 u S.
// D asks x = |Xf| >= S/64 (no severe cancellation in X), so
//   |sqrt|Xf| - sqrt|Xe|| <= 36 u S / sqrt|Xf| <= 36 * 64 u sqrt|Xf| (+ 2u for
//   v_sqrt_f32), and V is within 3.1u in both paths: |f - e| <= ~2310 u e < 2^-12 e.
// D also asks 2^-30 <= rho <= 2^30, every other input finite with sum of
// magnitudes <= 2^30 (no overflow, accurate reciprocal; NaN and inf fail
// these tests) and S >= 2^-90 (underflow errors negligible); then e is
// finite, never NaN.
// A workgroup takes T = F (1 - 2^-8) with F the largest f of its cells: every
// cell whose e is the workgroup's maximum has f >= E (1 - 2^-12) > T, so the
// exact cfl_cell evaluated on the cells with f >= T alone gives the exact
// maximum; a workgroup with any cell outside D (F = +inf) evaluates all its
// cells exactly (fluid_group).  tests/test_cfl_screen.py checks the bound on
// the CPU.
__device__ __forceinline__ float cfl_screen(const float *q, const float *bb)
{
    const float gam = 5.0f / 3.0f;
    const float rho = q[0];
    const float rr = __builtin_amdgcn_rcpf(rho);
    const float a1 = fabsf(q[1]), a2 = fabsf(q[2]), a3 = fabsf(q[3]), a4 = fabsf(q[4]);
    const float v = fmaxf(fmaxf(a1, a2), a3) * rr;
    const float msq = fmaf(q[1], q[1], fmaf(q[2], q[2], q[3] * q[3]));
    const float bsq = fmaf(bb[0], bb[0], fmaf(bb[1], bb[1], bb[2] * bb[2]));
    const float ps = fmaf(fmaf(-0.5f * msq, rr, q[4]), gam - 1.0f, (1.0f - 0.5f * gam) * bsq);
    const float p = fmaf(-0.5f, bsq, ps);
    const float x = fabsf(fmaf(2.0f, bsq, gam * p) * rr);
    const float w = (2.0f / 3.0f) * (fmaf(0.5f * msq, rr, a4) + bsq);
    const float S = fmaf(gam, w, 2.0f * bsq) * rr;
    const float mag = ((a1 + a2) + (a3 + a4)) + ((fabsf(bb[0]) + fabsf(bb[1])) + fabsf(bb[2]));
    const bool in = rho >= 0x1p-30f && rho <= 0x1p30f && mag <= 0x1p30f && S >= 0x1p-90f &&
                    x >= 0x1p-6f * S;
    return in ? v + __builtin_amdgcn_sqrtf(x) : INFINITY;
}
constexpr float kCflScreenKeep = 1.0f - 0x1p-8f;  // T = F * kCflScreenKeep

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
