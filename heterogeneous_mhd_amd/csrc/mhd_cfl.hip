// mhd_cfl.hip -- calcfl: global maximum of the fast-speed bound, returned as
// 1/c, plus the driver's dt rule applied on the device.
//
// Reference: calcfl C/calcfl.c:13-63 (max_3num :65-89, max_2num :91-101);
// dt rule C/mhd.c:62-64.
//
// The reference keeps a running max_2num over the cells in memory order,
// starting from 0.  For non-NaN values that is an order-free maximum.  A NaN
// cell makes the running value NaN and the next cell then replaces it, so the
// result is the maximum of the cells after the last NaN (NaN if the last cell
// is NaN).  The reduction below reproduces exactly that: every workgroup owns
// one contiguous run of cells and records (offset of its last NaN or -1,
// max of its cells after that NaN or -inf); the last workgroup to finish
// (ticket counter, no second launch, nothing for the host to clear between
// calls) folds the records in order.  Per-cell arithmetic follows the
// reference's float/double pattern, so 1/c is bit-exact.
#include "mhd_internal.h"

namespace mhd {
namespace {

constexpr int kCB = 256;                 // threads per workgroup
constexpr int kCPer = 4;                 // cells per thread
constexpr int kRun = kCB * kCPer;        // cells per workgroup

struct CflRec {
    int last_nan;  // offset in the run of the last NaN, -1 if none
    float mx;      // max of the cells after last_nan, -inf if none
};

__device__ __forceinline__ float max3(float a, float b, float c)
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
    const float v = max3(t1, t2, t3);
    const float msq = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const float bsq = bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2];
    const float ps = (float)(((double)q[4] - (double)(msq / rho) / 2.0) * ((double)gam - 1.0)
                             + (2.0 - (double)gam) * (double)bsq / 2.0);
    const float p = (float)((double)ps - (double)bsq / 2.0);
    return (float)((double)v + sqrt(fabs(((double)bsq * 2.0 + (double)(gam * p)) / (double)rho)));
}

__device__ __forceinline__ int block_max_int(int x, int *sh)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        int o = __shfl_xor(x, off, 64);
        x = o > x ? o : x;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = x;
    __syncthreads();
    x = sh[0];
    for (int w = 1; w < kCB / 64; w++) x = sh[w] > x ? sh[w] : x;
    __syncthreads();
    return x;
}

// max of non-NaN values with -inf as the empty value
__device__ __forceinline__ float block_max_f(float x, float *sh)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float o = __shfl_xor(x, off, 64);
        x = o > x ? o : x;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = x;
    __syncthreads();
    x = sh[0];
    for (int w = 1; w < kCB / 64; w++) x = sh[w] > x ? sh[w] : x;
    __syncthreads();
    return x;
}

__global__ __launch_bounds__(kCB) void cfl_kernel(const float *__restrict__ u,
                                                  const float *__restrict__ b, size_t ncell,
                                                  CflRec *__restrict__ recs,
                                                  unsigned *__restrict__ ticket, float *inv_c,
                                                  double cfl, float *dt_dev, float *t_dev, float tf)
{
    __shared__ int shi[kCB / 64];
    __shared__ float shf[kCB / 64];
    __shared__ bool last;
    __shared__ float su[5 * kRun], sb[3 * kRun];
    const size_t base = (size_t)blockIdx.x * kRun;
    const int nrun = (int)(ncell - base < (size_t)kRun ? ncell - base : (size_t)kRun);
    // coalesced staging of the run's u and b through LDS
    for (int f = threadIdx.x; f < 5 * nrun; f += kCB) su[f] = u[base * 5 + f];
    for (int f = threadIdx.x; f < 3 * nrun; f += kCB) sb[f] = b[base * 3 + f];
    __syncthreads();
    float t[kCPer];
    int lnan = -1;
#pragma unroll
    for (int r = 0; r < kCPer; r++) {
        const int off = threadIdx.x + r * kCB;
        t[r] = -INFINITY;
        if (off < nrun) {
            t[r] = cfl_cell(su + off * 5, sb + off * 3);
            if (t[r] != t[r]) lnan = off;
        }
    }
    const int L = block_max_int(lnan, shi);
    float m = -INFINITY;
#pragma unroll
    for (int r = 0; r < kCPer; r++) {
        const int off = threadIdx.x + r * kCB;
        if (off > L && t[r] > m) m = t[r];   // cells beyond ncell hold -inf
    }
    m = block_max_f(m, shf);
    if (threadIdx.x == 0) {
        CflRec rec;
        rec.last_nan = L;
        rec.mx = m;
        recs[blockIdx.x] = rec;
        __threadfence();
        const unsigned tk = atomicAdd(ticket, 1u);
        last = (tk == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;

    // ---- the last workgroup folds the records in memory order
    __threadfence();
    const int nrec = (int)gridDim.x;
    int bstar = -1;  // last run that holds a NaN
    for (int r = threadIdx.x; r < nrec; r += kCB) {
        const int ln = __hip_atomic_load(&recs[r].last_nan, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ln >= 0) bstar = r;
    }
    bstar = block_max_int(bstar, shi);
    float mm = -INFINITY;
    for (int r = threadIdx.x; r < nrec; r += kCB) {
        if (r >= bstar) {
            const float v = __hip_atomic_load(&recs[r].mx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            mm = v > mm ? v : mm;
        }
    }
    mm = block_max_f(mm, shf);
    if (threadIdx.x == 0) {
        float c;
        if (bstar < 0) c = mm > 0.0f ? mm : 0.0f;   // max_2num from c = 0
        else c = (mm == -INFINITY) ? __builtin_nanf("") : mm;
        const float inv = 1 / c;                    // C/calcfl.c:61
        inv_c[0] = inv;
        if (dt_dev) {                               // C/mhd.c:62-64
            float dt = (float)(cfl * (double)inv);
            if (t_dev) {
                const float t0 = *t_dev;
                if (tf > 0.0f && dt > (tf - t0) / 2) dt = (tf - t0) / 2;
                *t_dev = t0 + 2 * dt;
            }
            *dt_dev = dt;
        }
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace

size_t cfl_workspace_bytes(size_t ncell)
{
    const size_t nrec = (ncell + kRun - 1) / kRun;
    return 256 + nrec * sizeof(CflRec);
}

int launch_calcfl(float *inv_c, const float *u, const float *b, int nx, int ny, int nz, void *ws,
                  double cfl, float *dt_dev, float *t_dev, float tf, hipStream_t s)
{
    if (!inv_c || !u || !b || !ws || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    const size_t ncell = (size_t)nx * ny * nz;
    const size_t nrec = (ncell + kRun - 1) / kRun;
    if (nrec > 0x7fffffff) return MHD_ERR_ARG;
    unsigned *ticket = (unsigned *)ws;
    CflRec *recs = (CflRec *)((char *)ws + 256);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL(cfl_kernel, dim3((unsigned)nrec), dim3(kCB), 0, s, u, b, ncell, recs, ticket,
                       inv_c, cfl, dt_dev, t_dev, tf);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

}  // namespace mhd
