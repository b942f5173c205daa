// mhd_cfl.hip -- calcfl: global maximum of the fast-speed bound, returned as
// 1/c, plus the driver's dt rule applied on the device.
//
// Reference: calcfl C/calcfl.c:13-63; dt rule C/mhd.c:62-64.  The per-cell
// bound and the order-aware maximum (exact NaN semantics of the reference's
// running max_2num) are described in mhd_cfl_common.h.  The producer kernel
// stages one run of 1024 cells per workgroup through LDS with 16-byte loads,
// all issued before the first LDS store, and writes one record; the fold
// kernel reduces the records in order (two levels, the second by the last
// workgroup to finish).  Per-cell arithmetic follows the reference's
// float/double pattern, so 1/c is bit-exact.
#include <mutex>

#include "mhd_cfl_common.h"

namespace mhd {
namespace {

constexpr int kCB = 256;                 // threads per workgroup
constexpr int kCPer = 4;                 // cells per thread
constexpr int kRun = kCB * kCPer;        // cells per workgroup

__global__ __launch_bounds__(kCB) void cfl_kernel(const float *__restrict__ u,
                                                  const float *__restrict__ b, size_t ncell,
                                                  void *ws)
{
    __shared__ __attribute__((aligned(16))) float su[5 * kRun], sb[3 * kRun];
    const size_t base = (size_t)blockIdx.x * kRun;
    const int nrun = (int)(ncell - base < (size_t)kRun ? ncell - base : (size_t)kRun);
    if (nrun == kRun && ((uintptr_t)u % 16 == 0) && ((uintptr_t)b % 16 == 0)) {
        constexpr int NU = 5 * kRun / 4 / kCB, NB = 3 * kRun / 4 / kCB;
        const float4 *gu = reinterpret_cast<const float4 *>(u + base * 5);
        const float4 *gb = reinterpret_cast<const float4 *>(b + base * 3);
        float4 ru[NU], rb[NB];
#pragma unroll
        for (int k = 0; k < NU; k++) ru[k] = gu[threadIdx.x + k * kCB];
#pragma unroll
        for (int k = 0; k < NB; k++) rb[k] = gb[threadIdx.x + k * kCB];
#pragma unroll
        for (int k = 0; k < NU; k++) reinterpret_cast<float4 *>(su)[threadIdx.x + k * kCB] = ru[k];
#pragma unroll
        for (int k = 0; k < NB; k++) reinterpret_cast<float4 *>(sb)[threadIdx.x + k * kCB] = rb[k];
    } else {
        for (int f = threadIdx.x; f < 5 * nrun; f += kCB) su[f] = u[base * 5 + f];
        for (int f = threadIdx.x; f < 3 * nrun; f += kCB) sb[f] = b[base * 3 + f];
    }
    __syncthreads();
    float t[kCPer];
    int off[kCPer];
    int lnan = -1;
#pragma unroll
    for (int r = 0; r < kCPer; r++) {
        off[r] = threadIdx.x + r * kCB;
        t[r] = -INFINITY;
        if (off[r] < nrun) {
            t[r] = cfl_cell(su + off[r] * 5, sb + off[r] * 3);
            if (t[r] != t[r]) lnan = off[r];
        }
    }
    cfl_block_record<kCB, kCPer>(t, off, lnan, ws, blockIdx.x);
}

__global__ __launch_bounds__(kFoldBlock) void cfl_fold_kernel(int nrec, CflOut o)
{
    __shared__ CflRec sh[kFoldBlock];
    __shared__ bool last;
    const CflRec ident = {0, -INFINITY};
    const CflRec *l1 = cfl_level1(o.ws);
    CflRec *l2 = cfl_level2(o.ws);
    unsigned *ticket = (unsigned *)o.ws;
    // level 1: workgroup g folds records [g*chunk, (g+1)*chunk), thread t a
    // contiguous slice of that, in order
    const int chunk = (nrec + (int)gridDim.x - 1) / (int)gridDim.x;
    const int per = (chunk + kFoldBlock - 1) / kFoldBlock;
    const int lo = blockIdx.x * chunk;
    const int hi = nrec < lo + chunk ? nrec : lo + chunk;
    const int a = lo + threadIdx.x * per;
    const int e = hi < a + per ? hi : a + per;
    CflRec acc = ident;
    for (int r = a; r < e; r++) acc = cfl_combine(acc, l1[r]);
    acc = ordered_block_fold(acc, sh);
    if (threadIdx.x == 0) {
        l2[blockIdx.x] = acc;
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    // level 2, by the last workgroup to finish
    __threadfence();
    CflRec r2 = ident;
    if (threadIdx.x < gridDim.x) {  // one level-2 record per workgroup of this launch
        r2.has_nan = __hip_atomic_load(&l2[threadIdx.x].has_nan, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        r2.after = __hip_atomic_load(&l2[threadIdx.x].after, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    r2 = ordered_block_fold(r2, sh);
    if (threadIdx.x == 0) {
        const float c = cfl_finish(r2);
        const float inv = 1 / c;                                         // C/calcfl.c:61
        o.inv_c[0] = inv;
        if (o.dt) {                                                      // C/mhd.c:62-64
            float dt = (float)(o.cfl * (double)inv);
            if (o.t) {
                const float t0 = *o.t;
                if (o.tf > 0.0f && dt > (o.tf - t0) / 2) dt = (o.tf - t0) / 2;
                *o.t = t0 + 2 * dt;
            }
            *o.dt = dt;
        }
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// zeroes the workspace header (the fold's ticket counter) with vector
// stores: a kernel, not a memset node, so that a captured graph replays it
__global__ void cfl_clear_kernel(void *ws)
{
    ((unsigned *)ws)[threadIdx.x] = 0u;  // 64 x 4 bytes = kCflHeader
}
static_assert(kCflHeader == 64 * sizeof(unsigned), "one thread per header word");

}  // namespace

static_assert(kFoldGroups == kFoldBlock, "level 2 folds one record per thread");

int launch_cfl_fold(int nrec, const CflOut &o, hipStream_t s)
{
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    // few records (small boxes): one workgroup, at most 16 records per
    // thread, instead of 256 workgroups and a ticket wait for the last one
    const unsigned groups = nrec <= 16 * kFoldBlock ? 1u : (unsigned)kFoldGroups;
    hipLaunchKernelGGL(cfl_fold_kernel, dim3(groups), dim3(kFoldBlock), 0, s, nrec, o);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

size_t cfl_workspace_bytes(int nx, int ny, int nz)
{
    // one level-1 record per workgroup of the widest producer: this kernel,
    // or the x-sweep fluid kernel that fuses the reduction into the last
    // sweep of a step
    const size_t ncell = (size_t)nx * ny * nz;
    size_t nrec = (ncell + kRun - 1) / kRun;
    const size_t nf = fluid_workgroups(nx, (long)ny * nz);
    if (nf > nrec) nrec = nf;
    return kCflHeader + (kFoldGroups + nrec) * sizeof(CflRec);
}

int launch_calcfl(float *inv_c, const float *u, const float *b, int nx, int ny, int nz, void *ws,
                  double cfl, float *dt_dev, float *t_dev, float tf, hipStream_t s, bool clear_ticket)
{
    if (!inv_c || !u || !b || !ws || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    const size_t ncell = (size_t)nx * ny * nz;
    const size_t nrec = (ncell + kRun - 1) / kRun;
    if (nrec > 0x7fffffff) return MHD_ERR_ARG;
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    if (clear_ticket) hipLaunchKernelGGL(cfl_clear_kernel, dim3(1), dim3(64), 0, s, ws);
    hipLaunchKernelGGL(cfl_kernel, dim3((unsigned)nrec), dim3(kCB), 0, s, u, b, ncell, ws);
    if (hipGetLastError() != hipSuccess) return MHD_ERR_HIP;
    CflOut o = {ws, inv_c, dt_dev, t_dev, cfl, tf};
    return launch_cfl_fold((int)nrec, o, s);
}

// ---- calcfl with no allocation ----------------------------------------------
// For callers that cannot allocate (code under HIP graph capture, serving
// loops): the workspace of launch_calcfl is a file-scope device array of a
// fixed size, one copy per GPU, which the library holds until the process
// ends.  In host code `scratch` names a host-side stand-in, not device
// memory: its device address comes from hipGetSymbolAddress, once per
// device with that device current, and only that address is used.

namespace {

constexpr size_t kNoAllocMaxCells = (size_t)1 << 30;  // 1024^3
constexpr size_t kScratchBytes =
    kCflHeader + (kFoldGroups + (kNoAllocMaxCells + kRun - 1) / kRun) * sizeof(CflRec);  // 8 MiB + 2304 B
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
void *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

}  // namespace

size_t noalloc_scratch_bytes() { return kScratchBytes; }
size_t noalloc_max_cells() { return kNoAllocMaxCells; }

// The current device's scratch address (nullptr and the HIP error on failure).
void *noalloc_scratch(hipError_t *err)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *err = e != hipSuccess ? e : hipErrorInvalidDevice;
        return nullptr;
    }
    std::call_once(g_scratch_once[dev], [dev] {
        g_scratch_err[dev] = hipGetSymbolAddress(&g_scratch_ptr[dev], HIP_SYMBOL(scratch));
    });
    *err = g_scratch_err[dev];
    return g_scratch_err[dev] == hipSuccess ? g_scratch_ptr[dev] : nullptr;
}

}  // namespace mhd
