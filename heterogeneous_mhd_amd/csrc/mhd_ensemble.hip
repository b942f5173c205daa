// mhd_ensemble.hip -- batched ensembles of independent periodic boxes of
// different sizes, advanced together: one persistent launch per sweep covers
// every box, each box has its own CFL dt, and the whole step can be captured
// into one HIP graph.  No counterpart in the reference (one box per run).
//
// Boxes are cubic (side n), so every stage of the split step
// (C/mhd.c:59-124) has dims (n, n, n) for each box and the ensemble walks the
// same stage sequence as the single-box plan.  All boxes' u (resp. b) are
// packed into one buffer at 16-byte aligned offsets, with a second pair of
// buffers for the ping-pong.  A stage's work list is the concatenation of
// every box's work items (pencil groups for the sweeps, 32x32 tiles for the
// transposes), boxes ordered by decreasing size; a fixed grid of workgroups
// strides through the list, finds the item's box by binary search over a
// prefix table and calls the very device functions the single-box kernels
// use, so every box evolves bitwise as it would alone.  calcfl is fused into
// each box's last fluid sweep and folded per box (one workgroup per box,
// records in memory order).
//
// The workgroup shape (threads, cells per thread, threads per pencil) and the
// persistent grid size are chosen by mhd_ensemble_tune, which times whole
// steps of the actual ensemble for each candidate and keeps the fastest; the
// state is saved before and restored after, so tuning does not move it.
#include "../../include/mhd_hip.h"
#include "mhd_advect_body.h"
#include "mhd_advect_transpose_body.h"
#include "mhd_fluid_body.h"
#include "mhd_transpose_advect_body.h"
#include "mhd_transpose_body.h"

#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

namespace mhd {
namespace {

struct EnsDev {
    int nbox;
    const int *n;            // side of each box
    const long long *uoff;   // float offset of each box in the u buffers
    const long long *boff;   // float offset of each box in the b buffers
    const int *gpre;         // prefix sums of sweep pencil groups per box (nbox+1)
    const int *tpre;         // prefix sums of transpose tiles per box (nbox+1)
    const int *lpre;         // prefix sums of advectbyzx+transpose tiles per box (nbox+1)
    float *t, *dt, *dtn, *inv;
};

// largest b with pre[b] <= item (pre[0] = 0)
__device__ __forceinline__ int find_box(const int *pre, int nbox, int item)
{
    int lo = 0, hi = nbox;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int BLOCK, int M, int TP, bool CFL, bool FAST, bool FIT = false>
__global__ __launch_bounds__(BLOCK) void ens_fluid_kernel(EnsDev E, float *Uout, const float *Uin,
                                                          const float *Bf, void *ws, int first,
                                                          int end)
{
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ float red[2 * (BLOCK / 64)];
    for (int item = first + blockIdx.x; item < end; item += gridDim.x) {
        const int box = find_box(E.gpre, E.nbox, item);
        const int n = E.n[box];
        __syncthreads();  // the previous item's last LDS reads are done
        // record index = item: each box's records are contiguous, in memory order
        fluid_group<BLOCK, M, CFL, FAST, false, TP, FIT>((lds_f *)lds, (lds_f *)red, Uout + E.uoff[box], Uin + E.uoff[box],
                                                    Bf + E.boff[box], n, n, n, TP,
                                                    item - E.gpre[box], E.dt[box], ws, item);
    }
}

template <int BLOCK, int M, int TP, bool FAST, bool FIT = false>
__global__ __launch_bounds__(BLOCK) void ens_advect_kernel(EnsDev E, float *Bout, const float *Bmain,
                                                           const float *U, const float *Bin, int first,
                                                           int end)
{
    extern __shared__ float lds[];
    for (int item = first + blockIdx.x; item < end; item += gridDim.x) {
        const int box = find_box(E.gpre, E.nbox, item);
        const int n = E.n[box];
        const long long bo = E.boff[box];
        __syncthreads();
        float keep[M][5];
        advect_group<BLOCK, M, FAST, TP, FIT>((lds_f *)lds, Bout + bo, Bmain + bo, U + E.uoff[box], Bin + bo, n, n,
                                         n, TP, item - E.gpre[box], E.dt[box], keep);
    }
}

// advectbyzx + fluidx of one sweep in one pass (see mhd_fused.hip); u goes
// to Uout, which must not be Uin
template <int BLOCK, int M, int TP, bool CFL, bool FAST, bool FIT = false>
__global__ __launch_bounds__(BLOCK) void ens_fused_kernel(EnsDev E, float *Uout, float *Bout,
                                                          const float *Uin, const float *Bin, void *ws,
                                                          int first, int end)
{
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ float red[2 * (BLOCK / 64)];
    for (int item = first + blockIdx.x; item < end; item += gridDim.x) {
        const int box = find_box(E.gpre, E.nbox, item);
        const int n = E.n[box];
        const long long bo = E.boff[box], uo = E.uoff[box];
        const int grp = item - E.gpre[box];
        __syncthreads();  // the previous item's last LDS reads are done
        float keep[M][5];
        advect_group<BLOCK, M, FAST, TP, FIT>((lds_f *)lds, Bout + bo, Bin + bo, Uin + uo, Bin + bo, n, n, n, TP,
                                         grp, E.dt[box], keep);
        __syncthreads();
        fluid_group<BLOCK, M, CFL, FAST, true, TP, FIT>((lds_f *)lds, (lds_f *)red, Uout + uo, Uin + uo, Bin + bo, n, n, n, TP,
                                                   grp, E.dt[box], ws, item, keep);
    }
}

// t13 = false: transpose12 (a = y), true: transpose13 (a = z); cubic boxes
template <int NC>
__global__ __launch_bounds__(kTB) void ens_transpose_kernel(EnsDev E, float *out, const float *in,
                                                           int t13, Perm perm)
{
    __shared__ __attribute__((aligned(16))) float tile[kT * lds_row<NC>()];
    const long long *off = NC == 5 ? E.uoff : E.boff;
    const int total = E.tpre[E.nbox];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        const int box = find_box(E.tpre, E.nbox, item);
        const int n = E.n[box];
        const int nt = (n + kT - 1) / kT;
        const int local = item - E.tpre[box];
        const int ti = local % nt, ta = (local / nt) % nt;
        const size_t zz = (size_t)(local / (nt * nt));
        const size_t n1 = (size_t)n, n2 = (size_t)n * n;
        const size_t sa_in = t13 ? n2 : n1, sz_in = t13 ? n1 : n2;
        const size_t si_out = t13 ? n2 : n1, sz_out = t13 ? n1 : n2;
        const int i0 = ti * kT, a0 = ta * kT;
        const int ni = n - i0 < kT ? n - i0 : kT, na = n - a0 < kT ? n - a0 : kT;
        float *o = out + off[box];
        const float *s = in + off[box];
        __syncthreads();
        if (n % 4 == 0 && ni == kT && na == kT)
            tile_fast<NC>(o, s, tile, i0, a0, zz, sa_in, sz_in, si_out, sz_out, perm);
        else
            tile_scalar<NC>(o, s, tile, i0, a0, ni, na, zz, sa_in, sz_in, si_out, sz_out, perm);
    }
}

// The tiles of the fused advectbyzx+transpose (first half) and
// transpose+advectbyzx (second half) kernels, for boxes whose side is a
// multiple of kTP: 56 cells along the sweep axis x 16 pencils x a run of at
// most 64 steps of the remaining axis.  Each run starts its walk from the
// inputs alone, so how a box is cut into runs does not change a bit of its
// result: every box evolves as it would alone.
__host__ __device__ inline int ens_tile_qlen(int n) { return n < 64 ? n : 64; }
__host__ __device__ inline int ens_tiles(int n)
{
    return (n + at_tile::kTI - 1) / at_tile::kTI * (n / at_tile::kTP) * ((n + ens_tile_qlen(n) - 1) / ens_tile_qlen(n));
}

template <bool FAST, int WHICH>
__global__ __launch_bounds__(512) void ens_advect_transpose_kernel(EnsDev E, float *Uo, float *Bo, const float *Ui,
                                                                   const float *Bi)
{
    const int total = E.lpre[E.nbox];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        const int box = find_box(E.lpre, E.nbox, item);
        const int n = E.n[box];
        at_tile::AtGeom g;
        g.n1 = g.n2 = g.n3 = g.np = g.nq = n;
        g.nti = (n + at_tile::kTI - 1) / at_tile::kTI;
        g.npb = n / at_tile::kTP;
        g.qlen = ens_tile_qlen(n);
        g.nqc = (n + g.qlen - 1) / g.qlen;
        g.nwg = 0;
        const long long uo = E.uoff[box], bo = E.boff[box];
        __syncthreads();  // the previous item's last LDS reads are done
        at_tile::advect_transpose_tile<512, FAST, WHICH>(Uo + uo, Bo + bo, Ui + uo, Bi + bo, g,
                                                         (unsigned)(item - E.lpre[box]), E.dt[box]);
    }
}

template <bool FAST, int WHICH>
__global__ __launch_bounds__(512) void ens_transpose_advect_kernel(EnsDev E, float *Uo, float *Bo, const float *Ui,
                                                                   const float *Bi)
{
    const int total = E.lpre[E.nbox];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        const int box = find_box(E.lpre, E.nbox, item);
        const int n = E.n[box];
        ta_tile::TaGeom g;
        g.ns = g.np = g.nq = n;
        g.nti = (n + ta_tile::kTI - 1) / ta_tile::kTI;
        g.npb = n / ta_tile::kTP;
        g.qlen = ens_tile_qlen(n);
        g.nqc = (n + g.qlen - 1) / g.qlen;
        g.nwg = 0;
        const long long uo = E.uoff[box], bo = E.boff[box];
        __syncthreads();  // the previous item's last LDS reads are done
        ta_tile::transpose_advect_tile<512, FAST, WHICH>(Uo + uo, Bo + bo, Ui + uo, Bi + bo, g,
                                                         (unsigned)(item - E.lpre[box]), E.dt[box]);
    }
}

// one workgroup per box: ordered fold of its records -> 1/c and the next dt
__global__ __launch_bounds__(kFoldBlock) void ens_cfl_fold_kernel(EnsDev E, const void *ws, double cfl)
{
    __shared__ CflRec sh[kFoldBlock];
    const int box = blockIdx.x;
    const CflRec *l1 = cfl_level1(const_cast<void *>(ws));
    const int lo = E.gpre[box], hi = E.gpre[box + 1];
    const int per = (hi - lo + kFoldBlock - 1) / kFoldBlock;
    const int a = lo + threadIdx.x * per;
    const int e = hi < a + per ? hi : a + per;
    CflRec acc = {0, -INFINITY};
    for (int r = a; r < e; r++) acc = cfl_combine(acc, l1[r]);
    acc = ordered_block_fold(acc, sh);
    if (threadIdx.x == 0) {
        const float inv = 1 / cfl_finish(acc);     // C/calcfl.c:61
        E.inv[box] = inv;
        E.dtn[box] = (float)(cfl * (double)inv);   // C/mhd.c:62
    }
}

// dt = dtn, clamped to (tf - t)/2 when tf > 0; t += 2 dt   (C/mhd.c:63-64)
__global__ void ens_begin_kernel(EnsDev E, float tf)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= E.nbox) return;
    float d = E.dtn[b];
    const float t0 = E.t[b];
    if (tf > 0.0f && d > (tf - t0) / 2) d = (tf - t0) / 2;
    E.t[b] = t0 + 2 * d;
    E.dt[b] = d;
}

// ---- launch shapes -------------------------------------------------------

struct Shape {
    int block, m, tp;
};
// Candidate sweep shapes; valid for an ensemble when its largest side <= m*tp.
constexpr Shape kShapes[] = {{256, 2, 64}, {128, 2, 64}, {256, 1, 128}, {512, 2, 64},
                             {256, 2, 128}, {512, 2, 256}, {1024, 2, 512}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);
constexpr int kGridMult[] = {2, 4, 8, 16};  // persistent workgroups per CU
constexpr int kNumGrids = sizeof(kGridMult) / sizeof(kGridMult[0]);

template <int BLOCK, int M, int TP, bool CFL, bool FAST, bool FIT>
int launch_fluid_k(const EnsDev &E, float *Uo, const float *Ui, const float *B, size_t lds, void *ws,
                   int grid, hipStream_t s, int first, int end)
{
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)ens_fluid_kernel<BLOCK, M, TP, CFL, FAST, FIT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();
    hipLaunchKernelGGL((ens_fluid_kernel<BLOCK, M, TP, CFL, FAST, FIT>), dim3(grid), dim3(BLOCK), lds, s, E,
                       Uo, Ui, B, ws, first, end);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

template <int BLOCK, int M, int TP, bool FAST, bool FIT>
int launch_advect_k(const EnsDev &E, float *Bo, const float *Bm, const float *U, const float *Bi,
                    size_t lds, int grid, hipStream_t s, int first, int end)
{
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)ens_advect_kernel<BLOCK, M, TP, FAST, FIT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();
    hipLaunchKernelGGL((ens_advect_kernel<BLOCK, M, TP, FAST, FIT>), dim3(grid), dim3(BLOCK), lds, s, E, Bo,
                       Bm, U, Bi, first, end);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

template <int BLOCK, int M, int TP, bool CFL, bool FAST, bool FIT>
int launch_fused_k(const EnsDev &E, float *Uo, float *Bo, const float *Ui, const float *Bi, size_t lds,
                   void *ws, int grid, hipStream_t s, int first, int end)
{
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)ens_fused_kernel<BLOCK, M, TP, CFL, FAST, FIT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();
    hipLaunchKernelGGL((ens_fused_kernel<BLOCK, M, TP, CFL, FAST, FIT>), dim3(grid), dim3(BLOCK), lds, s, E,
                       Uo, Bo, Ui, Bi, ws, first, end);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

// One sweep kernel of one launch shape.  kind 0: fluid (Uo <- Ui, B0);
// 1: advect (B0 -> Bo with b_main Bm, u U); 2: fused (Uo, Bo <- Ui, Bi).
template <int BLOCK, int M, int TP, bool FIT = false>
int launch_sweep(int kind, bool cfl, bool fast, const EnsDev &E, float *Uo, float *Bo, const float *Ui,
                 const float *Bi, const float *Bm, int nmax, void *ws, int grid, hipStream_t s, int first,
                 int end)
{
    constexpr int pb = BLOCK / TP;
    if (kind == 1) {
        const size_t lds = sizeof(float) * 2 * kLines * (size_t)nmax * pb;
        return fast ? launch_advect_k<BLOCK, M, TP, true, FIT>(E, Bo, Bm, Ui, Bi, lds, grid, s, first, end)
                    : launch_advect_k<BLOCK, M, TP, false, FIT>(E, Bo, Bm, Ui, Bi, lds, grid, s, first, end);
    }
    if (kind == 0) {
        const size_t lds = sizeof(float) * 10 * (size_t)nmax * pb;
        if (cfl)
            return fast ? launch_fluid_k<BLOCK, M, TP, true, true, FIT>(E, Uo, Ui, Bi, lds, ws, grid, s, first, end)
                        : launch_fluid_k<BLOCK, M, TP, true, false, FIT>(E, Uo, Ui, Bi, lds, ws, grid, s, first, end);
        return fast ? launch_fluid_k<BLOCK, M, TP, false, true, FIT>(E, Uo, Ui, Bi, lds, ws, grid, s, first, end)
                    : launch_fluid_k<BLOCK, M, TP, false, false, FIT>(E, Uo, Ui, Bi, lds, ws, grid, s, first, end);
    }
    const size_t lds = sizeof(float) * 10 * (size_t)nmax * pb;  // >= the advect's 2*kLines
    if (cfl)
        return fast ? launch_fused_k<BLOCK, M, TP, true, true, FIT>(E, Uo, Bo, Ui, Bi, lds, ws, grid, s, first, end)
                    : launch_fused_k<BLOCK, M, TP, true, false, FIT>(E, Uo, Bo, Ui, Bi, lds, ws, grid, s, first, end);
    return fast ? launch_fused_k<BLOCK, M, TP, false, true, FIT>(E, Uo, Bo, Ui, Bi, lds, ws, grid, s, first, end)
                : launch_fused_k<BLOCK, M, TP, false, false, FIT>(E, Uo, Bo, Ui, Bi, lds, ws, grid, s, first, end);
}

// The items of boxes that fit one cell per thread (n <= TP) form the tail of
// the work list (boxes are packed by decreasing size); for a two-cell shape
// they run as a second launch of the one-cell shape with the same block and
// threads per pencil, so the items, groups and CFL records are the same and
// no thread of a small box's pencil is idle.  split = the first such item
// (total: none); nsmall = their largest side.
// The head of the list, boxes of side exactly M*TP (items [0, nfit)), runs
// the compile-time pencil-length form (immediate LDS offsets, no bounds tests).
template <int BLOCK, int M, int TP>
int sweep_split(int kind, bool cfl, bool fast, const EnsDev &E, float *Uo, float *Bo, const float *Ui,
                const float *Bi, const float *Bm, int nmax, int nsmall, int nfit, int split, int nfit1,
                int total, void *ws, int grid, hipStream_t s)
{
    int rc = MHD_OK;
    int from = 0;
    if (nfit > 0) {
        rc = launch_sweep<BLOCK, M, TP, true>(kind, cfl, fast, E, Uo, Bo, Ui, Bi, Bm, M * TP, ws, grid, s,
                                              0, nfit);
        from = nfit;
    }
    if (rc != MHD_OK || from >= total) return rc;
    if (M == 1 || split >= total)
        return launch_sweep<BLOCK, M, TP>(kind, cfl, fast, E, Uo, Bo, Ui, Bi, Bm, nmax, ws, grid, s, from,
                                          total);
    if (split > from)
        rc = launch_sweep<BLOCK, M, TP>(kind, cfl, fast, E, Uo, Bo, Ui, Bi, Bm, nmax, ws, grid, s, from,
                                        split);
    from = split;
    if (rc == MHD_OK && nfit1 > from) {  // small boxes of side exactly TP: the one-cell fixed form
        rc = launch_sweep<BLOCK, 1, TP, true>(kind, cfl, fast, E, Uo, Bo, Ui, Bi, Bm, TP, ws, grid, s, from,
                                              nfit1);
        from = nfit1;
    }
    if (rc == MHD_OK && from < total)
        rc = launch_sweep<BLOCK, 1, TP>(kind, cfl, fast, E, Uo, Bo, Ui, Bi, Bm, nsmall, ws, grid, s, from,
                                        total);
    return rc;
}

int sweep(int shape, int kind, bool cfl, bool fast, const EnsDev &E, float *Uo, float *Bo,
          const float *Ui, const float *Bi, const float *Bm, int nmax, int nsmall, int nfit, int split,
          int nfit1, int total, void *ws, int grid, hipStream_t s)
{
#define SW(B, M_, T) sweep_split<B, M_, T>(kind, cfl, fast, E, Uo, Bo, Ui, Bi, Bm, nmax, nsmall, nfit, split, nfit1, total, ws, grid, s)
    switch (shape) {  // the entries of kShapes
    case 0: return SW(256, 2, 64);
    case 1: return SW(128, 2, 64);
    case 2: return SW(256, 1, 128);
    case 3: return SW(512, 2, 64);
    case 4: return SW(256, 2, 128);
    case 5: return SW(512, 2, 256);
    default: return SW(1024, 2, 512);
    }
#undef SW
}

const int kPU12[5] = {0, 2, 1, 3, 4}, kPB12[3] = {1, 0, 2};
const int kPU13[5] = {0, 3, 2, 1, 4}, kPB13[3] = {2, 1, 0};

Perm make_perm(const int *p, int nc)
{
    Perm q = {};
    for (int c = 0; c < nc; c++) {
        q.p[c] = p[c];
        q.inv[p[c]] = c;
    }
    return q;
}

int launch_transpose(const EnsDev &E, float *Uo, float *Bo, const float *Ui, const float *Bi, bool t13,
                     int grid, hipStream_t s)
{
    const Perm pu = make_perm(t13 ? kPU13 : kPU12, 5), pb = make_perm(t13 ? kPB13 : kPB12, 3);
    (void)hipGetLastError();
    hipLaunchKernelGGL((ens_transpose_kernel<5>), dim3(grid), dim3(kTB), 0, s, E, Uo, Ui, (int)t13, pu);
    hipLaunchKernelGGL((ens_transpose_kernel<3>), dim3(grid), dim3(kTB), 0, s, E, Bo, Bi, (int)t13, pb);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

// the fused tile kernels: first = true for advectbyzx+transpose (first
// half), false for transpose+advectbyzx (second half); grid: persistent
// workgroups (two fit per CU)
int launch_tiles(const EnsDev &E, bool first, bool fast, int which, float *Uo, float *Bo, const float *Ui,
                 const float *Bi, int grid, hipStream_t s)
{
    (void)hipGetLastError();
    auto kernel = [&]() -> const void * {
        if (first && which == 12)
            return fast ? (const void *)ens_advect_transpose_kernel<true, 12> : (const void *)ens_advect_transpose_kernel<false, 12>;
        if (first)
            return fast ? (const void *)ens_advect_transpose_kernel<true, 13> : (const void *)ens_advect_transpose_kernel<false, 13>;
        if (which == 12)
            return fast ? (const void *)ens_transpose_advect_kernel<true, 12> : (const void *)ens_transpose_advect_kernel<false, 12>;
        return fast ? (const void *)ens_transpose_advect_kernel<true, 13> : (const void *)ens_transpose_advect_kernel<false, 13>;
    };
    void *args[] = {(void *)&E, (void *)&Uo, (void *)&Bo, (void *)&Ui, (void *)&Bi};
    if (hipLaunchKernel(kernel(), dim3(grid), dim3(512), args, 0, s) != hipSuccess) return MHD_ERR_HIP;
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

}  // namespace
}  // namespace mhd

using namespace mhd;

struct mhd_ensemble {
    int nbox, nmax, ncu;
    double cfl;
    float tf;
    hipStream_t s;
    std::vector<int> n;                 // per box, in the caller's order
    std::vector<int> order;             // pack order (decreasing size) -> caller index
    std::vector<long long> uoff, boff;  // per box (caller's order)
    size_t ufloats, bfloats;
    int shape, grid;                    // current launch shape and persistent grid size
    int split, nsmall, nfit, nfit1, total;  // see sweep_split (set with the shape)
    bool fast;                          // FP32 tolerance mode of the sweeps
    bool tiles;                         // every side a multiple of 16: the fused tile schedule
    int tile_items;                     // work items of one tile launch
    // device tables
    void *tab;
    EnsDev E;
    void *ws;                           // CFL records
    float *U0, *B0, *U1, *B1;           // bound state buffers
    bool dt_ready;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

namespace {

int ens_error(hipError_t e) { return e == hipSuccess ? MHD_OK : MHD_ERR_HIP; }

void drop_graph(mhd_ensemble *e)
{
    if (e->exec) (void)hipStreamSynchronize(e->s);  // the graph may still be running
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

int ens_step(mhd_ensemble *e)
{
    const EnsDev &E = e->E;
    const hipStream_t s = e->s;
    float *U0 = e->U0, *B0 = e->B0, *U1 = e->U1, *B1 = e->B1;
    const int sh = e->shape, g = e->grid, nm = e->nmax;
    const int ns = e->nsmall, nf = e->nfit, sp = e->split, nf1 = e->nfit1, tot = e->total;
    const bool f = e->fast;
    int rc = MHD_OK;
#define RUN(x)                      \
    do {                            \
        if (rc == MHD_OK) rc = (x); \
    } while (0)
    (void)hipGetLastError();
    hipLaunchKernelGGL(ens_begin_kernel, dim3((e->nbox + 63) / 64), dim3(64), 0, s, E, e->tf);  // :63-64
    RUN(ens_error(hipGetLastError()));
    if (e->tiles) {
        // the order of C/mhd.c:71-117 with every transpose fused with the
        // advectbyzx beside it, as the single-box plan runs it; the step ends
        // in (U0, B0)
        const int tg = e->tile_items < 2 * e->ncu ? e->tile_items : 2 * e->ncu;
        RUN(sweep(sh, 0, false, f, E, U1, nullptr, U0, B0, nullptr, nm, ns, nf, sp, nf1, tot, nullptr, g, s));  // :71
        RUN(launch_tiles(E, true, f, 12, U0, B1, U1, B0, tg, s));                                          // :74, :78
        RUN(sweep(sh, 0, false, f, E, U0, nullptr, U0, B1, nullptr, nm, ns, nf, sp, nf1, tot, nullptr, g, s));  // :80
        RUN(launch_tiles(E, true, f, 13, U1, B0, U0, B1, tg, s));                                          // :84, :88
        RUN(sweep(sh, 0, false, f, E, U1, nullptr, U1, B0, nullptr, nm, ns, nf, sp, nf1, tot, nullptr, g, s));  // :91
        RUN(sweep(sh, 1, false, f, E, nullptr, B1, U1, B0, B0, nm, ns, nf, sp, nf1, tot, nullptr, g, s));       // :93
        RUN(sweep(sh, 2, false, f, E, U0, B0, U1, B1, nullptr, nm, ns, nf, sp, nf1, tot, nullptr, g, s));       // :95-97
        RUN(launch_tiles(E, false, f, 13, U1, B1, U0, B0, tg, s));                                         // :102, :105
        RUN(sweep(sh, 0, false, f, E, U1, nullptr, U1, B1, nullptr, nm, ns, nf, sp, nf1, tot, nullptr, g, s));  // :107
        RUN(launch_tiles(E, false, f, 12, U0, B0, U1, B1, tg, s));                                         // :112, :115
        RUN(sweep(sh, 0, true, f, E, U0, nullptr, U0, B0, nullptr, nm, ns, nf, sp, nf1, tot, e->ws, g, s));    // :117 + next calcfl
        if (rc == MHD_OK) {
            hipLaunchKernelGGL(ens_cfl_fold_kernel, dim3(e->nbox), dim3(kFoldBlock), 0, s, E, e->ws, e->cfl);
            rc = ens_error(hipGetLastError());
        }
        return rc;
    }
    // the order of C/mhd.c:71-117; the second half's advectbyzx + fluidx pairs
    // are fused, and u ping-pongs so that the step ends in (U0, B0)
    RUN(sweep(sh, 0, false, f, E, U1, nullptr, U0, B0, nullptr, nm, ns, nf, sp, nf1, tot, nullptr, g, s));   // :71
    RUN(sweep(sh, 1, false, f, E, nullptr, B1, U1, B0, B0, nm, ns, nf, sp, nf1, tot, nullptr, g, s));        // :74
    RUN(launch_transpose(E, U0, B0, U1, B1, false, g, s));                             // :78
    RUN(sweep(sh, 0, false, f, E, U0, nullptr, U0, B0, nullptr, nm, ns, nf, sp, nf1, tot, nullptr, g, s));   // :80
    RUN(sweep(sh, 1, false, f, E, nullptr, B1, U0, B0, B0, nm, ns, nf, sp, nf1, tot, nullptr, g, s));        // :84
    RUN(launch_transpose(E, U1, B0, U0, B1, true, g, s));                              // :88
    RUN(sweep(sh, 0, false, f, E, U1, nullptr, U1, B0, nullptr, nm, ns, nf, sp, nf1, tot, nullptr, g, s));   // :91
    RUN(sweep(sh, 1, false, f, E, nullptr, B1, U1, B0, B0, nm, ns, nf, sp, nf1, tot, nullptr, g, s));        // :93
    RUN(sweep(sh, 2, false, f, E, U0, B0, U1, B1, nullptr, nm, ns, nf, sp, nf1, tot, nullptr, g, s));        // :95-97
    RUN(launch_transpose(E, U1, B1, U0, B0, true, g, s));                              // :102
    RUN(sweep(sh, 2, false, f, E, U0, B0, U1, B1, nullptr, nm, ns, nf, sp, nf1, tot, nullptr, g, s));        // :105-107
    RUN(launch_transpose(E, U1, B1, U0, B0, false, g, s));                             // :112
    RUN(sweep(sh, 2, true, f, E, U0, B0, U1, B1, nullptr, nm, ns, nf, sp, nf1, tot, e->ws, g, s));           // :115-117 + next calcfl
    if (rc == MHD_OK) {
        hipLaunchKernelGGL(ens_cfl_fold_kernel, dim3(e->nbox), dim3(kFoldBlock), 0, s, E, e->ws, e->cfl);
        rc = ens_error(hipGetLastError());
    }
#undef RUN
    return rc;
}

// calcfl of every box's current state (first step of a run)
int ens_initial_cfl(mhd_ensemble *e)
{
    for (int q = 0; q < e->nbox; q++) {  // device scalars are in pack order
        const int b = e->order[q];
        const int n = e->n[b];
        // the stand-alone kernel: its own workgroups and records
        const int rc = launch_calcfl(e->E.inv + q, e->U0 + e->uoff[b], e->B0 + e->boff[b], n, n, n,
                                     e->ws, e->cfl, e->E.dtn + q, nullptr, 0.0f, e->s);
        if (rc != MHD_OK) return rc;
    }
    e->dt_ready = true;
    return MHD_OK;
}

}  // namespace

extern "C" {

int mhd_ensemble_create(mhd_ensemble **out, int nbox, const int *sides, double cfl, float tf,
                        mhd_stream_t stream)
{
    if (!out || nbox < 1 || !sides) return MHD_ERR_ARG;
    mhd_ensemble *e = new mhd_ensemble();
    e->nbox = nbox;
    e->cfl = cfl;
    e->tf = tf;
    e->s = (hipStream_t)stream;
    e->n.assign(sides, sides + nbox);
    e->nmax = *std::max_element(e->n.begin(), e->n.end());
    if (*std::min_element(e->n.begin(), e->n.end()) < 1) {
        delete e;
        return MHD_ERR_ARG;
    }
    if (e->nmax > 1024) {
        delete e;
        return MHD_ERR_TOO_LONG;
    }
    // pack in decreasing size: large boxes first in every work list
    e->order.resize(nbox);
    std::iota(e->order.begin(), e->order.end(), 0);
    std::stable_sort(e->order.begin(), e->order.end(), [&](int a, int b) { return e->n[a] > e->n[b]; });
    e->uoff.assign(nbox, 0);
    e->boff.assign(nbox, 0);
    long long uo = 0, bo = 0;
    for (int q = 0; q < nbox; q++) {
        const int b = e->order[q];
        const long long c = (long long)e->n[b] * e->n[b] * e->n[b];
        e->uoff[b] = uo;
        e->boff[b] = bo;
        uo += (5 * c + 3) / 4 * 4;  // keep every box 16-byte aligned
        bo += (3 * c + 3) / 4 * 4;
    }
    e->ufloats = (size_t)uo;
    e->bfloats = (size_t)bo;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        delete e;
        return MHD_ERR_HIP;
    }
    e->ncu = prop.multiProcessorCount;
    e->shape = -1;
    for (int c = 0; c < kNumShapes; c++)
        if (e->nmax <= kShapes[c].m * kShapes[c].tp) {
            e->shape = c;
            break;
        }
    e->grid = 4 * e->ncu;
    // device tables, in pack order for the work lists
    const size_t tb = sizeof(int) * nbox + 2 * sizeof(long long) * nbox + 3 * sizeof(int) * (nbox + 1) +
                      4 * sizeof(float) * nbox + 256;
    if (hipMalloc(&e->tab, tb) != hipSuccess) {
        delete e;
        return MHD_ERR_NOMEM;
    }
    char *p = (char *)e->tab;
    long long *d_uoff = (long long *)p;
    long long *d_boff = d_uoff + nbox;
    int *d_n = (int *)(d_boff + nbox);
    int *d_gpre = d_n + nbox;
    int *d_tpre = d_gpre + nbox + 1;
    int *d_lpre = d_tpre + nbox + 1;
    float *d_f = (float *)(d_lpre + nbox + 1);
    std::vector<long long> hu(nbox), hb(nbox);
    std::vector<int> hn(nbox), ht(nbox + 1, 0), hl(nbox + 1, 0);
    e->tiles = true;
    for (int q = 0; q < nbox; q++) {
        const int b = e->order[q];
        hu[q] = e->uoff[b];
        hb[q] = e->boff[b];
        hn[q] = e->n[b];
        const int nt = (hn[q] + 31) / 32;
        ht[q + 1] = ht[q] + nt * nt * hn[q];
        if (hn[q] % at_tile::kTP) e->tiles = false;
        hl[q + 1] = hl[q] + (hn[q] % at_tile::kTP ? 0 : ens_tiles(hn[q]));
    }
    hipError_t er = hipMemcpy(d_uoff, hu.data(), sizeof(long long) * nbox, hipMemcpyHostToDevice);
    if (er == hipSuccess) er = hipMemcpy(d_boff, hb.data(), sizeof(long long) * nbox, hipMemcpyHostToDevice);
    if (er == hipSuccess) er = hipMemcpy(d_n, hn.data(), sizeof(int) * nbox, hipMemcpyHostToDevice);
    if (er == hipSuccess) er = hipMemcpy(d_tpre, ht.data(), sizeof(int) * (nbox + 1), hipMemcpyHostToDevice);
    if (er == hipSuccess) er = hipMemcpy(d_lpre, hl.data(), sizeof(int) * (nbox + 1), hipMemcpyHostToDevice);
    e->tile_items = hl[nbox];
    if (er == hipSuccess) er = hipMemset(d_f, 0, 4 * sizeof(float) * nbox);
    e->E.nbox = nbox;
    e->E.n = d_n;
    e->E.uoff = d_uoff;
    e->E.boff = d_boff;
    e->E.gpre = d_gpre;
    e->E.tpre = d_tpre;
    e->E.lpre = d_lpre;
    e->E.t = d_f;
    e->E.dt = d_f + nbox;
    e->E.dtn = d_f + 2 * nbox;
    e->E.inv = d_f + 3 * nbox;
    // CFL records: the largest user is either the fused sweep (one record per
    // pencil group, smallest groups = 1 pencil) or the stand-alone calcfl of
    // the largest box
    size_t nrec = 0;
    for (int b = 0; b < nbox; b++) nrec += (size_t)e->n[b] * e->n[b];
    const size_t big = cfl_workspace_bytes(e->nmax, e->nmax, e->nmax);
    size_t wsb = kCflHeader + (kFoldGroups + nrec) * sizeof(CflRec);
    if (big > wsb) wsb = big;
    if (er == hipSuccess) er = hipMalloc(&e->ws, wsb);
    if (er == hipSuccess) er = hipMemset(e->ws, 0, wsb);
    if (er == hipSuccess) er = hipDeviceSynchronize();
    if (er != hipSuccess || e->shape < 0) {
        mhd_ensemble_destroy(e);
        return er != hipSuccess ? MHD_ERR_HIP : MHD_ERR_TOO_LONG;
    }
    *out = e;
    return mhd_ensemble_set_config(e, e->shape, e->grid);
}

void mhd_ensemble_destroy(mhd_ensemble *e)
{
    if (!e) return;
    (void)hipStreamSynchronize(e->s);  // queued steps still use the tables and workspace
    drop_graph(e);
    if (e->tab) (void)hipFree(e->tab);
    if (e->ws) (void)hipFree(e->ws);
    delete e;
}

int mhd_ensemble_sizes(mhd_ensemble *e, size_t *u_floats, size_t *b_floats, long long *u_off,
                       long long *b_off)
{
    if (!e) return MHD_ERR_ARG;
    if (u_floats) *u_floats = e->ufloats;
    if (b_floats) *b_floats = e->bfloats;
    if (u_off) memcpy(u_off, e->uoff.data(), sizeof(long long) * e->nbox);
    if (b_off) memcpy(b_off, e->boff.data(), sizeof(long long) * e->nbox);
    return MHD_OK;
}

int mhd_ensemble_num_configs(void) { return kNumShapes * kNumGrids; }

int mhd_ensemble_set_config(mhd_ensemble *e, int shape, int grid)
{
    if (!e || shape < 0 || shape >= kNumShapes || grid < 1) return MHD_ERR_ARG;
    const Shape &c = kShapes[shape];
    if (e->nmax > c.m * c.tp) return MHD_ERR_TOO_LONG;
    const int pb = c.block / c.tp;
    std::vector<int> hg(e->nbox + 1, 0);
    for (int q = 0; q < e->nbox; q++) {
        const int n = e->n[e->order[q]];
        hg[q + 1] = hg[q] + (n * n + pb - 1) / pb;
    }
    hipError_t er = hipMemcpy((void *)e->E.gpre, hg.data(), sizeof(int) * (e->nbox + 1), hipMemcpyHostToDevice);
    if (er != hipSuccess) return MHD_ERR_HIP;
    // boxes are packed by decreasing size: the ones with n <= tp are a tail
    e->total = hg[e->nbox];
    e->split = e->total;
    e->nsmall = 0;
    e->nfit = 0;  // the head boxes of side exactly m*tp
    for (int q = 0; q < e->nbox && e->n[e->order[q]] == c.m * c.tp; q++) e->nfit = hg[q + 1];
    for (int q = 0; q < e->nbox; q++) {
        const int n = e->n[e->order[q]];
        if (n <= c.tp) {
            e->split = hg[q];
            e->nsmall = n;
            break;
        }
    }
    e->nfit1 = e->split;  // then the small boxes of side exactly tp
    for (int q = 0; q < e->nbox; q++)
        if (hg[q] >= e->split && e->n[e->order[q]] == c.tp) e->nfit1 = hg[q + 1];
    e->shape = shape;
    e->grid = grid;
    drop_graph(e);
    return MHD_OK;
}

int mhd_ensemble_set_precision(mhd_ensemble *e, int mode)
{
    if (!e || (mode != MHD_PREC_EXACT && mode != MHD_PREC_FAST)) return MHD_ERR_ARG;
    if (e->fast != (mode == MHD_PREC_FAST)) {
        drop_graph(e);
        e->dt_ready = false;  // the next run starts from calcfl of the state, as the reference does
    }
    e->fast = mode == MHD_PREC_FAST;
    return MHD_OK;
}

int mhd_ensemble_get_config(mhd_ensemble *e, int *shape, int *grid, int *block, int *cells_per_thread,
                            int *threads_per_pencil)
{
    if (!e) return MHD_ERR_ARG;
    const Shape &c = kShapes[e->shape];
    if (shape) *shape = e->shape;
    if (grid) *grid = e->grid;
    if (block) *block = c.block;
    if (cells_per_thread) *cells_per_thread = c.m;
    if (threads_per_pencil) *threads_per_pencil = c.tp;
    return MHD_OK;
}

int mhd_ensemble_bind(mhd_ensemble *e, float *u, float *b, float *u2, float *b2)
{
    if (!e || !u || !b || !u2 || !b2) return MHD_ERR_ARG;
    for (const void *p : {(const void *)u, (const void *)b, (const void *)u2, (const void *)b2})
        if ((uintptr_t)p % 16) return MHD_ERR_ARG;
    if (u != e->U0 || b != e->B0 || u2 != e->U1 || b2 != e->B1) {
        e->U0 = u, e->B0 = b, e->U1 = u2, e->B1 = b2;
        e->dt_ready = false;
        drop_graph(e);
    }
    return MHD_OK;
}

int mhd_ensemble_init_box(mhd_ensemble *e, int box, uint64_t seed, double amp_v, double amp_a,
                          double rho0, double p0, double gamma, int kmax, double *scratch_dev)
{
    if (!e || box < 0 || box >= e->nbox || !e->U0 || !scratch_dev) return MHD_ERR_ARG;
    const int n = e->n[box];
    const int rc = ic_generate_device(e->U0 + e->uoff[box], e->B0 + e->boff[box], n, n, n, seed, amp_v,
                                      amp_a, rho0, p0, gamma, kmax, scratch_dev, e->s);
    e->dt_ready = false;
    return rc;
}

int mhd_ensemble_reset_time(mhd_ensemble *e)
{
    if (!e) return MHD_ERR_ARG;
    hipError_t er = hipMemsetAsync(e->E.t, 0, sizeof(float) * e->nbox, e->s);
    if (er == hipSuccess) er = hipStreamSynchronize(e->s);
    e->dt_ready = false;
    return ens_error(er);
}

int mhd_ensemble_run(mhd_ensemble *e, int nsteps, int use_graph)
{
    if (!e || nsteps < 0 || !e->U0) return MHD_ERR_ARG;
    if (!e->dt_ready) {
        const int rc = ens_initial_cfl(e);
        if (rc != MHD_OK) return rc;
    }
    if (!use_graph) {
        for (int i = 0; i < nsteps; i++) {
            const int rc = ens_step(e);
            if (rc != MHD_OK) return rc;
        }
        return MHD_OK;
    }
    if (!e->exec) {
        hipError_t er = hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal);
        if (er != hipSuccess) return MHD_ERR_HIP;
        const int rc = ens_step(e);
        er = hipStreamEndCapture(e->s, &e->graph);
        if (rc != MHD_OK) return rc;
        if (er != hipSuccess) return MHD_ERR_HIP;
        er = hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0);
        if (er != hipSuccess) return MHD_ERR_HIP;
    }
    for (int i = 0; i < nsteps; i++)
        if (hipGraphLaunch(e->exec, e->s) != hipSuccess) return MHD_ERR_HIP;
    return MHD_OK;
}

int mhd_ensemble_get_times(mhd_ensemble *e, float *t, float *dt)
{
    if (!e) return MHD_ERR_ARG;
    std::vector<float> h(2 * e->nbox);
    hipError_t er = hipMemcpyAsync(h.data(), e->E.t, sizeof(float) * 2 * e->nbox, hipMemcpyDeviceToHost, e->s);
    if (er == hipSuccess) er = hipStreamSynchronize(e->s);
    if (er != hipSuccess) return MHD_ERR_HIP;
    // device tables are in pack order
    for (int q = 0; q < e->nbox; q++) {
        const int b = e->order[q];
        if (t) t[b] = h[q];
        if (dt) dt[b] = h[e->nbox + q];
    }
    return MHD_OK;
}

int mhd_ensemble_tune(mhd_ensemble *e, int reps, float *best_ms)
{
    if (!e || reps < 1 || !e->U0) return MHD_ERR_ARG;
    // save the state (and scalars) so that tuning leaves it untouched
    void *save = nullptr;
    const size_t ub = e->ufloats * sizeof(float), bb = e->bfloats * sizeof(float);
    const size_t sb = 4 * sizeof(float) * e->nbox;
    if (hipMalloc(&save, ub + bb + sb) != hipSuccess) return MHD_ERR_NOMEM;
    char *sp = (char *)save;
    const bool was_ready = e->dt_ready;
    if (!was_ready && ens_initial_cfl(e) != MHD_OK) {
        (void)hipFree(save);
        return MHD_ERR_HIP;
    }
    hipError_t er = hipMemcpyAsync(sp, e->U0, ub, hipMemcpyDeviceToDevice, e->s);
    if (er == hipSuccess) er = hipMemcpyAsync(sp + ub, e->B0, bb, hipMemcpyDeviceToDevice, e->s);
    if (er == hipSuccess) er = hipMemcpyAsync(sp + ub + bb, e->E.t, sb, hipMemcpyDeviceToDevice, e->s);
    hipEvent_t ev0, ev1;
    if (er == hipSuccess) er = hipEventCreate(&ev0);
    if (er == hipSuccess) er = hipEventCreate(&ev1);
    float best = 1e30f;
    int bs = e->shape, bg = e->grid;
    for (int c = 0; c < kNumShapes && er == hipSuccess; c++) {
        if (e->nmax > kShapes[c].m * kShapes[c].tp) continue;
        for (int gi = 0; gi < kNumGrids && er == hipSuccess; gi++) {
            const int grid = kGridMult[gi] * e->ncu;
            if (mhd_ensemble_set_config(e, c, grid) != MHD_OK) continue;
            if (ens_step(e) != MHD_OK) continue;  // warm-up
            er = hipEventRecord(ev0, e->s);
            int rc = MHD_OK;
            for (int r = 0; r < reps && rc == MHD_OK; r++) rc = ens_step(e);
            if (er == hipSuccess) er = hipEventRecord(ev1, e->s);
            if (er == hipSuccess) er = hipEventSynchronize(ev1);
            float ms = 0;
            if (er == hipSuccess && rc == MHD_OK && hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess &&
                ms / reps < best) {
                best = ms / reps;
                bs = c;
                bg = grid;
            }
        }
    }
    // restore
    if (er == hipSuccess) er = hipMemcpyAsync(e->U0, sp, ub, hipMemcpyDeviceToDevice, e->s);
    if (er == hipSuccess) er = hipMemcpyAsync(e->B0, sp + ub, bb, hipMemcpyDeviceToDevice, e->s);
    if (er == hipSuccess) er = hipMemcpyAsync(e->E.t, sp + ub + bb, sb, hipMemcpyDeviceToDevice, e->s);
    if (er == hipSuccess) er = hipStreamSynchronize(e->s);
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    (void)hipFree(save);
    if (er != hipSuccess) return MHD_ERR_HIP;
    if (best_ms) *best_ms = best;
    e->dt_ready = true;  // the restored scalars include the state's calcfl
    return mhd_ensemble_set_config(e, bs, bg);
}

}  // extern "C"
