// mhd_fluid_body.h -- fluidx: the relaxing-TVD fluid sweep along the first axis
// (device code shared by the single-box and the ensemble kernels).
//
// Reference: fluidx C/fluidx.c:19-102, tvd1 C/tvd1.c:8-118,
// mhdflux/find_max C/mhdflux.c:9-72.
//
// Mapping.  One pencil (all nx cells of one (j,k)) is owned by `tp` threads
// (a multiple of the 64-wide wavefront); a workgroup owns BLOCK/tp
// consecutive pencils, i.e. one contiguous run of u in HBM, which it stages
// into LDS with 16-byte loads (all issued before the first LDS store) and
// writes back the same way.  Thread t of a pencil holds cells t, t+tp, ... in
// registers.  Stencil neighbours go through two LDS arrays of 5 floats per
// cell, cell-major: a stride of 5 words is odd, so a wavefront's per-cell
// reads are free of bank conflicts.  The pencil-global freezing speed c
// (find_max, twice per sweep) is a wavefront butterfly plus, when a pencil
// spans several wavefronts, one LDS exchange.
//
// Arithmetic.  Every result is bitwise the reference's (float storage,
// double wherever a double literal or a libm double routine appears, no FMA
// contraction: the library is built with -ffp-contract=off).  The sweep is
// VALU-bound, so the reference's operations are regrouped where the rounded
// results are provably identical:
//  * abXYc(L, R, R, c, 0, 0) = c*R + 0*R + 0 (C/includefunc_spu.h:80) is
//    computed as c*R + (R - R): both give c*R with -0 mapped to +0 for finite
//    R and NaN for infinite or NaN R.
//  * dfrp(i) and dfrm(i+1) (C/tvd1.c:60-64) are the same rounded expression
//    of fr(i+1), fr(i); likewise dflm(i+1) and dflp(i) (:86-90); and
//    fl(i) = c*wl(i+1) (:83-84).  fr and fl go through LDS once per stage;
//    cell i evaluates dfrm and dflm at i and at i+1 (and the stage-1 flux at
//    i and i-1) itself from the shared values, with the same expressions,
//    instead of waiting for a further LDS round trip.
//  * x / c for the pencil-constant c (C/mhdflux.c:54) is x*r corrected by one
//    FMA residual step with r = RN(1/c); with |x| in [2^-60, 2^60] and c in
//    [2^-30, 2^30] that is the correctly rounded quotient (Markstein's
//    theorem for an exactly rounded reciprocal); zeros, tiny, huge and
//    non-finite x and any c outside that range use the IEEE division.
// The limiter 2ab/(a+b) (:78, :104) and the double square root stay IEEE.
//
// FAST (the tolerance mode) keeps the scheme and its operation order but
// computes in FP32 only: hardware reciprocal and square root (v_rcp_f32,
// v_sqrt_f32, 1 ulp), FMA, and c*(w +- v/c) evaluated as c*w +- v without the
// division.  Its results differ from the reference's by FP32 rounding noise;
// the tests state the tolerance.
//
// With a CFL workspace the kernel also computes calcfl (C/calcfl.c:13-63) of
// the state it writes, so the last sweep of a step yields the next step's dt
// without another pass over HBM.  The result is bitwise calcfl's in both
// modes; FAST screens the cells with an FP32 estimate of proven accuracy and
// evaluates the exact bound only where it can be the maximum (cfl_screen).
#pragma once
#include "mhd_cfl_common.h"

namespace mhd {
namespace {


// mhdflux for one cell (C/mhdflux.c:23-45).  Returns the cell's freezing
// speed candidate; v receives the un-normalised flux.
__device__ __forceinline__ float flux_cell(const float q[5], const float bc[3], float v[5])
{
    const float gam = (float)(5.0 / 3.0);
    const float rho = q[0], mx = q[1], my = q[2], mz = q[3], e = q[4];
    const float b1 = bc[0], b2 = bc[1], b3 = bc[2];
    const float vx = mx / rho;
    const float msq = mx * mx + my * my + mz * mz;
    const float bsq = b1 * b1 + b2 * b2 + b3 * b3;
    const float ps = (float)(((double)e - (double)(msq / rho) / 2.0) * ((double)gam - 1.0)
                             + (2.0 - (double)gam) * (double)bsq / 2.0);
    v[0] = mx;
    v[1] = mx * vx + ps - b1 * b1;
    v[2] = my * vx - b2 * b1;
    v[3] = mz * vx - b3 * b1;
    v[4] = (e + ps) * vx - b1 * (b1 * mx + b2 * my + b3 * mz) / rho;
    const float p = (float)((double)ps - (double)bsq / 2.0);
    return (float)(fabs((double)vx) + sqrt(fabs((double)((bsq + gam * p) / rho))));
}

// flux_cell in FP32 with hardware reciprocal / square root (tolerance mode).
__device__ __forceinline__ float flux_cell_fast(const float q[5], const float bc[3], float v[5])
{
    const float gam = 5.0f / 3.0f;
    const float rho = q[0], mx = q[1], my = q[2], mz = q[3], e = q[4];
    const float b1 = bc[0], b2 = bc[1], b3 = bc[2];
    const float rr = __builtin_amdgcn_rcpf(rho);
    const float vx = mx * rr;
    const float msq = fmaf(mx, mx, fmaf(my, my, mz * mz));
    const float bsq = fmaf(b1, b1, fmaf(b2, b2, b3 * b3));
    const float ps = fmaf(fmaf(-0.5f * msq, rr, e), gam - 1.0f, (1.0f - 0.5f * gam) * bsq);
    v[0] = mx;
    v[1] = fmaf(-b1, b1, fmaf(mx, vx, ps));
    v[2] = fmaf(my, vx, -b2 * b1);
    v[3] = fmaf(mz, vx, -b3 * b1);
    const float bm = fmaf(b1, mx, fmaf(b2, my, b3 * mz));
    v[4] = fmaf(e + ps, vx, -b1 * bm * rr);
    const float p = fmaf(-0.5f, bsq, ps);
    return fabsf(vx) + __builtin_amdgcn_sqrtf(fabsf(fmaf(gam, p, bsq) * rr));
}

// 2ab/(a+b) where a and b have the same sign, else 0, in FP32 with the
// hardware reciprocal, as (a|b| + |a|b) / (|a| + |b|): the numerator is 0
// when the signs differ (abs is a free source modifier), and the denominator
// is kept >= FLT_MIN so that a = b = 0 gives 0; there is no compare or select.
// Where |a| + |b| < FLT_MIN the numerator underflows to at most ~1e-38.
__device__ __forceinline__ float limiter_fast(float a, float b)
{
    const float n = fmaf(a, fabsf(b), fabsf(a) * b);
    return n * __builtin_amdgcn_rcpf(fmaxf(fabsf(a) + fabsf(b), 0x1p-126f));
}

// C/tvd1.c:75-79: 2ab/(a+b) in double where the float product a*b > 0.
__device__ __forceinline__ float limiter(float a, float b)
{
    float r = 0.0f;
    if (a * b > 0.0f) r = (float)(2.0 * (double)a * (double)b / (double)(a + b));
    return r;
}

// Global access at a byte offset from a wavefront-uniform base: the
// global_load / global_store form with a scalar base and a 32-bit offset.
template <class T>
__device__ __forceinline__ T ld_at(const float *base, unsigned off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
// Output stores are nontemporal: a sweep's output is not read again by the
// same launch, so it should not displace the neighbour pencils that later
// workgroups of the launch read from L2.
template <class T>
__device__ __forceinline__ void st_at(float *base, unsigned off, T v)
{
    __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off));
}

// c*R + 0*R + 0 of the reference, see the file header.
__device__ __forceinline__ float scale0(float c, float x) { return c * x + (x - x); }

// x / c, correctly rounded; r = 1/c, fast = c in [2^-30, 2^30] (uniform).
__device__ __forceinline__ float div_c(float x, float c, float r, bool fast)
{
    if (!fast) return x / c;
    const float q0 = x * r;
    float q = fmaf(fmaf(-q0, c, x), r, q0);
    const float ax = fabsf(x);
    if (__builtin_expect(!(ax >= 0x1p-60f && ax <= 0x1p60f), 0)) q = (x == 0.0f) ? q0 : x / c;
    return q;
}

// One DPP lane exchange of a 32-bit value (CTRL: a gfx9 DPP control code).
template <int CTRL>
__device__ __forceinline__ unsigned dpp_u32(unsigned v)
{
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

// Max over the threads of one pencil.  x >= 0 and never NaN (the callers
// start from 0 and keep the running max with a strict '>', as find_max does,
// so a NaN candidate never enters), so the bit patterns order as unsigned
// integers and the max is exact in any order: a DPP butterfly within each
// row of 16 lanes (quad swaps, half-row and row mirrors), two lane swaps
// across rows, and, when a pencil spans several wavefronts, one LDS
// exchange.  Returns after a workgroup barrier that follows every LDS read
// this thread made before the call, so the caller may then overwrite the
// stencil arrays: the exchange's own barrier when a pencil spans several
// wavefronts, else one at the end.
template <int BLOCK>
__device__ __forceinline__ float pencil_max(float xf, int tp, lds_f *red)
{
    unsigned x = __float_as_uint(xf);
    x = max(x, dpp_u32<0xB1>(x));   // quad_perm [1,0,3,2]
    x = max(x, dpp_u32<0x4E>(x));   // quad_perm [2,3,0,1]
    x = max(x, dpp_u32<0x141>(x));  // row_half_mirror
    x = max(x, dpp_u32<0x140>(x));  // row_mirror
    x = max(x, (unsigned)__shfl_xor((int)x, 16, 64));
    x = max(x, (unsigned)__shfl_xor((int)x, 32, 64));
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (tp > 64) {
        lds_u *r = (lds_u *)red;
        if (lane == 0) r[wave] = x;
        __syncthreads();
        const int wpp = tp >> 6, first = (threadIdx.x / tp) * wpp;
        x = r[first];
        for (int w = 1; w < wpp; w++) x = max(x, r[first + w]);
    } else {
        __syncthreads();
    }
    return __uint_as_float(x);
}

// Copies na floats a_src -> a_dst and nb floats b_src -> b_dst (global ->
// LDS).  With vec, 16-byte vectors: all loads of both arrays are issued
// before the first LDS store.  Loads and stores use an index clamped into
// range (a clamped lane rewrites the last vector with its own, identical
// value), so there is no branch for the compiler to sink the loads into.
template <int BLOCK, int KA, int KB>
__device__ __forceinline__ void stage_in(lds_f *a_dst, const float *a_src, int na, lds_f *b_dst,
                                         const float *b_src, int nb, bool vec)
{
    if (vec) {
        const int na4 = na >> 2, nb4 = nb >> 2;
        v4f ra[KA], rb[KB > 0 ? KB : 1];
#pragma unroll
        for (int k = 0; k < KA; k++) {
            const int f = threadIdx.x + k * BLOCK;
            ra[k] = ld_at<v4f>(a_src, 16u * (unsigned)(f < na4 ? f : na4 - 1));
        }
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const int f = threadIdx.x + k * BLOCK;
            rb[k] = ld_at<v4f>(b_src, 16u * (unsigned)(f < nb4 ? f : nb4 - 1));
        }
#pragma unroll
        for (int k = 0; k < KA; k++) {
            const int f = threadIdx.x + k * BLOCK;
            ((lds_v4 *)a_dst)[f < na4 ? f : na4 - 1] = ra[k];
        }
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const int f = threadIdx.x + k * BLOCK;
            ((lds_v4 *)b_dst)[f < nb4 ? f : nb4 - 1] = rb[k];
        }
    } else {
        for (int f = threadIdx.x; f < na; f += BLOCK) a_dst[f] = a_src[f];
        for (int f = threadIdx.x; f < nb; f += BLOCK) b_dst[f] = b_src[f];
    }
}

// FUSED: the face field is not read from bu but taken from the registers of
// a preceding advectbyzx of the same pencils (keep, see advect_group); the
// caller has synchronised the workgroup since its last LDS read.
// TPC / FIT: as for advect_group (compile-time threads per pencil and nx).
template <int BLOCK, int M, bool CFL, bool FAST, bool FUSED = false, int TPC = 0, bool FIT = false>
__device__ __forceinline__ void fluid_group(lds_f *lds, lds_f *red, float *u_out, const float *uu,
                                            const float *__restrict__ bu, int nx_, int ny, int nz,
                                            int tp_, int grp, float dt, void *cfl_ws, int rec,
                                            const float (*keep)[5] = nullptr)
{
    static_assert(!FIT || TPC > 0, "FIT needs the threads per pencil at compile time");
    const int tp = TPC > 0 ? TPC : tp_;
    const int nx = FIT ? M * TPC : nx_;
    const int pb = BLOCK / tp;                 // pencils per workgroup
    // this thread's pencil in the group: uniform over the wavefront
    // ONE: one pencil per workgroup and one workgroup per pencil (the
    // launchers' grid), so every thread's cells exist and no test is needed
    constexpr bool ONE = FIT && BLOCK == TPC;
    const int lp = ONE ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x / tp);
    const int lt = threadIdx.x - lp * tp;      // this thread's rank in its pencil
    if constexpr (FIT) __builtin_assume(lt >= 0 && lt < TPC);
    const int npen = ny * nz;
    const int pen0 = grp * pb;
    const int npb = npen - pen0 < pb ? npen - pen0 : pb;  // pencils present
    const int pen = pen0 + lp;
    const bool live = ONE || lp < npb;
    const int j = pen % ny, k = pen / ny;
    const int jp = j + 1 == ny ? 0 : j + 1, kp = k + 1 == nz ? 0 : k + 1;  // C/fluidx.c:52-53

    lds_f *A = lds;                            // 5*nx*pb floats
    lds_f *B = lds + 5 * nx * pb;              // 5*nx*pb floats
    lds_f *Ap = A + 5 * nx * lp, *Bp = B + 5 * nx * lp;

    const float z = 0.0f, h = 0.5f;
    constexpr int KU = (5 * M + 3) / 4, KB = (3 * M + 3) / 4;  // 16-byte loads per thread

    // ---- stage the group's u (5 floats/cell) and b (3/cell) into LDS; the
    // neighbour faces b2(j+1), b3(k+1) go straight to registers
    const float *gu = uu + (size_t)pen0 * nx * 5;
    const float *gb = bu + (size_t)pen0 * nx * 3;
    const int nu = 5 * nx * npb, nb = 3 * nx * npb;
    const bool vec = ((uintptr_t)gu % 16 == 0) && ((uintptr_t)gb % 16 == 0) && nu % 4 == 0 &&
                     nb % 4 == 0 && (5 * nx * pb) % 4 == 0;
    float b2jp[M], b3kp[M];
    if constexpr (FUSED) {
        stage_in<BLOCK, KU, 0>(A, gu, nu, B, gb, 0, vec);
#pragma unroll
        for (int r = 0; r < M; r++) {
            const int i = lt + r * tp;
            if (live && i < nx) {
#pragma unroll
                for (int c = 0; c < 3; c++) B[3 * nx * lp + 3 * i + c] = keep[r][c];
                b2jp[r] = keep[r][3];
                b3kp[r] = keep[r][4];
            }
        }
    } else {
        const float *bjp = bu + ((size_t)k * ny + jp) * nx * 3, *bkp = bu + ((size_t)kp * ny + j) * nx * 3;
#pragma unroll
        for (int r = 0; r < M; r++) {
            const int i = lt + r * tp;
            if (live && i < nx) {
                b2jp[r] = ld_at<float>(bjp, 12u * (unsigned)i + 4u);
                b3kp[r] = ld_at<float>(bkp, 12u * (unsigned)i + 8u);
            }
        }
        stage_in<BLOCK, KU, KB>(A, gu, nu, B, gb, nb, vec);
    }
    __syncthreads();

    float q[M][5], bc[M][3], v[M][5], bf[CFL ? M : 1][3];
    float cmax = 0.0f;
    const lds_f *Bs = B + 3 * nx * lp;         // b staging of this pencil (3 floats/cell)
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
            const int ip = i + 1 == nx ? 0 : i + 1;
#pragma unroll
            for (int c = 0; c < 5; c++) q[r][c] = Ap[5 * i + c];
            if (CFL)
                for (int c = 0; c < 3; c++) bf[r][c] = Bs[3 * i + c];
            // cell-centred B, C/fluidx.c:71-81
            const float h1 = h * Bs[3 * i], h1p = h * Bs[3 * ip];
            bc[r][0] = h1 + h1p;
            bc[r][1] = h * Bs[3 * i + 1] + b2jp[r] / 2;
            bc[r][2] = h * Bs[3 * i + 2] + b3kp[r] / 2;
            const float cf = FAST ? flux_cell_fast(q[r], bc[r], v[r]) : flux_cell(q[r], bc[r], v[r]);
            cmax = cf > cmax ? cf : cmax;
        }
    }
    const float c1 = pencil_max<BLOCK>(cmax, tp, red);

    // ---- stage 1: half-step predictor, C/tvd1.c:36-49
    // s[][] holds, in turn, fr of stage 1, the predictor u1, fr of stage 2 and
    // the final flux; q keeps the sweep's input u, the base of the update (:116).
    // Each exchange writes both wave halves, and each thread then reads every
    // neighbour value it needs in one round after one barrier, evaluating the
    // stencil terms of its neighbours (flux(i-1); dfrm, dflm at i+1) itself,
    // with the same expressions, instead of a further LDS round trip.
    float s[M][5], t[M][5];
    {
        const bool norm = c1 > 0.0f;
        const bool fast = c1 >= 0x1p-30f && c1 <= 0x1p30f;
        const float rc = norm ? 1.0f / c1 : 0.0f;
        const float g = norm ? 1.0f : 0.0f;
#pragma unroll
        for (int r = 0; r < M; r++) {
            const int i = lt + r * tp;
            if (live && i < nx) {
#pragma unroll
                for (int c = 0; c < 5; c++) {
                    if constexpr (FAST) {
                        // c*(u +- v/c) = c*u +- v; a pencil with c = 0 has zero
                        // fluxes, which the pencil-uniform factor g gives
                        const float vg = v[r][c] * g;
                        s[r][c] = fmaf(c1, q[r][c], vg);
                        t[r][c] = fmaf(c1, q[r][c], -vg);
                    } else {
                        const float vv = norm ? div_c(v[r][c], c1, rc, fast) : v[r][c];
                        const float wr = q[r][c] + vv + z;
                        const float wl = q[r][c] - vv + z;
                        s[r][c] = scale0(c1, wr);          // fr
                        t[r][c] = scale0(c1, wl);          // fl(i-1) = c*wl(i)
                    }
                    Ap[5 * i + c] = t[r][c];
                    Bp[5 * i + c] = s[r][c];
                }
            }
        }
    }
    __syncthreads();
    const float hdt = (float)(-1.0 * (double)dt / 2.0);
    // FAST keeps the stage fluxes, the limiter inputs and the final flux at
    // twice their value and halves the time step instead: power-of-two
    // scaling is exact and the limiter is homogeneous of degree one, so the
    // results are the same (outside the denormal and overflow ranges) with
    // 20 fewer multiplies per cell.
    const float hdt2 = 0.5f * hdt;
    cmax = 0.0f;
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
            const int ip = i + 1 == nx ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
#pragma unroll
            for (int c = 0; c < 5; c++) {
                // flux(i) = (fr(i) - fl(i+1)) / 2 and flux(i-1)
                const float fl1 = Ap[5 * ip + c], frm = Bp[5 * im + c];
                if constexpr (FAST) {
                    const float f0 = s[r][c] - fl1, fm = frm - t[r][c];  // twice the fluxes, see hdt
                    s[r][c] = fmaf(hdt2, f0 - fm, q[r][c]);
                } else {
                    const float f0 = h * s[r][c] + (-h) * fl1 + z;
                    const float fm = h * frm + (-h) * t[r][c] + z;
                    const float d = f0 - fm + z;
                    s[r][c] = q[r][c] + hdt * d + z;           // u1
                }
            }
            const float cf = FAST ? flux_cell_fast(s[r], bc[r], v[r]) : flux_cell(s[r], bc[r], v[r]);
            cmax = cf > cmax ? cf : cmax;
        }
    }
    const float c2 = pencil_max<BLOCK>(cmax, tp, red + BLOCK / 64);

    // ---- stage 2: corrector with the van Leer limiter, C/tvd1.c:54-116
    {
        const bool norm = c2 > 0.0f;
        const bool fast = c2 >= 0x1p-30f && c2 <= 0x1p30f;
        const float rc = norm ? 1.0f / c2 : 0.0f;
        const float g = norm ? 1.0f : 0.0f;
#pragma unroll
        for (int r = 0; r < M; r++) {
            const int i = lt + r * tp;
            if (live && i < nx) {
#pragma unroll
                for (int c = 0; c < 5; c++) {
                    if constexpr (FAST) {
                        const float u1 = s[r][c], vg = v[r][c] * g;
                        s[r][c] = fmaf(c2, u1, vg);
                        t[r][c] = fmaf(c2, u1, -vg);
                    } else {
                        const float vv = norm ? div_c(v[r][c], c2, rc, fast) : v[r][c];
                        const float wr = s[r][c] + vv + z;
                        const float wl = s[r][c] - vv + z;
                        s[r][c] = scale0(c2, wr);          // fr
                        t[r][c] = scale0(c2, wl);          // fl(i-1)
                    }
                    Ap[5 * i + c] = s[r][c];
                    Bp[5 * i + c] = t[r][c];
                }
            }
        }
    }
    __syncthreads();
    // D(i) = dfrm(i) = dfrp(i-1); E(i) = dflm(i) = dflp(i-1); t1 = fr - fl
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
            const int ip = i + 1 == nx ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
            const int ip2 = ip + 1 == nx ? 0 : ip + 1;
#pragma unroll
            for (int c = 0; c < 5; c++) {
                const float fr = s[r][c], fl = t[r][c];
                const float frm = Ap[5 * im + c], frp = Ap[5 * ip + c];
                const float flp = Bp[5 * ip + c], flp2 = Bp[5 * ip2 + c];
                if constexpr (FAST) {
                    const float D = fr - frm, Dp = frp - fr;    // twice dfrm at i, i+1
                    const float E = fl - flp, Ep = flp - flp2;  // twice dflm at i, i+1
                    const float dfr = limiter_fast(Dp, D);
                    const float dfl = limiter_fast(Ep, E);
                    s[r][c] = fmaf(0.5f, dfr - dfl, fr - flp);  // twice the flux
                } else {
                    const float D = h * fr + (-h) * frm + z, Dp = h * frp + (-h) * fr + z;
                    const float E = h * fl + (-h) * flp + z, Ep = h * flp + (-h) * flp2 + z;
                    const float t1 = fr - flp + z;
                    const float dfr = limiter(Dp, D);   // (dfrp, dfrm)
                    const float dfl = limiter(Ep, E);   // (dflp, dflm)
                    const float t2 = dfr - dfl + z;
                    s[r][c] = h * t1 + h * t2 + z;      // flux
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
#pragma unroll
            for (int c = 0; c < 5; c++) Ap[5 * i + c] = s[r][c];
        }
    }
    __syncthreads();
    const float fdt = (float)(-1.0 * (double)dt);
    const float fdt2 = 0.5f * fdt;
    float tcf[M];
    int toff[M];
    int lnan = -1;
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        tcf[r] = -INFINITY;
        toff[r] = lp * nx + i;
        if (live && i < nx) {
            const int im = i == 0 ? nx - 1 : i - 1;
            float un[5];
#pragma unroll
            for (int c = 0; c < 5; c++) {
                if constexpr (FAST) {
                    un[c] = fmaf(fdt2, s[r][c] - Ap[5 * im + c], q[r][c]);
                } else {
                    const float d = s[r][c] - Ap[5 * im + c] + z;
                    un[c] = q[r][c] + fdt * d + z;
                }
                Bp[5 * i + c] = un[c];
            }
            if (CFL) {
                if constexpr (FAST) {
                    tcf[r] = cfl_screen(un, bf[r]);  // >= 0 or +inf, see below
                } else {
                    tcf[r] = cfl_cell(un, bf[r]);
                    if (tcf[r] != tcf[r]) lnan = toff[r] > lnan ? toff[r] : lnan;
                }
            }
        }
    }
    __syncthreads();
    if constexpr (CFL && FAST) {
        // The exact maximum from the screened estimates (cfl_screen): the
        // exact cfl_cell runs only on cells within 2^-8 of the workgroup's
        // largest estimate, or on every cell when one lies outside the
        // screened domain (F = +inf; NaN cells are found there).  The new u
        // is read back from the staged output in B.
        float fm = 0.0f;
#pragma unroll
        for (int r = 0; r < M; r++) fm = tcf[r] > fm ? tcf[r] : fm;  // -inf for absent cells
        __shared__ float shs[BLOCK / 64];
        const float F = cfl_block_max_f<BLOCK>(fm, shs);
        const float T = F == INFINITY ? -INFINITY : F * kCflScreenKeep;
#pragma unroll
        for (int r = 0; r < M; r++) {
            const int i = lt + r * tp;
            if (live && i < nx && tcf[r] >= T) {
                float un[5];
#pragma unroll
                for (int c = 0; c < 5; c++) un[c] = Bp[5 * i + c];
                tcf[r] = cfl_cell(un, bf[r]);
                if (tcf[r] != tcf[r]) lnan = toff[r] > lnan ? toff[r] : lnan;
            } else {
                tcf[r] = -INFINITY;
            }
        }
    }
    {
        float *go = u_out + (size_t)pen0 * nx * 5;
        if (vec) {
            const lds_v4 *s4 = (const lds_v4 *)B;
            const int n4 = nu >> 2;
#pragma unroll
            for (int kk = 0; kk < KU; kk++) {
                const int f = threadIdx.x + kk * BLOCK;
                if (f < n4) st_at<v4f>(go, 16u * (unsigned)f, s4[f]);
            }
        } else {
            for (int f = threadIdx.x; f < nu; f += BLOCK) go[f] = B[f];
        }
    }
    if (CFL) cfl_block_record<BLOCK, M>(tcf, toff, lnan, cfl_ws, rec);
}

}  // namespace
}  // namespace mhd
