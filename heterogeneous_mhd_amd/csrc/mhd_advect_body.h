// mhd_advect_body.h -- advectbyzx: constrained-transport advection of the two
// transverse face fields along the sweep axis (device code shared by the
// single-box and the ensemble kernels).
//
// Reference: advectbyzx C/advectbyzx.c:14-23 (A1, A2, B1, B2 in order),
// advectbyzxA1 C/advectbyzxA1.c:18-63, A2 C/advectbyzxA2.c:9-28,
// B1 C/advectbyzxB1.c:19-62, B2 C/advectbyzxB2.c:9-28, tvdb C/tvdb.c:10-91.
//
// The reference runs four full passes and scatters the edge flux of pencil
// (j,k) into b1 of pencil (j-1,k) (A2) and (j,k-1) (B2) through a global tmp
// array, so that per cell
//   b1 <- (((b1 - FA(j,k)) + FA(j+1,k)) - FB(j,k)) + FB(j,k+1)     (at i-1)
// Here one workgroup owns pencil (j,k) and evaluates the four tvdb lines it
// needs -- A at (j,k) and (j+1,k), B at (j,k) and (j,k+1) -- itself, so
// there is no tmp array, no scatter, no atomics and a single pass; b1 is
// written once in the reference's addition order.  The output b must not
// alias update_b (neighbour pencils read update_b while this one writes b);
// b_main is read only at the cell being written, so it may alias b.
//
// FAST (the tolerance mode): the same scheme in FP32 only, with the hardware
// reciprocal (v_rcp_f32) for the face velocity and the limiter and FMA.
#pragma once
#include "mhd_fluid_body.h"  // limiter_fast

namespace mhd {
namespace {

constexpr int kLines = 4;  // A(j), A(j+1), B(k), B(k+1)

// keep[r] receives, for the thread's r-th cell, the new (b1, b2, b3) of this
// pencil and the new b2 of pencil (j+1, k) and b3 of pencil (j, k+1) -- the
// lines A(j+1) and B(k+1) this workgroup evaluates are the ones those
// pencils' own b2 and b3 come from, computed in the same order -- so a fused
// fluid sweep can follow without reading b back (see mhd_fused.hip).
//
// TPC > 0 fixes the threads per pencil at compile time; FIT (with TPC) also
// fixes nx = M*TPC, so that LDS offsets are immediates and there is no
// bounds test.  The pencil index is uniform over a wavefront (tp is a
// multiple of 64), so it is taken from the first lane and the per-pencil
// address arithmetic runs on the scalar unit.
//
// NPASS = 2: two advectbyzx calls with the same u (C/mhd.c:93 and :95, the
// end of the first half and the start of the second).  The face velocities
// depend on u only, so they are computed once; the second pass takes its
// fields from the first pass's keep (the new b of this pencil, b2 of (j+1,k)
// and b3 of (j,k+1): every value the second call reads), so the
// intermediate field never goes to HBM.  Only the second pass writes b_out;
// the results are bitwise those of the two calls.
template <int BLOCK, int M, bool FAST, int TPC = 0, bool FIT = false, int NPASS = 1>
__device__ __forceinline__ void advect_group(lds_f *lds, float *b_out, const float *b_main,
                                             const float *__restrict__ uu,
                                             const float *__restrict__ bu, int nx_, int ny, int nz,
                                             int tp_, int grp, float dt, float (&keep)[M][5])
{
    static_assert(!FIT || TPC > 0, "FIT needs the threads per pencil at compile time");
    const int tp = TPC > 0 ? TPC : tp_;
    const int nx = FIT ? M * TPC : nx_;
    const int pb = BLOCK / tp;
    constexpr bool ONE = FIT && BLOCK == TPC;  // one pencil per workgroup, as in fluid_group
    const int lp = ONE ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x / tp);
    const int lt = threadIdx.x - lp * tp;
    if constexpr (FIT) __builtin_assume(lt >= 0 && lt < TPC);
    const int npen = ny * nz;
    const int pen = grp * pb + lp;
    const bool live = ONE || pen < npen;
    const int pc = live ? pen : 0;             // in-range pencil for the clamped loads
    const int j = pc % ny, k = pc / ny;
    const int jm = j == 0 ? ny - 1 : j - 1, jp = j + 1 == ny ? 0 : j + 1;
    const int km = k == 0 ? nz - 1 : k - 1, kp = k + 1 == nz ? 0 : k + 1;
    const float z = 0.0f, h = 0.5f;

    // two stencil arrays of kLines*nx floats per pencil, line-major
    lds_f *X = lds + lp * 2 * kLines * nx;
    lds_f *Y = X + kLines * nx;

    // pencil (row) offsets in cells: jm, c, jp, km, kp
    const size_t row = (size_t)nx;
    const size_t op[5] = {((size_t)k * ny + jm) * row, ((size_t)k * ny + j) * row,
                          ((size_t)k * ny + jp) * row, ((size_t)km * ny + j) * row,
                          ((size_t)kp * ny + j) * row};
    // line t: face velocity from pencils (lo, hi); its field comes from pencil hi
    constexpr int lo[kLines] = {0, 1, 3, 1}, hi[kLines] = {1, 2, 1, 4};
    constexpr int comp[kLines] = {1, 1, 2, 2};

    // ---- every global load first (indices clamped into range so that the
    // loads are unconditional and all in flight together), then the raw face
    // velocity vx = (m1(lo)+m1(hi))/(rho(lo)+rho(hi)), A1 :34 / B1 :34
    float rho[M][5], m1[M][5];
    float bl[M][kLines];  // the advected face field, per line
    float b1m[M];         // main b1 of this pencil
    // pencil bases are uniform over the wavefront and the in-pencil byte
    // offset is shared by all of a cell's loads: uniform base + 32-bit
    // offset, the global_load form with a scalar base
    const float *ub[5], *bb5[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        ub[q] = uu + op[q] * 5;
        bb5[q] = bu + op[q] * 3;
    }
    const float *bm = b_main + op[1] * 3;
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        const unsigned ic = (unsigned)(i < nx ? i : nx - 1);
        const unsigned ou = ic * 20u, ob = ic * 12u;  // bytes into a pencil of u, b
#pragma unroll
        for (int q = 0; q < 5; q++) {
            rho[r][q] = ld_at<float>(ub[q], ou);
            m1[r][q] = ld_at<float>(ub[q], ou + 4u);
        }
#pragma unroll
        for (int t = 0; t < kLines; t++) bl[r][t] = ld_at<float>(bb5[hi[t]], ob + 4u * comp[t]);
        b1m[r] = ld_at<float>(bm, ob);
    }
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
#pragma unroll
            for (int t = 0; t < kLines; t++)
                X[t * nx + i] = FAST ? (m1[r][lo[t]] + m1[r][hi[t]]) *
                                           __builtin_amdgcn_rcpf(rho[r][lo[t]] + rho[r][hi[t]])
                                     : (m1[r][lo[t]] + m1[r][hi[t]]) / (rho[r][lo[t]] + rho[r][hi[t]]);
        }
    }
    __syncthreads();
    // ---- smoothing vx = (vx(i-1)+vx(i+1)+2vx)/4 in double, A1 :37-41
    float vg[M][kLines];
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
            const int im = i == 0 ? nx - 1 : i - 1, ip = i + 1 == nx ? 0 : i + 1;
#pragma unroll
            for (int t = 0; t < kLines; t++) {
                const float s = X[t * nx + im] + X[t * nx + ip];
                vg[r][t] = FAST ? 0.25f * fmaf(2.0f, X[t * nx + i], s)
                                : (float)(((double)s + 2.0 * (double)X[t * nx + i]) / 4.0);
                Y[t * nx + i] = vg[r][t];
            }
        }
    }
    __syncthreads();
    // ---- tvdb, C/tvdb.c:26-46: vh, and b*vg staged for the upwind pick
    float vh[M][kLines];
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
            const int ip = i + 1 == nx ? 0 : i + 1;
#pragma unroll
            for (int t = 0; t < kLines; t++) {
                vh[r][t] = FAST ? 0.5f * (vg[r][t] + Y[t * nx + ip]) : h * vg[r][t] + h * Y[t * nx + ip] + z;
                X[t * nx + i] = bl[r][t] * vg[r][t];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int pass = 0; pass < NPASS; pass++) {
        if (pass > 0) {
            // the second call: fields from the first one's results, same vg
            __syncthreads();  // the first pass's last reads of Y are done
#pragma unroll
            for (int r = 0; r < M; r++) {
                const int i = lt + r * tp;
                if (live && i < nx) {
                    bl[r][0] = keep[r][1];  // b2 of this pencil
                    bl[r][1] = keep[r][3];  // b2 of (j+1, k)
                    bl[r][2] = keep[r][2];  // b3 of this pencil
                    bl[r][3] = keep[r][4];  // b3 of (j, k+1)
                    b1m[r] = keep[r][0];
#pragma unroll
                    for (int t = 0; t < kLines; t++) {
                        Y[t * nx + i] = vg[r][t];
                        X[t * nx + i] = bl[r][t] * vg[r][t];
                    }
                }
            }
            __syncthreads();
        }
        // ---- half step b1 = b - (flux1 - flux1(i-1))*dt/2, C/tvdb.c:48-52, with
        // the upwind flux1 = vh > 0 ? b*vg : (b*vg)(i+1) of C/tvdb.c:44-46 formed
        // in registers for i and i-1 (vh(i-1) by the same expression its owner
        // uses), so no third LDS array is needed; X is rewritten after a barrier
        float bh[M][kLines];
#pragma unroll
        for (int r = 0; r < M; r++) {
            const int i = lt + r * tp;
            if (live && i < nx) {
                const int ip = i + 1 == nx ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
#pragma unroll
                for (int t = 0; t < kLines; t++) {
                    const float vhm = FAST ? 0.5f * (Y[t * nx + im] + Y[t * nx + i])
                                           : h * Y[t * nx + im] + h * Y[t * nx + i] + z;
                    const float f1 = vh[r][t] > 0.0f ? X[t * nx + i] : X[t * nx + ip];
                    const float f1m = vhm > 0.0f ? X[t * nx + im] : X[t * nx + i];
                    if constexpr (FAST) {
                        bh[r][t] = fmaf(-0.5f * dt, f1 - f1m, bl[r][t]);
                    } else {
                        const float d = (f1 - f1m) * dt;
                        bh[r][t] = (float)((double)bl[r][t] - (double)d / 2.0);
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
                for (int t = 0; t < kLines; t++) X[t * nx + i] = bh[r][t];
            }
        }
        __syncthreads();
        // ---- limited flux, C/tvdb.c:54-78
        float fx[M][kLines];
#pragma unroll
        for (int r = 0; r < M; r++) {
            const int i = lt + r * tp;
            if (live && i < nx) {
                const int ip = i + 1 == nx ? 0 : i + 1, ipp = ip + 1 == nx ? 0 : ip + 1;
                const int im = i == 0 ? nx - 1 : i - 1;
#pragma unroll
                for (int t = 0; t < kLines; t++) {
                    const lds_f *bb = X + t * nx, *vv = Y + t * nx;
                    float w;
                    if constexpr (FAST) {
                        // the upwind pair {wm, wp} is {(w1-w0)/2, (w-wa)/2} for vh > 0 and
                        // {(w0-w1)/2, (w-wa)/2} otherwise; the limiter is symmetric and
                        // homogeneous, so dw = limiter(+-(w1-w0), w-wa)/2
                        const bool up = vh[r][t] > 0.0f;
                        const float w0 = vv[i] * bb[i], w1 = vv[ip] * bb[ip];
                        const int ia = up ? im : ipp;
                        const float wa = vv[ia] * bb[ia];
                        w = up ? w0 : w1;
                        const float d1 = w1 - w0;
                        fx[r][t] = fmaf(0.5f, limiter_fast(up ? d1 : -d1, w - wa), w) * dt;
                    } else {
                        float wp, wm;
                        if (vh[r][t] > 0.0f) {
                            w = vv[i] * bb[i];
                            wp = (float)((double)(vv[ip] * bb[ip] - w) / 2.0);
                            wm = (float)((double)(w - vv[im] * bb[im]) / 2.0);
                        } else {
                            w = vv[ip] * bb[ip];
                            wp = (float)((double)(w - vv[ipp] * bb[ipp]) / 2.0);
                            wm = (float)((double)(vv[i] * bb[i] - w) / 2.0);
                        }
                        float dw = 0.0f;
                        if (wm * wp > 0.0f) dw = (float)(2.0 * (double)wm * (double)wp / (double)(wm + wp));
                        fx[r][t] = (w + dw) * dt;
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
                for (int t = 0; t < kLines; t++) Y[t * nx + i] = fx[r][t];  // vg is restaged from registers for a second pass
            }
        }
        __syncthreads();
        // ---- b -= flux - flux(i-1) (C/tvdb.c:80-88); b1 gets the four edge
        // fluxes F(i-1) in the order A1, A2, B1, B2 apply them.
#pragma unroll
        for (int r = 0; r < M; r++) {
            const int i = lt + r * tp;
            if (live && i < nx) {
                const int im = i == 0 ? nx - 1 : i - 1;
                float fm[kLines];
#pragma unroll
                for (int t = 0; t < kLines; t++) fm[t] = Y[t * nx + im];
                const float b2n = bl[r][0] - (fx[r][0] - fm[0]);
                const float b3n = bl[r][2] - (fx[r][2] - fm[2]);
                const float b1n = (((b1m[r] - fm[0]) + fm[1]) - fm[2]) + fm[3];
                if (pass == NPASS - 1) {
                    float *o = b_out + op[1] * 3;
                    const unsigned ob = 12u * (unsigned)i;
                    st_at<float>(o, ob, b1n);
                    st_at<float>(o, ob + 4u, b2n);
                    st_at<float>(o, ob + 8u, b3n);
                }
                keep[r][0] = b1n;
                keep[r][1] = b2n;
                keep[r][2] = b3n;
                keep[r][3] = bl[r][1] - (fx[r][1] - fm[1]);  // b2 of pencil (j+1, k)
                keep[r][4] = bl[r][3] - (fx[r][3] - fm[3]);  // b3 of pencil (j, k+1)
            }
        }
    }
}

}  // namespace
}  // namespace mhd
