// mhd_corrector.hip -- the learned coarse-grid corrector: a 4-layer 3x3x3
// periodic convolution network (32 channels, GELU, residual output) trained
// and run on MI355X with hand-written BF16 MFMA kernels accumulating in FP32,
// plus the downsampling that makes its training data and the fallback checks
// that guard its use.  No counterpart in the reference.
//
// Layouts.  A voxel index is v = (k*ny + j)*nx + i (the state's cell order).
// Activations are bf16 [V][32], channel fastest, so the 32 input channels of
// one voxel are one 64-byte run; the 8 physical channels (rho, m1..m3, E,
// cell-centred B1..B3) are zero-padded to 32 so every layer is 32 -> 32.
// Convolution weights are bf16 [27][32 out][32 in], tap t = (dz+1)*9 +
// (dy+1)*3 + (dx+1) for the neighbour offset (dx,dy,dz); boundaries are
// periodic, as the box is.
//
// The first layer (8 data channels) has its own packed form, see
// conv_first_kernel.
//
// Forward and input-gradient kernel (implicit GEMM).  A wavefront owns 64
// consecutive voxels of one x-row (M; two adjacent y-rows when ny is even,
// sharing the A loads) and all 32 output channels (N).  Each
// of the 27 taps is one K-step of 32 input channels: the A operand (16
// voxels x 32 channels) is one 16-byte load per lane straight from HBM/L2,
// the B operand (32 channels x 16 outputs) one 16-byte LDS read per lane from
// the weights staged once per 8-wavefront workgroup; v_mfma_f32_16x16x32_bf16
// does the product, 8 per tap (4 voxel blocks x 2 output blocks).  The A
// loads run two taps ahead of the MFMAs from a wave-uniform row base.  The
// input
// gradient is the same convolution with taps mirrored and the weight matrix
// transposed (mhd_corr_weight_prep writes both forms).  Epilogues: bias +
// GELU (storing the FP32 pre-activation when a backward pass follows), bias only
// (output layer), or multiplication by GELU'(z) of the previous layer (the
// backward pass through the activation).
//
// Weight gradient.  dW[t][o][c] = sum_v dY[v][o] * X[v + off_t][c] is a GEMM
// whose K is the voxel index.  A workgroup walks x-rows; for each it stages
// dY of the row and X of the 9 neighbour rows (with one halo voxel each
// side) channel-major in LDS, so each lane's 8 consecutive voxels are
// contiguous; each wavefront owns one 16x16 (o, c) block for all 27 taps.
// Partial sums are written per workgroup and folded in a fixed order, so
// the gradient is deterministic.
#include "mhd_internal.h"
#include "mhd_hip.h"

#include <math.h>

namespace mhd {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr int kC = 32;           // channels of every layer
constexpr int kTaps = 27;
constexpr int kWElems = kTaps * kC * kC;
constexpr int kParams = kWElems + kC;  // weights + bias of one layer
constexpr int kDwGroups = 256;   // workgroups of the weight-gradient kernel

__device__ __forceinline__ u16 f2bf(float x)
{
    unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                        // nearest even
    return (u16)(u >> 16);
}
// the same rounding by the gfx950 conversion instruction (finite inputs)
__device__ __forceinline__ u16 bf16_rne(float x) { return __builtin_bit_cast(u16, (__bf16)x); }
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Phi(z) = 0.5 erfc(-z/sqrt2), the standard normal CDF, in the form of
// Numerical Recipes' erfcc: erfc(x) = t exp(-x^2 + P(t)), t = 1/(1 + x/2),
// x >= 0, with our own degree-9 fit of P = log(exp(x^2) erfc(x) / t) on
// t in [0, 1] (heterogeneous_mhd_amd/phi_fit.py; max |error in P| 1.0e-7,
// so erfc to ~1e-7 relative before FP32 rounding of -x^2).  About 15 operations and one exp against ~60 for
// erfcf; the activations it feeds are rounded to bf16.
constexpr float kPhiP[10] = {
    -1.265512220e+00f, 1.000022661e+00f, 3.741248111e-01f,
    9.639016016e-02f, -1.839756770e-01f, 2.713769287e-01f,
    -1.121139396e+00f, 1.473290611e+00f, -8.133419797e-01f,
    1.687640158e-01f};

struct PhiExp {
    float phi;  // Phi(z)
    float g;    // exp(-z^2/2)
};
template <bool WANT_G>
__device__ __forceinline__ PhiExp phi_of(float z)
{
    const float x = fabsf(z) * 0.70710678118654752f;
    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(0.5f, x, 1.0f));
    float p = kPhiP[9];
#pragma unroll
    for (int q = 8; q >= 0; q--) p = __builtin_fmaf(p, t, kPhiP[q]);
    const float x2 = x * x;
    const float half_erfc = 0.5f * t * __expf(p - x2);
    return PhiExp{z < 0.0f ? half_erfc : 1.0f - half_erfc, WANT_G ? __expf(-x2) : 0.0f};
}
__device__ __forceinline__ float gelu(float z) { return z * phi_of<false>(z).phi; }
__device__ __forceinline__ float gelu_grad(float z)
{
    const PhiExp r = phi_of<true>(z);
    return r.phi + z * r.g * 0.39894228040143268f;  // Phi + z phi(z)
}

enum { kModeHidden = 0, kModeOutput = 1, kModeGrad = 2, kModeFirst = 3 };

// the network output limited to [-lim, lim] (lim <= 0: no limit)
__device__ __forceinline__ float ylim(float y, float lim)
{
    return lim > 0.0f ? fminf(fmaxf(y, -lim), lim) : y;
}

// Epilogue of one 64-voxel output row whose first element is at element
// offset vb: lane holds rows (voxels) 16*mb + 4*lg + q, column (channel)
// 16*nb + lm, i.e. element vb + (16*mb + 4*lg + q)*32 + o.
template <int MODE>
__device__ __forceinline__ void conv_store(const f32x4 (&acc)[4][2], size_t vb, int lm, int lg,
                                           const float *__restrict__ bias, float *__restrict__ zout,
                                           u16 *__restrict__ aout, const float *__restrict__ zprev)
{
    float *zw = zout ? zout + vb : nullptr;
    u16 *aw16 = aout ? aout + vb : nullptr;
    const float *zp = zprev ? zprev + vb : nullptr;
#pragma unroll
    for (int nb = 0; nb < 2; nb++) {
        const int o = 16 * nb + lm;
        const float bo = MODE == kModeGrad ? 0.0f : bias[o];
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const unsigned e = (unsigned)((16 * mb + 4 * lg + q) * kC + o);
                const float x = acc[mb][nb][q];
                if (MODE == kModeHidden) {
                    const float z = x + bo;
                    if (zw) zw[e] = z;  // kept for a backward pass only
                    aw16[e] = bf16_rne(gelu(z));
                } else if (MODE == kModeOutput) {
                    zw[e] = x + bo;
                } else {
                    aw16[e] = bf16_rne(x * gelu_grad(zp[e]));
                }
            }
    }
}

constexpr int kConvWaves = 8;    // wavefronts per forward-kernel workgroup

// A wavefront owns R adjacent y-rows (one z plane) of one 64-voxel x-chunk.
// It walks the 3 (R + 2) (3 dz x R + 2 input rows) neighbour rows once per
// x shift; each loaded A operand feeds every output row it is a tap of (up
// to 3), so A traffic per output row falls from 27 to 9 (R + 2) / R loads.
// Each output row still accumulates its taps in the order t = 0..26, so the
// result does not depend on R.  PF steps of A loads are in flight ahead of
// the MFMAs.
template <int MODE, int R, int PF>
__global__ __launch_bounds__(64 * kConvWaves) __attribute__((amdgpu_waves_per_eu(4)))
void conv_kernel(const u16 *__restrict__ X, const u16 *__restrict__ W,
                 const float *__restrict__ bias, int nx, int ny, int nz,
                 float *__restrict__ zout, u16 *__restrict__ aout,
                 const float *__restrict__ zprev)
{
    __shared__ __attribute__((aligned(16))) u16 wl[kWElems];
    for (int f = threadIdx.x; f < kWElems / 8; f += 64 * kConvWaves)
        reinterpret_cast<uint4 *>(wl)[f] = reinterpret_cast<const uint4 *>(W)[f];
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int lm = lane & 15, lg = lane >> 4;
    const int cx = nx / 64, nyr = ny / R;
    const long g = (long)blockIdx.x * kConvWaves + wave;  // this wavefront's R x 64 voxels
    if (g >= (long)cx * nyr * nz) return;                 // tail workgroup: no work left
    const long pr = g / cx;
    const int i0 = __builtin_amdgcn_readfirstlane((int)(g - pr * cx) * 64);
    const int j = __builtin_amdgcn_readfirstlane((int)(pr % nyr) * R);
    const int k = __builtin_amdgcn_readfirstlane((int)(pr / nyr));

    // A operand: 16 voxels x 32 input channels per voxel block mb.  The
    // neighbour row is a wave-uniform base and the lane offsets are computed
    // once, so a step costs no vector address arithmetic.  A shift of one
    // voxel is +-64 bytes (an immediate offset) except where block 0 (shift
    // -1) or block 3 (shift +1) wraps around the periodic box.
    unsigned ac[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) ac[mb] = 2u * (unsigned)((i0 + 16 * mb + lm) * kC + 8 * lg);
    const unsigned aw = 2u * (unsigned)(wrap(i0 + lm - 1, nx) * kC + 8 * lg);
    const unsigned ae = 2u * (unsigned)(wrap(i0 + 48 + lm + 1, nx) * kC + 8 * lg);
    constexpr int NJ = R + 2, NS = 9 * NJ;  // input rows per plane, load steps
    // step s: plane dz = s / (3 NJ) - 1, input row j - 1 + (s / 3) % NJ, shift s % 3 - 1
    auto load_step = [&](bf16x8 (&av)[4], int s) {
        const int dz = s / (3 * NJ) - 1, jj = (s / 3) % NJ, dx = s % 3 - 1;
        const char *rb = reinterpret_cast<const char *>(
            X + ((size_t)wrap(k + dz, nz) * ny + wrap(j - 1 + jj, ny)) * nx * kC);
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const char *p = dx < 0 && mb == 0 ? rb + aw
                          : dx > 0 && mb == 3 ? rb + ae
                          : rb + 2 * kC * dx + ac[mb];
            av[mb] = *reinterpret_cast<const bf16x8 *>(p);
        }
    };

    f32x4 acc[R][4][2];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int nb = 0; nb < 2; nb++) acc[r][mb][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // ring of PF + 1 steps, fully unrolled so the slots are static registers
    bf16x8 av[PF + 1][4];
#pragma unroll
    for (int s = 0; s < PF; s++) load_step(av[s], s);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (s + PF < NS) load_step(av[(s + PF) % (PF + 1)], s + PF);
        const int slot = s % (PF + 1);
        const int dz = s / (3 * NJ), jj = (s / 3) % NJ, dx = s % 3;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int dy = jj - r;  // tap row index 0..2 of output row j + r
            if (dy < 0 || dy > 2) continue;
            const int t = dz * 9 + dy * 3 + dx;
            bf16x8 bw[2];
#pragma unroll
            for (int nb = 0; nb < 2; nb++)
                bw[nb] = *reinterpret_cast<const bf16x8 *>(wl + (t * kC + 16 * nb + lm) * kC + 8 * lg);
#pragma unroll
            for (int mb = 0; mb < 4; mb++)
#pragma unroll
                for (int nb = 0; nb < 2; nb++)
                    acc[r][mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[slot][mb], bw[nb],
                                                                             acc[r][mb][nb], 0, 0, 0);
        }
        // keep the scheduler from hoisting later steps' loads past the ring
        // (they would need registers beyond the 4-waves-per-SIMD budget)
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int r = 0; r < R; r++)
        conv_store<MODE>(acc[r], (((size_t)k * ny + j + r) * nx + i0) * kC, lm, lg, bias, zout, aout,
                         zprev);
}

constexpr int kConvPF1 = 2;  // A load steps in flight, one row per wavefront
constexpr int kConvPF2 = 1;  // the same, two rows per wavefront (64 more accumulator registers)

template <int R, int PF>
void launch_conv(int mode, const u16 *x, const u16 *w, const float *bias, int nx, int ny, int nz,
                 float *z_out, u16 *a_out, const float *z_prev, hipStream_t s)
{
    const long waves = (long)nx * ny * nz / (64 * R);
    const dim3 grid((unsigned)((waves + kConvWaves - 1) / kConvWaves)), block(64 * kConvWaves);
    if (mode == kModeHidden)
        hipLaunchKernelGGL((conv_kernel<kModeHidden, R, PF>), grid, block, 0, s, x, w, bias, nx, ny, nz,
                           z_out, a_out, nullptr);
    else if (mode == kModeOutput)
        hipLaunchKernelGGL((conv_kernel<kModeOutput, R, PF>), grid, block, 0, s, x, w, bias, nx, ny, nz,
                           z_out, nullptr, nullptr);
    else
        hipLaunchKernelGGL((conv_kernel<kModeGrad, R, PF>), grid, block, 0, s, x, w, nullptr, nx, ny, nz,
                           nullptr, a_out, z_prev);
}

// First layer: only kIn0 = 8 of the 32 input channels carry data (the
// physical fields; pack zero-pads the rest).  Instead of 27 K-steps of which
// 3/4 multiply zeros, each K-step packs 4 taps x 8 channels: lane group lg
// loads the 8 channels of tap 4g + lg, and the weights are stored in that
// order (mhd_corr_weight_prep_first; the 28th tap slot is zero).  7 K-steps,
// 14 KB of staged weights; hidden-layer epilogue.
constexpr int kIn0 = 8;
constexpr int kSteps0 = (kTaps + 3) / 4;
constexpr int kW0Elems = kSteps0 * kC * kC;

__global__ __launch_bounds__(64 * kConvWaves) __attribute__((amdgpu_waves_per_eu(4)))
void conv_first_kernel(const u16 *__restrict__ X, const u16 *__restrict__ W,
                       const float *__restrict__ bias, int nx, int ny, int nz,
                       float *__restrict__ zout, u16 *__restrict__ aout)
{
    __shared__ __attribute__((aligned(16))) u16 wl[kW0Elems];
    for (int f = threadIdx.x; f < kW0Elems / 8; f += 64 * kConvWaves)
        reinterpret_cast<uint4 *>(wl)[f] = reinterpret_cast<const uint4 *>(W)[f];
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int lm = lane & 15, lg = lane >> 4;
    const long v0 = ((long)blockIdx.x * kConvWaves + wave) * 64;  // this wavefront's 64 voxels
    if (v0 >= (long)nx * ny * nz) return;                         // tail workgroup: no work left
    const long row = v0 / nx;
    const int i0 = __builtin_amdgcn_readfirstlane((int)(v0 - row * nx));
    const int j = __builtin_amdgcn_readfirstlane((int)(row % ny));
    const int k = __builtin_amdgcn_readfirstlane((int)(row / ny));

    // A operand of K-step g for this lane: voxel block mb's 16 voxels (lm),
    // channels 0..7 of the neighbour along tap 4g + lg (zero for the pad slot)
    auto load_step = [&](uint4 (&av)[4], int g) {
        const int t = 4 * g + lg;
        const int tt = t < kTaps ? t : kTaps - 1;
        const int dz = tt / 9 - 1, dy = (tt / 3) % 3 - 1, dx = tt % 3 - 1;
        const u16 *rb = X + ((size_t)wrap(k + dz, nz) * ny + wrap(j + dy, ny)) * nx * kC;
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const uint4 a = *reinterpret_cast<const uint4 *>(rb + (size_t)wrap(i0 + 16 * mb + lm + dx, nx) * kC);
            av[mb] = t < kTaps ? a : make_uint4(0u, 0u, 0u, 0u);
        }
    };

    f32x4 acc[4][2];
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int nb = 0; nb < 2; nb++) acc[mb][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    uint4 av[2][4];
    load_step(av[0], 0);
#pragma unroll
    for (int g = 0; g < kSteps0; g++) {
        if (g + 1 < kSteps0) load_step(av[(g + 1) & 1], g + 1);
        bf16x8 bw[2];
#pragma unroll
        for (int nb = 0; nb < 2; nb++)
            bw[nb] = *reinterpret_cast<const bf16x8 *>(wl + (g * kC + 16 * nb + lm) * kC + 8 * lg);
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int nb = 0; nb < 2; nb++)
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(bf16x8, av[g & 1][mb]), bw[nb], acc[mb][nb], 0, 0, 0);
    }
    conv_store<kModeHidden>(acc, (size_t)v0 * kC, lm, lg, bias, zout, aout, nullptr);
}

// Weight gradient, see the file header.  ws[g] (kParams floats per
// workgroup g) receives (or, with accumulate, adds) the group's partial dW
// and bias gradient.
__global__ __launch_bounds__(256) void conv_dw_kernel(const u16 *__restrict__ dY,
                                                      const u16 *__restrict__ X, int nx, int ny,
                                                      int nz, float *__restrict__ ws, int accumulate)
{
    extern __shared__ __attribute__((aligned(16))) u16 sm[];
    const int wn = nx + 2;                 // staged row length with halo
    u16 *dyt = sm;                         // [32][nx]
    u16 *xt = sm + kC * nx;                // [9][32][nx + 2]
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lm = lane & 15, lg = lane >> 4;
    const int ob = wave >> 1, cb = wave & 1;  // this wavefront's (o, c) block
    const long nrow = (long)ny * nz;
    f32x4 acc[kTaps];
#pragma unroll
    for (int t = 0; t < kTaps; t++) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float db = 0.0f;  // bias gradient of channel threadIdx.x (< 32)

    for (long row = blockIdx.x; row < nrow; row += gridDim.x) {
        const int j = (int)(row % ny), k = (int)(row / ny);
        __syncthreads();  // the previous row's LDS reads are done
        // dY of the row, channel-major
        for (int f = threadIdx.x; f < nx * 4; f += 256) {
            const int i = f >> 2, c8 = (f & 3) * 8;
            const uint4 q = *reinterpret_cast<const uint4 *>(dY + (row * nx + i) * kC + c8);
            const u16 *h = reinterpret_cast<const u16 *>(&q);
#pragma unroll
            for (int e = 0; e < 8; e++) dyt[(c8 + e) * nx + i] = h[e];
        }
        // X of the 9 neighbour rows, voxels -1..nx
        for (int f = threadIdx.x; f < 9 * wn * 4; f += 256) {
            const int q9 = f / (wn * 4), rem = f - q9 * wn * 4;
            const int ih = rem >> 2, c8 = (rem & 3) * 8;
            const int dz = q9 / 3 - 1, dy = q9 % 3 - 1;
            const long src = ((long)wrap(k + dz, nz) * ny + wrap(j + dy, ny)) * nx + wrap(ih - 1, nx);
            const uint4 q = *reinterpret_cast<const uint4 *>(X + src * kC + c8);
            const u16 *h = reinterpret_cast<const u16 *>(&q);
#pragma unroll
            for (int e = 0; e < 8; e++) xt[(q9 * kC + c8 + e) * wn + ih] = h[e];
        }
        __syncthreads();
        if (threadIdx.x < kC)
            for (int i = 0; i < nx; i++) db += bf2f(dyt[threadIdx.x * nx + i]);
        for (int kc = 0; kc < nx / 32; kc++) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(dyt + (16 * ob + lm) * nx + 32 * kc + 8 * lg);
#pragma unroll
            for (int t = 0; t < kTaps; t++) {
                const int q9 = t / 3, dx = t % 3 - 1;
                const u16 *p = xt + (q9 * kC + 16 * cb + lm) * wn + 32 * kc + 8 * lg + dx + 1;
                bf16x8 b;
#pragma unroll
                for (int e = 0; e < 8; e++) b[e] = __builtin_bit_cast(__bf16, p[e]);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
            }
        }
    }
    float *out = ws + (size_t)blockIdx.x * kParams;
#pragma unroll
    for (int t = 0; t < kTaps; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int o = 16 * ob + 4 * lg + r, c = 16 * cb + lm;
            float *d = out + (t * kC + o) * kC + c;
            *d = accumulate ? *d + acc[t][r] : acc[t][r];
        }
    if (threadIdx.x < kC) {
        float *d = out + kWElems + threadIdx.x;
        *d = accumulate ? *d + db : db;
    }
}

// grad[e] += sum over the workgroups' partials, in workgroup order
__global__ void dw_reduce_kernel(const float *__restrict__ ws, int groups, float *__restrict__ grad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kParams) return;
    float s = 0.0f;
    for (int g = 0; g < groups; g++) s += ws[(size_t)g * kParams + e];
    grad[e] += s;
}

// FP32 master weights [27][32 o][32 c] + bias -> bf16 forward form
// [t][o][c] and input-gradient form [26 - t][c][o].
__global__ void weight_prep_kernel(const float *__restrict__ w, u16 *__restrict__ wf, u16 *__restrict__ wb)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kWElems) return;
    const int t = e / (kC * kC), o = (e / kC) % kC, c = e % kC;
    const u16 h = f2bf(w[e]);
    wf[e] = h;
    wb[((kTaps - 1 - t) * kC + c) * kC + o] = h;
}

// first-layer form: wp[g][o][8 q + c] = w[4 g + q][o][c] (c < 8), 0 past tap 26
__global__ void weight_prep_first_kernel(const float *__restrict__ w, u16 *__restrict__ wp)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kW0Elems) return;
    const int g = e / (kC * kC), o = (e / kC) % kC, kk = e % kC;
    const int t = 4 * g + kk / kIn0, c = kk % kIn0;
    wp[e] = t < kTaps ? f2bf(w[(t * kC + o) * kC + c]) : (u16)0;
}

__global__ void adam_kernel(float *__restrict__ p, float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                            float c1, float c2, float gscale)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gr = g[e] * gscale;
    const float mm = b1 * m[e] + (1.0f - b1) * gr;
    const float vv = b2 * v[e] + (1.0f - b2) * gr * gr;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    g[e] = 0.0f;
}

__device__ __forceinline__ long vid(int i, int j, int k, int nx, int ny)
{
    return ((long)k * ny + j) * nx + i;
}

// state (u AoS 5, b faces 3) -> normalised bf16 input [V][32]
__global__ void pack_kernel(const float *__restrict__ u, const float *__restrict__ b, int nx, int ny,
                            int nz, const float *__restrict__ norm, u16 *__restrict__ x)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long)nx * ny * nz) return;
    const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long)nx * ny));
    float f[8];
#pragma unroll
    for (int c = 0; c < 5; c++) f[c] = u[v * 5 + c];
    f[5] = 0.5f * (b[v * 3 + 0] + b[vid(wrap(i + 1, nx), j, k, nx, ny) * 3 + 0]);
    f[6] = 0.5f * (b[v * 3 + 1] + b[vid(i, wrap(j + 1, ny), k, nx, ny) * 3 + 1]);
    f[7] = 0.5f * (b[v * 3 + 2] + b[vid(i, j, wrap(k + 1, nz), nx, ny) * 3 + 2]);
    u16 h[kC];
#pragma unroll
    for (int c = 0; c < 8; c++) h[c] = f2bf((f[c] - norm[c]) * norm[8 + c]);
#pragma unroll
    for (int c = 8; c < kC; c++) h[c] = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
        reinterpret_cast<uint4 *>(x + v * kC)[q] = reinterpret_cast<const uint4 *>(h)[q];
}

// The 8 normalised inputs at cell (i, j, k): (f - mean) * inv_std, FP32.
__device__ __forceinline__ void norm_features(const float *__restrict__ u, const float *__restrict__ b,
                                              int i, int j, int k, int nx, int ny, int nz,
                                              const float *__restrict__ norm, float f[8])
{
    const long v = vid(i, j, k, nx, ny);
#pragma unroll
    for (int c = 0; c < 5; c++) f[c] = u[v * 5 + c];
    f[5] = 0.5f * (b[v * 3 + 0] + b[vid(wrap(i + 1, nx), j, k, nx, ny) * 3 + 0]);
    f[6] = 0.5f * (b[v * 3 + 1] + b[vid(i, wrap(j + 1, ny), k, nx, ny) * 3 + 1]);
    f[7] = 0.5f * (b[v * 3 + 2] + b[vid(i, j, wrap(k + 1, nz), nx, ny) * 3 + 2]);
#pragma unroll
    for (int c = 0; c < 8; c++) f[c] = (f[c] - norm[c]) * norm[8 + c];
}

// Network input with derived channels (KIND 1: the 6-neighbour Laplacian of
// each normalised input; 2: its second differences along x, y and z; 3: the
// part of each normalised input that the bf16 rounding of channels 0..7
// drops, times 256).  The derived values are formed in FP32 from the FP32
// normalised inputs, so a bf16 channel carries a curvature to 8 significant
// bits even where it is a small difference of O(1) values -- which the first
// layer's bf16 products could not resolve from channels 0..7 alone.
// Channel 8 + q holds derived value q times norm[16 + q].
template <int KIND>
__global__ void pack_features_kernel(const float *__restrict__ u, const float *__restrict__ b, int nx,
                                     int ny, int nz, const float *__restrict__ norm,
                                     u16 *__restrict__ x)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long)nx * ny * nz) return;
    const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long)nx * ny));
    float f0[8];
    norm_features(u, b, i, j, k, nx, ny, nz, norm, f0);
    u16 h[kC];
#pragma unroll
    for (int c = 0; c < 8; c++) h[c] = f2bf(f0[c]);
#pragma unroll
    for (int c = 8; c < kC; c++) h[c] = 0;
    if constexpr (KIND == 3) {
#pragma unroll
        for (int c = 0; c < 8; c++) h[8 + c] = f2bf((f0[c] - bf2f(h[c])) * 256.0f);
    } else if constexpr (KIND == 1 || KIND == 2) {
        float lap[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            float fm[8], fp[8];
            norm_features(u, b, d == 0 ? wrap(i - 1, nx) : i, d == 1 ? wrap(j - 1, ny) : j,
                          d == 2 ? wrap(k - 1, nz) : k, nx, ny, nz, norm, fm);
            norm_features(u, b, d == 0 ? wrap(i + 1, nx) : i, d == 1 ? wrap(j + 1, ny) : j,
                          d == 2 ? wrap(k + 1, nz) : k, nx, ny, nz, norm, fp);
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const float s = (fm[c] + fp[c]) - 2.0f * f0[c];
                if constexpr (KIND == 1) lap[c] += s;
                else h[8 + 3 * c + d] = f2bf(s * norm[16 + 3 * c + d]);
            }
        }
        if constexpr (KIND == 1) {
#pragma unroll
            for (int c = 0; c < 8; c++) h[8 + c] = f2bf(lap[c] * norm[16 + c]);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
        reinterpret_cast<uint4 *>(x + v * kC)[q] = reinterpret_cast<const uint4 *>(h)[q];
}

// b = curl A with A on edges (A1 at (i, j-1/2, k-1/2), ...), forward differences;
// A is read from channels 5..7 of a [V][ld] array, each value limited to
// [-lim, lim] when lim > 0.
__device__ __forceinline__ void curl_at(const float *A, int ld, int i, int j, int k, int nx, int ny,
                                        int nz, float out[3], float lim = 0.0f)
{
    const float *p0 = A + vid(i, j, k, nx, ny) * ld + 5;
    const float *pi = A + vid(wrap(i + 1, nx), j, k, nx, ny) * ld + 5;
    const float *pj = A + vid(i, wrap(j + 1, ny), k, nx, ny) * ld + 5;
    const float *pk = A + vid(i, j, wrap(k + 1, nz), nx, ny) * ld + 5;
    float a0[3], ai[3], aj[3], ak[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        a0[d] = ylim(p0[d], lim);
        ai[d] = ylim(pi[d], lim);
        aj[d] = ylim(pj[d], lim);
        ak[d] = ylim(pk[d], lim);
    }
    out[0] = (aj[2] - a0[2]) - (ak[1] - a0[1]);
    out[1] = (ak[0] - a0[0]) - (ai[2] - a0[2]);
    out[2] = (ai[1] - a0[1]) - (aj[0] - a0[0]);
}

// R[v][0..4] = y - target (u part); R[v][5..7] = curl(y A-channels) - target
// (B part); lossp[block] = sum of R^2 over the block (double).
__global__ __launch_bounds__(256) void residual_kernel(const float *__restrict__ y,
                                                       const float *__restrict__ tgt, int nx, int ny,
                                                       int nz, float *__restrict__ R,
                                                       double *__restrict__ lossp)
{
    __shared__ double red[256];
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (v < (long)nx * ny * nz) {
        const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long)nx * ny));
#pragma unroll
        for (int c = 0; c < 5; c++) {
            const float r = y[v * kC + c] - tgt[v * 8 + c];
            R[v * 8 + c] = r;
            s += (double)r * r;
        }
        float cb[3];
        curl_at(y, kC, i, j, k, nx, ny, nz, cb);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const float r = cb[d] - tgt[v * 8 + 5 + d];
            R[v * 8 + 5 + d] = r;
            s += (double)r * r;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

// dY = d(mean over voxels of sum R^2)/dy: 2R/V on the u channels, the curl
// adjoint of 2R_B/V on the A channels, 0 on the padding channels.
__global__ void loss_grad_kernel(const float *__restrict__ R, int nx, int ny, int nz, float gscale,
                                 u16 *__restrict__ dy)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long)nx * ny * nz) return;
    const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long)nx * ny));
    const float *r0 = R + v * 8 + 5;
    const float *rim = R + vid(wrap(i - 1, nx), j, k, nx, ny) * 8 + 5;
    const float *rjm = R + vid(i, wrap(j - 1, ny), k, nx, ny) * 8 + 5;
    const float *rkm = R + vid(i, j, wrap(k - 1, nz), nx, ny) * 8 + 5;
    float g[8];
#pragma unroll
    for (int c = 0; c < 5; c++) g[c] = gscale * R[v * 8 + c];
    g[5] = gscale * ((rkm[1] - r0[1]) - (rjm[2] - r0[2]));
    g[6] = gscale * ((r0[0] - rkm[0]) + (rim[2] - r0[2]));
    g[7] = gscale * ((rjm[0] - r0[0]) - (rim[1] - r0[1]));
    u16 h[kC];
#pragma unroll
    for (int c = 0; c < 8; c++) h[c] = f2bf(g[c]);
#pragma unroll
    for (int c = 8; c < kC; c++) h[c] = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
        reinterpret_cast<uint4 *>(dy + v * kC)[q] = reinterpret_cast<const uint4 *>(h)[q];
}

// Block averages of a fine state (factor r): u over r^3 cells, each face
// component over the r^2 fine faces on the coarse face (keeps div b = 0).
__global__ void downsample_kernel(const float *__restrict__ u, const float *__restrict__ b, int nx,
                                  int ny, int nz, int r, float *__restrict__ uc, float *__restrict__ bc)
{
    const int cx = nx / r, cy = ny / r, cz = nz / r;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long)cx * cy * cz) return;
    const int I = (int)(v % cx), J = (int)((v / cx) % cy), K = (int)(v / ((long)cx * cy));
    double su[5] = {0, 0, 0, 0, 0}, sb[3] = {0, 0, 0};
    for (int c = 0; c < r; c++)
        for (int bq = 0; bq < r; bq++) {
            for (int a = 0; a < r; a++) {
                const float *q = u + vid(r * I + a, r * J + bq, r * K + c, nx, ny) * 5;
#pragma unroll
                for (int m = 0; m < 5; m++) su[m] += q[m];
            }
            sb[0] += b[vid(r * I, r * J + bq, r * K + c, nx, ny) * 3 + 0];
            sb[1] += b[vid(r * I + bq, r * J, r * K + c, nx, ny) * 3 + 1];
            sb[2] += b[vid(r * I + bq, r * J + c, r * K, nx, ny) * 3 + 2];
        }
    const double iv = 1.0 / ((double)r * r * r), ia = 1.0 / ((double)r * r);
#pragma unroll
    for (int m = 0; m < 5; m++) uc[v * 5 + m] = (float)(su[m] * iv);
#pragma unroll
    for (int d = 0; d < 3; d++) bc[v * 3 + d] = (float)(sb[d] * ia);
}

// Per block: sums of the u channels of the (limited) correction (double).
__global__ __launch_bounds__(256) void ysum_kernel(const float *__restrict__ y, long n, float lim,
                                                   double *__restrict__ part)
{
    __shared__ double red[5][256];
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int c = 0; c < 5; c++) red[c][threadIdx.x] = v < n ? (double)ylim(y[v * kC + c], lim) : 0.0;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
#pragma unroll
            for (int c = 0; c < 5; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 5) part[5 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// One workgroup: the box means of the u channels, folded in block order.
__global__ __launch_bounds__(256) void ymean_kernel(const double *__restrict__ part, int nblk, long n,
                                                    float *__restrict__ mean)
{
    __shared__ double red[5][256];
#pragma unroll
    for (int c = 0; c < 5; c++) {
        double a = 0.0;
        for (int q = threadIdx.x; q < nblk; q += 256) a += part[5 * q + c];
        red[c][threadIdx.x] = a;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
#pragma unroll
            for (int c = 0; c < 5; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 5) mean[threadIdx.x] = (float)(red[threadIdx.x][0] / (double)n);
}

// candidate = state + correction: u += s_c * (y_c - mean y_c) (c < 5; the
// scheme conserves the box totals of u and so does the fine solution, so the
// correction is projected to zero total), b += s_B * curl(y_A)
__global__ void apply_kernel(const float *__restrict__ u, const float *__restrict__ b,
                             const float *__restrict__ y, int nx, int ny, int nz,
                             const float *__restrict__ scale, const float *__restrict__ ymean,
                             float lim, float *__restrict__ uo, float *__restrict__ bo)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long)nx * ny * nz) return;
    const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long)nx * ny));
#pragma unroll
    for (int c = 0; c < 5; c++)
        uo[v * 5 + c] = u[v * 5 + c] + scale[c] * (ylim(y[v * kC + c], lim) - ymean[c]);
    float cb[3];
    curl_at(y, kC, i, j, k, nx, ny, nz, cb, lim);
#pragma unroll
    for (int d = 0; d < 3; d++) bo[v * 3 + d] = b[v * 3 + d] + scale[5] * cb[d];
}

// Per block (double): the kinetic plus magnetic energy of the state and of
// the candidate, the state's total energy E, and whether every candidate
// cell has rho > 0 and p > 0 (p from E, m and the cell-centred B, gamma =
// 5/3).  The correction leaves the box total of E unchanged (its u part has
// zero box mean), so the change of kinetic plus magnetic energy is the
// thermal energy it converts; the decision bounds that.
__device__ __forceinline__ double kin_mag(const float *q, const float *b, long v, int i, int j, int k, int nx,
                                          int ny, int nz)
{
    const float b1 = 0.5f * (b[v * 3] + b[vid(wrap(i + 1, nx), j, k, nx, ny) * 3]);
    const float b2 = 0.5f * (b[v * 3 + 1] + b[vid(i, wrap(j + 1, ny), k, nx, ny) * 3 + 1]);
    const float b3 = 0.5f * (b[v * 3 + 2] + b[vid(i, j, wrap(k + 1, nz), nx, ny) * 3 + 2]);
    const double kin = 0.5 * ((double)q[1] * q[1] + (double)q[2] * q[2] + (double)q[3] * q[3]) / (double)q[0];
    return kin + 0.5 * ((double)b1 * b1 + (double)b2 * b2 + (double)b3 * b3);
}

__global__ __launch_bounds__(256) void check_kernel(const float *__restrict__ u, const float *__restrict__ b,
                                                    const float *__restrict__ uc,
                                                    const float *__restrict__ bc, int nx, int ny,
                                                    int nz, double *__restrict__ part)
{
    __shared__ double e0[256], e1[256], et[256];
    __shared__ int okb[256];
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, c = 0.0, e = 0.0;
    int ok = 1;
    if (v < (long)nx * ny * nz) {
        const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long)nx * ny));
        a = kin_mag(u + v * 5, b, v, i, j, k, nx, ny, nz);
        c = kin_mag(uc + v * 5, bc, v, i, j, k, nx, ny, nz);
        e = u[v * 5 + 4];
        const double rho = uc[v * 5];
        const double p = (2.0 / 3.0) * ((double)uc[v * 5 + 4] - c);
        ok = rho > 0.0 && p > 0.0;  // false for NaN too
    }
    e0[threadIdx.x] = a;
    e1[threadIdx.x] = c;
    et[threadIdx.x] = e;
    okb[threadIdx.x] = ok;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            e0[threadIdx.x] += e0[threadIdx.x + w];
            e1[threadIdx.x] += e1[threadIdx.x + w];
            et[threadIdx.x] += et[threadIdx.x + w];
            okb[threadIdx.x] &= okb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x] = e0[0];
        part[4 * blockIdx.x + 1] = e1[0];
        part[4 * blockIdx.x + 2] = et[0];
        part[4 * blockIdx.x + 3] = okb[0];
    }
}

// One workgroup: fold the block partials in order, decide, count.
// stats[0] += accepted, stats[1] += rejected; flag[0] = accepted.
__global__ __launch_bounds__(256) void decide_kernel(const double *__restrict__ part, int nblk,
                                                     double etol, int *__restrict__ flag,
                                                     int *__restrict__ stats)
{
    __shared__ double e0[256], e1[256], et[256];
    __shared__ int okb[256];
    double a = 0.0, c = 0.0, e = 0.0;
    int ok = 1;
    for (int q = threadIdx.x; q < nblk; q += 256) {
        a += part[4 * q];
        c += part[4 * q + 1];
        e += part[4 * q + 2];
        ok &= part[4 * q + 3] != 0.0;
    }
    e0[threadIdx.x] = a;
    e1[threadIdx.x] = c;
    et[threadIdx.x] = e;
    okb[threadIdx.x] = ok;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            e0[threadIdx.x] += e0[threadIdx.x + w];
            e1[threadIdx.x] += e1[threadIdx.x + w];
            et[threadIdx.x] += et[threadIdx.x + w];
            okb[threadIdx.x] &= okb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const bool acc = okb[0] && fabs(e1[0] - e0[0]) <= etol * fabs(et[0]);
        flag[0] = acc;
        stats[acc ? 0 : 1] += 1;
    }
}

__global__ void commit_kernel(float *__restrict__ u, float *__restrict__ b, const float *__restrict__ uc,
                              const float *__restrict__ bc, long n, const int *__restrict__ flag)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n || !flag[0]) return;
#pragma unroll
    for (int c = 0; c < 5; c++) u[v * 5 + c] = uc[v * 5 + c];
#pragma unroll
    for (int d = 0; d < 3; d++) b[v * 3 + d] = bc[v * 3 + d];
}

int done() { return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP; }
unsigned blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
bool conv_shape_ok(int nx, int ny, int nz)
{
    const long n = (long)nx * ny * nz;
    return nx >= 64 && nx % 64 == 0 && ny >= 1 && nz >= 1 && n % 256 == 0 && n / 256 < 0x7fffffffL;
}

}  // namespace
}  // namespace mhd

using namespace mhd;

extern "C" {

int mhd_corr_num_params(void) { return kParams; }

int mhd_corr_conv(const uint16_t *x, const uint16_t *w, const float *bias, int nx, int ny, int nz,
                  int mode, float *z_out, uint16_t *a_out, const float *z_prev, mhd_stream_t stream)
{
    if (!x || !w || !conv_shape_ok(nx, ny, nz)) return MHD_ERR_ARG;
    switch (mode) {
    case kModeHidden: if (!bias || !a_out) return MHD_ERR_ARG; break;
    case kModeOutput: if (!bias || !z_out) return MHD_ERR_ARG; break;
    case kModeGrad: if (!a_out || !z_prev) return MHD_ERR_ARG; break;
    case kModeFirst: if (!bias || !a_out) return MHD_ERR_ARG; break;
    default: return MHD_ERR_ARG;
    }
    (void)hipGetLastError();
    const hipStream_t s = (hipStream_t)stream;
    if (mode == kModeFirst) {
        const long waves = (long)nx * ny * nz / 64;
        hipLaunchKernelGGL(conv_first_kernel, dim3((unsigned)((waves + kConvWaves - 1) / kConvWaves)),
                           dim3(64 * kConvWaves), 0, s, x, w, bias, nx, ny, nz, z_out, a_out);
    } else if (ny % 2 == 0) {
        launch_conv<2, kConvPF2>(mode, x, w, bias, nx, ny, nz, z_out, a_out, z_prev, s);
    } else {
        launch_conv<1, kConvPF1>(mode, x, w, bias, nx, ny, nz, z_out, a_out, z_prev, s);
    }
    return done();
}

int mhd_corr_first_weight_elems(void) { return kW0Elems; }

int mhd_corr_weight_prep_first(const float *w, uint16_t *wp, mhd_stream_t stream)
{
    if (!w || !wp) return MHD_ERR_ARG;
    (void)hipGetLastError();
    hipLaunchKernelGGL(weight_prep_first_kernel, dim3(blocks(kW0Elems, 256)), dim3(256), 0,
                       (hipStream_t)stream, w, wp);
    return done();
}

size_t mhd_corr_dw_workspace_bytes(void) { return sizeof(float) * (size_t)kDwGroups * kParams; }

int mhd_corr_conv_dw(const uint16_t *dy, const uint16_t *x, int nx, int ny, int nz, float *ws,
                     int accumulate, mhd_stream_t stream)
{
    if (!dy || !x || !ws || !conv_shape_ok(nx, ny, nz)) return MHD_ERR_ARG;
    const size_t lds = sizeof(u16) * ((size_t)kC * nx + 9 * (size_t)kC * (nx + 2));
    if (lds > 160 * 1024) return MHD_ERR_TOO_LONG;
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)conv_dw_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();
    hipLaunchKernelGGL(conv_dw_kernel, dim3(kDwGroups), dim3(256), lds, (hipStream_t)stream, dy, x,
                       nx, ny, nz, ws, accumulate);
    return done();
}

int mhd_corr_dw_reduce(const float *ws, float *grad, mhd_stream_t stream)
{
    if (!ws || !grad) return MHD_ERR_ARG;
    (void)hipGetLastError();
    hipLaunchKernelGGL(dw_reduce_kernel, dim3(blocks(kParams, 256)), dim3(256), 0, (hipStream_t)stream,
                       ws, kDwGroups, grad);
    return done();
}

int mhd_corr_weight_prep(const float *w, uint16_t *wf, uint16_t *wb, mhd_stream_t stream)
{
    if (!w || !wf || !wb) return MHD_ERR_ARG;
    (void)hipGetLastError();
    hipLaunchKernelGGL(weight_prep_kernel, dim3(blocks(kWElems, 256)), dim3(256), 0,
                       (hipStream_t)stream, w, wf, wb);
    return done();
}

int mhd_corr_adam(float *p, float *g, float *m, float *v, int n, float lr, float beta1, float beta2,
                  float eps, int step, float gscale, mhd_stream_t stream)
{
    if (!p || !g || !m || !v || n < 1 || step < 1) return MHD_ERR_ARG;
    const float c1 = (float)(1.0 / (1.0 - pow((double)beta1, step)));
    const float c2 = (float)(1.0 / (1.0 - pow((double)beta2, step)));
    (void)hipGetLastError();
    hipLaunchKernelGGL(adam_kernel, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, p, g, m,
                       v, n, lr, beta1, beta2, eps, c1, c2, gscale);
    return done();
}

int mhd_corr_pack(const float *u, const float *b, int nx, int ny, int nz, const float *norm,
                  uint16_t *x, mhd_stream_t stream)
{
    if (!u || !b || !norm || !x || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    (void)hipGetLastError();
    hipLaunchKernelGGL(pack_kernel, dim3(blocks((long)nx * ny * nz, 256)), dim3(256), 0,
                       (hipStream_t)stream, u, b, nx, ny, nz, norm, x);
    return done();
}

int mhd_corr_pack_features(const float *u, const float *b, int nx, int ny, int nz, const float *norm,
                           int kind, uint16_t *x, mhd_stream_t stream)
{
    if (!u || !b || !norm || !x || nx < 1 || ny < 1 || nz < 1 || kind < 0 || kind > 3)
        return MHD_ERR_ARG;
    const dim3 g(blocks((long)nx * ny * nz, 256)), bs(256);
    const hipStream_t s = (hipStream_t)stream;
    (void)hipGetLastError();
    switch (kind) {
    case 0: hipLaunchKernelGGL(pack_kernel, g, bs, 0, s, u, b, nx, ny, nz, norm, x); break;
    case 1: hipLaunchKernelGGL(pack_features_kernel<1>, g, bs, 0, s, u, b, nx, ny, nz, norm, x); break;
    case 2: hipLaunchKernelGGL(pack_features_kernel<2>, g, bs, 0, s, u, b, nx, ny, nz, norm, x); break;
    default: hipLaunchKernelGGL(pack_features_kernel<3>, g, bs, 0, s, u, b, nx, ny, nz, norm, x); break;
    }
    return done();
}

size_t mhd_corr_loss_partials(int nx, int ny, int nz) { return blocks((long)nx * ny * nz, 256); }

int mhd_corr_loss_grad(const float *y, const float *target, int nx, int ny, int nz, float *r_scratch,
                       double *loss_partials, uint16_t *dy, mhd_stream_t stream)
{
    if (!y || !target || !r_scratch || !loss_partials || !dy || nx < 1 || ny < 1 || nz < 1)
        return MHD_ERR_ARG;
    const long n = (long)nx * ny * nz;
    const hipStream_t s = (hipStream_t)stream;
    (void)hipGetLastError();
    hipLaunchKernelGGL(residual_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, y, target, nx, ny, nz,
                       r_scratch, loss_partials);
    hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, r_scratch, nx, ny, nz,
                       (float)(2.0 / (double)n), dy);
    return done();
}

int mhd_downsample(const float *u, const float *b, int nx, int ny, int nz, int r, float *uc,
                   float *bc, mhd_stream_t stream)
{
    if (!u || !b || !uc || !bc || r < 1 || nx % r || ny % r || nz % r || nx < 1 || ny < 1 || nz < 1)
        return MHD_ERR_ARG;
    const long n = (long)(nx / r) * (ny / r) * (nz / r);
    (void)hipGetLastError();
    hipLaunchKernelGGL(downsample_kernel, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, u,
                       b, nx, ny, nz, r, uc, bc);
    return done();
}

size_t mhd_corr_check_bytes(int nx, int ny, int nz)
{
    return 5 * sizeof(double) * blocks((long)nx * ny * nz, 256) + 64;
}

int mhd_corr_apply(float *u, float *b, const float *y, int nx, int ny, int nz, const float *scale,
                   float limit, double etol, float *u_cand, float *b_cand, void *check_ws,
                   int *flag_stats, mhd_stream_t stream)
{
    if (!u || !b || !y || !scale || !u_cand || !b_cand || !check_ws || !flag_stats || nx < 1 ||
        ny < 1 || nz < 1)
        return MHD_ERR_ARG;
    const long n = (long)nx * ny * nz;
    const unsigned nb = blocks(n, 256);
    const hipStream_t s = (hipStream_t)stream;
    double *part = (double *)check_ws;  // 5 doubles per block, reused (4 per block) by the check
    float *ymean = (float *)(part + 5 * (size_t)nb);
    (void)hipGetLastError();
    hipLaunchKernelGGL(ysum_kernel, dim3(nb), dim3(256), 0, s, y, n, limit, part);
    hipLaunchKernelGGL(ymean_kernel, dim3(1), dim3(256), 0, s, part, (int)nb, n, ymean);
    hipLaunchKernelGGL(apply_kernel, dim3(nb), dim3(256), 0, s, u, b, y, nx, ny, nz, scale, ymean,
                       limit, u_cand, b_cand);
    hipLaunchKernelGGL(check_kernel, dim3(nb), dim3(256), 0, s, u, b, u_cand, b_cand, nx, ny, nz, part);
    hipLaunchKernelGGL(decide_kernel, dim3(1), dim3(256), 0, s, part, (int)nb, etol, flag_stats,
                       flag_stats + 1);
    hipLaunchKernelGGL(commit_kernel, dim3(nb), dim3(256), 0, s, u, b, u_cand, b_cand, n, flag_stats);
    return done();
}

}  // extern "C"
