// mhd_fluid.hip -- fluidx: the relaxing-TVD fluid sweep along the first axis.
//
// Reference: fluidx C/fluidx.c:19-102, tvd1 C/tvd1.c:8-118,
// mhdflux/find_max C/mhdflux.c:9-72.
//
// Mapping.  One pencil (all nx cells of one (j,k)) is owned by `tp` threads
// (a multiple of the 64-wide wavefront); a 256- or 512-thread workgroup owns
// blockDim/tp consecutive pencils, i.e. one contiguous run of u in HBM, which
// it loads and stores fully coalesced through LDS.  Thread t of a pencil holds
// cells t, t+tp, t+2tp, ... in registers.  Stencil neighbours (radius 3 over
// the two TVD stages) go through two LDS arrays of 5 floats per cell, stored
// cell-major: a stride of 5 words is odd, so the per-cell reads of a
// wavefront are free of bank conflicts.  The pencil-global freezing speed c
// (find_max, twice per sweep) is a wavefront butterfly plus, when a pencil
// spans several wavefronts, one LDS exchange.
//
// Arithmetic.  Every expression repeats the reference's float/double
// pattern (float storage, double wherever a double literal or libm double
// routine appears) and the library is built with -ffp-contract=off, so the
// result is bitwise that of the reference compiled by gcc on x86-64.
// Each "+ z" is the reference's abXYc "+ c" with c = 0 (C/includefunc_spu.h:80);
// it maps -0 to +0 and is kept on purpose.
#include "mhd_internal.h"

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

// C/tvd1.c:75-79: 2ab/(a+b) in double where the float product a*b > 0.
__device__ __forceinline__ float limiter(float a, float b)
{
    float r = 0.0f;
    if (a * b > 0.0f) r = (float)(2.0 * (double)a * (double)b / (double)(a + b));
    return r;
}

// Max over the threads of one pencil.  x >= 0 and never NaN (the callers
// start from 0 and keep the running max with a strict '>', as find_max does,
// so a NaN candidate never enters).  Ends with a workgroup barrier, which the
// caller relies on to recycle the LDS stencil arrays.
template <int BLOCK>
__device__ __forceinline__ float pencil_max(float x, int tp, float *red)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float o = __shfl_xor(x, off, 64);
        x = o > x ? o : x;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (tp > 64) {
        if (lane == 0) red[wave] = x;
        __syncthreads();
        const int wpp = tp >> 6, first = (threadIdx.x / tp) * wpp;
        x = red[first];
        for (int w = 1; w < wpp; w++) {
            float o = red[first + w];
            x = o > x ? o : x;
        }
    }
    __syncthreads();
    return x;
}

template <int BLOCK, int M>
__global__ __launch_bounds__(BLOCK) void fluidx_kernel(float *u_out,
                                                       const float *uu,
                                                       const float *__restrict__ bu, int nx, int ny,
                                                       int nz, int tp, const float *__restrict__ dtp,
                                                       float dtv)
{
    extern __shared__ float lds[];
    __shared__ float red[2][BLOCK / 64];
    const int pb = BLOCK / tp;                 // pencils per workgroup
    const int lp = threadIdx.x / tp;           // this thread's pencil in the group
    const int lt = threadIdx.x - lp * tp;      // this thread's rank in its pencil
    const long npen = (long)ny * nz;
    const long pen0 = (long)blockIdx.x * pb;
    const int npb = (int)((npen - pen0) < pb ? (npen - pen0) : pb);  // pencils present
    const long pen = pen0 + lp;
    const bool live = lp < npb;
    const int j = (int)(pen % ny), k = (int)(pen / ny);
    const int jp = j + 1 == ny ? 0 : j + 1, kp = k + 1 == nz ? 0 : k + 1;  // C/fluidx.c:52-53

    float *A = lds;                            // 5*nx*pb floats
    float *B = lds + 5 * nx * pb;              // 5*nx*pb floats
    float *Ap = A + 5 * nx * lp, *Bp = B + 5 * nx * lp;

    const float dt = dtp ? *dtp : dtv;
    const float z = 0.0f, h = 0.5f;

    // ---- coalesced load of the group's u (5 floats/cell) and b (3/cell)
    {
        const float *gu = uu + (size_t)pen0 * nx * 5;
        const float *gb = bu + (size_t)pen0 * nx * 3;
        const int nu = 5 * nx * npb, nb = 3 * nx * npb;
        for (int f = threadIdx.x; f < nu; f += BLOCK) A[f] = gu[f];
        for (int f = threadIdx.x; f < nb; f += BLOCK) B[f] = gb[f];
    }
    __syncthreads();

    float q[M][5], bc[M][3], v[M][5];
    float cmax = 0.0f;
    const float *Bs = B + 3 * nx * lp;         // b staging of this pencil (3 floats/cell)
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
            const int ip = i + 1 == nx ? 0 : i + 1;
#pragma unroll
            for (int c = 0; c < 5; c++) q[r][c] = Ap[5 * i + c];
            // cell-centred B, C/fluidx.c:71-81
            const float h1 = h * Bs[3 * i], h1p = h * Bs[3 * ip];
            const float b2jp = bu[(((size_t)k * ny + jp) * nx + i) * 3 + 1];
            const float b3kp = bu[(((size_t)kp * ny + j) * nx + i) * 3 + 2];
            bc[r][0] = h1 + h1p;
            bc[r][1] = h * Bs[3 * i + 1] + b2jp / 2;
            bc[r][2] = h * Bs[3 * i + 2] + b3kp / 2;
            const float cf = flux_cell(q[r], bc[r], v[r]);
            cmax = cf > cmax ? cf : cmax;
        }
    }
    const float c1 = pencil_max<BLOCK>(cmax, tp, red[0]);

    // ---- stage 1: half-step predictor, C/tvd1.c:36-49
    // s[][] holds, in turn, fr of stage 1, the predictor u1, and the final
    // flux; q keeps the sweep's input u, the base of the final update (:116).
    float s[M][5];
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
#pragma unroll
            for (int c = 0; c < 5; c++) {
                float vv = v[r][c];
                if (c1 > 0.0f) vv = vv / c1;
                const float wr = q[r][c] + vv + z;
                Ap[5 * i + c] = q[r][c] - vv + z;           // wl
                s[r][c] = c1 * wr + z * wr + z;   // fr
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
            const int ip = i + 1 == nx ? 0 : i + 1;
#pragma unroll
            for (int c = 0; c < 5; c++) {
                const float t = Ap[5 * ip + c];
                const float fl = c1 * t + z * t + z;
                Bp[5 * i + c] = h * s[r][c] + (-h) * fl + z;  // flux
            }
        }
    }
    __syncthreads();
    const float hdt = (float)(-1.0 * (double)dt / 2.0);
    cmax = 0.0f;
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
            const int im = i == 0 ? nx - 1 : i - 1;
#pragma unroll
            for (int c = 0; c < 5; c++) {
                const float d = Bp[5 * i + c] - Bp[5 * im + c] + z;
                s[r][c] = q[r][c] + hdt * d + z;           // u1
            }
            const float cf = flux_cell(s[r], bc[r], v[r]);
            cmax = cf > cmax ? cf : cmax;
        }
    }
    const float c2 = pencil_max<BLOCK>(cmax, tp, red[1]);

    // ---- stage 2: corrector with the van Leer limiter, C/tvd1.c:54-116
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
#pragma unroll
            for (int c = 0; c < 5; c++) {
                float vv = v[r][c];
                if (c2 > 0.0f) vv = vv / c2;
                const float wr = s[r][c] + vv + z;
                Ap[5 * i + c] = c2 * wr + z * wr + z;      // fr
                Bp[5 * i + c] = s[r][c] - vv + z;          // wl
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
            const int ip = i + 1 == nx ? 0 : i + 1;
            const int ipp = ip + 1 == nx ? 0 : ip + 1;
            const int im = i == 0 ? nx - 1 : i - 1;
#pragma unroll
            for (int c = 0; c < 5; c++) {
                const float frm = Ap[5 * im + c], fr0 = Ap[5 * i + c], frp = Ap[5 * ip + c];
                const float dfrp = h * frp + (-h) * fr0 + z;
                const float dfrm = h * fr0 + (-h) * frm + z;
                const float dfr = limiter(dfrp, dfrm);
                const float w0 = Bp[5 * i + c], w1 = Bp[5 * ip + c], w2 = Bp[5 * ipp + c];
                const float flm = c2 * w0 + z * w0 + z;     // fl(i-1)
                const float fl0 = c2 * w1 + z * w1 + z;     // fl(i)
                const float flp = c2 * w2 + z * w2 + z;     // fl(i+1)
                const float dflp = h * fl0 + (-h) * flp + z;
                const float dflm = h * flm + (-h) * fl0 + z;
                const float dfl = limiter(dflp, dflm);
                const float t1 = fr0 - fl0 + z;
                const float t2 = dfr - dfl + z;
                s[r][c] = h * t1 + h * t2 + z;              // flux
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
#pragma unroll
    for (int r = 0; r < M; r++) {
        const int i = lt + r * tp;
        if (live && i < nx) {
            const int im = i == 0 ? nx - 1 : i - 1;
#pragma unroll
            for (int c = 0; c < 5; c++) {
                const float d = Ap[5 * i + c] - Ap[5 * im + c] + z;
                Bp[5 * i + c] = q[r][c] + fdt * d + z;
            }
        }
    }
    __syncthreads();
    {
        float *gu = u_out + (size_t)pen0 * nx * 5;
        const int nu = 5 * nx * npb;
        for (int f = threadIdx.x; f < nu; f += BLOCK) gu[f] = B[f];
    }
}

template <int BLOCK, int M>
int launch_cfg(float *u_out, const float *uu, const float *bu, int nx, int ny, int nz, int tp,
               const float *dtp, float dtv, hipStream_t s)
{
    const int pb = BLOCK / tp;
    const long npen = (long)ny * nz;
    const long nblk = (npen + pb - 1) / pb;
    const size_t lds = sizeof(float) * 10 * (size_t)nx * pb;
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)fluidx_kernel<BLOCK, M>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL((fluidx_kernel<BLOCK, M>), dim3((unsigned)nblk), dim3(BLOCK), lds, s, u_out,
                       uu, bu, nx, ny, nz, tp, dtp, dtv);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

}  // namespace

int launch_fluidx(float *u_out, const float *uu, const float *bu, int nx, int ny, int nz,
                  const float *dt_dev, float dt_val, hipStream_t s)
{
    if (!u_out || !uu || !bu || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    if (nx > kMaxPencil) return MHD_ERR_TOO_LONG;
    if (nx <= 64) return launch_cfg<256, 1>(u_out, uu, bu, nx, ny, nz, 64, dt_dev, dt_val, s);
    if (nx <= 128) return launch_cfg<256, 2>(u_out, uu, bu, nx, ny, nz, 64, dt_dev, dt_val, s);
    if (nx <= 256) return launch_cfg<256, 4>(u_out, uu, bu, nx, ny, nz, 64, dt_dev, dt_val, s);
    if (nx <= 512) return launch_cfg<256, 4>(u_out, uu, bu, nx, ny, nz, 128, dt_dev, dt_val, s);
    if (nx <= 1024) return launch_cfg<256, 4>(u_out, uu, bu, nx, ny, nz, 256, dt_dev, dt_val, s);
    return launch_cfg<512, 4>(u_out, uu, bu, nx, ny, nz, 512, dt_dev, dt_val, s);
}

}  // namespace mhd
