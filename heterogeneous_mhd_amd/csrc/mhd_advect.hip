// mhd_advect.hip -- advectbyzx launchers (kernel body: mhd_advect_body.h).
#include "mhd_advect_body.h"

namespace mhd {
namespace {

template <int BLOCK, int M, bool FAST, int TP, bool FIT>
__global__ __launch_bounds__(BLOCK) void advect_kernel(float *b_out, const float *b_main,
                                                       const float *__restrict__ uu,
                                                       const float *__restrict__ bu, int nx,
                                                       int ny, int nz, int tp,
                                                       const float *__restrict__ dtp, float dtv)
{
    extern __shared__ float lds[];
    const int pb = BLOCK / tp;
    const int grp = sweep_group(blockIdx.x, gridDim.x, ny % pb == 0 ? ny / pb : 0, nz);
    float keep[M][5];
    advect_group<BLOCK, M, FAST, TP, FIT>((lds_f *)lds, b_out, b_main, uu, bu, nx, ny, nz, tp, grp,
                                          dtp ? *dtp : dtv, keep);
}

template <int BLOCK, int M, int TP, bool FAST>
int launch_cfg(float *b_out, const float *b_main, const float *uu, const float *bu, int nx, int ny,
               int nz, const float *dtp, float dtv, hipStream_t s)
{
    constexpr int pb = BLOCK / TP;
    const long npen = (long)ny * nz;
    const long nblk = (npen + pb - 1) / pb;
    const size_t lds = sizeof(float) * 2 * kLines * (size_t)nx * pb;
    auto k = nx == M * TP ? advect_kernel<BLOCK, M, FAST, TP, true> : advect_kernel<BLOCK, M, FAST, TP, false>;
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL(k, dim3((unsigned)nblk), dim3(BLOCK), lds, s, b_out, b_main, uu, bu, nx, ny, nz,
                       TP, dtp, dtv);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

template <bool FAST>
int launch_any(float *b_out, const float *b_main, const float *uu, const float *bu, int nx, int ny,
               int nz, const float *dt_dev, float dt_val, hipStream_t s)
{
    const SweepCfg c = sweep_cfg(nx);
    switch (c.tp * 4 + c.m) {
    case 64 * 4 + 1: return launch_cfg<256, 1, 64, FAST>(b_out, b_main, uu, bu, nx, ny, nz, dt_dev, dt_val, s);
    case 64 * 4 + 2: return launch_cfg<256, 2, 64, FAST>(b_out, b_main, uu, bu, nx, ny, nz, dt_dev, dt_val, s);
    case 128 * 4 + 2: return launch_cfg<256, 2, 128, FAST>(b_out, b_main, uu, bu, nx, ny, nz, dt_dev, dt_val, s);
    case 256 * 4 + 2: return launch_cfg<256, 2, 256, FAST>(b_out, b_main, uu, bu, nx, ny, nz, dt_dev, dt_val, s);
    case 512 * 4 + 2: return launch_cfg<512, 2, 512, FAST>(b_out, b_main, uu, bu, nx, ny, nz, dt_dev, dt_val, s);
    default: return launch_cfg<1024, 2, 1024, FAST>(b_out, b_main, uu, bu, nx, ny, nz, dt_dev, dt_val, s);
    }
}

}  // namespace

int launch_advectbyzx(float *b_out, const float *b_main, const float *uu, const float *bu, int nx,
                      int ny, int nz, const float *dt_dev, float dt_val, hipStream_t s, bool fast)
{
    if (!b_out || !b_main || !uu || !bu || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    if (nx > kMaxPencil) return MHD_ERR_TOO_LONG;
    if ((long)ny * nz > 0x7fffffffL) return MHD_ERR_ARG;
    return fast ? launch_any<true>(b_out, b_main, uu, bu, nx, ny, nz, dt_dev, dt_val, s)
                : launch_any<false>(b_out, b_main, uu, bu, nx, ny, nz, dt_dev, dt_val, s);
}

}  // namespace mhd
