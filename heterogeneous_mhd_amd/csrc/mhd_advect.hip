// mhd_advect.hip -- advectbyzx launchers (kernel body: mhd_advect_body.h).
#include "mhd_advect_body.h"

namespace mhd {
namespace {

template <int BLOCK, int M, bool FAST>
__global__ __launch_bounds__(BLOCK) void advect_kernel(float *b_out, const float *b_main,
                                                       const float *__restrict__ uu,
                                                       const float *__restrict__ bu, int nx,
                                                       int ny, int nz, int tp,
                                                       const float *__restrict__ dtp, float dtv)
{
    extern __shared__ float lds[];
    const int pb = BLOCK / tp;
    const int grp = sweep_group(blockIdx.x, gridDim.x, ny % pb == 0 ? ny / pb : 0, nz);
    advect_group<BLOCK, M, FAST>(lds, b_out, b_main, uu, bu, nx, ny, nz, tp, grp, dtp ? *dtp : dtv);
}

template <int BLOCK, int M, bool FAST>
int launch_cfg(float *b_out, const float *b_main, const float *uu, const float *bu, int nx, int ny,
               int nz, int tp, const float *dtp, float dtv, hipStream_t s)
{
    const int pb = BLOCK / tp;
    const long npen = (long)ny * nz;
    const long nblk = (npen + pb - 1) / pb;
    const size_t lds = sizeof(float) * 3 * kLines * (size_t)nx * pb;
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)advect_kernel<BLOCK, M, FAST>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL((advect_kernel<BLOCK, M, FAST>), dim3((unsigned)nblk), dim3(BLOCK), lds, s, b_out,
                       b_main, uu, bu, nx, ny, nz, tp, dtp, dtv);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

template <bool FAST>
int launch_any(float *b_out, const float *b_main, const float *uu, const float *bu, int nx, int ny,
               int nz, const float *dt_dev, float dt_val, hipStream_t s)
{
    if (nx <= 64) return launch_cfg<256, 1, FAST>(b_out, b_main, uu, bu, nx, ny, nz, 64, dt_dev, dt_val, s);
    if (nx <= 128) return launch_cfg<256, 2, FAST>(b_out, b_main, uu, bu, nx, ny, nz, 64, dt_dev, dt_val, s);
    if (nx <= 256) return launch_cfg<256, 2, FAST>(b_out, b_main, uu, bu, nx, ny, nz, 128, dt_dev, dt_val, s);
    if (nx <= 512) return launch_cfg<256, 2, FAST>(b_out, b_main, uu, bu, nx, ny, nz, 256, dt_dev, dt_val, s);
    if (nx <= 1024) return launch_cfg<512, 2, FAST>(b_out, b_main, uu, bu, nx, ny, nz, 512, dt_dev, dt_val, s);
    return launch_cfg<1024, 2, FAST>(b_out, b_main, uu, bu, nx, ny, nz, 1024, dt_dev, dt_val, s);
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
