// mhd_fluid.hip -- fluidx launchers (kernel body: mhd_fluid_body.h).
#include "mhd_fluid_body.h"

namespace mhd {
namespace {

// The FP32 sweep with the fused calcfl and the exact sweep without it are
// held to 80 VGPRs (>= 6 waves per SIMD): three workgroups per CU, as LDS
// allows, instead of the two their 81-82 VGPRs would leave room for (the
// exact sweep: 25.6 -> 23.7 ms at 1024^3, no spills).
template <int BLOCK, int M, bool CFL, bool FAST, int TP, bool FIT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu((CFL == FAST) && BLOCK <= 512 ? 6 : 1)))
void fluidx_kernel(float *u_out, const float *uu, const float *__restrict__ bu, int nx, int ny, int nz, int tp,
                   const float *__restrict__ dtp, float dtv, CflOut co)
{
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ float red[2 * (BLOCK / 64)];
    const int pb = BLOCK / tp;
    const int grp = sweep_group(blockIdx.x, gridDim.x, ny % pb == 0 ? ny / pb : 0, nz);
    const float dt = dtp ? *dtp : dtv;
    fluid_group<BLOCK, M, CFL, FAST, false, TP, FIT>((lds_f *)lds, (lds_f *)red, u_out, uu, bu, nx, ny, nz, tp, grp, dt,
                                                     co.ws, grp);
}

template <int BLOCK, int M, int TP, bool CFL, bool FAST>
int launch_cfg(float *u_out, const float *uu, const float *bu, int nx, int ny, int nz,
               const float *dtp, float dtv, const CflOut &co, hipStream_t s)
{
    constexpr int pb = BLOCK / TP;
    const long npen = (long)ny * nz;
    const long nblk = (npen + pb - 1) / pb;
    const size_t lds = sizeof(float) * 10 * (size_t)nx * pb;
    // nx == M*TP: the compile-time pencil length (no bounds tests, immediate offsets)
    auto k = nx == M * TP ? fluidx_kernel<BLOCK, M, CFL, FAST, TP, true>
                          : fluidx_kernel<BLOCK, M, CFL, FAST, TP, false>;
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL(k, dim3((unsigned)nblk), dim3(BLOCK), lds, s, u_out, uu, bu, nx, ny, nz, TP, dtp,
                       dtv, co);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

template <bool CFL, bool FAST>
int launch_any(float *u_out, const float *uu, const float *bu, int nx, int ny, int nz,
               const float *dtp, float dtv, const CflOut &co, hipStream_t s)
{
    const SweepCfg c = sweep_cfg(nx);
    switch (c.tp * 4 + c.m) {
    case 64 * 4 + 1: return launch_cfg<256, 1, 64, CFL, FAST>(u_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    case 64 * 4 + 2: return launch_cfg<256, 2, 64, CFL, FAST>(u_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    case 128 * 4 + 2: return launch_cfg<256, 2, 128, CFL, FAST>(u_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    case 256 * 4 + 2: return launch_cfg<256, 2, 256, CFL, FAST>(u_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    case 512 * 4 + 2: return launch_cfg<512, 2, 512, CFL, FAST>(u_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    default: return launch_cfg<1024, 2, 1024, CFL, FAST>(u_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    }
}

template <bool CFL>
int launch_mode(bool fast, float *u_out, const float *uu, const float *bu, int nx, int ny, int nz,
                const float *dtp, float dtv, const CflOut &co, hipStream_t s)
{
    return fast ? launch_any<CFL, true>(u_out, uu, bu, nx, ny, nz, dtp, dtv, co, s)
                : launch_any<CFL, false>(u_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
}

}  // namespace

size_t fluid_workgroups(int nx, long npen)
{
    const SweepCfg c = sweep_cfg(nx);
    const int pb = c.block / c.tp;
    return (size_t)((npen + pb - 1) / pb);
}

int launch_fluidx(float *u_out, const float *uu, const float *bu, int nx, int ny, int nz,
                  const float *dt_dev, float dt_val, hipStream_t s, void *cfl_ws, float *inv_c,
                  float *dt_next, double cfl, bool fast)
{
    if (!u_out || !uu || !bu || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    if (nx > kMaxPencil) return MHD_ERR_TOO_LONG;
    if ((long)ny * nz > 0x7fffffffL) return MHD_ERR_ARG;
    CflOut co = {cfl_ws, inv_c, dt_next, nullptr, cfl, 0.0f};
    if (cfl_ws) {
        if (!inv_c) return MHD_ERR_ARG;
        const int rc = launch_mode<true>(fast, u_out, uu, bu, nx, ny, nz, dt_dev, dt_val, co, s);
        if (rc != MHD_OK) return rc;
        return launch_cfl_fold((int)fluid_workgroups(nx, (long)ny * nz), co, s);
    }
    return launch_mode<false>(fast, u_out, uu, bu, nx, ny, nz, dt_dev, dt_val, co, s);
}

namespace {
__global__ void begin_step_kernel(float *t, float *dt, const float *dt_next, float tf)
{
    float d = *dt_next;
    const float t0 = *t;
    if (tf > 0.0f && d > (tf - t0) / 2) d = (tf - t0) / 2;   // C/mhd.c:63
    *t = t0 + 2 * d;                                         // C/mhd.c:64
    *dt = d;
}
}  // namespace

int launch_begin_step(float *t, float *dt, const float *dt_next, float tf, hipStream_t s)
{
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL(begin_step_kernel, dim3(1), dim3(1), 0, s, t, dt, dt_next, tf);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

}  // namespace mhd
