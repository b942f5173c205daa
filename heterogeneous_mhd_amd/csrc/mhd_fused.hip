// mhd_fused.hip -- advectbyzx followed by fluidx of the same sweep in one
// kernel.  In the second half of the split step the reference calls
// advectbyzx and then fluidx on the same pencils (C/mhd.c:95-97, :105-107,
// :115-117).  The fluid sweep of pencil (j,k) needs the new b of that pencil
// and the new b2 of (j+1,k) and b3 of (j,k+1); the advect workgroup of (j,k)
// already evaluates the tvdb lines those come from (A(j+1) and B(k+1)), in
// the same operation order, so it keeps all five in registers and runs the
// fluid sweep straight away: one pass over HBM instead of two, and the
// latency-bound field update overlaps the arithmetic-bound fluid update
// across workgroups.  Results are bitwise those of the two separate calls.
// u is written to a buffer other than the one read, because neighbouring
// workgroups still read the old rho and m1 of this pencil.
// npass = 2 runs advectbyzx twice before the fluid sweep: the pair at the
// middle of the step (C/mhd.c:93 and :95-97), whose two advectbyzx calls use
// the same u (see advect_group).
#include "mhd_advect_body.h"

namespace mhd {
namespace {

// The two-pass kernel is held to >= 6 waves per SIMD (<= 80 VGPRs, no
// spills): three 512-thread workgroups per CU instead of the two its 86
// VGPRs would allow, measured 30.5 -> 24.1 ms per 1024^3 launch.  So is the
// FP32 one-pass kernel with the fused calcfl (81 VGPRs unbounded).  The
// other one-pass kernels keep the compiler's choice (64 VGPRs for a fourth
// workgroup gave no gain).
template <int BLOCK, int M, bool CFL, bool FAST, int TP, bool FIT, int NPASS>
__global__ __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(NPASS == 2 || (CFL && FAST && BLOCK <= 512) ? 6 : 1)))
void advect_fluid_kernel(float *u_out, float *b_out,
                                                             const float *__restrict__ uu,
                                                             const float *__restrict__ bu, int nx,
                                                             int ny, int nz, int tp,
                                                             const float *__restrict__ dtp,
                                                             float dtv, CflOut co)
{
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ float red[2 * (BLOCK / 64)];
    const int pb = BLOCK / tp;
    const int grp = sweep_group(blockIdx.x, gridDim.x, ny % pb == 0 ? ny / pb : 0, nz);
    const float dt = dtp ? *dtp : dtv;
    float keep[M][5];
    advect_group<BLOCK, M, FAST, TP, FIT, NPASS>((lds_f *)lds, b_out, bu, uu, bu, nx, ny, nz, tp, grp, dt,
                                                 keep);
    __syncthreads();  // the advect's last LDS reads are done: the fluid sweep reuses LDS
    fluid_group<BLOCK, M, CFL, FAST, true, TP, FIT>((lds_f *)lds, (lds_f *)red, u_out, uu, bu, nx, ny, nz, tp, grp, dt,
                                                    co.ws, grp, keep);
}

template <int BLOCK, int M, int TP, bool CFL, bool FAST, int NP>
int launch_cfg(float *u_out, float *b_out, const float *uu, const float *bu, int nx, int ny, int nz,
               const float *dtp, float dtv, const CflOut &co, hipStream_t s)
{
    constexpr int pb = BLOCK / TP;
    const long npen = (long)ny * nz;
    const long nblk = (npen + pb - 1) / pb;
    // the fluid sweep's 10 floats per cell exceed the advect's 2*kLines
    const size_t lds = sizeof(float) * 10 * (size_t)nx * pb;
    auto k = nx == M * TP ? advect_fluid_kernel<BLOCK, M, CFL, FAST, TP, true, NP>
                          : advect_fluid_kernel<BLOCK, M, CFL, FAST, TP, false, NP>;
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL(k, dim3((unsigned)nblk), dim3(BLOCK), lds, s, u_out, b_out, uu, bu, nx, ny, nz,
                       TP, dtp, dtv, co);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

template <bool CFL, bool FAST, int NP>
int launch_any(float *u_out, float *b_out, const float *uu, const float *bu, int nx, int ny, int nz,
               const float *dtp, float dtv, const CflOut &co, hipStream_t s)
{
    const SweepCfg c = sweep_cfg(nx);
    switch (c.tp * 4 + c.m) {
    case 64 * 4 + 1: return launch_cfg<256, 1, 64, CFL, FAST, NP>(u_out, b_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    case 64 * 4 + 2: return launch_cfg<256, 2, 64, CFL, FAST, NP>(u_out, b_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    case 128 * 4 + 2: return launch_cfg<256, 2, 128, CFL, FAST, NP>(u_out, b_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    case 256 * 4 + 2: return launch_cfg<256, 2, 256, CFL, FAST, NP>(u_out, b_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    case 512 * 4 + 2: return launch_cfg<512, 2, 512, CFL, FAST, NP>(u_out, b_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    default: return launch_cfg<1024, 2, 1024, CFL, FAST, NP>(u_out, b_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
    }
}

template <bool CFL, int NP>
int launch_mode(bool fast, float *u_out, float *b_out, const float *uu, const float *bu, int nx,
                int ny, int nz, const float *dtp, float dtv, const CflOut &co, hipStream_t s)
{
    return fast ? launch_any<CFL, true, NP>(u_out, b_out, uu, bu, nx, ny, nz, dtp, dtv, co, s)
                : launch_any<CFL, false, NP>(u_out, b_out, uu, bu, nx, ny, nz, dtp, dtv, co, s);
}

bool overlap(const float *a, size_t na, const float *b, size_t nb)
{
    return a < b + nb && b < a + na;
}

}  // namespace

int launch_advect_fluid(float *u_out, float *b_out, const float *uu, const float *bu, int nx, int ny,
                        int nz, const float *dt_dev, float dt_val, hipStream_t s, void *cfl_ws,
                        float *inv_c, float *dt_next, double cfl, bool fast, int npass)
{
    if (npass != 1 && !(npass == 2 && !cfl_ws)) return MHD_ERR_ARG;
    if (!u_out || !b_out || !uu || !bu || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    if (nx > kMaxPencil) return MHD_ERR_TOO_LONG;
    if ((long)ny * nz > 0x7fffffffL) return MHD_ERR_ARG;
    const size_t n = (size_t)nx * ny * nz;
    if (overlap(u_out, 5 * n, uu, 5 * n) || overlap(b_out, 3 * n, bu, 3 * n)) return MHD_ERR_ARG;
    CflOut co = {cfl_ws, inv_c, dt_next, nullptr, cfl, 0.0f};
    if (cfl_ws) {
        if (!inv_c) return MHD_ERR_ARG;
        const int rc = launch_mode<true, 1>(fast, u_out, b_out, uu, bu, nx, ny, nz, dt_dev, dt_val, co, s);
        if (rc != MHD_OK) return rc;
        return launch_cfl_fold((int)fluid_workgroups(nx, (long)ny * nz), co, s);
    }
    if (npass == 2) return launch_mode<false, 2>(fast, u_out, b_out, uu, bu, nx, ny, nz, dt_dev, dt_val, co, s);
    return launch_mode<false, 1>(fast, u_out, b_out, uu, bu, nx, ny, nz, dt_dev, dt_val, co, s);
}

}  // namespace mhd
