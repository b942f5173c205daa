// mhd_advect_transpose.hip -- the single-box launcher of advect_transpose_tile
// (mhd_advect_transpose_body.h): one workgroup per tile, each XCD taking a
// contiguous share of the tiles.
#include "mhd_advect_transpose_body.h"

namespace mhd {
namespace {
using namespace at_tile;

template <int BLOCK, bool FAST, int WHICH>
__global__ __launch_bounds__(BLOCK) void advect_transpose_kernel(float *__restrict__ u_out, float *__restrict__ b_out,
                                                                 const float *__restrict__ uu,
                                                                 const float *__restrict__ bu, AtGeom g,
                                                                 const float *__restrict__ dtp, float dtv)
{
    unsigned w = blockIdx.x;
    if (g.nwg % 8 == 0) w = (w % 8) * (g.nwg / 8) + w / 8;
    advect_transpose_tile<BLOCK, FAST, WHICH>(u_out, b_out, uu, bu, g, w, dtp ? *dtp : dtv);
}

template <bool FAST, int WHICH>
int launch_at(float *u_out, float *b_out, const float *uu, const float *bu, const AtGeom &g, const float *dtp,
              float dtv, hipStream_t s)
{
    constexpr int kBlock = 64 * kTP / 2;
    (void)hipGetLastError();
    hipLaunchKernelGGL((advect_transpose_kernel<kBlock, FAST, WHICH>), dim3(g.nwg), dim3(kBlock), 0, s, u_out,
                       b_out, uu, bu, g, dtp, dtv);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

bool at_overlap(const float *a, size_t na, const float *b, size_t nb) { return a < b + nb && b < a + na; }

}  // namespace

int launch_advect_transpose(float *u_out, float *b_out, const float *uu, const float *bu, int n1, int n2, int n3,
                            const float *dt_dev, float dt_val, hipStream_t s, bool fast, int which)
{
    if (which != 12 && which != 13) return MHD_ERR_ARG;
    if (!u_out || !b_out || !uu || !bu || n1 < 1 || n2 < 1 || n3 < 1) return MHD_ERR_ARG;
    const size_t n = (size_t)n1 * n2 * n3;
    if (at_overlap(u_out, 5 * n, uu, 5 * n) || at_overlap(b_out, 3 * n, bu, 3 * n) ||
        at_overlap(u_out, 5 * n, bu, 3 * n) || at_overlap(b_out, 3 * n, uu, 5 * n) ||
        at_overlap(u_out, 5 * n, b_out, 3 * n))
        return MHD_ERR_ARG;
    if (n1 > (1 << 26)) return MHD_ERR_ARG;  // byte offsets within a pencil are 32-bit
    AtGeom g;
    g.n1 = n1, g.n2 = n2, g.n3 = n3;
    g.np = which == 12 ? n2 : n3;
    g.nq = which == 12 ? n3 : n2;
    g.nti = (n1 + kTI - 1) / kTI;
    g.npb = (g.np + kTP - 1) / kTP;
    // runs of at most 64 steps of q, and enough workgroups to fill the chip
    // (each run starts with one extra set of N lines)
    const long cols = (long)g.nti * g.npb;
    long nqc = (g.nq + 63) / 64;
    const long want = (2048 + cols - 1) / cols;
    if (nqc < want) nqc = want;
    if (nqc > g.nq) nqc = g.nq;
    g.qlen = (int)((g.nq + nqc - 1) / nqc);
    g.nqc = (g.nq + g.qlen - 1) / g.qlen;
    // the per-lane 32-bit byte offsets of the kernel: the tile's pencils
    // (u and b) and its output rows (u)
    const unsigned long long pspan = which == 12 ? (unsigned long long)kTP : (unsigned long long)kTP * n2;
    const unsigned long long ospan = which == 12 ? (unsigned long long)g.np : (unsigned long long)g.nq * g.np;
    if (pspan * n1 * 20 >= (1ull << 32) || (ospan * kTI + kTP) * 20 >= (1ull << 32)) return MHD_ERR_TOO_LONG;
    const unsigned long long nwg = (unsigned long long)cols * g.nqc;
    if (nwg > 0x7fffffffull) return MHD_ERR_ARG;
    g.nwg = (unsigned)nwg;
    if (which == 12)
        return fast ? launch_at<true, 12>(u_out, b_out, uu, bu, g, dt_dev, dt_val, s)
                    : launch_at<false, 12>(u_out, b_out, uu, bu, g, dt_dev, dt_val, s);
    return fast ? launch_at<true, 13>(u_out, b_out, uu, bu, g, dt_dev, dt_val, s)
                : launch_at<false, 13>(u_out, b_out, uu, bu, g, dt_dev, dt_val, s);
}

}  // namespace mhd
