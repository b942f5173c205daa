// mhd_transpose_advect.hip -- the single-box launcher of transpose_advect_tile
// (mhd_transpose_advect_body.h): one workgroup per tile, each XCD taking a
// contiguous share of the tiles.
#include "mhd_transpose_advect_body.h"

namespace mhd {
namespace {
using namespace ta_tile;

template <int BLOCK, bool FAST, int WHICH>
__global__ __launch_bounds__(BLOCK) void transpose_advect_kernel(float *__restrict__ u_out, float *__restrict__ b_out,
                                                                 const float *__restrict__ uu,
                                                                 const float *__restrict__ bu, TaGeom g,
                                                                 const float *__restrict__ dtp, float dtv)
{
    unsigned w = blockIdx.x;
    if (g.nwg % 8 == 0) w = (w % 8) * (g.nwg / 8) + w / 8;
    transpose_advect_tile<BLOCK, FAST, WHICH>(u_out, b_out, uu, bu, g, w, dtp ? *dtp : dtv);
}

template <bool FAST, int WHICH>
int launch_ta(float *u_out, float *b_out, const float *uu, const float *bu, const TaGeom &g, const float *dtp,
              float dtv, hipStream_t s)
{
    constexpr int kBlock = 64 * kTP / 2;
    (void)hipGetLastError();
    hipLaunchKernelGGL((transpose_advect_kernel<kBlock, FAST, WHICH>), dim3(g.nwg), dim3(kBlock), 0, s, u_out,
                       b_out, uu, bu, g, dtp, dtv);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

bool ta_overlap(const float *a, size_t na, const float *b, size_t nb) { return a < b + nb && b < a + na; }

}  // namespace

// n1, n2, n3: the INPUT dims (the transpose's argument dims).  transpose12:
// input (np, ns, nq); transpose13 (needs n1 == n3 in the reference form):
// input (np, nq, ns).
bool transpose_advect_supported(const float *u_out, const float *uu, const float *bu, int n1, int n2, int n3,
                                int which)
{
    const int ns = which == 12 ? n2 : n3;
    return n1 % kTP == 0 && ns % 4 == 0 && (uintptr_t)uu % 16 == 0 && (uintptr_t)bu % 16 == 0 &&
           (uintptr_t)u_out % 16 == 0 && n2 >= 1 && n3 >= 1;
}

int launch_transpose_advect(float *u_out, float *b_out, const float *uu, const float *bu, int n1, int n2,
                            int n3, const float *dt_dev, float dt_val, hipStream_t s, bool fast, int which)
{
    if (which != 12 && which != 13) return MHD_ERR_ARG;
    if (!u_out || !b_out || !uu || !bu || n1 < 1 || n2 < 1 || n3 < 1) return MHD_ERR_ARG;
    if (!transpose_advect_supported(u_out, uu, bu, n1, n2, n3, which)) return MHD_ERR_SHAPE;
    const size_t n = (size_t)n1 * n2 * n3;
    if (ta_overlap(u_out, 5 * n, uu, 5 * n) || ta_overlap(b_out, 3 * n, bu, 3 * n) ||
        ta_overlap(u_out, 5 * n, bu, 3 * n) || ta_overlap(b_out, 3 * n, uu, 5 * n) ||
        ta_overlap(u_out, 5 * n, b_out, 3 * n))
        return MHD_ERR_ARG;
    TaGeom g;
    g.np = n1;
    g.ns = which == 12 ? n2 : n3;
    g.nq = which == 12 ? n3 : n2;
    g.nti = (g.ns + kTI - 1) / kTI;
    g.npb = g.np / kTP;
    // runs of at most 64 steps of q, and enough workgroups to fill the chip
    // (each run starts with one extra set of N lines)
    const long cols = (long)g.nti * g.npb;
    long nqc = (g.nq + 63) / 64;
    const long want = (2048 + cols - 1) / cols;
    if (nqc < want) nqc = want;
    if (nqc > g.nq) nqc = g.nq;
    g.qlen = (int)((g.nq + nqc - 1) / nqc);
    g.nqc = (g.nq + g.qlen - 1) / g.qlen;
    const unsigned long long nwg = (unsigned long long)cols * g.nqc;
    if (nwg > 0x7fffffffull) return MHD_ERR_ARG;
    g.nwg = (unsigned)nwg;
    if (which == 12)
        return fast ? launch_ta<true, 12>(u_out, b_out, uu, bu, g, dt_dev, dt_val, s)
                    : launch_ta<false, 12>(u_out, b_out, uu, bu, g, dt_dev, dt_val, s);
    return fast ? launch_ta<true, 13>(u_out, b_out, uu, bu, g, dt_dev, dt_val, s)
                : launch_ta<false, 13>(u_out, b_out, uu, bu, g, dt_dev, dt_val, s);
}

}  // namespace mhd
