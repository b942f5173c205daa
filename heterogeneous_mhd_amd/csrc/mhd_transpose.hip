// mhd_transpose.hip -- transpose12 / transpose13 launchers, u and b in one
// launch (tile code: mhd_transpose_body.h).
#include "mhd_transpose_body.h"

namespace mhd {
namespace {

// One launch transposes u (5 components) and b (3 components): the plane
// index runs over 2 x nplanes, the first half u's planes, the second b's (a
// workgroup-uniform branch; the LDS tile is sized for u).  plane_fast:
// blockIdx.x runs over the planes and (y, z) over the tiles, so the
// workgroups in flight share tile rows; transpose13, whose rows lie a whole
// x-y plane apart in both arrays, uses it.
struct PairArgs {
    float *out[2];
    const float *in[2];
    Perm perm[2];
    int aligned[2];
};

__global__ __launch_bounds__(kTB) void transpose_kernel(PairArgs pa, int ni, int na, int nplanes,
                                                       size_t sa_in, size_t sz_in, size_t si_out,
                                                       size_t sz_out, int plane_fast)
{
    __shared__ __attribute__((aligned(16))) float tile[kT * lds_row<5>()];
    const int i0 = (plane_fast ? blockIdx.y : blockIdx.x) * kT;
    const int a0 = (plane_fast ? blockIdx.z : blockIdx.y) * kT;
    const int pz = plane_fast ? blockIdx.x : blockIdx.z;
    const int which = pz >= nplanes;                // 0: u, 1: b
    const size_t zz = (size_t)(pz - which * nplanes);
    const int ti = ni - i0 < kT ? ni - i0 : kT;   // cells present along i
    const int ta = na - a0 < kT ? na - a0 : kT;   // cells present along a
    const bool fast = pa.aligned[which] && ti == kT && ta == kT;
    if (which == 0) {
        if (fast)
            tile_fast<5>(pa.out[0], pa.in[0], tile, i0, a0, zz, sa_in, sz_in, si_out, sz_out,
                         pa.perm[0]);
        else
            tile_scalar<5>(pa.out[0], pa.in[0], tile, i0, a0, ti, ta, zz, sa_in, sz_in, si_out,
                           sz_out, pa.perm[0]);
    } else {
        if (fast)
            tile_fast<3>(pa.out[1], pa.in[1], tile, i0, a0, zz, sa_in, sz_in, si_out, sz_out,
                         pa.perm[1]);
        else
            tile_scalar<3>(pa.out[1], pa.in[1], tile, i0, a0, ti, ta, zz, sa_in, sz_in, si_out,
                           sz_out, pa.perm[1]);
    }
}

template <int NC>
Perm make_perm(const int *p)
{
    Perm perm = {};
    for (int c = 0; c < NC; c++) {
        perm.p[c] = p[c];
        perm.inv[p[c]] = c;
    }
    return perm;
}

// 16-byte vectors need every row start 16-byte aligned: strides multiple of 4 cells
int aligned_rows(const float *out, const float *in, size_t sa_in, size_t sz_in, size_t si_out,
                 size_t sz_out)
{
    return ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0) && sa_in % 4 == 0 &&
           sz_in % 4 == 0 && si_out % 4 == 0 && sz_out % 4 == 0;
}

int launch_pair(float *u_out, float *b_out, const float *uu, const float *bu, int ni, int na,
                int nplanes, size_t sa_in, size_t sz_in, size_t si_out, size_t sz_out,
                const int *pu, const int *pb, hipStream_t s, bool plane_fast = false)
{
    const unsigned ti = (ni + kT - 1) / kT, ta = (na + kT - 1) / kT;
    const long np2 = 2L * nplanes;
    if (np2 > 65535) plane_fast = true;  // only x takes more than 65535 groups
    if (plane_fast ? (ti > 65535 || ta > 65535 || np2 > 0x7fffffffL) : ta > 65535)
        return MHD_ERR_ARG;
    PairArgs pa;
    pa.out[0] = u_out;
    pa.out[1] = b_out;
    pa.in[0] = uu;
    pa.in[1] = bu;
    pa.perm[0] = make_perm<5>(pu);
    pa.perm[1] = make_perm<3>(pb);
    pa.aligned[0] = aligned_rows(u_out, uu, sa_in, sz_in, si_out, sz_out);
    pa.aligned[1] = aligned_rows(b_out, bu, sa_in, sz_in, si_out, sz_out);
    const dim3 grid = plane_fast ? dim3((unsigned)np2, ti, ta) : dim3(ti, ta, (unsigned)np2);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(kTB), 0, s, pa, ni, na, nplanes, sa_in, sz_in,
                       si_out, sz_out, (int)plane_fast);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

const int kPermU12[5] = {0, 2, 1, 3, 4}, kPermB12[3] = {1, 0, 2};
const int kPermU13[5] = {0, 3, 2, 1, 4}, kPermB13[3] = {2, 1, 0};

}  // namespace

int launch_transpose12(float *u_out, float *b_out, const float *uu, const float *bu, int nx, int ny,
                       int nz, hipStream_t s)
{
    if (!u_out || !b_out || !uu || !bu || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    const size_t plane = (size_t)nx * ny;
    return launch_pair(u_out, b_out, uu, bu, nx, ny, nz, (size_t)nx, plane, (size_t)ny, plane,
                       kPermU12, kPermB12, s);
}

int launch_transpose13(float *u_out, float *b_out, const float *uu, const float *bu, int nx, int ny,
                       int nz, hipStream_t s, bool any_shape)
{
    if (!u_out || !b_out || !uu || !bu || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    if (nx != nz && !any_shape) return MHD_ERR_SHAPE;  // C/transpose13.c:16
    // i = x, a = z, plane index = y.  in cell = i + z*nx*ny + y*nx;
    // out dims (nz, ny, nx): out cell = z + y*nz + i*nz*ny.
    const size_t sa_in = (size_t)nx * ny, sz_in = (size_t)nx;
    const size_t si_out = (size_t)nz * ny, sz_out = (size_t)nz;
    return launch_pair(u_out, b_out, uu, bu, nx, nz, ny, sa_in, sz_in, si_out, sz_out, kPermU13,
                       kPermB13, s, true);
}

}  // namespace mhd
