// mhd_transpose.hip -- transpose12 / transpose13 launchers (tile code:
// mhd_transpose_body.h).
#include "mhd_transpose_body.h"

namespace mhd {
namespace {

// plane_fast: blockIdx.x runs over the planes and (y, z) over the tiles, so
// the workgroups in flight share tile rows; transpose13, whose rows lie a
// whole x-y plane apart in both arrays, uses it.
template <int NC>
__global__ __launch_bounds__(kTB) void transpose_kernel(float *__restrict__ out,
                                                       const float *__restrict__ in, int ni,
                                                       int na, size_t sa_in, size_t sz_in,
                                                       size_t si_out, size_t sz_out, Perm perm,
                                                       int aligned, int plane_fast)
{
    __shared__ __attribute__((aligned(16))) float tile[kT * lds_row<NC>()];
    const int i0 = (plane_fast ? blockIdx.y : blockIdx.x) * kT;
    const int a0 = (plane_fast ? blockIdx.z : blockIdx.y) * kT;
    const size_t zz = plane_fast ? blockIdx.x : blockIdx.z;
    const int ti = ni - i0 < kT ? ni - i0 : kT;   // cells present along i
    const int ta = na - a0 < kT ? na - a0 : kT;   // cells present along a
    if (aligned && ti == kT && ta == kT)
        tile_fast<NC>(out, in, tile, i0, a0, zz, sa_in, sz_in, si_out, sz_out, perm);
    else
        tile_scalar<NC>(out, in, tile, i0, a0, ti, ta, zz, sa_in, sz_in, si_out, sz_out, perm);
}

template <int NC>
int launch_one(float *out, const float *in, int ni, int na, int nz_planes, size_t sa_in,
               size_t sz_in, size_t si_out, size_t sz_out, const int *p, hipStream_t s,
               bool plane_fast = false)
{
    if (nz_planes > 65535) return MHD_ERR_ARG;
    Perm perm = {};
    for (int c = 0; c < NC; c++) {
        perm.p[c] = p[c];
        perm.inv[p[c]] = c;
    }
    // 16-byte vectors need every row start 16-byte aligned: strides multiple of 4 cells
    const int aligned = ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0) && sa_in % 4 == 0 &&
                        sz_in % 4 == 0 && si_out % 4 == 0 && sz_out % 4 == 0;
    const unsigned ti = (ni + kT - 1) / kT, ta = (na + kT - 1) / kT;
    if (plane_fast && (ti > 65535 || ta > 65535)) return MHD_ERR_ARG;
    const dim3 grid = plane_fast ? dim3(nz_planes, ti, ta) : dim3(ti, ta, nz_planes);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL((transpose_kernel<NC>), grid, dim3(kTB), 0, s, out, in, ni, na, sa_in,
                       sz_in, si_out, sz_out, perm, aligned, (int)plane_fast);
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
    int rc = launch_one<5>(u_out, uu, nx, ny, nz, (size_t)nx, plane, (size_t)ny, plane, kPermU12, s);
    if (rc) return rc;
    return launch_one<3>(b_out, bu, nx, ny, nz, (size_t)nx, plane, (size_t)ny, plane, kPermB12, s);
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
    int rc = launch_one<5>(u_out, uu, nx, nz, ny, sa_in, sz_in, si_out, sz_out, kPermU13, s, true);
    if (rc) return rc;
    return launch_one<3>(b_out, bu, nx, nz, ny, sa_in, sz_in, si_out, sz_out, kPermB13, s, true);
}

}  // namespace mhd
