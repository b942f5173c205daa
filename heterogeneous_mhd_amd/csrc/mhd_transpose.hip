// mhd_transpose.hip -- transpose12 / transpose13: axis swaps with a component
// permutation, so that every sweep runs along the contiguous axis.
//
// Reference: transpose12 C/transpose12.c:7-47 (out dims (ny,nx,nz); u comps
// (1,3,2,4,5), b comps (2,1,3)), transpose13 C/transpose13.c:9-62 (needs
// nx == nz; out dims (nz,ny,nx); u comps (1,4,3,2,5), b comps (3,2,1)).
//
// Both are one kernel: for every plane z, the (i, a) matrix of NC-float cells
// is transposed through an LDS tile of T x T cells.  Input rows (fixed a,
// running i) and output rows (fixed i, running a) are both T*NC contiguous
// floats, so loads and stores are coalesced; the LDS row stride T*NC+1 is odd,
// so the column-wise reads of the store phase are free of bank conflicts.
// Pure data movement: bit-exact by construction.
#include "mhd_internal.h"

namespace mhd {
namespace {

constexpr int kT = 32;        // tile edge in cells
constexpr int kTB = 256;      // threads per workgroup

struct Perm {
    int p[5];
};

// in cell  (i, a, z) = i + a*sa_in  + z*sz_in
// out cell (a, i, z) = a + i*si_out + z*sz_out;  out[c] = in[perm[c]]
// FULL: the tile is kT x kT cells, so row lengths are compile-time constants
// and the per-element index split is a multiply-shift, not a division.
template <int NC, bool FULL>
__device__ __forceinline__ void transpose_tile(float *__restrict__ out,
                                               const float *__restrict__ in, float *tile, int i0,
                                               int a0, int ti, int ta, size_t zz, size_t sa_in,
                                               size_t sz_in, size_t si_out, size_t sz_out,
                                               const Perm &perm)
{
    constexpr int RS = kT * NC + 1;        // LDS row stride (odd)
    // load: ta rows of ti*NC contiguous floats
    {
        const int rowlen = FULL ? kT * NC : ti * NC;
        const int total = FULL ? kT * kT * NC : ta * rowlen;
        const size_t base = (size_t)i0 + (size_t)a0 * sa_in + zz * sz_in;
        for (int f = threadIdx.x; f < total; f += kTB) {
            const int a = f / rowlen, col = f - a * rowlen;
            tile[a * RS + col] = in[(base + (size_t)a * sa_in) * NC + col];
        }
    }
    __syncthreads();
    // store: ti rows (one per i) of ta*NC contiguous floats, permuted comps
    {
        const int rowlen = FULL ? kT * NC : ta * NC;
        const int total = FULL ? kT * kT * NC : ti * rowlen;
        const size_t base = (size_t)a0 + (size_t)i0 * si_out + zz * sz_out;
        for (int f = threadIdx.x; f < total; f += kTB) {
            const int i = f / rowlen, rem = f - i * rowlen;
            const int a = rem / NC, c = rem - a * NC;
            out[(base + (size_t)i * si_out) * NC + rem] = tile[a * RS + i * NC + perm.p[c]];
        }
    }
}

template <int NC>
__global__ __launch_bounds__(kTB) void transpose_kernel(float *__restrict__ out,
                                                       const float *__restrict__ in, int ni,
                                                       int na, size_t sa_in, size_t sz_in,
                                                       size_t si_out, size_t sz_out, Perm perm)
{
    __shared__ float tile[kT * (kT * NC + 1)];
    const int i0 = blockIdx.x * kT, a0 = blockIdx.y * kT;
    const size_t zz = blockIdx.z;
    const int ti = ni - i0 < kT ? ni - i0 : kT;   // cells present along i
    const int ta = na - a0 < kT ? na - a0 : kT;   // cells present along a
    if (ti == kT && ta == kT)
        transpose_tile<NC, true>(out, in, tile, i0, a0, ti, ta, zz, sa_in, sz_in, si_out, sz_out, perm);
    else
        transpose_tile<NC, false>(out, in, tile, i0, a0, ti, ta, zz, sa_in, sz_in, si_out, sz_out, perm);
}

template <int NC>
int launch_one(float *out, const float *in, int ni, int na, int nz_planes, size_t sa_in,
               size_t sz_in, size_t si_out, size_t sz_out, Perm perm, hipStream_t s)
{
    if (nz_planes > 65535) return MHD_ERR_ARG;
    dim3 grid((ni + kT - 1) / kT, (na + kT - 1) / kT, nz_planes);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL((transpose_kernel<NC>), grid, dim3(kTB), 0, s, out, in, ni, na, sa_in,
                       sz_in, si_out, sz_out, perm);
    return hipGetLastError() == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

}  // namespace

int launch_transpose12(float *u_out, float *b_out, const float *uu, const float *bu, int nx, int ny,
                       int nz, hipStream_t s)
{
    if (!u_out || !b_out || !uu || !bu || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    const size_t plane = (size_t)nx * ny;
    const Perm pu = {{0, 2, 1, 3, 4}}, pb = {{1, 0, 2, 0, 0}};
    int rc = launch_one<5>(u_out, uu, nx, ny, nz, (size_t)nx, plane, (size_t)ny, plane, pu, s);
    if (rc) return rc;
    return launch_one<3>(b_out, bu, nx, ny, nz, (size_t)nx, plane, (size_t)ny, plane, pb, s);
}

int launch_transpose13(float *u_out, float *b_out, const float *uu, const float *bu, int nx, int ny,
                       int nz, hipStream_t s)
{
    if (!u_out || !b_out || !uu || !bu || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    if (nx != nz) return MHD_ERR_SHAPE;
    // i = x, a = z, plane index = y.  in cell = i + z*nx*ny + y*nx;
    // out dims (nz, ny, nx): out cell = z + y*nz + i*nz*ny.
    const Perm pu = {{0, 3, 2, 1, 4}}, pb = {{2, 1, 0, 0, 0}};
    const size_t sa_in = (size_t)nx * ny, sz_in = (size_t)nx;
    const size_t si_out = (size_t)nz * ny, sz_out = (size_t)nz;
    int rc = launch_one<5>(u_out, uu, nx, nz, ny, sa_in, sz_in, si_out, sz_out, pu, s);
    if (rc) return rc;
    return launch_one<3>(b_out, bu, nx, nz, ny, sa_in, sz_in, si_out, sz_out, pb, s);
}

}  // namespace mhd
