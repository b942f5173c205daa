// mhd_transpose.hip -- transpose12 / transpose13: axis swaps with a component
// permutation, so that every sweep runs along the contiguous axis.
//
// Reference: transpose12 C/transpose12.c:7-47 (out dims (ny,nx,nz); u comps
// (1,3,2,4,5), b comps (2,1,3)), transpose13 C/transpose13.c:9-62 (needs
// nx == nz; out dims (nz,ny,nx); u comps (1,4,3,2,5), b comps (3,2,1)).
//
// Both are one kernel: for every plane z, the (i, a) matrix of NC-float cells
// is transposed through an LDS tile of 32 x 32 cells.  Input rows (fixed a,
// running i) and output rows (fixed i, running a) are both 32*NC contiguous
// floats.  On the fast path (row starts 16-byte aligned, full tile) every
// thread issues all of its 16-byte global loads before touching LDS, so a
// workgroup keeps its whole 20 KB (u) or 12 KB (b) tile in flight; the
// component permutation is applied while scattering into LDS, which is laid
// out in output order, so the store phase reads 16-byte LDS vectors and writes
// 16-byte global vectors.  Partial tiles and unaligned shapes take a scalar
// path.  Pure data movement: bit-exact by construction.
#include "mhd_internal.h"

namespace mhd {
namespace {

constexpr int kT = 32;        // tile edge in cells
constexpr int kTB = 256;      // threads per workgroup

struct Perm {
    int p[5];   // out comp c <- in comp p[c]
    int inv[5]; // in comp c -> out comp inv[c]
};

template <int NC>
constexpr int lds_row() { return kT * NC + 4; }  // output-ordered row, 16-byte multiple

// in cell  (i, a, z) = i + a*sa_in  + z*sz_in
// out cell (a, i, z) = a + i*si_out + z*sz_out;  out[c] = in[perm[c]]
template <int NC>
__device__ __forceinline__ void tile_fast(float *__restrict__ out, const float *__restrict__ in,
                                          float *tile, int i0, int a0, size_t zz, size_t sa_in,
                                          size_t sz_in, size_t si_out, size_t sz_out,
                                          const Perm &perm)
{
    constexpr int R4 = kT * NC / 4;          // float4 per row
    constexpr int TOT = kT * R4;             // float4 per tile
    constexpr int PER = TOT / kTB;           // float4 per thread
    static_assert(TOT % kTB == 0, "tile must split evenly");
    constexpr int RS = lds_row<NC>();
    const size_t ibase = (size_t)i0 + (size_t)a0 * sa_in + zz * sz_in;
    float4 v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int f = threadIdx.x + k * kTB;
        const int a = f / R4, c4 = f - a * R4;
        v[k] = *reinterpret_cast<const float4 *>(in + (ibase + (size_t)a * sa_in) * NC + 4 * c4);
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int f = threadIdx.x + k * kTB;
        const int a = f / R4, c4 = f - a * R4;
        const float e[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int col = 4 * c4 + q;            // input float within row a
            const int i = col / NC, c = col - i * NC;
            tile[i * RS + a * NC + perm.inv[c]] = e[q];
        }
    }
    __syncthreads();
    const size_t obase = (size_t)a0 + (size_t)i0 * si_out + zz * sz_out;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int f = threadIdx.x + k * kTB;
        const int i = f / R4, c4 = f - i * R4;
        const float4 w = *reinterpret_cast<const float4 *>(tile + i * RS + 4 * c4);
        *reinterpret_cast<float4 *>(out + (obase + (size_t)i * si_out) * NC + 4 * c4) = w;
    }
}

template <int NC>
__device__ __forceinline__ void tile_scalar(float *__restrict__ out, const float *__restrict__ in,
                                            float *tile, int i0, int a0, int ti, int ta, size_t zz,
                                            size_t sa_in, size_t sz_in, size_t si_out,
                                            size_t sz_out, const Perm &perm)
{
    constexpr int RS = lds_row<NC>();
    {
        const int rowlen = ti * NC, total = ta * rowlen;
        const size_t base = (size_t)i0 + (size_t)a0 * sa_in + zz * sz_in;
        for (int f = threadIdx.x; f < total; f += kTB) {
            const int a = f / rowlen, col = f - a * rowlen;
            const int i = col / NC, c = col - i * NC;
            tile[i * RS + a * NC + perm.inv[c]] = in[(base + (size_t)a * sa_in) * NC + col];
        }
    }
    __syncthreads();
    {
        const int rowlen = ta * NC, total = ti * rowlen;
        const size_t base = (size_t)a0 + (size_t)i0 * si_out + zz * sz_out;
        for (int f = threadIdx.x; f < total; f += kTB) {
            const int i = f / rowlen, rem = f - i * rowlen;
            out[(base + (size_t)i * si_out) * NC + rem] = tile[i * RS + rem];
        }
    }
}

template <int NC>
__global__ __launch_bounds__(kTB) void transpose_kernel(float *__restrict__ out,
                                                       const float *__restrict__ in, int ni,
                                                       int na, size_t sa_in, size_t sz_in,
                                                       size_t si_out, size_t sz_out, Perm perm,
                                                       int aligned)
{
    __shared__ __attribute__((aligned(16))) float tile[kT * lds_row<NC>()];
    const int i0 = blockIdx.x * kT, a0 = blockIdx.y * kT;
    const size_t zz = blockIdx.z;
    const int ti = ni - i0 < kT ? ni - i0 : kT;   // cells present along i
    const int ta = na - a0 < kT ? na - a0 : kT;   // cells present along a
    if (aligned && ti == kT && ta == kT)
        tile_fast<NC>(out, in, tile, i0, a0, zz, sa_in, sz_in, si_out, sz_out, perm);
    else
        tile_scalar<NC>(out, in, tile, i0, a0, ti, ta, zz, sa_in, sz_in, si_out, sz_out, perm);
}

template <int NC>
int launch_one(float *out, const float *in, int ni, int na, int nz_planes, size_t sa_in,
               size_t sz_in, size_t si_out, size_t sz_out, const int *p, hipStream_t s)
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
    dim3 grid((ni + kT - 1) / kT, (na + kT - 1) / kT, nz_planes);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL((transpose_kernel<NC>), grid, dim3(kTB), 0, s, out, in, ni, na, sa_in,
                       sz_in, si_out, sz_out, perm, aligned);
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
                       int nz, hipStream_t s)
{
    if (!u_out || !b_out || !uu || !bu || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    if (nx != nz) return MHD_ERR_SHAPE;
    // i = x, a = z, plane index = y.  in cell = i + z*nx*ny + y*nx;
    // out dims (nz, ny, nx): out cell = z + y*nz + i*nz*ny.
    const size_t sa_in = (size_t)nx * ny, sz_in = (size_t)nx;
    const size_t si_out = (size_t)nz * ny, sz_out = (size_t)nz;
    int rc = launch_one<5>(u_out, uu, nx, nz, ny, sa_in, sz_in, si_out, sz_out, kPermU13, s);
    if (rc) return rc;
    return launch_one<3>(b_out, bu, nx, nz, ny, sa_in, sz_in, si_out, sz_out, kPermB13, s);
}

}  // namespace mhd
