// mhd_transpose_body.h -- transpose12 / transpose13: axis swaps with a component
// permutation, so that every sweep runs along the contiguous axis (device
// code shared by the single-box and the ensemble kernels).
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
// 16-byte global vectors, nontemporal (the output is not read again in the
// launch; measured 1-2 ms faster per 1024^3 step).  Partial tiles and
// unaligned shapes take a scalar path.  Pure data movement: bit-exact by
// construction.
#pragma once
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
        const v4f w = *reinterpret_cast<const v4f *>(tile + i * RS + 4 * c4);
        __builtin_nontemporal_store(w, reinterpret_cast<v4f *>(out + (obase + (size_t)i * si_out) * NC + 4 * c4));
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

}  // namespace
}  // namespace mhd
