// mhd_advect_transpose_body.h -- advectbyzx followed by the transpose the
// reference calls next, as one kernel: C/mhd.c:74 then :78 (advectbyzx along
// x, transpose12) and :84 then :88 (advectbyzx along y, transpose13).
//
// Unlike fluidx, advectbyzx has no pencil-global quantity: the new b of cell
// i depends only on cells i-4 .. i+4 of the tvdb lines it reads
// (C/advectbyzxA1.c:34-57, C/tvdb.c:26-88).  So it can be evaluated on tiles
// of kTI cells along the sweep axis with a 4-cell halo on each side.  A tile
// of kTP pencils along the axis the transpose swaps with the sweep axis (the
// partner axis: j for transpose12, k for transpose13) is what an LDS-tiled
// transpose writes: kTI output rows of kTP cells.  One workgroup owns such a
// tile at one value q of the remaining axis, and
//  * reads the tile's u once (the transposed u, and with the lines' other
//    inputs the face velocities): the separate calls read u twice, the
//    advect's reads of whole 20-byte cells for rho and m1 included, and b
//    twice, with a write of b in between;
//  * evaluates the 3 kTP + 1 tvdb lines the tile needs, one line per
//    wavefront with one lane per cell of the 64-cell window, stencil
//    neighbours by DPP lane shifts:
//      S(p), p = p0 .. p0 + kTP  along the partner axis (shared by the
//                                neighbouring pencils p - 1 and p of a tile);
//      O(p), N(p), p in the tile along the other axis, at q and q + 1;
//  * combines them into the new b of each cell, in the order A1, A2, B1, B2
//    apply the edge fluxes (C/advectbyzx.c:17-20), and writes u and b in
//    the transposed layout with the transpose's component permutation.
// Each line's arithmetic is advect_group's, expression for expression, in
// both modes, so the results are bitwise those of the two calls.
//
// Which reference line is which: for transpose12 (p = j, q = k) the A lines
// (C/advectbyzxA1.c, face velocity from pencils j-1, j; field b2) are S and
// the B lines (C/advectbyzxB1.c, k-1, k; b3) are O (at k) and N (at k+1).
// For transpose13 (p = k, q = j) S are the B lines and O, N the A lines.
#pragma once
#include "mhd_at_line.h"  // at_line, LineIn, limiter_fast, ld_at

namespace mhd {
namespace {
namespace at_tile {

constexpr int kTI = 56;           // tile cells along the sweep axis
constexpr int kWin = kTI + 8;     // the window with its halos: one lane per cell
constexpr int kTP = 16;           // tile pencils along the partner axis
constexpr int kLn = 3 * kTP + 1;  // tvdb lines per tile
constexpr int kPad = kTI + 1;     // LDS row pitch (odd: the pencil-fastest reads are conflict-free)
static_assert(kWin == 64, "one lane per window cell");

struct AtGeom {
    int n1, n2, n3;  // input dims, sweep axis first
    int np, nq;      // extents of the partner axis and of the remaining axis
    int nti, npb;    // tiles along the sweep axis, pencil blocks along the partner axis
    int qlen, nqc;   // steps of q per workgroup, runs of q
    unsigned nwg;
};

__device__ __forceinline__ int wrapi(int a, int n)
{
    a %= n;
    return a < 0 ? a + n : a;
}

// pencil index k * n2 + j of in-range partner coordinate p and remaining
// coordinate q
template <int WHICH>
__device__ __forceinline__ int at_pencil_in(int p, int q, const AtGeom &g)
{
    return WHICH == 12 ? q * g.n2 + p : p * g.n2 + q;
}

// output row (in cells / np) of input cell i at q: transpose12 out dims
// (n2, n1, n3), cell (k*n1 + i)*n2 + j; transpose13 out dims (n3, n2, n1),
// cell (i*n2 + j)*n3 + k
template <int WHICH>
__device__ __forceinline__ size_t at_out_row(int i, int q, const AtGeom &g)
{
    return WHICH == 12 ? (size_t)q * g.n1 + i : (size_t)i * g.nq + q;
}

// input component of output component c of u (C/transpose12.c:22-29,
// C/transpose13.c:24-31)
template <int WHICH>
__device__ __forceinline__ int at_perm_u(int c)
{
    if (WHICH == 12) return c == 1 ? 2 : (c == 2 ? 1 : c);
    return c == 1 ? 3 : (c == 3 ? 1 : c);
}

// Store of one value of an output row segment [seg0, seg1) (bytes).  A
// segment's first or last 128-byte line is shared with the segment of the
// neighbouring pencil block, which the workgroup beside it writes at about
// the same time.  Those lines get ordinary stores, which the L2 merges
// before it writes the line back; with nontemporal stores of both halves
// the tile wrote 38-42 instead of 32 bytes per cell (FP32, 1024^3).  The
// lines the segment covers alone get nontemporal stores (same traffic as
// ordinary stores throughout; with this walk order transpose13 gained 3.3%
// this way against 1.6% with ordinary stores, each against the same base).
template <class T>
__device__ __forceinline__ void st_seg(float *base, unsigned off, T v, uintptr_t seg0, uintptr_t seg1)
{
    T *const p = reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off);
    const uintptr_t line = reinterpret_cast<uintptr_t>(p) & ~(uintptr_t)127;
    if (line < seg0 || line + 128 > seg1)
        *p = v;
    else
        __builtin_nontemporal_store(v, p);
}

// Everything one step of the walk reads of the pencils at one q (row q of
// the tile's pencil block), issued two steps before its output: rho, m1 and
// both transverse fields over the window for the lines, the halo pencils
// p0 - 1 and p0 + kTP, b1 and the tile's u.  Each row is read once, in one
// go, so its cache lines are fetched from HBM once.
template <int BLOCK>
struct RowIn {
    static constexpr int kU = (kTP * kTI * 5 / 4 + BLOCK - 1) / BLOCK;
    static constexpr int kPer = (kTP * kTI + BLOCK - 1) / BLOCK;
    float r[2], m[2], bo[2], bs[2];  // this wavefront's pencils: rho, m1, field of O/N and of S lines
    float hr, hm, hb;                // a halo pencil (wavefronts 0 and kTP/2 - 1)
    v4f uv[kU];                      // the tile's u, 16-byte vectors
    float b1[kPer];                  // b1 of the tile's cells
};

template <int BLOCK, bool FAST, int WHICH>
__device__ __forceinline__ void advect_transpose_tile(float *__restrict__ u_out, float *__restrict__ b_out,
                                            const float *__restrict__ uu, const float *__restrict__ bu,
                                            const AtGeom &g, unsigned w, float dt)
{
    static_assert(BLOCK == 64 * kTP / 2, "each wavefront owns two pencils of the tile");
    constexpr int kCells = kTP * kTI;
    constexpr int kV = kTI * 5 / 4;            // 16-byte vectors of u per tile pencil
    using In = RowIn<BLOCK>;
    constexpr int cs = WHICH == 12 ? 1 : 2;    // field of the S lines (b2 / b3)
    constexpr int co = WHICH == 12 ? 2 : 1;    // field of the O and N lines (b3 / b2)
    // Double buffers: step q evaluates N(q) (between q and q + 1) and
    // S(q + 1), and writes the output of q, which needs S(q) and N(q - 1)
    // (= O(q)) from the step before.
    __shared__ float s_eS[2][(kTP + 1) * kPad];  // edge flux F(i-1) of S(p0 .. p0 + kTP)
    __shared__ float s_fS[2][kTP * kPad];        // new field of S(p), p in the tile
    __shared__ float s_eN[2][kTP * kPad];        // the same for N(p)
    __shared__ float s_fN[2][kTP * kPad];
    __shared__ float s_r[(kTP + 2) * kWin];      // rho of pencils p0 - 1 .. p0 + kTP at q + 1
    __shared__ float s_m[(kTP + 2) * kWin];      // m1, likewise
    __shared__ __attribute__((aligned(16))) float s_u[2][kCells * 5];  // u of the tile at q (pencil-major)

    // tile w -> column (pencil block fastest, then i-tile) and run of q:
    // the tiles that run side by side share their halo pencils and the
    // output lines at their block edges (4.8% / 3.3% faster for
    // transpose12 / 13 at 1024^3 than i-tile fastest, same process)
    const int p0 = (int)(w % (unsigned)g.npb) * kTP;
    const int rest = (int)(w / (unsigned)g.npb);
    const int it = rest % g.nti;
    const int qa = (rest / g.nti) * g.qlen;
    const int nsteps = min(g.qlen, g.nq - qa);
    const int i0 = it * kTI;
    const int n1 = g.n1, np = g.np, nq = g.nq;
    const int nti = min(kTI, n1 - i0);         // cells of the tile along the sweep axis
    const int ntp = min(kTP, np - p0);         // pencils of the tile
    const size_t row = (size_t)n1;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int ii = wrapi(i0 - 4 + lane, n1);   // this lane's cell of the window
    const unsigned ou = 20u * (unsigned)ii, ob = 12u * (unsigned)ii;
    const bool keep = lane >= 4 && lane < 4 + kTI;
    const int pl[2] = {wave, wave + kTP / 2};  // this wavefront's tile pencils
    // their partner coordinates; a partial block's missing pencils wrap (their
    // lines are evaluated, their output is not written)
    const int pk[2] = {wrapi(p0 + pl[0], np), wrapi(p0 + pl[1], np)};
    const int ph_lo = wrapi(p0 - 1, np), ph_hi = wrapi(p0 + kTP, np);
    const bool vec = n1 % 4 == 0 && nti == kTI && (uintptr_t)uu % 16 == 0;
    auto qw = [nq](int q) { return wrapi(q, nq); };  // q + 2 exceeds 2 nq when nq = 1
    // pencil index = pp(p) + qq(q) (k * n2 + j)
    auto pp = [&](int p) { return WHICH == 12 ? p : p * g.n2; };
    auto qq = [&](int q) { return WHICH == 12 ? q * g.n2 : q; };
    const size_t u_pen = row * 5, b_pen = row * 3;  // floats per pencil

    // ---- per-lane byte offsets, fixed for the walk, from per-step uniform
    // bases (the global_load / global_store form with a scalar base and a
    // 32-bit vector offset; the launcher checks they fit)
    // the tile's u, 16-byte vectors f = pl * kV + v
    unsigned off_u[In::kU];
    bool ok_u[In::kU];
#pragma unroll
    for (int r = 0; r < In::kU; r++) {
        const int f = threadIdx.x + r * BLOCK;
        const int p = f / kV, v = f % kV;
        ok_u[r] = vec && f < kTP * kV && p < ntp;
        off_u[r] = ok_u[r] ? (unsigned)((size_t)(pp(p0 + p) - pp(p0)) * u_pen * 4) + 16u * (unsigned)v : 0u;
    }
    // b1 and the new b of cell (p, il) of the tile: consecutive threads take
    // consecutive pencils of one output row
    const size_t ostride = WHICH == 12 ? (size_t)np : (size_t)nq * np;  // output cells from row il to il + 1
    unsigned off_b1[In::kPer], off_ob[In::kPer];
    bool ok_b[In::kPer];
#pragma unroll
    for (int r = 0; r < In::kPer; r++) {
        const int idx = threadIdx.x + r * BLOCK;
        const int p = idx % kTP, il = idx / kTP;
        ok_b[r] = idx < kCells && p < ntp && il < nti;
        off_b1[r] = ok_b[r] ? (unsigned)(((size_t)(pp(p0 + p) - pp(p0)) * b_pen + (size_t)il * 3) * 4) : 0u;
        off_ob[r] = ok_b[r] ? (unsigned)(((size_t)il * ostride + p) * 12) : 0u;
    }
    // uniform bases of row q
    auto u_row = [&](int p, int q) { return uu + (size_t)(pp(p) + qq(q)) * u_pen; };
    auto b_row = [&](int p, int q) { return bu + (size_t)(pp(p) + qq(q)) * b_pen; };
    // output base of row q: cell (q*n1 + i0)*np + p0 (transpose12), (i0*nq + q)*np + p0 (transpose13)
    auto out_cell = [&](int q) {
        return (WHICH == 12 ? ((size_t)q * n1 + i0) : ((size_t)i0 * nq + q)) * np + p0;
    };

    // every load of row q (outputs: whether its u and b1 are needed)
    auto load = [&](In &x, int q, bool outputs) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const float *u0 = u_row(pk[k], q);
            const float *b0 = b_row(pk[k], q);
            x.r[k] = ld_at<float>(u0, ou);
            x.m[k] = ld_at<float>(u0, ou + 4u);
            x.bo[k] = ld_at<float>(b0, ob + 4u * co);
            x.bs[k] = ld_at<float>(b0, ob + 4u * cs);
        }
        x.hr = x.hm = x.hb = 0.0f;
        if (wave == 0 || wave == kTP / 2 - 1) {
            const int ph = wave == 0 ? ph_lo : ph_hi;
            const float *uh = u_row(ph, q);
            x.hr = ld_at<float>(uh, ou);
            x.hm = ld_at<float>(uh, ou + 4u);
            x.hb = ld_at<float>(b_row(ph, q), ob + 4u * cs);
        }
        const float *ut = u_row(p0, q) + (size_t)i0 * 5;
        const float *bt = b_row(p0, q) + (size_t)i0 * 3;
#pragma unroll
        for (int r = 0; r < In::kU; r++)
            if (outputs && ok_u[r]) x.uv[r] = ld_at<v4f>(ut, off_u[r]);
#pragma unroll
        for (int r = 0; r < In::kPer; r++) x.b1[r] = outputs && ok_b[r] ? ld_at<float>(bt, off_b1[r]) : 0.0f;
    };
    // the tile's u of row q into s_u[buf]
    auto stage_u = [&](const In &x, int buf, int q) {
        if (vec) {
#pragma unroll
            for (int r = 0; r < In::kU; r++) {
                const int f = threadIdx.x + r * BLOCK;
                if (ok_u[r]) ((lds_v4 *)s_u[buf])[f] = x.uv[r];
            }
        } else {
            for (int f = threadIdx.x; f < kCells * 5; f += BLOCK) {
                const int p = f / (kTI * 5), e = f % (kTI * 5);
                if (p < ntp && e < nti * 5) s_u[buf][f] = u_row(p0 + p, q)[(size_t)i0 * 5 + e];
            }
        }
    };
    // one tvdb line into the edge-flux array e and (unless null) the field array f
    auto line = [&](float rl, float ml, float rh, float mh, float bl, lds_f *e, lds_f *f, int slot) {
        const LineIn in = {rl, ml, rh, mh, bl};
        float fm;
        const float bnew = at_line<FAST>(in, dt, fm);
        if (keep) {
            e[slot * kPad + lane - 4] = fm;
            if (f) f[slot * kPad + lane - 4] = bnew;
        }
    };

    // ---- prologue: rows qa - 1 and qa; S(qa) and N(qa - 1) come from a
    // warm-up step
    float rq[2], mq[2];  // rho, m1 of this wavefront's pencils at the step's q
    {
        const int qm = qw(qa - 1);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            rq[k] = ld_at<float>(u_row(pk[k], qm), ou);
            mq[k] = ld_at<float>(u_row(pk[k], qm), ou + 4u);
        }
    }
    In ra, rb;  // rows q + 1 and q + 2 of a step, in alternating roles
    load(ra, qa, true);
    float b1q[In::kPer];  // b1 of the step's row (from the load two steps back)
    stage_u(ra, qa & 1, qa);
    int sb = 0, nb = 0;  // s_eS[sb]: S(q); s_eN[nb]: N(q), s_eN[nb ^ 1]: N(q - 1)

    // One step at q = qa + step: cur holds row q + 1, ahead receives row q + 2.
    auto step_fn = [&](int step, In &cur, In &ahead) {
        const int q = qw(qa + step);
        const bool out = step >= 0;
        // look-ahead: every load of row q + 2 (its lines run next step, its output two steps on)
        const bool more = step + 2 <= nsteps;
        if (more) load(ahead, qw(q + 2), step + 2 < nsteps);
        // rho, m1 at q + 1 of pencils p0 - 1 + r, r = 0 .. kTP + 1, for S(q + 1)
#pragma unroll
        for (int k = 0; k < 2; k++) {
            s_r[(1 + pl[k]) * kWin + lane] = cur.r[k];
            s_m[(1 + pl[k]) * kWin + lane] = cur.m[k];
        }
        if (wave == 0) {
            s_r[lane] = cur.hr;
            s_m[lane] = cur.hm;
        }
        if (wave == kTP / 2 - 1) {
            s_r[(kTP + 1) * kWin + lane] = cur.hr;
            s_m[(kTP + 1) * kWin + lane] = cur.hm;
        }
        __syncthreads();  // the S lines' inputs are in LDS
        // N(q): pencils (p, q), (p, q + 1); field of (p, q + 1)
#pragma unroll
        for (int k = 0; k < 2; k++)
            line(rq[k], mq[k], cur.r[k], cur.m[k], cur.bo[k], (lds_f *)s_eN[nb], (lds_f *)s_fN[nb], pl[k]);
        // S(q + 1): pencils (p - 1, q + 1), (p, q + 1); field of (p, q + 1)
        if (step + 1 < nsteps) {
#pragma unroll
            for (int k = 0; k < 2; k++)
                line(s_r[pl[k] * kWin + lane], s_m[pl[k] * kWin + lane], cur.r[k], cur.m[k], cur.bs[k],
                     (lds_f *)s_eS[sb ^ 1], (lds_f *)s_fS[sb ^ 1], pl[k]);
            if (wave == kTP / 2 - 1)  // S(p0 + kTP), for the last pencil's b1
                line(s_r[kTP * kWin + lane], s_m[kTP * kWin + lane], cur.hr, cur.hm, cur.hb, (lds_f *)s_eS[sb ^ 1],
                     nullptr, kTP);
        }
        __syncthreads();  // every line of the step is in LDS
        if (out) {
            // new b of row q, written transposed
            const lds_f *eS = (lds_f *)s_eS[sb], *fS = (lds_f *)s_fS[sb];
            const lds_f *eO = (lds_f *)s_eN[nb ^ 1], *fO = (lds_f *)s_fN[nb ^ 1], *eN = (lds_f *)s_eN[nb];
            float *ob_base = b_out + out_cell(q) * 3;
#pragma unroll
            for (int r = 0; r < In::kPer; r++) {
                const int idx = threadIdx.x + r * BLOCK;
                const int p = idx % kTP, il = idx / kTP;
                if (ok_b[r]) {
                    const uintptr_t seg0 = reinterpret_cast<uintptr_t>(ob_base) + (uintptr_t)il * ostride * 12;
                    const uintptr_t seg1 = seg0 + (uintptr_t)ntp * 12;
                    const float eS0 = eS[p * kPad + il], eS1 = eS[(p + 1) * kPad + il];
                    const float eo = eO[p * kPad + il], en = eN[p * kPad + il];
                    const float fs = fS[p * kPad + il], fo = fO[p * kPad + il];
                    if constexpr (WHICH == 12) {
                        // b1: A1, A2 (the S lines), then B1, B2 (O, N); out (b2, b1, b3)
                        const float b1n = (((b1q[r] - eS0) + eS1) - eo) + en;
                        st_seg<float>(ob_base, off_ob[r], fs, seg0, seg1);
                        st_seg<float>(ob_base, off_ob[r] + 4u, b1n, seg0, seg1);
                        st_seg<float>(ob_base, off_ob[r] + 8u, fo, seg0, seg1);
                    } else {
                        // b1: A1, A2 (O, N), then B1, B2 (the S lines); out (b3, b2, b1)
                        const float b1n = (((b1q[r] - eo) + en) - eS0) + eS1;
                        st_seg<float>(ob_base, off_ob[r], fs, seg0, seg1);
                        st_seg<float>(ob_base, off_ob[r] + 4u, fo, seg0, seg1);
                        st_seg<float>(ob_base, off_ob[r] + 8u, b1n, seg0, seg1);
                    }
                }
            }
            // u written transposed with the component permutation
            // (transpose12: rho, m2, m1, m3, E; transpose13: rho, m3, m2, m1, E)
            const lds_f *su = (lds_f *)s_u[(qa + step) & 1];  // the run's q does not wrap
            auto src = [&](int il, int f) {
                const int p = f / 5, c = f % 5;
                return su[(p * kTI + il) * 5 + at_perm_u<WHICH>(c)];
            };
            float *ou_base = u_out + out_cell(q) * 5;
            if (ntp == kTP && np % 4 == 0 && (uintptr_t)u_out % 16 == 0) {
                constexpr int kO = kTP * 5 / 4;  // 16-byte vectors per output row
                for (int f = threadIdx.x; f < nti * kO; f += BLOCK) {
                    const int il = f / kO, v = f % kO;
                    v4f wv;
                    wv.x = src(il, 4 * v);
                    wv.y = src(il, 4 * v + 1);
                    wv.z = src(il, 4 * v + 2);
                    wv.w = src(il, 4 * v + 3);
                    const uintptr_t seg0 = reinterpret_cast<uintptr_t>(ou_base) + (uintptr_t)il * ostride * 20;
                    st_seg<v4f>(ou_base, (unsigned)((size_t)il * ostride * 20) + 16u * (unsigned)v, wv, seg0,
                                seg0 + kTP * 20);
                }
            } else {
                for (int f = threadIdx.x; f < nti * ntp * 5; f += BLOCK) {
                    const int il = f / (ntp * 5), e = f % (ntp * 5);
                    ou_base[(size_t)il * ostride * 5 + e] = src(il, e);
                }
            }
        }
        __syncthreads();  // the step's LDS reads are done
        // rotate: row q + 1 becomes the step's row, row q + 2 the look-ahead
#pragma unroll
        for (int k = 0; k < 2; k++) {
            rq[k] = cur.r[k];
            mq[k] = cur.m[k];
        }
#pragma unroll
        for (int r = 0; r < In::kPer; r++) b1q[r] = cur.b1[r];
        if (more && step + 2 < nsteps) stage_u(ahead, (qa + step + 2) & 1, qw(q + 2));  // read at step + 2
        sb ^= 1;
        nb ^= 1;
    };
    // the rows alternate between ra and rb instead of being copied
    for (int step = -1; step < nsteps; step += 2) {
        step_fn(step, ra, rb);
        if (step + 1 < nsteps) step_fn(step + 1, rb, ra);
    }
}

}  // namespace at_tile
}  // namespace
}  // namespace mhd
