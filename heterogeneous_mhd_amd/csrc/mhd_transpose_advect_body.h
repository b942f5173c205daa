// mhd_transpose_advect_body.h -- a transpose and the advectbyzx the reference
// calls right after it, as one kernel: C/mhd.c:102 then :105 (transpose13,
// advectbyzx along y) and :112 then :115 (transpose12, advectbyzx along x).
//
// The mirror image of mhd_advect_transpose_body.h.  There the sweep axis is the
// input's contiguous axis and the output is written transposed; here the
// sweep axis is the first axis of the OUTPUT, an axis of the input with a
// large stride, and the partner axis (the one the transpose swaps with it) is
// the input's contiguous axis.  Everything below is said in the output frame
// (dims ns x n2 x n3, sweep axis first): the partner axis p (size np) and the
// remaining axis q (size nq) are the output's second and third axes in some
// order.  A workgroup owns a tile of kTI cells along the sweep axis (a window
// of 64 with the 4-cell halos) and kTP pencils along p, at one q, and walks q:
//  * it reads the input row of every window cell at q + 1 -- the kTP
//    pencils' 8 floats are one contiguous run of the input -- with 16-byte
//    loads one step ahead, into LDS, plus rho, m1 (and the field) of the halo
//    pencils p0 - 1 and p0 + kTP;
//  * one wavefront per tvdb line reads its line's inputs down a column of
//    that LDS row (one lane per window cell) and evaluates it with the
//    arithmetic of mhd_at_line.h: S(p), the lines along p (shared by
//    neighbouring pencils), and N(q), the lines between q and q + 1, whose
//    predecessor N(q - 1) is the O line of row q;
//  * the transposed u of the row goes out at once, from the same LDS row;
//    the new b of row q follows one step later, when its four edge fluxes
//    are known, combined in the order A1, A2, B1, B2 apply them
//    (C/advectbyzx.c:17-20); each pencil's 56 cells are one contiguous run
//    of the output.
// The output is bitwise that of transpose12/13 followed by advectbyzx, in
// both modes.
//
// Which reference line is which: transpose12 (output dims (ns, np, nq), p =
// j, q = k): the S lines are the A lines (field b2), O and N the B lines (b3).
// transpose13 (output dims (ns, nq, np), p = k, q = j): S are the B lines
// (b3), O and N the A lines (b2).  In the input, whose first axis is p, the S
// field is component 0, the O/N field component 2 (transpose12) or 1
// (transpose13), and the output's b1 component 1 or 2.
//
// The launcher takes the fused form only for pencil blocks that tile the
// partner axis (np a multiple of kTP) and 16-byte aligned rows; the plan
// falls back to the separate transpose for other shapes.
#pragma once
#include "mhd_at_line.h"

namespace mhd {
namespace {
namespace ta_tile {

constexpr int kTI = 56;           // tile cells along the sweep axis
constexpr int kWin = kTI + 8;     // the window with its halos: one lane per cell
constexpr int kTP = 16;           // tile pencils along the partner axis
constexpr int kPad = kTI + 1;     // pitch of the per-line LDS arrays
constexpr int kRowF = kTP * 8;    // floats of one window cell's row: u of the kTP pencils, then b
constexpr int kRP = kRowF + 1;    // LDS pitch per window cell (odd: column reads are conflict-free)
constexpr int kV4 = kRowF / 4;    // 16-byte vectors per window cell (20 of u, 12 of b)
static_assert(kWin == 64, "one lane per window cell");

struct TaGeom {
    int ns, np, nq;  // output frame: sweep, partner and remaining extents
    int nti, npb;    // tiles along the sweep axis, pencil blocks along the partner axis
    int qlen, nqc;   // steps of q per workgroup, runs of q
    unsigned nwg;
};

__device__ __forceinline__ int twrap(int a, int n)
{
    a %= n;
    return a < 0 ? a + n : a;
}

// input component of output component c of u (C/transpose12.c:22-29,
// C/transpose13.c:24-31)
template <int WHICH>
__host__ __device__ constexpr int ta_perm_u(int c)
{
    if (WHICH == 12) return c == 1 ? 2 : (c == 2 ? 1 : c);
    return c == 1 ? 3 : (c == 3 ? 1 : c);
}

// input cell (in cells) of window cell I of pencil (p, q): transpose12 input
// dims (np, ns, nq), transpose13 (np, nq, ns)
template <int WHICH>
__device__ __forceinline__ size_t ta_in_cell(int I, int p, int q, const TaGeom &g)
{
    return (WHICH == 12 ? ((size_t)q * g.ns + I) : ((size_t)I * g.nq + q)) * g.np + p;
}

// output cell of (I, p, q): transpose12 output (ns, np, nq), transpose13 (ns, nq, np)
template <int WHICH>
__device__ __forceinline__ size_t ta_out_cell(int I, int p, int q, const TaGeom &g)
{
    return (WHICH == 12 ? ((size_t)q * g.np + p) : ((size_t)p * g.nq + q)) * g.ns + I;
}

// One row (one q) of the walk in registers: this thread's 16-byte vectors of
// the window cells' runs, and one halo value.
struct TaRow {
    static constexpr int kPer = kWin * kV4 / 512;
    v4f v[kPer];
    float h;
};

template <int BLOCK, bool FAST, int WHICH>
__device__ __forceinline__ void transpose_advect_tile(float *__restrict__ u_out, float *__restrict__ b_out,
                                            const float *__restrict__ uu, const float *__restrict__ bu,
                                            const TaGeom &g, unsigned w, float dt)
{
    static_assert(BLOCK == 64 * kTP / 2, "each wavefront owns two pencils of the tile");
    static_assert(TaRow::kPer * BLOCK == kWin * kV4, "the row's vectors split evenly");
    constexpr int pm = ta_perm_u<WHICH>(1);   // input component of the line's m1
    constexpr int cS = 0;                     // input component of the S lines' field
    constexpr int cO = WHICH == 12 ? 2 : 1;   // of the O and N lines' field
    constexpr int c1 = WHICH == 12 ? 1 : 2;   // of the output's b1
    // Double buffers: step q evaluates N(q) and S(q + 1) and writes the new b
    // of q, which needs S(q) and N(q - 1) (= O(q)) from the step before.
    __shared__ float s_eS[2][(kTP + 1) * kPad];  // edge flux F(i-1) of S(p0 .. p0 + kTP)
    __shared__ float s_fS[2][kTP * kPad];        // new field of S(p), p in the tile
    __shared__ float s_eN[2][kTP * kPad];        // the same for N(p)
    __shared__ float s_fN[2][kTP * kPad];
    __shared__ float s_row[kWin * kRP];          // row q + 1: [window cell][pencil u (5) ..., pencil b (3) ...]
    __shared__ float s_hl[2 * kWin];             // halo pencil p0 - 1: rho, m1
    __shared__ float s_hh[3 * kWin];             // halo pencil p0 + kTP: rho, m1, S field

    // tile w -> pencil block fastest, then i-tile, then run of q: the blocks
    // beside each other along p, whose rows hold each other's halo pencils,
    // run at the same time on one XCD and walk q together, so those halo
    // values come from L2 (transpose13+advect 19.1 -> 18.4 ms at 1024^3; the
    // advect+transpose tiles walk in the same order)
    const int p0 = (int)(w % (unsigned)g.npb) * kTP;
    const int rest = (int)(w / (unsigned)g.npb);
    const int it = rest % g.nti;
    const int qa = (rest / g.nti) * g.qlen;
    const int nsteps = min(g.qlen, g.nq - qa);
    const int i0 = it * kTI;
    const int ns = g.ns, np = g.np, nq = g.nq;
    const int nti = min(kTI, ns - i0);         // cells of the tile along the sweep axis
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const bool keep = lane >= 4 && lane < 4 + nti;
    const int pl[2] = {wave, wave + kTP / 2};  // this wavefront's tile pencils
    const int ph_lo = twrap(p0 - 1, np), ph_hi = twrap(p0 + kTP, np);
    auto qw = [nq](int q) { return twrap(q, nq); };

    // ---- the loads of one row: vector f = threadIdx + r * BLOCK is vector
    // f % kV4 of window cell f / kV4 (u of the tile's pencils, then their b);
    // threads 0 .. 5 * kWin - 1 also load one halo value
    int vx[TaRow::kPer], vv[TaRow::kPer];
#pragma unroll
    for (int r = 0; r < TaRow::kPer; r++) {
        const int f = threadIdx.x + r * BLOCK;
        vx[r] = f / kV4;
        vv[r] = f % kV4;
    }
    const int ht = threadIdx.x;
    const bool hlo = ht < 2 * kWin, hhi = !hlo && ht < 5 * kWin;
    const int hx = hlo ? ht >> 1 : (ht - 2 * kWin) / 3;
    const int hw = hlo ? ht & 1 : (ht - 2 * kWin) % 3;  // 0 rho, 1 m1, 2 the S field
    auto load = [&](TaRow &x, int q) {
#pragma unroll
        for (int r = 0; r < TaRow::kPer; r++) {
            const int I = twrap(i0 - 4 + vx[r], ns);
            const size_t c0 = ta_in_cell<WHICH>(I, p0, q, g);
            x.v[r] = vv[r] < kTP * 5 / 4 ? *reinterpret_cast<const v4f *>(uu + c0 * 5 + 4 * vv[r])
                                          : *reinterpret_cast<const v4f *>(bu + c0 * 3 + 4 * (vv[r] - kTP * 5 / 4));
        }
        x.h = 0.0f;
        if (hlo || hhi) {
            const int I = twrap(i0 - 4 + hx, ns);
            const size_t c = ta_in_cell<WHICH>(I, hlo ? ph_lo : ph_hi, q, g);
            x.h = hw == 2 ? bu[c * 3 + cS] : uu[c * 5 + (hw == 0 ? 0 : pm)];
        }
    };
    auto stage = [&](const TaRow &x) {
#pragma unroll
        for (int r = 0; r < TaRow::kPer; r++) {
            float *d = s_row + vx[r] * kRP + 4 * vv[r];
            d[0] = x.v[r].x;
            d[1] = x.v[r].y;
            d[2] = x.v[r].z;
            d[3] = x.v[r].w;
        }
        if (hlo) s_hl[hw * kWin + hx] = x.h;
        if (hhi) s_hh[hw * kWin + hx] = x.h;
    };
    // one tvdb line into the edge-flux array e and (unless null) the field array f
    auto line = [&](float rl, float ml, float rh, float mh, float bl, float *e, float *f, int slot) {
        const LineIn in = {rl, ml, rh, mh, bl};
        float fm;
        const float bnew = at_line<FAST>(in, dt, fm);
        if (keep) {
            e[slot * kPad + lane - 4] = fm;
            if (f) f[slot * kPad + lane - 4] = bnew;
        }
    };
    const float *lrow = s_row + lane * kRP;  // this lane's window cell
    auto ru = [&](int p, int c) { return lrow[p * 5 + c]; };
    auto rb = [&](int p, int c) { return lrow[kTP * 5 + p * 3 + c]; };

    // ---- prologue: rho, m1 of this wavefront's pencils at qa - 1 (for N(qa - 1)),
    // and row qa into registers
    float rq[2], mq[2], b1q[2] = {0.0f, 0.0f};
    {
        const int I = twrap(i0 - 4 + lane, ns);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const size_t c = ta_in_cell<WHICH>(I, p0 + pl[k], qw(qa - 1), g);
            rq[k] = uu[c * 5];
            mq[k] = uu[c * 5 + pm];
        }
    }
    TaRow ra, rb2;
    load(ra, qa);
    int sb = 0, nb = 0;  // s_eS[sb]: S(q); s_eN[nb]: N(q), s_eN[nb ^ 1]: N(q - 1)
    const bool vec_out = ns % 4 == 0 && (uintptr_t)u_out % 16 == 0;

    // For transpose12 the new b of row q is stored at the start of step
    // q + 1, after the look-ahead loads: the next wait for the vector memory
    // counter (in order) then comes a step later, not right behind the
    // stores (transpose12+advect 5.6% faster at 1024^3, same process).  For
    // transpose13 the same order was 3.2% slower, so it stores at once.
    constexpr bool kDefer = WHICH == 12;
    float bst[2][3];
    int bq = -1;  // row of the pending b stores, -1 for none
    auto store_b = [&]() {
        if (bq >= 0 && keep) {
            const int il = lane - 4;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                float *o = b_out + ta_out_cell<WHICH>(i0 + il, p0 + pl[k], bq, g) * 3;
                __builtin_nontemporal_store(bst[k][0], o);
                __builtin_nontemporal_store(bst[k][1], o + 1);
                __builtin_nontemporal_store(bst[k][2], o + 2);
            }
        }
        bq = -1;
    };

    // One step at q = qa + step: cur holds row q + 1, ahead receives row q + 2.
    auto step_fn = [&](int step, TaRow &cur, TaRow &ahead) {
        const int q = qw(qa + step), q1 = qw(qa + step + 1);
        stage(cur);
        if (step + 2 <= nsteps) load(ahead, qw(q + 2));
        store_b();
        __syncthreads();  // row q + 1 is in LDS
        float b1n[2];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int p = pl[k];
            const float rh = ru(p, 0), mh = ru(p, pm);
            // N(q): pencils (p, q), (p, q + 1); field of (p, q + 1)
            line(rq[k], mq[k], rh, mh, rb(p, cO), s_eN[nb], s_fN[nb], p);
            // S(q + 1): pencils (p - 1, q + 1), (p, q + 1); field of (p, q + 1)
            if (step + 1 < nsteps) {
                const float rl = p == 0 ? s_hl[lane] : ru(p - 1, 0);
                const float ml = p == 0 ? s_hl[kWin + lane] : ru(p - 1, pm);
                line(rl, ml, rh, mh, rb(p, cS), s_eS[sb ^ 1], s_fS[sb ^ 1], p);
            }
            rq[k] = rh;
            mq[k] = mh;
            b1n[k] = rb(p, c1);
        }
        if (step + 1 < nsteps && wave == kTP / 2 - 1)  // S(p0 + kTP), for the last pencil's b1
            line(ru(kTP - 1, 0), ru(kTP - 1, pm), s_hh[lane], s_hh[kWin + lane], s_hh[2 * kWin + lane],
                 s_eS[sb ^ 1], nullptr, kTP);
        // u of row q + 1, transposed: pencil p's cells i0 .. i0 + nti - 1 are one run
        if (step + 1 < nsteps) {
            const int nv = nti * 5 / 4;  // 16-byte vectors per pencil (nti is a multiple of 4)
            if (vec_out) {
                for (int f = threadIdx.x; f < kTP * nv; f += BLOCK) {
                    const int p = f / nv, v = f % nv;
                    float e[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int idx = 4 * v + j, il = idx / 5, c = idx % 5;
                        e[j] = s_row[(il + 4) * kRP + p * 5 + ta_perm_u<WHICH>(c)];
                    }
                    v4f o;
                    o.x = e[0];
                    o.y = e[1];
                    o.z = e[2];
                    o.w = e[3];
                    __builtin_nontemporal_store(o, reinterpret_cast<v4f *>(u_out + ta_out_cell<WHICH>(i0, p0 + p, q1, g) * 5) + v);
                }
            } else {
                for (int f = threadIdx.x; f < kTP * nti * 5; f += BLOCK) {
                    const int p = f / (nti * 5), idx = f % (nti * 5), il = idx / 5, c = idx % 5;
                    u_out[ta_out_cell<WHICH>(i0, p0 + p, q1, g) * 5 + idx] =
                        s_row[(il + 4) * kRP + p * 5 + ta_perm_u<WHICH>(c)];
                }
            }
        }
        __syncthreads();  // every line of the step is in LDS; row q + 1 is read
        if (step >= 0 && keep) {
            // new b of row q: b1 from the four edge fluxes, b2 and b3 from the lines
            const float *eS = s_eS[sb], *fS = s_fS[sb];
            const float *eO = s_eN[nb ^ 1], *fO = s_fN[nb ^ 1], *eN = s_eN[nb];
            const int il = lane - 4;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int p = pl[k];
                const float eS0 = eS[p * kPad + il], eS1 = eS[(p + 1) * kPad + il];
                const float eo = eO[p * kPad + il], en = eN[p * kPad + il];
                const float fs = fS[p * kPad + il], fo = fO[p * kPad + il];
                if constexpr (WHICH == 12) {
                    // A1, A2 (the S lines), then B1, B2 (O, N); out (b1, b2 = S, b3 = O)
                    bst[k][0] = (((b1q[k] - eS0) + eS1) - eo) + en;
                    bst[k][1] = fs;
                    bst[k][2] = fo;
                } else {
                    // A1, A2 (O, N), then B1, B2 (the S lines); out (b1, b2 = O, b3 = S)
                    bst[k][0] = (((b1q[k] - eo) + en) - eS0) + eS1;
                    bst[k][1] = fo;
                    bst[k][2] = fs;
                }
            }
        }
        if (step >= 0) bq = q;
        if constexpr (!kDefer) store_b();
        // the next step's LDS writes come after its first barrier, which
        // every thread reaches only after the reads above
        b1q[0] = b1n[0];
        b1q[1] = b1n[1];
        sb ^= 1;
        nb ^= 1;
    };
    // the rows alternate between ra and rb2 instead of being copied
    for (int step = -1; step < nsteps; step += 2) {
        step_fn(step, ra, rb2);
        if (step + 1 < nsteps) step_fn(step + 1, rb2, ra);
    }
    store_b();
}

}  // namespace ta_tile
}  // namespace
}  // namespace mhd
