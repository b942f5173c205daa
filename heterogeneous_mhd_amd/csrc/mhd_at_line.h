// mhd_at_line.h -- one tvdb line of advectbyzx held across a wavefront (one
// lane per cell of a 64-cell window, stencil neighbours by DPP lane shifts),
// shared by the fused advect+transpose and transpose+advect kernels.
//
// Reference: C/advectbyzxA1.c:34-46 (face velocity and its smoothing),
// C/tvdb.c:26-88 (half step, limited flux, update).  The arithmetic is
// advect_group's (mhd_advect_body.h), expression for expression, in both
// modes, so a line gives bitwise the values advectbyzx gives.
#pragma once
#include "mhd_fluid_body.h"  // limiter_fast, ld_at, st_at

namespace mhd {
namespace {

struct LineIn {
    float rl, ml, rh, mh, bl;  // rho, m1 of pencils lo, hi; the advected field
};

// Window neighbours by whole-wavefront DPP shifts (a VALU move, no LDS
// traffic): wave_shr:1 gives lane x the value of lane x - 1, wave_shl:1 that
// of lane x + 1.  The lanes at the ends receive 0; they are halo lanes whose
// values no kept lane depends on.
__device__ __forceinline__ float up1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float dn1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ float dn2(float v) { return dn1(dn1(v)); }

// One tvdb line (C/advectbyzxA1.c:34-46, C/tvdb.c:26-88) over the window,
// lane x <-> cell i0 - 4 + x.  Returns the new field; fm receives the edge
// flux F(i-1).  Valid for lanes 4 .. 4 + kTI - 1 (the halo lanes feed them).
template <bool FAST>
__device__ __forceinline__ float at_line(const LineIn &in, float dt, float &fm)
{
    const float z = 0.0f, h = 0.5f;
    // raw face velocity, A1 :34
    const float vx = FAST ? (in.ml + in.mh) * __builtin_amdgcn_rcpf(in.rl + in.rh) : (in.ml + in.mh) / (in.rl + in.rh);
    // smoothing, A1 :37-41
    const float s = up1(vx) + dn1(vx);
    const float vg = FAST ? 0.25f * fmaf(2.0f, vx, s) : (float)(((double)s + 2.0 * (double)vx) / 4.0);
    // tvdb :26-46: vh, b*vg, upwind flux1, half step b1
    const float vgp = dn1(vg);
    const float vh = FAST ? 0.5f * (vg + vgp) : h * vg + h * vgp + z;
    const float xb = in.bl * vg;
    const float xbp = dn1(xb), xbm = up1(xb), vhm = up1(vh);
    const float f1 = vh > 0.0f ? xb : xbp;
    const float f1m = vhm > 0.0f ? xbm : xb;
    float bh;
    if constexpr (FAST) {
        bh = fmaf(-0.5f * dt, f1 - f1m, in.bl);
    } else {
        const float d = (f1 - f1m) * dt;
        bh = (float)((double)in.bl - (double)d / 2.0);
    }
    // limited flux, tvdb :54-78
    const float bhm = up1(bh), bhp = dn1(bh), bhpp = dn2(bh);
    const float vgm = up1(vg), vgpp = dn2(vg);
    float fx;
    if constexpr (FAST) {
        const bool up = vh > 0.0f;
        const float w0 = vg * bh, w1 = vgp * bhp;
        const float wa = up ? vgm * bhm : vgpp * bhpp;
        const float w = up ? w0 : w1;
        const float d1 = w1 - w0;
        fx = fmaf(0.5f, limiter_fast(up ? d1 : -d1, w - wa), w) * dt;
    } else {
        float w, wp, wm;
        if (vh > 0.0f) {
            w = vg * bh;
            wp = (float)((double)(vgp * bhp - w) / 2.0);
            wm = (float)((double)(w - vgm * bhm) / 2.0);
        } else {
            w = vgp * bhp;
            wp = (float)((double)(w - vgpp * bhpp) / 2.0);
            wm = (float)((double)(vg * bh - w) / 2.0);
        }
        float dw = 0.0f;
        if (wm * wp > 0.0f) dw = (float)(2.0 * (double)wm * (double)wp / (double)(wm + wp));
        fx = (w + dw) * dt;
    }
    // b -= flux - flux(i-1), tvdb :80-88
    fm = up1(fx);
    return in.bl - (fx - fm);
}

}  // namespace
}  // namespace mhd
