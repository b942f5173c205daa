/*
 * mhd_ref.c -- CPU restatement of the reference's serial C path for the
 * relaxing-TVD split MHD step.  TEST INFRASTRUCTURE ONLY: tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it as the checker; the
 * product library never links or calls it.
 *
 * Compiled twice by oracle/Makefile: REAL=float (bit-identical to the
 * reference built with gcc on x86-64) and REAL=double (the "FP64 build":
 * the reference compiled with -Dfloat=double).  Built with -O2
 * -ffp-contract=off and no -march, so no FMA contraction happens.
 *
 * Arithmetic pattern: the reference stores every intermediate in float but
 * promotes to double wherever a double literal (2.0, 1.0, 4.0) or a libm
 * double routine (sqrt, fabs) meets a float.  Each expression below states
 * that pattern with explicit (double) casts; in the REAL=double build the
 * casts are no-ops, as they are in the reference's -Dfloat=double build.
 *
 * Layout (C/array_definition.h:12-14): u[((k*ny+j)*nx+i)*5+c],
 * b[((k*ny+j)*nx+i)*3+c]; every neighbour index is periodic.
 *
 * Pinning (what checks each function; all in tests/):
 *   flux/tvd/fluidx   golden vectors from oracle/_ref (the compiled
 *                     reference), uniform-state invariance, mass telescoping
 *   tvdb/advectbyzx   golden vectors, v=0 -> no change, div.B preservation
 *   transpose12/13    golden vectors + closed-form permutation + involution
 *   calcfl            golden vectors + closed form on a uniform state +
 *                     brute force on tiny boxes
 *   step driver       per-step SHA-256 and dt of golden trajectories
 *   CUDA semantics    (mhdref_set_cuda_semantics) the two index/averaging
 *                     changes only; checked to differ from the C path in
 *                     exactly the fields and steps they touch
 *   transpose13_any,  (no reference counterpart: boxes with ny != nz)
 *   step_any          closed-form index formula + involution; equal to
 *                     transpose13 / the reference step where ny == nz;
 *                     conserved mass and energy, div.B at round-off
 *   diagnostics       (no reference counterpart) sums against math.fsum;
 *                     div.B against a numpy forward-difference divergence,
 *                     zero for a forward curl, planted-face closed forms
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

#define U5(u, nx, ny, c, i, j, k) (u)[((((size_t)(k) * (ny) + (j)) * (nx) + (i)) * 5) + (c)]
#define B3(b, nx, ny, c, i, j, k) (b)[((((size_t)(k) * (ny) + (j)) * (nx) + (i)) * 3) + (c)]

static int g_threads = 1;

void mhdref_set_threads(int t) { g_threads = t < 1 ? 1 : t; }

/* The reference's CUDA path (Cuda/) instead of its C path, for the
 * cross-check of how far the two drift apart.  Two semantic divergences:
 *  - advectbyzx2_kernel reads the upwind pencil's rho, m1 at (i, km, k), the
 *    j slot, instead of (i, j, km) (Cuda/cuda_advectbyzx.cu:204,212 against
 *    C/advectbyzxB1.c:34), i.e. the B1 line's neighbour is pencil (k-1, k);
 *  - cfl_kernel uses the face-averaged B in ps and p (Cuda/cuda_cfl.cu:102-121)
 *    where C uses the face values (C/calcfl.c:50-54).
 * Its other defects (uninitialised cfl_c entries, host-side leaks) are not
 * reproducible semantics.  Off by default. */
static int g_cuda = 0;
void mhdref_set_cuda_semantics(int on) { g_cuda = on != 0; }
int mhdref_real_size(void) { return (int)sizeof(real); }

/* --------------------------------------------------------------------- */
/* mhdflux + find_max (C/mhdflux.c:9-72).  u: pencil (5,n) AoS, bc: (3,n)  */
/* cell-centred B.  Writes v (5,n) and returns the freezing speed c.      */
/* --------------------------------------------------------------------- */
static real ref_flux(int n, const real *u, const real *bc, real *v)
{
    const real gam = (real)(5.0 / 3.0); /* C/mhdflux.c:17-18: float gamma */
    real c = (real)0.0;                 /* find_max starts from 0 (:65) */
    for (int i = 0; i < n; i++) {
        const real *q = u + 5 * i;
        const real *bb = bc + 3 * i;
        real rho = q[0], mx = q[1], my = q[2], mz = q[3], e = q[4];
        real b1 = bb[0], b2 = bb[1], b3 = bb[2];
        real vx = mx / rho;                                   /* :23 */
        real msq = mx * mx + my * my + mz * mz;
        real bsq = b1 * b1 + b2 * b2 + b3 * b3;
        /* :28  ps=(u5-msq/u1/2.0)*(gamma-1.0)+(2.0-gamma)*bsq/2.0 */
        real ps = (real)(((double)e - (double)(msq / rho) / 2.0) * ((double)gam - 1.0)
                         + (2.0 - (double)gam) * (double)bsq / 2.0);
        real *w = v + 5 * i;                                  /* :33-37 */
        w[0] = mx;
        w[1] = mx * vx + ps - b1 * b1;
        w[2] = my * vx - b2 * b1;
        w[3] = mz * vx - b3 * b1;
        w[4] = (e + ps) * vx - b1 * (b1 * mx + b2 * my + b3 * mz) / rho;
        real p = (real)((double)ps - (double)bsq / 2.0);     /* :41 */
        /* :45  fabs(vx)+sqrt(fabs((bsq+gamma*p)/u1)) in double */
        real cf = (real)(fabs((double)vx) + sqrt(fabs((double)((bsq + gam * p) / rho))));
        if (cf > c) c = cf;                                   /* :68 */
    }
    if (c > 0) {                                              /* :48-57 */
        for (int i = 0; i < 5 * n; i++) v[i] = v[i] / c;
    }
    return c;
}

/* limiter of C/tvd1.c:75-79: 0, or 2ab/(a+b) in double when a*b>0 (float test) */
static inline real ref_limiter(real a, real b)
{
    real r = (real)0.0;
    if (a * b > 0) r = (real)(2.0 * (double)a * (double)b / (double)(a + b));
    return r;
}

/* --------------------------------------------------------------------- */
/* tvd1 (C/tvd1.c:8-118) on one pencil, in place.  Every abXYc of the     */
/* reference is a*Ra + b*Rb + c in float (C/includefunc_spu.h:73-84);     */
/* with a=1, b=+-1 that is Ra+-Rb, and the trailing "+ 0" turns -0 into    */
/* +0, so it is kept.                                                      */
/* --------------------------------------------------------------------- */
static void ref_tvd1(int n, real *u, const real *bc, real dt, real *scratch)
{
    const real z = (real)0.0, h = (real)0.5;
    real *v = scratch, *wl = v + 5 * n, *fr = wl + 5 * n, *flux = fr + 5 * n,
         *u1 = flux + 5 * n, *fl = u1 + 5 * n, *dfr = fl + 5 * n, *dfl = dfr + 5 * n;
#define IX(ii, i) (5 * (i) + (ii))
#define IP(i) (((i) + 1) % n)
#define IM(i) (((i) + n - 1) % n)
    real c = ref_flux(n, u, bc, v);                                  /* :34 */
    for (int i = 0; i < n; i++)
        for (int ii = 0; ii < 5; ii++) {
            real wr = u[IX(ii, i)] + v[IX(ii, i)] + z;               /* :36 */
            wl[IX(ii, i)] = u[IX(ii, i)] - v[IX(ii, i)] + z;         /* :38 */
            fr[IX(ii, i)] = c * wr + z * wr + z;                     /* :40 */
        }
    for (int i = 0; i < n; i++)
        for (int ii = 0; ii < 5; ii++) {
            real t = wl[IX(ii, IP(i))];                              /* :42-43 */
            real f = c * t + z * t + z;
            flux[IX(ii, i)] = h * fr[IX(ii, i)] + (-h) * f + z;       /* :45 */
        }
    const real hdt = (real)(-1.0 * (double)dt / 2.0);                /* :49 */
    for (int i = 0; i < n; i++)
        for (int ii = 0; ii < 5; ii++) {
            real d = flux[IX(ii, i)] - flux[IX(ii, IM(i))] + z;     /* :47-48 */
            u1[IX(ii, i)] = u[IX(ii, i)] + hdt * d + z;             /* :49 */
        }
    c = ref_flux(n, u1, bc, v);                                      /* :52 */
    for (int i = 0; i < n; i++)
        for (int ii = 0; ii < 5; ii++) {
            real wr = u1[IX(ii, i)] + v[IX(ii, i)] + z;              /* :54 */
            wl[IX(ii, i)] = u1[IX(ii, i)] - v[IX(ii, i)] + z;        /* :56 */
            fr[IX(ii, i)] = c * wr + z * wr + z;                     /* :58 */
        }
    for (int i = 0; i < n; i++)
        for (int ii = 0; ii < 5; ii++) {
            real dfrp = h * fr[IX(ii, IP(i))] + (-h) * fr[IX(ii, i)] + z;  /* :60-61 */
            real dfrm = h * fr[IX(ii, i)] + (-h) * fr[IX(ii, IM(i))] + z;  /* :63-64 */
            dfr[IX(ii, i)] = ref_limiter(dfrp, dfrm);                       /* :71-81 */
            real t = wl[IX(ii, IP(i))];                                     /* :83-84 */
            fl[IX(ii, i)] = c * t + z * t + z;
        }
    for (int i = 0; i < n; i++)
        for (int ii = 0; ii < 5; ii++) {
            real dflp = h * fl[IX(ii, i)] + (-h) * fl[IX(ii, IP(i))] + z;  /* :86-87 */
            real dflm = h * fl[IX(ii, IM(i))] + (-h) * fl[IX(ii, i)] + z;  /* :89-90 */
            dfl[IX(ii, i)] = ref_limiter(dflp, dflm);                       /* :97-107 */
        }
    for (int i = 0; i < n; i++)
        for (int ii = 0; ii < 5; ii++) {
            real t1 = fr[IX(ii, i)] - fl[IX(ii, i)] + z;             /* :110 */
            real t2 = dfr[IX(ii, i)] - dfl[IX(ii, i)] + z;           /* :111 */
            flux[IX(ii, i)] = h * t1 + h * t2 + z;                   /* :112 */
        }
    const real fdt = (real)(-1.0 * (double)dt);                      /* :116 */
    for (int i = 0; i < n; i++)
        for (int ii = 0; ii < 5; ii++) {
            real d = flux[IX(ii, i)] - flux[IX(ii, IM(i))] + z;     /* :114-115 */
            u[IX(ii, i)] = u[IX(ii, i)] + fdt * d + z;              /* :116 */
        }
#undef IX
#undef IP
#undef IM
}

/* --------------------------------------------------------------------- */
/* fluidx (C/fluidx.c:19-102): for every x-pencil, cell-centred B from    */
/* face B at i+1, j+1, k+1, then tvd1.  Reads update_u/update_b, writes u. */
/* --------------------------------------------------------------------- */
void mhdref_fluidx(real *u, const real *b, int nx, int ny, int nz, real dt,
                   const real *uu, const real *bu)
{
    (void)b;
    size_t npen = (size_t)ny * nz;
#pragma omp parallel num_threads(g_threads)
    {
        real *pu = (real *)malloc(sizeof(real) * (size_t)nx * (5 + 3 + 40));
        real *pb = pu + 5 * nx, *scratch = pb + 3 * nx;
#pragma omp for schedule(static)
        for (size_t pk = 0; pk < npen; pk++) {
            int j = (int)(pk % ny), k = (int)(pk / ny);
            int jp = (j + 1) % ny, kp = (k + 1) % nz;            /* :52-53 */
            for (int i = 0; i < nx; i++)
                for (int c = 0; c < 5; c++) pu[5 * i + c] = U5(uu, nx, ny, c, i, j, k);
            for (int i = 0; i < nx; i++) {                        /* :71-81 */
                int ip = (i + 1) % nx;
                real h1 = (real)0.5 * B3(bu, nx, ny, 0, i, j, k);
                real h1p = (real)0.5 * B3(bu, nx, ny, 0, ip, j, k);
                pb[3 * i + 0] = h1 + h1p;
                pb[3 * i + 1] = (real)0.5 * B3(bu, nx, ny, 1, i, j, k) + B3(bu, nx, ny, 1, i, jp, k) / 2;
                pb[3 * i + 2] = (real)0.5 * B3(bu, nx, ny, 2, i, j, k) + B3(bu, nx, ny, 2, i, j, kp) / 2;
            }
            ref_tvd1(nx, pu, pb, dt, scratch);                    /* :83 */
            for (int i = 0; i < nx; i++)
                for (int c = 0; c < 5; c++) U5(u, nx, ny, c, i, j, k) = pu[5 * i + c];
        }
        free(pu);
    }
}

/* --------------------------------------------------------------------- */
/* tvdb (C/tvdb.c:10-91): 2nd-order upwind limited flux of b*v on one     */
/* pencil.  b is updated in place; flux receives the edge fluxes.         */
/* --------------------------------------------------------------------- */
static void ref_tvdb(int n, real *flux, real *b, const real *vg, real dt, real *scratch)
{
    real *vh = scratch, *bv = vh + n, *flux1 = bv + n, *b1 = flux1 + n;
    const real z = (real)0.0, h = (real)0.5;
    for (int i = 0; i < n; i++) vh[i] = h * vg[i] + h * vg[(i + 1) % n] + z;  /* :26-27 */
    for (int i = 0; i < n; i++) bv[i] = b[i] * vg[i];                         /* :33 */
    for (int i = 0; i < n; i++) flux1[i] = vh[i] > 0 ? bv[i] : bv[(i + 1) % n]; /* :36-46 */
    for (int i = 0; i < n; i++) {                                              /* :48-52 */
        real d = (flux1[i] - flux1[(i + n - 1) % n]) * dt;
        b1[i] = (real)((double)b[i] - (double)d / 2.0);
    }
    for (int i = 0; i < n; i++) {                                              /* :54-78 */
        int ip = (i + 1) % n, ipp = (ip + 1) % n, im = (i + n - 1) % n;
        real w, wp, wm, dw;
        if (vh[i] > 0) {
            w = vg[i] * b1[i];
            wp = (real)((double)(vg[ip] * b1[ip] - w) / 2.0);
            wm = (real)((double)(w - vg[im] * b1[im]) / 2.0);
        } else {
            w = vg[ip] * b1[ip];
            wp = (real)((double)(w - vg[ipp] * b1[ipp]) / 2.0);
            wm = (real)((double)(vg[i] * b1[i] - w) / 2.0);
        }
        dw = (real)0.0;
        if (wm * wp > 0) dw = (real)(2.0 * (double)wm * (double)wp / (double)(wm + wp));
        flux[i] = (w + dw) * dt;
    }
    for (int i = 0; i < n; i++) b[i] = b[i] - (flux[i] - flux[(i + n - 1) % n]); /* :80-88 */
}

/* One A1/B1 pencil (C/advectbyzxA1.c:26-57, C/advectbyzxB1.c:26-56).     */
/* (jo,ko) is the upwind neighbour pencil (j-1,k) or (j,k-1); comp is the */
/* advected face component (1 for b2, 2 for b3).                          */
static void ref_advect_pencil(real *b, int nx, int ny, real dt, const real *uu, const real *bu,
                              real *tmp, int j, int k, int jo, int ko, int comp, real *scratch)
{
    real *vx = scratch, *vs = vx + nx, *bx = vs + nx, *fx = bx + nx, *sc = fx + nx;
    for (int i = 0; i < nx; i++)                                       /* A1 :31-35 */
        vx[i] = (U5(uu, nx, ny, 1, i, jo, ko) + U5(uu, nx, ny, 1, i, j, k))
                / (U5(uu, nx, ny, 0, i, jo, ko) + U5(uu, nx, ny, 0, i, j, k));
    for (int i = 0; i < nx; i++) {                                     /* A1 :37-44 */
        real s = vx[(i + nx - 1) % nx] + vx[(i + 1) % nx];
        vs[i] = (real)(((double)s + 2.0 * (double)vx[i]) / 4.0);
        bx[i] = B3(bu, nx, ny, comp, i, j, k);
    }
    ref_tvdb(nx, fx, bx, vs, dt, sc);                                  /* A1 :46 */
    for (int i = 0; i < nx; i++) {                                     /* A1 :49-57 */
        real fm = fx[(i + nx - 1) % nx];
        B3(b, nx, ny, comp, i, j, k) = bx[i];
        B3(b, nx, ny, 0, i, j, k) = B3(b, nx, ny, 0, i, j, k) - fm;
        tmp[((size_t)k * ny + j) * nx + i] = fm;
    }
}

/* --------------------------------------------------------------------- */
/* advectbyzx (C/advectbyzx.c:14-23): A1, A2, B1, B2 in this order,       */
/* sharing one tmp array, so b1 receives                                 */
/*   (((b1 - tA(j)) + tA(j+1)) - tB(k)) + tB(k+1).                        */
/* --------------------------------------------------------------------- */
void mhdref_advectbyzx(real *b, int nx, int ny, int nz, real dt, const real *uu, const real *bu)
{
    size_t ncell = (size_t)nx * ny * nz, npen = (size_t)ny * nz;
    real *tmp = (real *)malloc(sizeof(real) * ncell);
    for (int pass = 0; pass < 2; pass++) {
        int comp = pass == 0 ? 1 : 2;
#pragma omp parallel num_threads(g_threads)
        {
            real *scratch = (real *)malloc(sizeof(real) * (size_t)nx * 12);
#pragma omp for schedule(static)
            for (size_t pk = 0; pk < npen; pk++) {
                int j = (int)(pk % ny), k = (int)(pk / ny);
                int jo = pass == 0 ? (j + ny - 1) % ny : j;          /* A1 :30 */
                int ko = pass == 0 ? k : (k + nz - 1) % nz;          /* B1 :30 */
                if (g_cuda && pass == 1) {          /* Cuda/cuda_advectbyzx.cu:204 */
                    jo = ((k + nz - 1) % nz) % ny;
                    ko = k;
                }
                ref_advect_pencil(b, nx, ny, dt, uu, bu, tmp, j, k, jo, ko, comp, scratch);
            }
            free(scratch);
        }
        /* A2 (C/advectbyzxA2.c:16-26) / B2 (C/advectbyzxB2.c:16-26) */
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (size_t pk = 0; pk < npen; pk++) {
            int j = (int)(pk % ny), k = (int)(pk / ny);
            int jd = pass == 0 ? (j + ny - 1) % ny : j;
            int kd = pass == 0 ? k : (k + nz - 1) % nz;
            for (int i = 0; i < nx; i++)
                B3(b, nx, ny, 0, i, jd, kd) = B3(b, nx, ny, 0, i, jd, kd) + tmp[((size_t)k * ny + j) * nx + i];
        }
    }
    free(tmp);
}

/* transpose12 (C/transpose12.c:7-47): out dims (ny,nx,nz); u comps       */
/* (1,3,2,4,5), b comps (2,1,3).                                           */
void mhdref_transpose12(real *u, real *b, const real *uu, const real *bu, int nx, int ny, int nz)
{
    static const int pu[5] = {0, 2, 1, 3, 4}, pb[3] = {1, 0, 2};
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                for (int c = 0; c < 5; c++) U5(u, ny, nx, c, j, i, k) = U5(uu, nx, ny, pu[c], i, j, k);
                for (int c = 0; c < 3; c++) B3(b, ny, nx, c, j, i, k) = B3(bu, nx, ny, pb[c], i, j, k);
            }
}

void mhdref_transpose13_any(real *u, real *b, const real *uu, const real *bu, int nx, int ny, int nz);

/* transpose13 (C/transpose13.c:9-62): needs nx==nz; out dims (nz,ny,nx); */
/* u comps (1,4,3,2,5), b comps (3,2,1).  Returns -1 (and writes nothing) */
/* when nx != nz, where the reference prints and returns.                 */
int mhdref_transpose13(real *u, real *b, const real *uu, const real *bu, int nx, int ny, int nz)
{
    if (nx != nz) return -1;
    mhdref_transpose13_any(u, b, uu, bu, nx, ny, nz);
    return 0;
}

/* The same x<->z swap for any nx, nz (no reference counterpart: the      */
/* reference refuses nx != nz).  Closed-form permutation; pinned by the   */
/* index formula and by transpose13_any twice being the identity.         */
void mhdref_transpose13_any(real *u, real *b, const real *uu, const real *bu, int nx, int ny, int nz)
{
    static const int pu[5] = {0, 3, 2, 1, 4}, pb[3] = {2, 1, 0};
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++) {
                for (int c = 0; c < 5; c++) U5(u, nz, ny, c, k, j, i) = U5(uu, nx, ny, pu[c], i, j, k);
                for (int c = 0; c < 3; c++) B3(b, nz, ny, c, k, j, i) = B3(bu, nx, ny, pb[c], i, j, k);
            }
}

/* max_3num (C/calcfl.c:65-89) */
static inline real ref_max3(real a, real b, real c)
{
    if (a > b) return a > c ? a : c;
    return b > c ? b : c;
}

/* Per-cell CFL speed of C/calcfl.c:39-54 (face B, not averaged). */
real mhdref_cfl_cell(const real *q, const real *bb)
{
    const real gam = (real)(5.0 / 3.0);
    real rho = q[0];
    real t1 = (real)fabs((double)(q[1] / rho));
    real t2 = (real)fabs((double)(q[2] / rho));
    real t3 = (real)fabs((double)(q[3] / rho));
    real v = ref_max3(t1, t2, t3);
    real msq = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    real bsq = bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2];
    real ps = (real)(((double)q[4] - (double)(msq / rho) / 2.0) * ((double)gam - 1.0)
                     + (2.0 - (double)gam) * (double)bsq / 2.0);
    real p = (real)((double)ps - (double)bsq / 2.0);
    return (real)((double)v + sqrt(fabs(((double)bsq * 2.0 + (double)(gam * p)) / (double)rho)));
}

/* calcfl (C/calcfl.c:13-63): running max_2num over cells in k,j,i order  */
/* starting from 0, returns 1/c.  Serial so the NaN semantics of the       */
/* running max (a NaN cell resets the max to the next cell's value) hold.  */
real mhdref_calcfl(const real *u, const real *b, int nx, int ny, int nz)
{
    real c = (real)0.0;
    size_t ncell = (size_t)nx * ny * nz;
    for (size_t s = 0; s < ncell; s++) {
        real t;
        if (g_cuda) {  /* face-averaged B, Cuda/cuda_cfl.cu:102-108 */
            const int i = (int)(s % nx), j = (int)((s / nx) % ny), k = (int)(s / ((size_t)nx * ny));
            real ba[3];
            ba[0] = (real)((double)(B3(b, nx, ny, 0, i, j, k) + B3(b, nx, ny, 0, (i + 1) % nx, j, k)) / 2.0);
            ba[1] = (real)((double)(B3(b, nx, ny, 1, i, j, k) + B3(b, nx, ny, 1, i, (j + 1) % ny, k)) / 2.0);
            ba[2] = (real)((double)(B3(b, nx, ny, 2, i, j, k) + B3(b, nx, ny, 2, i, j, (k + 1) % nz)) / 2.0);
            t = mhdref_cfl_cell(u + 5 * s, ba);
        } else {
            t = mhdref_cfl_cell(u + 5 * s, b + 3 * s);
        }
        c = c > t ? c : t;                                         /* max_2num :93-100 */
    }
    return 1 / c;                                                  /* :61 */
}

/* copy_matrix (C/copy_matrix.c:9-38): update <- main */
void mhdref_copy_matrix(const real *u, const real *b, int nx, int ny, int nz, real *uu, real *bu)
{
    size_t ncell = (size_t)nx * ny * nz;
    memcpy(uu, u, sizeof(real) * 5 * ncell);
    memcpy(bu, b, sizeof(real) * 3 * ncell);
}

/* --------------------------------------------------------------------- */
/* One full step with the sweep order and dt rule of C/mhd.c:59-124.      */
/* (u,b) is the main state, (uu,bu) the update copy; both hold the state  */
/* on entry and on exit.  dt = (real)(cfl * calcfl); if tf > 0 the        */
/* reference's clamp dt <= (tf-t)/2 applies; *t advances by 2dt.          */
/* Returns dt.                                                            */
/* --------------------------------------------------------------------- */
static real step_body(real *u, real *b, real *uu, real *bu, int nx, int ny, int nz, double cfl,
                      real tf, real *t, int any_shape);

real mhdref_step(real *u, real *b, real *uu, real *bu, int nx, int ny, int nz, double cfl,
                 real tf, real *t)
{
    return step_body(u, b, uu, bu, nx, ny, nz, cfl, tf, t, 0);
}

/* The same step for ny != nz, with transpose13_any (no reference          */
/* counterpart: the reference's transpose13 refuses such boxes).           */
real mhdref_step_any(real *u, real *b, real *uu, real *bu, int nx, int ny, int nz, double cfl,
                     real tf, real *t)
{
    return step_body(u, b, uu, bu, nx, ny, nz, cfl, tf, t, 1);
}

static real step_body(real *u, real *b, real *uu, real *bu, int nx, int ny, int nz, double cfl,
                      real tf, real *t, int any_shape)
{
    real dt = (real)(cfl * (double)mhdref_calcfl(u, b, nx, ny, nz));  /* :62 */
    if (tf > 0 && dt > (tf - *t) / 2) dt = (tf - *t) / 2;              /* :63 */
    *t = *t + 2 * dt;                                                  /* :64 */
#define CP(a, bb_, c) mhdref_copy_matrix(u, b, a, bb_, c, uu, bu)
    mhdref_fluidx(u, b, nx, ny, nz, dt, uu, bu);            CP(nx, ny, nz);   /* :71-72 */
    mhdref_advectbyzx(b, nx, ny, nz, dt, uu, bu);           CP(nx, ny, nz);   /* :74-75 */
    mhdref_transpose12(u, b, uu, bu, nx, ny, nz);           CP(ny, nx, nz);   /* :78-79 */
    mhdref_fluidx(u, b, ny, nx, nz, dt, uu, bu);            CP(ny, nx, nz);   /* :80-81 */
    mhdref_advectbyzx(b, ny, nx, nz, dt, uu, bu);           CP(ny, nx, nz);   /* :84-85 */
    if (any_shape) mhdref_transpose13_any(u, b, uu, bu, ny, nx, nz);
    else mhdref_transpose13(u, b, uu, bu, ny, nx, nz);
    CP(nz, nx, ny);                                                            /* :88-89 */
    mhdref_fluidx(u, b, nz, nx, ny, dt, uu, bu);            CP(nz, nx, ny);   /* :91-92 */
    mhdref_advectbyzx(b, nz, nx, ny, dt, uu, bu);           CP(nz, nx, ny);   /* :93-94 */
    mhdref_advectbyzx(b, nz, nx, ny, dt, uu, bu);           CP(nz, nx, ny);   /* :95-96 */
    mhdref_fluidx(u, b, nz, nx, ny, dt, uu, bu);            CP(nz, nx, ny);   /* :97-98 */
    if (any_shape) mhdref_transpose13_any(u, b, uu, bu, nz, nx, ny);
    else mhdref_transpose13(u, b, uu, bu, nz, nx, ny);
    CP(ny, nx, nz);                                                            /* :102-103 */
    mhdref_advectbyzx(b, ny, nx, nz, dt, uu, bu);           CP(ny, nx, nz);   /* :105-106 */
    mhdref_fluidx(u, b, ny, nx, nz, dt, uu, bu);            CP(ny, nx, nz);   /* :107-108 */
    mhdref_transpose12(u, b, uu, bu, ny, nx, nz);           CP(nx, ny, nz);   /* :112-113 */
    mhdref_advectbyzx(b, nx, ny, nz, dt, uu, bu);           CP(nx, ny, nz);   /* :115-116 */
    mhdref_fluidx(u, b, nx, ny, nz, dt, uu, bu);            CP(nx, ny, nz);   /* :117-118 */
#undef CP
    return dt;
}

/* --------------------------------------------------------------------- */
/* FP64 diagnostics checker: conserved sums (mass, 3 momenta, energy) and */
/* div.B, accumulated in long double.  sums[0..4]; divb[0]=max|div|,      */
/* divb[1]=sqrt(mean div^2).  Face b_d at index i sits on the low face.   */
/* --------------------------------------------------------------------- */
void mhdref_diagnostics(const real *u, const real *b, int nx, int ny, int nz, double *sums, double *divb)
{
    long double s[5] = {0, 0, 0, 0, 0}, d2 = 0, dmax = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                for (int c = 0; c < 5; c++) s[c] += (long double)U5(u, nx, ny, c, i, j, k);
                long double d = ((long double)B3(b, nx, ny, 0, (i + 1) % nx, j, k) - B3(b, nx, ny, 0, i, j, k))
                                + ((long double)B3(b, nx, ny, 1, i, (j + 1) % ny, k) - B3(b, nx, ny, 1, i, j, k))
                                + ((long double)B3(b, nx, ny, 2, i, j, (k + 1) % nz) - B3(b, nx, ny, 2, i, j, k));
                d2 += d * d;
                if (fabsl(d) > dmax) dmax = fabsl(d);
            }
    for (int c = 0; c < 5; c++) sums[c] = (double)s[c];
    divb[0] = (double)dmax;
    divb[1] = (double)sqrtl(d2 / ((long double)nx * ny * nz));
}
