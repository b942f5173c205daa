/*
 * ref_harness.c -- drives the reference's own C kernels (compiled from
 * /root/reference/C by oracle/Makefile into oracle/_ref/) so that golden
 * vectors can be written from the real reference.  TEST INFRASTRUCTURE ONLY.
 *
 * The reference's main (C/mhd.c) cannot serve: it includes the Cell SDK
 * header libspe2.h and reads an IC file that is not shipped.  This harness
 * calls the same kernels in the order of C/mhd.c:59-124 instead.
 *
 * The reference kernels keep static pencil buffers sized by the compile-time
 * constant n=100 (C/parameter.h:6, C/fluidx.c:9-15, C/advectbyzx.c:7), so
 * every dimension must be <= 100 and nx*ny*nz <= 100^3; the entry points
 * return -1 otherwise.  When built with -Dfloat=double every "float" below
 * and in the reference becomes double (the FP64 build).
 */
#include <stdio.h>

void fluidx(float *u, float *b, int *nx, int *ny, int *nz, float *dt, float *uu, float *bu, int iter);
void copy_matrix(float *u, float *b, int *nx, int *ny, int *nz, float *uu, float *bu);
void advectbyzx(float *u, float *b, int *nx, int *ny, int *nz, float *dt, float *uu, float *bu);
void transpose12(float *u, float *b, float *uu, float *bu, int *nx, int *ny, int *nz, int iter);
void transpose13(float *u, float *b, float *uu, float *bu, int *nx, int *ny, int *nz);
float calcfl(float *u, float *b, int *nx, int *ny, int *nz);

#define REF_NMAX 100

static int dims_ok(int nx, int ny, int nz)
{
    if (nx < 1 || ny < 1 || nz < 1) return 0;
    if (nx > REF_NMAX || ny > REF_NMAX || nz > REF_NMAX) return 0;
    return 1;
}

int refc_real_size(void) { return (int)sizeof(float); }

int refc_fluidx(float *u, float *b, int nx, int ny, int nz, float dt, float *uu, float *bu)
{
    if (!dims_ok(nx, ny, nz)) return -1;
    fluidx(u, b, &nx, &ny, &nz, &dt, uu, bu, 0);
    return 0;
}

int refc_advectbyzx(float *u, float *b, int nx, int ny, int nz, float dt, float *uu, float *bu)
{
    if (!dims_ok(nx, ny, nz)) return -1;
    advectbyzx(u, b, &nx, &ny, &nz, &dt, uu, bu);
    return 0;
}

int refc_transpose12(float *u, float *b, float *uu, float *bu, int nx, int ny, int nz)
{
    if (!dims_ok(nx, ny, nz)) return -1;
    transpose12(u, b, uu, bu, &nx, &ny, &nz, 0);
    return 0;
}

int refc_transpose13(float *u, float *b, float *uu, float *bu, int nx, int ny, int nz)
{
    if (!dims_ok(nx, ny, nz)) return -1;
    transpose13(u, b, uu, bu, &nx, &ny, &nz);
    return 0;
}

int refc_calcfl(float *u, float *b, int nx, int ny, int nz, float *out)
{
    if (!dims_ok(nx, ny, nz)) return -1;
    *out = calcfl(u, b, &nx, &ny, &nz);
    return 0;
}

int refc_copy_matrix(float *u, float *b, int nx, int ny, int nz, float *uu, float *bu)
{
    if (!dims_ok(nx, ny, nz)) return -1;
    copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
    return 0;
}

/* One step of C/mhd.c:59-124 with the CFL factor as a parameter (the
 * reference hard-codes 0.9) and the tf clamp applied only when tf > 0. */
int refc_step(float *u, float *b, float *uu, float *bu, int nx, int ny, int nz, double cfl,
              float tf, float *t, float *dt_out)
{
    if (!dims_ok(nx, ny, nz) || ny != nz) return -1;
    float dt = cfl * calcfl(u, b, &nx, &ny, &nz);
    if (tf > 0 && dt > (tf - *t) / 2) dt = (tf - *t) / 2;
    *t = *t + 2 * dt;
    fluidx(u, b, &nx, &ny, &nz, &dt, uu, bu, 0);       copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
    advectbyzx(u, b, &nx, &ny, &nz, &dt, uu, bu);      copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
    transpose12(u, b, uu, bu, &nx, &ny, &nz, 0);       copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
    fluidx(u, b, &ny, &nx, &nz, &dt, uu, bu, 0);       copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
    advectbyzx(u, b, &ny, &nx, &nz, &dt, uu, bu);      copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
    transpose13(u, b, uu, bu, &ny, &nx, &nz);          copy_matrix(u, b, &nz, &nx, &ny, uu, bu);
    fluidx(u, b, &nz, &nx, &ny, &dt, uu, bu, 0);       copy_matrix(u, b, &nz, &nx, &ny, uu, bu);
    advectbyzx(u, b, &nz, &nx, &ny, &dt, uu, bu);      copy_matrix(u, b, &nz, &nx, &ny, uu, bu);
    advectbyzx(u, b, &nz, &nx, &ny, &dt, uu, bu);      copy_matrix(u, b, &nz, &nx, &ny, uu, bu);
    fluidx(u, b, &nz, &nx, &ny, &dt, uu, bu, 0);       copy_matrix(u, b, &nz, &nx, &ny, uu, bu);
    transpose13(u, b, uu, bu, &nz, &nx, &ny);          copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
    advectbyzx(u, b, &ny, &nx, &nz, &dt, uu, bu);      copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
    fluidx(u, b, &ny, &nx, &nz, &dt, uu, bu, 0);       copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
    transpose12(u, b, uu, bu, &ny, &nx, &nz, 0);       copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
    advectbyzx(u, b, &nx, &ny, &nz, &dt, uu, bu);      copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
    fluidx(u, b, &nx, &ny, &nz, &dt, uu, bu, 0);       copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
    *dt_out = dt;
    return 0;
}
