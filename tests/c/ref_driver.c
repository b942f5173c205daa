/*
 * ref_driver.c -- a reference-style C driver: the step loop of C/mhd.c:59-124
 * (with CFL 0.7 and a fixed step count), written only against the
 * reference's prototypes (C/spu_function.h), on plain host arrays, and
 * linked against libmhd_hip.so instead of the reference's C files.
 *
 *   ref_driver <in.bin> <out.bin> <nx> <ny> <nz> <nsteps>
 *
 * in.bin/out.bin hold u (5*N floats) then b (3*N floats).  Prints one line
 * per step: "step <i> dt <dt as %a> t <t as %a>".
 */
#include <stdio.h>
#include <stdlib.h>

#include "mhd_hip.h"

int main(int argc, char **argv)
{
    if (argc != 7) {
        fprintf(stderr, "usage: %s in.bin out.bin nx ny nz nsteps\n", argv[0]);
        return 2;
    }
    int nx = atoi(argv[3]), ny = atoi(argv[4]), nz = atoi(argv[5]), nsteps = atoi(argv[6]);
    size_t n = (size_t)nx * ny * nz;
    float *u = malloc(sizeof(float) * 5 * n), *b = malloc(sizeof(float) * 3 * n);
    float *uu = malloc(sizeof(float) * 5 * n), *bu = malloc(sizeof(float) * 3 * n);
    FILE *f = fopen(argv[1], "rb");
    if (!f || fread(u, sizeof(float), 5 * n, f) != 5 * n || fread(b, sizeof(float), 3 * n, f) != 3 * n) {
        fprintf(stderr, "cannot read %s\n", argv[1]);
        return 1;
    }
    fclose(f);
    copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
    float t = 0, dt;
    for (int iter = 1; iter <= nsteps; iter++) {
        dt = 0.7 * calcfl(u, b, &nx, &ny, &nz);
        t = t + 2 * dt;
        fluidx(u, b, &nx, &ny, &nz, &dt, uu, bu, iter);       copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
        advectbyzx(u, b, &nx, &ny, &nz, &dt, uu, bu);         copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
        transpose12(u, b, uu, bu, &nx, &ny, &nz, iter);       copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
        fluidx(u, b, &ny, &nx, &nz, &dt, uu, bu, iter);       copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
        advectbyzx(u, b, &ny, &nx, &nz, &dt, uu, bu);         copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
        transpose13(u, b, uu, bu, &ny, &nx, &nz);             copy_matrix(u, b, &nz, &nx, &ny, uu, bu);
        fluidx(u, b, &nz, &nx, &ny, &dt, uu, bu, iter);       copy_matrix(u, b, &nz, &nx, &ny, uu, bu);
        advectbyzx(u, b, &nz, &nx, &ny, &dt, uu, bu);         copy_matrix(u, b, &nz, &nx, &ny, uu, bu);
        advectbyzx(u, b, &nz, &nx, &ny, &dt, uu, bu);         copy_matrix(u, b, &nz, &nx, &ny, uu, bu);
        fluidx(u, b, &nz, &nx, &ny, &dt, uu, bu, iter);       copy_matrix(u, b, &nz, &nx, &ny, uu, bu);
        transpose13(u, b, uu, bu, &nz, &nx, &ny);             copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
        advectbyzx(u, b, &ny, &nx, &nz, &dt, uu, bu);         copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
        fluidx(u, b, &ny, &nx, &nz, &dt, uu, bu, iter);       copy_matrix(u, b, &ny, &nx, &nz, uu, bu);
        transpose12(u, b, uu, bu, &ny, &nx, &nz, iter);       copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
        advectbyzx(u, b, &nx, &ny, &nz, &dt, uu, bu);         copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
        fluidx(u, b, &nx, &ny, &nz, &dt, uu, bu, iter);       copy_matrix(u, b, &nx, &ny, &nz, uu, bu);
        printf("step %d dt %a t %a\n", iter, dt, t);
    }
    f = fopen(argv[2], "wb");
    if (!f) return 1;
    fwrite(u, sizeof(float), 5 * n, f);
    fwrite(b, sizeof(float), 3 * n, f);
    fclose(f);
    free(u); free(b); free(uu); free(bu);
    return 0;
}
