// mhd_internal.h -- shared declarations of the gfx950 HIP kernels' host-side
// launchers.  Everything here is internal to libmhd_hip.so; the public C ABI
// is include/mhd_hip.h.
//
// State layout (the reference's Fortran order, C/array_definition.h:12-14):
//   u[((k*ny + j)*nx + i)*5 + c]  c = rho, m1, m2, m3, E
//   b[((k*ny + j)*nx + i)*3 + c]  c = b1, b2, b3 (face-staggered: b_d at index
//                                  i sits on the low face of cell i along d)
// Each operator sweeps along its own first axis (nx); callers permute the
// dimensions after transposes exactly as C/mhd.c does.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define MHD_OK 0
#define MHD_ERR_ARG -1        // bad dimensions or null pointer
#define MHD_ERR_SHAPE -2      // transpose13 with nx != nz
#define MHD_ERR_HIP -3        // a HIP runtime call failed
#define MHD_ERR_TOO_LONG -4   // pencil longer than the kernel's LDS plan allows
#define MHD_ERR_NOMEM -5      // device allocation failed

namespace mhd {

// LDS-qualified pointee types for device code: index arithmetic on them is
// 32-bit, so the constant part of an offset folds into the ds_* immediate.
// v4f is a builtin 16-byte vector (no class operators, so it can live in LDS).
using lds_f = __attribute__((address_space(3))) float;
typedef float v4f __attribute__((ext_vector_type(4)));
using lds_v4 = __attribute__((address_space(3))) v4f;
using lds_u = __attribute__((address_space(3))) unsigned;

// Largest pencil (first-axis length) the sweep kernels accept.
constexpr int kMaxPencil = 2048;

// Workgroup -> pencil-group schedule of the sweep kernels.  A group is `pb`
// consecutive pencils along j (one contiguous run of cells); there are
// G = gj * nz groups with gj = ny / pb.  Workgroups are dispatched round-robin
// over the 8 XCDs of an MI355X, each with its own L2, so (a) every XCD gets a
// contiguous share of the schedule, and (b) that share is walked so that the
// neighbour pencils j+-1 and k+-1 a sweep reads were read by the same XCD a
// short while before and still sit in its L2:
//  * strips: kStripW groups wide in j, walked down all of k (j fastest), when
//    gj is a multiple of 8 * kStripW.  Row k+-1 of a strip is kStripW groups
//    away; only the two strip edges read from outside.
//  * else 8 x 8 tiles of groups in (j, k) when both divide;
//  * else the plain order.
// Returns the group index in memory order.
#ifndef MHD_STRIP_W
#define MHD_STRIP_W 32
#endif
constexpr int kStripW = MHD_STRIP_W;
__device__ __forceinline__ int sweep_group(int bid, int G, int gj, int nz)
{
    constexpr int kXcd = 8, kT = 8;
    int s = bid;
    if (G % kXcd == 0) s = (bid % kXcd) * (G / kXcd) + bid / kXcd;
    if (G != gj * nz) return s;
    if (kStripW > 0 && gj % (kXcd * kStripW) == 0) {
        const int strip = s / (kStripW * nz), w = s % (kStripW * nz);
        return (w / kStripW) * gj + strip * kStripW + w % kStripW;
    }
    if (gj % kT == 0 && nz % kT == 0) {
        const int tile = s / (kT * kT), w = s % (kT * kT);
        const int tiles_j = gj / kT;
        const int tj = tile % tiles_j, tk = tile / tiles_j;
        return (tk * kT + w / kT) * gj + tj * kT + w % kT;
    }
    return s;
}

// Launch shape of the pencil sweeps (fluidx, advectbyzx and the fused pair
// share it, so their workgroup -> pencil-group maps and CFL record counts
// agree): threads per workgroup, cells per thread, threads per pencil.
struct SweepCfg {
    int block, m, tp;
};
inline SweepCfg sweep_cfg(int nx)
{
    if (nx <= 64) return {256, 1, 64};
    if (nx <= 128) return {256, 2, 64};
    if (nx <= 256) return {256, 2, 128};
    if (nx <= 512) return {256, 2, 256};
    if (nx <= 1024) return {512, 2, 512};
    return {1024, 2, 1024};
}

// fast selects the FP32 tolerance mode of the sweep kernels (hardware
// reciprocal and square root, FMA); otherwise results are bitwise the
// reference's.
// When cfl_ws is non-null the kernel also computes calcfl of its output
// state (u_out with the face field bu) and writes inv_c / dt_next exactly as
// launch_calcfl does (no clamp, no t update).
int launch_fluidx(float *u_out, const float *uu, const float *bu, int nx, int ny, int nz,
                  const float *dt_dev, float dt_val, hipStream_t s, void *cfl_ws = nullptr,
                  float *inv_c = nullptr, float *dt_next = nullptr, double cfl = 0.0,
                  bool fast = false);

// dt = dt_next, clamped to (tf - t)/2 when tf > 0; t += 2 dt  (C/mhd.c:63-64)
int launch_begin_step(float *t, float *dt, const float *dt_next, float tf, hipStream_t s);

int launch_advectbyzx(float *b_out, const float *b_main, const float *uu, const float *bu, int nx,
                      int ny, int nz, const float *dt_dev, float dt_val, hipStream_t s,
                      bool fast = false);

// advectbyzx (b_out from bu, b_main = bu) followed by fluidx (u_out from uu
// and the new b), bitwise equal to the two launches; u_out must not alias uu
// nor b_out bu.  cfl_ws etc. as for launch_fluidx.  npass = 2: advectbyzx
// twice with the same uu, then fluidx (no CFL).
int launch_advect_fluid(float *u_out, float *b_out, const float *uu, const float *bu, int nx, int ny,
                        int nz, const float *dt_dev, float dt_val, hipStream_t s, void *cfl_ws,
                        float *inv_c, float *dt_next, double cfl, bool fast, int npass = 1);

// advectbyzx (b_main = bu) followed by transpose12 (which = 12, out dims
// (n2, n1, n3)) or the general transpose13 (which = 13, out dims (n3, n2,
// n1)) of uu and the new b, in one kernel (mhd_advect_transpose.hip); bitwise
// equal to the two launches.  No output may overlap an input.
int launch_advect_transpose(float *u_out, float *b_out, const float *uu, const float *bu, int n1, int n2, int n3,
                            const float *dt_dev, float dt_val, hipStream_t s, bool fast, int which);
// transpose12 (which = 12, out dims (n2, n1, n3)) or the general transpose13
// (which = 13, out dims (n3, n2, n1)) of (uu, bu) followed by advectbyzx of
// the transposed state (b_main = the transposed b), in one kernel
// (mhd_transpose_advect.hip); bitwise equal to the two launches.  Needs
// transpose_advect_supported (n1 a multiple of 16, the output's first dim a
// multiple of 4, 16-byte aligned u and b); MHD_ERR_SHAPE otherwise.  No
// output may overlap an input.
bool transpose_advect_supported(const float *u_out, const float *uu, const float *bu, int n1, int n2, int n3,
                                int which);
int launch_transpose_advect(float *u_out, float *b_out, const float *uu, const float *bu, int n1, int n2,
                            int n3, const float *dt_dev, float dt_val, hipStream_t s, bool fast, int which);

int launch_transpose12(float *u_out, float *b_out, const float *uu, const float *bu, int nx, int ny,
                       int nz, hipStream_t s);

// any_shape: nx != nz allowed (out dims (nz, ny, nx)); the reference's
// transpose13 refuses it, so its C-ABI entry points keep any_shape false.
int launch_transpose13(float *u_out, float *b_out, const float *uu, const float *bu, int nx, int ny,
                       int nz, hipStream_t s, bool any_shape = false);

// CFL reduction workspace: a ticket counter + one record per workgroup.  The
// counter must be zero before the first launch; the kernels leave it zero.
// Sized for both the stand-alone kernel and the x-sweep fluid kernel that
// fuses the reduction (launch_fluidx with cfl != null).
size_t cfl_workspace_bytes(int nx, int ny, int nz);
// Workgroups of launch_fluidx for pencils of length nx.
size_t fluid_workgroups(int nx, long npen);
// Writes inv_c[0] = 1/c (float, the reference's calcfl return value).  When
// dt_dev is non-null also applies the driver's dt rule on the device:
//   dt = (float)(cfl * (double)inv_c); if (tf > 0 && dt > (tf - t)/2) dt = (tf - t)/2;
//   t = t + 2*dt     (t_dev may be null: then no clamp and no t update)
// clear_ticket: zero the workspace header first (a kernel), for a workspace
// whose counter is not known to be zero.
int launch_calcfl(float *inv_c, const float *u, const float *b, int nx, int ny, int nz, void *ws,
                  double cfl, float *dt_dev, float *t_dev, float tf, hipStream_t s,
                  bool clear_ticket = false);

// The no-allocation calcfl's fixed workspace (mhd_cfl.hip): its size, the
// largest box it serves, and its address on the current device (nullptr, with
// *err set, if the lookup failed).
size_t noalloc_scratch_bytes();
size_t noalloc_max_cells();
void *noalloc_scratch(hipError_t *err);

// copy_matrix on the device (mhd_copy.hip): dst0 <- src0 (n0 floats) and
// dst1 <- src1 (n1 floats) in one streaming launch.
int stream_copy2(float *dst0, const float *src0, size_t n0, float *dst1, const float *src1, size_t n1,
                 hipStream_t s);

// Seeded initial condition (mhd_ic.hip).
int ic_mode_count(int kmax);
int ic_generate_host(float *u, float *b, int nx, int ny, int nz, uint64_t seed, double amp_v,
                     double amp_a, double rho0, double p0, double gamma, int kmax);
int ic_generate_device(float *u, float *b, int nx, int ny, int nz, uint64_t seed, double amp_v,
                       double amp_a, double rho0, double p0, double gamma, int kmax, double *A,
                       hipStream_t s);

// FP64 diagnostics (mhd_ic.hip).
size_t diag_workspace_bytes(size_t ncell);
int diagnostics(const float *u, const float *b, int nx, int ny, int nz, void *ws, double *out,
                hipStream_t s);

}  // namespace mhd
