/*
 * mhd_hip.h -- C ABI of libmhd_hip.so, an MI355X (gfx950) implementation of
 * the relaxing-TVD, dimensionally split, constrained-transport MHD step of
 * uelipen/Heterogeneous_MHD.
 *
 * Part 1 is the reference's own call surface: the prototypes of
 * C/spu_function.h:1-15 with C linkage and identical argument meaning, so
 * a C driver written against the reference (C/mhd.c) links against this
 * library unchanged.  State pointers may be device pointers (hipMalloc, the
 * fast path) or ordinary host pointers (staged through device memory).
 * These calls are synchronous, like the reference's.
 *
 * Part 2 adds what the reference lacks: status codes, explicit HIP streams,
 * a device-resident dt, the full step driven on the device (optionally
 * replayed from a captured HIP graph), the seeded initial condition and the
 * FP64 diagnostics.
 *
 * State layout (C/array_definition.h:12): u[((k*ny+j)*nx+i)*5+c] holds
 * (rho, m1, m2, m3, E); b[((k*ny+j)*nx+i)*3+c] holds the face-staggered
 * field, b_d at index i on the low face of cell i along axis d.  FP32.
 */
#ifndef MHD_HIP_H
#define MHD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Part 1: the reference interface (C/spu_function.h:1-15)             */
/* ------------------------------------------------------------------ */

/* replaces fluidx, C/fluidx.c:19 -- fluid sweep along the first axis:
 * reads update_u/update_b, writes u.  iter is accepted and unused, as in
 * the reference. */
void fluidx(float *u, float *b, int *nx, int *ny, int *nz, float *dt, float *update_u,
            float *update_b, int iter);

/* replaces copy_matrix, C/copy_matrix.c:9 -- update <- main, all 8 comps. */
void copy_matrix(float *u, float *b, int *nx, int *ny, int *nz, float *update_u, float *update_b);

/* replaces advectbyzx, C/advectbyzx.c:14 -- CT advection of b2, b3 along
 * the first axis: reads update_u, update_b(2,3) and main b(1), writes b. */
void advectbyzx(float *u, float *b, int *nx, int *ny, int *nz, float *dt, float *update_u,
                float *update_b);

/* replaces transpose12, C/transpose12.c:7 -- x<->y swap of update_* into
 * u, b (out dims (ny,nx,nz)).  iter is unused. */
void transpose12(float *u, float *b, float *update_u, float *update_b, int *nx, int *ny, int *nz,
                 int iter);

/* replaces transpose13, C/transpose13.c:9 -- x<->z swap (out dims
 * (nz,ny,nx)); needs nx == nz, otherwise prints the reference's message
 * and writes nothing. */
void transpose13(float *u, float *b, float *update_u, float *update_b, int *nx, int *ny, int *nz);

/* replaces calcfl, C/calcfl.c:13 -- returns 1/c, c the global maximum of
 * the fast-speed bound (bit-exact). */
float calcfl(float *u, float *b, int *nx, int *ny, int *nz);

/* ------------------------------------------------------------------ */
/* Part 2: extensions                                                  */
/* ------------------------------------------------------------------ */

#define MHD_OK 0
#define MHD_ERR_ARG -1       /* bad dimensions or null pointer */
#define MHD_ERR_SHAPE -2     /* transpose13 (the reference form) with nx != nz */
#define MHD_ERR_HIP -3       /* a HIP runtime call failed */
#define MHD_ERR_TOO_LONG -4  /* pencil longer than mhd_max_pencil() */
#define MHD_ERR_NOMEM -5     /* device allocation failed */
#define MHD_ERR_IO -6        /* a file could not be opened, read or written */
#define MHD_ERR_SCRATCH -7   /* the box exceeds the fixed scratch of mhd_calcfl_noalloc_async */

typedef void *mhd_stream_t; /* a hipStream_t; NULL is the null stream */

/* Arithmetic of the sweep kernels in plans and ensembles.  EXACT (the
 * default) reproduces the reference bitwise.  FAST keeps the scheme and
 * computes in FP32 only (hardware reciprocal and square root, FMA); its
 * fields differ from the reference's by FP32 rounding noise, within the
 * relative tolerance the library's tests state.  Transposes and calcfl are
 * exact in both modes, and so is the calcfl that plans and ensembles fuse
 * into the last sweep of a step (the next step's dt): in FAST it screens the
 * cells in FP32 and evaluates the reference's bound, bitwise, on every cell
 * that can hold the maximum.  Switching the mode of a plan or ensemble
 * recomputes calcfl of its state before the next step. */
#define MHD_PREC_EXACT 0
#define MHD_PREC_FAST 1

/* Library version string and the longest pencil the sweeps accept. */
const char *mhd_version(void);
int mhd_max_pencil(void);
/* Last HIP error text seen by the library ("" if none). */
const char *mhd_last_error(void);

/* Asynchronous device-pointer forms of Part 1.  dt_dev, when non-null, is a
 * device pointer the kernels read dt from (so a captured graph picks up the
 * current step's dt); otherwise dt is used.  advectbyzx's b must not alias
 * update_b; b_main may alias b.  Transposes must not alias in and out. */
int mhd_fluidx_async(float *u, const float *update_u, const float *update_b, int nx, int ny, int nz,
                     const float *dt_dev, float dt, mhd_stream_t stream);
int mhd_advectbyzx_async(float *b, const float *b_main, const float *update_u,
                         const float *update_b, int nx, int ny, int nz, const float *dt_dev,
                         float dt, mhd_stream_t stream);
/* The same two sweeps in MHD_PREC_FAST arithmetic (no reference counterpart:
 * the reference has only the exact form). */
int mhd_fluidx_fast_async(float *u, const float *update_u, const float *update_b, int nx, int ny,
                          int nz, const float *dt_dev, float dt, mhd_stream_t stream);
int mhd_advectbyzx_fast_async(float *b, const float *b_main, const float *update_u,
                              const float *update_b, int nx, int ny, int nz, const float *dt_dev,
                              float dt, mhd_stream_t stream);
/* advectbyzx (b from update_b, with b_main = update_b) followed by fluidx
 * (u from update_u and the new b) in one kernel, bitwise equal to the two
 * calls -- the pairs at C/mhd.c:95-97, :105-107, :115-117 (no reference
 * counterpart as one call).  u must not alias update_u, b not update_b.
 * precision: MHD_PREC_EXACT or MHD_PREC_FAST. */
int mhd_advect_fluidx_async(float *u, float *b, const float *update_u, const float *update_b, int nx,
                            int ny, int nz, const float *dt_dev, float dt, int precision,
                            mhd_stream_t stream);
/* advectbyzx twice with the same update_u (b from update_b, then from that
 * result), then fluidx (u from update_u and the final b), in one kernel,
 * bitwise equal to the three calls -- the middle of the step, C/mhd.c:93
 * and :95-97 (no reference counterpart as one call).  The intermediate field
 * is not written.  u must not alias update_u, b not update_b.
 * precision: MHD_PREC_EXACT or MHD_PREC_FAST. */
int mhd_advect2_fluidx_async(float *u, float *b, const float *update_u, const float *update_b,
                             int nx, int ny, int nz, const float *dt_dev, float dt, int precision,
                             mhd_stream_t stream);
/* advectbyzx (the new b from update_b, with b_main = update_b) followed by
 * transpose12 of update_u and that b, in one kernel, bitwise equal to the
 * two calls -- C/mhd.c:74 and :78 (no reference counterpart as one call).
 * u, b receive the transposed state, dims (ny, nx, nz); the new b in the
 * input layout is not written.  No output may overlap an input.
 * precision: MHD_PREC_EXACT or MHD_PREC_FAST. */
int mhd_advect_transpose12_async(float *u, float *b, const float *update_u, const float *update_b,
                                 int nx, int ny, int nz, const float *dt_dev, float dt, int precision,
                                 mhd_stream_t stream);
/* The same with transpose13 for any nx, nz (out dims (nz, ny, nx), the
 * permutation of mhd_transpose13_any_async) -- C/mhd.c:84 and :88. */
int mhd_advect_transpose13_async(float *u, float *b, const float *update_u, const float *update_b,
                                 int nx, int ny, int nz, const float *dt_dev, float dt, int precision,
                                 mhd_stream_t stream);
/* transpose12 of (update_u, update_b) followed by advectbyzx of the
 * transposed state along its first axis (b_main = the transposed b), in one
 * kernel, bitwise equal to the two calls -- C/mhd.c:112 and :115 (no
 * reference counterpart as one call).  u receives the transposed u and b
 * the new field, dims (ny, nx, nz).  Needs nx a multiple of 16, ny a
 * multiple of 4 and 16-byte aligned pointers (MHD_ERR_SHAPE otherwise).  No
 * output may overlap an input.  precision: MHD_PREC_EXACT or MHD_PREC_FAST. */
int mhd_transpose12_advect_async(float *u, float *b, const float *update_u, const float *update_b,
                                 int nx, int ny, int nz, const float *dt_dev, float dt, int precision,
                                 mhd_stream_t stream);
/* The same with the general transpose13 (out dims (nz, ny, nx)) --
 * C/mhd.c:102 and :105.  Needs nx a multiple of 16 and nz of 4. */
int mhd_transpose13_advect_async(float *u, float *b, const float *update_u, const float *update_b,
                                 int nx, int ny, int nz, const float *dt_dev, float dt, int precision,
                                 mhd_stream_t stream);
int mhd_transpose12_async(float *u, float *b, const float *update_u, const float *update_b, int nx,
                          int ny, int nz, mhd_stream_t stream);
int mhd_transpose13_async(float *u, float *b, const float *update_u, const float *update_b, int nx,
                          int ny, int nz, mhd_stream_t stream);
/* transpose13 for any nx, nz: out dims (nz, ny, nx) with the same component
 * permutation.  No reference counterpart: C/transpose13.c:16 refuses nx != nz
 * (so the box of C/mhd.c needs ny == nz); the device plan uses this one, so it
 * also steps boxes with ny != nz. */
int mhd_transpose13_any_async(float *u, float *b, const float *update_u, const float *update_b,
                              int nx, int ny, int nz, mhd_stream_t stream);
int mhd_copy_matrix_async(const float *u, const float *b, int nx, int ny, int nz, float *update_u,
                          float *update_b, mhd_stream_t stream);

/* calcfl on the device.  workspace: mhd_calcfl_workspace_bytes() bytes of
 * device memory whose first 4 bytes are zero before its first use (the
 * kernel leaves them zero).  Writes inv_c_dev[0] = 1/c.  When dt_dev is
 * non-null it also applies the driver's rule of C/mhd.c:62-64:
 * dt = (float)(cfl * 1/c); if t_dev is non-null and tf > 0, dt is clamped to
 * (tf - t)/2, and t += 2 dt. */
size_t mhd_calcfl_workspace_bytes(int nx, int ny, int nz);
int mhd_calcfl_async(float *inv_c_dev, const float *u, const float *b, int nx, int ny, int nz,
                     void *workspace, double cfl, float *dt_dev, float *t_dev, float tf,
                     mhd_stream_t stream);
/* calcfl that allocates nothing (no counterpart in the reference): for
 * callers that cannot allocate -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Same results and
 * arguments as mhd_calcfl_async without the workspace: the library uses a
 * fixed device array of 8 MiB + 2304 bytes per GPU (held until the process
 * ends), which serves boxes of up to 2^30 cells (1024^3); larger boxes return
 * MHD_ERR_SCRATCH with a message in mhd_last_error().  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (two streams, two
 * threads) race; calls ordered on one stream do not.  Calls no hipMalloc,
 * hipFree or allocating library routine. */
int mhd_calcfl_noalloc_async(float *inv_c_dev, const float *u, const float *b, int nx, int ny, int nz,
                             double cfl, float *dt_dev, float *t_dev, float tf, mhd_stream_t stream);

/* ---- full step on the device (the sweep order of C/mhd.c:59-124) ---- */
typedef struct mhd_plan mhd_plan;

/* A plan for periodic boxes of nx*ny*nz, every axis <= 2048.  The reference
 * needs ny == nz (its transpose13 at the driver's call sites); the plan uses
 * mhd_transpose13_any and also takes ny != nz.  cfl is the CFL factor (the
 * reference uses 0.9); tf > 0 enables the reference's end-time clamp.  Owns
 * the device scalars t and dt, the CFL workspace and a third face-field
 * buffer (12 bytes per cell). */
int mhd_plan_create(mhd_plan **plan, int nx, int ny, int nz, double cfl, float tf,
                    mhd_stream_t stream);
void mhd_plan_destroy(mhd_plan *plan);
/* Device pointers of the plan's t and dt (floats). */
float *mhd_plan_t_dev(mhd_plan *plan);
float *mhd_plan_dt_dev(mhd_plan *plan);
/* Sets t (device copy, synchronous). */
int mhd_plan_set_time(mhd_plan *plan, float t);
/* Advances (u,b) by nsteps full steps.  (u2,b2) are scratch buffers of the
 * same sizes.  The state is in (u,b) on return.  use_graph != 0 captures one
 * step into a HIP graph on first use and replays it.  Asynchronous on the
 * plan's stream. */
int mhd_plan_run(mhd_plan *plan, float *u, float *b, float *u2, float *b2, int nsteps, int use_graph);
/* Selects MHD_PREC_EXACT or MHD_PREC_FAST for the plan's later steps. */
int mhd_plan_set_precision(mhd_plan *plan, int mode);
/* Changes the end time of the clamp (tf > 0: dt <= (tf - t)/2, so t lands on
 * tf and later steps have dt = 0). */
int mhd_plan_set_tf(mhd_plan *plan, float tf);
/* Tells the plan that (u, b) was modified outside its steps: the next run
 * recomputes calcfl before its first step. */
int mhd_plan_state_changed(mhd_plan *plan);
/* Copies t and dt to the host (synchronises the plan's stream). */
int mhd_plan_get_time(mhd_plan *plan, float *t, float *dt);
/* Synchronises the plan's stream. */
int mhd_plan_sync(mhd_plan *plan);

/* ---- seeded initial condition ---- */
/* rho = rho0, p = p0, gamma; velocity and vector potential A are sums of
 * random-phase Fourier modes with integer wavevector 0 < |m| <= kmax
 * (one of each +-m pair), per-component RMS amp_v (velocity, in units of
 * the cell) and amp_a (A, in units where the box's first side is 2*pi);
 * b = discrete curl of A on the staggered faces (div b = 0 to round-off);
 * E = p0/(gamma-1) + |m|^2/(2 rho) + |Bc|^2/2 with Bc the cell-centred face
 * average the fluid sweep uses.  Host and device forms give bitwise equal
 * arrays.  scratch_dev: 3*nx*ny*nz doubles of device memory (for A). */
int mhd_ic_mode_count(int kmax);
int mhd_ic_generate_host(float *u, float *b, int nx, int ny, int nz, uint64_t seed, double amp_v,
                         double amp_a, double rho0, double p0, double gamma, int kmax);
int mhd_ic_generate(float *u_dev, float *b_dev, int nx, int ny, int nz, uint64_t seed, double amp_v,
                    double amp_a, double rho0, double p0, double gamma, int kmax,
                    double *scratch_dev, mhd_stream_t stream);

/* ---- FP64 diagnostics on the device ---- */
/* out[0..4]: sums of rho, m1, m2, m3, E in double; out[5] = max |div b|;
 * out[6] = RMS div b.  workspace: mhd_diag_workspace_bytes() bytes.
 * Deterministic (fixed reduction order).  Synchronous: out is host memory. */
size_t mhd_diag_workspace_bytes(int nx, int ny, int nz);
int mhd_diagnostics(const float *u_dev, const float *b_dev, int nx, int ny, int nz, void *workspace,
                    double *out, mhd_stream_t stream);

/* ---- batched ensembles of cubic boxes (no counterpart in the reference) ---- */
/* Boxes of sides sides[0..nbox-1] (<= 1024) advanced together: one
 * persistent launch per sweep covers all boxes, each box keeps its own t and
 * CFL dt, and a step can be replayed from one captured HIP graph.  Each box
 * evolves bitwise as a single-box plan of the same size would. */
typedef struct mhd_ensemble mhd_ensemble;
int mhd_ensemble_create(mhd_ensemble **ens, int nbox, const int *sides, double cfl, float tf,
                        mhd_stream_t stream);
void mhd_ensemble_destroy(mhd_ensemble *ens);
/* Packed state sizes (floats) and each box's float offset into them. */
int mhd_ensemble_sizes(mhd_ensemble *ens, size_t *u_floats, size_t *b_floats, long long *u_off,
                       long long *b_off);
/* Binds the packed device state (u, b) and the scratch pair (u2, b2), all
 * 16-byte aligned, of the sizes mhd_ensemble_sizes reports. */
int mhd_ensemble_bind(mhd_ensemble *ens, float *u, float *b, float *u2, float *b2);
/* Seeded IC of one box into the bound state (see mhd_ic_generate);
 * scratch_dev holds 3*n^3 doubles for that box's side n. */
int mhd_ensemble_init_box(mhd_ensemble *ens, int box, uint64_t seed, double amp_v, double amp_a,
                          double rho0, double p0, double gamma, int kmax, double *scratch_dev);
/* Sets every box's t to 0 (synchronous). */
int mhd_ensemble_reset_time(mhd_ensemble *ens);
/* nsteps steps of every box (async on the stream); use_graph replays one
 * captured step. */
int mhd_ensemble_run(mhd_ensemble *ens, int nsteps, int use_graph);
/* Per-box t and dt of the last step (synchronises). */
int mhd_ensemble_get_times(mhd_ensemble *ens, float *t, float *dt);
/* Launch shape (threads per workgroup, cells per thread, threads per
 * pencil) and persistent grid size.  mhd_ensemble_tune times `reps` steps of
 * every valid candidate on the bound state, keeps the fastest and restores
 * the state; *best_ms receives its time per step. */
int mhd_ensemble_num_configs(void);
int mhd_ensemble_set_config(mhd_ensemble *ens, int shape, int grid);
int mhd_ensemble_get_config(mhd_ensemble *ens, int *shape, int *grid, int *block,
                            int *cells_per_thread, int *threads_per_pencil);
int mhd_ensemble_tune(mhd_ensemble *ens, int reps, float *best_ms);
/* Selects MHD_PREC_EXACT or MHD_PREC_FAST for the ensemble's later steps. */
int mhd_ensemble_set_precision(mhd_ensemble *ens, int mode);

/* ---- the reference's text formats (host arrays, no GPU needed) ---- */
/* replaces the IC read of C/init.c:34-52: 8 whitespace-separated floats
 * per cell (rho, m1, m2, m3, E, b1, b2, b3), cells with i outermost and k
 * innermost, converted as fscanf("%f") converts them. */
int mhd_read_source_init(const char *path, float *u, float *b, int nx, int ny, int nz);
/* The inverse, with 9 significant digits, so every float reads back as
 * itself (no reference counterpart: the reference only reads these files). */
int mhd_write_source_init(const char *path, const float *u, const float *b, int nx, int ny, int nz);
/* replaces the dumps of C/check_value.c:53-69: dir/C_version_check_u1..u5.dat
 * and _b1..b3.dat, one "%8.6e" value per line, cells with k outermost and i
 * innermost. */
int mhd_write_check_dumps(const char *dir, const float *u, const float *b, int nx, int ny, int nz);

/* ------------------------------------------------------------------ */
/* Part 3: learned coarse-grid corrector (no counterpart in the reference) */
/* ------------------------------------------------------------------ */
/* A 4-layer periodic 3x3x3 convolution network, 32 channels, GELU, whose
 * output corrects a coarse step: u += scale[c] * y[c] (c < 5) and
 * b += scale[5] * curl(y[5..7]) (a vector-potential correction, so div b
 * stays at round-off).  Activations are bf16 [V][32] (channel fastest, V =
 * nx*ny*nz in state order, physical channels rho, m1..m3, E, cell-centred
 * B1..B3, zero-padded); weights bf16 [27][32 out][32 in] with tap
 * (dz+1)*9 + (dy+1)*3 + (dx+1).  Products on BF16 MFMA, FP32 accumulate.
 * Convolutions need nx a multiple of 64 and nx*ny*nz a multiple of 256. */

/* floats per layer: 27*32*32 weights followed by 32 biases */
int mhd_corr_num_params(void);
/* mode 0: z_out = conv(x) + bias (FP32 [V][32]), a_out = bf16(GELU(z));
 *         z_out may be NULL when no backward pass follows (not stored);
 * mode 1: z_out = conv(x) + bias;
 * mode 2: a_out = bf16(conv(x) * GELU'(z_prev)) -- with the input-gradient
 *         weights of mhd_corr_weight_prep, the backward pass of a layer;
 * mode 3: as mode 0 (z_out may be NULL) for the first layer, whose input
 *         carries data in channels 0..7 only (mhd_corr_pack): reads those 8
 *         channels of x, with the packed weights of
 *         mhd_corr_weight_prep_first -- 4 taps per 32-wide MFMA K-step, 7
 *         K-steps instead of 27. */
int mhd_corr_conv(const uint16_t *x, const uint16_t *w, const float *bias, int nx, int ny, int nz,
                  int mode, float *z_out, uint16_t *a_out, const float *z_prev, mhd_stream_t stream);
/* Weight gradient of one layer: partial sums of dW[t][o][c] = sum_v dy[v][o] *
 * x[v + off_t][c] and of the bias gradient, per workgroup, into ws
 * (mhd_corr_dw_workspace_bytes()); accumulate != 0 adds to what ws holds.
 * mhd_corr_dw_reduce adds the folded sums to grad (num_params floats).
 * Deterministic. */
size_t mhd_corr_dw_workspace_bytes(void);
int mhd_corr_conv_dw(const uint16_t *dy, const uint16_t *x, int nx, int ny, int nz, float *ws,
                     int accumulate, mhd_stream_t stream);
int mhd_corr_dw_reduce(const float *ws, float *grad, mhd_stream_t stream);
/* FP32 master weights -> bf16 forward form wf and input-gradient form wb. */
int mhd_corr_weight_prep(const float *w, uint16_t *wf, uint16_t *wb, mhd_stream_t stream);
/* FP32 master weights of the first layer -> the packed bf16 form of mode 3
 * (mhd_corr_first_weight_elems() elements: [7][32 out][4 taps x 8 channels],
 * the slot past tap 26 zero). */
int mhd_corr_first_weight_elems(void);
int mhd_corr_weight_prep_first(const float *w, uint16_t *wp, mhd_stream_t stream);
/* Adam step on n parameters (gradient scaled by gscale, then zeroed). */
int mhd_corr_adam(float *p, float *g, float *m, float *v, int n, float lr, float beta1, float beta2,
                  float eps, int step, float gscale, mhd_stream_t stream);
/* State -> network input; norm[0..7] means, norm[8..15] inverse std devs. */
int mhd_corr_pack(const float *u, const float *b, int nx, int ny, int nz, const float *norm,
                  uint16_t *x, mhd_stream_t stream);
/* As mhd_corr_pack, plus derived channels formed in FP32 from the FP32
 * normalised inputs f_c (channels 0..7) before their bf16 rounding:
 *   kind 0: none (mhd_corr_pack);
 *   kind 1: channel 8+c = (sum over the 6 neighbours of f_c - 6 f_c, as
 *           the sum over d of (f_c(-e_d) + f_c(+e_d)) - 2 f_c) * norm[16+c];
 *   kind 2: channel 8+3c+d = ((f_c(-e_d) + f_c(+e_d)) - 2 f_c) * norm[16+3c+d];
 *   kind 3: channel 8+c = (f_c - bf16(f_c)) * 256.
 * Periodic neighbours; norm holds 16 + (8 or 24) floats for kinds 1 and 2. */
int mhd_corr_pack_features(const float *u, const float *b, int nx, int ny, int nz, const float *norm,
                           int kind, uint16_t *x, mhd_stream_t stream);
/* Loss = mean over voxels of |y_u - t_u|^2 + |curl(y_A) - t_B|^2 with
 * target [V][8] (normalised u correction, normalised b correction); writes
 * per-block loss sums (mhd_corr_loss_partials() doubles), the residual
 * (r_scratch, [V][8] floats) and dy = dLoss/dy (bf16 [V][32]). */
size_t mhd_corr_loss_partials(int nx, int ny, int nz);
int mhd_corr_loss_grad(const float *y, const float *target, int nx, int ny, int nz, float *r_scratch,
                       double *loss_partials, uint16_t *dy, mhd_stream_t stream);
/* Block average by r of a fine state: u over r^3 cells, each face component
 * over the r^2 fine faces of the coarse face (div b stays 0). */
int mhd_downsample(const float *u, const float *b, int nx, int ny, int nz, int r, float *uc,
                   float *bc, mhd_stream_t stream);
/* Applies the correction y (FP32 [V][32] output layer) to (u, b) -- each
 * output value limited to [-limit, limit] when limit > 0, the u part
 * projected to zero box total, as the scheme conserves the totals of rho, m
 * and E -- unless the corrected state fails a check: rho > 0 and p > 0 in
 * every cell, and the box total of kinetic plus magnetic energy (cell-centred
 * B) moves by at most etol times the box total of E, i.e. the correction
 * converts at most that share of the energy to or from heat (FP64 sums; the
 * total of E itself is unchanged by the zero-total projection).  u_cand,
 * b_cand: scratch of the state's sizes; check_ws: mhd_corr_check_bytes();
 * flag_stats: 3 device ints: [0] = this call accepted, [1] += accepted,
 * [2] += rejected.  Entirely on the device (graph-capturable). */
size_t mhd_corr_check_bytes(int nx, int ny, int nz);
int mhd_corr_apply(float *u, float *b, const float *y, int nx, int ny, int nz, const float *scale,
                   float limit, double etol, float *u_cand, float *b_cand, void *check_ws,
                   int *flag_stats, mhd_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* MHD_HIP_H */
