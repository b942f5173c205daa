// mhd_ic.hip -- seeded initial condition (host and device, bitwise equal) and
// the FP64 diagnostics (conserved sums, div b).
//
// Neither has a counterpart in the reference, whose IC is a text file that
// is not shipped (C/init.c:28-52) and which has no diagnostics beyond printf.
//
// IC.  Modes: every integer m with 0 < |m|^2 <= kmax^2, one of each +-m pair
// (first non-zero component positive), enumerated mz, my, mx ascending.  A
// splitmix64 stream seeded with `seed` draws six phases per mode (three for
// v, three for A), phi = 2*pi * (next >> 11) * 2^-53.  Each field is
//   f(x) = scale * sum_m cos(2*pi*(mx*x/nx + my*y/ny + mz*z/nz) + phi_m)
// with scale = amp*sqrt(2/Nm), so its grid RMS is amp when |m_d| < n_d/2.
// v is sampled at cell centres; A1 at (i, j-1/2, k-1/2), A2 at (i-1/2, j,
// k-1/2), A3 at (i-1/2, j-1/2, k) (cell edges), and scaled by nx/(2 pi) so
// that b = curl A has an amplitude independent of the resolution.  Each cos
// is the real part of a product of three per-axis unit phasors taken from
// tables (sin/cos evaluated once per table entry on the host), so the per-cell
// work is only double multiply/add in a fixed order, compiled without FMA
// contraction on both host and device: the two forms give the same bits.
#include "mhd_internal.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

namespace mhd {
namespace {

struct Cplx {
    double re, im;
};

__host__ __device__ inline Cplx cmul(Cplx a, Cplx b)
{
    Cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

struct Mode {
    int mx, my, mz, pad;
    double cph[6], sph[6];  // cos/sin of the 6 phases: v1 v2 v3 A1 A2 A3
};

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

std::vector<Mode> make_modes(int kmax, uint64_t seed)
{
    std::vector<Mode> modes;
    uint64_t s = seed;
    for (int mz = -kmax; mz <= kmax; mz++)
        for (int my = -kmax; my <= kmax; my++)
            for (int mx = -kmax; mx <= kmax; mx++) {
                const int r2 = mx * mx + my * my + mz * mz;
                if (r2 == 0 || r2 > kmax * kmax) continue;
                const bool canon = mx > 0 || (mx == 0 && my > 0) || (mx == 0 && my == 0 && mz > 0);
                if (!canon) continue;
                Mode m;
                m.mx = mx, m.my = my, m.mz = mz, m.pad = 0;
                for (int f = 0; f < 6; f++) {
                    const double ph = 2.0 * M_PI * (double)(splitmix64(s) >> 11) * 0x1.0p-53;
                    m.cph[f] = cos(ph);
                    m.sph[f] = sin(ph);
                }
                modes.push_back(m);
            }
    return modes;
}

// Per-axis phasor table: tab[(m+kmax)*2n + 2x + h] = exp(i*2*pi*m*(x - h/2)/n),
// h = 0 (cell centre) or 1 (low face).  Angle reduced exactly: pi*q/n with
// q = m*(2x - h) mod 2n.
std::vector<Cplx> make_axis(int n, int kmax)
{
    std::vector<Cplx> t((size_t)(2 * kmax + 1) * 2 * n);
    for (int m = -kmax; m <= kmax; m++)
        for (int x = 0; x < n; x++)
            for (int h = 0; h < 2; h++) {
                long q = (long)m * (2 * x - h) % (2L * n);
                if (q < 0) q += 2L * n;
                const double th = M_PI * (double)q / (double)n;
                t[(size_t)(m + kmax) * 2 * n + 2 * x + h] = {cos(th), sin(th)};
            }
    return t;
}

struct IcParams {
    int nx, ny, nz, kmax, nmodes;
    double scale_v, scale_a, rho0;
};

// Velocity at the cell centre and A at the three edges of cell (i,j,k).
__host__ __device__ inline void ic_cell(const IcParams &P, const Mode *modes, const Cplx *tx,
                                        const Cplx *ty, const Cplx *tz, int i, int j, int k,
                                        double v[3], double a[3])
{
    double sv[3] = {0.0, 0.0, 0.0}, sa[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < P.nmodes; q++) {
        const Mode &m = modes[q];
        const Cplx *ex = tx + (size_t)(m.mx + P.kmax) * 2 * P.nx + 2 * i;
        const Cplx *ey = ty + (size_t)(m.my + P.kmax) * 2 * P.ny + 2 * j;
        const Cplx *ez = tz + (size_t)(m.mz + P.kmax) * 2 * P.nz + 2 * k;
        const Cplx pc = cmul(cmul(ex[0], ey[0]), ez[0]);  // centre
        const Cplx p1 = cmul(cmul(ex[0], ey[1]), ez[1]);  // A1 edge
        const Cplx p2 = cmul(cmul(ex[1], ey[0]), ez[1]);  // A2 edge
        const Cplx p3 = cmul(cmul(ex[1], ey[1]), ez[0]);  // A3 edge
        for (int c = 0; c < 3; c++) sv[c] = sv[c] + (m.cph[c] * pc.re - m.sph[c] * pc.im);
        sa[0] = sa[0] + (m.cph[3] * p1.re - m.sph[3] * p1.im);
        sa[1] = sa[1] + (m.cph[4] * p2.re - m.sph[4] * p2.im);
        sa[2] = sa[2] + (m.cph[5] * p3.re - m.sph[5] * p3.im);
    }
    for (int c = 0; c < 3; c++) {
        v[c] = P.scale_v * sv[c];
        a[c] = P.scale_a * sa[c];
    }
}

__host__ __device__ inline size_t cidx(int i, int j, int k, int nx, int ny)
{
    return ((size_t)k * ny + j) * nx + i;
}

// b = curl A on the faces of cell (i,j,k).
__host__ __device__ inline void ic_curl(const double *A, int i, int j, int k, int nx, int ny, int nz,
                                        float *bo)
{
    const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1, kp = k + 1 == nz ? 0 : k + 1;
    const double *a0 = A + 3 * cidx(i, j, k, nx, ny);
    const double *ai = A + 3 * cidx(ip, j, k, nx, ny);
    const double *aj = A + 3 * cidx(i, jp, k, nx, ny);
    const double *ak = A + 3 * cidx(i, j, kp, nx, ny);
    bo[0] = (float)((aj[2] - a0[2]) - (ak[1] - a0[1]));
    bo[1] = (float)((ak[0] - a0[0]) - (ai[2] - a0[2]));
    bo[2] = (float)((ai[1] - a0[1]) - (aj[0] - a0[0]));
}

// Total energy of cell (i,j,k) from its stored rho, m and the cell-centred
// B of the fluid sweep (C/fluidx.c:71-81).
__host__ __device__ inline float ic_energy(const float *u, const float *b, int i, int j, int k, int nx,
                                           int ny, int nz, double p0, double gamma)
{
    const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1, kp = k + 1 == nz ? 0 : k + 1;
    const float *q = u + 5 * cidx(i, j, k, nx, ny);
    const float *b0 = b + 3 * cidx(i, j, k, nx, ny);
    const float bc1 = 0.5f * b0[0] + 0.5f * b[3 * cidx(ip, j, k, nx, ny) + 0];
    const float bc2 = 0.5f * b0[1] + b[3 * cidx(i, jp, k, nx, ny) + 1] / 2;
    const float bc3 = 0.5f * b0[2] + b[3 * cidx(i, j, kp, nx, ny) + 2] / 2;
    const double msq = (double)q[1] * q[1] + (double)q[2] * q[2] + (double)q[3] * q[3];
    const double bsq = (double)bc1 * bc1 + (double)bc2 * bc2 + (double)bc3 * bc3;
    return (float)(p0 / (gamma - 1.0) + 0.5 * msq / (double)q[0] + 0.5 * bsq);
}

__global__ void ic_pass1(IcParams P, const Mode *modes, const Cplx *tx, const Cplx *ty, const Cplx *tz,
                         float *u, double *A)
{
    const size_t n = (size_t)P.nx * P.ny * P.nz;
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = (int)(s % P.nx), j = (int)((s / P.nx) % P.ny), k = (int)(s / ((size_t)P.nx * P.ny));
    double v[3], a[3];
    ic_cell(P, modes, tx, ty, tz, i, j, k, v, a);
    float *q = u + 5 * s;
    q[0] = (float)P.rho0;
    for (int c = 0; c < 3; c++) q[1 + c] = (float)(P.rho0 * v[c]);
    for (int c = 0; c < 3; c++) A[3 * s + c] = a[c];
}

__global__ void ic_pass2(const double *A, float *b, int nx, int ny, int nz)
{
    const size_t n = (size_t)nx * ny * nz;
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = (int)(s % nx), j = (int)((s / nx) % ny), k = (int)(s / ((size_t)nx * ny));
    ic_curl(A, i, j, k, nx, ny, nz, b + 3 * s);
}

__global__ void ic_pass3(float *u, const float *b, int nx, int ny, int nz, double p0, double gamma)
{
    const size_t n = (size_t)nx * ny * nz;
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = (int)(s % nx), j = (int)((s / nx) % ny), k = (int)(s / ((size_t)nx * ny));
    u[5 * s + 4] = ic_energy(u, b, i, j, k, nx, ny, nz, p0, gamma);
}

IcParams make_params(int nx, int ny, int nz, int kmax, int nmodes, double amp_v, double amp_a,
                     double rho0)
{
    IcParams P;
    P.nx = nx, P.ny = ny, P.nz = nz, P.kmax = kmax, P.nmodes = nmodes;
    const double norm = nmodes > 0 ? sqrt(2.0 / (double)nmodes) : 0.0;
    P.scale_v = amp_v * norm;
    P.scale_a = amp_a * norm * (double)nx / (2.0 * M_PI);
    P.rho0 = rho0;
    return P;
}

template <class F>
void parallel_cells(size_t n, F f)
{
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    const size_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        const size_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([=] {
            for (size_t s = lo; s < hi; s++) f(s);
        });
    }
    for (auto &x : th) x.join();
}

}  // namespace

int ic_mode_count(int kmax)
{
    if (kmax < 0) return 0;
    return (int)make_modes(kmax, 0).size();
}

int ic_generate_host(float *u, float *b, int nx, int ny, int nz, uint64_t seed, double amp_v,
                     double amp_a, double rho0, double p0, double gamma, int kmax)
{
    if (!u || !b || nx < 1 || ny < 1 || nz < 1 || kmax < 0) return MHD_ERR_ARG;
    const std::vector<Mode> modes = make_modes(kmax, seed);
    const std::vector<Cplx> tx = make_axis(nx, kmax), ty = make_axis(ny, kmax), tz = make_axis(nz, kmax);
    const IcParams P = make_params(nx, ny, nz, kmax, (int)modes.size(), amp_v, amp_a, rho0);
    const size_t n = (size_t)nx * ny * nz;
    std::vector<double> A(3 * n);
    parallel_cells(n, [&](size_t s) {
        const int i = (int)(s % nx), j = (int)((s / nx) % ny), k = (int)(s / ((size_t)nx * ny));
        double v[3], a[3];
        ic_cell(P, modes.data(), tx.data(), ty.data(), tz.data(), i, j, k, v, a);
        float *q = u + 5 * s;
        q[0] = (float)rho0;
        for (int c = 0; c < 3; c++) q[1 + c] = (float)(rho0 * v[c]);
        for (int c = 0; c < 3; c++) A[3 * s + c] = a[c];
    });
    parallel_cells(n, [&](size_t s) {
        const int i = (int)(s % nx), j = (int)((s / nx) % ny), k = (int)(s / ((size_t)nx * ny));
        ic_curl(A.data(), i, j, k, nx, ny, nz, b + 3 * s);
    });
    parallel_cells(n, [&](size_t s) {
        const int i = (int)(s % nx), j = (int)((s / nx) % ny), k = (int)(s / ((size_t)nx * ny));
        u[5 * s + 4] = ic_energy(u, b, i, j, k, nx, ny, nz, p0, gamma);
    });
    return MHD_OK;
}

int ic_generate_device(float *u, float *b, int nx, int ny, int nz, uint64_t seed, double amp_v,
                       double amp_a, double rho0, double p0, double gamma, int kmax, double *A,
                       hipStream_t s)
{
    if (!u || !b || !A || nx < 1 || ny < 1 || nz < 1 || kmax < 0) return MHD_ERR_ARG;
    const std::vector<Mode> modes = make_modes(kmax, seed);
    const std::vector<Cplx> tx = make_axis(nx, kmax), ty = make_axis(ny, kmax), tz = make_axis(nz, kmax);
    const IcParams P = make_params(nx, ny, nz, kmax, (int)modes.size(), amp_v, amp_a, rho0);
    const size_t bytes_m = modes.size() * sizeof(Mode);
    const size_t bx = tx.size() * sizeof(Cplx), by = ty.size() * sizeof(Cplx), bz = tz.size() * sizeof(Cplx);
    char *tab = nullptr;
    if (hipMalloc(&tab, bytes_m + bx + by + bz + 64) != hipSuccess) return MHD_ERR_NOMEM;
    Mode *dm = (Mode *)tab;
    Cplx *dx = (Cplx *)(tab + bytes_m), *dy = (Cplx *)(tab + bytes_m + bx), *dz = (Cplx *)(tab + bytes_m + bx + by);
    int rc = MHD_OK;
    if (hipMemcpyAsync(dm, modes.data(), bytes_m, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dx, tx.data(), bx, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dy, ty.data(), by, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dz, tz.data(), bz, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = MHD_ERR_HIP;
    const size_t n = (size_t)nx * ny * nz;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (rc == MHD_OK) {
        (void)hipGetLastError();  // clear a stale error so the check below sees this launch
        hipLaunchKernelGGL(ic_pass1, dim3(nb), dim3(256), 0, s, P, dm, dx, dy, dz, u, A);
        (void)hipGetLastError();  // clear a stale error so the check below sees this launch
        hipLaunchKernelGGL(ic_pass2, dim3(nb), dim3(256), 0, s, A, b, nx, ny, nz);
        (void)hipGetLastError();  // clear a stale error so the check below sees this launch
        hipLaunchKernelGGL(ic_pass3, dim3(nb), dim3(256), 0, s, u, b, nx, ny, nz, p0, gamma);
        if (hipGetLastError() != hipSuccess) rc = MHD_ERR_HIP;
    }
    // the tables must outlive the kernels
    if (hipStreamSynchronize(s) != hipSuccess) rc = MHD_ERR_HIP;
    (void)hipFree(tab);
    return rc;
}

// ---------------------------------------------------------------------------
// FP64 diagnostics.  Workgroup w owns cells [w*kDRun, (w+1)*kDRun) and writes
// 7 doubles (5 sums, sum of div^2, max |div|) reduced in a fixed tree order;
// one more workgroup folds the records in index order.  Deterministic.
// ---------------------------------------------------------------------------
namespace {
constexpr int kDB = 256, kDPer = 16, kDRun = kDB * kDPer, kDOut = 7;

__global__ __launch_bounds__(kDB) void diag_partial(const float *__restrict__ u,
                                                    const float *__restrict__ b, int nx, int ny,
                                                    int nz, double *__restrict__ rec)
{
    __shared__ double sh[kDOut][kDB];
    const size_t n = (size_t)nx * ny * nz;
    double acc[kDOut] = {0, 0, 0, 0, 0, 0, 0};
    for (int r = 0; r < kDPer; r++) {
        const size_t s = (size_t)blockIdx.x * kDRun + r * kDB + threadIdx.x;
        if (s >= n) break;
        const int i = (int)(s % nx), j = (int)((s / nx) % ny), k = (int)(s / ((size_t)nx * ny));
        const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1, kp = k + 1 == nz ? 0 : k + 1;
        for (int c = 0; c < 5; c++) acc[c] += (double)u[5 * s + c];
        const double d = ((double)b[3 * cidx(ip, j, k, nx, ny) + 0] - (double)b[3 * s + 0])
                         + ((double)b[3 * cidx(i, jp, k, nx, ny) + 1] - (double)b[3 * s + 1])
                         + ((double)b[3 * cidx(i, j, kp, nx, ny) + 2] - (double)b[3 * s + 2]);
        acc[5] += d * d;
        acc[6] = fmax(acc[6], fabs(d));
    }
    for (int c = 0; c < kDOut; c++) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = kDB / 2; w >= 1; w >>= 1) {
        if (threadIdx.x < w) {
            for (int c = 0; c < 6; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
            sh[6][threadIdx.x] = fmax(sh[6][threadIdx.x], sh[6][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < kDOut) rec[(size_t)blockIdx.x * kDOut + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(kDB) void diag_final(const double *__restrict__ rec, int nrec,
                                                  double *__restrict__ out)
{
    __shared__ double sh[kDOut][kDB];
    double acc[kDOut] = {0, 0, 0, 0, 0, 0, 0};
    for (int r = threadIdx.x; r < nrec; r += kDB) {
        for (int c = 0; c < 6; c++) acc[c] += rec[(size_t)r * kDOut + c];
        acc[6] = fmax(acc[6], rec[(size_t)r * kDOut + 6]);
    }
    for (int c = 0; c < kDOut; c++) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = kDB / 2; w >= 1; w >>= 1) {
        if (threadIdx.x < w) {
            for (int c = 0; c < 6; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
            sh[6][threadIdx.x] = fmax(sh[6][threadIdx.x], sh[6][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < kDOut) out[threadIdx.x] = sh[threadIdx.x][0];
}
}  // namespace

size_t diag_workspace_bytes(size_t ncell)
{
    const size_t nrec = (ncell + kDRun - 1) / kDRun;
    return (nrec + 1) * kDOut * sizeof(double);
}

int diagnostics(const float *u, const float *b, int nx, int ny, int nz, void *ws, double *out,
                hipStream_t s)
{
    if (!u || !b || !ws || !out || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    const size_t n = (size_t)nx * ny * nz;
    const int nrec = (int)((n + kDRun - 1) / kDRun);
    double *rec = (double *)ws;
    double *res = rec + (size_t)nrec * kDOut;
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL(diag_partial, dim3(nrec), dim3(kDB), 0, s, u, b, nx, ny, nz, rec);
    (void)hipGetLastError();  // clear a stale error so the check below sees this launch
    hipLaunchKernelGGL(diag_final, dim3(1), dim3(kDB), 0, s, rec, nrec, res);
    if (hipGetLastError() != hipSuccess) return MHD_ERR_HIP;
    double h[kDOut];
    if (hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess) return MHD_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return MHD_ERR_HIP;
    for (int c = 0; c < 5; c++) out[c] = h[c];
    out[5] = h[6];
    out[6] = sqrt(h[5] / (double)n);
    return MHD_OK;
}

}  // namespace mhd
