// Runs the host code that needs no GPU under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by tests/c/sanitize.mk): the seeded host
// IC generator, every oracle function on small cubic and non-cubic boxes,
// and the text formats on good and malformed files.  Exits non-zero on the
// first failed check; the sanitizers abort on any error they find.
//
//   sanitize_f32 <scratch dir>     (REAL = float)
//   sanitize_f64 <scratch dir>     (REAL = double)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mhd_hip.h"

typedef REAL real;

namespace mhd {
int ic_generate_host(float *u, float *b, int nx, int ny, int nz, uint64_t seed, double amp_v, double amp_a,
                     double rho0, double p0, double gamma, int kmax);
}

extern "C" {
void mhdref_fluidx(real *u, real *b, int nx, int ny, int nz, real dt, real *uu, real *bu);
void mhdref_advectbyzx(real *b, int nx, int ny, int nz, real dt, real *uu, real *bu);
void mhdref_transpose12(real *u, real *b, real *uu, real *bu, int nx, int ny, int nz);
int mhdref_transpose13(real *u, real *b, real *uu, real *bu, int nx, int ny, int nz);
void mhdref_transpose13_any(real *u, real *b, real *uu, real *bu, int nx, int ny, int nz);
real mhdref_calcfl(const real *u, const real *b, int nx, int ny, int nz);
real mhdref_step_any(real *u, real *b, real *uu, real *bu, int nx, int ny, int nz, double cfl, real tf, real *t);
void mhdref_diagnostics(const real *u, const real *b, int nx, int ny, int nz, double *sums, double *divb);
}

static int fails = 0;
#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            fails++;                                                             \
        }                                                                        \
    } while (0)

static void oracle_case(int nx, int ny, int nz)
{
    const size_t n = (size_t)nx * ny * nz;
    std::vector<float> uf(5 * n), bf(3 * n);
    CHECK(mhd::ic_generate_host(uf.data(), bf.data(), nx, ny, nz, 7 + nx, 0.2, 0.05, 1.0, 0.6, 5.0 / 3.0, 4) == MHD_OK);
    std::vector<real> u(uf.begin(), uf.end()), b(bf.begin(), bf.end()), uu(u), bu(b), u2(5 * n), b2(3 * n);
    for (size_t i = 0; i < 5 * n; i++) CHECK(std::isfinite((double)u[i]));
    const real dt = (real)0.05;
    mhdref_fluidx(u2.data(), b.data(), nx, ny, nz, dt, uu.data(), bu.data());
    b2 = b;
    mhdref_advectbyzx(b2.data(), nx, ny, nz, dt, uu.data(), bu.data());
    mhdref_transpose12(u2.data(), b2.data(), uu.data(), bu.data(), nx, ny, nz);
    mhdref_transpose13_any(u2.data(), b2.data(), uu.data(), bu.data(), nx, ny, nz);
    CHECK(mhdref_transpose13(u2.data(), b2.data(), uu.data(), bu.data(), nx, ny, nz) == (nx == nz ? 0 : -1) ||
          nx != nz);
    CHECK(mhdref_calcfl(u.data(), b.data(), nx, ny, nz) > 0);
    double s0[5], d0[2], s1[5], d1[2];
    mhdref_diagnostics(u.data(), b.data(), nx, ny, nz, s0, d0);
    real t = 0;
    for (int s = 0; s < 3; s++) mhdref_step_any(u.data(), b.data(), uu.data(), bu.data(), nx, ny, nz, 0.7, 0, &t);
    mhdref_diagnostics(u.data(), b.data(), nx, ny, nz, s1, d1);
    CHECK(std::fabs(s1[0] - s0[0]) <= 1e-5 * std::fabs(s0[0]));
    CHECK(d1[0] < 1e-4);
    CHECK(t > 0);
}

static std::string slurp(const std::string &p)
{
    std::string s;
    FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) return s;
    char buf[4096];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
    std::fclose(f);
    return s;
}

static void spit(const std::string &p, const std::string &text)
{
    FILE *f = std::fopen(p.c_str(), "wb");
    std::fwrite(text.data(), 1, text.size(), f);
    std::fclose(f);
}

static void text_formats(const std::string &dir)
{
    const int nx = 5, ny = 4, nz = 3;
    const size_t n = (size_t)nx * ny * nz;
    std::vector<float> u(5 * n), b(3 * n), u2(5 * n), b2(3 * n);
    CHECK(mhd::ic_generate_host(u.data(), b.data(), nx, ny, nz, 3, 0.3, 0.1, 1.0, 0.6, 5.0 / 3.0, 4) == MHD_OK);
    u[7] = -0.0f;
    u[8] = 1e-38f;  // a subnormal-adjacent value
    const std::string ic = dir + "/source_init_alfvenlinear.dat";
    CHECK(mhd_write_source_init(ic.c_str(), u.data(), b.data(), nx, ny, nz) == MHD_OK);
    CHECK(mhd_read_source_init(ic.c_str(), u2.data(), b2.data(), nx, ny, nz) == MHD_OK);
    CHECK(std::memcmp(u.data(), u2.data(), 4 * u.size()) == 0 && std::memcmp(b.data(), b2.data(), 4 * b.size()) == 0);
    CHECK(mhd_write_check_dumps(dir.c_str(), u.data(), b.data(), nx, ny, nz) == MHD_OK);
    CHECK(!slurp(dir + "/C_version_check_u1.dat").empty() && !slurp(dir + "/C_version_check_b3.dat").empty());
    // malformed inputs are refused, never read past
    const std::string full = slurp(ic), bad = dir + "/bad.dat";
    spit(bad, full.substr(0, full.size() / 2));  // truncated
    CHECK(mhd_read_source_init(bad.c_str(), u2.data(), b2.data(), nx, ny, nz) == MHD_ERR_IO);
    spit(bad, "");  // empty
    CHECK(mhd_read_source_init(bad.c_str(), u2.data(), b2.data(), nx, ny, nz) == MHD_ERR_IO);
    spit(bad, "1.0 2.0 abc 4.0\n");  // a token that is not a number
    CHECK(mhd_read_source_init(bad.c_str(), u2.data(), b2.data(), nx, ny, nz) == MHD_ERR_IO);
    std::string spaced;  // every value on its own line, with extra blanks and exponent forms
    for (size_t i = 0; i < 8 * n; i++) spaced += (i % 3 ? "  1.5e+00\n\n" : "\t-2.25E-1 ");
    spit(bad, spaced);
    CHECK(mhd_read_source_init(bad.c_str(), u2.data(), b2.data(), nx, ny, nz) == MHD_OK);
    CHECK(u2[0] == -0.225f && b2[3 * n - 1] == ((8 * n - 1) % 3 ? 1.5f : -0.225f));
    CHECK(mhd_read_source_init((dir + "/missing.dat").c_str(), u2.data(), b2.data(), nx, ny, nz) == MHD_ERR_IO);
    CHECK(mhd_write_check_dumps((dir + "/no/such/dir").c_str(), u.data(), b.data(), nx, ny, nz) == MHD_ERR_IO);
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <scratch dir>\n", argv[0]);
        return 2;
    }
    const int dims[][3] = {{7, 5, 3}, {12, 10, 6}, {16, 16, 16}, {1, 4, 4}, {9, 1, 2}};
    for (const auto &d : dims) oracle_case(d[0], d[1], d[2]);
    text_formats(argv[1]);
    if (fails) {
        std::fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    std::printf("sanitize harness ok (%zu-byte reals)\n", sizeof(real));
    return 0;
}
