// mhd_io.hip -- the reference's text formats for initial conditions and state
// dumps (host code only).
//
// Reference: the IC reader of C/init.c:34-52 (8 floats per cell read with
// fscanf("%f"): rho, m1, m2, m3, E, b1, b2, b3; cells with i outermost and k
// innermost) and the dumps of C/check_value.c:53-69 (one file per component,
// C_version_check_u1..u5.dat and _b1..b3.dat, one "%8.6e" value per line,
// cells with k outermost and i innermost, i.e. memory order).
//
// The reader tokenises the file and converts each token with strtof in the C
// locale (strtof_l, so no global locale changes), which is the conversion
// fscanf("%f") performs, so the floats are the same.  The writers print in
// the C locale the process runs in (printf's default "C" unless the caller
// changed it), as the reference does.
// Files are streamed through large buffers; a 1024^3 IC (8.6 G values) is
// never held as text.
#include "mhd_hip.h"
#include "mhd_internal.h"

#include <cctype>
#include <clocale>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Reader {
    FILE *f;
    std::vector<char> buf;
    size_t pos = 0, len = 0;
    bool eof = false;
    locale_t cloc;
    explicit Reader(FILE *fp) : f(fp), buf(1 << 22), cloc(newlocale(LC_NUMERIC_MASK, "C", (locale_t)0)) {}
    ~Reader()
    {
        if (cloc) freelocale(cloc);
    }
    bool refill(size_t keep)
    {
        if (eof) return false;
        memmove(buf.data(), buf.data() + pos, keep);
        pos = 0;
        len = keep + fread(buf.data() + keep, 1, buf.size() - keep - 1, f);
        if (len == keep) eof = true;
        buf[len] = '\0';
        return len > keep;
    }
    // next whitespace-separated token, converted like fscanf("%f")
    bool next(float *out)
    {
        for (;;) {
            while (pos < len && isspace((unsigned char)buf[pos])) pos++;
            if (pos < len) break;
            if (!refill(0)) return false;
        }
        size_t end = pos;
        while (end < len && !isspace((unsigned char)buf[end])) end++;
        if (end == len && !eof) {  // the token may continue in the next chunk
            refill(len - pos);
            return next(out);
        }
        const char save = buf[end];
        buf[end] = '\0';
        char *stop = nullptr;
        *out = strtof_l(buf.data() + pos, &stop, cloc);
        buf[end] = save;
        if (stop == buf.data() + pos) return false;
        pos = end;
        return true;
    }
};

inline size_t cell(int i, int j, int k, int nx, int ny) { return ((size_t)k * ny + j) * nx + i; }

}  // namespace

extern "C" {

int mhd_read_source_init(const char *path, float *u, float *b, int nx, int ny, int nz)
{
    if (!path || !u || !b || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    FILE *f = fopen(path, "r");
    if (!f) return MHD_ERR_IO;
    Reader rd(f);
    if (!rd.cloc) {
        fclose(f);
        return MHD_ERR_IO;
    }
    int rc = MHD_OK;
    for (int i = 0; i < nx && rc == MHD_OK; i++)      // C/init.c:34
        for (int j = 0; j < ny && rc == MHD_OK; j++)
            for (int k = 0; k < nz && rc == MHD_OK; k++) {
                const size_t c = cell(i, j, k, nx, ny);
                for (int q = 0; q < 5 && rc == MHD_OK; q++)
                    if (!rd.next(&u[5 * c + q])) rc = MHD_ERR_IO;
                for (int q = 0; q < 3 && rc == MHD_OK; q++)
                    if (!rd.next(&b[3 * c + q])) rc = MHD_ERR_IO;
            }
    fclose(f);
    return rc;
}

int mhd_write_source_init(const char *path, const float *u, const float *b, int nx, int ny, int nz)
{
    if (!path || !u || !b || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    FILE *f = fopen(path, "w");
    if (!f) return MHD_ERR_IO;
    std::vector<char> obuf(1 << 22);
    setvbuf(f, obuf.data(), _IOFBF, obuf.size());
    bool ok = true;
    for (int i = 0; i < nx && ok; i++)
        for (int j = 0; j < ny && ok; j++)
            for (int k = 0; k < nz && ok; k++) {
                const size_t c = cell(i, j, k, nx, ny);
                // 9 significant digits: every float reads back as itself
                ok = fprintf(f, "%.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", u[5 * c], u[5 * c + 1],
                             u[5 * c + 2], u[5 * c + 3], u[5 * c + 4], b[3 * c], b[3 * c + 1],
                             b[3 * c + 2]) > 0;
            }
    ok = (fclose(f) == 0) && ok;
    return ok ? MHD_OK : MHD_ERR_IO;
}

int mhd_write_check_dumps(const char *dir, const float *u, const float *b, int nx, int ny, int nz)
{
    if (!dir || !u || !b || nx < 1 || ny < 1 || nz < 1) return MHD_ERR_ARG;
    static const char *names[8] = {"u1", "u2", "u3", "u4", "u5", "b1", "b2", "b3"};
    const size_t n = (size_t)nx * ny * nz;
    for (int q = 0; q < 8; q++) {
        const std::string path = std::string(dir) + "/C_version_check_" + names[q] + ".dat";
        FILE *f = fopen(path.c_str(), "w");
        if (!f) return MHD_ERR_IO;
        std::vector<char> obuf(1 << 22);
        setvbuf(f, obuf.data(), _IOFBF, obuf.size());
        const float *src = q < 5 ? u + q : b + (q - 5);
        const int stride = q < 5 ? 5 : 3;
        bool ok = true;
        for (size_t c = 0; c < n && ok; c++)  // k outer, i inner: memory order (C/check_value.c:53-55)
            ok = fprintf(f, "%8.6e\n", src[c * stride]) > 0;
        ok = (fclose(f) == 0) && ok;
        if (!ok) return MHD_ERR_IO;
    }
    return MHD_OK;
}

}  // extern "C"
