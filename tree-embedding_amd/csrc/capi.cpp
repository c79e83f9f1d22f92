// capi.cpp -- error state and device queries of the C ABI.
#include "tsvd_internal.h"

#include <fstream>

#include <cstring>

namespace tsvd {

static thread_local char g_err[1024] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace tsvd

extern "C" const char *tsvd_last_error(void) { return tsvd::g_err; }

extern "C" int tsvd_device_info(char *name, int name_len, int64_t *hbm_total, int64_t *hbm_free)
{
    return tsvd::guard([&] {
        int dev = 0;
        TSVD_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        TSVD_HIP(hipGetDeviceProperties(&prop, dev));
        if (name && name_len > 0) {
            snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
        }
        size_t fr = 0, tot = 0;
        TSVD_HIP(hipMemGetInfo(&fr, &tot));
        if (hbm_total) *hbm_total = (int64_t)tot;
        if (hbm_free) *hbm_free = (int64_t)fr;
    });
}

// ------------------------------------------------------------ edge files
// The snapshot-file loop of Graph_dynamic.h:127-130 / 539-542:
//   while (infile.good()) { infile >> from >> to; edge_vec.push_back({from, to}); }
// An extraction that meets end-of-file before a digit fails and leaves the
// variable unchanged, so a file ending in a newline yields its last edge
// twice; dup_last = 0 drops that duplicate.
namespace {

struct IntReader {
    FILE *f;
    // 0: value read, more input follows; 1: value read, end of file; 2: failed
    int next(int32_t *out)
    {
        int c;
        do {
            c = fgetc(f);
        } while (c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f');
        if (c == EOF) return 2;
        bool neg = false;
        if (c == '-' || c == '+') {
            neg = c == '-';
            c = fgetc(f);
        }
        if (c < '0' || c > '9') {
            if (c != EOF) ungetc(c, f);
            return 2;
        }
        int64_t v = 0;
        while (c >= '0' && c <= '9') {
            v = v * 10 + (c - '0');
            c = fgetc(f);
        }
        *out = (int32_t)(neg ? -v : v);
        if (c == EOF) return 1;
        ungetc(c, f);
        return 0;
    }
};

}  // namespace

extern "C" int tsvd_read_edge_file(const char *path, int dup_last, int32_t *src, int32_t *dst, int64_t cap,
                                   int64_t *count)
{
    return tsvd::guard([&] {
        TSVD_REQUIRE(path && count, "null argument");
        FILE *f = fopen(path, "r");
        TSVD_REQUIRE(f, "cannot open %s", path);
        IntReader rd{f};
        int64_t cnt = 0;
        int32_t a = 0, b = 0;
        bool first = true;
        for (;;) {
            const int ra = rd.next(&a);
            int rb = 2;
            if (ra == 0) rb = rd.next(&b);
            const bool failed = ra != 0 || rb == 2;
            if (failed && first) break;   // empty file: the reference would push an unset pair
            if (!failed || dup_last) {
                if (src && cnt < cap) {
                    src[cnt] = a;
                    dst[cnt] = b;
                }
                cnt++;
            }
            first = false;
            if (failed || rb == 1) break;
        }
        fclose(f);
        *count = cnt;
    });
}

// ------------------------------------------------------ driver file formats
// The subset file of the drivers (Tree_u_NC.cpp:363-371):
//   while (infile2 >> node_number) { ... }   -- integers until extraction fails.
extern "C" int tsvd_read_subset_file(const char *path, int32_t *ids, int64_t cap, int64_t *count)
{
    return tsvd::guard([&] {
        TSVD_REQUIRE(path && count, "null argument");
        FILE *f = fopen(path, "r");
        TSVD_REQUIRE(f, "cannot open %s", path);
        IntReader rd{f};
        int64_t cnt = 0;
        int32_t v = 0;
        for (;;) {
            const int r = rd.next(&v);
            if (r == 2) break;
            if (ids && cnt < cap) ids[cnt] = v;
            cnt++;
            if (r == 1) break;
        }
        fclose(f);
        *count = cnt;
    });
}

// config.txt of the drivers (Tree_u_NC.cpp:338-351): one path per line, read
// with getline, so a final newline adds no line and a final line without one
// still counts.  Lines go to buf NUL-terminated one after another; *needed is
// the size that holds them all.
extern "C" int tsvd_read_lines(const char *path, char *buf, int64_t len, int32_t *count, int64_t *needed)
{
    return tsvd::guard([&] {
        TSVD_REQUIRE(path && count, "null argument");
        FILE *f = fopen(path, "rb");
        TSVD_REQUIRE(f, "cannot open %s", path);
        std::string all, line;
        int32_t n = 0;
        int c;
        bool pending = false;
        while ((c = fgetc(f)) != EOF) {
            if (c == '\n') {
                all += line;
                all += '\0';
                line.clear();
                n++;
                pending = false;
            } else {
                line += (char)c;
                pending = true;
            }
        }
        if (pending) {
            all += line;
            all += '\0';
            n++;
        }
        fclose(f);
        *count = n;
        if (needed) *needed = (int64_t)all.size();
        if (buf) {
            TSVD_REQUIRE((int64_t)all.size() <= len, "buffer of %lld bytes is too small (%zu needed)", (long long)len,
                         all.size());
            memcpy(buf, all.data(), all.size());
        }
    });
}

// The embedding file of the drivers (Tree_u_NC.cpp:63, 898-909): Eigen's
// IOFormat(StreamPrecision, DontAlignCols, ", ", "\n") of the row-major
// rows x cols matrix -- every coefficient through an ostream at its default
// precision, ", " between coefficients, "\n" between rows, none at the end.
extern "C" int tsvd_write_embedding_csv(const char *path, const double *U, int64_t rows, int32_t cols)
{
    return tsvd::guard([&] {
        TSVD_REQUIRE(path && (U || rows == 0) && rows >= 0 && cols >= 0, "bad arguments");
        std::ofstream out(path);
        TSVD_REQUIRE(out.good(), "cannot write %s", path);
        for (int64_t i = 0; i < rows; i++) {
            if (i) out << "\n";
            for (int32_t j = 0; j < cols; j++) {
                if (j) out << ", ";
                out << U[i * cols + j];
            }
        }
        out.close();
        TSVD_REQUIRE(!out.fail(), "write to %s failed", path);
    });
}

// ------------------------------------------------------- device buffers
// Plumbing for callers that stage inputs in HBM (no reference counterpart).
extern "C" int tsvd_device_alloc(int64_t bytes, void **out)
{
    return tsvd::guard([&] {
        TSVD_REQUIRE(out && bytes > 0, "bad allocation request");
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, (size_t)bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw tsvd::Error{tsvd::set_error(TSVD_ENOMEM, "hipMalloc(%lld) failed", (long long)bytes)};
        }
        *out = p;
    });
}

extern "C" int tsvd_device_free(void *p)
{
    return tsvd::guard([&] { TSVD_HIP(hipFree(p)); });
}

extern "C" int tsvd_memcpy(void *dst, const void *src, int64_t bytes, int kind)
{
    return tsvd::guard([&] {
        TSVD_REQUIRE(dst && src && bytes >= 0 && kind >= 0 && kind <= 2, "bad copy request");
        const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                             : hipMemcpyDeviceToDevice;
        TSVD_HIP(hipMemcpy(dst, src, (size_t)bytes, k));
    });
}
