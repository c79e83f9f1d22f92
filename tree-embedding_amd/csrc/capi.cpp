// capi.cpp -- error state and device queries of the C ABI.
#include "tsvd_internal.h"

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
