// Exercise of the libscarl_h5 C ABI for a sanitizer build (AddressSanitizer + UndefinedBehavior
// Sanitizer, host code only): strided and narrowing block writes, attribute replacement, reads
// into buffers smaller than the data, truncating name buffers, and every error path. Compiled and
// run by tests/test_h5_sanitizers.py; exits non-zero on any unexpected result.
#include "../include/scarl_h5.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

static int g_bad = 0;

#define OK(x)                                                                            \
    do {                                                                                 \
        const int rc_ = (x);                                                             \
        if (rc_) {                                                                       \
            char b_[256];                                                                \
            scarl_h5_last_error(b_, sizeof b_);                                          \
            std::printf("FAIL rc=%d at line %d: %s (%s)\n", rc_, __LINE__, #x, b_);      \
            ++g_bad;                                                                     \
        }                                                                                \
    } while (0)

#define MUST_FAIL(x)                                                                     \
    do {                                                                                 \
        const int rc_ = (x);                                                             \
        char b_[8];                                                                      \
        scarl_h5_last_error(b_, sizeof b_);                                              \
        if (!rc_) {                                                                      \
            std::printf("FAIL: %s at line %d succeeded\n", #x, __LINE__);                \
            ++g_bad;                                                                     \
        }                                                                                \
    } while (0)

#define CHECK(c)                                                                         \
    do {                                                                                 \
        if (!(c)) {                                                                      \
            std::printf("FAIL: %s at line %d\n", #c, __LINE__);                          \
            ++g_bad;                                                                     \
        }                                                                                \
    } while (0)

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    const std::string path = dir + "/exercise.h5";
    scarl_h5* f = nullptr;
    OK(scarl_h5_create(path.c_str(), &f));
    const int32_t dims[3] = {5, 4, 3};
    OK(scarl_h5_create_dataset(f, "samples", dims, 64));
    OK(scarl_h5_create_dataset(f, "s32", dims, 32));

    // an FP32 block with padded rows and planes (ld1 7, ld2 42) into the FP64 dataset
    std::vector<float> a(7 * 6 * 3);
    for (size_t i = 0; i < a.size(); ++i) a[i] = static_cast<float>(i);
    const int32_t off[3] = {0, 0, 0}, cnt[3] = {5, 4, 3};
    OK(scarl_h5_write_block(f, "samples", a.data(), 32, 7, 42, off, cnt));
    // FP64 into the FP32 dataset (values exact in FP32)
    std::vector<double> d(5 * 4 * 3);
    for (size_t i = 0; i < d.size(); ++i) d[i] = 0.5 * static_cast<double>(i);
    OK(scarl_h5_write_block(f, "s32", d.data(), 64, 5, 20, off, cnt));
    MUST_FAIL(scarl_h5_write_block(f, "samples", d.data(), 64, 5, 20, dims, cnt));  // past the extent
    MUST_FAIL(scarl_h5_write_block(f, "nope", d.data(), 64, 5, 20, off, cnt));
    MUST_FAIL(scarl_h5_write_block(f, "samples", d.data(), 64, 4, 20, off, cnt));  // ld1 < count[0]
    MUST_FAIL(scarl_h5_write_block(f, "samples", d.data(), 64, 5, 19, off, cnt));  // ld2 < ld1 count[1]
    MUST_FAIL(scarl_h5_write_block(f, "samples", d.data(), 16, 5, 20, off, cnt));  // precision
    MUST_FAIL(scarl_h5_create_dataset(f, "bad", dims, 16));

    const int32_t iv[3] = {1, 2, 3};
    const double dv[3] = {0.5, -1.0, 1e300};
    OK(scarl_h5_attr_i32(f, "", "nDim", iv, 0));
    OK(scarl_h5_attr_i32(f, "", "L", iv, 3));
    OK(scarl_h5_attr_f64(f, "", "xStep", dv, 3));
    OK(scarl_h5_attr_f64(f, "samples", "xMin", dv, 3));
    OK(scarl_h5_attr_f64(f, "", "xStep", dv, 2));  // replaced by a shorter vector
    MUST_FAIL(scarl_h5_attr_f64(f, "missing", "x", dv, 3));

    char labels[3 * SCARL_H5_LABEL_LEN];
    std::memset(labels, ' ', sizeof labels);
    std::memcpy(labels, "Read_input", 10);
    std::memcpy(labels + 2 * SCARL_H5_LABEL_LEN, "Total_Wall_time", 15);
    const double times[3 * 2] = {1, 2, 3, 4, 5, 6};
    OK(scarl_h5_write_times(f, labels, times, 3, 2));
    OK(scarl_h5_close(f));

    OK(scarl_h5_open(path.c_str(), 0, &f));
    int32_t di[3] = {0, 0, 0}, fp = 0;
    OK(scarl_h5_dataset_info(f, "samples", di, &fp));
    CHECK(di[0] == 5 && di[1] == 4 && di[2] == 3 && fp == 64);
    OK(scarl_h5_dataset_info(f, "s32", di, &fp));
    CHECK(fp == 32);
    std::vector<double> r(9 * 4 * 3, -1.0);  // read into padded rows (ld1 9, ld2 36)
    OK(scarl_h5_read_block(f, "samples", r.data(), 64, 9, 36, off, cnt));
    for (int z = 0; z < 3; ++z)
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 9; ++x)
                CHECK(r[x + 9 * y + 36 * z] == (x < 5 ? static_cast<double>(a[x + 7 * y + 42 * z]) : -1.0));
    std::vector<float> rf(5 * 4 * 3);
    OK(scarl_h5_read_block(f, "s32", rf.data(), 32, 5, 20, off, cnt));
    for (size_t i = 0; i < rf.size(); ++i) CHECK(rf[i] == static_cast<float>(d[i]));
    const int32_t off2[3] = {3, 2, 1}, cnt2[3] = {2, 2, 2};
    double sub[8];
    OK(scarl_h5_read_block(f, "samples", sub, 64, 2, 4, off2, cnt2));
    CHECK(sub[7] == static_cast<double>(a[4 + 7 * 3 + 42 * 2]));
    MUST_FAIL(scarl_h5_read_block(f, "samples", sub, 64, 2, 4, off2, cnt));

    int32_t n = 0, ib[1] = {0};
    double db[1] = {0.0};
    OK(scarl_h5_read_attr_i32(f, "", "L", ib, 1, &n));  // cap below the length
    CHECK(n == 3 && ib[0] == 1);
    OK(scarl_h5_read_attr_f64(f, "", "xStep", db, 1, &n));
    CHECK(n == 2 && db[0] == 0.5);
    OK(scarl_h5_read_attr_i32(f, "", "nDim", ib, 1, &n));
    CHECK(n == 1 && ib[0] == 1);
    MUST_FAIL(scarl_h5_read_attr_f64(f, "", "absent", db, 1, &n));
    int32_t na = 0;
    OK(scarl_h5_attr_count(f, "", &na));
    CHECK(na == 3);
    for (int i = 0; i < na; ++i) {
        char nm[3];  // truncating buffer
        OK(scarl_h5_attr_name(f, "", i, nm, sizeof nm));
        CHECK(std::strlen(nm) <= 2);
    }
    MUST_FAIL(scarl_h5_attr_name(f, "", na + 5, nullptr, 0));
    double t2[2] = {0, 0};
    char l2[SCARL_H5_LABEL_LEN];
    int32_t nt = 0, np = 0;
    OK(scarl_h5_read_times(f, t2, 2, l2, 1, &nt, &np));  // caps below the data
    CHECK(nt == 3 && np == 2 && t2[0] == 1.0 && t2[1] == 2.0 && std::memcmp(l2, "Read_input", 10) == 0);
    OK(scarl_h5_close(f));
    MUST_FAIL(scarl_h5_open((dir + "/absent/absent.h5").c_str(), 0, &f));

    const double x0[3] = {0, 0, 0}, x1[3] = {10, 20, 30};
    const int32_t L[3] = {11, 21, 31};
    OK(scarl_xmf_write_elements(dir.c_str(), "exercise.xmf", "exercise.h5", ".", "samples", x0, x1, L, 64));
    const char* h5l[2] = {"a_proc_00000.h5", "a_proc_00001.h5"};
    const double c0[6] = {0, 0, 0, 5, 0, 0};
    OK(scarl_xmf_write_global((dir + "/global.xmf").c_str(), h5l, c0, x1, L, 2));

    struct {
        double v;
        const char* s;
    } fx[] = {{0.0, "             0.0000000000"}, {-0.0, "            -0.0000000000"},
              {1.25, "             1.2500000000"}, {-123456.789, "       -123456.7890000000"},
              {1e30, "*************************"}, {NAN, "                      NaN"},
              {INFINITY, "                 Infinity"}, {-INFINITY, "                -Infinity"},
              {5e-11, "             0.0000000001"}};
    for (const auto& e : fx) {
        char b[26];
        OK(scarl_fortran_fixed(e.v, 25, 10, b, sizeof b));
        CHECK(std::strcmp(b, e.s) == 0);
    }
    char small[4];
    MUST_FAIL(scarl_fortran_fixed(1.0, 25, 10, small, sizeof small));
    char tiny[2];
    scarl_h5_last_error(tiny, sizeof tiny);
    scarl_h5_last_error(tiny, 0);

    std::printf("%s (%d failures)\n", g_bad ? "FAILED" : "OK", g_bad);
    return g_bad ? 1 : 0;
}
