// scarl_runtime.hip -- host runtime and C ABI of libscarl_hip.so.
//
// Planner (radix split, threads per pencil, pencils per tile, LDS pitch), twiddle pools
// (computed in long double, rounded once to the kernel precision), field descriptors, tile
// tables, per-batch hipGraphs, event timing and the extern "C" entry points of scarl_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scarl_hip.h"
#include "scarl_device.h"
#include "scarl_internal.h"

namespace scarl {

// launch wrappers (scarl_kernels.hip)
hipError_t launch_h_strided_f32(int axis, bool gen, int spec, int ecap, const FieldH* fields, const Tile* tiles,
                                int tile0, int ntiles, const PlanDev* plans, const void* pool, size_t lds,
                                hipStream_t s);
hipError_t launch_h_strided_f64(int axis, bool gen, int spec, int ecap, const FieldH* fields, const Tile* tiles,
                                int tile0, int ntiles, const PlanDev* plans, const void* pool, size_t lds,
                                hipStream_t s);
hipError_t launch_h_kx0_f32(const FieldH* fields, const int* ids, int nfields, int64_t max_pts, hipStream_t s);
hipError_t launch_h_kx0_f64(const FieldH* fields, const int* ids, int nfields, int64_t max_pts, hipStream_t s);
hipError_t launch_h_xpass_f32(int spec, int ecap, bool blend, bool full, const FieldH* fields, const Tile* tiles, int tile0,
                              int ntiles, const PlanDev* plans, const void* pool, MomPartial* partials, size_t lds,
                              hipStream_t s);
hipError_t launch_h_xpass_f64(int spec, int ecap, bool blend, bool full, const FieldH* fields, const Tile* tiles, int tile0,
                              int ntiles, const PlanDev* plans, const void* pool, MomPartial* partials, size_t lds,
                              hipStream_t s);

static hipError_t launch_h_strided(bool fp64, int axis, bool gen, int spec, int ecap, const FieldH* fields,
                                   const Tile* tiles, int tile0, int ntiles, const PlanDev* plans, const void* pool,
                                   size_t lds, hipStream_t s) {
    return fp64 ? launch_h_strided_f64(axis, gen, spec, ecap, fields, tiles, tile0, ntiles, plans, pool, lds, s)
                : launch_h_strided_f32(axis, gen, spec, ecap, fields, tiles, tile0, ntiles, plans, pool, lds, s);
}
static hipError_t launch_h_xpass(bool fp64, int spec, int ecap, bool blend, bool full, const FieldH* fields, const Tile* tiles,
                                 int tile0, int ntiles, const PlanDev* plans, const void* pool, MomPartial* partials,
                                 size_t lds, hipStream_t s) {
    return fp64 ? launch_h_xpass_f64(spec, ecap, blend, full, fields, tiles, tile0, ntiles, plans, pool, partials, lds, s)
                : launch_h_xpass_f32(spec, ecap, blend, full, fields, tiles, tile0, ntiles, plans, pool, partials, lds, s);
}
hipError_t launch_mom(const FieldH* fields, int nfields, const int2* field_tiles, const MomPartial* partials,
                      MomPartial* field_mom, const double* slot_count, int nslots, double* slot_out, hipStream_t s);
hipError_t launch_spectrum_sum(const FieldH& f, MomPartial* out, int nblocks, hipStream_t s);
hipError_t launch_r_pass(int axis, int spec, const FieldR& f, const PlanDev* plans, const void* pool, int nblocks,
                         size_t lds_bytes, hipStream_t s);
hipError_t launch_marginal(double* x, int64_t n, int margi, double mu, double sd, hipStream_t s);
hipError_t launch_pou(double* x, int n1, int n2, int n3, int3 lo, int3 hi, hipStream_t s);
hipError_t launch_sums(const double* rf, int64_t n, int nmc, int nblocks, MomPartial* partial, hipStream_t s);
hipError_t launch_lag(bool fp64, const LagJob* jobs, int njobs, const LagBlock* blocks, int64_t nblocks, size_t lds,
                      const LagBlock* fblocks, int64_t nfblocks, size_t flds, const PlanDev* plans, const void* pool,
                      int nlag_max, double* out, hipStream_t s);
hipError_t launch_box_sums(const double* x, const int n[3], const int box[3], double centre, MomPartial* partial,
                           int nblocks, hipStream_t s);
hipError_t launch_normalize(double* x, int64_t n, double a, double b, hipStream_t s);
hipError_t launch_lcg_uniform(int32_t seed0, int64_t n, double* r, int32_t* seeds, hipStream_t s);
hipError_t launch_target_corr(int model, double nu, const double* rho, int n, double* out, hipStream_t s);
hipError_t launch_kv(double nu, const double* x, int n, double* out, hipStream_t s);
hipError_t launch_build_sk(const void* params, void* sk, int64_t tot, hipStream_t s);
hipError_t launch_shino_place(const int* dims9, const void* sk, void* c, hipStream_t s);
hipError_t launch_c2c_xcrop(const PlanDev* plans, int plan, int T, const void* pool, const void* c, int c1, int c2,
                            int np1, int np2, int np3, double* out, size_t lds, hipStream_t s);
// static scratch of the allocation-free entry point (scarl_noalloc.hip)
hipError_t noalloc_setup(int device);
void* noalloc_scratch(int device, size_t* bytes);
const char* noalloc_error(int device);
// host mirror of the kernel parameter block of k_build_sk (scarl_k_r.hip)
struct SkParamsHost {
    int s1, sz2, sz3;
    int gaussian;
    int32_t seed;
    double de[3];
    double scale;
};


// ------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(SCARL_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                    \
    } while (0)

#define SCARL_TRY(expr)          \
    do {                         \
        int rc_ = (expr);        \
        if (rc_) return rc_;     \
    } while (0)

// ------------------------------------------------------------------------------------------
// context: device, stream, timing, scratch buffers for the plain entry points
// ------------------------------------------------------------------------------------------
struct TimedLaunch {
    std::string name;
    hipEvent_t a, b;
};

struct Context {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t gen_stream = nullptr;  // generating passes of a pipelined batch (default priority)
    // launches of one pass beside each other in batches that do not pipeline (default priority;
    // with the two above, the four hardware queues of the process)
    hipStream_t side[2] = {nullptr, nullptr};
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::map<std::string, std::pair<long, double>> totals;
    std::map<std::string, std::pair<void*, size_t>> scratch;  // named grow-only device buffers
    int margi = 0;
    double m_avg = 0.0, m_std = 1.0;
    int fp32_ecap = 16;  // register budget (complex elements per thread) of new FP32 plans
    int mfma_dft16 = 0;       // FP32 radix-16 first stages on the matrix cores: 1 all passes, 2 x passes
    bool autotune = false;    // time every plan's candidates at batch creation
    bool lag_fft = true;      // periodic lag sums by FFT where the length plans (else direct products)
    int schedule = 0;         // batches of several groups: 0 by cache fit (autotuned), 1 pipelined, 2 sequential
    void* stage[2] = {nullptr, nullptr};         // pinned staging buffers for caller arrays
    hipEvent_t stage_ev[2] = {nullptr, nullptr};  // last DMA that used each staging buffer
};

static Context g_ctx;

static int dmalloc(void** p, size_t bytes) {
    const char* fi = getenv("SCARL_FAULT_INJECT");
    if (fi && strcmp(fi, "alloc") == 0) return fail(SCARL_E_NOMEM, "device allocation failed (fault injection)");
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(SCARL_E_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    return 0;
}

static int ensure_init() {
    if (g_ctx.ready) return 0;
    const char* fi = getenv("SCARL_FAULT_INJECT");
    if (fi && strcmp(fi, "init") == 0) return fail(SCARL_E_HIP, "no HIP device (fault injection)");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(SCARL_E_HIP, "no HIP device available");
    if (g_ctx.device >= n) return fail(SCARL_E_INVALID, "device %d out of range (%d devices)", g_ctx.device, n);
    HIP_TRY(hipSetDevice(g_ctx.device));
    // the allocation-free entry point's scratch address on this device; a failure here is kept
    // and reported by that entry point only
    (void)noalloc_setup(g_ctx.device);
    HIP_TRY(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&g_ctx.gen_stream, hipStreamNonBlocking));
    for (hipStream_t& st : g_ctx.side) HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    g_ctx.ready = true;
    return 0;
}

// Copies between caller-owned (pageable) host arrays and device memory go through two pinned
// staging buffers: the DMA of one chunk overlaps the host copy of the other (up to 8 threads),
// about 4x the throughput of a direct pageable copy.  Both calls are ordered on the library stream
// and return when the host side is done (copy_to_device: the data may still be in flight to the
// device, ahead of any later work on the stream).
constexpr size_t kStageBytes = size_t(32) << 20;
constexpr size_t kStageMin = size_t(4) << 20;

static int stage_ready() {
    if (g_ctx.stage[0]) return 0;
    for (int i = 0; i < 2; ++i) {
        HIP_TRY(hipHostMalloc(&g_ctx.stage[i], kStageBytes, hipHostMallocDefault));
        HIP_TRY(hipEventCreateWithFlags(&g_ctx.stage_ev[i], hipEventDisableTiming));
    }
    return 0;
}

static void host_copy(void* dst, const void* src, size_t bytes) {
    const int nt = bytes >= (size_t(8) << 20) ? 8 : 1;
    if (nt == 1) {
        memcpy(dst, src, bytes);
        return;
    }
    std::vector<std::thread> th;
    const size_t part = (bytes + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const size_t o = (size_t)t * part;
        if (o >= bytes) break;
        const size_t n = std::min(part, bytes - o);
        th.emplace_back([=] { memcpy(static_cast<char*>(dst) + o, static_cast<const char*>(src) + o, n); });
    }
    for (auto& x : th) x.join();
}

static int copy_to_host(void* h, const void* d, size_t bytes) {
    if (bytes < kStageMin) {
        HIP_TRY(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, g_ctx.stream));
        HIP_TRY(hipStreamSynchronize(g_ctx.stream));
        return 0;
    }
    SCARL_TRY(stage_ready());
    const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
    auto enqueue = [&](size_t c) -> int {
        const size_t o = c * kStageBytes, n = std::min(kStageBytes, bytes - o);
        HIP_TRY(hipMemcpyAsync(g_ctx.stage[c & 1], static_cast<const char*>(d) + o, n, hipMemcpyDeviceToHost,
                               g_ctx.stream));
        HIP_TRY(hipEventRecord(g_ctx.stage_ev[c & 1], g_ctx.stream));
        return 0;
    };
    SCARL_TRY(enqueue(0));
    for (size_t c = 0; c < nch; ++c) {
        // the other buffer's host copy finished in the previous iteration
        if (c + 1 < nch) SCARL_TRY(enqueue(c + 1));
        HIP_TRY(hipEventSynchronize(g_ctx.stage_ev[c & 1]));
        const size_t o = c * kStageBytes;
        host_copy(static_cast<char*>(h) + o, g_ctx.stage[c & 1], std::min(kStageBytes, bytes - o));
    }
    return 0;
}

static int copy_to_device(void* d, const void* h, size_t bytes) {
    if (bytes < kStageMin) {
        HIP_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, g_ctx.stream));
        HIP_TRY(hipStreamSynchronize(g_ctx.stream));
        return 0;
    }
    SCARL_TRY(stage_ready());
    const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
    for (size_t c = 0; c < nch; ++c) {
        const size_t o = c * kStageBytes, n = std::min(kStageBytes, bytes - o);
        HIP_TRY(hipEventSynchronize(g_ctx.stage_ev[c & 1]));  // the buffer's previous DMA is done
        host_copy(g_ctx.stage[c & 1], static_cast<const char*>(h) + o, n);
        HIP_TRY(hipMemcpyAsync(static_cast<char*>(d) + o, g_ctx.stage[c & 1], n, hipMemcpyHostToDevice,
                               g_ctx.stream));
        HIP_TRY(hipEventRecord(g_ctx.stage_ev[c & 1], g_ctx.stream));
    }
    return 0;
}

static int scratch_get(const std::string& name, size_t bytes, void** p) {
    auto it = g_ctx.scratch.find(name);
    if (it != g_ctx.scratch.end() && it->second.second >= bytes) {
        *p = it->second.first;
        return 0;
    }
    if (it != g_ctx.scratch.end()) {
        HIP_TRY(hipStreamSynchronize(g_ctx.stream));
        HIP_TRY(hipFree(it->second.first));
        g_ctx.scratch.erase(it);
    }
    SCARL_TRY(dmalloc(p, bytes));
    g_ctx.scratch[name] = {*p, bytes};
    return 0;
}

// kernel launch bracketed by events when timing is on
template <class F>
static int timed(const char* name, F&& launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (g_ctx.timing) {
        HIP_TRY(hipEventCreate(&a));
        HIP_TRY(hipEventCreate(&b));
        HIP_TRY(hipEventRecord(a, g_ctx.stream));
    }
    hipError_t e = launch();
    if (e != hipSuccess) return fail(SCARL_E_HIP, "launch of %s failed: %s", name, hipGetErrorString(e));
    if (g_ctx.timing) {
        HIP_TRY(hipEventRecord(b, g_ctx.stream));
        g_ctx.pending.push_back({name, a, b});
    }
    return 0;
}

static int timing_collect() {
    if (g_ctx.pending.empty()) return 0;
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    for (auto& t : g_ctx.pending) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
        auto& tot = g_ctx.totals[t.name];
        tot.first += 1;
        tot.second += ms;
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    g_ctx.pending.clear();
    return 0;
}

// ------------------------------------------------------------------------------------------
// planner and twiddle pools
// ------------------------------------------------------------------------------------------
static const long double kPiL = 3.14159265358979323846264338327950288L;
// reference PI literal (SRC_F/constants_ScaRL.f90:5)
static const double kPiRefHost = 3.1415926535898;

// exp(+2 pi i num/den) in long double
static void cis_ld(long long num, long long den, long double& c, long double& s) {
    num %= den;
    if (num < 0) num += den;
    const long double a = 2.0L * kPiL * (long double)num / (long double)den;
    c = cosl(a);
    s = sinl(a);
}

// One way to plan a transform length (the autotuner's search space).  kind 0 reproduces the
// fixed rules: the compile-time planned kernel of SCARL_FIXED_SPECS when one matches (or the MFMA
// kernels when the set asks for them), else the runtime plan with the set's register budget.
enum PlanVariant { PV_DEFAULT = 0, PV_FIXED = 1, PV_GENERIC = 2, PV_MFMA = 3, PV_DIRECT = 4 };
struct PlanChoice {
    int variant = PV_DEFAULT;
    int spec = 0;   // PV_FIXED: SCARL_FIXED_SPECS id
    int ecap = 0;   // PV_GENERIC: complex elements per thread and stage (16 or 32; 0 = the set's)
    int order = 0;  // PV_GENERIC: 0 = largest radix first, 1 = smallest first
    int tdiv = 1;   // PV_GENERIC / PV_MFMA: pencils per tile divided by this; PV_GENERIC 0: rounded
                    // down to a multiple of 16 (whole 128-byte rows in a pencil-fastest tile)
    int skew = 1;   // PV_GENERIC / PV_MFMA, butterfly-fastest layout: one LDS element of skew per 16
    int comp = 0;   // PV_GENERIC: 1 = fewest stages with the composite radices 6, 9, 12 allowed
};

static std::string choice_name(const PlanChoice& c) {
    char b[96];
    switch (c.variant) {
    case PV_FIXED: snprintf(b, sizeof b, "fixed:%d", c.spec); break;
    case PV_GENERIC:
        snprintf(b, sizeof b, "generic:e%d:o%d:t%d:s%d:c%d", c.ecap, c.order, c.tdiv, c.skew, c.comp);
        break;
    case PV_MFMA: snprintf(b, sizeof b, "mfma:t%d:s%d", c.tdiv, c.skew); break;
    case PV_DIRECT: snprintf(b, sizeof b, "direct"); break;
    default: snprintf(b, sizeof b, "default"); break;
    }
    return b;
}

static bool parse_choice(const char* txt, PlanChoice& c) {
    c = PlanChoice();
    if (strcmp(txt, "default") == 0) return true;
    if (strcmp(txt, "direct") == 0) {
        c.variant = PV_DIRECT;
        return true;
    }
    if (sscanf(txt, "fixed:%d", &c.spec) == 1) { c.variant = PV_FIXED; return true; }
    if (sscanf(txt, "generic:e%d:o%d:t%d:s%d:c%d", &c.ecap, &c.order, &c.tdiv, &c.skew, &c.comp) == 5) {
        c.variant = PV_GENERIC;
        return (c.ecap == 16 || c.ecap == 32) && c.tdiv >= 0 && c.tdiv <= 8;
    }
    if (sscanf(txt, "mfma:t%d:s%d", &c.tdiv, &c.skew) == 2) {
        c.variant = PV_MFMA;
        return c.tdiv >= 1 && c.tdiv <= 8;
    }
    return false;
}

// Backward DFT (exp(+2 pi i jk/n)) of a 2^a 3^b 5^c length in long double, recursive mixed radix.
static void dft_ld(std::vector<long double>& re, std::vector<long double>& im) {
    const int n = (int)re.size();
    if (n == 1) return;
    const int p = n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : 5;
    const int m = n / p;
    std::vector<std::vector<long double>> sr(p, std::vector<long double>(m)), si(p, std::vector<long double>(m));
    for (int r = 0; r < p; ++r)
        for (int j = 0; j < m; ++j) {
            sr[r][j] = re[(size_t)p * j + r];
            si[r][j] = im[(size_t)p * j + r];
        }
    for (int r = 0; r < p; ++r) dft_ld(sr[r], si[r]);
    for (int k = 0; k < n; ++k) {
        long double ar = 0.0L, ai = 0.0L;
        for (int r = 0; r < p; ++r) {
            long double c, sn;
            cis_ld((long long)r * k, n, c, sn);
            const long double xr = sr[r][k % m], xi = si[r][k % m];
            ar += xr * c - xi * sn;
            ai += xr * sn + xi * c;
        }
        re[k] = ar;
        im[k] = ai;
    }
}

// Inner length of a Bluestein plan: the 2^a 3^b 5^c length M >= 2N - 1 (up to twice that) with the
// least M x (Stockham stages at register budget ec).
static int bluestein_length(int N, int ec);

struct PlanSet {
    std::vector<PlanDev> plans;
    std::vector<long double> pool;  // interleaved re/im, long double; rounded once at upload
    std::map<std::string, int> index;
    std::vector<std::string> keys;     // key of every plan (index order)
    std::vector<PlanChoice> choices;   // how every plan was made
    std::map<std::string, PlanChoice> overrides;  // key -> choice (the autotuner's picks)
    PlanDev* d_plans = nullptr;
    void* d_pool = nullptr;
    bool fp64 = false;
    int mfma = 0;        // MFMA radix-16 stages where the length allows: 1 every FP32 pass, 2 the x passes
    int ecap = kEcap64;  // register budget of the default runtime plans (matches a kernel instantiation)
    size_t max_lds = 0;

    ~PlanSet() {
        if (d_plans) hipFree(d_plans);
        if (d_pool) hipFree(d_pool);
    }

    int pool_push(long double c, long double s) {
        const int off = (int)(pool.size() / 2);
        pool.push_back(c);
        pool.push_back(s);
        return off;
    }

    // radix split: powers of two in chunks no larger than the register capacity, then the odd
    // primes 3, 5, 7, 11, 13; anything else -> direct DFT.
    static bool factor(int N, int ecap, std::vector<int>& rad) {
        rad.clear();
        int a = 0, n = N;
        while (n % 2 == 0) { n /= 2; ++a; }
        std::vector<int> odd;
        for (int p : {3, 5, 7, 11, 13})
            while (n % p == 0) { n /= p; odd.push_back(p); }
        if (n != 1) return false;
        const int maxr = ecap >= 32 ? 32 : 16;
        const int lmax = maxr == 32 ? 5 : 4;
        // split a into chunks of at most lmax bits, as evenly as possible, largest first
        if (a > 0) {
            const int k = (a + lmax - 1) / lmax;
            std::vector<int> bits(k, a / k);
            for (int i = 0; i < a % k; ++i) bits[i] += 1;
            for (int b : bits) rad.push_back(1 << b);
        }
        for (int p : odd) rad.push_back(p);
        return true;
    }

    // Fewest Stockham stages over the radices 2 ... 16 (32 with ec >= 32) including the composite
    // codelets 6, 9 and 12; among those, the least sum of per-point butterfly costs; largest first.
    static bool factor_composite(int N, int ec, std::vector<int>& rad) {
        rad.clear();
        std::vector<int> R = {2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 16};
        if (ec >= 32) R.push_back(32);
        auto cost = [](int r) {  // real operations per point of one radix-r butterfly, roughly
            switch (r) {
            case 2: return 2.0; case 3: return 4.0; case 4: return 4.0; case 5: return 6.8; case 6: return 5.3;
            case 7: return 9.0; case 8: return 5.3; case 9: return 7.0; case 11: return 13.0; case 12: return 6.0;
            case 13: return 14.0; case 16: return 6.2; default: return 7.0;
            }
        };
        std::vector<int> st(N + 1, 1 << 20), last(N + 1, 0);
        std::vector<double> cs(N + 1, 1e30);
        st[1] = 0;
        cs[1] = 0.0;
        for (int m = 2; m <= N; ++m) {
            if (N % m) continue;
            for (int r : R)
                if (m % r == 0 && st[m / r] < (1 << 20)) {
                    const int s2 = st[m / r] + 1;
                    const double c2 = cs[m / r] + cost(r);
                    if (s2 < st[m] || (s2 == st[m] && c2 < cs[m] - 1e-9)) {
                        st[m] = s2;
                        cs[m] = c2;
                        last[m] = r;
                    }
                }
        }
        if (N < 2 || st[N] >= (1 << 20)) return false;
        for (int m = N; m > 1; m /= last[m]) rad.push_back(last[m]);
        std::sort(rad.begin(), rad.end(), [](int a, int b) { return a > b; });
        return true;
    }

    // the compile-time planned kernel for (N, precision, layout), or 0
    int fixed_spec(int N, int layout, int* r0, int* r1, int* tp) const {
        const int prec = fp64 ? 8 : 4;
        int spec = 0;
#define SCARL_MATCH(ID, P, L, NN, A0, A1, TP)                                   \
        if (spec == 0 && P == prec && L == layout && NN == N) {                 \
            spec = ID;                                                          \
            *r0 = A0;                                                           \
            *r1 = A1;                                                           \
            *tp = TP;                                                           \
        }
        SCARL_FIXED_SPECS(SCARL_MATCH)
#undef SCARL_MATCH
        return spec;
    }

    // every compile-time planned kernel for (N, precision, layout)
    std::vector<int> fixed_specs(int N, int layout) const {
        const int prec = fp64 ? 8 : 4;
        std::vector<int> ids;
#define SCARL_ALL(ID, P, L, NN, A0, A1, TP) \
        if (P == prec && L == layout && NN == N) ids.push_back(ID);
        SCARL_FIXED_SPECS(SCARL_ALL)
#undef SCARL_ALL
        return ids;
    }
    static void fixed_spec_radices(int id, int* r0, int* r1, int* tp) {
#define SCARL_RAD(ID, P, L, NN, A0, A1, TP) \
        if (ID == id) {                     \
            *r0 = A0;                       \
            *r1 = A1;                       \
            *tp = TP;                       \
        }
        SCARL_FIXED_SPECS(SCARL_RAD)
#undef SCARL_RAD
    }

    // Stockham stages of radices rad with their twiddle tables; tpp from the register budget ec
    // (the first stage excluded when it runs on the matrix cores)
    void stages(PlanDev& pl, const std::vector<int>& rad, int ec, bool mfma_first) {
        pl.nst = (int)rad.size();
        int Ns = 1, tpp = 1;
        for (int s = 0; s < pl.nst; ++s) {
            const int R = rad[s];
            pl.R[s] = R;
            pl.Ns[s] = Ns;
            pl.tw_off[s] = 0;
            if (Ns > 1) {
                pl.tw_off[s] = (int)(pool.size() / 2);
                for (int k = 0; k < Ns; ++k)
                    for (int r = 1; r < R; ++r) {
                        long double c, sn;
                        cis_ld((long long)k * r, (long long)Ns * R, c, sn);
                        pool_push(c, sn);
                    }
            }
            if (!(mfma_first && s == 0)) {
                const int nb = pl.N / R, bmax = std::max(1, ec / R);
                tpp = std::max(tpp, (nb + bmax - 1) / bmax);
            }
            Ns *= R;
        }
        pl.tpp = tpp;
    }

    // Builds the plan of one choice.  kind: 0 none, 1 C2R post-twiddle (exp(2 pi i k/(2N)), k<N),
    // 2 DCT pre-twiddle (exp(i pi j/(2N))).  Returns 0, or an error code without touching the set
    // when the choice does not apply to this length (unknown kernel, LDS, tile size).
    int make(int N, int layout, int kind, const PlanChoice& ch, PlanDev& pl) {
        const size_t pool0 = pool.size();
        auto undo = [&](int code) {
            pool.resize(pool0);
            return code;
        };
        const int csz = fp64 ? 16 : 8;
        memset(&pl, 0, sizeof pl);
        pl.N = N;
        pl.ecap = ecap;
        pl.direct_off = -1;
        pl.pre_off = -1;
        pl.skew = 1;
        std::vector<int> rad;
        int sr0 = 0, sr1 = 0, stpp = 0;
        const int fspec = fixed_spec(N, layout, &sr0, &sr1, &stpp);
        const bool mfma_ok = !fp64 && kind != 2 && N >= 32 && N % 16 == 0 && N <= kMfmaTileElems;
        int variant = ch.variant;
        if (variant == PV_DEFAULT) {
            if (mfma && (mfma == 1 || layout == LAYOUT_BF) && mfma_ok && !(fspec && layout == LAYOUT_BF && kind == 1 && N == 256 && sr0 == 16 &&
                                     sr1 == 16 && stpp == 16))
                variant = PV_MFMA;
            else variant = fspec ? PV_FIXED : PV_GENERIC;
        }
        int tdiv = ch.variant == PV_DEFAULT ? 1 : ch.tdiv;
        if (variant == PV_DIRECT && factor(N, ecap, rad)) return undo(SCARL_E_UNSUPPORTED);
        if (variant == PV_FIXED) {
            const int spec = ch.variant == PV_FIXED ? ch.spec : fspec;
            const std::vector<int> all = fixed_specs(N, layout);
            if (!spec || std::find(all.begin(), all.end(), spec) == all.end()) return SCARL_E_UNSUPPORTED;
            fixed_spec_radices(spec, &sr0, &sr1, &stpp);
            pl.spec = spec;
            if (ch.variant == PV_DEFAULT && mfma && !fp64 && layout == LAYOUT_BF && kind == 1 && N == 256 && sr0 == 16 &&
                sr1 == 16 && stpp == 16)
                pl.spec = kSpecMfma256;  // same radices, tile and twiddle tables: only the kernel differs
            pl.nst = sr1 > 1 ? 2 : 1;
            pl.R[0] = sr0;
            pl.Ns[0] = 1;
            if (sr1 > 1) {
                pl.R[1] = sr1;
                pl.Ns[1] = sr0;
                pl.tw_off[1] = (int)(pool.size() / 2);
                for (int k = 0; k < sr0; ++k)
                    for (int r = 1; r < sr1; ++r) {
                        long double c, sn;
                        cis_ld((long long)k * r, (long long)sr0 * sr1, c, sn);
                        pool_push(c, sn);
                    }
            }
            pl.tpp = stpp;
            tdiv = 1;
        } else if (variant == PV_MFMA) {
            if (!mfma_ok || !factor(N / 16, 16, rad)) return undo(SCARL_E_UNSUPPORTED);
            rad.insert(rad.begin(), 16);
            pl.spec = kSpecMfmaGeneric;
            pl.ecap = 16;
            if (ch.variant == PV_MFMA) pl.skew = ch.skew;
            stages(pl, rad, 16, true);
        } else if (N >= 2 && factor(N, ch.variant == PV_GENERIC ? ch.ecap : ecap, rad)) {
            const int ec = ch.variant == PV_GENERIC ? ch.ecap : ecap;
            // the composite radices when asked for, and by default when they save a stage
            std::vector<int> crad;
            if (factor_composite(N, ec, crad) &&
                ((ch.variant == PV_GENERIC && ch.comp) || (ch.variant == PV_DEFAULT && crad.size() < rad.size())))
                rad = crad;
            else if (ch.variant == PV_GENERIC && ch.comp)
                return undo(SCARL_E_UNSUPPORTED);
            if (fp64 && ec != kEcap64) return undo(SCARL_E_UNSUPPORTED);
            if (ch.variant == PV_GENERIC) {
                if (ch.order) std::reverse(rad.begin(), rad.end());
                pl.skew = ch.skew;
            }
            pl.spec = 0;
            for (int r : rad)
                if (r == 6 || r == 9 || r == 12) pl.spec = kSpecGenericComp;
            pl.ecap = ec;
            stages(pl, rad, ec, false);
        } else if (N >= 32 && ch.variant != PV_DIRECT && ch.variant != PV_FIXED && ch.variant != PV_MFMA) {
            // a prime factor above 13: Bluestein over the inner length M
            const int ec = ch.variant == PV_GENERIC ? ch.ecap : ecap;
            if (fp64 && ec != kEcap64) return undo(SCARL_E_UNSUPPORTED);
            const int M = bluestein_length(N, ec);
            if (M <= 0 || !factor(M, ec, rad)) return -fail(SCARL_E_UNSUPPORTED, "no Bluestein length for %d", N);
            if (ch.variant == PV_GENERIC) {
                if (ch.order) std::reverse(rad.begin(), rad.end());
                pl.skew = ch.skew;
            }
            pl.spec = kSpecNonSmooth;
            pl.ecap = ec;
            pl.N = M;  // the stages and threads per pencil are those of M
            stages(pl, rad, ec, false);
            pl.N = N;
            pl.blu_M = M;
            pl.chirp_off = (int)(pool.size() / 2);
            for (long long j = 0; j < N; ++j) {
                long double c, sn;
                cis_ld((j * j) % (2LL * N), 2LL * N, c, sn);  // exp(+i pi j^2/N)
                pool_push(c, sn);
            }
            // E = DFT_M^-1(d)/M, d_m = exp(-i pi m^2/N) for |m| < N (wrapped mod M), 0 elsewhere
            std::vector<long double> dr(M, 0.0L), di(M, 0.0L);
            for (long long m = 0; m < N; ++m) {
                long double c, sn;
                cis_ld((m * m) % (2LL * N), 2LL * N, c, sn);
                dr[m] = c;
                di[m] = -sn;
                if (m) {
                    dr[M - m] = c;
                    di[M - m] = -sn;
                }
            }
            dft_ld(dr, di);
            pl.kern_off = (int)(pool.size() / 2);
            for (int m = 0; m < M; ++m) pool_push(dr[m] / M, di[m] / M);
        } else {
            if (ch.variant != PV_DEFAULT && ch.variant != PV_GENERIC && ch.variant != PV_DIRECT)
                return undo(SCARL_E_UNSUPPORTED);
            pl.nst = 0;  // direct DFT (N == 1, or a short length with a prime factor above 13)
            pl.spec = kSpecNonSmooth;
            if (N > 4096) return -fail(SCARL_E_UNSUPPORTED, "length %d has a prime factor above 13 and exceeds 4096", N);
            pl.direct_off = (int)(pool.size() / 2);
            for (int t = 0; t < N; ++t) {
                long double c, sn;
                cis_ld(t, N, c, sn);
                pool_push(c, sn);
            }
            pl.tpp = std::min(N, 64);
        }
        if (pl.tpp > kBlock) {
            if (ch.variant != PV_DEFAULT) return undo(SCARL_E_UNSUPPORTED);
            return -fail(SCARL_E_UNSUPPORTED, "transform length %d too large for one tile", N);
        }
        pl.T = kBlock / pl.tpp;
        const int NL = pl.blu_M > 0 ? pl.blu_M : N;  // elements of one pencil in the LDS tile
        pl.ldp = NL + (pl.skew ? NL / 16 : 0) + 1;
        // keep the LDS tile within 64 KiB so two tiles fit on a CU.  Every plan gets an LDS
        // tile, because passes that stage their input or output use it even for one stage.
        // (compile-time planned kernels fix T = 256 / TPP; the spec table respects the bound)
        const size_t per = (size_t)(layout == LAYOUT_PF ? NL : pl.ldp) * csz;
        const int Tmax = (int)std::max<size_t>(1, (64 * 1024) / per);
        const bool runtime_tile = pl.spec == 0 || pl.spec == kSpecMfmaGeneric || pl.spec == kSpecNonSmooth ||
                                  pl.spec == kSpecGenericComp;
        if (runtime_tile) pl.T = std::min(pl.T, Tmax);
        if (pl.spec == kSpecMfmaGeneric) pl.T = std::max(1, std::min(pl.T, kMfmaTileElems / N));
        // tdiv = 0: a tile of 16 or more pencils cut to a multiple of 16, so that every access of a
        // pencil-fastest tile covers whole 128-byte rows (the autotuner picks it for config 4's
        // 144-point y pass: 21 -> 16 pencils, 0.449 -> 0.355 ms); any other tdiv divides T
        if (runtime_tile) pl.T = tdiv == 0 ? (pl.T >= 16 ? pl.T / 16 * 16 : pl.T) : std::max(1, pl.T / tdiv);
        pl.lds_elems = (int)(pl.T * (layout == LAYOUT_PF ? NL : pl.ldp));
        if (kind == 1) {
            pl.pre_off = (int)(pool.size() / 2);
            for (int k = 0; k < N; ++k) {
                long double c, sn;
                cis_ld(k, 2LL * N, c, sn);
                pool_push(c, sn);
            }
        } else if (kind == 2) {
            pl.pre_off = (int)(pool.size() / 2);
            for (int j = 0; j < N; ++j) {
                long double c, sn;
                cis_ld(j, 4LL * N, c, sn);  // exp(i pi j/(2N)) = exp(2 pi i j/(4N))
                pool_push(c, sn);
            }
        }
        return 0;
    }

    static std::string key_of(int N, int layout, int kind, int role) {
        char key[64];
        snprintf(key, sizeof key, "%d/%d/%d/%d", N, layout, kind, role);
        return key;
    }

    // Plan index for (N, layout, kind, role); role separates passes that share a length but may
    // prefer different plans (0: strided, not generating; 1: generating; 2: x pass).  The choice
    // is the autotuner's pick for the key when there is one, else the default rules.
    int get(int N, int layout, int kind, int role = 0) {
        const std::string key = key_of(N, layout, kind, role);
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        PlanChoice ch;
        auto ov = overrides.find(key);
        if (ov != overrides.end()) ch = ov->second;
        PlanDev pl;
        int rc = make(N, layout, kind, ch, pl);
        if (rc > 0 && ch.variant != PV_DEFAULT) {
            ch = PlanChoice();  // a cached pick that no longer applies: default rules
            rc = make(N, layout, kind, ch, pl);
        }
        if (rc) return rc < 0 ? rc : -rc;
        max_lds = std::max(max_lds, (size_t)pl.lds_elems * (fp64 ? 16 : 8));
        plans.push_back(pl);
        keys.push_back(key);
        choices.push_back(ch);
        const int id = (int)plans.size() - 1;
        index[key] = id;
        return id;
    }

    int upload() {
        if (d_plans) { hipFree(d_plans); d_plans = nullptr; }
        if (d_pool) { hipFree(d_pool); d_pool = nullptr; }
        SCARL_TRY(dmalloc((void**)&d_plans, plans.size() * sizeof(PlanDev)));
        HIP_TRY(hipMemcpy(d_plans, plans.data(), plans.size() * sizeof(PlanDev), hipMemcpyHostToDevice));
        const size_t nc = std::max<size_t>(1, pool.size() / 2);
        if (fp64) {
            std::vector<double> h(2 * nc, 0.0);
            for (size_t i = 0; i < pool.size(); ++i) h[i] = (double)pool[i];
            SCARL_TRY(dmalloc(&d_pool, h.size() * sizeof(double)));
            HIP_TRY(hipMemcpy(d_pool, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            std::vector<float> h(2 * nc, 0.f);
            for (size_t i = 0; i < pool.size(); ++i) h[i] = (float)pool[i];
            SCARL_TRY(dmalloc(&d_pool, h.size() * sizeof(float)));
            HIP_TRY(hipMemcpy(d_pool, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        return 0;
    }
};

static int bluestein_length(int N, int ec) {
    const long long lo = 2LL * N - 1;
    int best = -1;
    long long best_cost = 0;
    std::vector<int> rad;
    for (long long M = lo; M <= 2 * lo && M <= 16384; ++M) {
        long long r = M;
        for (int p : {2, 3, 5})
            while (r % p == 0) r /= p;
        if (r != 1 || !PlanSet::factor((int)M, ec, rad)) continue;
        const long long cost = M * (long long)rad.size();
        if (best < 0 || cost < best_cost) {
            best = (int)M;
            best_cost = cost;
        }
    }
    return best;
}

static int check_lds(size_t bytes) {
    if (bytes > 160 * 1024) return fail(SCARL_E_UNSUPPORTED, "tile needs %zu bytes of LDS", bytes);
    return 0;
}

// ------------------------------------------------------------------------------------------
// marginal parameters (SRC_F/main_ScaRL.f90:1367-1387)
// ------------------------------------------------------------------------------------------
static int marginal_params(int margi, double avg, double std_dev, double* mu, double* sd) {
    double var;
    if (margi == FOM_NONE) {
        *mu = 0.0;
        *sd = 1.0;
        return 0;
    }
    if (margi == FOM_GAUSSIAN) {
        var = std_dev * std_dev;
        *mu = avg;
    } else if (margi == FOM_LOGNORMAL) {
        if (avg <= 0.0) return fail(SCARL_E_INVALID, "ERROR - when using lognormal fieldAvg should be a positive number");
        var = log(1.0 + (std_dev / avg) * (std_dev / avg));
        *mu = log(avg) - var / 2.0;
    } else {
        return fail(SCARL_E_INVALID, "First-order marginal not detected (margiFirst = %d)", margi);
    }
    *sd = sqrt(var);
    return 0;
}

// ------------------------------------------------------------------------------------------
// batch of mode H fields
// ------------------------------------------------------------------------------------------
struct FieldHost {
    scarl_field_desc d;
    FieldH f;
    double spec_sum = 0.0;
    int64_t out_elems = 0;
    void* noise_buf = nullptr;
};

}  // namespace scarl

struct scarl_batch {
    bool fp64 = false;
    bool global = false;
    // groups run pipelined (the next group's z pass beside this group's y and x passes) or one
    // after the other with every pass's launches fanned out; see choose_schedule
    bool pipeline = true;
    int gdim[3] = {0, 0, 0};
    std::vector<scarl::FieldHost> fields;
    scarl::PlanSet ps;
    std::vector<scarl::Tile> tz, ty, tx;     // tiles of the z, y and x passes
    struct TileClass {
        int plan, t0, t1;                    // tiles [t0, t1) of one plan (one kernel launch)
        int color = 0;                       // x pass of a blend: the colour of its subdomains
        bool gen = false;                    // y pass: the fields are 2D (this pass generates)
    };
    // A group of fields runs its z, y and x passes in order; in a pipelined batch the z pass of
    // group g + 1 (compute-bound generation) overlaps the y and x passes of group g (HBM-bound).
    struct Group {
        std::vector<TileClass> cz, cy, cx;   // launches of each pass (x: colour-major for the blend)
        int kx0_off = 0, kx0_cnt = 0;        // its 3D fields in kx0_ids, for k_h_kx0
        int64_t kx0_max = 0;                 // their largest kx = 0 plane (points)
    };
    std::vector<Group> groups;
    std::vector<int> kx0_ids;                // the 3D fields of every group, group by group
    int* d_kx0_ids = nullptr;
    std::vector<hipEvent_t> gev;             // fork event, then one per group (pipelined batches)
    hipEvent_t xev[4] = {nullptr, nullptr, nullptr, nullptr};  // fan-out of a pass: fork, three joins
    std::vector<int2> field_xtiles;          // per-field [begin, end) in tx
    scarl::FieldH* d_fields = nullptr;
    scarl::Tile *d_tz = nullptr, *d_ty = nullptr, *d_tx = nullptr;
    int2* d_field_xtiles = nullptr;
    scarl::MomPartial* d_partials = nullptr;
    scarl::MomPartial* d_field_mom = nullptr;
    double* d_slot_count = nullptr;
    double* d_slot_out = nullptr;
    int nslots = 0;
    void* d_ws = nullptr;       // all half-spectrum workspaces
    void* d_out = nullptr;      // all outputs (or the global field)
    void* d_kx0 = nullptr;      // the kx = 0 planes' inputs of the 3D fields (k_h_kx0)
    int64_t kx0_max = 0;        // largest n2 n3 among them (0: no 3D field)
    int64_t bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<std::string> tune_notes;     // per plan: the autotuner's timings (empty if untuned)

    ~scarl_batch() {
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        for (hipEvent_t e : gev) hipEventDestroy(e);
        for (hipEvent_t e : xev)
            if (e) hipEventDestroy(e);
        for (auto& f : fields)
            if (f.noise_buf) hipFree(f.noise_buf);
        void* ptrs[] = {d_fields, d_tz, d_ty, d_tx, d_field_xtiles, d_partials, d_field_mom, d_slot_count, d_slot_out,
                        d_ws, d_out, d_kx0, d_kx0_ids};
        for (void* p : ptrs)
            if (p) hipFree(p);
    }
};

namespace scarl {

static int validate_desc(const scarl_field_desc& d) {
    for (int a = 0; a < 3; ++a)
        if (d.n[a] < 1) return fail(SCARL_E_INVALID, "extent n[%d] = %d must be >= 1", a, d.n[a]);
    if (d.n[0] < 2) return fail(SCARL_E_INVALID, "x extent must be >= 2");
    if (d.n[2] > 1 && d.n[1] < 2) return fail(SCARL_E_INVALID, "a 3D field needs n[1] >= 2");
    if (d.n[1] < 2) return fail(SCARL_E_UNSUPPORTED, "1D fields are not supported (n[1] = 1)");
    if (d.model != CM_GAUSSIAN && d.model != CM_VON_KARMAN && d.model != CM_EXPONENTIAL)
        return fail(SCARL_E_INVALID, "unknown correlation model %d", d.model);
    for (int a = 0; a < 3; ++a) {
        if (!(d.dx[a] > 0.0)) return fail(SCARL_E_INVALID, "dx[%d] must be positive", a);
        if (!(d.corrL[a] > 0.0)) return fail(SCARL_E_INVALID, "corrL[%d] must be positive", a);
    }
    if (d.model == CM_VON_KARMAN && !(d.nu > 0.0)) return fail(SCARL_E_INVALID, "von Karman needs nu > 0");
    if (!(d.sigma >= 0.0)) return fail(SCARL_E_INVALID, "sigma must be >= 0");
    return 0;
}

// Build the FieldH of one descriptor (spectrum constants; amp0 is finished after the FP64
// spectrum sum).  Returns 0 or an error.
static int build_field(scarl_batch* b, FieldHost& fh) {
    const scarl_field_desc& d = fh.d;
    FieldH& f = fh.f;
    memset(&f, 0, sizeof f);
    for (int a = 0; a < 3; ++a) f.n[a] = d.n[a];
    f.h1 = d.n[0] / 2 + 1;
    // even n1: rows of n1/2 plus the Nyquist plane; odd n1: rows of h1 (no Nyquist column)
    const bool odd = (d.n[0] & 1) != 0;
    f.wp = odd ? f.h1 : d.n[0] / 2;
    f.nyq = odd ? 0 : 1;
    const int dim = d.n[2] > 1 ? 3 : 2;
    f.first_axis = (dim == 3) ? 2 : 1;
    f.src_mode = SRC_DEVICE_NOISE;
    f.model = d.model;
    for (int a = 0; a < 3; ++a) f.alpha[a] = (a < dim) ? 2.0 * M_PI * d.corrL[a] / (d.n[a] * d.dx[a]) : 0.0;
    const double nu = (d.model == CM_EXPONENTIAL) ? 0.5 : d.nu;
    f.expo = -(nu + 0.5 * dim) / 2.0;
    f.seed = d.seed;
    f.stream = d.stream;
    f.margi = d.margi;
    SCARL_TRY(marginal_params(d.margi, d.avg, d.std_dev, &f.mu, &f.sd));
    f.mom_slot = -1;
    // plans
    // role 1: the generating pass (z, or y for a 2D field); role 0: the in-place y pass
    const int kz = (dim == 3) ? b->ps.get(d.n[2], LAYOUT_PF, 0, 1) : 0;
    if (kz < 0) return -kz;
    const int ky = b->ps.get(d.n[1], LAYOUT_PF, 0, dim == 3 ? 0 : 1);
    if (ky < 0) return -ky;
    // x pass: half-length C2R (even n1, with its pre-twiddle table) or a full-length C2C (odd n1)
    const int kx = odd ? b->ps.get(d.n[0], LAYOUT_BF, 0, 2) : b->ps.get(d.n[0] / 2, LAYOUT_BF, 1, 2);
    if (kx < 0) return -kx;
    f.plan_z = kz;
    f.plan_y = ky;
    f.plan_x = kx;
    return 0;
}

static double spectrum_prefactor(const scarl_field_desc& d, int dim) {
    // c such that S = c * g(q): Gaussian prod L/(2 sqrt pi); Matern Gamma(e)/(Gamma(nu) pi^(dim/2)) prod L
    double pL = 1.0;
    for (int a = 0; a < dim; ++a) pL *= d.corrL[a];
    if (d.model == CM_GAUSSIAN) return pL / pow(2.0 * sqrt(M_PI), dim);
    const double nu = (d.model == CM_EXPONENTIAL) ? 0.5 : d.nu;
    const double e = nu + 0.5 * dim;
    return tgamma(e) / (tgamma(nu) * pow(M_PI, 0.5 * dim)) * pL;
}

// FP64 device reduction of the spectrum over every field; fills amp0.
static int finish_amplitudes(scarl_batch* b) {
    int maxblocks = 1;
    for (auto& fh : b->fields) {
        const int64_t cols = (int64_t)fh.f.h1 * fh.f.n[1];
        maxblocks = std::max<int>(maxblocks, (int)((cols + kBlock - 1) / kBlock));
    }
    MomPartial* d_part = nullptr;
    SCARL_TRY(dmalloc((void**)&d_part, sizeof(MomPartial) * maxblocks));
    std::vector<MomPartial> h(maxblocks);
    std::map<std::string, double> cache;
    int rc = 0;
    for (auto& fh : b->fields) {
        const FieldH& f = fh.f;
        char key[256];
        snprintf(key, sizeof key, "%d %d %d %d %.17g %.17g %.17g %.17g", f.n[0], f.n[1], f.n[2], f.model, f.alpha[0],
                 f.alpha[1], f.alpha[2], f.expo);
        auto it = cache.find(key);
        double G;
        if (it != cache.end()) {
            G = it->second;
        } else {
            const int64_t cols = (int64_t)f.h1 * f.n[1];
            const int nblocks = (int)((cols + kBlock - 1) / kBlock);
            hipError_t e = launch_spectrum_sum(f, d_part, nblocks, g_ctx.stream);
            if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_part, sizeof(MomPartial) * nblocks,
                                                    hipMemcpyDeviceToHost, g_ctx.stream);
            if (e == hipSuccess) e = hipStreamSynchronize(g_ctx.stream);
            if (e != hipSuccess) {
                rc = fail(SCARL_E_HIP, "spectrum sum: %s", hipGetErrorString(e));
                break;
            }
            // Neumaier combine of the block partials on the host
            double s = 0.0, c = 0.0;
            for (int i = 0; i < nblocks; ++i) {
                const double v = h[i].s;
                const double t = s + v;
                c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
                s = t;
                c += h[i].c;
            }
            G = s + c;
            cache[key] = G;
        }
        const int dim = f.n[2] > 1 ? 3 : 2;
        const double cpre = spectrum_prefactor(fh.d, dim);
        fh.spec_sum = cpre * G;  // sum of S over the full discrete spectrum
        if (fh.d.normalize) {
            fh.f.amp0 = fh.d.sigma / sqrt(G);
        } else {
            double dk = 1.0;
            for (int a = 0; a < dim; ++a) dk *= 2.0 * M_PI / (fh.d.n[a] * fh.d.dx[a]);
            fh.f.amp0 = fh.d.sigma * sqrt(cpre * dk);
        }
    }
    hipFree(d_part);
    return rc;
}

// Tiles of one strided pass over a field: the nother pencils of the Nyquist plane (kind 1), the
// regular tiles that hold a kx = 0 pencil (kind 2) and the other regular tiles of the workspace
// rows (kind 0).  The first two carry the self-conjugate noise, whose unpaired draws make them
// costlier in the generating pass.
static void strided_tiles(int field, const FieldH& f, int nother, int T, int kind, std::vector<Tile>& out) {
    if (kind == 1) {
        if (f.nyq)
            for (int p = 0; p < nother; p += T) out.push_back({field, p, 1, 0});
        return;
    }
    for (int64_t p = 0; p < (int64_t)f.wp * nother; p += T) {
        const bool has_kx0 = (p % f.wp == 0) || (p / f.wp != (p + std::min<int64_t>(T, (int64_t)f.wp * nother - p) - 1) / f.wp);
        if (has_kx0 == (kind == 2)) out.push_back({field, (int)p, 0, 0});
    }
}

// Minimum z-pass tiles of a pipeline group: enough workgroups to fill the GPU several times over.
constexpr int64_t kGroupTiles = 4096;

// Tile tables.  Fields are cut into consecutive groups of at least kGroupTiles z-pass tiles, so
// that a group's workspace (about 170 MB for 4096 tiles) is still in the 256 MB Infinity Cache
// when its y and x passes read it.  A blend's subdomains are ordered by colour and a group holds
// one colour only: the x passes then run in colour order, as the owner order of the overlap
// points requires, and the groups pipeline like an ensemble's.  Within a group every pass has one
// launch per plan (for a blend, per plan of the group's colour), each with its own kernel and LDS
// size; the costlier
// self-conjugate tiles go first so that the round-robin dispatch of workgroups spreads them evenly
// over the XCDs (in a row-major order every row's kx = 0 tile would fall on the same XCD whenever
// the tiles per row are a multiple of 8).
static void make_tiles(scarl_batch* b) {
    b->tz.clear();
    b->ty.clear();
    b->tx.clear();
    b->groups.clear();
    b->kx0_ids.clear();
    b->field_xtiles.assign(b->fields.size(), make_int2(0, 0));
    std::vector<std::vector<int>> members(1);
    int64_t acc = 0;
    // fields of the same plans next to each other, so that a group holds few distinct plans and each
    // of its launches covers many tiles (the order of the outputs does not change)
    std::vector<int> order(b->fields.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int u, int v) {
        const FieldH &fu = b->fields[u].f, &fv = b->fields[v].f;
        if (b->global && (fu.color & 7) != (fv.color & 7)) return (fu.color & 7) < (fv.color & 7);
        if (fu.plan_z != fv.plan_z) return fu.plan_z < fv.plan_z;
        if (fu.plan_y != fv.plan_y) return fu.plan_y < fv.plan_y;
        return fu.plan_x < fv.plan_x;
    });
    for (int i : order) {
        const FieldH& f = b->fields[i].f;
        const bool new_colour = b->global && !members.back().empty() &&
                                (b->fields[members.back().back()].f.color & 7) != (f.color & 7);
        if (acc >= kGroupTiles || new_colour) {
            members.emplace_back();
            acc = 0;
        }
        members.back().push_back(i);
        const int pz = f.n[2] > 1 ? f.plan_z : f.plan_y;
        acc += ((int64_t)(f.wp + 1) * (f.n[2] > 1 ? f.n[1] : f.n[2])) / b->ps.plans[pz].T + 1;
    }
    for (const auto& mem : members) {
        scarl_batch::Group g;
        g.kx0_off = (int)b->kx0_ids.size();
        for (int i : mem) {
            const FieldH& f = b->fields[i].f;
            if (f.n[2] > 1) {
                b->kx0_ids.push_back(i);
                g.kx0_max = std::max<int64_t>(g.kx0_max, (int64_t)f.n[1] * f.n[2]);
            }
        }
        g.kx0_cnt = (int)b->kx0_ids.size() - g.kx0_off;
        std::map<int, std::vector<int>> zp;
        std::map<std::pair<int, int>, std::vector<int>> yp;  // (plan, 2D): 2D fields generate in y
        for (int i : mem) {
            const FieldH& f = b->fields[i].f;
            if (f.n[2] > 1) zp[f.plan_z].push_back(i);
            yp[{f.plan_y, f.n[2] == 1 ? 1 : 0}].push_back(i);
        }
        for (auto& kv : zp) {
            const int t0 = (int)b->tz.size(), T = b->ps.plans[kv.first].T;
            for (int kind : {1, 2, 0})
                for (int i : kv.second) strided_tiles(i, b->fields[i].f, b->fields[i].f.n[1], T, kind, b->tz);
            g.cz.push_back({kv.first, t0, (int)b->tz.size()});
        }
        for (auto& kv : yp) {
            const int plan = kv.first.first, t0 = (int)b->ty.size(), T = b->ps.plans[plan].T;
            for (int kind : {1, 2, 0})
                for (int i : kv.second) strided_tiles(i, b->fields[i].f, b->fields[i].f.n[2], T, kind, b->ty);
            scarl_batch::TileClass cl{plan, t0, (int)b->ty.size()};
            cl.gen = kv.first.second != 0;
            g.cy.push_back(cl);
        }
        for (int c = 0; c < 8; ++c) {
            std::map<int, std::vector<int>> xp;
            for (int i : mem)
                if ((b->fields[i].f.color & 7) == c) xp[b->fields[i].f.plan_x].push_back(i);
            for (auto& kv : xp) {
                const int t0 = (int)b->tx.size(), T = b->ps.plans[kv.first].T;
                for (int i : kv.second) {
                    const FieldH& f = b->fields[i].f;
                    // odd n1: a pencil of the x pass is a pair of rows (one complex transform)
                    const int64_t rows = (f.n[0] & 1) ? ((int64_t)f.n[1] * f.n[2] + 1) / 2 : (int64_t)f.n[1] * f.n[2];
                    b->field_xtiles[i].x = (int)b->tx.size();
                    // last rows first: the y pass wrote the high kz planes last, so part of them is
                    // still in the 256 MB Infinity Cache when the x pass starts (1.3 % on the x pass)
                    for (int64_t t = (rows + T - 1) / T - 1; t >= 0; --t) b->tx.push_back({i, (int)(t * T)});
                    b->field_xtiles[i].y = (int)b->tx.size();
                }
                g.cx.push_back({kv.first, t0, (int)b->tx.size(), c});
            }
        }
        b->groups.push_back(std::move(g));
    }
}

template <class V>
static int upload_vec(const V& v, void** d) {
    using E = typename V::value_type;
    SCARL_TRY(dmalloc(d, std::max<size_t>(1, v.size()) * sizeof(E)));
    if (!v.empty()) HIP_TRY(hipMemcpy(*d, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice));
    return 0;
}

// Device copies of the plan set, the tile tables, the moment partials and the pipeline events;
// called again whenever the plans change (autotuning), so it releases what it replaces.
static int upload_tables(scarl_batch* b) {
    const size_t rsz = b->fp64 ? 8 : 4;
    SCARL_TRY(b->ps.upload());
    make_tiles(b);
    for (hipEvent_t e : b->gev) hipEventDestroy(e);
    b->gev.clear();
    if (b->groups.size() > 1)
        for (size_t g = 0; g <= b->groups.size(); ++g) {
            hipEvent_t e;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            b->gev.push_back(e);
        }
    if (!b->xev[0])
        for (hipEvent_t& e : b->xev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (const PlanDev& pl : b->ps.plans) SCARL_TRY(check_lds((size_t)pl.lds_elems * 2 * rsz));
    void** tabs[] = {(void**)&b->d_tz, (void**)&b->d_ty, (void**)&b->d_tx, (void**)&b->d_field_xtiles,
                     (void**)&b->d_partials, (void**)&b->d_kx0_ids};
    for (void** t : tabs)
        if (*t) {
            hipFree(*t);
            *t = nullptr;
        }
    SCARL_TRY(upload_vec(b->tz, (void**)&b->d_tz));
    SCARL_TRY(upload_vec(b->ty, (void**)&b->d_ty));
    SCARL_TRY(upload_vec(b->tx, (void**)&b->d_tx));
    SCARL_TRY(upload_vec(b->field_xtiles, (void**)&b->d_field_xtiles));
    SCARL_TRY(upload_vec(b->kx0_ids, (void**)&b->d_kx0_ids));
    SCARL_TRY(dmalloc((void**)&b->d_partials, std::max<size_t>(1, b->tx.size()) * sizeof(MomPartial)));
    return 0;
}

// Infinity Cache (MALL) of the MI355X: 256 MB in front of HBM.
constexpr int64_t kMallBytes = int64_t(256) << 20;

// Default schedule of a batch of several groups.  When groups fit the Infinity Cache, a group's
// y and x passes read what its z pass left there, and a pipelined z pass of the next group
// (writing its own workspace beside them) evicts it: a 256-field ensemble of 64-192^3 (groups of
// 62-282 MB, 150 MB on average) runs 4.57 ms per step pipelined and 4.15 ms one group after
// another.  Groups larger than the cache (a 512^3 field each, 540 MB) have nothing to lose and
// pipeline (2.23 against 2.26 ms).  A blend pipelines whatever its group size: its colour order
// already serialises the x passes, and overlapping the next colour's generation with them pays
// more than the cache (8 subdomains of 241^3: 3.81 against 4.03 ms; 1024 x 1024 x 256 from 128
// subdomains: 2.51 against 2.53 ms).  The autotuner times both schedules and keeps the faster.
static void choose_schedule(scarl_batch* b) {
    const size_t rsz = b->fp64 ? 8 : 4;
    int64_t total = 0;
    for (const auto& fh : b->fields) {
        const FieldH& f = fh.f;
        total += (f.wz * f.n[2] + (int64_t)f.nyq * f.n[1] * f.n[2]) * 2 * (int64_t)rsz;
    }
    const int64_t mean = total / std::max<int64_t>(1, (int64_t)b->groups.size());
    b->pipeline = g_ctx.schedule == 0 ? (b->global || mean > kMallBytes) : g_ctx.schedule == 1;
}

static int finalize_batch(scarl_batch* b, const std::vector<double>& slot_count) {
    const size_t rsz = b->fp64 ? 8 : 4;
    int64_t ws_elems = 0, out_elems = 0;
    for (auto& fh : b->fields) {
        // kz planes of the rows are padded by 4 KiB when a plane is a multiple of 64 KiB: power-of-two
        // plane strides put every row a z-pass tile touches on the same HBM channels
        FieldH& f = fh.f;
        const int64_t plane = (int64_t)f.wp * f.n[1];
        f.wz = plane + ((f.n[2] > 1 && (plane * 2 * rsz) % 65536 == 0) ? 4096 / (2 * (int64_t)rsz) : 0);
        // the strided passes step along a pencil by a 32-bit byte pitch
        if (f.wz * 2 * (int64_t)rsz >= (int64_t(1) << 32))
            return fail(SCARL_E_UNSUPPORTED, "a kz plane of %lld elements exceeds 4 GiB", (long long)f.wz);
        ws_elems += f.wz * f.n[2] + (int64_t)f.nyq * f.n[1] * f.n[2];
    }
    if (!b->global)
        for (auto& fh : b->fields) out_elems += (int64_t)fh.f.n[0] * fh.f.n[1] * fh.f.n[2];
    else
        out_elems = (int64_t)b->gdim[0] * b->gdim[1] * b->gdim[2];
    SCARL_TRY(dmalloc(&b->d_ws, (size_t)ws_elems * 2 * rsz));
    SCARL_TRY(dmalloc(&b->d_out, (size_t)out_elems * rsz));
    int64_t kx0_elems = 0;
    b->kx0_max = 0;
    for (auto& fh : b->fields)
        if (fh.f.n[2] > 1) {
            kx0_elems += (int64_t)fh.f.n[1] * fh.f.n[2];
            b->kx0_max = std::max<int64_t>(b->kx0_max, (int64_t)fh.f.n[1] * fh.f.n[2]);
        }
    if (kx0_elems > 0) {
        SCARL_TRY(dmalloc(&b->d_kx0, (size_t)kx0_elems * 2 * rsz));
        int64_t ko = 0;
        for (auto& fh : b->fields)
            if (fh.f.n[2] > 1) {
                fh.f.kx0 = static_cast<char*>(b->d_kx0) + ko * 2 * rsz;
                ko += (int64_t)fh.f.n[1] * fh.f.n[2];
            }
    }
    b->bytes = ws_elems * 2 * rsz + out_elems * rsz;
    int64_t wo = 0, oo = 0;
    for (auto& fh : b->fields) {
        FieldH& f = fh.f;
        f.ws = static_cast<char*>(b->d_ws) + wo * 2 * rsz;
        wo += f.wz * f.n[2] + (int64_t)f.nyq * f.n[1] * f.n[2];
        if (!b->global) {
            f.out = static_cast<char*>(b->d_out) + oo * rsz;
            f.out_ld1 = f.n[0];
            f.out_ld2 = (int64_t)f.n[0] * f.n[1];
            fh.out_elems = (int64_t)f.n[0] * f.n[1] * f.n[2];
            oo += fh.out_elems;
        }
    }
    SCARL_TRY(upload_tables(b));
    choose_schedule(b);
    std::vector<FieldH> hf;
    for (auto& fh : b->fields) hf.push_back(fh.f);
    SCARL_TRY(upload_vec(hf, (void**)&b->d_fields));
    SCARL_TRY(dmalloc((void**)&b->d_field_mom, std::max<size_t>(1, b->fields.size()) * kMomChunks * sizeof(MomPartial)));
    b->nslots = (int)slot_count.size();
    SCARL_TRY(upload_vec(slot_count, (void**)&b->d_slot_count));
    SCARL_TRY(dmalloc((void**)&b->d_slot_out, std::max<size_t>(1, slot_count.size()) * 5 * sizeof(double)));
    return 0;
}

// Passes of one group: z on stream zs (the generating pass of 3D fields), then y and x on s.
// only_plan >= 0 launches only the tile classes of that plan (the autotuner's timing).
static int enqueue_group(scarl_batch* b, const scarl_batch::Group& g, hipStream_t zs, hipStream_t s,
                         hipEvent_t z_done, int only_plan = -1) {
    const bool f64 = b->fp64;
    const size_t csz = f64 ? 16 : 8;
    // The launches of one pass (one per plan; in a blend's x pass, those of one colour, since a
    // later owner of an overlap point adds to what the earlier ones stored) touch disjoint fields
    // or subdomains.  In a batch that does not pipeline they run side by side on the stream of the
    // pass and up to three more (the generating stream is free there), which fills the GPU
    // through each launch's tail.  With event timing on, every launch runs alone.
    // In a pipelined batch the generating stream carries the next group's z pass, so only the x
    // launches fan out, over the two side streams.
    const bool pipelined = z_done != nullptr;
    const bool fan = !g_ctx.timing && only_plan < 0 && b->xev[0] != nullptr;
    using TC = scarl_batch::TileClass;
    auto fan_out = [&](const std::vector<TC>& cls, size_t i, size_t e, hipStream_t home, auto&& launch) -> int {
        hipStream_t side3[3] = {g_ctx.gen_stream, g_ctx.side[0], g_ctx.side[1]};
        hipStream_t* side = pipelined ? side3 + 1 : side3;
        const bool use = fan && (!pipelined || &cls == &g.cx);
        const int nside = use ? (int)std::min<size_t>(pipelined ? 2 : 3, e - i - 1) : 0;
        if (nside > 0) {
            HIP_TRY(hipEventRecord(b->xev[0], home));
            for (int k = 0; k < nside; ++k) HIP_TRY(hipStreamWaitEvent(side[k], b->xev[0], 0));
        }
        for (size_t c = i; c < e; ++c) {
            if (only_plan >= 0 && cls[c].plan != only_plan) continue;
            const int lane = nside > 0 ? (int)((c - i) % (size_t)(nside + 1)) : 0;
            SCARL_TRY(launch(cls[c], lane == 0 ? home : side[lane - 1]));
        }
        for (int k = 0; k < nside; ++k) {
            HIP_TRY(hipEventRecord(b->xev[1 + k], side[k]));
            HIP_TRY(hipStreamWaitEvent(home, b->xev[1 + k], 0));
        }
        return 0;
    };
    // the kx = 0 planes of the group's 3D fields, on the stream of its generating pass just before
    // it: in a pipelined batch only the first group's then lengthens the step
    if (g.kx0_cnt > 0 && only_plan < 0) {
        SCARL_TRY(timed("h_kx0", [&] {
            const int* ids = b->d_kx0_ids + g.kx0_off;
            return f64 ? launch_h_kx0_f64(b->d_fields, ids, g.kx0_cnt, g.kx0_max, zs)
                       : launch_h_kx0_f32(b->d_fields, ids, g.kx0_cnt, g.kx0_max, zs);
        }));
    }
    SCARL_TRY(fan_out(g.cz, 0, g.cz.size(), zs, [&](const TC& cl, hipStream_t st) {
        const PlanDev& pl = b->ps.plans[cl.plan];
        return timed("h_zpass", [&] {
            return launch_h_strided(f64, 2, true, pl.spec, pl.ecap, b->d_fields, b->d_tz, cl.t0, cl.t1 - cl.t0,
                                    b->ps.d_plans, b->ps.d_pool, (size_t)pl.lds_elems * csz, st);
        });
    }));
    if (z_done) {
        HIP_TRY(hipEventRecord(z_done, zs));
        HIP_TRY(hipStreamWaitEvent(s, z_done, 0));
    }
    SCARL_TRY(fan_out(g.cy, 0, g.cy.size(), s, [&](const TC& cl, hipStream_t st) {
        const PlanDev& pl = b->ps.plans[cl.plan];
        return timed("h_ypass", [&] {
            return launch_h_strided(f64, 1, cl.gen, pl.spec, pl.ecap, b->d_fields, b->d_ty, cl.t0, cl.t1 - cl.t0,
                                    b->ps.d_plans, b->ps.d_pool, (size_t)pl.lds_elems * csz, st);
        });
    }));
    for (size_t i = 0; i < g.cx.size();) {
        size_t e = i + 1;
        while (e < g.cx.size() && g.cx[e].color == g.cx[i].color) ++e;
        SCARL_TRY(fan_out(g.cx, i, e, s, [&](const TC& cl, hipStream_t st) {
            const PlanDev& pl = b->ps.plans[cl.plan];
            // x-pass plans of odd n1 are plain C2C (no pre-twiddle table); a blend keeps its x
            // weights (n1 reals) in LDS after the transform's tile
            const bool full = pl.pre_off < 0;
            const size_t lds =
                (size_t)pl.lds_elems * csz + (b->global ? (size_t)(full ? pl.N : 2 * pl.N) * (csz / 2) : 0);
            return timed("h_xpass", [&] {
                return launch_h_xpass(f64, pl.spec, pl.ecap, b->global, full, b->d_fields, b->d_tx, cl.t0,
                                      cl.t1 - cl.t0, b->ps.d_plans, b->ps.d_pool, b->d_partials, lds, st);
            });
        }));
        i = e;
    }
    return 0;
}

// One step of a batch on the library stream.  Several groups pipeline over two streams (the z
// passes on gen_stream, each group's y and x passes on the library stream after its z pass), except
// with event timing on, where every launch runs alone on the library stream.
static int enqueue_batch(scarl_batch* b) {
    hipStream_t s = g_ctx.stream;
    const bool pipe = b->groups.size() > 1 && !g_ctx.timing && b->pipeline;
    if (pipe) {
        HIP_TRY(hipEventRecord(b->gev[0], s));
        HIP_TRY(hipStreamWaitEvent(g_ctx.gen_stream, b->gev[0], 0));
    }
    for (size_t g = 0; g < b->groups.size(); ++g)
        SCARL_TRY(enqueue_group(b, b->groups[g], pipe ? g_ctx.gen_stream : s, s, pipe ? b->gev[g + 1] : nullptr));
    if (b->nslots > 0) {
        SCARL_TRY(timed("moments", [&] {
            return launch_mom(b->d_fields, (int)b->fields.size(), b->d_field_xtiles, b->d_partials, b->d_field_mom,
                              b->d_slot_count, b->nslots, b->d_slot_out, s);
        }));
    }
    return 0;
}

static int upload_fields(scarl_batch* b) {
    std::vector<FieldH> hf;
    for (auto& fh : b->fields) hf.push_back(fh.f);
    HIP_TRY(hipMemcpy(b->d_fields, hf.data(), hf.size() * sizeof(FieldH), hipMemcpyHostToDevice));
    return 0;
}

static void drop_graph(scarl_batch* b) {
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

// The batch's step captured from the library stream (and the second stream it forks to) into a
// hipGraph, instantiated once; the tables it points to stay fixed until the next drop_graph.
static int ensure_graph(scarl_batch* b) {
    if (b->gexec) return 0;
    HIP_TRY(hipStreamBeginCapture(g_ctx.stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_batch(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(g_ctx.stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(SCARL_E_HIP, "graph capture: %s", hipGetErrorString(e));
    b->graph = g;
    e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return fail(SCARL_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    return 0;
}

// ------------------------------------------------------------------------------------------
// autotuner: per plan of a batch (length, layout, kind, role), every candidate of the search
// space (compile-time kernel, runtime plans with 16 or 32 elements per thread, radix order,
// pencils per tile, LDS skew, MFMA first stage) is timed on the batch's own launches of that plan
// with HIP events.  That isolated timing ranks the candidates; the default and the best three are
// then timed again in graph replays of the batch's whole step, since a plan's speed there also
// depends on what the previous pass leaves in L2 and the Infinity Cache and on the z passes
// running beside it on the second stream (an x plan 11 % faster alone made the step of a
// 256-field ensemble 3 % slower).  Picks are cached for the process per precision, plan key and
// batch geometry, and in the file named by SCARL_TUNE_CACHE when it is set (one "key choice" per
// line; a key without the "@geometry" suffix applies to every batch, which is how choices are
// forced).  The reference's counterpart is FFTW's planner (SRC_F/fftw_ScaRL.f90:378,
// FFTW_ESTIMATE).
// ------------------------------------------------------------------------------------------
struct TunePick {
    PlanChoice choice;
    double ms = 0.0, default_ms = 0.0;            // isolated launches of the plan
    double step_ms = 0.0, default_step_ms = 0.0;  // whole steps (0 when one candidate was left)
    int tried = 0;
};
static std::map<std::string, TunePick> g_tune_cache;
static bool g_tune_file_loaded = false;

static void tune_file_load() {
    if (g_tune_file_loaded) return;
    g_tune_file_loaded = true;
    const char* path = getenv("SCARL_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = fopen(path, "r");
    if (!f) return;
    char key[128], ch[128];
    while (fscanf(f, "%127s %127s", key, ch) == 2) {
        TunePick tp;
        if (parse_choice(ch, tp.choice)) g_tune_cache[key] = tp;
    }
    fclose(f);
}

static void tune_file_append(const std::string& key, const PlanChoice& c) {
    const char* path = getenv("SCARL_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s %s\n", key.c_str(), choice_name(c).c_str());
    fclose(f);
}

static std::vector<PlanChoice> plan_candidates(const PlanSet& ps, int N, int layout) {
    std::vector<PlanChoice> c(1);  // the default rules first
    for (int fs : ps.fixed_specs(N, layout)) {
        PlanChoice f;
        f.variant = PV_FIXED;
        f.spec = fs;
        c.push_back(f);
    }
    const std::vector<int> skews = layout == LAYOUT_BF ? std::vector<int>{1, 0} : std::vector<int>{1};
    for (int ec : ps.fp64 ? std::vector<int>{16} : std::vector<int>{16, 32})
        for (int comp : {0, 1})
            for (int order : {0, 1})
                for (int tdiv : layout == LAYOUT_PF ? std::vector<int>{1, 2, 0} : std::vector<int>{1, 2})
                    for (int sk : skews) {
                        PlanChoice g;
                        g.variant = PV_GENERIC;
                        g.ecap = ec;
                        g.order = order;
                        g.tdiv = tdiv;
                        g.skew = sk;
                        g.comp = comp;
                        c.push_back(g);
                    }
    {
        PlanChoice d;
        d.variant = PV_DIRECT;  // applies only to lengths with a prime factor above 13
        c.push_back(d);
    }
    if (!ps.fp64)
        for (int tdiv : {1, 2})
            for (int sk : skews) {
                PlanChoice m;
                m.variant = PV_MFMA;
                m.tdiv = tdiv;
                m.skew = sk;
                c.push_back(m);
            }
    return c;
}

// what distinguishes two plans' kernels and tiles (candidates with the same signature are one)
static std::string plan_signature(const PlanDev& p) {
    std::string s = std::to_string(p.spec) + "/" + std::to_string(p.ecap) + "/" + std::to_string(p.T) + "/" +
                    std::to_string(p.tpp) + "/" + std::to_string(p.ldp) + "/" + std::to_string(p.skew) + "/";
    for (int i = 0; i < p.nst; ++i) s += std::to_string(p.R[i]) + ",";
    return s + "/" + std::to_string(p.blu_M);
}

// Average time of the batch's launches of one plan (every group), eager on the library stream.
static int time_plan(scarl_batch* b, int plan, int reps, double* ms) {
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = 0;
    for (int r = -1; r < reps && !rc; ++r) {
        if (r == 0 && hipEventRecord(e0, g_ctx.stream) != hipSuccess) rc = fail(SCARL_E_HIP, "event record");
        for (const auto& g : b->groups) {
            if (rc) break;
            rc = enqueue_group(b, g, g_ctx.stream, g_ctx.stream, nullptr, plan);
        }
    }
    float t = 0.f;
    if (!rc) {
        hipError_t e = hipEventRecord(e1, g_ctx.stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e != hipSuccess) rc = fail(SCARL_E_HIP, "autotuner timing: %s", hipGetErrorString(e));
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms = t / reps;
    return rc;
}

// Average time of a whole step of the batch as a graph replay (the z passes pipelined beside the
// y and x passes of the previous group), after one untimed replay.  The graph is captured afresh
// for the current tables.  Eager launches of the same step order the two streams' work otherwise:
// a y plan measured faster there made the replayed step of a 256-field ensemble 0.8 % slower.
static int time_step(scarl_batch* b, int reps, double* ms) {
    drop_graph(b);
    SCARL_TRY(ensure_graph(b));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = 0;
    for (int r = -1; r < reps && !rc; ++r) {
        if (r == 0 && hipEventRecord(e0, g_ctx.stream) != hipSuccess) rc = fail(SCARL_E_HIP, "event record");
        if (!rc && hipGraphLaunch(b->gexec, g_ctx.stream) != hipSuccess) rc = fail(SCARL_E_HIP, "graph launch");
    }
    float t = 0.f;
    if (!rc) {
        hipError_t e = hipEventRecord(e1, g_ctx.stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e != hipSuccess) rc = fail(SCARL_E_HIP, "autotuner timing: %s", hipGetErrorString(e));
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms = t / reps;
    return rc;
}

// "@" and 16 hex digits of an FNV-1a hash over what fixes a batch's passes: precision, MFMA
// stages, global extents, and the sorted field extents.
static std::string tune_context(const scarl_batch* b) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](int64_t v) {
        for (int i = 0; i < 8; ++i) {
            h ^= (uint64_t)(v >> (8 * i)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    mix(b->fp64);
    mix(b->ps.mfma);
    mix(b->global);
    for (int d = 0; d < 3; ++d) mix(b->global ? b->gdim[d] : 0);
    std::vector<std::array<int64_t, 3>> shapes;
    for (const auto& fh : b->fields) shapes.push_back({fh.f.n[0], fh.f.n[1], fh.f.n[2]});
    std::sort(shapes.begin(), shapes.end());
    for (const auto& n : shapes)
        for (int64_t v : n) mix(v);
    char buf[24];
    snprintf(buf, sizeof buf, "@%016llx", (unsigned long long)h);
    return buf;
}

static std::string plan_line(const PlanSet& ps, int i) {
    const PlanDev& p = ps.plans[i];
    int N, layout, kind, role;
    sscanf(ps.keys[i].c_str(), "%d/%d/%d/%d", &N, &layout, &kind, &role);
    static const char* roles[3] = {"y", "gen", "x"};
    char buf[256];
    std::string rad;
    for (int s = 0; s < p.nst; ++s) rad += (s ? "x" : "") + std::to_string(p.R[s]);
    if (p.nst == 0) rad = "direct";
    if (p.blu_M > 0) rad = "bluestein" + std::to_string(p.blu_M) + ":" + rad;
    const char* kern = p.spec == kSpecMfmaGeneric ? "mfma-generic" : p.spec == kSpecMfma256 ? "mfma-256"
                       : p.spec == kSpecNonSmooth ? (p.blu_M > 0 ? "bluestein" : "direct")
                       : p.spec == kSpecGenericComp ? "generic-composite"
                       : p.spec ? "fixed" : "generic";
    snprintf(buf, sizeof buf, "pass=%s N=%d kernel=%s spec=%d radices=%s ecap=%d T=%d tpp=%d ldp=%d skew=%d choice=%s",
             role >= 0 && role < 3 ? roles[role] : "?", N, kern, p.spec, rad.c_str(), p.ecap, p.T, p.tpp, p.ldp,
             p.skew, choice_name(ps.choices[i]).c_str());
    return buf;
}

static void swap_plansets(PlanSet& a, PlanSet& b) {
    std::swap(a.plans, b.plans);
    std::swap(a.pool, b.pool);
    std::swap(a.index, b.index);
    std::swap(a.keys, b.keys);
    std::swap(a.choices, b.choices);
    std::swap(a.overrides, b.overrides);
    std::swap(a.d_plans, b.d_plans);
    std::swap(a.d_pool, b.d_pool);
    std::swap(a.fp64, b.fp64);
    std::swap(a.mfma, b.mfma);
    std::swap(a.ecap, b.ecap);
    std::swap(a.max_lds, b.max_lds);
}

static int autotune_batch(scarl_batch* b) {
    drop_graph(b);
    tune_file_load();
    const bool saved_timing = g_ctx.timing;
    g_ctx.timing = false;
    struct Restore {
        bool v;
        ~Restore() { g_ctx.timing = v; }
    } restore{saved_timing};
    // one generation with the default plans, so that every workspace holds field data
    SCARL_TRY(enqueue_batch(b));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    const std::string prec = b->fp64 ? "fp64/" : "fp32/";
    const std::string where = tune_context(b);
    const int nplans = (int)b->ps.plans.size();
    b->tune_notes.assign(nplans, "");
    std::map<std::string, PlanChoice> picks;
    struct Cand {
        PlanChoice c;
        PlanDev pd;
        double ms;
    };
    std::vector<std::vector<Cand>> shortlist(nplans);  // [0] is the default
    std::vector<TunePick> best(nplans);
    std::vector<int> order;
    for (int pi = 0; pi < nplans; ++pi) {
        const std::string key = b->ps.keys[pi];
        int N, layout, kind, role;
        sscanf(key.c_str(), "%d/%d/%d/%d", &N, &layout, &kind, &role);
        auto hit = g_tune_cache.find(prec + key + where);
        if (hit == g_tune_cache.end()) hit = g_tune_cache.find(prec + key);
        if (hit != g_tune_cache.end()) {
            picks[key] = hit->second.choice;
            b->tune_notes[pi] = " cached";
            PlanDev pd;  // the steps of the second phase run with the cached pick
            if (b->ps.make(N, layout, kind, hit->second.choice, pd) == 0) b->ps.plans[pi] = pd;
            continue;
        }
        const PlanDev keep = b->ps.plans[pi];
        std::vector<Cand> cands;
        std::map<std::string, bool> seen;
        for (const PlanChoice& c : plan_candidates(b->ps, N, layout)) {
            PlanDev pd;
            if (b->ps.make(N, layout, kind, c, pd) != 0) continue;
            const std::string sig = plan_signature(pd);
            if (seen.count(sig)) continue;
            seen[sig] = true;
            if (check_lds((size_t)pd.lds_elems * (b->fp64 ? 16 : 8)) != 0) continue;
            b->ps.plans[pi] = pd;
            SCARL_TRY(upload_tables(b));
            double ms = 0.0;
            SCARL_TRY(time_plan(b, pi, 3, &ms));
            cands.push_back({c, pd, ms});
        }
        b->ps.plans[pi] = keep;
        if (cands.empty()) continue;
        TunePick& tp = best[pi];
        tp.tried = (int)cands.size();
        int di = -1;
        for (int k = 0; k < (int)cands.size(); ++k)
            if (cands[k].c.variant == PV_DEFAULT) di = k;
        if (di < 0) {  // no default plan among the candidates: the isolated fastest
            int bi = 0;
            for (int k = 1; k < (int)cands.size(); ++k)
                if (cands[k].ms < cands[bi].ms) bi = k;
            tp.choice = cands[bi].c;
            tp.ms = tp.default_ms = cands[bi].ms;
            b->ps.plans[pi] = cands[bi].pd;
            continue;
        }
        tp.choice = cands[di].c;
        tp.ms = tp.default_ms = cands[di].ms;
        shortlist[pi].push_back(cands[di]);
        std::vector<int> idx;
        for (int k = 0; k < (int)cands.size(); ++k)
            if (k != di && cands[k].ms < cands[di].ms) idx.push_back(k);
        std::sort(idx.begin(), idx.end(), [&](int a, int c) { return cands[a].ms < cands[c].ms; });
        for (int k = 0; k < (int)idx.size() && k < 3; ++k) shortlist[pi].push_back(cands[idx[k]]);
        if (shortlist[pi].size() > 1) order.push_back(pi);
    }
    // second phase: plan by plan, the one with the longest default launches first, the shortlisted
    // candidates in replayed steps, taken in turn three times over (so that clock drift falls on
    // every candidate alike), each run about 40 ms of replays; a candidate's time is its fastest
    // run, and it replaces the default only when the step gets faster by more than 0.3 %
    std::sort(order.begin(), order.end(), [&](int a, int c) { return best[a].default_ms > best[c].default_ms; });
    int reps = 4;
    if (!order.empty() || (b->groups.size() > 1 && g_ctx.schedule == 0)) {
        double t = 0.0;
        SCARL_TRY(upload_tables(b));
        SCARL_TRY(time_step(b, 2, &t));
        reps = std::max(4, std::min(50, (int)std::ceil(40.0 / std::max(t, 1e-3))));
    }
    for (int pi : order) {
        const auto& cl = shortlist[pi];
        std::vector<double> tk(cl.size(), 1e30);
        for (int round = 0; round < 3; ++round)
            for (int k = 0; k < (int)cl.size(); ++k) {
                b->ps.plans[pi] = cl[k].pd;
                SCARL_TRY(upload_tables(b));
                double t = 0.0;
                SCARL_TRY(time_step(b, reps, &t));
                tk[k] = std::min(tk[k], t);
            }
        int pick = 0;
        for (int k = 1; k < (int)cl.size(); ++k)
            if (tk[k] < tk[0] * 0.997 && tk[k] < tk[pick]) pick = k;
        b->ps.plans[pi] = cl[pick].pd;
        best[pi].choice = cl[pick].c;
        best[pi].ms = cl[pick].ms;
        best[pi].step_ms = tk[pick];
        best[pi].default_step_ms = tk[0];
    }
    // the schedule of a batch of several groups: pipelined or one group after another, timed
    // like the plans (in turn, three times over, fastest run each) with the picked plans
    if (b->groups.size() > 1 && g_ctx.schedule == 0) {
        SCARL_TRY(upload_tables(b));
        const bool dflt = b->pipeline;
        double ts[2] = {1e30, 1e30};
        for (int round = 0; round < 3; ++round)
            for (int k = 0; k < 2; ++k) {
                b->pipeline = k == 0 ? dflt : !dflt;
                double t = 0.0;
                SCARL_TRY(time_step(b, reps, &t));
                ts[k] = std::min(ts[k], t);
            }
        b->pipeline = ts[1] < ts[0] * 0.997 ? !dflt : dflt;
    }
    for (int pi = 0; pi < nplans; ++pi) {
        const TunePick& tp = best[pi];
        if (tp.tried == 0) continue;
        const std::string key = b->ps.keys[pi];
        picks[key] = tp.choice;
        g_tune_cache[prec + key + where] = tp;
        tune_file_append(prec + key + where, tp.choice);
        char note[192];
        if (tp.step_ms > 0.0)
            snprintf(note, sizeof note, " tuned_ms=%.4f default_ms=%.4f step_ms=%.4f default_step_ms=%.4f candidates=%d",
                     tp.ms, tp.default_ms, tp.step_ms, tp.default_step_ms, tp.tried);
        else
            snprintf(note, sizeof note, " tuned_ms=%.4f default_ms=%.4f candidates=%d", tp.ms, tp.default_ms, tp.tried);
        b->tune_notes[pi] = note;
    }
    // a compact plan set with the picks, same indices (same keys in the same order)
    PlanSet fresh;
    fresh.fp64 = b->ps.fp64;
    fresh.mfma = b->ps.mfma;
    fresh.ecap = b->ps.ecap;
    fresh.overrides = picks;
    for (int pi = 0; pi < nplans; ++pi) {
        int N, layout, kind, role;
        sscanf(b->ps.keys[pi].c_str(), "%d/%d/%d/%d", &N, &layout, &kind, &role);
        const int id = fresh.get(N, layout, kind, role);
        if (id != pi) return fail(SCARL_E_HIP, "autotuner: plan %d rebuilt as %d", pi, id < 0 ? -id : id);
    }
    swap_plansets(b->ps, fresh);
    drop_graph(b);  // the timing graphs point to the tables replaced here
    SCARL_TRY(upload_tables(b));
    return 0;
}

// ------------------------------------------------------------------------------------------
// mode R plain path (one field per call, FP64)
// ------------------------------------------------------------------------------------------
struct ModeRCache {
    int n[3] = {0, 0, 0}, m[3] = {0, 0, 0};
    std::unique_ptr<PlanSet> ps;
    int plan[3] = {0, 0, 0};
    size_t lds[3] = {0, 0, 0};
};
static ModeRCache g_r;
// FP64 plans of the FFT path of the lag sums (lengths P), kept for the process and uploaded when
// they grow.
static std::unique_ptr<PlanSet> g_lag_ps;
static size_t g_lag_uploaded = 0;

static int mode_r_prepare(const int32_t n[3], const int32_t m[3]) {
    if (g_r.ps && memcmp(g_r.n, n, sizeof g_r.n) == 0 && memcmp(g_r.m, m, sizeof g_r.m) == 0) return 0;
    std::unique_ptr<PlanSet> ps(new PlanSet());
    ps->fp64 = true;
    int plan[3];
    for (int a = 0; a < 3; ++a) {
        plan[a] = ps->get(n[a], a == 0 ? LAYOUT_BF : LAYOUT_PF, 2);
        if (plan[a] < 0) return -plan[a];
    }
    SCARL_TRY(ps->upload());
    for (int a = 0; a < 3; ++a) {
        g_r.lds[a] = (size_t)ps->plans[plan[a]].lds_elems * 16;
        SCARL_TRY(check_lds(g_r.lds[a]));
        g_r.plan[a] = plan[a];
        g_r.n[a] = n[a];
        g_r.m[a] = m[a];
    }
    g_r.ps = std::move(ps);
    return 0;
}

// The FieldR of one call (SRC_F/fftw_ScaRL.f90:300-302 k spacing); plans and workspaces are
// filled in by the caller.
static int mode_r_desc(FieldR& f, double* d_out, const int32_t n[3], const int32_t m[3], const double xRange[3],
                       const double corrL[3], int32_t corrMod, uint64_t seed, const double* d_phi, int32_t flags) {
    if (corrMod != CM_GAUSSIAN)
        return fail(SCARL_E_UNSUPPORTED, "gen_std_gauss_FFT implements corrMod = cm_GAUSSIAN only (got %d)", corrMod);
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 1) return fail(SCARL_E_INVALID, "Np(%d) = %d must be >= 1", a + 1, n[a]);
        if (!(corrL[a] > 0.0)) return fail(SCARL_E_INVALID, "corrL(%d) must be positive", a + 1);
    }
    memset(&f, 0, sizeof f);
    for (int a = 0; a < 3; ++a) {
        f.n[a] = n[a];
        f.m[a] = m[a];
        const double delta_x = xRange[a] / (double)(n[a] - 1);
        const double k_max = 1.0 / delta_x;
        f.dk[a] = k_max / (double)(n[a] - 1);
        f.corrL[a] = corrL[a];
    }
    f.src_mode = d_phi ? SRC_HOST_NOISE : SRC_DEVICE_NOISE;
    f.phi = d_phi;
    f.seed = seed;
    f.stream = 0;
    f.margi = (flags & 1) ? g_ctx.margi : FOM_NONE;
    SCARL_TRY(marginal_params(f.margi, g_ctx.m_avg, g_ctx.m_std, &f.mu, &f.sd));
    f.out = d_out;
    return 0;
}

// Workspace elements (doubles) of the x and y passes: tmp1 (n1, m2, m3), tmp2 (n1, n2, m3).
static void mode_r_tmp_elems(const int32_t n[3], const int32_t m[3], int64_t* t1, int64_t* t2) {
    *t1 = std::max<int64_t>(1, (int64_t)n[0] * m[1] * m[2]);
    *t2 = std::max<int64_t>(1, (int64_t)n[0] * n[1] * m[2]);
}

// The three passes x, y, z of a filled FieldR.  host_plans: the plan table (host copy) whose
// device copy is d_plans; use_timed = 0 launches without events (no event objects are created).
static int mode_r_passes(const FieldR& f, const PlanDev* host_plans, const PlanDev* d_plans, const void* d_pool,
                         const size_t lds[3], bool use_timed) {
    // the descriptor travels by value in the kernel arguments (no host buffer outlives the call)
    const int64_t npen_real[3] = {(int64_t)f.m[1] * f.m[2], (int64_t)f.n[0] * f.m[2], (int64_t)f.n[0] * f.n[1]};
    static const char* names[3] = {"r_xpass", "r_ypass", "r_zpass"};
    for (int a = 0; a < 3; ++a) {
        const PlanDev& pl = host_plans[f.plan[a]];
        const int64_t ncplx = (npen_real[a] + 1) / 2;
        const int64_t nblocks = (ncplx + pl.T - 1) / pl.T;
        if (nblocks > 0x7fffffff) return fail(SCARL_E_UNSUPPORTED, "grid too large");
        if (a == 0 && (f.m[1] == 0 || f.m[2] == 0)) {
            // empty octant: x pass has no pencils; tmp1 content is irrelevant (m2 or m3 = 0)
            continue;
        }
        auto launch = [&] { return launch_r_pass(a, pl.spec, f, d_plans, d_pool, (int)nblocks, lds[a], g_ctx.stream); };
        if (use_timed) {
            SCARL_TRY(timed(names[a], launch));
        } else {
            const hipError_t e = launch();
            if (e != hipSuccess) return fail(SCARL_E_HIP, "launch of %s failed: %s", names[a], hipGetErrorString(e));
        }
    }
    return 0;
}

// d_out, d_phi device pointers; m = octant support
static int mode_r_run(double* d_out, const int32_t n[3], const int32_t m[3], const double xRange[3],
                      const double corrL[3], int32_t corrMod, uint64_t seed, const double* d_phi, int32_t flags) {
    FieldR f;
    SCARL_TRY(mode_r_desc(f, d_out, n, m, xRange, corrL, corrMod, seed, d_phi, flags));
    SCARL_TRY(mode_r_prepare(n, m));
    for (int a = 0; a < 3; ++a) f.plan[a] = g_r.plan[a];
    int64_t t1, t2;
    mode_r_tmp_elems(n, m, &t1, &t2);
    void *p1, *p2;
    SCARL_TRY(scratch_get("r_tmp1", (size_t)t1 * 8, &p1));
    SCARL_TRY(scratch_get("r_tmp2", (size_t)t2 * 8, &p2));
    f.tmp1 = static_cast<double*>(p1);
    f.tmp2 = static_cast<double*>(p2);
    return mode_r_passes(f, g_r.ps->plans.data(), g_r.ps->d_plans, g_r.ps->d_pool, g_r.lds, true);
}

// The same passes with every device buffer (plan table, twiddle pool, tmp1, tmp2) carved out of
// the library's static scratch array (scarl_noalloc.hip, noalloc_scratch()): no device memory is
// allocated.  When the shape differs from the last call's on this device, the plan table and the
// pool are copied into the scratch in stream order and the host waits for that copy; the passes
// themselves run asynchronously, as in the _dev call.
// Per device: the shape whose plan table and twiddles the scratch holds (a repeated shape skips
// the planning, the copy and its wait).
struct NoAllocTables {
    bool valid = false;
    int n[3] = {0, 0, 0}, m[3] = {0, 0, 0};
    std::vector<PlanDev> plans;
    int plan[3] = {0, 0, 0};
    size_t lds[3] = {0, 0, 0};
    size_t o_pool = 0, o_t1 = 0, o_t2 = 0;
};
static NoAllocTables g_noalloc_tables[64];

static int mode_r_run_noalloc(double* d_out, const int32_t n[3], const int32_t m[3], const double xRange[3],
                              const double corrL[3], int32_t corrMod, uint64_t seed, const double* d_phi,
                              int32_t flags) {
    size_t cap = 0;
    char* base = static_cast<char*>(noalloc_scratch(g_ctx.device, &cap));
    if (!base) return fail(SCARL_E_HIP, "%s", noalloc_error(g_ctx.device));
    if (g_ctx.device < 0 || g_ctx.device >= 64) return fail(SCARL_E_INVALID, "device index out of range");
    FieldR f;
    SCARL_TRY(mode_r_desc(f, d_out, n, m, xRange, corrL, corrMod, seed, d_phi, flags));
    NoAllocTables& tb = g_noalloc_tables[g_ctx.device];
    const bool same = tb.valid && memcmp(tb.n, n, sizeof tb.n) == 0 && memcmp(tb.m, m, sizeof tb.m) == 0;
    if (!same) {
        tb.valid = false;
        PlanSet ps;  // host side only: upload() is never called, so it owns no device memory
        ps.fp64 = true;
        for (int a = 0; a < 3; ++a) {
            tb.plan[a] = ps.get(n[a], a == 0 ? LAYOUT_BF : LAYOUT_PF, 2);
            if (tb.plan[a] < 0) return -tb.plan[a];
        }
        for (int a = 0; a < 3; ++a) {
            tb.lds[a] = (size_t)ps.plans[tb.plan[a]].lds_elems * 16;
            SCARL_TRY(check_lds(tb.lds[a]));
        }
        // layout: [plan table | twiddle pool (FP64 complex) | tmp1 | tmp2], each 256-byte aligned
        auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
        std::vector<double> pool(std::max<size_t>(2, ps.pool.size()), 0.0);
        for (size_t i = 0; i < ps.pool.size(); ++i) pool[i] = (double)ps.pool[i];
        const size_t plan_bytes = ps.plans.size() * sizeof(PlanDev), pool_bytes = pool.size() * sizeof(double);
        int64_t t1, t2;
        mode_r_tmp_elems(n, m, &t1, &t2);
        tb.o_pool = up(plan_bytes);
        tb.o_t1 = tb.o_pool + up(pool_bytes);
        tb.o_t2 = tb.o_t1 + up((size_t)t1 * 8);
        const size_t need = tb.o_t2 + up((size_t)t2 * 8);
        if (need > cap)
            return fail(SCARL_E_INVALID,
                        "Np = (%d, %d, %d) needs %zu bytes of scratch; the allocation-free entry point holds %zu "
                        "(workspaces up to 96 MiB, e.g. 256^3 points)", n[0], n[1], n[2], need, cap);
        // in stream order, after any earlier call's kernels have read the old tables; the host
        // waits so that the host tables may die on return
        HIP_TRY(hipMemcpyAsync(base, ps.plans.data(), plan_bytes, hipMemcpyHostToDevice, g_ctx.stream));
        HIP_TRY(hipMemcpyAsync(base + tb.o_pool, pool.data(), pool_bytes, hipMemcpyHostToDevice, g_ctx.stream));
        HIP_TRY(hipStreamSynchronize(g_ctx.stream));
        tb.plans = ps.plans;
        memcpy(tb.n, n, sizeof tb.n);
        memcpy(tb.m, m, sizeof tb.m);
        tb.valid = true;
    }
    for (int a = 0; a < 3; ++a) f.plan[a] = tb.plan[a];
    f.tmp1 = reinterpret_cast<double*>(base + tb.o_t1);
    f.tmp2 = reinterpret_cast<double*>(base + tb.o_t2);
    return mode_r_passes(f, tb.plans.data(), reinterpret_cast<const PlanDev*>(base), base + tb.o_pool, tb.lds, false);
}

}  // namespace scarl

using namespace scarl;

// ==========================================================================================
// extern "C" entry points
// ==========================================================================================
extern "C" {

int scarl_hip_init(int32_t device) {
    if (g_ctx.ready && device != g_ctx.device) return fail(SCARL_E_INVALID, "already initialised on device %d", g_ctx.device);
    g_ctx.device = device;
    return ensure_init();
}

int scarl_hip_finalize(void) {
    if (!g_ctx.ready) return 0;
    hipStreamSynchronize(g_ctx.stream);
    for (auto& kv : g_ctx.scratch) hipFree(kv.second.first);
    g_ctx.scratch.clear();
    for (auto& t : g_ctx.pending) {
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    g_ctx.pending.clear();
    g_r.ps.reset();
    memset(g_r.n, 0, sizeof g_r.n);
    g_lag_ps.reset();
    g_lag_uploaded = 0;
    for (int i = 0; i < 2; ++i) {
        if (g_ctx.stage_ev[i]) {
            hipEventSynchronize(g_ctx.stage_ev[i]);
            hipEventDestroy(g_ctx.stage_ev[i]);
        }
        if (g_ctx.stage[i]) hipHostFree(g_ctx.stage[i]);
        g_ctx.stage[i] = nullptr;
        g_ctx.stage_ev[i] = nullptr;
    }
    hipStreamDestroy(g_ctx.stream);
    hipStreamDestroy(g_ctx.gen_stream);
    for (hipStream_t& st : g_ctx.side) {
        hipStreamDestroy(st);
        st = nullptr;
    }
    g_ctx.stream = nullptr;
    g_ctx.gen_stream = nullptr;
    g_ctx.ready = false;
    return 0;
}

int scarl_hip_last_error(char* buf, int32_t len) {
    if (!buf || len <= 0) return SCARL_E_INVALID;
    snprintf(buf, (size_t)len, "%s", g_err.c_str());
    return 0;
}

int scarl_hip_set_timing(int32_t on) {
    SCARL_TRY(ensure_init());
    g_ctx.timing = on != 0;
    return 0;
}

int scarl_hip_timing_report(char* buf, int32_t len) {
    SCARL_TRY(ensure_init());
    SCARL_TRY(timing_collect());
    std::string s;
    char line[256];
    for (auto& kv : g_ctx.totals) {
        snprintf(line, sizeof line, "%s %ld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
        s += line;
    }
    if (!buf || len <= 0) return SCARL_E_INVALID;
    snprintf(buf, (size_t)len, "%s", s.c_str());
    return 0;
}

int scarl_hip_timing_reset(void) {
    SCARL_TRY(ensure_init());
    SCARL_TRY(timing_collect());
    g_ctx.totals.clear();
    return 0;
}

int scarl_hip_get_stream(void** stream) {
    SCARL_TRY(ensure_init());
    if (!stream) return fail(SCARL_E_INVALID, "null stream pointer");
    *stream = (void*)g_ctx.stream;
    return 0;
}

int scarl_hip_set_tuning(int32_t key, int32_t value) {
    if (key == SCARL_TUNE_FP32_ELEMS) {
        if (value != 16 && value != 32) return fail(SCARL_E_INVALID, "FP32 elements per thread must be 16 or 32");
        g_ctx.fp32_ecap = value;
        return 0;
    }
    if (key == SCARL_TUNE_MFMA_DFT16) {
        if (value < 0 || value > 2) return fail(SCARL_E_INVALID, "SCARL_TUNE_MFMA_DFT16 takes 0, 1 or 2");
        g_ctx.mfma_dft16 = value;
        return 0;
    }
    if (key == SCARL_TUNE_LAG_FFT) {
        if (value != 0 && value != 1) return fail(SCARL_E_INVALID, "SCARL_TUNE_LAG_FFT takes 0 or 1");
        g_ctx.lag_fft = value != 0;
        return 0;
    }
    if (key == SCARL_TUNE_AUTOTUNE) {
        if (value != 0 && value != 1) return fail(SCARL_E_INVALID, "SCARL_TUNE_AUTOTUNE takes 0 or 1");
        g_ctx.autotune = value != 0;
        return 0;
    }
    if (key == SCARL_TUNE_SCHEDULE) {
        if (value < 0 || value > 2) return fail(SCARL_E_INVALID, "SCARL_TUNE_SCHEDULE takes 0, 1 or 2");
        g_ctx.schedule = value;
        return 0;
    }
    return fail(SCARL_E_INVALID, "unknown tuning key %d", key);
}

int scarl_hip_set_marginal(int32_t margiFirst, double avg, double std_dev) {
    double mu, sd;
    SCARL_TRY(marginal_params(margiFirst, avg, std_dev, &mu, &sd));
    g_ctx.margi = margiFirst;
    g_ctx.m_avg = avg;
    g_ctx.m_std = std_dev;
    return 0;
}

// ---- fftw_ScaRL --------------------------------------------------------------------------
static int gen_r_host(double* h_out, const int32_t n[3], const int32_t m[3], const double xRange[3],
                      const double corrL[3], int32_t corrMod, int32_t seed, int32_t rank, const double* h_phi,
                      int32_t flags) {
    SCARL_TRY(ensure_init());
    if (!h_out) return fail(SCARL_E_INVALID, "null output");
    const int64_t N = (int64_t)n[0] * n[1] * n[2];
    const int64_t P = (int64_t)m[0] * m[1] * m[2];
    void *d_out, *d_phi = nullptr;
    SCARL_TRY(scratch_get("r_out", (size_t)std::max<int64_t>(1, N) * 8, &d_out));
    if (h_phi) {
        SCARL_TRY(scratch_get("r_phi", (size_t)std::max<int64_t>(1, P) * 8, &d_phi));
        if (P > 0) SCARL_TRY(copy_to_device(d_phi, h_phi, (size_t)P * 8));
    }
    const uint64_t key = (uint64_t)(int64_t)((int64_t)seed + rank);
    SCARL_TRY(mode_r_run((double*)d_out, n, m, xRange, corrL, corrMod, key, (const double*)d_phi, flags));
    SCARL_TRY(copy_to_host(h_out, d_out, (size_t)N * 8));
    return 0;
}

int scarl_hip_gen_std_gauss_fft(double* h_out, const int32_t Np[3], const double xRange[3], const double corrL[3],
                                int32_t corrMod, int32_t seed, int32_t rank, const double* h_phi, int32_t flags) {
    if (!Np || !xRange || !corrL) return fail(SCARL_E_INVALID, "null argument");
    const int32_t m[3] = {Np[0] / 2, Np[1] / 2, Np[2] / 2};
    return gen_r_host(h_out, Np, m, xRange, corrL, corrMod, seed, rank, h_phi, flags);
}

int scarl_hip_gen_std_gauss_fft_mpi(double* h_out, const double xRange[3], const double corrL[3], int32_t corrMod,
                                    int32_t seed, int32_t rank, const int32_t L[3], const int32_t Np[3],
                                    int32_t comm_size, const double* h_phi, int32_t flags) {
    if (!Np || !xRange || !corrL || !L) return fail(SCARL_E_INVALID, "null argument");
    if (comm_size != 1)
        return fail(SCARL_E_UNSUPPORTED, "one GPU holds the whole field: comm_size must be 1 (got %d)", comm_size);
    if (Np[0] != L[0] || Np[1] != L[1] || Np[2] != L[2]) return fail(SCARL_E_INVALID, "Error in Np definition inside FFT_MPI");
    const int32_t m[3] = {L[0] / 2, L[1] / 2, L[2]};
    return gen_r_host(h_out, L, m, xRange, corrL, corrMod, seed, rank, h_phi, flags);
}

int scarl_hip_gen_std_gauss_fft_dev(double* d_out, const int32_t Np[3], const double xRange[3],
                                    const double corrL[3], int32_t corrMod, int32_t seed, int32_t rank,
                                    const double* d_phi, int32_t flags) {
    SCARL_TRY(ensure_init());
    if (!d_out || !Np || !xRange || !corrL) return fail(SCARL_E_INVALID, "null argument");
    const int32_t m[3] = {Np[0] / 2, Np[1] / 2, Np[2] / 2};
    const uint64_t key = (uint64_t)(int64_t)((int64_t)seed + rank);
    return mode_r_run(d_out, Np, m, xRange, corrL, corrMod, key, d_phi, flags);
}

int scarl_hip_gen_std_gauss_fft_dev_noalloc(double* d_out, const int32_t Np[3], const double xRange[3],
                                            const double corrL[3], int32_t corrMod, int32_t seed, int32_t rank,
                                            const double* d_phi, int32_t flags) {
    SCARL_TRY(ensure_init());
    if (!d_out || !Np || !xRange || !corrL) return fail(SCARL_E_INVALID, "null argument");
    const int32_t m[3] = {Np[0] / 2, Np[1] / 2, Np[2] / 2};
    const uint64_t key = (uint64_t)(int64_t)((int64_t)seed + rank);
    return mode_r_run_noalloc(d_out, Np, m, xRange, corrL, corrMod, key, d_phi, flags);
}

// ---- normal_ScaRL ------------------------------------------------------------------------
int scarl_hip_r8vec_uniform_01(int32_t n, int32_t* seed, double* h_r) {
    SCARL_TRY(ensure_init());
    if (!seed || (n > 0 && !h_r)) return fail(SCARL_E_INVALID, "null argument");
    if (n <= 0) return 0;
    void *d_r, *d_s;
    SCARL_TRY(scratch_get("lcg_r", (size_t)n * 8, &d_r));
    SCARL_TRY(scratch_get("lcg_s", (size_t)n * 4, &d_s));
    SCARL_TRY(timed("lcg_uniform", [&] {
        return launch_lcg_uniform(*seed, n, (double*)d_r, (int32_t*)d_s, g_ctx.stream);
    }));
    int32_t last = 0;
    SCARL_TRY(copy_to_host(h_r, d_r, (size_t)n * 8));
    SCARL_TRY(copy_to_host(&last, (int32_t*)d_s + (n - 1), 4));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    *seed = last;
    return 0;
}

int scarl_hip_r8_uniform_01(int32_t* seed, double* r) { return scarl_hip_r8vec_uniform_01(1, seed, r); }

int scarl_hip_r8vec_normal_01(int32_t n, int32_t* seed, double* h_x) {
    // Box-Muller on the reference's pair layout (SRC_F/normal_ScaRL.f90:1186-1254)
    if (!seed || (n > 0 && !h_x)) return fail(SCARL_E_INVALID, "null argument");
    if (n <= 0) return 0;
    const double r8_pi = 3.141592653589793;
    if (n == 1) {
        double r[2];
        SCARL_TRY(scarl_hip_r8vec_uniform_01(2, seed, r));
        if (r[0] == 0.0) return fail(SCARL_E_INVALID, "R8VEC_NORMAL_01 - Fatal error: R8_UNIFORM_01 returned a value of 0.");
        h_x[0] = sqrt(-2.0 * log(r[0])) * cos(2.0 * r8_pi * r[1]);
        return 0;
    }
    const int32_t npair = n / 2, m = (n % 2 == 0) ? npair : npair + 1;
    std::vector<double> r((size_t)2 * m);
    SCARL_TRY(scarl_hip_r8vec_uniform_01(2 * m, seed, r.data()));
    for (int32_t p = 0; p < npair; ++p) {
        const double a = sqrt(-2.0 * log(r[2 * p]));
        h_x[2 * p] = a * cos(2.0 * r8_pi * r[2 * p + 1]);
        h_x[2 * p + 1] = a * sin(2.0 * r8_pi * r[2 * p + 1]);
    }
    if (n % 2 == 1) h_x[n - 1] = sqrt(-2.0 * log(r[2 * m - 2])) * cos(2.0 * r8_pi * r[2 * m - 1]);
    return 0;
}

// ---- main_ScaRL element routines ---------------------------------------------------------
int scarl_hip_apply_unity_partition(double* h_field, const int32_t Np[3], const int32_t Np_ovlp[3],
                                    int32_t partition_type, const int32_t topo_pos[3], const int32_t topo_shape[3]) {
    SCARL_TRY(ensure_init());
    if (!h_field || !Np || !Np_ovlp || !topo_pos || !topo_shape) return fail(SCARL_E_INVALID, "null argument");
    if (partition_type != 1) return fail(SCARL_E_UNSUPPORTED, "partitionType %d (only 1, cosine, exists)", partition_type);
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        if (Np[a] < 1) return fail(SCARL_E_INVALID, "Np(%d) must be >= 1", a + 1);
        if (Np_ovlp[a] > Np[a]) return fail(SCARL_E_INVALID, "Np_ovlp(%d) > Np(%d)", a + 1, a + 1);
        // a neighbour exists on a side iff topo_pos -+ 1 stays inside the topology (:1257-1260)
        lo[a] = (topo_pos[a] - 1 >= 0 && Np_ovlp[a] >= 2) ? Np_ovlp[a] : 0;
        hi[a] = (topo_pos[a] + 1 <= topo_shape[a] - 1 && Np_ovlp[a] >= 2) ? Np_ovlp[a] : 0;
    }
    const int64_t N = (int64_t)Np[0] * Np[1] * Np[2];
    void* d;
    SCARL_TRY(scratch_get("pou", (size_t)N * 8, &d));
    SCARL_TRY(copy_to_device(d, h_field, (size_t)N * 8));
    SCARL_TRY(timed("pou", [&] {
        return launch_pou((double*)d, Np[0], Np[1], Np[2], make_int3(lo[0], lo[1], lo[2]),
                          make_int3(hi[0], hi[1], hi[2]), g_ctx.stream);
    }));
    SCARL_TRY(copy_to_host(h_field, d, (size_t)N * 8));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    return 0;
}

int scarl_hip_multivariate_transformation(double avg, double std_dev, int32_t margiFirst, double* h_field, int64_t n) {
    SCARL_TRY(ensure_init());
    double mu, sd;
    if (margiFirst != FOM_GAUSSIAN && margiFirst != FOM_LOGNORMAL)
        return fail(SCARL_E_INVALID, "First-order marginal not detected (margiFirst = %d)", margiFirst);
    SCARL_TRY(marginal_params(margiFirst, avg, std_dev, &mu, &sd));
    if (n <= 0) return 0;
    if (!h_field) return fail(SCARL_E_INVALID, "null field");
    void* d;
    SCARL_TRY(scratch_get("marg", (size_t)n * 8, &d));
    SCARL_TRY(copy_to_device(d, h_field, (size_t)n * 8));
    SCARL_TRY(timed("marginal", [&] { return launch_marginal((double*)d, n, margiFirst, mu, sd, g_ctx.stream); }));
    SCARL_TRY(copy_to_host(h_field, d, (size_t)n * 8));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    return 0;
}

// ---- Stat moments ------------------------------------------------------------------------
int scarl_hip_moments(const double* h_rf, int64_t n, int32_t nmc, double* evnt_avg, double* evnt_std, double* g_avg,
                      double* g_std) {
    SCARL_TRY(ensure_init());
    if (n <= 0 || nmc <= 0) return fail(SCARL_E_INVALID, "n and nmc must be positive");
    if (!h_rf || !evnt_avg || !evnt_std || !g_avg || !g_std) return fail(SCARL_E_INVALID, "null argument");
    const int nblocks = (int)std::min<int64_t>(std::max<int64_t>(1, n / (kBlock * 16)), 1024);
    void *d, *dp;
    SCARL_TRY(scratch_get("mom_in", (size_t)(n * nmc) * 8, &d));
    SCARL_TRY(scratch_get("mom_part", (size_t)nblocks * nmc * sizeof(MomPartial), &dp));
    SCARL_TRY(copy_to_device(d, h_rf, (size_t)(n * nmc) * 8));
    SCARL_TRY(timed("sums", [&] { return launch_sums((const double*)d, n, nmc, nblocks, (MomPartial*)dp, g_ctx.stream); }));
    std::vector<MomPartial> h((size_t)nblocks * nmc);
    SCARL_TRY(copy_to_host(h.data(), dp, h.size() * sizeof(MomPartial)));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    auto nadd_h = [](double& s, double& c, double v) {
        const double t = s + v;
        c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
        s = t;
    };
    double ts = 0, tc = 0, ts2 = 0, tc2 = 0;
    for (int e = 0; e < nmc; ++e) {
        double s = 0, c = 0, s2 = 0, c2 = 0;
        for (int b = 0; b < nblocks; ++b) {
            const MomPartial& m = h[(size_t)e * nblocks + b];
            nadd_h(s, c, m.s);
            c += m.c;
            nadd_h(s2, c2, m.s2);
            c2 += m.c2;
        }
        const double S = s + c, S2 = s2 + c2;
        // SRC_F/Stat_ScaRL/statistics_RF.f90:80-87
        evnt_avg[e] = S / (double)n;
        evnt_std[e] = sqrt(S2 / (double)n - evnt_avg[e] * evnt_avg[e]);
        nadd_h(ts, tc, S);
        nadd_h(ts2, tc2, S2);
    }
    const double tot = (double)n * nmc;
    *g_avg = (ts + tc) / tot;
    *g_std = sqrt((ts2 + tc2) / tot - (*g_avg) * (*g_avg));
    return 0;
}

// ---- lag products: per-axis covariance and the corrL estimator ------------------------------
// The FFT length and plan of a periodic lag job (the corrL estimator): P = n; -1 when the direct
// products must be used (the switch is off, fewer than 8 lags, bounded pairs, a length only
// Bluestein could plan, or P > 4096).  Bounded pairs (the per-axis covariance, a few dozen lags)
// stay on the direct products: zero padding to P >= n + nlag - 1 made them slower than the products.
static int lag_fft_plan(const LagJob& J, int* P_out) {
    const int n = J.n[J.axis];
    if (!g_ctx.lag_fft || J.nlag < 8 || !J.circular) return -1;
    if (!g_lag_ps) {
        g_lag_ps.reset(new PlanSet());
        g_lag_ps->fp64 = true;
    }
    std::vector<int> rad;
    const int P = n;
    if (P < 2 || P > 4096 || !PlanSet::factor(P, kEcap64, rad)) return -1;
    const int id = g_lag_ps->get(P, LAYOUT_PF, 0, 3);
    if (id < 0 || g_lag_ps->plans[id].spec == kSpecNonSmooth) return -1;
    *P_out = P;
    return id;
}

// Runs every job in one pass per method (block tables over jobs) and writes [job][nlag_max] to d_out.
// Direct products: pencils per block the largest power of two (at most 32) whose (n + nlag) T doubles
// fit 64 KiB.  FFT: 2 T real pencils per block, T the plan's complex pencils per tile.
static int run_lag_jobs(std::vector<LagJob>& jobs, bool fp64, int nlag_max, double* d_out) {
    size_t lds = 0, flds = 0;
    int64_t tot_partial = 0;
    std::vector<LagBlock> blocks, fblocks;
    for (size_t j = 0; j < jobs.size(); ++j) {
        LagJob& J = jobs[j];
        const int n = J.n[J.axis];
        const int64_t npen = (int64_t)J.n[0] * J.n[1] * J.n[2] / n;
        if (npen > INT32_MAX) return fail(SCARL_E_UNSUPPORTED, "too many pencils (%lld)", (long long)npen);
        int P = 0;
        J.fft_plan = lag_fft_plan(J, &P);
        int T;
        if (J.fft_plan >= 0) {
            const PlanDev& pl = g_lag_ps->plans[J.fft_plan];
            J.P = P;
            T = 2 * pl.T;
            flds = std::max(flds, (size_t)pl.T * P * 16);
        } else {
            if ((size_t)(n + J.nlag) * 8 > 65536)
                return fail(SCARL_E_UNSUPPORTED, "a pencil of %d points with %d lags exceeds the 64 KiB tile", n, J.nlag);
            T = 32;
            while (T > 1 && (size_t)(n + J.nlag) * T * 8 > 65536) T >>= 1;
            lds = std::max(lds, (size_t)(n + J.nlag) * T * 8);
        }
        J.T = T;
        J.nblocks = (npen + T - 1) / T;
        for (int64_t p = 0; p < npen; p += T) (J.fft_plan >= 0 ? fblocks : blocks).push_back({(int)j, (int)p});
        tot_partial += J.nblocks * J.nlag;
    }
    if (g_lag_ps && g_lag_ps->plans.size() != g_lag_uploaded) {
        SCARL_TRY(g_lag_ps->upload());
        g_lag_uploaded = g_lag_ps->plans.size();
    }
    void *d_part, *d_jobs, *d_blocks, *d_fblocks;
    SCARL_TRY(scratch_get("lag_partial", (size_t)std::max<int64_t>(1, tot_partial) * 8, &d_part));
    int64_t off = 0;
    for (LagJob& J : jobs) {
        J.partial = static_cast<double*>(d_part) + off;
        off += J.nblocks * J.nlag;
    }
    SCARL_TRY(scratch_get("lag_jobs", jobs.size() * sizeof(LagJob), &d_jobs));
    SCARL_TRY(scratch_get("lag_blocks", std::max<size_t>(1, blocks.size()) * sizeof(LagBlock), &d_blocks));
    SCARL_TRY(scratch_get("lag_fblocks", std::max<size_t>(1, fblocks.size()) * sizeof(LagBlock), &d_fblocks));
    SCARL_TRY(copy_to_device(d_jobs, jobs.data(), jobs.size() * sizeof(LagJob)));
    if (!blocks.empty())
        HIP_TRY(hipMemcpyAsync(d_blocks, blocks.data(), blocks.size() * sizeof(LagBlock), hipMemcpyHostToDevice,
                               g_ctx.stream));
    if (!fblocks.empty())
        HIP_TRY(hipMemcpyAsync(d_fblocks, fblocks.data(), fblocks.size() * sizeof(LagBlock), hipMemcpyHostToDevice,
                               g_ctx.stream));
    SCARL_TRY(check_lds(flds));
    SCARL_TRY(timed("lag", [&] {
        return launch_lag(fp64, (const LagJob*)d_jobs, (int)jobs.size(), (const LagBlock*)d_blocks,
                          (int64_t)blocks.size(), lds, (const LagBlock*)d_fblocks, (int64_t)fblocks.size(), flds,
                          g_lag_ps ? g_lag_ps->d_plans : nullptr, g_lag_ps ? g_lag_ps->d_pool : nullptr, nlag_max,
                          d_out, g_ctx.stream);
    }));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));  // the host job and block tables die on return
    return 0;
}

static LagJob lag_job(const void* x, const int32_t n[3], int axis, int nlag, int circular, int window, double centre,
                      const double* d_centre, double scale) {
    LagJob J;
    memset(&J, 0, sizeof J);
    J.x = x;
    for (int a = 0; a < 3; ++a) J.n[a] = n[a];
    J.axis = axis;
    J.ld1 = n[0];
    J.ld2 = (int64_t)n[0] * n[1];
    J.nlag = nlag;
    J.circular = circular;
    J.window = window;
    J.centre = d_centre;
    J.cval = centre;
    J.scale = scale;
    return J;
}

// SRC_F/Stat_ScaRL/statistics_RF.f90:258-462 rebuild_Sk_FFT: Sk_part(l) = Re IFFT(mean over the
// pencils along dir of |FFT(u)/N|^2), u = (x - avg)/stdDev tapered along dir when nDim = 3, which is
// the periodic autocorrelation (1/N) sum_m u_m u_{(m+l) mod N} averaged over the pencils; computed
// directly in FP64 with compensated block combination.
int scarl_hip_rebuild_sk_fft(const double* h_field, const int32_t xNStep[3], int32_t dir, int32_t nDim, double avg,
                             double stdDev, double* h_sk_part) {
    using namespace scarl;
    SCARL_TRY(ensure_init());
    if (!h_field || !xNStep || !h_sk_part) return fail(SCARL_E_INVALID, "null argument");
    if (nDim < 1 || nDim > 3 || dir < 1 || dir > nDim) return fail(SCARL_E_INVALID, "bad nDim %d / dir %d", nDim, dir);
    int32_t n[3] = {1, 1, 1};
    for (int a = 0; a < nDim; ++a) {
        if (xNStep[a] < 1) return fail(SCARL_E_INVALID, "xNStep(%d) = %d", a + 1, xNStep[a]);
        n[a] = xNStep[a];
    }
    if (!(stdDev != 0.0)) return fail(SCARL_E_INVALID, "stdDev must be nonzero");
    const int64_t tot = (int64_t)n[0] * n[1] * n[2];
    const int nd = n[dir - 1];
    void *d_x, *d_out;
    SCARL_TRY(scratch_get("stat_field", (size_t)tot * 8, &d_x));
    SCARL_TRY(scratch_get("stat_sk", (size_t)nd * 8, &d_out));
    SCARL_TRY(copy_to_device(d_x, h_field, (size_t)tot * 8));
    std::vector<LagJob> jobs(1, lag_job(d_x, n, dir - 1, nd, 1, nDim == 3, avg, nullptr, 1.0 / stdDev));
    SCARL_TRY(run_lag_jobs(jobs, true, nd, (double*)d_out));
    HIP_TRY(hipMemcpy(h_sk_part, d_out, (size_t)nd * 8, hipMemcpyDeviceToHost));
    return 0;
}

// SRC_F/Stat_ScaRL/statistics_RF.f90:468-486 rebuild_corrL: corrL(d) = 2 (sum(Sk_d) - (first + last)/2)
// xStep(d) over the nhalf(d) kept lags of each direction (host arithmetic).
int scarl_hip_rebuild_corrl(const double* sk_tot, const int32_t nhalf[3], int32_t nDim, const double xStep[3],
                            double* corrL) {
    if (!sk_tot || !nhalf || !xStep || !corrL) return scarl::fail(SCARL_E_INVALID, "null argument");
    if (nDim < 1 || nDim > 3) return scarl::fail(SCARL_E_INVALID, "bad nDim %d", nDim);
    int64_t off = 0;
    for (int d = 0; d < nDim; ++d) {
        if (nhalf[d] < 1) return scarl::fail(SCARL_E_INVALID, "nhalf(%d) = %d", d + 1, nhalf[d]);
        double s = 0.0;
        for (int l = 0; l < nhalf[d]; ++l) s += sk_tot[off + l];
        corrL[d] = 2.0 * (s - (sk_tot[off] + sk_tot[off + nhalf[d] - 1]) / 2.0) * xStep[d];
        off += nhalf[d];
    }
    return 0;
}

// Per-axis centred covariance of one field at lags 0..nlag-1: cov[a][l] = sum over the pairs
// (p, p + l e_a) inside the box of (x_p - mean)(x_{p+l} - mean) / (number of pairs); NaN for l >= n_a.
int scarl_hip_lag_covariance(const double* h_field, const int32_t n[3], double mean, int32_t nlag, double* h_cov) {
    using namespace scarl;
    SCARL_TRY(ensure_init());
    if (!h_field || !n || !h_cov) return fail(SCARL_E_INVALID, "null argument");
    if (nlag < 1) return fail(SCARL_E_INVALID, "nlag must be >= 1");
    for (int a = 0; a < 3; ++a)
        if (n[a] < 1) return fail(SCARL_E_INVALID, "n[%d] = %d", a, n[a]);
    const int64_t tot = (int64_t)n[0] * n[1] * n[2];
    void *d_x, *d_out;
    SCARL_TRY(scratch_get("stat_field", (size_t)tot * 8, &d_x));
    SCARL_TRY(scratch_get("stat_cov", (size_t)3 * nlag * 8, &d_out));
    SCARL_TRY(copy_to_device(d_x, h_field, (size_t)tot * 8));
    std::vector<LagJob> jobs;
    for (int a = 0; a < 3; ++a) jobs.push_back(lag_job(d_x, n, a, nlag, 0, 0, mean, nullptr, 1.0));
    SCARL_TRY(run_lag_jobs(jobs, true, nlag, (double*)d_out));
    HIP_TRY(hipMemcpy(h_cov, d_out, (size_t)3 * nlag * 8, hipMemcpyDeviceToHost));
    return 0;
}

// Every field of an ensemble batch (after scarl_hip_batch_run): per-axis covariance at lags
// 0..nlag-1 centred on the field's own sample mean (read on the device from the batch moments).
// h_cov: [field][axis][lag].
int scarl_hip_batch_lag_cov(scarl_batch* b, int32_t nlag, double* h_cov) {
    using namespace scarl;
    SCARL_TRY(ensure_init());
    if (!b || !h_cov) return fail(SCARL_E_INVALID, "null argument");
    if (nlag < 1) return fail(SCARL_E_INVALID, "nlag must be >= 1");
    if (b->global)
        return fail(SCARL_E_UNSUPPORTED, "a blended batch has one global field: use scarl_hip_lag_covariance on it");
    std::vector<LagJob> jobs;
    for (size_t i = 0; i < b->fields.size(); ++i) {
        const FieldH& f = b->fields[i].f;
        const double* mean = b->d_slot_out + 5 * f.mom_slot + 3;
        for (int a = 0; a < 3; ++a) jobs.push_back(lag_job(f.out, f.n, a, nlag, 0, 0, 0.0, mean, 1.0));
    }
    void* d_out;
    SCARL_TRY(scratch_get("batch_cov", jobs.size() * (size_t)nlag * 8, &d_out));
    SCARL_TRY(run_lag_jobs(jobs, b->fp64, nlag, (double*)d_out));
    HIP_TRY(hipMemcpy(h_cov, d_out, jobs.size() * (size_t)nlag * 8, hipMemcpyDeviceToHost));
    return 0;
}

// SRC_F/main_ScaRL.f90:903-953 normalize_field for one rank (communicator of size one): the mean
// over the non-duplicated box 1:maxPos (maxPos = Np - Np_ovlp on sides that are not last), divided
// by xNTotal = prod(topo_shape (Np - Np_ovlp) + Np_ovlp); then the centred sum of squares over the
// same box; x <- (x - avg) / std_dev over the whole field.  Compensated FP64 sums.
int scarl_hip_normalize_field(double* h_field, const int32_t Np[3], const int32_t Np_ovlp[3], const int32_t topo_pos[3],
                              const int32_t topo_shape[3], double* avg_out, double* std_out) {
    using namespace scarl;
    SCARL_TRY(ensure_init());
    if (!h_field || !Np || !Np_ovlp || !topo_pos || !topo_shape) return fail(SCARL_E_INVALID, "null argument");
    int n[3], box[3];
    double xnt = 1.0;
    for (int a = 0; a < 3; ++a) {
        if (Np[a] < 1 || Np_ovlp[a] < 0 || Np_ovlp[a] >= Np[a] || topo_shape[a] < 1 || topo_pos[a] < 0 ||
            topo_pos[a] >= topo_shape[a])
            return fail(SCARL_E_INVALID, "bad geometry on axis %d", a);
        if (topo_shape[a] != 1)
            return fail(SCARL_E_UNSUPPORTED,
                        "normalize_field over %d ranks needs the global sums of every rank; one GPU holds one rank",
                        topo_shape[0] * topo_shape[1] * topo_shape[2]);
        n[a] = Np[a];
        box[a] = (topo_pos[a] != topo_shape[a] - 1) ? Np[a] - Np_ovlp[a] : Np[a];
        xnt *= (double)topo_shape[a] * (Np[a] - Np_ovlp[a]) + Np_ovlp[a];
    }
    const int64_t tot = (int64_t)n[0] * n[1] * n[2];
    const int nblocks = (int)std::min<int64_t>(std::max<int64_t>(1, tot / (kBlock * 16)), 1024);
    void *d_x, *d_part;
    SCARL_TRY(scratch_get("norm_field", (size_t)tot * 8, &d_x));
    SCARL_TRY(scratch_get("norm_part", (size_t)nblocks * sizeof(MomPartial), &d_part));
    SCARL_TRY(copy_to_device(d_x, h_field, (size_t)tot * 8));
    std::vector<MomPartial> h(nblocks);
    auto box_sum = [&](double centre, bool squares, double* out) -> int {
        SCARL_TRY(timed("box_sums", [&] {
            return launch_box_sums((const double*)d_x, n, box, centre, (MomPartial*)d_part, nblocks, g_ctx.stream);
        }));
        SCARL_TRY(copy_to_host(h.data(), d_part, h.size() * sizeof(MomPartial)));
        HIP_TRY(hipStreamSynchronize(g_ctx.stream));
        double s = 0.0, c = 0.0;
        for (const MomPartial& m : h) {
            const double v = squares ? m.s2 : m.s;
            const double t = s + v;
            c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
            s = t;
            c += squares ? m.c2 : m.c;
        }
        *out = s + c;
        return 0;
    };
    double S, S2;
    SCARL_TRY(box_sum(0.0, false, &S));
    const double avg = S / xnt;
    SCARL_TRY(box_sum(avg, true, &S2));
    const double sd = sqrt(S2 / xnt);
    if (!(sd > 0.0)) return fail(SCARL_E_INVALID, "the field is constant over the box: std_dev = 0");
    SCARL_TRY(timed("normalize", [&] { return launch_normalize((double*)d_x, tot, avg, sd, g_ctx.stream); }));
    SCARL_TRY(copy_to_host(h_field, d_x, (size_t)tot * 8));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    if (avg_out) *avg_out = avg;
    if (std_out) *std_out = sd;
    return 0;
}

// ---- special functions -------------------------------------------------------------------
static int run_fp64_map(int which, int32_t model, double nu, const double* h_x, int32_t n, double* h_out) {
    SCARL_TRY(ensure_init());
    if (n <= 0) return 0;
    if (!h_x || !h_out) return fail(SCARL_E_INVALID, "null argument");
    void *dx, *dy;
    SCARL_TRY(scratch_get("map_x", (size_t)n * 8, &dx));
    SCARL_TRY(scratch_get("map_y", (size_t)n * 8, &dy));
    SCARL_TRY(copy_to_device(dx, h_x, (size_t)n * 8));
    if (which == 0)
        SCARL_TRY(timed("kv", [&] { return launch_kv(nu, (const double*)dx, n, (double*)dy, g_ctx.stream); }));
    else
        SCARL_TRY(timed("target_corr", [&] {
            return launch_target_corr(model, nu, (const double*)dx, n, (double*)dy, g_ctx.stream);
        }));
    SCARL_TRY(copy_to_host(h_out, dy, (size_t)n * 8));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    return 0;
}

int scarl_hip_kv(double nu, const double* h_x, int32_t n, double* h_out) {
    if (!(nu >= 0.0)) return fail(SCARL_E_INVALID, "nu must be >= 0");
    return run_fp64_map(0, 0, nu, h_x, n, h_out);
}

int scarl_hip_target_corr(int32_t model, double nu, const double* h_rho, int32_t n, double* h_out) {
    if (model != CM_GAUSSIAN && model != CM_VON_KARMAN && model != CM_EXPONENTIAL)
        return fail(SCARL_E_INVALID, "unknown correlation model %d", model);
    if (model == CM_VON_KARMAN && !(nu > 0.0)) return fail(SCARL_E_INVALID, "von Karman needs nu > 0");
    return run_fp64_map(1, model, nu, h_rho, n, h_out);
}

// ---- batches -----------------------------------------------------------------------------
int scarl_hip_batch_create(const scarl_field_desc* fields, int32_t count, int32_t precision, scarl_batch** h) {
    SCARL_TRY(ensure_init());
    if (!fields || !h || count <= 0) return fail(SCARL_E_INVALID, "need at least one field descriptor");
    if (precision != SCARL_PREC_FP32 && precision != SCARL_PREC_FP64)
        return fail(SCARL_E_INVALID, "precision must be 32 or 64");
    *h = nullptr;
    std::unique_ptr<scarl_batch> b(new scarl_batch());
    b->fp64 = precision == SCARL_PREC_FP64;
    b->ps.fp64 = b->fp64;
    b->ps.ecap = b->fp64 ? kEcap64 : g_ctx.fp32_ecap;
    b->ps.mfma = g_ctx.mfma_dft16;
    b->fields.resize(count);
    std::vector<double> slot_count(count);
    for (int i = 0; i < count; ++i) {
        SCARL_TRY(validate_desc(fields[i]));
        b->fields[i].d = fields[i];
        SCARL_TRY(build_field(b.get(), b->fields[i]));
        b->fields[i].f.mom_slot = i;
        b->fields[i].f.color = 0;
        b->fields[i].f.blend = 0;
        slot_count[i] = (double)fields[i].n[0] * fields[i].n[1] * fields[i].n[2];
    }
    SCARL_TRY(finish_amplitudes(b.get()));
    SCARL_TRY(finalize_batch(b.get(), slot_count));
    if (g_ctx.autotune) SCARL_TRY(autotune_batch(b.get()));
    *h = b.release();
    return 0;
}

int scarl_hip_blend_create(const int32_t gdim[3], const int32_t nb[3], const int32_t* starts, const int32_t* exts,
                           const scarl_field_desc* params, int32_t precision, scarl_batch** h) {
    SCARL_TRY(ensure_init());
    if (!gdim || !nb || !starts || !exts || !params || !h) return fail(SCARL_E_INVALID, "null argument");
    if (precision != SCARL_PREC_FP32 && precision != SCARL_PREC_FP64)
        return fail(SCARL_E_INVALID, "precision must be 32 or 64");
    *h = nullptr;
    // validate the tensor-product tiling: positions cover [0, gdim) with overlaps only between
    // neighbours; overlaps of 1 point are rejected (no ramp exists, SRC_F/main_ScaRL.f90:1255)
    std::vector<int> st[3], ex[3], lo[3], hi[3];
    int off = 0;
    for (int a = 0; a < 3; ++a) {
        if (nb[a] < 1 || gdim[a] < 1) return fail(SCARL_E_INVALID, "bad grid on axis %d", a);
        for (int p = 0; p < nb[a]; ++p) {
            st[a].push_back(starts[off + p]);
            ex[a].push_back(exts[off + p]);
        }
        off += nb[a];
        lo[a].assign(nb[a], 0);
        hi[a].assign(nb[a], 0);
        if (st[a][0] != 0) return fail(SCARL_E_INVALID, "axis %d: first subdomain must start at 0", a);
        if (st[a].back() + ex[a].back() != gdim[a])
            return fail(SCARL_E_INVALID, "axis %d: last subdomain must end at %d", a, gdim[a]);
        for (int p = 0; p < nb[a]; ++p) {
            if (ex[a][p] < 1) return fail(SCARL_E_INVALID, "axis %d: empty subdomain", a);
            if (p > 0) {
                const int ov = st[a][p - 1] + ex[a][p - 1] - st[a][p];
                if (ov < 0) return fail(SCARL_E_INVALID, "axis %d: gap between subdomains %d and %d", a, p - 1, p);
                if (ov == 1) return fail(SCARL_E_UNSUPPORTED, "axis %d: overlap of one point has no PoU ramp", a);
                if (st[a][p] <= st[a][p - 1]) return fail(SCARL_E_INVALID, "axis %d: starts must increase", a);
                if (p > 1 && st[a][p] < st[a][p - 2] + ex[a][p - 2])
                    return fail(SCARL_E_INVALID, "axis %d: overlap reaches beyond the neighbour", a);
                if (ov > ex[a][p] || ov > ex[a][p - 1])
                    return fail(SCARL_E_INVALID, "axis %d: overlap %d larger than a subdomain", a, ov);
                lo[a][p] = ov;
                hi[a][p - 1] = ov;
            }
        }
        // the two ramps of a subdomain must not meet (the reference's expansion rule,
        // SRC_F/main_ScaRL.f90:161-162, guarantees it for its uniform tilings)
        for (int p = 0; p < nb[a]; ++p)
            if (lo[a][p] + hi[a][p] > ex[a][p])
                return fail(SCARL_E_INVALID, "axis %d: the overlaps of subdomain %d exceed its extent", a, p);
    }
    std::unique_ptr<scarl_batch> b(new scarl_batch());
    b->fp64 = precision == SCARL_PREC_FP64;
    b->ps.fp64 = b->fp64;
    b->ps.ecap = b->fp64 ? kEcap64 : g_ctx.fp32_ecap;
    b->ps.mfma = g_ctx.mfma_dft16;
    b->global = true;
    for (int a = 0; a < 3; ++a) b->gdim[a] = gdim[a];
    const int count = nb[0] * nb[1] * nb[2];
    b->fields.resize(count);
    int id = 0;
    for (int k = 0; k < nb[2]; ++k)
        for (int j = 0; j < nb[1]; ++j)
            for (int i = 0; i < nb[0]; ++i, ++id) {
                scarl_field_desc d = *params;
                d.n[0] = ex[0][i];
                d.n[1] = ex[1][j];
                d.n[2] = ex[2][k];
                d.stream = (uint32_t)id;
                SCARL_TRY(validate_desc(d));
                FieldHost& fh = b->fields[id];
                fh.d = d;
                SCARL_TRY(build_field(b.get(), fh));
                FieldH& f = fh.f;
                f.blend = 1;
                f.color = (i & 1) | ((j & 1) << 1) | ((k & 1) << 2);
                const int pos[3] = {i, j, k};
                for (int a = 0; a < 3; ++a) {
                    f.nov_lo[a] = lo[a][pos[a]];
                    f.nov_hi[a] = hi[a][pos[a]];
                }
                f.mom_slot = 0;
            }
    SCARL_TRY(finish_amplitudes(b.get()));
    std::vector<double> slot_count(1, (double)gdim[0] * gdim[1] * gdim[2]);
    SCARL_TRY(finalize_batch(b.get(), slot_count));
    // output placement inside the global field
    const size_t rsz = b->fp64 ? 8 : 4;
    id = 0;
    for (int k = 0; k < nb[2]; ++k)
        for (int j = 0; j < nb[1]; ++j)
            for (int i = 0; i < nb[0]; ++i, ++id) {
                FieldH& f = b->fields[id].f;
                const int64_t o = st[0][i] + (int64_t)gdim[0] * (st[1][j] + (int64_t)gdim[1] * st[2][k]);
                f.out = static_cast<char*>(b->d_out) + o * rsz;
                f.out_ld1 = gdim[0];
                f.out_ld2 = (int64_t)gdim[0] * gdim[1];
            }
    SCARL_TRY(upload_fields(b.get()));
    if (g_ctx.autotune) SCARL_TRY(autotune_batch(b.get()));
    *h = b.release();
    return 0;
}

int scarl_hip_batch_set_noise(scarl_batch* b, int32_t field, const void* h_noise) {
    SCARL_TRY(ensure_init());
    if (!b || field < 0 || field >= (int)b->fields.size()) return fail(SCARL_E_INVALID, "bad batch or field index");
    FieldHost& fh = b->fields[field];
    const size_t bytes = (size_t)fh.f.h1 * fh.f.n[1] * fh.f.n[2] * 2 * (b->fp64 ? 8 : 4);
    if (h_noise) {
        if (!fh.noise_buf) SCARL_TRY(dmalloc(&fh.noise_buf, bytes));
        SCARL_TRY(copy_to_device(fh.noise_buf, h_noise, bytes));
        fh.f.W = fh.noise_buf;
        fh.f.src_mode = SRC_HOST_NOISE;
    } else {
        fh.f.W = nullptr;
        fh.f.src_mode = SRC_DEVICE_NOISE;
    }
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    SCARL_TRY(upload_fields(b));
    return 0;
}

int scarl_hip_batch_run(scarl_batch* b, int32_t use_graph) {
    SCARL_TRY(ensure_init());
    if (!b) return fail(SCARL_E_INVALID, "null batch");
    if (!use_graph || g_ctx.timing) return enqueue_batch(b);
    SCARL_TRY(ensure_graph(b));
    HIP_TRY(hipGraphLaunch(b->gexec, g_ctx.stream));
    return 0;
}

int scarl_hip_batch_sync(scarl_batch* b) {
    SCARL_TRY(ensure_init());
    (void)b;
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    return 0;
}

int scarl_hip_batch_output(scarl_batch* b, int32_t field, void** d_ptr, int64_t dims[3]) {
    if (!b || !d_ptr || !dims) return fail(SCARL_E_INVALID, "null argument");
    if (field < 0) {
        if (!b->global) return fail(SCARL_E_INVALID, "batch has no global field");
        *d_ptr = b->d_out;
        for (int a = 0; a < 3; ++a) dims[a] = b->gdim[a];
        return 0;
    }
    if (field >= (int)b->fields.size()) return fail(SCARL_E_INVALID, "field index out of range");
    if (b->global) return fail(SCARL_E_INVALID, "blended subdomains are only available as the global field");
    *d_ptr = b->fields[field].f.out;
    for (int a = 0; a < 3; ++a) dims[a] = b->fields[field].f.n[a];
    return 0;
}

int scarl_hip_batch_copy_output(scarl_batch* b, int32_t field, void* h_dst, int64_t nbytes) {
    SCARL_TRY(ensure_init());
    void* p;
    int64_t dims[3];
    SCARL_TRY(scarl_hip_batch_output(b, field, &p, dims));
    const int64_t need = dims[0] * dims[1] * dims[2] * (b->fp64 ? 8 : 4);
    if (!h_dst || nbytes < need) return fail(SCARL_E_INVALID, "destination holds %lld bytes, need %lld",
                                             (long long)nbytes, (long long)need);
    SCARL_TRY(copy_to_host(h_dst, p, (size_t)need));
    return 0;
}

int scarl_hip_batch_moments(scarl_batch* b, double* h_out, int32_t max_slots, int32_t* count) {
    SCARL_TRY(ensure_init());
    if (!b || !h_out || !count) return fail(SCARL_E_INVALID, "null argument");
    const int n = std::min(max_slots, b->nslots);
    SCARL_TRY(copy_to_host(h_out, b->d_slot_out, (size_t)n * 5 * sizeof(double)));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    *count = b->nslots;
    return 0;
}

int scarl_hip_batch_field_info(scarl_batch* b, int32_t field, double* amp0, double* spectrum_sum) {
    if (!b || field < 0 || field >= (int)b->fields.size()) return fail(SCARL_E_INVALID, "bad batch or field index");
    if (amp0) *amp0 = b->fields[field].f.amp0;
    if (spectrum_sum) *spectrum_sum = b->fields[field].spec_sum;
    return 0;
}

int scarl_hip_batch_bytes(scarl_batch* b, int64_t* bytes) {
    if (!b || !bytes) return fail(SCARL_E_INVALID, "null argument");
    *bytes = b->bytes;
    return 0;
}

int scarl_hip_batch_plan_report(scarl_batch* b, char* buf, int32_t len) {
    if (!b || !buf || len <= 0) return fail(SCARL_E_INVALID, "null argument");
    std::string s;
    for (int i = 0; i < (int)b->ps.plans.size(); ++i) {
        s += plan_line(b->ps, i);
        if (i < (int)b->tune_notes.size()) s += b->tune_notes[i];
        s += b->groups.size() > 1 ? (b->pipeline ? " schedule=pipelined" : " schedule=sequential") : "";
        s += "\n";
    }
    snprintf(buf, (size_t)len, "%s", s.c_str());
    return 0;
}

int scarl_hip_batch_destroy(scarl_batch* b) {
    if (!b) return 0;
    if (g_ctx.ready) hipStreamSynchronize(g_ctx.stream);
    delete b;
    return 0;
}

// ---- spectra_ScaRL -----------------------------------------------------------------------
// sz_Sk = ceiling(k_cut/delta_k) with the k_cut fix-up (SRC_F/spectra_ScaRL.f90:42-72)
static int sk_shape(int32_t corrMod, const double xRange[3], const int32_t Np[3], const int32_t sz_c[3], int32_t sz[3],
                    double dk[3]) {
    if (corrMod != CM_GAUSSIAN && corrMod != CM_VON_KARMAN)
        return fail(SCARL_E_INVALID, "build_Sk_mtx: unknown corrMod %d", corrMod);
    double k_cut[3], k_max[3];
    bool any_lt = false;
    for (int d = 0; d < 3; ++d) {
        if (Np[d] < 2 || sz_c[d] < 2) return fail(SCARL_E_INVALID, "build_Sk_mtx: Np and sz_c must be >= 2");
        k_cut[d] = 7.355;
        const double delta_x = xRange[d] / (double)(Np[d] - 1);
        k_max[d] = 1.0 / delta_x;
        dk[d] = k_max[d] / (double)(sz_c[d] / 2);
        if (k_max[d] < k_cut[d]) any_lt = true;
    }
    if (any_lt)
        for (int d = 0; d < 3; ++d)
            if (k_max[d] > k_cut[d]) k_cut[d] = k_max[d];
    for (int d = 0; d < 3; ++d) sz[d] = (int32_t)ceil(k_cut[d] / dk[d]);
    return 0;
}

int scarl_hip_build_sk_query(int32_t corrMod, const double xRange[3], const double corrL[3], const int32_t Np[3],
                             const int32_t sz_c[3], int32_t seed, int32_t rank, int32_t out_shape[3]) {
    (void)corrL; (void)seed; (void)rank;
    if (!xRange || !Np || !sz_c || !out_shape) return fail(SCARL_E_INVALID, "null argument");
    int32_t sz[3];
    double dk[3];
    SCARL_TRY(sk_shape(corrMod, xRange, Np, sz_c, sz, dk));
    out_shape[0] = sz[0];
    out_shape[1] = 2 * sz[1];
    out_shape[2] = 2 * sz[2];
    return 0;
}

static int build_sk_device(int32_t corrMod, const double xRange[3], const int32_t Np[3], const int32_t sz_c[3],
                           int32_t seed, void** d_sk, int32_t shape[3]) {
    int32_t sz[3];
    double dk[3];
    SCARL_TRY(sk_shape(corrMod, xRange, Np, sz_c, sz, dk));
    shape[0] = sz[0];
    shape[1] = 2 * sz[1];
    shape[2] = 2 * sz[2];
    const int64_t tot = (int64_t)shape[0] * shape[1] * shape[2];
    if (tot > (int64_t)0x7fffffff) return fail(SCARL_E_UNSUPPORTED, "Sk_mtx larger than the int32 LCG stream");
    SCARL_TRY(scratch_get("sk", (size_t)std::max<int64_t>(1, tot) * 16, d_sk));
    SkParamsHost P;
    P.s1 = sz[0];
    P.sz2 = sz[1];
    P.sz3 = sz[2];
    P.gaussian = corrMod == CM_GAUSSIAN;
    P.seed = seed + 1000000;  // spectra_ScaRL.f90:78
    for (int d = 0; d < 3; ++d) P.de[d] = ((kPiRefHost * kPiRefHost) / 2.0) * (dk[d] * dk[d]);
    P.scale = 2.0 * sqrt(dk[0] * dk[1] * dk[2]);
    SCARL_TRY(timed("build_sk", [&] { return launch_build_sk(&P, *d_sk, tot, g_ctx.stream); }));
    return 0;
}

int scarl_hip_build_sk(int32_t corrMod, const double xRange[3], const double corrL[3], const int32_t Np[3],
                       const int32_t sz_c[3], int32_t seed, int32_t rank, double* h_sk) {
    (void)corrL; (void)rank;
    SCARL_TRY(ensure_init());
    if (!xRange || !Np || !sz_c || !h_sk) return fail(SCARL_E_INVALID, "null argument");
    void* d_sk;
    int32_t shape[3];
    SCARL_TRY(build_sk_device(corrMod, xRange, Np, sz_c, seed, &d_sk, shape));
    const size_t bytes = (size_t)shape[0] * shape[1] * shape[2] * 16;
    SCARL_TRY(copy_to_host(h_sk, d_sk, bytes));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    return 0;
}

// gen_std_gauss_Shino_FFT (SRC_F/fftw_ScaRL.f90:157-255, not called by the reference driver)
int scarl_hip_gen_std_gauss_shino_fft(double* h_out, const int32_t Np[3], const double xRange_in[3],
                                      const double corrL[3], const int32_t pointsPerCorrL[3], int32_t corrMod,
                                      int32_t seed, int32_t rank) {
    (void)pointsPerCorrL; (void)rank;
    SCARL_TRY(ensure_init());
    if (!h_out || !Np || !xRange_in || !corrL) return fail(SCARL_E_INVALID, "null argument");
    double xr[3];
    int32_t szc[3];
    for (int d = 0; d < 3; ++d) {
        if (!(corrL[d] > 0.0)) return fail(SCARL_E_INVALID, "corrL must be positive");
        xr[d] = xRange_in[d] / corrL[d];
        szc[d] = (int32_t)pow(2.0, ceil(log((double)(2 * Np[d])) / log(2.0)));
    }
    {
        // the quadrant copies (fftw_ScaRL.f90:217-233) must fit the padded grid without
        // overlapping; otherwise the reference writes out of bounds (k_max < k_cut regime)
        int32_t sz[3];
        double dk[3];
        SCARL_TRY(sk_shape(corrMod, xr, Np, szc, sz, dk));
        if (sz[0] / 2 > szc[0] || 2 * sz[1] > szc[1] || 2 * sz[2] > szc[2])
            return fail(SCARL_E_UNSUPPORTED,
                        "Sk quadrants (%d,%d,%d) exceed the padded grid (%d,%d,%d): mesh too coarse for this "
                        "representation", sz[0], 2 * sz[1], 2 * sz[2], szc[0], szc[1], szc[2]);
    }
    void* d_sk;
    int32_t shape[3];
    SCARL_TRY(build_sk_device(corrMod, xr, Np, szc, seed, &d_sk, shape));
    const int c1 = szc[0], c2 = szc[1], c3 = szc[2];
    const int64_t ctot = (int64_t)c1 * c2 * c3;
    void* d_c;
    SCARL_TRY(scratch_get("shino_c", (size_t)ctot * 16, &d_c));
    int dims9[9] = {shape[0], shape[1], shape[2], shape[0] / 2, shape[1] / 2, shape[2] / 2, c1, c2, c3};
    SCARL_TRY(timed("shino_place", [&] { return launch_shino_place(dims9, d_sk, d_c, g_ctx.stream); }));
    PlanSet ps;
    ps.fp64 = true;
    const int pz = ps.get(c3, LAYOUT_PF, 0), py = ps.get(c2, LAYOUT_PF, 0), px = ps.get(c1, LAYOUT_BF, 0);
    if (pz < 0 || py < 0 || px < 0) return SCARL_E_UNSUPPORTED;
    SCARL_TRY(ps.upload());
    FieldH f;
    memset(&f, 0, sizeof f);
    f.n[0] = c1; f.n[1] = c2; f.n[2] = c3;
    f.h1 = c1;
    f.wp = c1;
    f.wz = (int64_t)c1 * c2;
    f.nyq = 0;
    f.plan_z = pz; f.plan_y = py; f.plan_x = px;
    f.first_axis = -1;
    f.src_mode = SRC_WORKSPACE;
    f.ws = d_c;
    f.mom_slot = -1;
    std::vector<Tile> tz, ty;
    for (int kind : {2, 0}) {
        strided_tiles(0, f, c2, ps.plans[pz].T, kind, tz);
        strided_tiles(0, f, c3, ps.plans[py].T, kind, ty);
    }
    void *d_f = nullptr, *d_tz = nullptr, *d_ty = nullptr;
    std::vector<FieldH> hf(1, f);
    int rc = upload_vec(hf, &d_f);
    if (!rc) rc = upload_vec(tz, &d_tz);
    if (!rc) rc = upload_vec(ty, &d_ty);
    void* d_out = nullptr;
    if (!rc) rc = scratch_get("shino_out", (size_t)Np[0] * Np[1] * Np[2] * 8, &d_out);
    const size_t lz = (size_t)ps.plans[pz].lds_elems * 16, ly = (size_t)ps.plans[py].lds_elems * 16,
                 lx = (size_t)ps.plans[px].lds_elems * 16;
    if (!rc) rc = timed("shino_zpass", [&] {
        return launch_h_strided_f64(2, false, ps.plans[pz].spec, kEcap64, (FieldH*)d_f, (Tile*)d_tz, 0,
                                    (int)tz.size(), ps.d_plans, ps.d_pool, lz, g_ctx.stream);
    });
    if (!rc) rc = timed("shino_ypass", [&] {
        return launch_h_strided_f64(1, false, ps.plans[py].spec, kEcap64, (FieldH*)d_f, (Tile*)d_ty, 0,
                                    (int)ty.size(), ps.d_plans, ps.d_pool, ly, g_ctx.stream);
    });
    if (!rc) rc = timed("shino_xpass", [&] {
        return launch_c2c_xcrop(ps.d_plans, px, ps.plans[px].T, ps.d_pool, d_c, c1, c2, Np[0], Np[1], Np[2],
                                (double*)d_out, lx, g_ctx.stream);
    });
    if (!rc) {
        hipError_t e = hipMemcpyAsync(h_out, d_out, (size_t)Np[0] * Np[1] * Np[2] * 8, hipMemcpyDeviceToHost,
                                      g_ctx.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g_ctx.stream);
        if (e != hipSuccess) rc = fail(SCARL_E_HIP, "shino copy-out: %s", hipGetErrorString(e));
    } else {
        hipStreamSynchronize(g_ctx.stream);
    }
    if (d_f) hipFree(d_f);
    if (d_tz) hipFree(d_tz);
    if (d_ty) hipFree(d_ty);
    return rc;
}

int scarl_hip_gen_field_h(void* h_out, const scarl_field_desc* d, int32_t precision, const void* h_noise) {
    if (!h_out || !d) return fail(SCARL_E_INVALID, "null argument");
    scarl_batch* b = nullptr;
    SCARL_TRY(scarl_hip_batch_create(d, 1, precision, &b));
    std::unique_ptr<scarl_batch> guard(b);
    if (h_noise) SCARL_TRY(scarl_hip_batch_set_noise(b, 0, h_noise));
    SCARL_TRY(scarl_hip_batch_run(b, 0));
    const int64_t need = (int64_t)d->n[0] * d->n[1] * d->n[2] * (precision == SCARL_PREC_FP64 ? 8 : 4);
    SCARL_TRY(scarl_hip_batch_copy_output(b, 0, h_out, need));
    return 0;
}

}  // extern "C"
