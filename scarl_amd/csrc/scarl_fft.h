// scarl_fft.h -- one tile of pencils through a planned 1D backward DFT (device code).
//
// A tile is T pencils of one field handled by one 256-thread workgroup.  Stages run on
// register-resident butterflies; LDS holds the tile between stages (in place: every stage
// reads all its inputs, barriers, then writes).  Where the first stage can read its inputs
// straight from HBM (plain loads) and the last stage can write straight to HBM (plain stores),
// each pass costs exactly one HBM read and one HBM write.  Passes whose loader or epilogue is
// heavy (white-noise generation, blend) stage through LDS instead (SIN / SOUT), so that code is
// emitted once instead of once per unrolled element of every radix.
#pragma once
#include "scarl_device.h"
#include "scarl_internal.h"

#include <type_traits>

namespace scarl {

// reference PI literal (SRC_F/constants_ScaRL.f90:5)
constexpr double kPiRef = 3.1415926535898;

//   ctx = prep(p)          : per-thread context of tile pencil p (index decoding done once)
//   gld(ctx, n) -> cplx<T> : element n of the pencil (pass input)
//   gst(ctx, n, v)         : element n of the pencil (pass output)
//   gin(ctx, b, tpp, put)  : (SIN only) writes the pencil's N inputs through put(n, v); thread b of
//                            tpp threads; the default loops n = b, b + tpp, ... over gld
// SIN: load the whole tile into LDS first (one copy of the loader).  SOUT: finish the transform
// in LDS, then run gst from one loop.  EC: complex elements a thread holds per stage (the
// register budget; radices above EC are never planned for the instantiation).
struct DefaultStageIn {};

// NS: the plan is a length with a prime factor above 13 (Bluestein, or the direct O(N^2) DFT); those
// two paths live in their own instantiation so that they cost the Stockham kernels nothing.
template <typename T, int LAYOUT, bool SIN, bool SOUT, int EC, bool NS = false, bool COMP = false, class Prep,
          class GLoad, class GStore, class GIn = DefaultStageIn>
__device__ __forceinline__ void pencil_fft(const PlanDev& pl, const cplx<T>* __restrict__ pool, cplx<T>* lds,
                                           int npen, Prep&& prep, GLoad&& gld, GStore&& gst, GIn&& gin = GIn()) {
    const int tid = threadIdx.x;
    const int Tt = pl.T, tpp = pl.tpp, N = pl.N, ldp = pl.ldp, sk = pl.skew;
    int p, b;
    if constexpr (LAYOUT == LAYOUT_PF) {
        p = tid % Tt;
        b = tid / Tt;
    } else {
        b = tid % tpp;
        p = tid / tpp;
    }
    const bool active = (p < npen) && (p < Tt) && (b < tpp);
    auto addr = [&](int n) -> int {
        if constexpr (LAYOUT == LAYOUT_PF) return n * Tt + p;
        else return p * ldp + n + (n >> 4) * sk;
    };
    const auto ctx = prep(active ? p : 0);
    auto ld_g = [&](int n) -> cplx<T> { return gld(ctx, n); };
    auto st_g = [&](int n, cplx<T> v) { gst(ctx, n, v); };
    auto ld_l = [&](int n) -> cplx<T> { return lds[addr(n)]; };
    auto st_l = [&](int n, cplx<T> v) { lds[addr(n)] = v; };

    if constexpr (NS) {
    if (pl.blu_M > 0) {
        // Bluestein: X_k = c_k sum_j (x_j c_j) conj(c_{k-j}), c_j = exp(+i pi j^2/N), as the circular
        // convolution of length M >= 2N - 1: with a = (x c, zero padded) and E = DFT_M^-1(conj c)/M,
        // conv = conj(DFT_M^-1(conj(DFT_M^-1(a) E))); both transforms are the backward Stockham stages
        // of M, in place in the tile.
        const int M = pl.blu_M, S = pl.nst;
        const cplx<T>* ch = pool + pl.chirp_off;
        const cplx<T>* E = pool + pl.kern_off;
        auto put = [&](int n, cplx<T> v) { st_l(n, cmul(v, ch[n])); };
        if (active) {
            if constexpr (SIN && !std::is_same<typename std::decay<GIn>::type, DefaultStageIn>::value) {
                gin(ctx, b, tpp, put);
            } else {
#pragma unroll 1
                for (int n = b; n < N; n += tpp) put(n, ld_g(n));
            }
#pragma unroll 1
            for (int n = N + b; n < M; n += tpp) st_l(n, cplx<T>{T(0), T(0)});
        }
        __syncthreads();
        for (int pass = 0; pass < 2; ++pass) {
            for (int s = 0; s < S; ++s) {
                stage_dispatch<EC, T>(pl.R[s], M, pl.Ns[s], pool + pl.tw_off[s], b, tpp, active, true, ld_l, st_l);
                __syncthreads();
            }
            if (pass == 0) {
                if (active) {
#pragma unroll 1
                    for (int n = b; n < M; n += tpp) {
                        const cplx<T> v = cmul(ld_l(n), E[n]);
                        st_l(n, cplx<T>{v.x, -v.y});
                    }
                }
                __syncthreads();
            }
        }
        if (active) {
#pragma unroll 1
            for (int k = b; k < N; k += tpp) {
                const cplx<T> v = ld_l(k);
                st_g(k, cmul(cplx<T>{v.x, -v.y}, ch[k]));
            }
        }
        return;
    }
    {
        // direct O(N^2) DFT (short lengths with a prime factor above 13, and N = 1)
#pragma unroll 1
        for (int n = b; n < N; n += tpp)
            if (active) st_l(n, ld_g(n));
        __syncthreads();
        const cplx<T>* tw = pool + pl.direct_off;
#pragma unroll 1
        for (int k = b; k < N; k += tpp) {
            if (!active) break;
            cplx<T> acc = {T(0), T(0)};
            int t = 0;
#pragma unroll 1
            for (int j = 0; j < N; ++j) {
                acc = acc + cmul(ld_l(j), tw[t]);
                t += k;
                if (t >= N) t -= N;
            }
            st_g(k, acc);
        }
        return;
    }
    } else {
    if constexpr (SIN) {
        if (active) {
            if constexpr (std::is_same<typename std::decay<GIn>::type, DefaultStageIn>::value) {
#pragma unroll 1
                for (int n = b; n < N; n += tpp) st_l(n, ld_g(n));
            } else {
                gin(ctx, b, tpp, st_l);
            }
        }
        __syncthreads();
    }
    const int S = pl.nst;
    // stage 0
    if constexpr (SIN) {
        if (S == 1 && !SOUT) stage_dispatch<EC, T, COMP>(pl.R[0], N, 1, pool, b, tpp, active, true, ld_l, st_g);
        else stage_dispatch<EC, T, COMP>(pl.R[0], N, 1, pool, b, tpp, active, true, ld_l, st_l);
    } else {
        if (S == 1 && !SOUT) stage_dispatch<EC, T, COMP>(pl.R[0], N, 1, pool, b, tpp, active, false, ld_g, st_g);
        else stage_dispatch<EC, T, COMP>(pl.R[0], N, 1, pool, b, tpp, active, false, ld_g, st_l);
    }
    if (S > 1) {
        __syncthreads();
        for (int s = 1; s < S - 1; ++s) {
            stage_dispatch<EC, T, COMP>(pl.R[s], N, pl.Ns[s], pool + pl.tw_off[s], b, tpp, active, true, ld_l, st_l);
            __syncthreads();
        }
        const int s = S - 1;
        if constexpr (SOUT)
            stage_dispatch<EC, T, COMP>(pl.R[s], N, pl.Ns[s], pool + pl.tw_off[s], b, tpp, active, true, ld_l, st_l);
        else
            stage_dispatch<EC, T, COMP>(pl.R[s], N, pl.Ns[s], pool + pl.tw_off[s], b, tpp, active, false, ld_l, st_g);
    }
    if constexpr (SOUT) {
        __syncthreads();
#pragma unroll 1
        for (int n = b; n < N; n += tpp)
            if (active) st_g(n, ld_l(n));
    }
    }
}

// ------------------------------------------------------------------------------------------
// Compile-time planned variant: N = R0 * R1 (two stages) or N = R0 (one stage), TPP threads per
// pencil, 256 / TPP pencils per tile.  Strides, trip counts and index masks are constants, so the
// register budget is set by the element storage alone.
// ------------------------------------------------------------------------------------------
template <int R, int N, int NS, int TPP, typename T, class Load, class Store>
__device__ __forceinline__ void stage_fixed(const cplx<T>* __restrict__ tw, int b, bool active, bool sync, Load&& ld,
                                            Store&& st) {
    constexpr int nb = N / R;
    constexpr int BMAX = (nb + TPP - 1) / TPP;
    cplx<T> v[BMAX][R];
#pragma unroll
    for (int i = 0; i < BMAX; ++i) {
        const int j = b + TPP * i;
        if (active && (nb % TPP == 0 || j < nb)) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = ld(j + r * nb);
            if constexpr (NS > 1) {
                const int k = j % NS;
                const cplx<T>* t = tw + k * (R - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], t[r - 1]);
            }
            dft<R, T>(v[i]);
        }
    }
    if (sync) __syncthreads();
#pragma unroll
    for (int i = 0; i < BMAX; ++i) {
        const int j = b + TPP * i;
        if (active && (nb % TPP == 0 || j < nb)) {
            const int k = j % NS;
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) st(base + r * NS, v[i][r]);
        }
    }
}

template <typename T, int LAYOUT, bool SIN, bool SOUT, int N, int R0, int R1, int TPP, class Prep, class GLoad,
          class GStore, class GIn = DefaultStageIn>
__device__ __forceinline__ void pencil_fft_fixed(const PlanDev& pl, const cplx<T>* __restrict__ pool, cplx<T>* lds,
                                                 int npen, Prep&& prep, GLoad&& gld, GStore&& gst,
                                                 GIn&& gin = GIn()) {
    static_assert(R0 * R1 == N, "plan must multiply to N");
    constexpr int TT = kBlock / TPP;
    constexpr int LDP = N + N / 16 + 1;
    const int tid = threadIdx.x;
    int p, b;
    if constexpr (LAYOUT == LAYOUT_PF) {
        p = tid % TT;
        b = tid / TT;
    } else {
        b = tid % TPP;
        p = tid / TPP;
    }
    // When TPP does not divide the block (TT TPP < 256) the last threads get b = TPP (pencil-fastest)
    // or p = TT (butterfly-fastest): neither is a thread of the plan.  Had they run, a butterfly
    // j = TPP would read element N of its pencil, one past its end.
    const bool active = p < npen && (LAYOUT != LAYOUT_PF || b < TPP);
    auto addr = [&](int n) -> int {
        if constexpr (LAYOUT == LAYOUT_PF) return n * TT + p;
        else return p * LDP + n + (n >> 4);
    };
    const auto ctx = prep(active ? p : 0);
    auto ld_g = [&](int n) -> cplx<T> { return gld(ctx, n); };
    auto st_g = [&](int n, cplx<T> v) { gst(ctx, n, v); };
    auto ld_l = [&](int n) -> cplx<T> { return lds[addr(n)]; };
    auto st_l = [&](int n, cplx<T> v) { lds[addr(n)] = v; };
    if constexpr (SIN) {
        if (active) {
            if constexpr (std::is_same<typename std::decay<GIn>::type, DefaultStageIn>::value) {
#pragma unroll 1
                for (int n = b; n < N; n += TPP) st_l(n, ld_g(n));
            } else {
                gin(ctx, b, TPP, st_l);
            }
        }
        __syncthreads();
    }
    if constexpr (R1 == 1) {
        // single stage
        if constexpr (SIN && SOUT) stage_fixed<R0, N, 1, TPP>(pool, b, active, true, ld_l, st_l);
        else if constexpr (SIN) stage_fixed<R0, N, 1, TPP>(pool, b, active, true, ld_l, st_g);
        else if constexpr (SOUT) stage_fixed<R0, N, 1, TPP>(pool, b, active, false, ld_g, st_l);
        else stage_fixed<R0, N, 1, TPP>(pool, b, active, false, ld_g, st_g);
    } else {
        if constexpr (SIN) stage_fixed<R0, N, 1, TPP>(pool, b, active, true, ld_l, st_l);
        else stage_fixed<R0, N, 1, TPP>(pool, b, active, false, ld_g, st_l);
        __syncthreads();
        const cplx<T>* tw = pool + pl.tw_off[1];
        if constexpr (SOUT) stage_fixed<R1, N, R0, TPP>(tw, b, active, true, ld_l, st_l);
        else stage_fixed<R1, N, R0, TPP>(tw, b, active, false, ld_l, st_g);
    }
    if constexpr (SOUT) {
        __syncthreads();
#pragma unroll 1
        for (int n = b; n < N; n += TPP)
            if (active) st_g(n, ld_l(n));
    }
}

// A pair generator for the first stage: gp(ctx, q, has_hi, lo, hi) produces the inputs q and
// q + N/2 of a pencil.
template <class F>
struct GenPair {
    F& f;
    template <class C, typename T>
    __device__ __forceinline__ void operator()(const C& c, int q, bool has_hi, cplx<T>& lo, cplx<T>& hi) const {
        f(c, q, has_hi, lo, hi);
    }
};

// Pencil-fastest two-stage FFT whose inputs are generated, not loaded: an even-radix first stage
// holds the inputs q and q + N/2 in one butterfly (elements r and r + R0/2), so the generator
// produces both from one call, straight into the butterfly's registers. The first stage's
// outputs go through LDS to the second stage, which stores to global memory.
template <typename T, int N, int R0, int R1, int TPP, class Prep, class GStore, class GP>
__device__ __forceinline__ void pencil_fft_fixed_gen(const PlanDev& pl, const cplx<T>* __restrict__ pool,
                                                     cplx<T>* lds, int npen, Prep&& prep, GStore&& gst, GP&& gp) {
    static_assert(R0 * R1 == N && R0 % 2 == 0, "the first stage must have an even radix");
    constexpr int TT = kBlock / TPP;
    constexpr int nb = N / R0;
    constexpr int BMAX = (nb + TPP - 1) / TPP;
    const int tid = threadIdx.x;
    const int p = tid % TT, b = tid / TT;
    const bool active = p < npen && b < TPP;  // b = TPP: a thread past TT TPP (TPP not dividing 256)
    const auto ctx = prep(active ? p : 0);
    auto ld_l = [&](int n) -> cplx<T> { return lds[n * TT + p]; };
    auto st_l = [&](int n, cplx<T> v) { lds[n * TT + p] = v; };
    auto st_g = [&](int n, cplx<T> v) { gst(ctx, n, v); };
    cplx<T> v[BMAX][R0];
#pragma unroll
    for (int i = 0; i < BMAX; ++i) {
        const int j = b + TPP * i;
        if (active && (nb % TPP == 0 || j < nb)) {
#pragma unroll
            for (int r = 0; r < R0 / 2; ++r) gp(ctx, j + r * nb, true, v[i][r], v[i][r + R0 / 2]);
            dft<R0, T>(v[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < BMAX; ++i) {
        const int j = b + TPP * i;
        if (active && (nb % TPP == 0 || j < nb)) {
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                if constexpr (R1 == 1) st_g(j * R0 + r, v[i][r]);
                else st_l(j * R0 + r, v[i][r]);
            }
        }
    }
    if constexpr (R1 > 1) {
        __syncthreads();
        stage_fixed<R1, N, R0, TPP>(pool + pl.tw_off[1], b, active, false, ld_l, st_g);
    }
}

// Butterfly-fastest two-stage FFT for the x pass of an even-length real output (FP32, 16
// butterflies of the first stage, one per thread of a 16-lane row). Input k of the half-length
// transform combines X_k with X_{M-k} (M = N). For j >= 1 the partner of element r of butterfly j
// is element R0 - 1 - r of butterfly 16 - j, so each row loads every X_k once (aligned) and gets
// the partners from the mirrored lane (row_partner16). Lane 0's partners are its own elements
// R0 - r, and the Nyquist value for r = 0.
// xld(ctx, k) loads X_k, xnyq(ctx) the Nyquist value, xcomb(k, X_k, X_{M-k}) forms the input.
template <typename T, bool SOUT, int N, int R0, int R1, int TPP, class Prep, class XLoad, class XNyq, class XComb,
          class GStore>
__device__ __forceinline__ void pencil_fft_fixed_xpair(const PlanDev& pl, const cplx<T>* __restrict__ pool,
                                                       cplx<T>* lds, int npen, Prep&& prep, XLoad&& xld,
                                                       XNyq&& xnyq, XComb&& xcomb, GStore&& gst) {
    static_assert(sizeof(T) == 4 && TPP == 16 && N / R0 == 16 && R1 > 1, "16 first-stage butterflies per row");
    constexpr int nb = 16;
    constexpr int LDP = N + N / 16 + 1;
    const int tid = threadIdx.x, b = tid % TPP, p = tid / TPP;
    const bool active = p < npen;
    auto addr = [&](int n) -> int { return p * LDP + n + (n >> 4); };
    const auto ctx = prep(active ? p : 0);
    auto ld_l = [&](int n) -> cplx<T> { return lds[addr(n)]; };
    auto st_l = [&](int n, cplx<T> v) { lds[addr(n)] = v; };
    auto st_g = [&](int n, cplx<T> v) { gst(ctx, n, v); };
    cplx<T> a[R0], v[R0];
    cplx<T> xn = {T(0), T(0)};
    if (active) {
#pragma unroll
        for (int r = 0; r < R0; ++r) a[r] = xld(ctx, b + nb * r);
        if (b == 0) xn = xnyq(ctx);
    } else {
#pragma unroll
        for (int r = 0; r < R0; ++r) a[r] = cplx<T>{T(0), T(0)};
    }
#pragma unroll
    for (int r = 0; r < R0; ++r) {
        const cplx<T> m = a[R0 - 1 - r];
        cplx<T> q = {row_partner16(m.x), row_partner16(m.y)};
        if (b == 0) q = r ? a[(R0 - r) % R0] : xn;
        v[r] = xcomb(b + nb * r, a[r], q);
    }
    dft<R0, T>(v);
    if (active) {
#pragma unroll
        for (int r = 0; r < R0; ++r) st_l(b * R0 + r, v[r]);
    }
    __syncthreads();
    const cplx<T>* tw = pool + pl.tw_off[1];
    if constexpr (SOUT) {
        stage_fixed<R1, N, R0, TPP>(tw, b, active, true, ld_l, st_l);
        __syncthreads();
#pragma unroll 1
        for (int n = b; n < N; n += TPP)
            if (active) st_g(n, ld_l(n));
    } else {
        stage_fixed<R1, N, R0, TPP>(tw, b, active, false, ld_l, st_g);
    }
}

// FFT policies: the pass kernels take one of these as a template parameter.
template <int EC>
struct GenericFft {
    static constexpr bool kDirectGen = false;
    static constexpr bool kXPair = false;
    template <typename T, int LAYOUT, bool SIN, bool SOUT, class... A>
    __device__ __forceinline__ static void run(const PlanDev& pl, const cplx<T>* __restrict__ pool, cplx<T>* lds,
                                               int npen, A&&... a) {
        pencil_fft<T, LAYOUT, SIN, SOUT, EC, false>(pl, pool, lds, npen, static_cast<A&&>(a)...);
    }
};
// runtime plans with a composite radix 6, 9 or 12 (kSpecGenericComp)
template <int EC>
struct GenericCompFft {
    static constexpr bool kDirectGen = false;
    static constexpr bool kXPair = false;
    template <typename T, int LAYOUT, bool SIN, bool SOUT, class... A>
    __device__ __forceinline__ static void run(const PlanDev& pl, const cplx<T>* __restrict__ pool, cplx<T>* lds,
                                               int npen, A&&... a) {
        pencil_fft<T, LAYOUT, SIN, SOUT, EC, false, true>(pl, pool, lds, npen, static_cast<A&&>(a)...);
    }
};
// lengths with a prime factor above 13 (kSpecNonSmooth): Bluestein or the direct DFT
template <int EC>
struct NonSmoothFft {
    static constexpr bool kDirectGen = false;
    static constexpr bool kXPair = false;
    template <typename T, int LAYOUT, bool SIN, bool SOUT, class... A>
    __device__ __forceinline__ static void run(const PlanDev& pl, const cplx<T>* __restrict__ pool, cplx<T>* lds,
                                               int npen, A&&... a) {
        pencil_fft<T, LAYOUT, SIN, SOUT, EC, true>(pl, pool, lds, npen, static_cast<A&&>(a)...);
    }
};
template <int N, int R0, int R1, int TPP>
struct FixedFft {
    static constexpr bool kDirectGen = R0 % 2 == 0;
    static constexpr bool kXPair = TPP == 16 && N / R0 == 16 && R1 > 1;
    template <typename T, bool SOUT, class Prep, class XLoad, class XNyq, class XComb, class GStore>
    __device__ __forceinline__ static void run_xpair(const PlanDev& pl, const cplx<T>* __restrict__ pool,
                                                     cplx<T>* lds, int npen, Prep&& prep, XLoad&& xld, XNyq&& xnyq,
                                                     XComb&& xcomb, GStore&& gst) {
        if constexpr (kXPair && sizeof(T) == 4)
            pencil_fft_fixed_xpair<T, SOUT, N, R0, R1, TPP>(pl, pool, lds, npen, prep, xld, xnyq, xcomb, gst);
    }
    template <typename T, class Prep, class GStore, class GP>
    __device__ __forceinline__ static void run_gen(const PlanDev& pl, const cplx<T>* __restrict__ pool, cplx<T>* lds,
                                                   int npen, Prep&& prep, GStore&& gst, GP&& gp) {
        if constexpr (kDirectGen)
            pencil_fft_fixed_gen<T, N, R0, R1, TPP>(pl, pool, lds, npen, static_cast<Prep&&>(prep),
                                                    static_cast<GStore&&>(gst), static_cast<GP&&>(gp));
    }
    template <typename T, int LAYOUT, bool SIN, bool SOUT, class... A>
    __device__ __forceinline__ static void run(const PlanDev& pl, const cplx<T>* __restrict__ pool, cplx<T>* lds,
                                               int npen, A&&... a) {
        pencil_fft_fixed<T, LAYOUT, SIN, SOUT, N, R0, R1, TPP>(pl, pool, lds, npen, static_cast<A&&>(a)...);
    }
};

// Neumaier two-sum accumulation
__device__ __forceinline__ void nadd(double& s, double& c, double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
}

// Block reduction of plain per-thread sums (s, s2) of one tile: plain FP64 sums within each
// 64-lane wave (at most 64 partials, each of a few dozen values), then the four wave sums combined
// with compensation by one thread; red: 8 doubles of LDS.  One barrier, where block_reduce_mom
// takes eight.
__device__ __forceinline__ void block_reduce_sums(double s, double s2, double* red, MomPartial* out) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        s2 += __shfl_xor(s2, off);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[w] = s;
        red[4 + w] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red[0], ca = 0.0, b = red[4], cb = 0.0;
        for (int i = 1; i < kBlock / 64; ++i) {
            nadd(a, ca, red[i]);
            nadd(b, cb, red[4 + i]);
        }
        *out = {a, ca, b, cb};
    }
}

// Block reduction of compensated pairs; red: 4 * kBlock doubles of LDS.
__device__ __forceinline__ void block_reduce_mom(double s, double c, double s2, double c2, double* red,
                                                 MomPartial* out) {
    const int tid = threadIdx.x;
    red[tid] = s;
    red[tid + kBlock] = c;
    red[tid + 2 * kBlock] = s2;
    red[tid + 3 * kBlock] = c2;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) {
            double a = red[tid], ca = red[tid + kBlock];
            nadd(a, ca, red[tid + w]);
            ca += red[tid + w + kBlock];
            red[tid] = a;
            red[tid + kBlock] = ca;
            double a2 = red[tid + 2 * kBlock], ca2 = red[tid + 3 * kBlock];
            nadd(a2, ca2, red[tid + w + 2 * kBlock]);
            ca2 += red[tid + w + 3 * kBlock];
            red[tid + 2 * kBlock] = a2;
            red[tid + 3 * kBlock] = ca2;
        }
        __syncthreads();
    }
    if (tid == 0) *out = {red[0], red[kBlock], red[2 * kBlock], red[3 * kBlock]};
}

__device__ __forceinline__ int signed_freq(int m, int n) { return (m > n / 2) ? m - n : m; }

template <typename T>
__device__ __forceinline__ T pou_sqrt_weight(int i, int n, int nlo, int nhi) {
    T w = T(1);
    if (nlo >= 2 && i < nlo) {
        const T t = T(nlo - 1 - i) / T(nlo - 1);
        if constexpr (sizeof(T) == 4) w *= 0.5f * (1.0f + cospif(t));
        else w *= 0.5 * (1.0 + cos(kPiRef * t));
    }
    if (nhi >= 2 && i >= n - nhi) {
        const T t = T(i - (n - nhi)) / T(nhi - 1);
        if constexpr (sizeof(T) == 4) w *= 0.5f * (1.0f + cospif(t));
        else w *= 0.5 * (1.0 + cos(kPiRef * t));
    }
    return w;
}

// Square root of pou_sqrt_weight's product: sqrt(0.5 (1 + cos(pi t))) = |cos(pi t / 2)| for each
// ramp, so one cosine per ramp and no square root (the same angles, the reference's PI in FP64).
template <typename T>
__device__ __forceinline__ T pou_sqrt_factor(int i, int n, int nlo, int nhi) {
    T w = T(1);
    if (nlo >= 2 && i < nlo) {
        const T t = T(nlo - 1 - i) / T(nlo - 1);
        if constexpr (sizeof(T) == 4) w *= fabsf(cospif(0.5f * t));
        else w *= fabs(cos(0.5 * kPiRef * t));
    }
    if (nhi >= 2 && i >= n - nhi) {
        const T t = T(i - (n - nhi)) / T(nhi - 1);
        if constexpr (sizeof(T) == 4) w *= fabsf(cospif(0.5f * t));
        else w *= fabs(cos(0.5 * kPiRef * t));
    }
    return w;
}

__device__ __forceinline__ bool in_overlap(int i, int n, int nlo, int nhi) {
    return (nlo > 0 && i < nlo) || (nhi > 0 && i >= n - nhi);
}

template <typename T>
__device__ __forceinline__ T apply_margi(T v, int margi, T mu, T sd) {
    if (margi == FOM_GAUSSIAN) return v * sd + mu;
    if (margi == FOM_LOGNORMAL) {
        if constexpr (sizeof(T) == 4) return __builtin_amdgcn_exp2f((v * sd + mu) * 1.44269504088896340736f);
        else return exp(v * sd + mu);
    }
    return v;
}

}  // namespace scarl
