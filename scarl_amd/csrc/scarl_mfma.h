// scarl_mfma.h -- radix-16 DFT stages on the matrix cores (gfx950 v_mfma_f32_16x16x32_bf16) for
// the 256-point x-pass tile (256 = 16 x 16, FP32 data), the split-precision variant of the FP32
// Stockham pass.
//
// A backward 16-point DFT of 16 butterflies is the real product [Yr; Yi] = [[Fr, -Fi], [Fi, Fr]]
// [Xr; Xi] (F_qr = exp(+2 pi i q r / 16)): two 16 x 16 x 32 matrix products per pencil and stage.
// Every FP32 operand is split into three bf16 terms x = x0 + x1 + x2 (round to nearest; 24
// significant bits in all) and the six products with i + j <= 2 accumulate in FP32, smallest first.
//
// Per pencil (Stockham, as the FP32 kernel): stage 1 Y[q][j] = sum_r F_qr x[j + 16 r], stored at
// 16 j + q; stage 2 Z[q][j] = sum_r F_qr (y[j + 16 r] e^{2 pi i j r / 256}), stored at j + 16 q.
// Operand lanes (16x16x32): A row = lane % 16, k = 8 (lane / 16) + t; B column = lane % 16, same
// k; D column = lane % 16, rows 4 (lane / 16) + t.  Each wave transforms 4 pencils of the tile.
#pragma once
#include "scarl_fft.h"

namespace scarl {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct Split3 {
    bf16x8 t0, t1, t2;
};

__device__ __forceinline__ void split3_at(Split3& s, int t, float x) {
    const __bf16 a = (__bf16)x;
    const float r = x - (float)a;
    const __bf16 b = (__bf16)r;
    s.t0[t] = a;
    s.t1[t] = b;
    s.t2[t] = (__bf16)(r - (float)b);
}

__device__ __forceinline__ f32x4 mfma_split3(const Split3& A, const Split3& B) {
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.t2, B.t0, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.t1, B.t1, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.t0, B.t2, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.t1, B.t0, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.t0, B.t1, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.t0, B.t0, c, 0, 0, 0);
    return c;
}

// cos and sin of 2 pi m / 16
__device__ __forceinline__ float cos16(int m) {
    constexpr double c[16] = {1.0,
                              0.92387953251128675612818318939678829,
                              0.70710678118654752440084436210484904,
                              0.38268343236508977172845998403039887,
                              0.0,
                              -0.38268343236508977172845998403039887,
                              -0.70710678118654752440084436210484904,
                              -0.92387953251128675612818318939678829,
                              -1.0,
                              -0.92387953251128675612818318939678829,
                              -0.70710678118654752440084436210484904,
                              -0.38268343236508977172845998403039887,
                              0.0,
                              0.38268343236508977172845998403039887,
                              0.70710678118654752440084436210484904,
                              0.92387953251128675612818318939678829};
    return (float)c[m & 15];
}
__device__ __forceinline__ float sin16(int m) { return cos16(m - 4); }

// Lane operands of the backward 16-point DFT matrices: A_re = [Fr, -Fi], A_im = [Fi, Fr]
// (16 x 32, row q = lane % 16, k = 8 (lane / 16) + t, F_qr = exp(+2 pi i q r / 16)).
__device__ __forceinline__ void dft16_operands(Split3& Are, Split3& Aim) {
    const int lane = threadIdx.x & 63, row = lane & 15, g = lane >> 4;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int k = 8 * g + t, r = k & 15, m = row * r;
        const float c = cos16(m), s = sin16(m);
        split3_at(Are, t, k < 16 ? c : -s);
        split3_at(Aim, t, k < 16 ? s : c);
    }
}

// Most 16-column groups one wave of a 256-thread tile holds through the stage barrier: the
// planner keeps T * N <= 64 * 16 * kMfmaMaxGroups for plans whose first stage runs here.
constexpr int kMfmaMaxGroups = 4;

// First (radix-16, untwiddled) Stockham stage of a whole tile on the matrix cores, in place in LDS.
// Columns c = (p, j), p < npen pencils, j < M = N / 16: Y[p][16 j + q] = sum_r F_qr x[p][j + M r].
// Each wave takes the 16-column groups wave, wave + 4, ...: per group two products
// [A_re; A_im] x [Re x; Im x] (B column = lane % 16, k = 8 (lane / 16) + t; D rows 4 (lane / 16) + t),
// each as six split bf16 MFMAs.  Every thread of the block must call it (it holds two barriers).
template <class Addr>
__device__ __forceinline__ void mfma_stage16(cplx<float>* lds, int N, int npen, Addr&& addr) {
    const int M = N >> 4, ncol = npen * M, ngroups = (ncol + 15) >> 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 15, g = lane >> 4;
    Split3 Are, Aim;
    dft16_operands(Are, Aim);
    f32x4 yr[kMfmaMaxGroups], yi[kMfmaMaxGroups];
#pragma unroll
    for (int i = 0; i < kMfmaMaxGroups; ++i) {
        const int grp = wave + 4 * i;
        if (grp < ngroups) {  // uniform over the wave
            const int c = 16 * grp + col;
            const bool ok = c < ncol;
            const int p = ok ? c / M : 0, j = ok ? c - p * M : 0;
            Split3 B;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int k = 8 * g + t, r = k & 15;
                const cplx<float> v = ok ? lds[addr(p, j + M * r)] : cplx<float>{0.f, 0.f};
                split3_at(B, t, k < 16 ? v.x : v.y);
            }
            yr[i] = mfma_split3(Are, B);
            yi[i] = mfma_split3(Aim, B);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kMfmaMaxGroups; ++i) {
        const int grp = wave + 4 * i;
        const int c = 16 * grp + col;
        if (grp < ngroups && c < ncol) {
            const int p = c / M, j = c - p * M;
#pragma unroll
            for (int t = 0; t < 4; ++t) lds[addr(p, 16 * j + 4 * g + t)] = cplx<float>{yr[i][t], yi[i][t]};
        }
    }
    __syncthreads();
}

// Runtime-planned FP32 transform whose first stage (radix 16) runs on the matrix cores
// (mfma_stage16) and whose later stages are the register Stockham stages of pencil_fft.  The tile
// is always staged into LDS first (the loader, or the generator's gin).
template <int EC>
struct GenericMfmaFft {
    static constexpr bool kDirectGen = false;
    static constexpr bool kXPair = false;
    template <typename T, int LAYOUT, bool SIN, bool SOUT, class Prep, class GLoad, class GStore,
              class GIn = DefaultStageIn>
    __device__ __forceinline__ static void run(const PlanDev& pl, const cplx<T>* __restrict__ pool, cplx<T>* lds,
                                               int npen, Prep&& prep, GLoad&& gld, GStore&& gst, GIn&& gin = GIn()) {
        static_assert(sizeof(T) == 4, "the MFMA stage serves FP32 plans");
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
        auto addr = [&](int pp, int n) -> int {
            if constexpr (LAYOUT == LAYOUT_PF) return n * Tt + pp;
            else return pp * ldp + n + (n >> 4) * sk;
        };
        const auto ctx = prep(active ? p : 0);
        auto ld_l = [&](int n) -> cplx<T> { return lds[addr(p, n)]; };
        auto st_l = [&](int n, cplx<T> v) { lds[addr(p, n)] = v; };
        auto st_g = [&](int n, cplx<T> v) { gst(ctx, n, v); };
        if (active) {
            if constexpr (std::is_same<typename std::decay<GIn>::type, DefaultStageIn>::value) {
#pragma unroll 1
                for (int n = b; n < N; n += tpp) st_l(n, gld(ctx, n));
            } else {
                gin(ctx, b, tpp, st_l);
            }
        }
        __syncthreads();
        mfma_stage16(lds, N, npen, addr);
        const int S = pl.nst;
        for (int s = 1; s < S; ++s) {
            const bool last = s == S - 1;
            if (last && !SOUT) {
                stage_dispatch<EC, T>(pl.R[s], N, pl.Ns[s], pool + pl.tw_off[s], b, tpp, active, true, ld_l, st_g);
            } else {
                stage_dispatch<EC, T>(pl.R[s], N, pl.Ns[s], pool + pl.tw_off[s], b, tpp, active, true, ld_l, st_l);
                __syncthreads();
            }
        }
        if (S == 1 || SOUT) {
#pragma unroll 1
            for (int n = b; n < N; n += tpp)
                if (active) st_g(n, ld_l(n));
        }
    }
};

struct MfmaFft256 {
    static constexpr bool kDirectGen = false;
    static constexpr bool kXPair = false;
    template <typename T, int LAYOUT, bool SIN, bool SOUT, class Prep, class GLoad, class GStore,
              class GIn = DefaultStageIn>
    __device__ __forceinline__ static void run(const PlanDev& pl, const cplx<T>* __restrict__ pool, cplx<T>* lds,
                                               int npen, Prep&& prep, GLoad&& gld, GStore&& gst, GIn&& = GIn()) {
        static_assert(sizeof(T) == 4 && LAYOUT == LAYOUT_BF, "the MFMA stages serve the FP32 x pass");
        constexpr int N = 256, TPP = 16, LDP = N + N / 16 + 1;
        const int tid = threadIdx.x, b = tid % TPP, p = tid / TPP;
        const bool active = p < npen;
        auto addr = [](int pp, int n) { return pp * LDP + n + (n >> 4); };
        const auto ctx = prep(active ? p : 0);
        // stage in (every pencil of the tile; inactive pencils stay unused)
        if (active) {
#pragma unroll 1
            for (int n = b; n < N; n += TPP) lds[addr(p, n)] = gld(ctx, n);
        }
        // DFT matrix operands of this lane: row q = lane % 16, k = 8 g + t
        const int lane = tid & 63, wave = tid >> 6, col = lane & 15, g = lane >> 4;
        Split3 Are, Aim;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int k = 8 * g + t, r = k & 15, m = col * r;
            const float c = cos16(m), s = sin16(m);
            split3_at(Are, t, k < 16 ? c : -s);
            split3_at(Aim, t, k < 16 ? s : c);
        }
        const cplx<T>* __restrict__ tw = pool + pl.tw_off[1];
        __syncthreads();
        // stage 1
#pragma unroll 1
        for (int pi = 0; pi < 4; ++pi) {
            const int pp = wave * 4 + pi;
            Split3 B;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int k = 8 * g + t, r = k & 15;
                const cplx<T> v = lds[addr(pp, 16 * r + col)];
                split3_at(B, t, k < 16 ? v.x : v.y);
            }
            const f32x4 yr = mfma_split3(Are, B), yi = mfma_split3(Aim, B);
#pragma unroll
            for (int t = 0; t < 4; ++t) lds[addr(pp, 16 * col + 4 * g + t)] = cplx<T>{yr[t], yi[t]};
        }
        __syncthreads();
        // stage 2 (twiddle e^{2 pi i j r / 256} from the same long-double table as the FP32 kernel)
#pragma unroll 1
        for (int pi = 0; pi < 4; ++pi) {
            const int pp = wave * 4 + pi;
            Split3 B;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int k = 8 * g + t, r = k & 15;
                cplx<T> v = lds[addr(pp, col + 16 * r)];
                if (r) v = cmul(v, tw[col * 15 + r - 1]);
                split3_at(B, t, k < 16 ? v.x : v.y);
            }
            const f32x4 zr = mfma_split3(Are, B), zi = mfma_split3(Aim, B);
#pragma unroll
            for (int t = 0; t < 4; ++t) lds[addr(pp, col + 16 * (4 * g + t))] = cplx<T>{zr[t], zi[t]};
        }
        __syncthreads();
        if (active) {
#pragma unroll 1
            for (int n = b; n < N; n += TPP) gst(ctx, n, lds[addr(p, n)]);
        }
    }
};

}  // namespace scarl
