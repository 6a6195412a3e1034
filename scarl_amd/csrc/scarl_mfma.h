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
