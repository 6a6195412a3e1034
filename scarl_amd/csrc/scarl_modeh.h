// scarl_modeh.h -- Hermitian C2R generator passes (device code, templated on precision).
//
//   z pass : amp(k) * W(k) generated on the fly (device Philox noise, or host-exported noise)
//            -> C2C backward along z, into the half-spectrum workspace
//   y pass : C2C backward along y, in place on the workspace
//   x pass : C2R along x as a half-length complex FFT, fused with the marginal transform,
//            the partition-of-unity blend and compensated moments.
// For a 2D field (n3 = 1) the y pass generates.
#pragma once
#include "scarl_fft.h"
#include "scarl_mfma.h"

namespace scarl {

// amplitude sqrt(S(kappa) * scale) at half-spectrum point (kx, ky, kz)
template <typename T>
__device__ __forceinline__ T amplitude_h(const FieldH& f, int kx, int ky, int kz) {
    const T fx = T(kx), fy = T(signed_freq(ky, f.n[1])), fz = T(signed_freq(kz, f.n[2]));
    const T a0 = T(f.alpha[0]) * fx, a1 = T(f.alpha[1]) * fy, a2 = T(f.alpha[2]) * fz;
    const T q = a0 * a0 + a1 * a1 + a2 * a2;
    if (f.model == CM_GAUSSIAN) {
        if constexpr (sizeof(T) == 4) return T(f.amp0) * __builtin_amdgcn_exp2f(-0.18033688011112042f * q);
        else return f.amp0 * exp(-0.125 * q);
    } else {
        if constexpr (sizeof(T) == 4)
            return T(f.amp0) * __builtin_amdgcn_exp2f(T(f.expo) * __builtin_amdgcn_logf(1.0f + q));
        else return f.amp0 * pow(1.0 + q, f.expo);
    }
}

// Hermitian white noise at half-spectrum point (kx, ky, kz).  Off the self-conjugate planes
// (kx = 0, kx = n1/2) W = (g1 + i g2)/sqrt2 of the point itself; on them W(kx,-ky,-kz) =
// conj W(kx,ky,kz), drawn at the member of the pair with the lower linear index, and the
// self-conjugate points are real g1.  With h3 = ceil(n3/2), the draw of point (kx, ky, kz) is
// gauss_at with counter kx + h1 (ky + n2 (kz mod h3)) and half = (kz >= h3): the points kz and
// kz + h3 of a z pencil share one Philox block, and they are two inputs of the same butterfly of
// an even-radix first stage along z.
template <typename T>
__device__ __forceinline__ cplx<T> noise_h(const FieldH& f, int kx, int ky, int kz) {
    const T rs2 = T(0.70710678118654752440084436210484904);
    const int n1 = f.n[0], n2 = f.n[1], n3 = f.n[2];
    const int64_t h1 = f.h1;
    const int h3 = (n3 + 1) >> 1;
    const bool selfconj = (kx == 0) || (((n1 & 1) == 0) && kx == n1 / 2);
    if (!selfconj) {
        const int hz = kz >= h3;
        const cplx<T> g = gauss_at<T>(kx + h1 * (ky + (int64_t)n2 * (kz - hz * h3)), hz, f.seed, f.stream);
        return {g.x * rs2, g.y * rs2};
    }
    const int py = ky ? n2 - ky : 0, pz = kz ? n3 - kz : 0;
    const int64_t ls = ky + (int64_t)n2 * kz, lp = py + (int64_t)n2 * pz;
    const int cy = ls <= lp ? ky : py, cz = ls <= lp ? kz : pz;
    const int hc = cz >= h3;
    const cplx<T> g = gauss_at<T>(kx + h1 * (cy + (int64_t)n2 * (cz - hc * h3)), hc, f.seed, f.stream);
    if (ls == lp) return {g.x, T(0)};
    return {g.x * rs2, (ls < lp ? g.y : -g.y) * rs2};
}

// The two points (kx, ky, q) and (kx, ky, q + h3), q < h3, off the self-conjugate planes: one
// Philox call.
template <typename T>
__device__ __forceinline__ void noise_h_pair(const FieldH& f, int kx, int ky, int q, cplx<T>& w0, cplx<T>& w1) {
    const T rs2 = T(0.70710678118654752440084436210484904);
    uint32_t o[4];
    philox_at(kx + (int64_t)f.h1 * (ky + (int64_t)f.n[1] * q), f.seed, f.stream, o);
    const cplx<T> g0 = box_muller<T>(o[0], o[1]), g1 = box_muller<T>(o[2], o[3]);
    w0 = {g0.x * rs2, g0.y * rs2};
    w1 = {g1.x * rs2, g1.y * rs2};
}

template <typename T>
__device__ __forceinline__ cplx<T> generate_h(const FieldH& f, int kx, int ky, int kz) {
    cplx<T> w;
    if (f.src_mode == SRC_HOST_NOISE)
        w = gload(reinterpret_cast<const cplx<T>*>(f.W) + (kx + (int64_t)f.h1 * (ky + (int64_t)f.n[1] * kz)));
    else w = noise_h<T>(f, kx, ky, kz);
    const T a = amplitude_h<T>(f, kx, ky, kz);
    return {a * w.x, a * w.y};
}

// Strided pass along z (axis 2) or y (axis 1), in place on the half-spectrum workspace.
// GEN: the pass may generate its input (staged through LDS).
template <typename T, bool GEN, class Fft>
__global__ __launch_bounds__(kBlock) void k_h_strided(const FieldH* __restrict__ fields,
                                                       const Tile* __restrict__ tiles, int tile0,
                                                       const PlanDev* __restrict__ plans,
                                                       const cplx<T>* __restrict__ pool, int axis) {
    HIP_DYNAMIC_SHARED(char, smem)
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    const Tile tl = tiles[tile0 + blockIdx.x];
    const FieldH& f = fields[tl.field];
    const PlanDev& pl = plans[axis == 2 ? f.plan_z : f.plan_y];
    const int wp = f.wp, n2 = f.n[1], n3 = f.n[2];
    const int nother = (axis == 2) ? n2 : n3;
    const bool nyq = tl.nyq != 0;
    const int npen = min(pl.T, (nyq ? nother : wp * nother) - tl.pencil0);
    const int64_t wz = f.wz;
    const int64_t S = nyq ? ((axis == 2) ? (int64_t)n2 : 1) : ((axis == 2) ? wz : (int64_t)wp);
    cplx<T>* ws = reinterpret_cast<cplx<T>*>(f.ws) + (nyq ? wz * n3 : 0);
    const bool gen = GEN && (f.first_axis == axis) && (f.src_mode != SRC_WORKSPACE);
    // Element n of a pencil sits n * Sb bytes past the pencil's first element: one 32 x 32 -> 64-bit
    // multiply-add per access (the planner keeps a plane pitch below 4 GiB).
    const uint32_t Sb = (uint32_t)(S * (int64_t)sizeof(cplx<T>));
    // The field's FP32 generation constants, read once: per Philox call they would be reloaded
    // (and the FP64 ones converted again), since the descriptors sit in global memory that the
    // compiler cannot prove disjoint from the pass's stores.
    const bool gauss = f.model == CM_GAUSSIAN;
    const float g_a0 = float(f.alpha[0]), g_a1 = float(f.alpha[1]), g_az = float(f.alpha[2]);
    const float g_e2 = gauss ? -0.36067376022224085f : float(2.0 * f.expo);
    const float g_cc = float(-0.69314718055994530942 * f.amp0 * f.amp0);
    const uint64_t g_seed = f.seed;
    const uint32_t g_stream = f.stream;
    const int g_h1 = f.h1, g_N = pl.N;
    // Philox counters of a z pencil step by h1 n2 per kz (below 2^32: a kz plane of the workspace
    // stays under 4 GiB)
    const uint32_t g_cstride = (uint32_t)g_h1 * (uint32_t)n2;
    const cplx<T>* g_kx0 = reinterpret_cast<const cplx<T>*>(f.kx0);
    struct Ctx {
        char* row;     // the pencil's first element
        int kx, k2;    // k2 = ky (z pass) or kz (y pass)
        float qxy;     // z pass, FP32 generation: (alpha_x kx)^2 + (alpha_y f_y)^2 (+1 Matern)
        uint64_t ctr;  // z pass: Philox counter of the pencil's kz = 0 point
    };
    auto prep = [&](int p) -> Ctx {
        const int P = tl.pencil0 + p;
        Ctx c;
        if (nyq) {
            c = {reinterpret_cast<char*>(ws + ((axis == 2) ? (int64_t)P : (int64_t)n2 * P)), wp, P, 0.0f, 0};
        } else {
            const int kx = P % wp, q = P / wp;
            const int64_t base = (axis == 2) ? (int64_t)P : kx + wz * q;
            c = {reinterpret_cast<char*>(ws + base), kx, q, 0.0f, 0};
        }
        if constexpr (GEN && sizeof(T) == 4) {
            const float ax = g_a0 * float(c.kx), ay = g_a1 * float(signed_freq(c.k2, n2));
            c.qxy = fmaf(ax, ax, ay * ay) + (gauss ? 0.0f : 1.0f);
            c.ctr = (uint64_t)c.kx + (uint64_t)g_h1 * (uint64_t)c.k2;
        }
        return c;
    };
    auto at = [&](const Ctx& c, int n) { return reinterpret_cast<cplx<T>*>(c.row + (uint64_t)(uint32_t)n * Sb); };
    auto gld = [&](const Ctx& c, int n) -> cplx<T> {
        if constexpr (GEN) {
            if (gen) return (axis == 2) ? generate_h<T>(f, c.kx, c.k2, n) : generate_h<T>(f, c.kx, n, c.k2);
        }
        return gload(at(c, n));
    };
    auto gst = [&](const Ctx& c, int n, cplx<T> v) { gstore(at(c, n), v); };
    // Generation along z: the points q and q + h3 (h3 = ceil(n3/2)) of a pencil share one Philox
    // call (noise_h). gpair_dev produces both for device noise off the self-conjugate planes;
    // gpair falls back to one gld per point on those planes, for host-exported noise and for a
    // workspace source.
    auto gpair_dev = [&](const Ctx& c, int q, bool has_hi, cplx<T>& lo, cplx<T>& hi) {
        const int N = g_N, h3 = (N + 1) >> 1;
        {
            if constexpr (sizeof(T) == 4) {
                // amp(k) W(k) = sqrt(-ln2 amp0^2 log2(u1) P(q)) (cos 2 pi u2, sin 2 pi u2) with
                // P = (1 + s)^(2 expo) (Matern family) or 2^(-s / (4 ln 2)) (Gaussian), s = sum_d
                // (alpha_d f_d)^2: amplitude, the 1/sqrt2 of the complex draw and the Box-Muller
                // radius share one square root.  |f_z| = min(kz, n3 - kz) is q for kz = q < h3 and
                // n3 - h3 - q for kz = q + h3.
                auto root = [&](float fz, uint32_t u) {
                    const float t = g_az * fz;
                    const float s2 = fmaf(t, t, c.qxy);
                    const float P = __builtin_amdgcn_exp2f(g_e2 * (gauss ? s2 : __builtin_amdgcn_logf(s2)));
                    return __builtin_amdgcn_sqrtf(g_cc * __builtin_amdgcn_logf(u24<float>(u)) * P);
                };
                uint32_t o[4];
                philox_at(c.ctr + (uint64_t)g_cstride * (uint32_t)q, g_seed, g_stream, o);
                const float fq = float(q);
                const float r0 = root(fq, o[0]), v0 = u24<float>(o[1]);
                lo = cplx<T>{r0 * __builtin_amdgcn_cosf(v0), r0 * __builtin_amdgcn_sinf(v0)};
                if (has_hi) {
                    const float r1 = root(float(N - h3) - fq, o[2]), v1 = u24<float>(o[3]);
                    hi = cplx<T>{r1 * __builtin_amdgcn_cosf(v1), r1 * __builtin_amdgcn_sinf(v1)};
                }
            } else {
                cplx<T> w0, w1;
                noise_h_pair<T>(f, c.kx, c.k2, q, w0, w1);
                const T a0 = amplitude_h<T>(f, c.kx, c.k2, q);
                lo = cplx<T>{a0 * w0.x, a0 * w0.y};
                if (has_hi) {
                    const T a1 = amplitude_h<T>(f, c.kx, c.k2, q + h3);
                    hi = cplx<T>{a1 * w1.x, a1 * w1.y};
                }
            }
        }
        // a kx = 0 pencil (self-conjugate: the draw above pairs the wrong points) takes the inputs
        // k_h_kx0 made; every lane draws, so the generator stays out of any branch
        if (g_kx0 != nullptr && c.kx == 0) {
            const cplx<T>* z0 = g_kx0 + (int64_t)n3 * c.k2;
            lo = gload(z0 + q);
            if (has_hi) hi = gload(z0 + (q + h3));
        }
    };
    auto gpair = [&](const Ctx& c, int q, bool has_hi, cplx<T>& lo, cplx<T>& hi) {
        const int h3 = (pl.N + 1) >> 1;
        const bool paired = gen && axis == 2 && f.src_mode == SRC_DEVICE_NOISE && c.kx != 0 &&
                            !(((f.n[0] & 1) == 0) && c.kx == f.n[0] / 2);
        if (paired) {
            gpair_dev(c, q, has_hi, lo, hi);
        } else {
            lo = gld(c, q);
            if (has_hi) hi = gld(c, q + h3);
        }
    };
    if constexpr (GEN) {
        if (axis == 2) {
            // Row tiles with device noise: the first stage's butterflies hold the pairs
            // (q, q + N/2), generated straight into registers. A kx = 0 pencil (self-conjugate)
            // takes its inputs from the plane k_h_kx0 generated before this pass (or, without it,
            // the tile stages through LDS). The test is uniform over the tile, so only the paired
            // draw is unrolled into the butterflies. Nyquist tiles and host noise stage the
            // generated pencils through LDS.
            const int kx_first = tl.pencil0 % wp;
            const bool off_sc = f.kx0 != nullptr || (kx_first != 0 && kx_first + npen <= wp);
            const bool direct = gen && !nyq && f.src_mode == SRC_DEVICE_NOISE && off_sc;
            if constexpr (Fft::kDirectGen) {
                if (direct) {
                    Fft::template run_gen<T>(pl, pool, lds, npen, prep, gst, GenPair<decltype(gpair_dev)>{gpair_dev});
                    return;
                }
            }
            {
                // stage the generated pencil through LDS
                auto gin = [&](const Ctx& c, int b, int tpp, auto&& put) {
                    const int N = pl.N, h3 = (N + 1) >> 1;
#pragma unroll 1
                    for (int q = b; q < h3; q += tpp) {
                        cplx<T> lo, hi;
                        const bool has_hi = q + h3 < N;
                        gpair(c, q, has_hi, lo, hi);
                        put(q, lo);
                        if (has_hi) put(q + h3, hi);
                    }
                };
                Fft::template run<T, LAYOUT_PF, true, false>(pl, pool, lds, npen, prep, gld, gst, gin);
            }
            return;
        }
    }
    Fft::template run<T, LAYOUT_PF, GEN, false>(pl, pool, lds, npen, prep, gld, gst);
}

// The kx = 0 plane's inputs a(k) W(k) of every 3D field with device noise (self-conjugate: the
// draw of (0, ky, kz) is that of (0, -ky, -kz) conjugated), dense (n3, n2) with kz fastest (a z
// pencil's inputs are contiguous), for the z pass's direct generation.  Block (x, y): points
// [256 x, 256 x + 256) of field ids[y].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_h_kx0(const FieldH* __restrict__ fields, const int* __restrict__ ids) {
    const FieldH& f = fields[ids[blockIdx.y]];
    if (f.kx0 == nullptr || f.src_mode != SRC_DEVICE_NOISE) return;
    const int n2 = f.n[1], n3 = f.n[2];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)n2 * n3) return;
    const int kz = (int)(i % n3), ky = (int)(i / n3);
    gstore(reinterpret_cast<cplx<T>*>(const_cast<void*>(f.kx0)) + i, generate_h<T>(f, 0, ky, kz));
}

// x pass: for even n1 a half-length complex FFT of
// Z_k = (X_k + X*_{M-k}) + i e^{2 pi i k/n1} (X_k - X*_{M-k}) (imaginary parts of X_0 and X_M
// ignored), x_{2k} = Re z_k, x_{2k+1} = Im z_k; for odd n1 a full-length complex FFT of the
// Hermitian extension (X_k for k < h1, conj X_{n1-k} above; imaginary part of X_0 ignored),
// x_k = Re z_k; then
//   no blend: marginal, store, moments;
//   blend:    value * sqrt(w_PoU); first owner (colour order) stores, later owners add, the last
//             owner applies the marginal and takes the moments.
template <typename T, class Fft, bool BLEND, bool FULL>
__global__ __launch_bounds__(kBlock) void k_h_xpass(const FieldH* __restrict__ fields, const Tile* __restrict__ tiles,
                                                     int tile0, const PlanDev* __restrict__ plans,
                                                     const cplx<T>* __restrict__ pool,
                                                     MomPartial* __restrict__ partials) {
    HIP_DYNAMIC_SHARED(char, smem)
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    const int tix = tile0 + blockIdx.x;
    const Tile tl = tiles[tix];
    // a copy: the compiler then keeps the descriptor in registers across the epilogue's stores
    // instead of reloading it after each (they could alias it as far as it can tell)
    const FieldH f = fields[tl.field];
    const PlanDev& pl = plans[f.plan_x];
    const int n1 = f.n[0], n2 = f.n[1], n3 = f.n[2], M = n1 / 2, h1 = f.h1;
    constexpr bool full = FULL;  // odd n1 (full-length C2C over the Hermitian extension)
    const int64_t nrows = (int64_t)n2 * n3;
    // odd n1: a pencil is a pair of rows (2P, 2P + 1), transformed together as X_a + i X_b (both
    // outputs are real: row 2P is the real part, row 2P + 1 the imaginary part)
    const int64_t npencils = full ? (nrows + 1) / 2 : nrows;
    const int npen = (int)min((int64_t)pl.T, npencils - (int64_t)tl.pencil0);
    const cplx<T>* ws = reinterpret_cast<const cplx<T>*>(f.ws);
    const cplx<T>* pre = pool + pl.pre_off;
    T* out = reinterpret_cast<T*>(f.out);
    const T mu = T(f.mu), sd = T(f.sd);
    double ms = 0.0, ms2 = 0.0;
    const bool mom = f.mom_slot >= 0;

    struct Ctx {
        const cplx<T>* X;   // row (kx < wp)
        const cplx<T>* Xn;  // its kx = M element in the Nyquist plane (even n1)
        T* o;
        T swyz;    // square root of the row's PoU weight (blend)
        int myz;   // overlap mask bits of the row (blend): 2 y, 4 z
        // odd n1: the pair's second row (X2 null past the last row)
        const cplx<T>* X2;
        T* o2;
        T swyz2;
        int myz2;
    };
    // blend: the x factors of the PoU weight's square root for the tile's rows (all of one field),
    // in LDS after the transform's tile (1 off the ramps)
    T* wxl = reinterpret_cast<T*>(smem + (size_t)pl.lds_elems * sizeof(cplx<T>));
    if constexpr (BLEND) {
#pragma unroll 1
        for (int i = threadIdx.x; i < n1; i += kBlock) wxl[i] = pou_sqrt_factor<T>(i, n1, f.nov_lo[0], f.nov_hi[0]);
    }
    auto row_mask = [&](int row) {
        const int ky = row % n2, kz = row / n2;
        return (in_overlap(ky, n2, f.nov_lo[1], f.nov_hi[1]) ? 2 : 0) |
               (in_overlap(kz, n3, f.nov_lo[2], f.nov_hi[2]) ? 4 : 0);
    };
    auto row_ctx = [&](int row, const cplx<T>*& X, T*& o, T& swyz, int& myz) {
        const int ky = row % n2, kz = row / n2;
        X = ws + (int64_t)ky * f.wp + kz * f.wz;
        o = out + ky * f.out_ld1 + kz * f.out_ld2;
        swyz = T(1);
        myz = 0;
        if (BLEND) {
            swyz = pou_sqrt_factor<T>(ky, n2, f.nov_lo[1], f.nov_hi[1]) *
                   pou_sqrt_factor<T>(kz, n3, f.nov_lo[2], f.nov_hi[2]);
            myz = row_mask(row);
        }
    };
    auto prep = [&](int p) -> Ctx {
        Ctx c;
        if (full) {
            const int64_t row = 2 * ((int64_t)tl.pencil0 + p);
            row_ctx((int)row, c.X, c.o, c.swyz, c.myz);
            c.Xn = nullptr;
            if (row + 1 < nrows) {
                row_ctx((int)row + 1, c.X2, c.o2, c.swyz2, c.myz2);
            } else {
                c.X2 = nullptr;
                c.o2 = nullptr;
                c.swyz2 = T(1);
                c.myz2 = 0;
            }
            return c;
        }
        const int row = tl.pencil0 + p;
        row_ctx(row, c.X, c.o, c.swyz, c.myz);
        c.Xn = ws + f.wz * n3 + row;
        c.X2 = nullptr;
        c.o2 = nullptr;
        c.swyz2 = T(1);
        c.myz2 = 0;
        return c;
    };
    // element k of the Hermitian extension of a row (X_k below h1, conj X_{n1-k} above; the
    // imaginary part of X_0 ignored)
    auto hext = [&](const cplx<T>* X, int k) -> cplx<T> {
        if (k < h1) {
            cplx<T> a = gload(X + k);
            if (k == 0) a.y = T(0);
            return a;
        }
        const cplx<T> a = gload(X + (n1 - k));
        return {a.x, -a.y};
    };
    auto gld = [&](const Ctx& c, int k) -> cplx<T> {
        if (full) {
            const cplx<T> a = hext(c.X, k);
            const cplx<T> bb = c.X2 ? hext(c.X2, k) : cplx<T>{T(0), T(0)};
            return {a.x - bb.y, a.y + bb.x};
        }
        cplx<T> a = gload(c.X + k), bb = gload(k ? c.X + (M - k) : c.Xn);
        if (k == 0) {
            a.y = T(0);
            bb.y = T(0);
        }
        bb.y = -bb.y;
        const cplx<T> s = a + bb, d = a - bb;
        const cplx<T> wd = cmul(pre[k], d);
        return {s.x - wd.y, s.y + wd.x};
    };
    auto gst = [&](const Ctx& c, int k, cplx<T> z) {
        T* o = c.o;
        if constexpr (!BLEND) {
            if (full) {
                const T r = apply_margi<T>(z.x, f.margi, mu, sd);
                if (mom) {
                    const double d = r;
                    ms += d;
                    ms2 += d * d;
                }
                gstore1(o + k, r);
                if (c.o2) {
                    const T r2 = apply_margi<T>(z.y, f.margi, mu, sd);
                    if (mom) {
                        const double d = r2;
                        ms += d;
                        ms2 += d * d;
                    }
                    gstore1(c.o2 + k, r2);
                }
                return;
            }
            const T r0 = apply_margi<T>(z.x, f.margi, mu, sd), r1 = apply_margi<T>(z.y, f.margi, mu, sd);
            if (mom) {
                // a thread sums at most a few dozen values in plain FP64; the block, tile and
                // field combinations are compensated
                const double d0 = r0, d1 = r1;
                ms += d0 + d1;
                ms2 += d0 * d0 + d1 * d1;
            }
            gstore(reinterpret_cast<cplx<T>*>(o + 2 * k), cplx<T>{r0, r1});
        } else {
        // Branch-light: every point takes the weight (1 off the ramps: the row's y/z factor times
        // the x factor from the tile's LDS table), a point reads what an earlier owner wrote only
        // if it is not its first owner, and the marginal and the moments count only at its last
        // owner (a zero contribution elsewhere).  Even n1: both outputs of z go out as one
        // 2-element store when the row is aligned for it.
        // odd n1: element k of the pair's two rows (the second may be absent)
        const int ne = full ? (c.o2 ? 2 : 1) : 2;
        // even n1, aligned row: what the earlier owners stored at the element's two outputs comes
        // in one 8-byte load (the lanes' pairs are contiguous) wherever either output needs it
        const bool pair_ok = !full && (reinterpret_cast<uintptr_t>(o) % (2 * sizeof(T))) == 0;
        cplx<T> prev2{T(0), T(0)};
        if (pair_ok) {
            const int m0 = c.myz | (in_overlap(2 * k, n1, f.nov_lo[0], f.nov_hi[0]) ? 1 : 0);
            const int m1 = c.myz | (in_overlap(2 * k + 1, n1, f.nov_lo[0], f.nov_hi[0]) ? 1 : 0);
            if (f.color & (m0 | m1)) prev2 = gload(reinterpret_cast<const cplx<T>*>(o + 2 * k));
        }
        T val[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            if (e >= ne) break;
            const int i = full ? k : 2 * k + e;
            const bool r2 = full && e == 1;
            T* oe = r2 ? c.o2 : o;
            const int msk = (r2 ? c.myz2 : c.myz) | (in_overlap(i, n1, f.nov_lo[0], f.nov_hi[0]) ? 1 : 0);
            const T v0 = (e ? z.y : z.x) * ((r2 ? c.swyz2 : c.swyz) * wxl[i]);
            const T prev = pair_ok ? (e ? prev2.y : prev2.x) : (f.color & msk) ? gload1(oe + i) : T(0);
            const T v = (f.color & msk) ? v0 + prev : v0;  // not the first owner: add
            const bool last = (f.color & msk) == msk;
            const T r = apply_margi<T>(v, f.margi, mu, sd);
            val[e] = last ? r : v;
            if (mom) {
                const double d = last ? (double)r : 0.0;
                ms += d;
                ms2 += d * d;
            }
        }
        if (full) {
            gstore1(o + k, val[0]);
            if (ne == 2) gstore1(c.o2 + k, val[1]);
        } else if ((reinterpret_cast<uintptr_t>(o) % (2 * sizeof(T))) == 0) {
            gstore(reinterpret_cast<cplx<T>*>(o + 2 * k), cplx<T>{val[0], val[1]});
        } else {
            for (int e = 0; e < 2; ++e) gstore1(o + 2 * k + e, val[e]);
        }
        }
    };
    // the blend epilogue (PoU weights, owner order) is staged through LDS; the plain one stores from
    // registers in the last stage
    if constexpr (!FULL && Fft::kXPair && sizeof(T) == 4) {
        // every X_k loaded once, with the non-temporal hint (5 % on this pass); X_{M-k} comes from
        // the mirrored lane of the row
        auto xld = [&](const Ctx& c, int k) -> cplx<T> { return gload_nt(c.X + k); };
        auto xnyq = [&](const Ctx& c) -> cplx<T> { return gload(c.Xn); };
        auto xcomb = [&](int k, cplx<T> a, cplx<T> bb) -> cplx<T> {
            if (k == 0) {
                a.y = T(0);
                bb.y = T(0);
            }
            bb.y = -bb.y;
            const cplx<T> s = a + bb, d = a - bb;
            const cplx<T> wd = cmul(pre[k], d);
            return {s.x - wd.y, s.y + wd.x};
        };
        Fft::template run_xpair<T, BLEND>(pl, pool, lds, npen, prep, xld, xnyq, xcomb, gst);
    } else {
        Fft::template run<T, LAYOUT_BF, false, BLEND>(pl, pool, lds, npen, prep, gld, gst);
    }
    if (mom) {
        __syncthreads();
        block_reduce_sums(ms, ms2, reinterpret_cast<double*>(smem), partials + tix);
    }
}

// Launch dispatch for one precision: spec 0 runs the generic kernel with register budget E1 or E2,
// spec k > 0 the compile-time planned kernel k of SCARL_FIXED_SPECS (only those of precision PREC
// and the pass's layout are instantiated).
template <typename T, int PREC, int E1, int E2>
struct ModeHDispatch {
    template <bool GEN, class F>
    static void strided(int axis, const FieldH* fields, const Tile* tiles, int tile0, int ntiles,
                        const PlanDev* plans, const cplx<T>* p, size_t lds, hipStream_t s) {
        hipLaunchKernelGGL((k_h_strided<T, GEN, F>), dim3(ntiles), dim3(kBlock), lds, s, fields, tiles, tile0, plans,
                           p, axis);
    }
    template <bool BLEND, class F, bool FULL = false>
    static void xpass(const FieldH* fields, const Tile* tiles, int tile0, int ntiles, const PlanDev* plans,
                      const cplx<T>* p, MomPartial* partials, size_t lds, hipStream_t s) {
        hipLaunchKernelGGL((k_h_xpass<T, F, BLEND, FULL>), dim3(ntiles), dim3(kBlock), lds, s, fields, tiles, tile0,
                           plans, p, partials);
    }
    template <bool GEN>
    static hipError_t run_strided(int axis, int spec, int ecap, const FieldH* fields, const Tile* tiles,
                                  int tile0, int ntiles, const PlanDev* plans, const void* pool, size_t lds,
                                  hipStream_t s) {
        if (ntiles <= 0) return hipSuccess;
        const cplx<T>* p = static_cast<const cplx<T>*>(pool);
        if (spec == kSpecMfmaGeneric) {
            if constexpr (PREC == 4) strided<GEN, GenericMfmaFft<E1>>(axis, fields, tiles, tile0, ntiles, plans, p, lds, s);
            return hipGetLastError();
        }
        if (spec == kSpecNonSmooth) {
            if (ecap == E2) strided<GEN, NonSmoothFft<E2>>(axis, fields, tiles, tile0, ntiles, plans, p, lds, s);
            else strided<GEN, NonSmoothFft<E1>>(axis, fields, tiles, tile0, ntiles, plans, p, lds, s);
            return hipGetLastError();
        }
        if (spec == kSpecGenericComp) {
            if (ecap == E2) strided<GEN, GenericCompFft<E2>>(axis, fields, tiles, tile0, ntiles, plans, p, lds, s);
            else strided<GEN, GenericCompFft<E1>>(axis, fields, tiles, tile0, ntiles, plans, p, lds, s);
            return hipGetLastError();
        }
        switch (spec) {
#define SCARL_CASE(ID, P, L, N, R0, R1, TPP)                                                             \
    case ID:                                                                                             \
        if constexpr (P == PREC && L == LAYOUT_PF)                                                       \
            strided<GEN, FixedFft<N, R0, R1, TPP>>(axis, fields, tiles, tile0, ntiles, plans, p, lds, s); \
        break;
            SCARL_FIXED_SPECS(SCARL_CASE)
#undef SCARL_CASE
        default:
            if (ecap == E2) strided<GEN, GenericFft<E2>>(axis, fields, tiles, tile0, ntiles, plans, p, lds, s);
            else strided<GEN, GenericFft<E1>>(axis, fields, tiles, tile0, ntiles, plans, p, lds, s);
        }
        return hipGetLastError();
    }
    // full: odd n1, whose lengths only the generic kernel plans
    template <bool BLEND>
    static hipError_t run_xpass(int spec, int ecap, bool full, const FieldH* fields, const Tile* tiles,
                                int tile0, int ntiles, const PlanDev* plans, const void* pool, MomPartial* partials,
                                size_t lds, hipStream_t s) {
        if (ntiles <= 0) return hipSuccess;
        if (lds < 4 * kBlock * sizeof(double)) lds = 4 * kBlock * sizeof(double);
        const cplx<T>* p = static_cast<const cplx<T>*>(pool);
        if (full) {
            if (spec == kSpecNonSmooth) {
                if (ecap == E2) xpass<BLEND, NonSmoothFft<E2>, true>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
                else xpass<BLEND, NonSmoothFft<E1>, true>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            } else if (spec == kSpecGenericComp) {
                if (ecap == E2) xpass<BLEND, GenericCompFft<E2>, true>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
                else xpass<BLEND, GenericCompFft<E1>, true>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            } else if (ecap == E2) {
                xpass<BLEND, GenericFft<E2>, true>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            } else {
                xpass<BLEND, GenericFft<E1>, true>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            }
            return hipGetLastError();
        }
        if (spec == kSpecNonSmooth) {
            if (ecap == E2) xpass<BLEND, NonSmoothFft<E2>>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            else xpass<BLEND, NonSmoothFft<E1>>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            return hipGetLastError();
        }
        if (spec == kSpecGenericComp) {
            if (ecap == E2) xpass<BLEND, GenericCompFft<E2>>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            else xpass<BLEND, GenericCompFft<E1>>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            return hipGetLastError();
        }
        if (spec == kSpecMfma256) {
            if constexpr (PREC == 4) xpass<BLEND, MfmaFft256>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            return hipGetLastError();
        }
        if (spec == kSpecMfmaGeneric) {
            if constexpr (PREC == 4) xpass<BLEND, GenericMfmaFft<E1>>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            return hipGetLastError();
        }
        switch (spec) {
#define SCARL_CASE(ID, P, L, N, R0, R1, TPP)                                                                  \
    case ID:                                                                                                  \
        if constexpr (P == PREC && L == LAYOUT_BF)                                                            \
            xpass<BLEND, FixedFft<N, R0, R1, TPP>>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s); \
        break;
            SCARL_FIXED_SPECS(SCARL_CASE)
#undef SCARL_CASE
        default:
            if (ecap == E2) xpass<BLEND, GenericFft<E2>>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
            else xpass<BLEND, GenericFft<E1>>(fields, tiles, tile0, ntiles, plans, p, partials, lds, s);
        }
        return hipGetLastError();
    }
};

}  // namespace scarl

// Instantiate the launch wrappers for one precision T (SFX suffix, PREC bytes, register budgets).
// Launch wrappers of one precision, in four translation units that compile in parallel: the
// generating strided passes (SCARL_MODEH_STRIDED with GEN true), the other strided passes (GEN
// false), the blend's x pass (SCARL_MODEH_XPASS_PART with BLEND true), and the plain x pass with
// the kx = 0 plane and the dispatchers (SCARL_MODEH_XPASS and the part with BLEND false).
#define SCARL_MODEH_STRIDED(T, SFX, PREC, E1, E2, GEN, TAG)                                                 \
    namespace scarl {                                                                                      \
    hipError_t launch_h_strided_##TAG##_##SFX(int axis, int spec, int ecap, const FieldH* fields,          \
                                              const Tile* tiles, int tile0, int ntiles, const PlanDev* plans, \
                                              const void* pool, size_t lds, hipStream_t s) {               \
        return ModeHDispatch<T, PREC, E1, E2>::template run_strided<GEN>(axis, spec, ecap, fields, tiles, tile0, \
                                                                         ntiles, plans, pool, lds, s);     \
    }                                                                                                      \
    }
#define SCARL_MODEH_XPASS(T, SFX, PREC, E1, E2)                                                             \
    namespace scarl {                                                                                      \
    hipError_t launch_h_strided_gen_##SFX(int axis, int spec, int ecap, const FieldH* fields, const Tile* tiles, \
                                          int tile0, int ntiles, const PlanDev* plans, const void* pool,   \
                                          size_t lds, hipStream_t s);                                      \
    hipError_t launch_h_strided_plain_##SFX(int axis, int spec, int ecap, const FieldH* fields,            \
                                            const Tile* tiles, int tile0, int ntiles, const PlanDev* plans, \
                                            const void* pool, size_t lds, hipStream_t s);                  \
    hipError_t launch_h_strided_##SFX(int axis, bool gen, int spec, int ecap, const FieldH* fields,        \
                                      const Tile* tiles, int tile0, int ntiles, const PlanDev* plans,      \
                                      const void* pool, size_t lds, hipStream_t s) {                       \
        if (ntiles <= 0) return hipSuccess;                                                                \
        return gen ? launch_h_strided_gen_##SFX(axis, spec, ecap, fields, tiles, tile0, ntiles, plans, pool, lds, s) \
                   : launch_h_strided_plain_##SFX(axis, spec, ecap, fields, tiles, tile0, ntiles, plans, pool, lds, \
                                                  s);                                                      \
    }                                                                                                      \
    hipError_t launch_h_kx0_##SFX(const FieldH* fields, const int* ids, int nfields, int64_t max_pts,      \
                                  hipStream_t s) {                                                         \
        if (nfields <= 0 || max_pts <= 0) return hipSuccess;                                               \
        hipLaunchKernelGGL((k_h_kx0<T>), dim3((unsigned)((max_pts + kBlock - 1) / kBlock), (unsigned)nfields), \
                           dim3(kBlock), 0, s, fields, ids);                                                \
        return hipGetLastError();                                                                          \
    }                                                                                                      \
    hipError_t launch_h_xpass_blend_##SFX(int spec, int ecap, bool full, const FieldH* fields, const Tile* tiles, \
                                          int tile0, int ntiles, const PlanDev* plans, const void* pool,   \
                                          MomPartial* partials, size_t lds, hipStream_t s);                \
    hipError_t launch_h_xpass_plain_##SFX(int spec, int ecap, bool full, const FieldH* fields, const Tile* tiles, \
                                          int tile0, int ntiles, const PlanDev* plans, const void* pool,   \
                                          MomPartial* partials, size_t lds, hipStream_t s);                \
    hipError_t launch_h_xpass_##SFX(int spec, int ecap, bool blend, bool full, const FieldH* fields,       \
                                    const Tile* tiles, int tile0, int ntiles, const PlanDev* plans,        \
                                    const void* pool, MomPartial* partials, size_t lds, hipStream_t s) {   \
        return blend ? launch_h_xpass_blend_##SFX(spec, ecap, full, fields, tiles, tile0, ntiles, plans, pool, \
                                                  partials, lds, s)                                        \
                     : launch_h_xpass_plain_##SFX(spec, ecap, full, fields, tiles, tile0, ntiles, plans, pool, \
                                                  partials, lds, s);                                       \
    }                                                                                                      \
    }
#define SCARL_MODEH_XPASS_PART(T, SFX, PREC, E1, E2, BLEND, TAG)                                            \
    namespace scarl {                                                                                      \
    hipError_t launch_h_xpass_##TAG##_##SFX(int spec, int ecap, bool full, const FieldH* fields, const Tile* tiles, \
                                            int tile0, int ntiles, const PlanDev* plans, const void* pool, \
                                            MomPartial* partials, size_t lds, hipStream_t s) {             \
        return ModeHDispatch<T, PREC, E1, E2>::template run_xpass<BLEND>(spec, ecap, full, fields, tiles, tile0, \
                                                                         ntiles, plans, pool, partials, lds, s); \
    }                                                                                                      \
    }
