// scarl_k_r.hip -- REDFT01 (reference-parity) passes, FP64.
//
// DCT-III per axis as one complex FFT of length n per PAIR of real pencils:
// V_j = (X_j - i X_{n-j}) e^{i pi j/(2n)} is Hermitian for real X, so v = DFT^{-1}(V) is real and
// two pencils a, b travel as V(a) + i V(b);  y_{2m} = v_m, y_{2m+1} = v_{n-1-m}.
// Pass order x (octant amplitude generated from phi, staged through LDS) -> y -> z (marginal
// optionally fused); the x pass reads N/8 phases and writes N/4, the y pass N/4 -> N/2, the z
// pass N/2 -> N.
#include "scarl_fft.h"

namespace scarl {

// exp(-PI k_d^2) of the reference's Gaussian spectrum along axis d at octant index i
__device__ __forceinline__ double mode_r_gauss(const FieldR& f, int d, int i) {
    const double k = ((double)i * f.dk[d]) / f.corrL[d];
    return exp(-kPiRef * (k * k));
}

// SRC_F/fftw_ScaRL.f90:305-370 (kPoints, Sk product in axis order, sqrt(S) cos(2 PI phi)); e1, e2:
// mode_r_gauss along y and z, which are constant along an x pencil and evaluated once per pencil;
// the product keeps the reference's order S = ((corrL1 e0) corrL2 e1) corrL3 e2
__device__ __forceinline__ double mode_r_amp(const FieldR& f, int i, int j, int k, double e1, double e2) {
    const int64_t pidx = i + (int64_t)f.m[0] * (j + (int64_t)f.m[1] * k);
    const double ph = (f.src_mode == SRC_HOST_NOISE) ? gload1(f.phi + pidx)
                                                     : uniform53(static_cast<uint64_t>(pidx), f.seed, f.stream);
    double S = 1.0;
    S = S * f.corrL[0] * mode_r_gauss(f, 0, i);
    S = S * f.corrL[1] * e1;
    S = S * f.corrL[2] * e2;
    return 1.0 * sqrt(S) * cos(2.0 * kPiRef * ph);
}

template <int AXIS, class Fft>
__global__ __launch_bounds__(kBlock) void k_r_pass(const FieldR fp, const PlanDev* __restrict__ plans,
                                                    const cplx<double>* __restrict__ pool) {
    HIP_DYNAMIC_SHARED(char, smem)
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem);
    const FieldR& f = fp;
    const PlanDev& pl = plans[f.plan[AXIS]];
    const int n1 = f.n[0], n2 = f.n[1], n3 = f.n[2];
    const int m1 = f.m[0], m2 = f.m[1], m3 = f.m[2];
    const int n = f.n[AXIS], m = f.m[AXIS];
    // real pencils of this pass
    int64_t npen_real;
    if (AXIS == 0) npen_real = (int64_t)m2 * m3;
    else if (AXIS == 1) npen_real = (int64_t)n1 * m3;
    else npen_real = (int64_t)n1 * n2;
    const int64_t npen_cplx = (npen_real + 1) / 2;
    const int64_t c0 = (int64_t)blockIdx.x * pl.T;
    const int npen = (int)min((int64_t)pl.T, npen_cplx - c0);
    const cplx<double>* pre = pool + pl.pre_off;

    // per-thread context: the (up to) two real pencils of complex pencil p, decoded once
    struct Ctx {
        int64_t in[2], out[2];   // element-0 offsets of input / output of each real pencil
        int jy[2], kz[2];        // x pass: octant coordinates of the rows
        double e1[2], e2[2];     // x pass: the rows' y and z spectrum factors
        bool hasb;
    };
    auto prep = [&](int p) -> Ctx {
        Ctx c;
        const int64_t Pa = 2 * (c0 + p);
        c.hasb = Pa + 1 < npen_real;
        for (int e = 0; e < 2; ++e) {
            const int64_t P = (e == 0 || c.hasb) ? Pa + e : Pa;
            if constexpr (AXIS == 0) {
                c.jy[e] = (int)(P % m2);
                c.kz[e] = (int)(P / m2);
                c.e1[e] = mode_r_gauss(f, 1, c.jy[e]);
                c.e2[e] = mode_r_gauss(f, 2, c.kz[e]);
                c.in[e] = 0;
                c.out[e] = (int64_t)n1 * P;
            } else if constexpr (AXIS == 1) {
                const int64_t i = P % n1, k = P / n1;
                c.jy[e] = c.kz[e] = 0;
                c.e1[e] = c.e2[e] = 0.0;
                c.in[e] = i + (int64_t)n1 * m2 * k;
                c.out[e] = i + (int64_t)n1 * n2 * k;
            } else {
                c.jy[e] = c.kz[e] = 0;
                c.e1[e] = c.e2[e] = 0.0;
                c.in[e] = P;
                c.out[e] = P;
            }
        }
        return c;
    };
    auto rin = [&](const Ctx& c, int e, int j) -> double {
        if (j >= m) return 0.0;
        if constexpr (AXIS == 0) return mode_r_amp(f, j, c.jy[e], c.kz[e], c.e1[e], c.e2[e]);
        else if constexpr (AXIS == 1) return gload1(f.tmp1 + (c.in[e] + (int64_t)n1 * j));
        else return gload1(f.tmp2 + (c.in[e] + (int64_t)n1 * n2 * j));
    };
    auto rout = [&](const Ctx& c, int e, int o, double v) {
        if constexpr (AXIS == 0) {
            gstore1(f.tmp1 + (c.out[e] + o), v);
        } else if constexpr (AXIS == 1) {
            gstore1(f.tmp2 + (c.out[e] + (int64_t)n1 * o), v);
        } else {
            double r = v;
            if (f.margi == FOM_GAUSSIAN) r = r * f.sd + f.mu;
            else if (f.margi == FOM_LOGNORMAL) r = exp(r * f.sd + f.mu);
            gstore1(f.out + (c.out[e] + (int64_t)n1 * n2 * o), r);
        }
    };
    auto gld = [&](const Ctx& c, int j) -> cplx<double> {
        const cplx<double> w = pre[j];
        const double xa = rin(c, 0, j), xa2 = (j > 0) ? rin(c, 0, n - j) : 0.0;
        double xb = 0.0, xb2 = 0.0;
        if (c.hasb) {
            xb = rin(c, 1, j);
            xb2 = (j > 0) ? rin(c, 1, n - j) : 0.0;
        }
        // V(a) = (xa - i xa2) w ; V(b) likewise ; return V(a) + i V(b)
        const cplx<double> va = {xa * w.x + xa2 * w.y, xa * w.y - xa2 * w.x};
        const cplx<double> vb = {xb * w.x + xb2 * w.y, xb * w.y - xb2 * w.x};
        return {va.x - vb.y, va.y + vb.x};
    };
    auto gst = [&](const Ctx& c, int k, cplx<double> v) {
        const int o = (2 * k < n) ? 2 * k : 2 * (n - 1 - k) + 1;
        rout(c, 0, o, v.x);
        if (c.hasb) rout(c, 1, o, v.y);
    };
    // x pass: each octant amplitude X_j feeds V_j and V_{n-j}; one thread handles the pair (j, n-j)
    // and generates each amplitude once (gld would generate it for both)
    auto gin = [&](const Ctx& c, int b0, int tstep, auto&& put) {
        auto vpair = [&](int j, double xa, double xa2, double xb, double xb2) -> cplx<double> {
            const cplx<double> w = pre[j];
            const cplx<double> va = {xa * w.x + xa2 * w.y, xa * w.y - xa2 * w.x};
            const cplx<double> vb = {xb * w.x + xb2 * w.y, xb * w.y - xb2 * w.x};
            return {va.x - vb.y, va.y + vb.x};
        };
#pragma unroll 1
        for (int j = b0; 2 * j <= n; j += tstep) {
            const double xa = rin(c, 0, j), xb = c.hasb ? rin(c, 1, j) : 0.0;
            if (j == 0) {
                put(0, vpair(0, xa, 0.0, xb, 0.0));  // X_n = 0
            } else if (2 * j == n) {
                put(j, vpair(j, xa, xa, xb, xb));
            } else {
                const double ya = rin(c, 0, n - j), yb = c.hasb ? rin(c, 1, n - j) : 0.0;
                put(j, vpair(j, xa, ya, xb, yb));
                put(n - j, vpair(n - j, ya, xa, yb, xb));
            }
        }
    };
    if constexpr (AXIS == 0)
        Fft::template run<double, LAYOUT_BF, true, false>(pl, pool, lds, npen, prep, gld, gst, gin);
    else Fft::template run<double, LAYOUT_PF, false, false>(pl, pool, lds, npen, prep, gld, gst);
}



// ------------------------------------------------------------------------------------------
// build_Sk_mtx (SRC_F/spectra_ScaRL.f90:10-142), Gaussian branch, FP64.
// Element t = i + s1 (j + s2 k) of the (s1, 2 sz2, 2 sz3) matrix: phase r_t from the LCG
// jump-ahead seed' * 16807^(t+1) (r8vec_uniform_01(size, seed+1000000), :78), e^{2 PI i r} rounded
// to single precision (cmplx without KIND, :80), planes 1 and the last y/z planes zeroed
// (:81-85); for ii >= 2 the mirrored quadrants (:99-113) take sqrt(exp(-sum (n_d-1)^2 de_d)) of
// their source index; everything times sqrt(8) and 2 sqrt(prod dk) (:116-117).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mulmod_m31_r(uint64_t a, uint64_t b) {
    const uint64_t m = 2147483647ull;
    uint64_t p = a * b;
    p = (p & m) + (p >> 31);
    p = (p & m) + (p >> 31);
    return p >= m ? p - m : p;
}
__device__ __forceinline__ uint64_t powmod_m31_r(uint64_t base, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mulmod_m31_r(r, base);
        base = mulmod_m31_r(base, base);
        e >>= 1;
    }
    return r;
}

struct SkParams {
    int s1, sz2, sz3;      // matrix is (s1, 2 sz2, 2 sz3)
    int gaussian;          // 1: Gaussian branch; 0: phases only (von Karman branch is commented out)
    int32_t seed;          // seed + 1000000
    double de[3];          // (PI^2/2) dk_d^2
    double scale;          // 2 sqrt(prod dk) (applied after sqrt(8))
};

__global__ void k_build_sk(SkParams P, cplx<double>* __restrict__ sk) {
    const int s2 = 2 * P.sz2, s3 = 2 * P.sz3;
    const int64_t tot = (int64_t)P.s1 * s2 * s3;
    const int64_t s0 = ((int64_t)P.seed % 2147483647ll + 2147483647ll) % 2147483647ll;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % P.s1);
        const int64_t r = t / P.s1;
        const int j = (int)(r % s2), k = (int)(r / s2);
        const uint64_t sd = mulmod_m31_r((uint64_t)s0, powmod_m31_r(16807ull, (uint64_t)(t + 1)));
        const double u = (double)(int32_t)sd * 4.656612875e-10;
        const double th = 2.0 * kPiRef * u;
        double re = (double)(float)cos(th), im = (double)(float)sin(th);
        if (i == 0 || j == 0 || k == 0 || j == s2 - 1 || k == s3 - 1) re = im = 0.0;
        if (P.gaussian) {
            const int ii = i + 1, jj1 = j + 1, kk1 = k + 1;  // 1-based
            int jj = -1, kk = -1;
            if (jj1 >= 2 && jj1 <= P.sz2) jj = jj1;
            else if (jj1 >= P.sz2 + 1 && jj1 <= s2 - 1) jj = s2 - jj1 + 1;
            if (kk1 >= 2 && kk1 <= P.sz3) kk = kk1;
            else if (kk1 >= P.sz3 + 1 && kk1 <= s3 - 1) kk = s3 - kk1 + 1;
            if (ii >= 2 && jj >= 2 && kk >= 2) {
                const double a = (double)((ii - 1) * (ii - 1)) * P.de[0] + (double)((jj - 1) * (jj - 1)) * P.de[1] +
                                 (double)((kk - 1) * (kk - 1)) * P.de[2];
                const double v = sqrt(exp(-a));
                re *= v;
                im *= v;
            }
            const double r8 = sqrt(8.0);
            re = P.scale * (r8 * re);
            im = P.scale * (r8 * im);
        }
        sk[t] = {re, im};
    }
}

// Shinozuka consumer (SRC_F/fftw_ScaRL.f90:157-255): place the y/z quadrants of Sk into the
// zero-padded (c1, c2, c3) grid (x entries 1..sz1/2, :217-233).
struct ShinoPlace {
    int s1, s2, s3;        // Sk extents
    int hx, h2, h3;        // copied x extent (s1/2), half y, half z
    int c1, c2, c3;        // padded grid
};
__global__ void k_shino_place(ShinoPlace P, const cplx<double>* __restrict__ sk, cplx<double>* __restrict__ c) {
    const int64_t tot = (int64_t)P.c1 * P.c2 * P.c3;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % P.c1);
        const int64_t r = t / P.c1;
        const int jd = (int)(r % P.c2), kd = (int)(r / P.c2);
        cplx<double> v = {0.0, 0.0};
        int j = -1, k = -1;
        if (jd < P.h2) j = jd;
        else if (jd >= P.c2 - (P.s2 - P.h2)) j = jd - (P.c2 - P.s2);
        if (kd < P.h3) k = kd;
        else if (kd >= P.c3 - (P.s3 - P.h3)) k = kd - (P.c3 - P.s3);
        if (i < P.hx && j >= 0 && k >= 0) v = sk[i + (int64_t)P.s1 * (j + (int64_t)P.s2 * k)];
        c[t] = v;
    }
}

// x pass of the padded C2C backward transform; real part of the first (np1, np2, np3) points.
__global__ __launch_bounds__(kBlock) void k_c2c_xcrop(const PlanDev* __restrict__ plans, int plan,
                                                       const cplx<double>* __restrict__ pool,
                                                       const cplx<double>* __restrict__ c, int c1, int c2,
                                                       int np1, int np2, int np3, double* __restrict__ out) {
    HIP_DYNAMIC_SHARED(char, smem)
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem);
    const PlanDev& pl = plans[plan];
    const int64_t nrows = (int64_t)np2 * np3;
    const int64_t r0 = (int64_t)blockIdx.x * pl.T;
    const int npen = (int)min((int64_t)pl.T, nrows - r0);
    struct Ctx {
        int64_t in, out;
    };
    auto prep = [&](int p) -> Ctx {
        const int64_t P = r0 + p;
        const int64_t j = P % np2, k = P / np2;
        return {(j + (int64_t)c2 * k) * c1, (int64_t)np1 * P};
    };
    auto gld = [&](const Ctx& x, int n) -> cplx<double> { return c[x.in + n]; };
    auto gst = [&](const Ctx& x, int n, cplx<double> v) {
        if (n < np1) out[x.out + n] = v.x;
    };
    pencil_fft<double, LAYOUT_BF, false, true, kEcap64>(pl, pool, lds, npen, prep, gld, gst);
}

template <int AXIS, class F>
static void r_go(const FieldR& f, const PlanDev* plans, const cplx<double>* p, int nblocks, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_r_pass<AXIS, F>), dim3(nblocks), dim3(kBlock), lds, s, f, plans, p);
}

template <int AXIS>
static void r_dispatch(int spec, const FieldR& f, const PlanDev* plans, const cplx<double>* p, int nblocks,
                       size_t lds, hipStream_t s) {
    constexpr int LAY = AXIS == 0 ? LAYOUT_BF : LAYOUT_PF;
    switch (spec) {
#define SCARL_CASE(ID, P, L, N, R0, R1, TPP)                                                  \
    case ID:                                                                                  \
        if constexpr (P == 8 && L == LAY) r_go<AXIS, FixedFft<N, R0, R1, TPP>>(f, plans, p, nblocks, lds, s); \
        break;
        SCARL_FIXED_SPECS(SCARL_CASE)
#undef SCARL_CASE
    case kSpecNonSmooth: r_go<AXIS, NonSmoothFft<kEcap64>>(f, plans, p, nblocks, lds, s); break;
    case kSpecGenericComp: r_go<AXIS, GenericCompFft<kEcap64>>(f, plans, p, nblocks, lds, s); break;
    default: r_go<AXIS, GenericFft<kEcap64>>(f, plans, p, nblocks, lds, s);
    }
}

hipError_t launch_r_pass(int axis, int spec, const FieldR& f, const PlanDev* plans, const void* pool, int nblocks,
                         size_t lds_bytes, hipStream_t s) {
    const cplx<double>* p = static_cast<const cplx<double>*>(pool);
    if (nblocks <= 0) return hipSuccess;
    if (axis == 0) r_dispatch<0>(spec, f, plans, p, nblocks, lds_bytes, s);
    else if (axis == 1) r_dispatch<1>(spec, f, plans, p, nblocks, lds_bytes, s);
    else r_dispatch<2>(spec, f, plans, p, nblocks, lds_bytes, s);
    return hipGetLastError();
}


hipError_t launch_build_sk(const void* params, void* sk, int64_t tot, hipStream_t s) {
    const SkParams P = *static_cast<const SkParams*>(params);
    const int64_t nb = (tot + 255) / 256 < 8192 ? (tot + 255) / 256 : 8192;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_build_sk, dim3((unsigned)nb), dim3(256), 0, s, P, static_cast<cplx<double>*>(sk));
    return hipGetLastError();
}
hipError_t launch_shino_place(const int* dims9, const void* sk, void* c, hipStream_t s) {
    ShinoPlace P = {dims9[0], dims9[1], dims9[2], dims9[3], dims9[4], dims9[5], dims9[6], dims9[7], dims9[8]};
    const int64_t tot = (int64_t)P.c1 * P.c2 * P.c3;
    const int64_t nb = (tot + 255) / 256 < 8192 ? (tot + 255) / 256 : 8192;
    hipLaunchKernelGGL(k_shino_place, dim3((unsigned)nb), dim3(256), 0, s, P, static_cast<const cplx<double>*>(sk),
                       static_cast<cplx<double>*>(c));
    return hipGetLastError();
}
hipError_t launch_c2c_xcrop(const PlanDev* plans, int plan, int T, const void* pool, const void* c, int c1, int c2,
                            int np1, int np2, int np3, double* out, size_t lds, hipStream_t s) {
    const int64_t rows = (int64_t)np2 * np3;
    const int64_t nb = (rows + T - 1) / T;
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_c2c_xcrop, dim3((unsigned)nb), dim3(kBlock), lds, s, plans, plan,
                       static_cast<const cplx<double>*>(pool), static_cast<const cplx<double>*>(c), c1, c2, np1, np2,
                       np3, out);
    return hipGetLastError();
}
}  // namespace scarl
