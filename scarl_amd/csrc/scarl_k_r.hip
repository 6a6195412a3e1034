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

__device__ __forceinline__ double mode_r_amp(const FieldR& f, int i, int j, int k) {
    // SRC_F/fftw_ScaRL.f90:305-370 (kPoints, Sk product in axis order, sqrt(S) cos(2 PI phi))
    const int64_t pidx = i + (int64_t)f.m[0] * (j + (int64_t)f.m[1] * k);
    const double ph = (f.src_mode == SRC_HOST_NOISE) ? f.phi[pidx]
                                                     : uniform53(static_cast<uint64_t>(pidx), f.seed, f.stream);
    const double k0 = ((double)i * f.dk[0]) / f.corrL[0];
    const double k1 = ((double)j * f.dk[1]) / f.corrL[1];
    const double k2 = ((double)k * f.dk[2]) / f.corrL[2];
    double S = 1.0;
    S = S * f.corrL[0] * exp(-kPiRef * (k0 * k0));
    S = S * f.corrL[1] * exp(-kPiRef * (k1 * k1));
    S = S * f.corrL[2] * exp(-kPiRef * (k2 * k2));
    return 1.0 * sqrt(S) * cos(2.0 * kPiRef * ph);
}

template <int AXIS>
__global__ __launch_bounds__(kBlock) void k_r_pass(const FieldR* __restrict__ fp, const PlanDev* __restrict__ plans,
                                                    const cplx<double>* __restrict__ pool) {
    extern __shared__ char smem[];
    cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem);
    const FieldR& f = *fp;
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

    auto rin = [&](int64_t P, int j) -> double {
        if (j >= m) return 0.0;
        if constexpr (AXIS == 0) {
            const int jy = (int)(P % m2), kz = (int)(P / m2);
            return mode_r_amp(f, j, jy, kz);
        } else if constexpr (AXIS == 1) {
            const int64_t i = P % n1, k = P / n1;
            return f.tmp1[i + (int64_t)n1 * (j + (int64_t)m2 * k)];
        } else {
            return f.tmp2[P + (int64_t)n1 * n2 * j];
        }
    };
    auto rout = [&](int64_t P, int o, double v) {
        if constexpr (AXIS == 0) {
            f.tmp1[o + (int64_t)n1 * P] = v;
        } else if constexpr (AXIS == 1) {
            const int64_t i = P % n1, k = P / n1;
            f.tmp2[i + (int64_t)n1 * (o + (int64_t)n2 * k)] = v;
        } else {
            double r = v;
            if (f.margi == FOM_GAUSSIAN) r = r * f.sd + f.mu;
            else if (f.margi == FOM_LOGNORMAL) r = exp(r * f.sd + f.mu);
            f.out[P + (int64_t)n1 * n2 * o] = r;
        }
    };
    auto gld = [&](int p, int j) -> cplx<double> {
        const int64_t Pa = 2 * (c0 + p), Pb = Pa + 1;
        const cplx<double> w = pre[j];
        const double xa = rin(Pa, j), xa2 = (j > 0) ? rin(Pa, n - j) : 0.0;
        double xb = 0.0, xb2 = 0.0;
        if (Pb < npen_real) {
            xb = rin(Pb, j);
            xb2 = (j > 0) ? rin(Pb, n - j) : 0.0;
        }
        // V(a) = (xa - i xa2) w ; V(b) likewise ; return V(a) + i V(b)
        const cplx<double> va = {xa * w.x + xa2 * w.y, xa * w.y - xa2 * w.x};
        const cplx<double> vb = {xb * w.x + xb2 * w.y, xb * w.y - xb2 * w.x};
        return {va.x - vb.y, va.y + vb.x};
    };
    auto gst = [&](int p, int k, cplx<double> v) {
        const int64_t Pa = 2 * (c0 + p), Pb = Pa + 1;
        const int o = (2 * k < n) ? 2 * k : 2 * (n - 1 - k) + 1;
        rout(Pa, o, v.x);
        if (Pb < npen_real) rout(Pb, o, v.y);
    };
    if constexpr (AXIS == 0) pencil_fft<double, LAYOUT_BF, true, false>(pl, pool, lds, npen, gld, gst);
    else pencil_fft<double, LAYOUT_PF, false, false>(pl, pool, lds, npen, gld, gst);
}


hipError_t launch_r_pass(int axis, const FieldR* f, const PlanDev* plans, const void* pool, int nblocks,
                         size_t lds_bytes, hipStream_t s) {
    const cplx<double>* p = static_cast<const cplx<double>*>(pool);
    if (nblocks <= 0) return hipSuccess;
    if (axis == 0) hipLaunchKernelGGL(k_r_pass<0>, dim3(nblocks), dim3(kBlock), lds_bytes, s, f, plans, p);
    else if (axis == 1) hipLaunchKernelGGL(k_r_pass<1>, dim3(nblocks), dim3(kBlock), lds_bytes, s, f, plans, p);
    else hipLaunchKernelGGL(k_r_pass<2>, dim3(nblocks), dim3(kBlock), lds_bytes, s, f, plans, p);
    return hipGetLastError();
}

}  // namespace scarl
