// scarl_k_misc.hip -- reductions, element-wise operations, LCG jump-ahead, K_nu (gfx950).
#include "scarl_fft.h"

namespace scarl {

// Per-field reduction of tile partials in kMomChunks blocks per field -> field_mom[f * kMomChunks + g];
// then per slot -> slot_out[slot]
// slot_out[slot] = {sum, sum2, count, mean, std} with the reference's std formula
// sqrt(sum2/n - mean^2) (SRC_F/Stat_ScaRL/statistics_RF.f90:80-82).
__global__ __launch_bounds__(kBlock) void k_mom_field(const FieldH* __restrict__ fields,
                                                       const int2* __restrict__ field_tiles,
                                                       const MomPartial* __restrict__ partials,
                                                       MomPartial* __restrict__ field_mom) {
    __shared__ double red[4 * kBlock];
    const int fi = blockIdx.x, g = blockIdx.y;
    const FieldH& f = fields[fi];
    double s = 0, c = 0, s2 = 0, c2 = 0;
    if (f.mom_slot >= 0) {
        const int2 r = field_tiles[fi];
        for (int t = r.x + g * kBlock + threadIdx.x; t < r.y; t += kBlock * kMomChunks) {
            const MomPartial m = partials[t];
            nadd(s, c, m.s);
            c += m.c;
            nadd(s2, c2, m.s2);
            c2 += m.c2;
        }
    }
    block_reduce_mom(s, c, s2, c2, red, field_mom + fi * kMomChunks + g);
}

__global__ __launch_bounds__(kBlock) void k_mom_slot(const FieldH* __restrict__ fields, int nfields,
                                                      const MomPartial* __restrict__ field_mom,
                                                      const double* __restrict__ slot_count, int nslots,
                                                      double* __restrict__ slot_out) {
    __shared__ double red[4 * kBlock];
    const int sl = blockIdx.x;
    double s = 0, c = 0, s2 = 0, c2 = 0;
    for (int i = threadIdx.x; i < nfields * kMomChunks; i += kBlock) {
        if (fields[i / kMomChunks].mom_slot != sl) continue;
        const MomPartial m = field_mom[i];
        nadd(s, c, m.s);
        c += m.c;
        nadd(s2, c2, m.s2);
        c2 += m.c2;
    }
    __shared__ MomPartial res;
    block_reduce_mom(s, c, s2, c2, red, &res);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double n = slot_count[sl];
        const double S = res.s + res.c, S2 = res.s2 + res.c2;
        const double mean = S / n;
        slot_out[5 * sl + 0] = S;
        slot_out[5 * sl + 1] = S2;
        slot_out[5 * sl + 2] = n;
        slot_out[5 * sl + 3] = mean;
        slot_out[5 * sl + 4] = sqrt(S2 / n - mean * mean);
    }
}

// FP64 sum over the FULL spectrum of g(kappa) (Gaussian: exp(-q/4); Matern: (1+q)^(2 expo)),
// half spectrum with multiplicity 2 off the self-conjugate planes.  One thread per (kx, ky)
// column, loop over kz; Neumaier partials per block -> out[blockIdx.x].
__global__ __launch_bounds__(kBlock) void k_spectrum_sum(FieldH f, MomPartial* __restrict__ out) {
    __shared__ double red[4 * kBlock];
    const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int n1 = f.n[0], n2 = f.n[1], n3 = f.n[2], h1 = f.h1;
    double s = 0, c = 0;
    if (col < (int64_t)h1 * n2) {
        const int kx = (int)(col % h1), ky = (int)(col / h1);
        const double mult = (kx == 0 || ((n1 & 1) == 0 && kx == n1 / 2)) ? 1.0 : 2.0;
        const double a0 = f.alpha[0] * kx, a1 = f.alpha[1] * signed_freq(ky, n2);
        for (int kz = 0; kz < n3; ++kz) {
            const double a2 = f.alpha[2] * signed_freq(kz, n3);
            const double q = a0 * a0 + a1 * a1 + a2 * a2;
            const double g = (f.model == CM_GAUSSIAN) ? exp(-0.25 * q) : pow(1.0 + q, 2.0 * f.expo);
            nadd(s, c, mult * g);
        }
    }
    block_reduce_mom(s, c, 0.0, 0.0, red, out + blockIdx.x);
}

// Target correlation C(r)/sigma^2 (FP64, K_nu for von Karman): r[i] are anisotropic lags
// already divided by L, as rho values.
__global__ void k_target_corr(int model, double nu, const double* __restrict__ rho, int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = rho[i];
    double v;
    if (model == CM_GAUSSIAN) v = exp(-r * r);
    else if (model == CM_EXPONENTIAL) v = exp(-r);
    else v = (r == 0.0) ? 1.0 : pow(2.0, 1.0 - nu) / gamma_fp64(nu) * pow(r, nu) * kv_fp64(nu, r);
    out[i] = v;
}

__global__ void k_kv(double nu, const double* __restrict__ x, int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = kv_fp64(nu, x[i]);
}

// ------------------------------------------------------------------------------------------
// Standalone element-wise operations (FP64, device arrays)
// ------------------------------------------------------------------------------------------
// multiVariateTransformation (SRC_F/main_ScaRL.f90:1355-1395)
__global__ void k_marginal(double* __restrict__ x, int64_t n, int margi, double mu, double sd) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = x[i] * sd + mu;
        x[i] = (margi == FOM_LOGNORMAL) ? exp(v) : v;
    }
}

// apply_UnityPartition_mtx (SRC_F/main_ScaRL.f90:1216-1301): f *= sqrt(prod_d w_d). A block
// walks whole x rows, so the y and z weights are formed once per row and no point needs a 64-bit
// division; the product keeps the order (w_x w_y) w_z.
__global__ void k_pou(double* __restrict__ x, int n1, int n2, int n3, int3 lo, int3 hi) {
    const int64_t rows = (int64_t)n2 * n3;
    // rows of up to 4 * blockDim points: a thread visits the same x positions in every row, so their
    // weights are formed once and kept in registers
    constexpr int kRegX = 4;
    const bool cached = n1 <= kRegX * (int)blockDim.x;
    double wxr[kRegX];
#pragma unroll
    for (int m = 0; m < kRegX; ++m) {
        const int i = threadIdx.x + m * blockDim.x;
        wxr[m] = (cached && i < n1) ? pou_sqrt_weight<double>(i, n1, lo.x, hi.x) : 0.0;
    }
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const int j = (int)(r % n2), k = (int)(r / n2);
        const double wy = pou_sqrt_weight<double>(j, n2, lo.y, hi.y), wz = pou_sqrt_weight<double>(k, n3, lo.z, hi.z);
        double* row = x + r * n1;
        if (cached) {
#pragma unroll
            for (int m = 0; m < kRegX; ++m) {
                const int i = threadIdx.x + m * blockDim.x;
                if (i < n1) row[i] *= sqrt(wxr[m] * wy * wz);
            }
        } else {
            for (int i = threadIdx.x; i < n1; i += blockDim.x) {
                const double w = pou_sqrt_weight<double>(i, n1, lo.x, hi.x) * wy * wz;
                row[i] *= sqrt(w);
            }
        }
    }
}

// Compensated per-event sums of x and x^2: grid (blocks, nmc); partial per block
__global__ __launch_bounds__(kBlock) void k_sums(const double* __restrict__ rf, int64_t n,
                                                  MomPartial* __restrict__ partial) {
    __shared__ double red[4 * kBlock];
    const int e = blockIdx.y;
    const double* x = rf + (int64_t)e * n;
    double s = 0, c = 0, s2 = 0, c2 = 0;
    const int64_t G = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    // four loads in flight, added in the order of the grid-stride visit
    for (; i + 3 * G < n; i += 4 * G) {
        const double v0 = x[i], v1 = x[i + G], v2 = x[i + 2 * G], v3 = x[i + 3 * G];
        nadd(s, c, v0);
        nadd(s2, c2, v0 * v0);
        nadd(s, c, v1);
        nadd(s2, c2, v1 * v1);
        nadd(s, c, v2);
        nadd(s2, c2, v2 * v2);
        nadd(s, c, v3);
        nadd(s2, c2, v3 * v3);
    }
    for (; i < n; i += G) {
        const double v = x[i];
        nadd(s, c, v);
        nadd(s2, c2, v * v);
    }
    block_reduce_mom(s, c, s2, c2, red, partial + (int64_t)e * gridDim.x + blockIdx.x);
}

// LCG jump-ahead: seed_i = seed0 * 16807^(i+1) mod (2^31 - 1), r_i = seed_i * 4.656612875e-10
// (SRC_F/normal_ScaRL.f90:1497-1512); powers by binary exponentiation on 64-bit products.
__device__ __forceinline__ uint64_t mulmod_m31(uint64_t a, uint64_t b) {
    const uint64_t m = 2147483647ull;
    uint64_t p = a * b;                   // < 2^62
    p = (p & m) + (p >> 31);
    p = (p & m) + (p >> 31);
    return p >= m ? p - m : p;
}
__device__ __forceinline__ uint64_t powmod_m31(uint64_t base, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mulmod_m31(r, base);
        base = mulmod_m31(base, base);
        e >>= 1;
    }
    return r;
}
__global__ void k_lcg_uniform(int32_t seed0, int64_t n, double* __restrict__ r, int32_t* __restrict__ seeds) {
    const int64_t G = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // Schrage's step keeps seed_i == seed0 * 16807^(i+1) (mod 2^31-1) for any int32 seed0. A thread
    // forms its first seed by exponentiation and each later one (i + G) with one product by 16807^G.
    const int64_t s0 = ((int64_t)seed0 % 2147483647ll + 2147483647ll) % 2147483647ll;
    uint64_t s = mulmod_m31((uint64_t)s0, powmod_m31(16807ull, (uint64_t)(i + 1)));
    const uint64_t step = powmod_m31(16807ull, (uint64_t)G);
    for (; i < n; i += G) {
        r[i] = (double)(int32_t)s * 4.656612875e-10;
        if (seeds) seeds[i] = (int32_t)s;
        s = mulmod_m31(s, step);
    }
}


// ---- lag products (per-axis covariance and the corrL estimator) -------------------------------
// Taper of the corrL estimator as written in SRC_F/Stat_ScaRL/statistics_RF.f90:326-351: r is the
// default-real literal 0.1 widened to double, PI the literal 3.1415926535898, and every ratio
// above r takes the second cosine branch.
__device__ __forceinline__ double stat_taper(int m, int n) {
    const double r = 0.100000001490116119384765625, pi = 3.1415926535898;
    const double ratio = n > 1 ? (double)m / (double)(n - 1) : 0.0;
    if (ratio < r) return 0.5 * (1.0 + cos(2.0 * pi / r * (ratio - r / 2.0)));
    if (ratio > r) return 0.5 * (1.0 + cos(2.0 * pi / r * (ratio - 1.0 + r / 2.0)));
    return 1.0;
}

// One block: J.T pencils along J.axis staged in LDS as u[m * T + p] (FP64), then thread (p, g)
// sums lags l = g, g + G, ... of pencil p in plain FP64 (at most a pencil length of terms); the
// block's sum over its pencils goes to J.partial[block][l].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lag_partial(const LagJob* __restrict__ jobs,
                                                         const LagBlock* __restrict__ blocks) {
    HIP_DYNAMIC_SHARED(double, u)
    const LagBlock bl = blocks[blockIdx.x];
    const LagJob& J = jobs[bl.job];
    const int ax = J.axis, n = J.n[ax], TT = J.T, nlag = J.nlag;
    const int oa = ax == 0 ? 1 : 0, ob = ax == 2 ? 1 : 2;  // the other axes, oa fastest
    const int64_t st[3] = {1, J.ld1, J.ld2};
    const int64_t npen = (int64_t)J.n[oa] * J.n[ob];
    const int np = (int)min((int64_t)TT, npen - bl.pencil0);
    const T* __restrict__ x = static_cast<const T*>(J.x);
    const double c = J.centre ? *J.centre : J.cval;
    __shared__ int64_t pbase[32];  // element 0 of each pencil of the block
    if (threadIdx.x < TT) {
        const int64_t P = bl.pencil0 + threadIdx.x;
        pbase[threadIdx.x] = (P % J.n[oa]) * st[oa] + (P / J.n[oa]) * st[ob];
    }
    __syncthreads();
    auto put = [&](int p, int m) {
        double v = 0.0;
        if (p < np) {
            v = ((double)x[pbase[p] + (int64_t)m * st[ax]] - c) * J.scale;
            if (J.window) v *= stat_taper(m, n);
        }
        u[m * TT + p] = v;
    };
    if (ax == 0) {  // rows are contiguous: consecutive threads walk a row
        for (int p = 0; p < TT; ++p)
            for (int m = threadIdx.x; m < n; m += kBlock) put(p, m);
    } else {        // neighbouring pencils are neighbours in x
        const int G = kBlock / TT, p = threadIdx.x % TT;
        for (int m = threadIdx.x / TT; m < n && threadIdx.x < G * TT; m += G) put(p, m);
    }
    __syncthreads();
    double* red = u + (int64_t)n * TT;  // [l][p]
    const int G = kBlock / TT, p = threadIdx.x % TT, g = threadIdx.x / TT;
    for (int l = g; l < nlag; l += G) {
        double acc = 0.0;
        if (l < n) {
            const int mend = n - l;
            for (int m = 0; m < mend; ++m) acc += u[m * TT + p] * u[(m + l) * TT + p];
            if (J.circular)
                for (int m = mend; m < n; ++m) acc += u[m * TT + p] * u[(m + l - n) * TT + p];
        }
        red[l * TT + p] = acc;
    }
    __syncthreads();
    const int64_t bj = bl.pencil0 / TT;
    for (int l = threadIdx.x; l < nlag; l += kBlock) {
        double s = 0.0;
        for (int q = 0; q < TT; ++q) s += red[l * TT + q];
        J.partial[bj * nlag + l] = s;
    }
}

// The same partial sums by FFT, as the reference estimates them (SRC_F/Stat_ScaRL/statistics_RF.f90:
// 353-448): a block holds 2 T real pencils (T = the plan's pencils per tile) as T complex pencils
// z = u_a + i u_b of length P in LDS (zero padded past n when the pairs are bounded), transforms them,
// forms W_k = (|X_a,k|^2, |X_b,k|^2) from Z_k and Z_{P-k} (each pair handled by one thread), and
// transforms W: Re and Im of the result at l are P times the lag-l sums of u_a and u_b (circular over
// P, which equals the bounded sum when P >= n + nlag - 1).  FP64 throughout.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lag_fft(const LagJob* __restrict__ jobs, const LagBlock* __restrict__ blocks,
                                                     const PlanDev* __restrict__ plans,
                                                     const cplx<double>* __restrict__ pool) {
    HIP_DYNAMIC_SHARED(char, smem)
    cplx<double>* L = reinterpret_cast<cplx<double>*>(smem);
    double* Ld = reinterpret_cast<double*>(smem);
    const LagBlock bl = blocks[blockIdx.x];
    const LagJob& J = jobs[bl.job];
    const PlanDev& pl = plans[J.fft_plan];
    const int ax = J.axis, n = J.n[ax], nlag = J.nlag, P = J.P;
    const int Tc = pl.T, TR = 2 * Tc, tpp = pl.tpp;
    const int oa = ax == 0 ? 1 : 0, ob = ax == 2 ? 1 : 2;  // the other axes, oa fastest
    const int64_t st[3] = {1, J.ld1, J.ld2};
    const int64_t npen = (int64_t)J.n[oa] * J.n[ob];
    const int np = (int)min((int64_t)TR, npen - bl.pencil0);
    const T* __restrict__ x = static_cast<const T*>(J.x);
    const double c = J.centre ? *J.centre : J.cval;
    __shared__ int64_t pbase[2 * kBlock];  // element 0 of each real pencil of the block (TR <= 2 kBlock)
    for (int t = threadIdx.x; t < TR; t += kBlock) {
        const int64_t Pn = bl.pencil0 + t;
        pbase[t] = (Pn % J.n[oa]) * st[oa] + (Pn / J.n[oa]) * st[ob];
    }
    __syncthreads();
    auto put = [&](int rp, int m) {
        double v = 0.0;
        if (rp < np) {
            v = ((double)x[pbase[rp] + (int64_t)m * st[ax]] - c) * J.scale;
            if (J.window) v *= stat_taper(m, n);
        }
        Ld[2 * (m * Tc + (rp >> 1)) + (rp & 1)] = v;
    };
    if (ax == 0) {  // rows are contiguous: consecutive threads walk a row
        for (int rp = 0; rp < TR; ++rp)
            for (int m = threadIdx.x; m < n; m += kBlock) put(rp, m);
    } else if (TR <= kBlock) {  // neighbouring pencils are neighbours in x
        const int G = kBlock / TR, rp = threadIdx.x % TR;
        for (int m = threadIdx.x / TR; m < n && threadIdx.x < G * TR; m += G) put(rp, m);
    } else {
        for (int idx = threadIdx.x; idx < n * TR; idx += kBlock) put(idx % TR, idx / TR);
    }
    for (int idx = threadIdx.x; idx < (P - n) * Tc; idx += kBlock) L[n * Tc + idx] = cplx<double>{0.0, 0.0};
    __syncthreads();
    const int p = threadIdx.x % Tc, b = threadIdx.x / Tc;
    const bool active = b < tpp;
    auto ld_l = [&](int k) -> cplx<double> { return L[k * Tc + p]; };
    auto st_l = [&](int k, cplx<double> v) { L[k * Tc + p] = v; };
    for (int pass = 0; pass < 2; ++pass) {
        for (int s = 0; s < pl.nst; ++s) {
            stage_dispatch<kEcap64, double, true>(pl.R[s], P, pl.Ns[s], pool + pl.tw_off[s], b, tpp, active, true,
                                                  ld_l, st_l);
            __syncthreads();
        }
        if (pass == 0) {
            if (active) {
                for (int k = b; 2 * k <= P; k += tpp) {
                    const int kk = k ? P - k : 0;
                    const cplx<double> za = ld_l(k), zb = ld_l(kk);
                    const double sr = za.x + zb.x, si = za.y - zb.y;  // Z_k + conj Z_{-k} = 2 X_a
                    const double dr = za.x - zb.x, di = za.y + zb.y;  // Z_k - conj Z_{-k} = 2i X_b
                    const cplx<double> w = {0.25 * (sr * sr + si * si), 0.25 * (dr * dr + di * di)};
                    st_l(k, w);
                    st_l(kk, w);
                }
            }
            __syncthreads();
        }
    }
    const int64_t bj = bl.pencil0 / TR;
    const double inv = 1.0 / (double)P;
    for (int l = threadIdx.x; l < nlag; l += kBlock) {
        double s = 0.0;
        if (l < n)
            for (int q = 0; q < Tc; ++q) {
                const cplx<double> v = L[l * Tc + q];
                s += v.x + v.y;
            }
        J.partial[bj * nlag + l] = s * inv;
    }
}

// out[j][l] = (Neumaier sum over the job's blocks) / (pairs per pencil * pencils); undefined lags
// (l >= n) give NaN.  Grid (nlag_max, njobs).
__global__ __launch_bounds__(kBlock) void k_lag_finish(const LagJob* __restrict__ jobs, int nlag_max,
                                                        double* __restrict__ out) {
    __shared__ double red[4 * kBlock];
    __shared__ MomPartial res;
    const LagJob& J = jobs[blockIdx.y];
    const int l = blockIdx.x;
    if (l >= J.nlag) return;
    double s = 0, cs = 0, z = 0, cz = 0;
    for (int64_t b = threadIdx.x; b < J.nblocks; b += kBlock) nadd(s, cs, J.partial[b * J.nlag + l]);
    block_reduce_mom(s, cs, z, cz, red, &res);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int n = J.n[J.axis];
        const double npen = (double)J.n[0] * J.n[1] * J.n[2] / n;
        const double pairs = J.circular ? (double)n : (double)(n - l);
        out[(int64_t)blockIdx.y * nlag_max + l] = (l < n) ? (res.s + res.c) / (pairs * npen) : __builtin_nan("");
    }
}

// Compensated sums over the box [0, box) of (x - centre) and (x - centre)^2 -> partial[blockIdx.x].
__global__ __launch_bounds__(kBlock) void k_box_sums(const double* __restrict__ x, int n1, int n2, int b1, int b2,
                                                     int b3, double centre, MomPartial* __restrict__ partial) {
    __shared__ double red[4 * kBlock];
    double s = 0, c = 0, s2 = 0, c2 = 0;
    // a block walks whole x rows of the box: one 64-bit division per row, not per point
    const int64_t rows = (int64_t)b2 * b3;
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const int64_t b = r % b2, k = r / b2;
        const double* row = x + (int64_t)n1 * (b + (int64_t)n2 * k);
        for (int a = threadIdx.x; a < b1; a += kBlock) {
            const double v = row[a] - centre;
            nadd(s, c, v);
            nadd(s2, c2, v * v);
        }
    }
    block_reduce_mom(s, c, s2, c2, red, partial + blockIdx.x);
}

// x <- (x - a) / b over the whole array (two steps, as SRC_F/main_ScaRL.f90:930,938)
__global__ void k_normalize(double* __restrict__ x, int64_t n, double a, double b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (x[i] - a) / b;
}

// blocks[0, nblocks) run the direct products (LDS lds), fblocks[0, nfblocks) the FFT path (LDS flds,
// plans and pool of FP64 complex twiddles); then the compensated combination of every job.
hipError_t launch_lag(bool fp64, const LagJob* jobs, int njobs, const LagBlock* blocks, int64_t nblocks,
                      size_t lds, const LagBlock* fblocks, int64_t nfblocks, size_t flds, const PlanDev* plans,
                      const void* pool, int nlag_max, double* out, hipStream_t s) {
    if (nblocks > 0) {
        if (fp64) hipLaunchKernelGGL(k_lag_partial<double>, dim3((unsigned)nblocks), dim3(kBlock), lds, s, jobs, blocks);
        else hipLaunchKernelGGL(k_lag_partial<float>, dim3((unsigned)nblocks), dim3(kBlock), lds, s, jobs, blocks);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nfblocks > 0) {
        const cplx<double>* p = static_cast<const cplx<double>*>(pool);
        if (fp64)
            hipLaunchKernelGGL(k_lag_fft<double>, dim3((unsigned)nfblocks), dim3(kBlock), flds, s, jobs, fblocks, plans, p);
        else
            hipLaunchKernelGGL(k_lag_fft<float>, dim3((unsigned)nfblocks), dim3(kBlock), flds, s, jobs, fblocks, plans, p);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_lag_finish, dim3(nlag_max, njobs), dim3(kBlock), 0, s, jobs, nlag_max, out);
    return hipGetLastError();
}

hipError_t launch_box_sums(const double* x, const int n[3], const int box[3], double centre, MomPartial* partial,
                           int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_box_sums, dim3(nblocks), dim3(kBlock), 0, s, x, n[0], n[1], box[0], box[1], box[2], centre,
                       partial);
    return hipGetLastError();
}

hipError_t launch_normalize(double* x, int64_t n, double a, double b, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192;
    hipLaunchKernelGGL(k_normalize, dim3((unsigned)nb), dim3(256), 0, s, x, n, a, b);
    return hipGetLastError();
}

hipError_t launch_mom(const FieldH* fields, int nfields, const int2* field_tiles, const MomPartial* partials,
                      MomPartial* field_mom, const double* slot_count, int nslots, double* slot_out, hipStream_t s) {
    if (nfields <= 0 || nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mom_field, dim3(nfields, kMomChunks), dim3(kBlock), 0, s, fields, field_tiles, partials,
                       field_mom);
    hipLaunchKernelGGL(k_mom_slot, dim3(nslots), dim3(kBlock), 0, s, fields, nfields, field_mom, slot_count, nslots,
                       slot_out);
    return hipGetLastError();
}

hipError_t launch_spectrum_sum(const FieldH& f, MomPartial* out, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_spectrum_sum, dim3(nblocks), dim3(kBlock), 0, s, f, out);
    return hipGetLastError();
}

hipError_t launch_marginal(double* x, int64_t n, int margi, double mu, double sd, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192;
    hipLaunchKernelGGL(k_marginal, dim3((unsigned)nb), dim3(256), 0, s, x, n, margi, mu, sd);
    return hipGetLastError();
}

hipError_t launch_pou(double* x, int n1, int n2, int n3, int3 lo, int3 hi, hipStream_t s) {
    const int64_t n = (int64_t)n1 * n2 * n3;
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192;
    hipLaunchKernelGGL(k_pou, dim3((unsigned)nb), dim3(256), 0, s, x, n1, n2, n3, lo, hi);
    return hipGetLastError();
}

hipError_t launch_sums(const double* rf, int64_t n, int nmc, int nblocks, MomPartial* partial, hipStream_t s) {
    hipLaunchKernelGGL(k_sums, dim3(nblocks, nmc), dim3(kBlock), 0, s, rf, n, partial);
    return hipGetLastError();
}

hipError_t launch_lcg_uniform(int32_t seed0, int64_t n, double* r, int32_t* seeds, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
    hipLaunchKernelGGL(k_lcg_uniform, dim3((unsigned)nb), dim3(256), 0, s, seed0, n, r, seeds);
    return hipGetLastError();
}

hipError_t launch_target_corr(int model, double nu, const double* rho, int n, double* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_target_corr, dim3((n + 127) / 128), dim3(128), 0, s, model, nu, rho, n, out);
    return hipGetLastError();
}

hipError_t launch_kv(double nu, const double* x, int n, double* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kv, dim3((n + 127) / 128), dim3(128), 0, s, nu, x, n, out);
    return hipGetLastError();
}

}  // namespace scarl
