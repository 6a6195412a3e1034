/*
 * scarl_oracle.c -- CPU FP64 restatement of ScaRL's spectral random-field hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke entry and the
 * bench's cpu_baseline leg may load it; the product library (scarl_amd/csrc) never does.
 *
 * Layout convention: every 3D array is Fortran column-major, x fastest, exactly as the
 * reference stores k_mtx(Np1,Np2,Np3) (SRC_F/main_ScaRL.f90:405).  Complex arrays are
 * interleaved (re, im) doubles.
 *
 * Parity pins (each function says which pin covers it):
 *   - LCG:            reference known-answer test (SRC_F/normal_ScaRL.f90:762-769) and the
 *                     reference routines compiled from source into oracle/_ref.
 *   - FFT / REDFT01:  O(n^2) long-double brute force and scipy/pocketfft golden vectors.
 *   - Philox4x32-10:  Random123 known-answer vectors.
 *   - K_nu, Gamma:    scipy.special golden vectors, closed forms K_{1/2}, K_{3/2}, and the
 *                     reference's ikv compiled from source into oracle/_ref.
 *   - spectra:        closed forms (C(0)=sigma^2, exponential == Matern nu=1/2).
 *   - mode H field:   numpy.fft.irfftn on the same Hermitian half spectrum.
 *   - PoU / blend:    partition-of-unity invariants and brute-force topologies.
 *   - moments:        closed forms; Kahan vs naive on adversarial data.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* SRC_F/constants_ScaRL.f90:5 -- the reference's truncated PI literal (about 6.8e-15 below pi).
 * Mode R (reference parity) uses it everywhere the reference does. */
#define PI_REF 3.1415926535898
#define PI_EXACT 3.14159265358979323846264338327950288

enum { CM_GAUSSIAN = 1, CM_VON_KARMAN = 2, CM_EXPONENTIAL = 3 };
enum { FOM_GAUSSIAN = 1, FOM_LOGNORMAL = 2 };

/* ------------------------------------------------------------------------------------ */
/* Park-Miller LCG (SRC_F/normal_ScaRL.f90:746-841, 1435-1515)                           */
/* ------------------------------------------------------------------------------------ */

/* One Schrage step, SRC_F/normal_ScaRL.f90:827-838. */
double oracle_r8_uniform_01(int32_t *seed) {
    int32_t k = *seed / 127773;
    int32_t s = 16807 * (*seed - k * 127773) - k * 2836;
    if (s < 0) s += 2147483647;
    *seed = s;
    return (double)s * 4.656612875e-10;
}

/* SRC_F/normal_ScaRL.f90:1497-1512 */
void oracle_r8vec_uniform_01(int32_t n, int32_t *seed, double *r) {
    for (int32_t i = 0; i < n; ++i) r[i] = oracle_r8_uniform_01(seed);
}

/* Box-Muller pairs with the reference's odd-n tail, SRC_F/normal_ScaRL.f90:1186-1254.
 * Returns 1 (like the reference's stop 1) if a zero uniform would be logged. */
int oracle_r8vec_normal_01(int32_t n, int32_t *seed, double *x) {
    const double r8_pi = 3.141592653589793;
    if (n <= 0) return 0;
    if (n == 1) {
        double r1 = oracle_r8_uniform_01(seed);
        if (r1 == 0.0) return 1;
        double r2 = oracle_r8_uniform_01(seed);
        x[0] = sqrt(-2.0 * log(r1)) * cos(2.0 * r8_pi * r2);
        return 0;
    }
    int32_t npair = n / 2;               /* pairs written into x(1..2*npair) */
    int32_t m = (n % 2 == 0) ? npair : npair + 1;
    double *r = (double *)malloc(sizeof(double) * 2 * (size_t)m);
    oracle_r8vec_uniform_01(2 * m, seed, r);
    for (int32_t p = 0; p < npair; ++p) {
        double a = sqrt(-2.0 * log(r[2 * p]));
        x[2 * p] = a * cos(2.0 * r8_pi * r[2 * p + 1]);
        x[2 * p + 1] = a * sin(2.0 * r8_pi * r[2 * p + 1]);
    }
    if (n % 2 == 1) x[n - 1] = sqrt(-2.0 * log(r[2 * m - 2])) * cos(2.0 * r8_pi * r[2 * m - 1]);
    free(r);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Philox4x32-10 counter-based generator (Salmon et al., SC'11).  Pinned by the          */
/* Random123 known-answer vectors in the tests.                                           */
/* ------------------------------------------------------------------------------------ */
void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox_at(uint64_t idx, uint64_t seed, uint32_t stream, uint32_t o[4]) {
    uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), stream, 0u};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    oracle_philox4x32_10(c, k, o);
}

/* 24-bit midpoint uniform in (0,1): exactly representable in float and double. */
static double u24(uint32_t x) { return (double)(x >> 8) * 0x1p-24 + 0x1p-25; }

/* Box-Muller on two 24-bit uniforms. */
static void box_muller(uint32_t a, uint32_t b, double g[2]) {
    double u1 = u24(a), u2 = u24(b);
    double rad = sqrt(-2.0 * log(u1));
    g[0] = rad * cos(2.0 * PI_EXACT * u2);
    g[1] = rad * sin(2.0 * PI_EXACT * u2);
}

/* Standard-normal pair of half-spectrum point (kx, ky, kz). With h3 = ceil(n3/2), the points
 * kz and kz + h3 of one z pencil share the Philox block at counter kx + h1 (ky + n2 (kz mod h3)):
 * outputs (0,1) feed kz < h3, outputs (2,3) feed kz >= h3. */
void oracle_gauss_at(long kx, long ky, long kz, long h1, long n2, long n3, uint64_t seed, uint32_t stream,
                     double g[2]) {
    uint32_t o[4];
    const long h3 = (n3 + 1) / 2, hi = kz >= h3;
    philox_at((uint64_t)(kx + h1 * (ky + n2 * (kz - hi * h3))), seed, stream, o);
    if (hi) box_muller(o[2], o[3], g);
    else box_muller(o[0], o[1], g);
}

/* 53-bit uniform in [0,1) for counter idx (mode R phase phi). */
double oracle_uniform53(uint64_t idx, uint64_t seed, uint32_t stream) {
    uint32_t o[4];
    philox_at(idx, seed, stream, o);
    uint64_t b = ((uint64_t)o[1] << 32) | o[0];
    return (double)(b >> 11) * 0x1p-53;
}

/* ------------------------------------------------------------------------------------ */
/* Complex FFT (recursive mixed-radix decimation in time) and brute-force DFT            */
/* ------------------------------------------------------------------------------------ */
typedef struct { double re, im; } cplx;

static int smallest_factor(int n) {
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* tw[t] = exp(sign*2*pi*i*t/n), computed in long double and rounded once. */
static cplx *make_twiddles(int n, int sign) {
    cplx *tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int t = 0; t < n; ++t) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
        tw[t].re = (double)cosl(a);
        tw[t].im = (double)(sign * sinl(a));
    }
    return tw;
}

/* out[k] = sum_j in[j*stride] w_N^{jk}, where w_N = exp(sign 2 pi i / N); tw is the N-table
 * and tstep = N_total / n selects this level's root from it. */
static void fft_rec(int n, const cplx *in, int stride, cplx *out, const cplx *tw, int tstep, cplx *scratch) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; ++r) fft_rec(m, in + (size_t)r * stride, stride * p, out + (size_t)r * m, tw, tstep * p, scratch);
    /* combine: X[k + q m] = sum_r w_n^{r k} w_p^{r q} Y_r[k] */
    int N = n * tstep; /* length of the twiddle table */
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) {
            cplx y = out[(size_t)r * m + k];
            long t = ((long)r * k % n) * tstep;
            cplx w = tw[t % N];
            scratch[r].re = y.re * w.re - y.im * w.im;
            scratch[r].im = y.re * w.im + y.im * w.re;
        }
        if (p == 2) {
            cplx a = scratch[0], b = scratch[1];
            out[k].re = a.re + b.re; out[k].im = a.im + b.im;
            out[k + m].re = a.re - b.re; out[k + m].im = a.im - b.im;
        } else if (p == 4) {
            /* w_4 = tw[N/4] = (0, sign) */
            double s = tw[N / 4].im;
            cplx a = scratch[0], b = scratch[1], c = scratch[2], d = scratch[3];
            cplx t0 = {a.re + c.re, a.im + c.im}, t1 = {a.re - c.re, a.im - c.im};
            cplx t2 = {b.re + d.re, b.im + d.im}, t3 = {b.re - d.re, b.im - d.im};
            /* i*s*t3 */
            cplx it3 = {-s * t3.im, s * t3.re};
            out[k].re = t0.re + t2.re; out[k].im = t0.im + t2.im;
            out[k + m].re = t1.re + it3.re; out[k + m].im = t1.im + it3.im;
            out[k + 2 * m].re = t0.re - t2.re; out[k + 2 * m].im = t0.im - t2.im;
            out[k + 3 * m].re = t1.re - it3.re; out[k + 3 * m].im = t1.im - it3.im;
        } else {
            for (int q = 0; q < p; ++q) {
                double sr = 0, si = 0;
                for (int r = 0; r < p; ++r) {
                    cplx w = tw[((long)r * q % p) * m * tstep];
                    sr += scratch[r].re * w.re - scratch[r].im * w.im;
                    si += scratch[r].re * w.im + scratch[r].im * w.re;
                }
                out[k + (size_t)q * m].re = sr;
                out[k + (size_t)q * m].im = si;
            }
        }
    }
}

typedef struct { int n; cplx *tw; cplx *buf; cplx *scr; } fft_plan;

static void plan_init(fft_plan *pl, int n, int sign) {
    pl->n = n;
    pl->tw = make_twiddles(n, sign);
    pl->buf = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    pl->scr = (cplx *)malloc(sizeof(cplx) * (size_t)n);
}
static void plan_free(fft_plan *pl) { free(pl->tw); free(pl->buf); free(pl->scr); }

/* in-place transform of x (length n, contiguous) */
static void plan_exec(fft_plan *pl, cplx *x) {
    fft_rec(pl->n, x, 1, pl->buf, pl->tw, 1, pl->scr);
    memcpy(x, pl->buf, sizeof(cplx) * (size_t)pl->n);
}

/* Unnormalised complex DFT, sign +1 = backward (FFTW_BACKWARD convention), in place. */
void oracle_fft1d(int n, double *data, int sign) {
    fft_plan pl;
    plan_init(&pl, n, sign);
    plan_exec(&pl, (cplx *)data);
    plan_free(&pl);
}

/* Brute-force O(n^2) DFT in long double (pin for the FFT). */
void oracle_dft1d_direct(int n, const double *in, double *out, int sign) {
    const long double PIl = 3.14159265358979323846264338327950288L;
    for (int k = 0; k < n; ++k) {
        long double sr = 0, si = 0;
        for (int j = 0; j < n; ++j) {
            long double a = 2.0L * PIl * (long double)(((long)j * k) % n) / (long double)n;
            long double c = cosl(a), s = sign * sinl(a);
            sr += in[2 * j] * c - in[2 * j + 1] * s;
            si += in[2 * j] * s + in[2 * j + 1] * c;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

/* ------------------------------------------------------------------------------------ */
/* REDFT01 (FFTW's unnormalised DCT-III): Y_k = X_0 + 2 sum_{j>=1} X_j cos(pi j (k+1/2)/n) */
/* This is the transform of SRC_F/fftw_ScaRL.f90:378-381 and :562-567.                   */
/* ------------------------------------------------------------------------------------ */
void oracle_redft01_direct(int n, const double *x, double *y) {
    const long double PIl = 3.14159265358979323846264338327950288L;
    for (int k = 0; k < n; ++k) {
        long double s = x[0];
        for (int j = 1; j < n; ++j) {
            long a = (long)j * (2 * k + 1) % (4L * n); /* angle pi*a/(2n), reduced mod 2pi */
            s += 2.0L * x[j] * cosl(PIl * (long double)a / (2.0L * n));
        }
        y[k] = (double)s;
    }
}

/* Makhoul: V_j = (X_j - i X_{n-j}) e^{i pi j/(2n)}, v = backward DFT_n(V);
 * y_{2m} = Re v_m, y_{2m+1} = Re v_{n-1-m}.  pre[j] = e^{i pi j/(2n)}. */
typedef struct { fft_plan fp; cplx *pre; cplx *v; } dct_plan;

static void dct_init(dct_plan *d, int n) {
    plan_init(&d->fp, n, +1);
    d->pre = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    d->v = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < n; ++j) {
        long double a = 3.14159265358979323846264338327950288L * (long double)j / (2.0L * n);
        d->pre[j].re = (double)cosl(a);
        d->pre[j].im = (double)sinl(a);
    }
}
static void dct_free(dct_plan *d) { plan_free(&d->fp); free(d->pre); free(d->v); }

/* x: input with stride sx, only the first nz entries may be non-zero (the rest are zero). */
static void dct_exec(dct_plan *d, const double *x, long sx, int nz, double *y, long sy) {
    int n = d->fp.n;
    for (int j = 0; j < n; ++j) {
        double a = (j < nz) ? x[j * sx] : 0.0;
        int jj = n - j;
        double b = (j > 0 && jj < nz) ? x[jj * sx] : 0.0;
        /* (a - i b) * pre[j] */
        d->v[j].re = a * d->pre[j].re + b * d->pre[j].im;
        d->v[j].im = a * d->pre[j].im - b * d->pre[j].re;
    }
    plan_exec(&d->fp, d->v);
    for (int m = 0; 2 * m < n; ++m) y[(2 * m) * sy] = d->v[m].re;
    for (int m = 0; 2 * m + 1 < n; ++m) y[(2 * m + 1) * sy] = d->v[n - 1 - m].re;
}

void oracle_redft01_fast(int n, const double *x, double *y) {
    dct_plan d;
    dct_init(&d, n);
    dct_exec(&d, x, 1, n, y, 1);
    dct_free(&d);
}

static int nthreads_or_default(int nt) {
#ifdef _OPENMP
    return nt > 0 ? nt : omp_get_max_threads();
#else
    (void)nt;
    return 1;
#endif
}

/* 3D REDFT01 over all axes of a (n1,n2,n3) array whose non-zero support is the box
 * [0,m1)x[0,m2)x[0,m3).  in and out are both (n1,n2,n3) Fortran order; out may equal in.
 * The order is x, then y, then z -- any order gives the same separable transform. */
void oracle_redft01_3d(const int n[3], const int m[3], const double *in, double *out, int nthreads) {
    long n1 = n[0], n2 = n[1], n3 = n[2];
    int nt = nthreads_or_default(nthreads);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)(n1 * n2 * n3));
    /* x pass: rows (j<m2, k<m3) -> tmp, other rows zero */
    memset(tmp, 0, sizeof(double) * (size_t)(n1 * n2 * n3));
#pragma omp parallel num_threads(nt)
    {
        dct_plan d; dct_init(&d, (int)n1);
#pragma omp for collapse(2) schedule(static)
        for (long k = 0; k < m[2]; ++k)
            for (long j = 0; j < m[1]; ++j) {
                long off = (k * n2 + j) * n1;
                dct_exec(&d, in + off, 1, m[0], tmp + off, 1);
            }
        dct_free(&d);
    }
    /* y pass (k<m3) */
#pragma omp parallel num_threads(nt)
    {
        dct_plan d; dct_init(&d, (int)n2);
        double *col = (double *)malloc(sizeof(double) * (size_t)n2);
        double *res = (double *)malloc(sizeof(double) * (size_t)n2);
#pragma omp for collapse(2) schedule(static)
        for (long k = 0; k < m[2]; ++k)
            for (long i = 0; i < n1; ++i) {
                long off = k * n2 * n1 + i;
                for (long j = 0; j < m[1]; ++j) col[j] = tmp[off + j * n1];
                dct_exec(&d, col, 1, m[1], res, 1);
                for (long j = 0; j < n2; ++j) tmp[off + j * n1] = res[j];
            }
        free(col); free(res);
        dct_free(&d);
    }
    /* z pass */
#pragma omp parallel num_threads(nt)
    {
        dct_plan d; dct_init(&d, (int)n3);
        double *col = (double *)malloc(sizeof(double) * (size_t)n3);
        double *res = (double *)malloc(sizeof(double) * (size_t)n3);
#pragma omp for schedule(static)
        for (long ij = 0; ij < n1 * n2; ++ij) {
            for (long k = 0; k < m[2]; ++k) col[k] = tmp[ij + k * n1 * n2];
            dct_exec(&d, col, 1, m[2], res, 1);
            for (long k = 0; k < n3; ++k) out[ij + k * n1 * n2] = res[k];
        }
        free(col); free(res);
        dct_free(&d);
    }
    free(tmp);
}

/* ------------------------------------------------------------------------------------ */
/* Mode R generators (reference parity)                                                  */
/* ------------------------------------------------------------------------------------ */

/* Octant amplitude of SRC_F/fftw_ScaRL.f90:300-370:
 *   dx = xRange/(Np-1); kmax = 1/dx; dk = kmax/(Np-1); k_d(n) = n*dk_d/corrL_d
 *   S = prod_d corrL_d * exp(-PI k_d^2)   (multiplied in axis order, PI truncated)
 *   A = gamma*sqrt(S)*cos(2 PI phi), gamma = 1 (:367).
 * amp has shape (n1,n2,n3) Fortran; entries outside [0,m1)x[0,m2)x[0,m3) are set to 0
 * (the reference leaves them unwritten; zero is the intent of the comment at :331).
 * phi has shape (m1,m2,m3).  Returns 1 for an unsupported corrMod. */
int oracle_mode_r_amplitude(double *amp, const int32_t Np[3], const int32_t m[3], const double xRange[3],
                            const double corrL[3], int32_t corrMod, const double *phi) {
    if (corrMod != CM_GAUSSIAN) return 1;
    long n1 = Np[0], n2 = Np[1], n3 = Np[2];
    double dk[3];
    for (int d = 0; d < 3; ++d) {
        double delta_x = xRange[d] / (double)(Np[d] - 1);
        double k_max = 1.0 / delta_x;
        dk[d] = k_max / (double)(Np[d] - 1);
    }
    memset(amp, 0, sizeof(double) * (size_t)(n1 * n2 * n3));
    for (long k = 0; k < m[2]; ++k)
        for (long j = 0; j < m[1]; ++j)
            for (long i = 0; i < m[0]; ++i) {
                double kp[3] = {(double)i * dk[0] / corrL[0], (double)j * dk[1] / corrL[1], (double)k * dk[2] / corrL[2]};
                double S = 1.0;
                for (int d = 0; d < 3; ++d) S = S * corrL[d] * exp(-PI_REF * (kp[d] * kp[d]));
                double ph = phi[i + m[0] * (j + (long)m[1] * k)];
                amp[i + n1 * (j + n2 * k)] = 1.0 * sqrt(S) * cos(2.0 * PI_REF * ph);
            }
    return 0;
}

/* gen_std_gauss_FFT (SRC_F/fftw_ScaRL.f90:261-418) on a host-exported phi of shape
 * (Np1/2, Np2/2, Np3/2).  out: (Np1,Np2,Np3). */
int oracle_gen_std_gauss_fft(double *out, const int32_t Np[3], const double xRange[3], const double corrL[3],
                             int32_t corrMod, const double *phi, int nthreads) {
    int32_t m[3] = {Np[0] / 2, Np[1] / 2, Np[2] / 2};
    int n[3] = {Np[0], Np[1], Np[2]};
    int mi[3] = {m[0], m[1], m[2]};
    if (oracle_mode_r_amplitude(out, Np, m, xRange, corrL, corrMod, phi)) return 1;
    oracle_redft01_3d(n, mi, out, out, nthreads);
    return 0;
}

/* gen_Std_Gauss_FFT_MPI with one rank (SRC_F/fftw_ScaRL.f90:424-583): the octant is
 * restricted in x and y only, S is filled for all z (:489-510); the blocks the reference
 * leaves unwritten are zero.  phi shape (L1/2, L2/2, L3). */
int oracle_gen_std_gauss_fft_mpi(double *out, const int32_t L[3], const double xRange[3], const double corrL[3],
                                 int32_t corrMod, const double *phi, int nthreads) {
    int32_t m[3] = {L[0] / 2, L[1] / 2, L[2]};
    int n[3] = {L[0], L[1], L[2]};
    int mi[3] = {m[0], m[1], m[2]};
    if (oracle_mode_r_amplitude(out, L, m, xRange, corrL, corrMod, phi)) return 1;
    oracle_redft01_3d(n, mi, out, out, nthreads);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Gamma (Lanczos, g=7, n=9) and K_nu following the ikv scheme                           */
/* K_nu follows the routine IKV of Shanjie Zhang and Jianming Jin, "Computation of       */
/* Special Functions" (Wiley, 1996). IKV is copyrighted by Shanjie Zhang and Jianming    */
/* Jin, who give permission to incorporate it into a user program provided that the      */
/* copyright is acknowledged (SRC_F/bessel.f90:13522-13532).                             */
/* (SRC_F/bessel.f90:13522-13785): K_{v0} by series (x<=9) or asymptotic (x>9), I_{v0}   */
/* by series (x<=18) or asymptotic, I_{v0+1}/I_{v0} by backward recurrence, K_{v0+1} by  */
/* the Wronskian, then forward recurrence to K_nu.                                         */
/* ------------------------------------------------------------------------------------ */
double oracle_gamma(double x) {
    static const double c[9] = {0.99999999999980993, 676.5203681218851, -1259.1392167224028,
                                771.32342877765313, -176.61502916214059, 12.507343278686905,
                                -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};
    if (x < 0.5) return PI_EXACT / (sin(PI_EXACT * x) * oracle_gamma(1.0 - x));
    x -= 1.0;
    double a = c[0];
    double t = x + 7.5;
    for (int i = 1; i < 9; ++i) a += c[i] / (x + (double)i);
    return sqrt(2.0 * PI_EXACT) * pow(t, x + 0.5) * exp(-t) * a;
}

static double envj(int n, double x) { return 0.5 * log10(6.28 * n) - n * log10(1.36 * x / n); }

/* start order of the backward recurrence so that I_m/I_0 ~ 10^-mp (msta1, bessel.f90:20216) */
static int start_order_1(double x, int mp) {
    double a0 = fabs(x);
    int n0 = (int)(1.1 * a0) + 1;
    double f0 = envj(n0, a0) - mp;
    int n1 = n0 + 5;
    double f1 = envj(n1, a0) - mp;
    int nn = n1;
    for (int it = 0; it < 20; ++it) {
        nn = (int)(n1 - (n1 - n0) / (1.0 - f0 / f1));
        double f = envj(nn, a0) - mp;
        if (abs(nn - n1) < 1) break;
        n0 = n1; f0 = f1; n1 = nn; f1 = f;
    }
    return nn;
}

/* start order for n significant terms (msta2, bessel.f90:20295) */
static int start_order_2(double x, int n, int mp) {
    double a0 = fabs(x), hmp = 0.5 * mp, ejn = envj(n, a0), obj;
    int n0;
    if (ejn <= hmp) { obj = mp; n0 = (int)(1.1 * a0) + 1; }
    else { obj = hmp + ejn; n0 = n; }
    double f0 = envj(n0, a0) - obj;
    int n1 = n0 + 5;
    double f1 = envj(n1, a0) - obj;
    int nn = n1;
    for (int it = 0; it < 20; ++it) {
        nn = (int)(n1 - (n1 - n0) / (1.0 - f0 / f1));
        double f = envj(nn, a0) - obj;
        if (abs(nn - n1) < 1) break;
        n0 = n1; f0 = f1; n1 = nn; f1 = f;
    }
    return nn + 10;
}

/* K_nu(x) for nu >= 0, x > 0. */
double oracle_kv(double nu, double x) {
    if (!(x > 0.0)) return INFINITY;
    int n = (int)nu;
    double v0 = nu - n;
    double x2 = x * x;
    double piv = PI_EXACT * v0, vt = 4.0 * v0 * v0;
    int k0 = (x < 35.0) ? 14 : (x < 50.0 ? 10 : 8);
    double gap = oracle_gamma(1.0 + v0);
    double a1 = (v0 == 0.0) ? 1.0 : pow(0.5 * x, v0) / gap;
    /* I_{v0}(x) */
    double bi0;
    if (x <= 18.0) {
        double r = 1.0;
        bi0 = 1.0;
        for (int k = 1; k <= 30; ++k) {
            r = 0.25 * r * x2 / (k * (k + v0));
            bi0 += r;
            if (fabs(r / bi0) < 1e-15) break;
        }
        bi0 *= a1;
    } else {
        double ca = exp(x) / sqrt(2.0 * PI_EXACT * x), s = 1.0, r = 1.0;
        for (int k = 1; k <= k0; ++k) {
            r = -0.125 * r * (vt - (2.0 * k - 1.0) * (2.0 * k - 1.0)) / (k * x);
            s += r;
        }
        bi0 = ca * s;
    }
    /* I_{v0+1}/I_{v0} by backward recurrence from a safe start order */
    int ntop = 1;
    int mstart = start_order_1(x, 200);
    if (mstart < ntop) ntop = mstart; else mstart = start_order_2(x, ntop, 15);
    double f2 = 0.0, f1 = 1e-100, f = 0.0, b0 = 0.0, b1 = 0.0;
    for (int k = mstart; k >= 0; --k) {
        f = 2.0 * (v0 + k + 1.0) / x * f1 + f2;
        if (k == 1) b1 = f;
        if (k == 0) b0 = f;
        f2 = f1; f1 = f;
    }
    double cs = bi0 / f;
    double I0 = cs * b0, I1 = cs * b1;
    /* K_{v0}(x) */
    double bk0;
    if (x <= 9.0) {
        if (v0 == 0.0) {
            double ct = -log(0.5 * x) - 0.5772156649015329;
            double s = 0.0, w0 = 0.0, r = 1.0, ww = 0.0;
            for (int k = 1; k <= 50; ++k) {
                w0 += 1.0 / k;
                r = 0.25 * r / ((double)k * k) * x2;
                s += r * (w0 + ct);
                double wa = fabs(s);
                if (k > 1 && fabs((wa - ww) / wa) < 1e-15) break;
                ww = wa;
            }
            bk0 = ct + s;
        } else {
            double gan = oracle_gamma(1.0 - v0);
            double a2 = 1.0 / (gan * pow(0.5 * x, v0));
            double s = a2 - a1, r1 = 1.0, r2 = 1.0, ww = 0.0;
            for (int k = 1; k <= 120; ++k) {
                r1 = 0.25 * r1 * x2 / (k * (k - v0));
                r2 = 0.25 * r2 * x2 / (k * (k + v0));
                s += a2 * r1 - a1 * r2;
                double wa = fabs(s);
                if (k > 1 && fabs((wa - ww) / wa) < 1e-15) break;
                ww = wa;
            }
            bk0 = 0.5 * PI_EXACT * s / sin(piv);
        }
    } else {
        double cb = exp(-x) * sqrt(0.5 * PI_EXACT / x), s = 1.0, r = 1.0;
        for (int k = 1; k <= k0; ++k) {
            r = 0.125 * r * (vt - (2.0 * k - 1.0) * (2.0 * k - 1.0)) / (k * x);
            s += r;
        }
        bk0 = cb * s;
    }
    if (n == 0) return bk0;
    double bk1 = (1.0 / x - I1 * bk0) / I0; /* Wronskian */
    for (int k = 2; k <= n; ++k) {
        double bk2 = 2.0 * (v0 + k - 1.0) / x * bk1 + bk0;
        bk0 = bk1; bk1 = bk2;
    }
    return bk1;
}

/* ------------------------------------------------------------------------------------ */
/* Spectra (angular wavenumber, C(r) = integral S(kappa) e^{i kappa.r} dkappa)          */
/* ------------------------------------------------------------------------------------ */

/* rho2 = sum_d (kappa_d L_d)^2 over the dims that are present (dim count d). */
double oracle_spectrum(int model, int dim, const double kappa[3], const double L[3], double nu) {
    double pL = 1.0, rho2 = 0.0;
    for (int d = 0; d < dim; ++d) { pL *= L[d]; rho2 += kappa[d] * L[d] * kappa[d] * L[d]; }
    switch (model) {
    case CM_GAUSSIAN: {
        /* C = exp(-sum (r_d/L_d)^2):  S = prod_d L_d/(2 sqrt(pi)) exp(-kappa_d^2 L_d^2 / 4) */
        double s = 1.0;
        for (int d = 0; d < dim; ++d) s *= L[d] / (2.0 * sqrt(PI_EXACT)) * exp(-0.25 * kappa[d] * kappa[d] * L[d] * L[d]);
        return s;
    }
    case CM_EXPONENTIAL: {
        /* Matern nu = 1/2 closed forms */
        if (dim == 3) return pL / (PI_EXACT * PI_EXACT * (1.0 + rho2) * (1.0 + rho2));
        if (dim == 2) return pL / (2.0 * PI_EXACT * pow(1.0 + rho2, 1.5));
        return pL / (PI_EXACT * (1.0 + rho2));
    }
    case CM_VON_KARMAN: {
        double e = nu + 0.5 * dim;
        double c = oracle_gamma(e) / (oracle_gamma(nu) * pow(PI_EXACT, 0.5 * dim));
        return c * pL * pow(1.0 + rho2, -e);
    }
    }
    return 0.0;
}

/* Target correlation C(r)/sigma^2 at anisotropic lag r (in the same units as L). */
double oracle_target_corr(int model, int dim, const double r[3], const double L[3], double nu) {
    double rho2 = 0.0;
    for (int d = 0; d < dim; ++d) rho2 += (r[d] / L[d]) * (r[d] / L[d]);
    double rho = sqrt(rho2);
    switch (model) {
    case CM_GAUSSIAN: return exp(-rho2);
    case CM_EXPONENTIAL: return exp(-rho);
    case CM_VON_KARMAN:
        if (rho == 0.0) return 1.0;
        return pow(2.0, 1.0 - nu) / oracle_gamma(nu) * pow(rho, nu) * oracle_kv(nu, rho);
    }
    return 0.0;
}

/* ------------------------------------------------------------------------------------ */
/* Mode H: Hermitian half-spectrum white noise, amplitude, C2R                           */
/* Half spectrum H has shape (h1 = n1/2+1, n2, n3), Fortran order, complex.              */
/* ------------------------------------------------------------------------------------ */

static long signed_freq(long m, long n) { return (m > n / 2) ? m - n : m; }

/* Complex N(0,1) noise with Hermitian symmetry on the self-conjugate x planes
 * (kx = 0 and, for even n1, kx = n1/2): W(kx,-ky,-kz) = conj W(kx,ky,kz); self-conjugate
 * points are real N(0,1).  Counter = linear half-spectrum index of the canonical point. */
void oracle_noise_hermitian(double *W, const int n[3], uint64_t seed, uint32_t stream) {
    long n1 = n[0], n2 = n[1], n3 = n[2], h1 = n1 / 2 + 1;
    const double rs2 = 0.70710678118654752440;
    for (long kz = 0; kz < n3; ++kz)
        for (long ky = 0; ky < n2; ++ky)
            for (long kx = 0; kx < h1; ++kx) {
                long idx = kx + h1 * (ky + n2 * kz);
                double g[2];
                int selfconj_plane = (kx == 0) || ((n1 % 2 == 0) && kx == n1 / 2);
                if (!selfconj_plane) {
                    oracle_gauss_at(kx, ky, kz, h1, n2, n3, seed, stream, g);
                    W[2 * idx] = g[0] * rs2;
                    W[2 * idx + 1] = g[1] * rs2;
                    continue;
                }
                long py = (n2 - ky) % n2, pz = (n3 - kz) % n3;
                long lin_self = ky + n2 * kz, lin_part = py + n2 * pz;
                long cy = lin_self <= lin_part ? ky : py, cz = lin_self <= lin_part ? kz : pz;
                oracle_gauss_at(kx, cy, cz, h1, n2, n3, seed, stream, g);
                if (lin_self == lin_part) {
                    W[2 * idx] = g[0];
                    W[2 * idx + 1] = 0.0;
                } else {
                    double s = (lin_self < lin_part) ? 1.0 : -1.0;
                    W[2 * idx] = g[0] * rs2;
                    W[2 * idx + 1] = s * g[1] * rs2;
                }
            }
}

/* Sum over the FULL spectrum of S(kappa) (half spectrum with multiplicity 2 off the
 * self-conjugate planes), in long double.  dk = prod 2 pi/(n_d dx_d). */
double oracle_spectrum_sum(const int n[3], const double dx[3], int model, const double L[3], double nu) {
    long n1 = n[0], n2 = n[1], n3 = n[2], h1 = n1 / 2 + 1;
    int dim = (n3 > 1) ? 3 : ((n2 > 1) ? 2 : 1);
    long double s = 0.0L;
    for (long kz = 0; kz < n3; ++kz)
        for (long ky = 0; ky < n2; ++ky)
            for (long kx = 0; kx < h1; ++kx) {
                double kap[3] = {2.0 * PI_EXACT * (double)kx / (n1 * dx[0]),
                                 2.0 * PI_EXACT * (double)signed_freq(ky, n2) / (n2 * dx[1]),
                                 2.0 * PI_EXACT * (double)signed_freq(kz, n3) / (n3 * dx[2])};
                int mult = (kx == 0 || (n1 % 2 == 0 && kx == n1 / 2)) ? 1 : 2;
                s += (long double)mult * oracle_spectrum(model, dim, kap, L, nu);
            }
    return (double)s;
}

/* Amplitude a(k) = sigma*sqrt(S/sum S) (normalize=1: variance exactly sigma^2 over the
 * discrete spectrum) or sqrt(S*dkappa)*sigma (normalize=0).  amp: half-spectrum reals. */
void oracle_amplitude_h(double *amp, const int n[3], const double dx[3], int model, const double L[3], double nu,
                        double sigma, int normalize) {
    long n1 = n[0], n2 = n[1], n3 = n[2], h1 = n1 / 2 + 1;
    int dim = (n3 > 1) ? 3 : ((n2 > 1) ? 2 : 1);
    double dkap = 1.0;
    for (int d = 0; d < dim; ++d) dkap *= 2.0 * PI_EXACT / (n[d] * dx[d]);
    double scale = normalize ? 1.0 / oracle_spectrum_sum(n, dx, model, L, nu) : dkap;
    for (long kz = 0; kz < n3; ++kz)
        for (long ky = 0; ky < n2; ++ky)
            for (long kx = 0; kx < h1; ++kx) {
                double kap[3] = {2.0 * PI_EXACT * (double)kx / (n1 * dx[0]),
                                 2.0 * PI_EXACT * (double)signed_freq(ky, n2) / (n2 * dx[1]),
                                 2.0 * PI_EXACT * (double)signed_freq(kz, n3) / (n3 * dx[2])};
                amp[kx + h1 * (ky + n2 * kz)] = sigma * sqrt(oracle_spectrum(model, dim, kap, L, nu) * scale);
            }
}

/* Unnormalised 3D complex-to-real backward transform of a half spectrum (imaginary parts
 * of the kx=0 / kx=n1/2 entries are ignored, as FFTW and pocketfft do).  H is consumed. */
void oracle_c2r_3d(const int n[3], double *H, double *out, int nthreads) {
    long n1 = n[0], n2 = n[1], n3 = n[2], h1 = n1 / 2 + 1;
    int nt = nthreads_or_default(nthreads);
    cplx *h = (cplx *)H;
    if (n3 > 1) {
#pragma omp parallel num_threads(nt)
        {
            fft_plan pl; plan_init(&pl, (int)n3, +1);
            cplx *col = (cplx *)malloc(sizeof(cplx) * (size_t)n3);
#pragma omp for schedule(static)
            for (long ij = 0; ij < h1 * n2; ++ij) {
                for (long k = 0; k < n3; ++k) col[k] = h[ij + k * h1 * n2];
                plan_exec(&pl, col);
                for (long k = 0; k < n3; ++k) h[ij + k * h1 * n2] = col[k];
            }
            free(col); plan_free(&pl);
        }
    }
    if (n2 > 1) {
#pragma omp parallel num_threads(nt)
        {
            fft_plan pl; plan_init(&pl, (int)n2, +1);
            cplx *col = (cplx *)malloc(sizeof(cplx) * (size_t)n2);
#pragma omp for collapse(2) schedule(static)
            for (long k = 0; k < n3; ++k)
                for (long i = 0; i < h1; ++i) {
                    long off = k * h1 * n2 + i;
                    for (long j = 0; j < n2; ++j) col[j] = h[off + j * h1];
                    plan_exec(&pl, col);
                    for (long j = 0; j < n2; ++j) h[off + j * h1] = col[j];
                }
            free(col); plan_free(&pl);
        }
    }
#pragma omp parallel num_threads(nt)
    {
        fft_plan pl; plan_init(&pl, (int)n1, +1);
        cplx *row = (cplx *)malloc(sizeof(cplx) * (size_t)n1);
#pragma omp for schedule(static)
        for (long r = 0; r < n2 * n3; ++r) {
            const cplx *src = h + r * h1;
            for (long i = 0; i < h1; ++i) row[i] = src[i];
            row[0].im = 0.0;
            if (n1 % 2 == 0) row[n1 / 2].im = 0.0;
            for (long i = h1; i < n1; ++i) { row[i].re = src[n1 - i].re; row[i].im = -src[n1 - i].im; }
            plan_exec(&pl, row);
            for (long i = 0; i < n1; ++i) out[r * n1 + i] = row[i].re;
        }
        free(row); plan_free(&pl);
    }
}

/* ------------------------------------------------------------------------------------ */
/* Marginal transform (SRC_F/main_ScaRL.f90:1355-1395)                                   */
/* ------------------------------------------------------------------------------------ */
int oracle_marginal_params(int32_t margi, double avg, double std_dev, double *mu_n, double *sd_n) {
    double var;
    if (margi == FOM_GAUSSIAN) { var = std_dev * std_dev; *mu_n = avg; }
    else if (margi == FOM_LOGNORMAL) {
        if (avg <= 0.0) return 2;
        var = log(1.0 + (std_dev / avg) * (std_dev / avg));
        *mu_n = log(avg) - var / 2.0;
    } else return 1;
    *sd_n = sqrt(var);
    return 0;
}

int oracle_marginal(double *x, long n, int32_t margi, double avg, double std_dev) {
    double mu, sd;
    int rc = oracle_marginal_params(margi, avg, std_dev, &mu, &sd);
    if (rc) return rc;
    for (long i = 0; i < n; ++i) {
        double v = x[i] * sd + mu;
        x[i] = (margi == FOM_LOGNORMAL) ? exp(v) : v;
    }
    return 0;
}

/* Mode H single field: noise (W given, or Philox if W == NULL) * amplitude -> C2R ->
 * marginal.  out: (n1,n2,n3). */
int oracle_gen_mode_h(double *out, const int n[3], const double dx[3], int model, const double L[3], double nu,
                      int normalize, const double *W, uint64_t seed, uint32_t stream, int32_t margi, double avg,
                      double std_dev, int nthreads) {
    long n1 = n[0], n2 = n[1], n3 = n[2], h1 = n1 / 2 + 1, H = h1 * n2 * n3;
    double *amp = (double *)malloc(sizeof(double) * (size_t)H);
    double *X = (double *)malloc(sizeof(double) * 2 * (size_t)H);
    if (W) memcpy(X, W, sizeof(double) * 2 * (size_t)H);
    else oracle_noise_hermitian(X, n, seed, stream);
    oracle_amplitude_h(amp, n, dx, model, L, nu, 1.0, normalize);
    for (long i = 0; i < H; ++i) { X[2 * i] *= amp[i]; X[2 * i + 1] *= amp[i]; }
    free(amp);
    oracle_c2r_3d(n, X, out, nthreads);
    free(X);
    if (margi) return oracle_marginal(out, n1 * n2 * n3, margi, avg, std_dev);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Partition of unity (SRC_F/main_ScaRL.f90:1216-1301)                                   */
/* ------------------------------------------------------------------------------------ */

/* Per-axis weight at local index i (0-based) of an extent np axis with overlap nov on
 * the low side (has_lo) and/or high side (has_hi).  Cosine ramp t_k = k/(nov-1), reversed
 * on the low side; w = (1 + cos(pi t))/2 with the reference's PI literal. */
double oracle_pou_axis_weight(int i, int np, int nov_lo, int nov_hi, int has_lo, int has_hi) {
    double w = 1.0;
    if (has_lo && nov_lo >= 2 && i < nov_lo) {
        int k = nov_lo - 1 - i; /* reversed pattern */
        double t = (double)k / (double)(nov_lo - 1);
        w *= (1.0 + cos(PI_REF * t)) / 2.0;
    }
    if (has_hi && nov_hi >= 2 && i >= np - nov_hi) {
        int k = i - (np - nov_hi);
        double t = (double)k / (double)(nov_hi - 1);
        w *= (1.0 + cos(PI_REF * t)) / 2.0;
    }
    return w;
}

/* apply_UnityPartition_mtx: field *= sqrt(prod_axis w).  topo_pos/topo_shape as in the
 * reference; neighbour exists on a side iff topo_pos+-1 is inside the topology. */
void oracle_apply_pou(double *f, const int32_t Np[3], const int32_t Nov[3], const int32_t topo_pos[3],
                      const int32_t topo_shape[3]) {
    long n1 = Np[0], n2 = Np[1], n3 = Np[2];
    double *w[3];
    for (int d = 0; d < 3; ++d) {
        w[d] = (double *)malloc(sizeof(double) * (size_t)Np[d]);
        int lo = topo_pos[d] - 1 >= 0, hi = topo_pos[d] + 1 <= topo_shape[d] - 1;
        for (int i = 0; i < Np[d]; ++i) w[d][i] = oracle_pou_axis_weight(i, Np[d], Nov[d], Nov[d], lo, hi);
    }
    for (long k = 0; k < n3; ++k)
        for (long j = 0; j < n2; ++j)
            for (long i = 0; i < n1; ++i) f[i + n1 * (j + n2 * k)] *= sqrt(w[0][i] * w[1][j] * w[2][k]);
    for (int d = 0; d < 3; ++d) free(w[d]);
}

/* ------------------------------------------------------------------------------------ */
/* Moments (SRC_F/Stat_ScaRL/statistics_RF.f90:18-174), FP64 Kahan sums                  */
/* ------------------------------------------------------------------------------------ */
typedef struct { double s, c; } kahan;
static void kadd(kahan *k, double v) {
    double y = v - k->c;
    double t = k->s + y;
    k->c = (t - k->s) - y;
    k->s = t;
}

/* rf: (n, nmc) column-major; per event avg/std with the reference's formulas, global over
 * all events.  naive=1 reproduces the reference's plain summation. */
void oracle_moments(const double *rf, long n, int nmc, double *evnt_avg, double *evnt_std, double *g_avg,
                    double *g_std, int naive) {
    long double gs = 0, gs2 = 0;
    double tot_s = 0, tot_s2 = 0;
    kahan ks = {0, 0}, ks2 = {0, 0};
    for (int e = 0; e < nmc; ++e) {
        const double *x = rf + (size_t)e * n;
        double s, s2;
        if (naive) {
            s = 0; s2 = 0;
            for (long i = 0; i < n; ++i) { s += x[i]; s2 += x[i] * x[i]; }
        } else {
            kahan a = {0, 0}, b = {0, 0};
            for (long i = 0; i < n; ++i) { kadd(&a, x[i]); kadd(&b, x[i] * x[i]); }
            s = a.s; s2 = b.s;
        }
        evnt_avg[e] = s / (double)n;
        evnt_std[e] = sqrt(s2 / (double)n - evnt_avg[e] * evnt_avg[e]);
        if (naive) { tot_s += s; tot_s2 += s2; } else { kadd(&ks, s); kadd(&ks2, s2); }
        gs += s; gs2 += s2;
    }
    (void)gs; (void)gs2;
    double S = naive ? tot_s : ks.s, S2 = naive ? tot_s2 : ks2.s;
    *g_avg = S / ((double)n * nmc);
    *g_std = sqrt(S2 / ((double)n * nmc) - (*g_avg) * (*g_avg));
}

/* Per-axis lag sums: sums[axis*(nlag)+l] = sum over pairs (p, p + l e_axis) inside the box
 * of (x_p - centre)(x_{p+l} - centre); cnt[] = number of pairs.  Long double accumulation. */
void oracle_lag_sums(const double *x, const int n[3], int nlag, double centre, double *sums, double *cnt) {
    long n1 = n[0], n2 = n[1], n3 = n[2];
    long st[3] = {1, n1, n1 * n2};
    for (int a = 0; a < 3; ++a)
        for (int l = 0; l < nlag; ++l) {
            long double s = 0;
            long c = 0;
            if (l < n[a]) {
                for (long k = 0; k < n3; ++k)
                    for (long j = 0; j < n2; ++j)
                        for (long i = 0; i < n1; ++i) {
                            long idx[3] = {i, j, k};
                            if (idx[a] + l >= n[a]) continue;
                            long p = i + n1 * (j + n2 * k);
                            s += ((long double)x[p] - centre) * ((long double)x[p + l * st[a]] - centre);
                            ++c;
                        }
            }
            sums[a * nlag + l] = (double)s;
            cnt[a * nlag + l] = (double)c;
        }
}

/* ------------------------------------------------------------------------------------ */
/* build_Sk_mtx (SRC_F/spectra_ScaRL.f90:10-142), Gaussian branch, and its (dead)        */
/* consumer gen_std_gauss_Shino_FFT (SRC_F/fftw_ScaRL.f90:157-255).                      */
/* ------------------------------------------------------------------------------------ */

/* Shape query: sz_Sk = ceiling(k_cut/delta_k) (spectra_ScaRL.f90:42-72); the matrix is
 * (sz1, 2 sz2, 2 sz3).  Returns 1 for an unsupported corrMod. */
int oracle_build_sk_shape(int32_t corrMod, const double xRange[3], const int32_t Np[3], const int32_t sz_c[3],
                          int32_t sz_sk[3]) {
    if (corrMod != CM_GAUSSIAN && corrMod != CM_VON_KARMAN) return 1;
    double k_cut[3], delta_x[3], k_max[3], delta_k[3];
    int any_lt = 0;
    for (int d = 0; d < 3; ++d) {
        k_cut[d] = 7.355;
        delta_x[d] = xRange[d] / (double)(Np[d] - 1);
        k_max[d] = 1.0 / delta_x[d];
        delta_k[d] = k_max[d] / (double)(sz_c[d] / 2);
        if (k_max[d] < k_cut[d]) any_lt = 1;
    }
    if (any_lt)
        for (int d = 0; d < 3; ++d)
            if (k_max[d] > k_cut[d]) k_cut[d] = k_max[d];
    for (int d = 0; d < 3; ++d) sz_sk[d] = (int32_t)ceil(k_cut[d] / delta_k[d]);
    return 0;
}

/* Sk (complex, shape (sz1, 2 sz2, 2 sz3)).  Phase from r8vec_uniform_01(seed+1000000)
 * in Fortran order; planes 1 (all axes) and the last y/z planes zeroed; Gaussian factor
 * sqrt(exp(-sum n_d^2 (pi^2/2) dk_d^2)) applied to the 4 mirrored quadrants for
 * ii,jj,kk >= 2; then * sqrt(8) * 2 sqrt(prod dk).  von Karman is commented out in the
 * reference (:119-137): for it only the phase matrix is produced, unscaled. */
int oracle_build_sk(int32_t corrMod, const double xRange[3], const int32_t Np[3], const int32_t sz_c[3],
                    int32_t seed, double *sk) {
    int32_t sz[3];
    if (oracle_build_sk_shape(corrMod, xRange, Np, sz_c, sz)) return 1;
    long s1 = sz[0], s2 = 2L * sz[1], s3 = 2L * sz[2];
    long tot = s1 * s2 * s3;
    double *ph = (double *)malloc(sizeof(double) * (size_t)tot);
    int32_t sd = seed + 1000000;
    oracle_r8vec_uniform_01((int32_t)tot, &sd, ph);
    for (long i = 0; i < tot; ++i) {
        /* cmplx() without KIND yields default (single precision) complex (:80) */
        double t = 2.0 * PI_REF * ph[i];
        sk[2 * i] = (double)(float)cos(t);
        sk[2 * i + 1] = (double)(float)sin(t);
    }
    free(ph);
#define SK(i, j, k) (sk + 2 * ((i) + s1 * ((j) + s2 * (k))))
    for (long k = 0; k < s3; ++k)
        for (long j = 0; j < s2; ++j)
            for (long i = 0; i < s1; ++i)
                if (i == 0 || j == 0 || k == 0 || j == s2 - 1 || k == s3 - 1) { SK(i, j, k)[0] = 0; SK(i, j, k)[1] = 0; }
    if (corrMod == CM_GAUSSIAN) {
        double dk[3], de[3];
        for (int d = 0; d < 3; ++d) {
            double delta_x = xRange[d] / (double)(Np[d] - 1);
            dk[d] = (1.0 / delta_x) / (double)(sz_c[d] / 2);
            de[d] = ((PI_REF * PI_REF) / 2.0) * (dk[d] * dk[d]);
        }
        for (long ii = 2; ii <= sz[0]; ++ii)
            for (long jj = 2; jj <= sz[1]; ++jj)
                for (long kk = 2; kk <= sz[2]; ++kk) {
                    double a = (double)((ii - 1) * (ii - 1)) * de[0] + (double)((jj - 1) * (jj - 1)) * de[1] +
                               (double)((kk - 1) * (kk - 1)) * de[2];
                    double v = sqrt(exp(-a));
                    long i = ii - 1, j = jj - 1, k = kk - 1, jm = s2 - jj, km = s3 - kk;
                    double *p[4] = {SK(i, j, k), SK(i, jm, k), SK(i, j, km), SK(i, jm, km)};
                    for (int q = 0; q < 4; ++q) { p[q][0] *= v; p[q][1] *= v; }
                }
        double f = 2.0 * sqrt(dk[0] * dk[1] * dk[2]);
        double r8 = sqrt(8.0);
        for (long i = 0; i < tot; ++i) {
            sk[2 * i] = f * (r8 * sk[2 * i]);
            sk[2 * i + 1] = f * (r8 * sk[2 * i + 1]);
        }
    }
#undef SK
    return 0;
}

/* gen_std_gauss_Shino_FFT: xRange/corrL, pad to sz_c = 2^ceil(log2(2 Np)), place the four
 * y/z quadrants of Sk, 3D C2C BACKWARD of size sz_c, real part of the first Np. */
int oracle_gen_std_gauss_shino_fft(double *out, const int32_t Np[3], const double xRange_in[3], const double corrL[3],
                                   int32_t corrMod, int32_t seed, int nthreads) {
    double xr[3];
    int32_t szc[3], sz[3];
    for (int d = 0; d < 3; ++d) {
        xr[d] = xRange_in[d] / corrL[d];
        szc[d] = (int32_t)pow(2.0, ceil(log(2.0 * Np[d]) / log(2.0)));
    }
    if (oracle_build_sk_shape(corrMod, xr, Np, szc, sz)) return 1;
    /* quadrants that overflow or overlap the padded grid: the reference writes out of bounds */
    if (sz[0] / 2 > szc[0] || 2 * sz[1] > szc[1] || 2 * sz[2] > szc[2]) return 3;
    long s1 = sz[0], s2 = 2L * sz[1], s3 = 2L * sz[2];
    double *sk = (double *)malloc(sizeof(double) * 2 * (size_t)(s1 * s2 * s3));
    oracle_build_sk(corrMod, xr, Np, szc, seed, sk);
    long c1 = szc[0], c2 = szc[1], c3 = szc[2];
    cplx *c = (cplx *)calloc((size_t)(c1 * c2 * c3), sizeof(cplx));
    /* sz_Sk = shape(Sk_mtx)/2 (fftw_ScaRL.f90:208) halves the x extent too, so the
     * quadrant copies (:217-233) take x entries 1..sz1/2 (the left-hand extent). */
    long h2 = sz[1], h3 = sz[2], hx = s1 / 2;
    for (long k = 0; k < s3; ++k)
        for (long j = 0; j < s2; ++j)
            for (long i = 0; i < hx; ++i) {
                long jd = (j < h2) ? j : c2 - s2 + j;
                long kd = (k < h3) ? k : c3 - s3 + k;
                if (i >= c1 || jd < 0 || jd >= c2 || kd < 0 || kd >= c3) continue;
                cplx *dst = c + i + c1 * (jd + c2 * kd);
                const double *src = sk + 2 * (i + s1 * (j + s2 * k));
                dst->re = src[0]; dst->im = src[1];
            }
    free(sk);
    int nt = nthreads_or_default(nthreads);
    long dims[3] = {c1, c2, c3};
    long strides[3] = {1, c1, c1 * c2};
    for (int a = 0; a < 3; ++a) {
        long n = dims[a], st = strides[a];
        long npen = c1 * c2 * c3 / n;
#pragma omp parallel num_threads(nt)
        {
            fft_plan pl; plan_init(&pl, (int)n, +1);
            cplx *col = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
            for (long p = 0; p < npen; ++p) {
                /* pencil base: index with axis a removed */
                long rest = p, base = 0;
                for (int b = 0; b < 3; ++b) {
                    if (b == a) continue;
                    long coord = rest % dims[b];
                    rest /= dims[b];
                    base += coord * strides[b];
                }
                for (long t = 0; t < n; ++t) col[t] = c[base + t * st];
                plan_exec(&pl, col);
                for (long t = 0; t < n; ++t) c[base + t * st] = col[t];
            }
            free(col); plan_free(&pl);
        }
    }
    for (long k = 0; k < Np[2]; ++k)
        for (long j = 0; j < Np[1]; ++j)
            for (long i = 0; i < Np[0]; ++i) out[i + Np[0] * (j + (long)Np[1] * k)] = c[i + c1 * (j + c2 * k)].re;
    free(c);
    return 0;
}
