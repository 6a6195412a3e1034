/*
 * scarl_hip.h -- C ABI of libscarl_hip.so, the MI355X (gfx950) spectral random-field engine.
 *
 * Plain C: pointers, sizes and status codes only.  Every function returns 0 on success and a
 * non-zero SCARL_E_* code on failure; scarl_hip_last_error() returns the message.  Arrays are
 * Fortran column-major (x fastest), as the reference stores k_mtx(Np1,Np2,Np3)
 * (SRC_F/main_ScaRL.f90:405).  Host pointers (h_*) are owned by the caller; device pointers
 * (d_*) are owned by whoever allocated them.  Calls on one context are not re-entrant across
 * host threads; each call is synchronous unless it says otherwise.
 *
 * Constants follow SRC_F/constants_ScaRL.f90:5-27: cm_GAUSSIAN = 1, cm_VON_KARMAN = 2,
 * fom_GAUSSIAN = 1, fom_LOGNORMAL = 2.  The exponential model (3) is new here.
 */
#ifndef SCARL_HIP_H
#define SCARL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SCARL_OK = 0,
    SCARL_E_INVALID = 1,     /* bad argument (shape, enum, parameter domain)                */
    SCARL_E_HIP = 2,         /* HIP runtime error, or no GPU                                */
    SCARL_E_UNSUPPORTED = 3, /* valid request the engine does not implement                 */
    SCARL_E_NOMEM = 4        /* device allocation failed                                    */
};

enum { SCARL_CM_GAUSSIAN = 1, SCARL_CM_VON_KARMAN = 2, SCARL_CM_EXPONENTIAL = 3 };
enum { SCARL_FOM_NONE = 0, SCARL_FOM_GAUSSIAN = 1, SCARL_FOM_LOGNORMAL = 2 };

/* ---- lifetime, errors, timing ------------------------------------------------------------ */
int scarl_hip_init(int32_t device);
int scarl_hip_finalize(void);
/* copies the last error message of the calling thread (NUL terminated, truncated to len) */
int scarl_hip_last_error(char* buf, int32_t len);
/* 1 = bracket every kernel launch with HIP events on the library stream */
int scarl_hip_set_timing(int32_t on);
/* synchronises and writes "name count total_ms" lines for every timed kernel */
int scarl_hip_timing_report(char* buf, int32_t len);
int scarl_hip_timing_reset(void);
/* the library's HIP stream (hipStream_t), for callers that order their own work after it */
int scarl_hip_get_stream(void** stream);
/* plan tuning for batches created afterwards (no counterpart in the reference; the autotuner's
 * knob).  SCARL_TUNE_FP32_ELEMS: complex elements a thread holds per FFT stage, 16 or 32.
 * SCARL_TUNE_MFMA_DFT16: 1 runs every FP32 radix-16 first stage (lengths that are multiples of 16,
 * 32 to 4096; both stages of the 256-point x pass of 512-wide fields) as split-precision bf16 matrix
 * products on the MFMA units (three bf16 terms per operand, six products, FP32 accumulation);
 * 2 does so in the x passes only; 0 (default) keeps the FP32 Stockham butterflies.
 * SCARL_TUNE_AUTOTUNE: 1 makes batch creation time every plan's candidates (compile-time kernel,
 * runtime radix splits with 16 or 32 elements per thread in either order, pencils per tile, LDS
 * skew, MFMA first stage) on the batch's own passes, confirm the best in replays of the whole step
 * and keep the fastest (and, with SCARL_TUNE_SCHEDULE 0, the faster schedule); picks are cached
 * per process and batch geometry, and in the file named by the SCARL_TUNE_CACHE environment
 * variable when it is set (the reference's counterpart is FFTW's planner,
 * SRC_F/fftw_ScaRL.f90:378).
 * SCARL_TUNE_LAG_FFT: 1 (default) computes the periodic lag sums of the corrL estimator by FFT
 * (|FFT|^2 and back, FP64, as SRC_F/Stat_ScaRL/statistics_RF.f90:353-448 does) wherever the length
 * has a Stockham plan and there are at least 8 lags; 0 uses the direct products from LDS.  The
 * bounded lags of the per-axis covariance always use the direct products.
 * SCARL_TUNE_SCHEDULE: how a batch of several groups runs: 0 (default) pipelined for a blend or
 * when the mean group workspace exceeds the 256 MB Infinity Cache, else one group after another
 * (the autotuner times both and keeps the faster); 1 always pipelined (the next group's generating
 * pass on a second stream beside this group's y and x passes); 2 always one group after another. */
enum {
    SCARL_TUNE_FP32_ELEMS = 1,
    SCARL_TUNE_MFMA_DFT16 = 2,
    SCARL_TUNE_AUTOTUNE = 3,
    SCARL_TUNE_LAG_FFT = 4,
    SCARL_TUNE_SCHEDULE = 5
};
int scarl_hip_set_tuning(int32_t key, int32_t value);

/* ---- fftw_ScaRL drop-in ------------------------------------------------------------------ */
/* gen_std_gauss_FFT (SRC_F/fftw_ScaRL.f90:261-418): REDFT01 x REDFT01 x REDFT01 of
 * sqrt(S) cos(2 PI phi) on the octant [1:Np/2]^3, FP64.  h_out: Np1*Np2*Np3 doubles.
 * h_phi: (Np1/2)*(Np2/2)*(Np3/2) phases in [0,1) (host-exported noise), or NULL to draw phi on
 * the device from Philox4x32-10 with key (seed + rank).  flags: bit 0 = fuse the marginal
 * transform given by scarl_hip_set_marginal.  Octants the reference leaves unwritten are 0. */
int scarl_hip_gen_std_gauss_fft(double* h_out, const int32_t Np[3], const double xRange[3], const double corrL[3],
                                int32_t corrMod, int32_t seed, int32_t rank, const double* h_phi, int32_t flags);
/* gen_Std_Gauss_FFT_MPI (SRC_F/fftw_ScaRL.f90:424-583) for one rank holding the whole field
 * (comm_size must be 1): octant restricted in x and y, all z.  h_phi: (L1/2)*(L2/2)*L3. */
int scarl_hip_gen_std_gauss_fft_mpi(double* h_out, const double xRange[3], const double corrL[3], int32_t corrMod,
                                    int32_t seed, int32_t rank, const int32_t L[3], const int32_t Np[3],
                                    int32_t comm_size, const double* h_phi, int32_t flags);
/* device-resident variants (inputs and output already in HBM) */
int scarl_hip_gen_std_gauss_fft_dev(double* d_out, const int32_t Np[3], const double xRange[3],
                                    const double corrL[3], int32_t corrMod, int32_t seed, int32_t rank,
                                    const double* d_phi, int32_t flags);
/* Allocation-free form of scarl_hip_gen_std_gauss_fft_dev.  It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and results as the _dev
 * call (the same kernels), but every device buffer it needs (plan table, twiddles, the two
 * REDFT01 workspaces) is carved out of one static device array of 97 MiB held by the library on
 * each GPU it runs on: fields whose workspaces (n1 m2 m3 + n1 n2 m3 doubles, m = Np/2) fit in
 * 96 MiB are accepted (256^3 is), larger ones fail with SCARL_E_INVALID before any launch.  It calls no hipMalloc/hipFree.  All calls on one GPU share
 * that array: calls that overlap in time on one GPU race.  The call returns once its kernels are
 * queued on the library stream (scarl_hip_get_stream); the array keeps the last call's contents. */
int scarl_hip_gen_std_gauss_fft_dev_noalloc(double* d_out, const int32_t Np[3], const double xRange[3],
                                            const double corrL[3], int32_t corrMod, int32_t seed, int32_t rank,
                                            const double* d_phi, int32_t flags);
/* marginal fused by flags bit 0 (multiVariateTransformation, SRC_F/main_ScaRL.f90:1355-1395) */
int scarl_hip_set_marginal(int32_t margiFirst, double avg, double std_dev);

/* ---- spectra_ScaRL drop-in --------------------------------------------------------------- */
/* build_Sk_mtx (SRC_F/spectra_ScaRL.f90:10-142) allocates Sk_mtx inside; the C side splits it
 * into a shape query and a fill.  out_shape = (sz_Sk1, 2 sz_Sk2, 2 sz_Sk3). */
int scarl_hip_build_sk_query(int32_t corrMod, const double xRange[3], const double corrL[3], const int32_t Np[3],
                             const int32_t sz_c[3], int32_t seed, int32_t rank, int32_t out_shape[3]);
/* h_sk: complex (re, im) doubles of out_shape, Fortran order.  Phases from the bit-exact LCG
 * (seed + 1000000), rounded to single precision as cmplx() without KIND does (:80). */
int scarl_hip_build_sk(int32_t corrMod, const double xRange[3], const double corrL[3], const int32_t Np[3],
                       const int32_t sz_c[3], int32_t seed, int32_t rank, double* h_sk);
/* gen_std_gauss_Shino_FFT (SRC_F/fftw_ScaRL.f90:157-255; the reference driver does not call it):
 * Sk quadrants on the 2^ceil(log2(2 Np)) padded grid, 3D C2C backward, real part of the first Np. */
int scarl_hip_gen_std_gauss_shino_fft(double* h_out, const int32_t Np[3], const double xRange[3],
                                      const double corrL[3], const int32_t pointsPerCorrL[3], int32_t corrMod,
                                      int32_t seed, int32_t rank);

/* ---- normal_ScaRL drop-in (bit-exact Park-Miller LCG by jump-ahead) ---------------------- */
/* r8_uniform_01 (SRC_F/normal_ScaRL.f90:746-841) */
int scarl_hip_r8_uniform_01(int32_t* seed, double* r);
/* r8vec_uniform_01 (SRC_F/normal_ScaRL.f90:1435-1515); *seed is updated as the reference does */
int scarl_hip_r8vec_uniform_01(int32_t n, int32_t* seed, double* h_r);
/* r8vec_normal_01 (SRC_F/normal_ScaRL.f90:1138-1259), with the reference's odd-n tail */
int scarl_hip_r8vec_normal_01(int32_t n, int32_t* seed, double* h_x);

/* ---- main_ScaRL element / stencil routines (FP64, host arrays) --------------------------- */
/* apply_UnityPartition_mtx (SRC_F/main_ScaRL.f90:1216-1301), partitionType must be 1 */
int scarl_hip_apply_unity_partition(double* h_field, const int32_t Np[3], const int32_t Np_ovlp[3],
                                    int32_t partition_type, const int32_t topo_pos[3], const int32_t topo_shape[3]);
/* multiVariateTransformation (SRC_F/main_ScaRL.f90:1355-1395) */
int scarl_hip_multivariate_transformation(double avg, double std_dev, int32_t margiFirst, double* h_field, int64_t n);
/* normalize_field (SRC_F/main_ScaRL.f90:903-953) on one rank: the reference sums the non-duplicated
 * box of every rank with MPI_ALLREDUCE, so topo_shape must be (1,1,1) here (otherwise
 * SCARL_E_UNSUPPORTED).  h_field (Np, Fortran order) is normalised in place; avg_out / std_out
 * (nullable) receive the values used. */
int scarl_hip_normalize_field(double* h_field, const int32_t Np[3], const int32_t Np_ovlp[3], const int32_t topo_pos[3],
                              const int32_t topo_shape[3], double* avg_out, double* std_out);

/* ---- Stat_ScaRL moments ------------------------------------------------------------------ */
/* calculate_average_and_stdVar_MPI on one rank (SRC_F/Stat_ScaRL/statistics_RF.f90:18-174):
 * h_rf is (n, nmc); FP64 compensated sums on the device. */
int scarl_hip_moments(const double* h_rf, int64_t n, int32_t nmc, double* evnt_avg, double* evnt_std,
                      double* g_avg, double* g_std);
/* rebuild_Sk_FFT (SRC_F/Stat_ScaRL/statistics_RF.f90:258-462): h_sk_part receives xNStep(dir)
 * values (dir and nDim 1-based as in the reference; the taper applies when nDim = 3). */
int scarl_hip_rebuild_sk_fft(const double* h_field, const int32_t xNStep[3], int32_t dir, int32_t nDim, double avg,
                             double stdDev, double* h_sk_part);
/* rebuild_corrL (SRC_F/Stat_ScaRL/statistics_RF.f90:468-486): sk_tot holds nhalf(1) + ... +
 * nhalf(nDim) values (SkTot_Dir), corrL receives nDim values. */
int scarl_hip_rebuild_corrl(const double* sk_tot, const int32_t nhalf[3], int32_t nDim, const double xStep[3],
                            double* corrL);
/* Per-axis centred covariance at lags 0..nlag-1 (no counterpart in the reference: the empirical
 * covariance that checks a generated field against its target): h_cov[axis][lag], FP64 with
 * compensated combination, NaN for lags beyond the extent. */
int scarl_hip_lag_covariance(const double* h_field, const int32_t n[3], double mean, int32_t nlag, double* h_cov);

/* ---- spectra / special functions --------------------------------------------------------- */
/* K_nu(x) on the device in FP64 (the ikv scheme of SRC_F/bessel.f90:13522-13785) */
int scarl_hip_kv(double nu, const double* h_x, int32_t n, double* h_out);
/* target correlation C(rho)/sigma^2 for model (Gaussian exp(-rho^2), exponential exp(-rho),
 * von Karman 2^(1-nu)/Gamma(nu) rho^nu K_nu(rho)), rho = scaled lag |r/L| */
int scarl_hip_target_corr(int32_t model, double nu, const double* h_rho, int32_t n, double* h_out);

/* ---- Hermitian spectral generator (mode H), single field and batches ---------------------- */
typedef struct scarl_field_desc {
    int32_t n[3];          /* extents (x, y, z); n[2] = 1 for a 2D field; any n[0], odd or even  */
    double dx[3];          /* grid spacing                                                       */
    int32_t model;         /* SCARL_CM_*                                                         */
    double corrL[3];       /* correlation lengths, same units as dx                             */
    double nu;             /* von Karman smoothness                                              */
    double sigma;          /* standard deviation of the underlying Gaussian field (1 = standard) */
    int32_t margi;         /* SCARL_FOM_*                                                        */
    double avg, std_dev;   /* marginal parameters, as in the reference input deck                */
    uint64_t seed;         /* Philox key                                                          */
    uint32_t stream;       /* Philox stream (counter word 2)                                      */
    int32_t normalize;     /* 1 = scale so the discrete spectrum carries exactly sigma^2          */
} scarl_field_desc;

typedef struct scarl_batch scarl_batch;

enum { SCARL_PREC_FP32 = 32, SCARL_PREC_FP64 = 64 };

/* Independent fields (realisations) in one batch: every field gets its own output array and
 * its own moment slot (slot i = field i). */
int scarl_hip_batch_create(const scarl_field_desc* fields, int32_t count, int32_t precision, scarl_batch** h);
/* Overlapping subdomains of one global field (extents gdim): nb[d] positions per axis with
 * starts/extents (concatenated x list, then y, then z).  Subdomain (i,j,k) gets extents
 * (ext_x[i], ext_y[j], ext_z[k]), Philox stream = linear subdomain index, parameters from
 * *params (n is ignored).  Overlaps are blended with sqrt of cosine partition-of-unity
 * weights (SRC_F/main_ScaRL.f90:1216-1301) and summed (add_overlap, :984-1210); the marginal
 * is applied after the sum; moment slot 0 = the global field. */
int scarl_hip_blend_create(const int32_t gdim[3], const int32_t nb[3], const int32_t* starts, const int32_t* exts,
                           const scarl_field_desc* params, int32_t precision, scarl_batch** h);
/* host-exported Hermitian half-spectrum noise for one field: complex (n1/2+1)*n2*n3, interleaved
 * (re, im) in the batch precision; NULL returns the field to device Philox noise */
int scarl_hip_batch_set_noise(scarl_batch* h, int32_t field, const void* h_noise);
/* enqueue one generation of every field; use_graph = 1 replays a captured hipGraph */
int scarl_hip_batch_run(scarl_batch* h, int32_t use_graph);
int scarl_hip_batch_sync(scarl_batch* h);
/* device pointer and extents of an output (field index, or -1 for the blended global field) */
int scarl_hip_batch_output(scarl_batch* h, int32_t field, void** d_ptr, int64_t dims[3]);
int scarl_hip_batch_copy_output(scarl_batch* h, int32_t field, void* h_dst, int64_t nbytes);
/* per slot {sum, sum of squares, count, mean, std}; nslots returned via *count */
int scarl_hip_batch_moments(scarl_batch* h, double* h_out, int32_t max_slots, int32_t* count);
/* after scarl_hip_batch_run of an ensemble: per-axis covariance of every field at lags 0..nlag-1,
 * centred on the field's sample mean (from the batch moments, on the device); h_cov is
 * [field][axis][lag].  Blended batches: SCARL_E_UNSUPPORTED (use scarl_hip_lag_covariance). */
int scarl_hip_batch_lag_cov(scarl_batch* h, int32_t nlag, double* h_cov);
/* amplitude prefactor and the FP64 spectrum sum used for the variance normalisation */
int scarl_hip_batch_field_info(scarl_batch* h, int32_t field, double* amp0, double* spectrum_sum);
/* one line per plan of the batch: pass (gen / y / x), length, kernel, radices, pencils per tile,
 * threads per pencil, LDS pitch and skew, how it was chosen, and the autotuner's timings */
int scarl_hip_batch_plan_report(scarl_batch* h, char* buf, int32_t len);
/* bytes of device memory the batch holds */
int scarl_hip_batch_bytes(scarl_batch* h, int64_t* bytes);
int scarl_hip_batch_destroy(scarl_batch* h);

/* Single field convenience: Hermitian generator + marginal (h_noise may be NULL). */
int scarl_hip_gen_field_h(void* h_out, const scarl_field_desc* d, int32_t precision, const void* h_noise);

#ifdef __cplusplus
}
#endif
#endif /* SCARL_HIP_H */
