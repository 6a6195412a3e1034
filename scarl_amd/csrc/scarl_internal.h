// scarl_internal.h -- structures shared by the host runtime and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scarl {

constexpr int kBlock = 256;      // threads per workgroup for every FFT pass
constexpr int kMaxStages = 8;
constexpr int kEcap64 = 16;      // complex elements a thread holds per stage (FP64)

enum Layout { LAYOUT_PF = 0, LAYOUT_BF = 1 };  // pencil-fastest lanes / butterfly-fastest lanes

// One 1D transform length, planned for one precision and one lane layout.
struct PlanDev {
    int N;                 // complex transform length
    int nst;               // Stockham stages (0 = direct O(N^2) DFT fallback)
    int R[kMaxStages];     // radices
    int Ns[kMaxStages];    // product of the radices before each stage
    int tw_off[kMaxStages];// per-stage twiddle table offset (complex elements) in the pool
    int direct_off;        // direct-DFT table (N entries) offset, or -1
    int pre_off;           // pre-twiddle table offset (C2R: exp(2 pi i k/(2N)), k<N; DCT: exp(i pi j/(2N)), j<N)
    int tpp;               // threads per pencil
    int T;                 // pencils per workgroup
    int ldp;               // LDS pitch in complex elements (LAYOUT_BF)
    int lds_elems;         // complex elements of LDS the tile needs
    int ecap;              // complex elements a thread holds per stage (generic kernel instantiation)
    int spec;              // compile-time planned kernel id (SCARL_FIXED_SPECS), 0 = generic
    int skew;              // LAYOUT_BF runtime plans: 1 = one LDS element of skew per 16 (ldp includes it)
    int blu_M;             // > 0: Bluestein, the stages are those of the inner length blu_M >= 2N - 1
    int chirp_off;         // Bluestein chirp exp(+i pi j^2/N), j < N
    int kern_off;          // Bluestein kernel: the length-blu_M backward DFT of exp(-i pi m^2/N) (wrapped), / blu_M
};

enum SourceMode { SRC_WORKSPACE = 0, SRC_DEVICE_NOISE = 1, SRC_HOST_NOISE = 2 };
enum ModelCode { CM_GAUSSIAN = 1, CM_VON_KARMAN = 2, CM_EXPONENTIAL = 3 };
enum MarginCode { FOM_NONE = 0, FOM_GAUSSIAN = 1, FOM_LOGNORMAL = 2 };

// Mode H field (Hermitian half spectrum -> C2R) descriptor, device resident.
struct FieldH {
    int n[3];              // output extents x, y, z
    int h1;                // n1/2 + 1
    int plan_z, plan_y, plan_x;   // plan indices (z/y: LAYOUT_PF complex length n3/n2; x: LAYOUT_BF length n1/2)
    int first_axis;        // 2 = z pass generates, 1 = y pass generates (2D)
    int src_mode;          // SourceMode of the generating pass
    int model;             // ModelCode
    double alpha[3];       // kappa_d L_d per unit signed frequency = 2 pi L_d / (n_d dx_d)
    double expo;           // Matern: -(nu + dim/2)/2 ; Gaussian: unused
    double amp0;           // amplitude prefactor (FP64, includes the variance normalisation)
    uint64_t seed;
    uint32_t stream;
    int wp;                // workspace row length in complex elements: n1/2 (half spectrum) or n1
    int nyq;               // 1: the kx = n1/2 column is a separate (n2, n3) plane after the rows
    int pad0;
    const void* W;         // host-exported noise (device copy), complex, dense (h1, n2, n3) layout
    const void* kx0;       // 3D fields: a(k) W(k) of the kx = 0 plane, complex, dense (n3, n2), made
                           // by k_h_kx0 before the z pass (device noise), or null
    void* ws;              // complex workspace: rows of wp elements (planes wz apart), then the dense
                           // (n2, n3) Nyquist plane (nyq)
    void* out;             // output base (real)
    int64_t out_ld1, out_ld2;  // output strides in elements along y and z
    int64_t wz;            // workspace plane pitch (elements between kz planes of the rows, >= wp n2)
    // blend (partition of unity over overlapping subdomains, fused into the x pass)
    int color;             // bit d = parity of the subdomain position along axis d
    int blend;             // 0: plain store; 1: PoU-weighted first/middle/last accumulation
    int nov_lo[3], nov_hi[3];  // overlap lengths on sides that have a neighbour (0 = none)
    // marginal + moments
    int margi;
    int mom_slot;          // -1 = no moments
    double mu, sd;         // marginal parameters of the underlying Gaussian
};

// Mode R (REDFT01 reference parity, FP64) field descriptor.
struct FieldR {
    int n[3];              // output extents
    int m[3];              // support of the octant amplitude (m_d <= n_d)
    int plan[3];           // per-axis plan (x: LAYOUT_BF, y/z: LAYOUT_PF)
    double dk[3];          // reference k spacing delta_k_d / corrL_d
    double corrL[3];
    int src_mode;          // SRC_DEVICE_NOISE (phi from Philox) or SRC_HOST_NOISE (phi array)
    int margi;
    double mu, sd;
    uint64_t seed;
    uint32_t stream;
    int pad0;
    const double* phi;     // (m1,m2,m3) phases
    double* tmp1;          // (n1, m2, m3)
    double* tmp2;          // (n1, n2, m3)
    double* out;           // (n1, n2, n3)
};

// Pencils of a strided pass: regular tiles cover P = kx + wp q (kx < wp) over the workspace rows,
// Nyquist tiles cover q of the separate kx = n1/2 plane (q = ky for the z pass, kz for the y pass).
struct Tile {
    int field;
    int pencil0;
    int nyq;
    int pad;
};

// Compile-time planned pass kernels: X(id, prec, layout, N, R0, R1, TPP); id 0 = generic runtime
// plan.  prec: 4 = FP32, 8 = FP64; layout: LAYOUT_PF (strided y/z passes) or LAYOUT_BF (x pass);
// a tile holds 256/TPP pencils and every tile stays within 64 KiB of LDS.  The first kernel listed
// for a length and layout is the default; the autotuner tries the others.
#define SCARL_FIXED_SPECS(X)                \
    X(1, 4, LAYOUT_PF, 512, 16, 32, 16)     \
    X(2, 4, LAYOUT_PF, 256, 16, 16, 16)     \
    X(3, 4, LAYOUT_PF, 128, 16, 8, 8)       \
    X(4, 4, LAYOUT_PF, 64, 8, 8, 8)         \
    X(5, 4, LAYOUT_PF, 96, 32, 3, 3)        \
    X(6, 4, LAYOUT_PF, 160, 32, 5, 5)       \
    X(7, 4, LAYOUT_PF, 176, 16, 11, 11)     \
    X(8, 4, LAYOUT_BF, 256, 16, 16, 16)     \
    X(9, 4, LAYOUT_BF, 128, 16, 8, 8)       \
    X(10, 4, LAYOUT_BF, 64, 8, 8, 8)        \
    X(24, 4, LAYOUT_BF, 96, 16, 6, 6)       \
    X(11, 4, LAYOUT_BF, 96, 32, 3, 4)       \
    X(12, 4, LAYOUT_BF, 80, 16, 5, 5)       \
    X(13, 4, LAYOUT_BF, 88, 8, 11, 8)       \
    X(14, 4, LAYOUT_BF, 48, 16, 3, 3)       \
    X(15, 4, LAYOUT_BF, 32, 8, 4, 4)        \
    X(16, 8, LAYOUT_PF, 256, 16, 16, 16)    \
    X(17, 8, LAYOUT_BF, 256, 16, 16, 32)    \
    X(18, 8, LAYOUT_BF, 128, 16, 8, 16)     \
    X(19, 4, LAYOUT_PF, 144, 16, 9, 9)      \
    X(20, 4, LAYOUT_PF, 192, 16, 12, 12)    \
    X(21, 4, LAYOUT_BF, 72, 8, 9, 9)        \
    X(22, 4, LAYOUT_PF, 96, 16, 6, 6)       \
    X(23, 4, LAYOUT_BF, 144, 16, 9, 9)

// The FP32 x pass of a 512-wide field (N = 256 = 16 x 16) on the matrix cores: split-precision
// bf16 MFMA radix-16 stages (scarl_mfma.h), selected by SCARL_TUNE_MFMA_DFT16.
constexpr int kSpecMfma256 = 64;
// Runtime-planned FP32 transforms of a length that is a multiple of 16 (at least 32) whose first,
// radix-16 stage runs as split-precision bf16 MFMA products (GenericMfmaFft in scarl_mfma.h);
// selected by SCARL_TUNE_MFMA_DFT16 like kSpecMfma256.  The tile holds at most 4096 complex
// elements (four 16-column groups per wave).
constexpr int kSpecMfmaGeneric = 65;
constexpr int kMfmaTileElems = 4096;
// Runtime plans of lengths with a prime factor above 13: Bluestein (blu_M > 0) or the direct
// O(N^2) DFT (nst = 0), run by their own kernel instantiation (NonSmoothFft in scarl_fft.h).
constexpr int kSpecNonSmooth = 66;
// Runtime plans with a composite radix 6, 9 or 12 (GenericCompFft in scarl_fft.h).
constexpr int kSpecGenericComp = 67;

// Blocks per field of the tile-partial reduction (k_mom_field); the per-field buffer holds
// kMomChunks partials per field.
constexpr int kMomChunks = 32;

// Compensated (Neumaier) partial sums of x and x^2 for one tile.
struct MomPartial {
    double s, c, s2, c2;
};

// Lag products of one array (x fastest) along one axis: for every pencil along `axis` and lag
// l < nlag, sum_m u_m u_{m+l}, u = (x - centre) * scale (times the taper along the axis when
// `window` is set).  circular: m + l wraps around the pencil (periodic autocorrelation, all m);
// otherwise only pairs inside the pencil (m < n - l).
struct LagJob {
    const void* x;
    int n[3];
    int axis;
    int64_t ld1, ld2;       // element strides along y and z
    int nlag;
    int circular;
    int window;             // taper t_m of the reference's corrL estimator (r = 0.1, as written)
    int T;                  // pencils per block
    const double* centre;   // device scalar, or nullptr: cval
    double cval;
    double scale;
    double* partial;        // [block of this job][nlag]
    int64_t nblocks;
    int fft_plan;           // >= 0: lag sums by FFT (k_lag_fft) over the FP64 plan of length P; -1: direct
    int P;                  // FFT length: n (circular) or >= n + nlag - 1 (zero padded, bounded pairs)
};
struct LagBlock {
    int job;
    int pencil0;
};

}  // namespace scarl
