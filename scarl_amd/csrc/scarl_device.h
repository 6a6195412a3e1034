// scarl_device.h -- device-side building blocks of the gfx950 spectral random-field engine.
//
//  * cplx<T> and its arithmetic (T = float or double);
//  * backward (+i) DFT codelets of size 2,3,4,5,7,8,11,13,16,32 held in registers;
//  * one Stockham radix-R stage with a per-stage twiddle table (twiddles computed on the host
//    in long double and rounded once to T);
//  * Philox4x32-10 and Box-Muller white noise (counter = spectral index, so any element can be
//    produced independently by whichever thread needs it);
//  * FP64 K_nu(x) (series / asymptotic K_{v0}, Wronskian, forward recurrence -- the scheme of
//    SRC_F/bessel.f90:13522-13785) and Gamma (Lanczos).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scarl {

template <typename T>
struct cplx {
    T x, y;
};

// Global-memory accesses through address space 1.  Pointers held in descriptors are generic to
// the compiler, whose FLAT accesses also count against the LDS counter (every wait for one then
// waits for LDS traffic too); these compile to global_load / global_store.
#define SCARL_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ cplx<T> gload(const cplx<T>* p) {
    const SCARL_GLOBAL T* q = (const SCARL_GLOBAL T*)p;
    return {q[0], q[1]};
}
// The same load with the non-temporal hint, for data read exactly once (global_load ... nt)
template <typename T>
__device__ __forceinline__ cplx<T> gload_nt(const cplx<T>* p) {
    const SCARL_GLOBAL T* q = (const SCARL_GLOBAL T*)p;
    return {__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1)};
}
template <typename T>
__device__ __forceinline__ void gstore(cplx<T>* p, cplx<T> v) {
    SCARL_GLOBAL T* q = (SCARL_GLOBAL T*)p;
    q[0] = v.x;
    q[1] = v.y;
}
template <typename T>
__device__ __forceinline__ T gload1(const T* p) { return *(const SCARL_GLOBAL T*)p; }
template <typename T>
__device__ __forceinline__ void gstore1(T* p, T v) { *(SCARL_GLOBAL T*)p = v; }

template <typename T>
__device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by +i
template <typename T>
__device__ __forceinline__ cplx<T> mul_i(cplx<T> a) { return {-a.y, a.x}; }
template <typename T>
__device__ __forceinline__ cplx<T> conj(cplx<T> a) { return {a.x, -a.y}; }

// ------------------------------------------------------------------------------------------
// Backward DFT codelets: v[q] <- sum_r v[r] exp(+2 pi i r q / R), natural order in and out.
// ------------------------------------------------------------------------------------------
#define SCARL_C(v) static_cast<T>(v)

template <typename T>
__device__ __forceinline__ void dft2(cplx<T>* v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <typename T>
__device__ __forceinline__ void dft3(cplx<T>* v) {
    const T h = SCARL_C(0.86602540378443864676372317075293618);  // sin(2pi/3)
    cplx<T> t = v[1] + v[2];
    cplx<T> m = {v[0].x - SCARL_C(0.5) * t.x, v[0].y - SCARL_C(0.5) * t.y};
    cplx<T> d = v[1] - v[2];
    cplx<T> s = {-h * d.y, h * d.x};
    v[0] = v[0] + t;
    v[1] = m + s;
    v[2] = m - s;
}

template <typename T>
__device__ __forceinline__ void dft4(cplx<T>& a, cplx<T>& b, cplx<T>& c, cplx<T>& d) {
    cplx<T> t0 = a + c, t1 = a - c, t2 = b + d, t3 = mul_i(b - d);
    a = t0 + t2;
    b = t1 + t3;
    c = t0 - t2;
    d = t1 - t3;
}

template <typename T>
__device__ __forceinline__ void dft5(cplx<T>* v) {
    const T c1 = SCARL_C(0.30901699437494742410229341718281906);   // cos(2pi/5)
    const T c2 = SCARL_C(-0.80901699437494742410229341718281906);  // cos(4pi/5)
    const T s1 = SCARL_C(0.95105651629515357211643933337938214);   // sin(2pi/5)
    const T s2 = SCARL_C(0.58778525229247312916870595463907277);   // sin(4pi/5)
    cplx<T> a1 = v[1] + v[4], b1 = v[1] - v[4], a2 = v[2] + v[3], b2 = v[2] - v[3];
    cplx<T> r1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
    cplx<T> r2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
    cplx<T> i1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
    cplx<T> i2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
    v[0] = v[0] + a1 + a2;
    v[1] = r1 + mul_i(i1);
    v[4] = r1 - mul_i(i1);
    v[2] = r2 + mul_i(i2);
    v[3] = r2 - mul_i(i2);
}

// Odd prime R by symmetric pairs: X_q = v0 + sum_{j=1}^{(R-1)/2} [(v_j + v_{R-j}) cos(2 pi j q/R)
//                                        + i (v_j - v_{R-j}) sin(2 pi j q/R)].
template <typename T, int R>
struct PrimeTable;
template <typename T>
struct PrimeTable<T, 7> {
    __device__ static T c(int t) {
        constexpr double cs[7] = {1.0, 0.62348980185873353052500488400423981, -0.22252093395631440428890256449679476,
                                  -0.90096886790241912623610231950744505, -0.90096886790241912623610231950744505,
                                  -0.22252093395631440428890256449679476, 0.62348980185873353052500488400423981};
        return static_cast<T>(cs[t]);
    }
    __device__ static T s(int t) {
        constexpr double sn[7] = {0.0, 0.78183148246802980870844452667405775, 0.97492791218182360701813168299393122,
                                  0.43388373911755812047576833284835875, -0.43388373911755812047576833284835875,
                                  -0.97492791218182360701813168299393122, -0.78183148246802980870844452667405775};
        return static_cast<T>(sn[t]);
    }
};
template <typename T>
struct PrimeTable<T, 11> {
    __device__ static T c(int t) {
        constexpr double cs[11] = {1.0,
                                   0.84125353283118116886181164892859349,
                                   0.41541501300188642552927414923589369,
                                   -0.14231483827328514044379266862956055,
                                   -0.65486073394528506405692507247389985,
                                   -0.95949297361449738989036805707508091,
                                   -0.95949297361449738989036805707508091,
                                   -0.65486073394528506405692507247389985,
                                   -0.14231483827328514044379266862956055,
                                   0.41541501300188642552927414923589369,
                                   0.84125353283118116886181164892859349};
        return static_cast<T>(cs[t]);
    }
    __device__ static T s(int t) {
        constexpr double sn[11] = {0.0,
                                   0.54064081745559758210361401178489657,
                                   0.90963199535451837141171538308460987,
                                   0.98982144188093273237609203778065208,
                                   0.75574957435425828376591176710212287,
                                   0.28173255684142969771141791715034290,
                                   -0.28173255684142969771141791715034290,
                                   -0.75574957435425828376591176710212287,
                                   -0.98982144188093273237609203778065208,
                                   -0.90963199535451837141171538308460987,
                                   -0.54064081745559758210361401178489657};
        return static_cast<T>(sn[t]);
    }
};
template <typename T>
struct PrimeTable<T, 13> {
    __device__ static T c(int t) {
        constexpr double cs[13] = {1.0,
                                   0.88545602565320989813594019631559370,
                                   0.56806474673115580251180755912857090,
                                   0.12053668025532305334906768745254200,
                                   -0.35460488704253562596963250942040893,
                                   -0.74851074817110109863463059970135363,
                                   -0.97094181742605202715698227629379247,
                                   -0.97094181742605202715698227629379247,
                                   -0.74851074817110109863463059970135363,
                                   -0.35460488704253562596963250942040893,
                                   0.12053668025532305334906768745254200,
                                   0.56806474673115580251180755912857090,
                                   0.88545602565320989813594019631559370};
        return static_cast<T>(cs[t]);
    }
    __device__ static T s(int t) {
        constexpr double sn[13] = {0.0,
                                   0.46472317204376854566250792943246338,
                                   0.82298386589365639457961580177006006,
                                   0.99270887409805399280075885668953524,
                                   0.93501624268541482343565040160436127,
                                   0.66312265824079520237922745931226735,
                                   0.23931566428755776714875378576895700,
                                   -0.23931566428755776714875378576895700,
                                   -0.66312265824079520237922745931226735,
                                   -0.93501624268541482343565040160436127,
                                   -0.99270887409805399280075885668953524,
                                   -0.82298386589365639457961580177006006,
                                   -0.46472317204376854566250792943246338};
        return static_cast<T>(sn[t]);
    }
};

template <typename T, int R>
__device__ __forceinline__ void dft_prime(cplx<T>* v) {
    constexpr int H = (R - 1) / 2;
    cplx<T> a[H + 1], b[H + 1];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        a[j] = v[j] + v[R - j];
        b[j] = v[j] - v[R - j];
    }
    cplx<T> out[R];
    cplx<T> s0 = v[0];
#pragma unroll
    for (int j = 1; j <= H; ++j) s0 = s0 + a[j];
    out[0] = s0;
#pragma unroll
    for (int q = 1; q <= H; ++q) {
        cplx<T> re = v[0], im = {SCARL_C(0), SCARL_C(0)};
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const int t = (j * q) % R;
            const T c = PrimeTable<T, R>::c(t), s = PrimeTable<T, R>::s(t);
            re.x += c * a[j].x;
            re.y += c * a[j].y;
            im.x += s * b[j].x;
            im.y += s * b[j].y;
        }
        out[q] = re + mul_i(im);
        out[R - q] = re - mul_i(im);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
}

template <typename T>
__device__ __forceinline__ void dft8(cplx<T>* v) {
    const T r = SCARL_C(0.70710678118654752440084436210484904);
    cplx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cplx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    o1 = {r * (o1.x - o1.y), r * (o1.x + o1.y)};    // * e^{i pi/4}
    o2 = mul_i(o2);                                  // * e^{i pi/2}
    o3 = {-r * (o3.x + o3.y), r * (o3.x - o3.y)};   // * e^{3 i pi/4}
    v[0] = e0 + o0;
    v[4] = e0 - o0;
    v[1] = e1 + o1;
    v[5] = e1 - o1;
    v[2] = e2 + o2;
    v[6] = e2 - o2;
    v[3] = e3 + o3;
    v[7] = e3 - o3;
}

// exp(+2 pi i t / 16), t in [0,16)
template <typename T>
__device__ __forceinline__ cplx<T> w16(int t) {
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
    return {static_cast<T>(c[t & 15]), static_cast<T>(c[(t + 12) & 15])};
}

// 16 = 4 x 4: X_{q+4p} = sum_j w16^{jq} w4^{jp} DFT4_m(v_{j+4m})[q].  In place: after the first
// DFT4s v[j + 4q] holds A_j[q]; the second DFT4 over j runs on v[4q .. 4q+3], which leaves
// X_{q+4p} at v[4q + p]; a final 4x4 transpose restores natural order.
template <typename T>
__device__ __forceinline__ void dft16(cplx<T>* v) {
#pragma unroll
    for (int j = 0; j < 4; ++j) dft4(v[j], v[j + 4], v[j + 8], v[j + 12]);
#pragma unroll
    for (int j = 1; j < 4; ++j)
#pragma unroll
        for (int q = 1; q < 4; ++q) v[j + 4 * q] = cmul(v[j + 4 * q], w16<T>(j * q));
#pragma unroll
    for (int q = 0; q < 4; ++q) dft4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = a + 1; b < 4; ++b) {
            const cplx<T> t = v[4 * a + b];
            v[4 * a + b] = v[4 * b + a];
            v[4 * b + a] = t;
        }
}

// exp(+2 pi i t / 32)
template <typename T>
__device__ __forceinline__ cplx<T> w32(int t) {
    constexpr double c[32] = {1.0,
                              0.98078528040323044912618223613423903,
                              0.92387953251128675612818318939678829,
                              0.83146961230254523707878837761790576,
                              0.70710678118654752440084436210484904,
                              0.55557023301960222474283081394853287,
                              0.38268343236508977172845998403039887,
                              0.19509032201612826784828486847702224,
                              0.0,
                              -0.19509032201612826784828486847702224,
                              -0.38268343236508977172845998403039887,
                              -0.55557023301960222474283081394853287,
                              -0.70710678118654752440084436210484904,
                              -0.83146961230254523707878837761790576,
                              -0.92387953251128675612818318939678829,
                              -0.98078528040323044912618223613423903,
                              -1.0,
                              -0.98078528040323044912618223613423903,
                              -0.92387953251128675612818318939678829,
                              -0.83146961230254523707878837761790576,
                              -0.70710678118654752440084436210484904,
                              -0.55557023301960222474283081394853287,
                              -0.38268343236508977172845998403039887,
                              -0.19509032201612826784828486847702224,
                              0.0,
                              0.19509032201612826784828486847702224,
                              0.38268343236508977172845998403039887,
                              0.55557023301960222474283081394853287,
                              0.70710678118654752440084436210484904,
                              0.83146961230254523707878837761790576,
                              0.92387953251128675612818318939678829,
                              0.98078528040323044912618223613423903};
    return {static_cast<T>(c[t & 31]), static_cast<T>(c[(t + 24) & 31])};
}

// 32 = 8 x 4: A_j = DFT8_m(v_{j+4m}), A_j[q] *= w32^{jq}, X_{q+8p} = DFT4_j(A_j[q])[p].
// In place through an 8-element gather per j (registers are renamed, not copied, once unrolled).
template <typename T>
__device__ __forceinline__ void dft32(cplx<T>* v) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        cplx<T> a[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) a[m] = v[j + 4 * m];
        dft8(a);
#pragma unroll
        for (int q = 0; q < 8; ++q) v[j + 4 * q] = (j && q) ? cmul(a[q], w32<T>(j * q)) : a[q];
    }
    // v[j + 4q] = A_j[q] (twiddled); DFT4 over j for each q -> X_{q + 8p}
#pragma unroll
    for (int q = 0; q < 8; ++q) dft4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    // now v[4q + p] = X_{q + 8p}: transpose the 8 x 4 index map into natural order
    cplx<T> o[32];
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) o[q + 8 * p] = v[4 * q + p];
#pragma unroll
    for (int i = 0; i < 32; ++i) v[i] = o[i];
}

// exp(+2 pi i m / R) for the composite codelets (R = 6, 9, 12; m < R), as literals
template <int R, typename T>
__device__ __forceinline__ cplx<T> unit_root(int m) {
    constexpr double h = 0.86602540378443864676372317075293618;  // sin(60 deg)
    if constexpr (R == 6) {
        constexpr double c[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
        constexpr double s[6] = {0.0, h, h, 0.0, -h, -h};
        return {SCARL_C(c[m]), SCARL_C(s[m])};
    } else if constexpr (R == 9) {
        constexpr double c40 = 0.76604444311897803520239265055541667, s40 = 0.64278760968653932632264340990726343;
        constexpr double c80 = 0.17364817766693034885171662676931480, s80 = 0.98480775301220805936674302458952301;
        constexpr double c160 = -0.93969262078590838405410927732473146, s160 = 0.34202014332566873304409961468225958;
        constexpr double c[9] = {1.0, c40, c80, -0.5, c160, c160, -0.5, c80, c40};
        constexpr double s[9] = {0.0, s40, s80, h, s160, -s160, -h, -s80, -s40};
        return {SCARL_C(c[m]), SCARL_C(s[m])};
    } else {
        static_assert(R == 12, "composite codelet roots");
        constexpr double c[12] = {1.0, h, 0.5, 0.0, -0.5, -h, -1.0, -h, -0.5, 0.0, 0.5, h};
        constexpr double s[12] = {0.0, 0.5, h, 1.0, h, 0.5, 0.0, -0.5, -h, -1.0, -h, -0.5};
        return {SCARL_C(c[m]), SCARL_C(s[m])};
    }
}

template <int R, typename T>
__device__ __forceinline__ void dft(cplx<T>* v);

// Backward DFT of R = R1 R2 points in registers (Cooley-Tukey): R2 DFTs of R1 over the inputs
// n = R2 n1 + n2, twiddles exp(+2 pi i n2 k1 / R), then R1 DFTs of R2; output k1 + R1 k2.
template <typename T, int R1, int R2>
__device__ __forceinline__ void dft_ct(cplx<T>* v) {
    constexpr int R = R1 * R2;
    cplx<T> y[R2][R1];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) y[n2][n1] = v[R2 * n1 + n2];
        dft<R1, T>(y[n2]);
    }
#pragma unroll
    for (int n2 = 1; n2 < R2; ++n2)
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) y[n2][k1] = cmul(y[n2][k1], unit_root<R, T>(n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
        cplx<T> b[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) b[n2] = y[n2][k1];
        dft<R2, T>(b);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
    }
}

template <int R, typename T>
__device__ __forceinline__ void dft(cplx<T>* v) {
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v);
    else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 5) dft5(v);
    else if constexpr (R == 7) dft_prime<T, 7>(v);
    else if constexpr (R == 8) dft8(v);
    else if constexpr (R == 11) dft_prime<T, 11>(v);
    else if constexpr (R == 13) dft_prime<T, 13>(v);
    else if constexpr (R == 16) dft16(v);
    else if constexpr (R == 32) dft32(v);
    else if constexpr (R == 6) dft_ct<T, 2, 3>(v);
    else if constexpr (R == 9) dft_ct<T, 3, 3>(v);
    else if constexpr (R == 12) dft_ct<T, 4, 3>(v);
}

// ------------------------------------------------------------------------------------------
// Stockham radix-R stage (autosort):
//   for j in [0, N/R): v_r = src[j + r N/R]; k = j mod Ns; v_r *= exp(+2 pi i k r/(Ns R));
//   V = DFT_R(v); dst[(j - k) R + k + r Ns] = V_r.
// The per-stage twiddle table holds exp(+2 pi i k r/(Ns R)) at [k (R-1) + r - 1].
// A thread owns the butterflies j = b, b + tpp, ... (at most ECAP/R of them).
// When `sync` is set, a block-wide barrier separates all reads from all writes, so the
// stage may run in place on an LDS tile.
// ------------------------------------------------------------------------------------------
template <int R, int ECAP, typename T, class Load, class Store>
__device__ __forceinline__ void stockham_stage(int N, int Ns, const cplx<T>* __restrict__ tw, int b, int tpp,
                                               bool active, bool sync, Load&& ld, Store&& st) {
    constexpr int BMAX = (ECAP / R) > 0 ? (ECAP / R) : 1;
    const int nb = N / R;
    cplx<T> v[BMAX][R];
#pragma unroll
    for (int i = 0; i < BMAX; ++i) {
        const int j = b + tpp * i;
        if (active && j < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = ld(j + r * nb);
            if (Ns > 1) {
                const int k = j % Ns;
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
        const int j = b + tpp * i;
        if (active && j < nb) {
            const int k = j % Ns;
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) st(base + r * Ns, v[i][r]);
        }
    }
}

// Dispatch one stage on a runtime radix.  Radices above ECAP are never planned; the composite
// radices 6, 9 and 12 only exist in the COMP instantiations (their plans have their own kernels, so
// that the other runtime-planned kernels do not carry their code).
template <int ECAP, typename T, bool COMP = false, class Load, class Store>
__device__ __forceinline__ void stage_dispatch(int R, int N, int Ns, const cplx<T>* __restrict__ tw, int b, int tpp,
                                               bool active, bool sync, Load&& ld, Store&& st) {
    if constexpr (COMP) {
        switch (R) {
        case 6: stockham_stage<6, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); return;
        case 9: stockham_stage<9, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); return;
        case 12: stockham_stage<12, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); return;
        default: break;
        }
    }
    switch (R) {
    case 2: stockham_stage<2, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); break;
    case 3: stockham_stage<3, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); break;
    case 4: stockham_stage<4, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); break;
    case 5: stockham_stage<5, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); break;
    case 7: stockham_stage<7, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); break;
    case 8: stockham_stage<8, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); break;
    case 11: stockham_stage<11, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); break;
    case 13: stockham_stage<13, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st); break;
    case 16:
        if constexpr (ECAP >= 16) stockham_stage<16, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st);
        break;
    case 32:
        if constexpr (ECAP >= 32) stockham_stage<32, ECAP, T>(N, Ns, tw, b, tpp, active, sync, ld, st);
        break;
    default: break;
    }
}

// ------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller on 24-bit uniforms.
// Counter = (idx_lo, idx_hi, stream, 0), key = (seed_lo, seed_hi).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32x32 -> 64 product per multiplier (v_mad_u64_u32) rather than separate low and high
        // multiplies: 37 % more Philox calls per second on gfx950
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
        // three-input XOR in one v_bitop3_b32 (truth table 0x96), which the compiler does not form
        // from a ^ b ^ c by itself
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32(static_cast<uint32_t>(p1 >> 32), c[1], k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32(static_cast<uint32_t>(p0 >> 32), c[3], k1, 0x96);
        c[0] = n0;
        c[1] = static_cast<uint32_t>(p1);
        c[2] = n2;
        c[3] = static_cast<uint32_t>(p0);
    }
}

__device__ __forceinline__ void philox_at(uint64_t idx, uint64_t seed, uint32_t stream, uint32_t o[4]) {
    o[0] = static_cast<uint32_t>(idx);
    o[1] = static_cast<uint32_t>(idx >> 32);
    o[2] = stream;
    o[3] = 0u;
    philox4x32_10(o, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
}

// Within each 16-lane row of a wave, lane j receives the value of lane (16 - j) mod 16: DPP
// row_mirror (lane j <- lane 15 - j), then DPP row_shr:1 (lane j <- lane j - 1). Lane 0 keeps the
// mirrored value of lane 15 and must be overridden by the caller. Every lane of the row must
// execute it.
__device__ __forceinline__ float row_partner16(float v) {
    const int m = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false);
    return __int_as_float(__builtin_amdgcn_update_dpp(m, m, 0x111, 0xF, 0xF, false));
}

template <typename T>
__device__ __forceinline__ T u24(uint32_t x) {
    return static_cast<T>(x >> 8) * static_cast<T>(5.9604644775390625e-08) + static_cast<T>(2.98023223876953125e-08);
}

// Box-Muller on two 24-bit midpoint uniforms: (sqrt(-2 ln u1) cos 2 pi u2, sqrt(-2 ln u1) sin 2 pi u2).
// FP32 uses the hardware transcendentals (v_log_f32 is log2; v_sin/v_cos take revolutions, so
// sin(2 pi u2) is one instruction); FP64 uses the correctly rounded library functions.
template <typename T>
__device__ __forceinline__ cplx<T> box_muller(uint32_t a, uint32_t b) {
    const T u1 = u24<T>(a), u2 = u24<T>(b);
    if constexpr (sizeof(T) == 4) {
        const float rad = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
        return {rad * __builtin_amdgcn_cosf(u2), rad * __builtin_amdgcn_sinf(u2)};
    } else {
        double s, c;
        const double rad = sqrt(-2.0 * log(u1));
        sincospi(2.0 * u2, &s, &c);
        return {rad * c, rad * s};
    }
}

// Standard-normal pair of a spectral point from the Philox block at counter ctr: outputs (0,1)
// for half = 0, outputs (2,3) for half = 1, so one call serves two points.
template <typename T>
__device__ __forceinline__ cplx<T> gauss_at(uint64_t ctr, int half, uint64_t seed, uint32_t stream) {
    uint32_t o[4];
    philox_at(ctr, seed, stream, o);
    return half ? box_muller<T>(o[2], o[3]) : box_muller<T>(o[0], o[1]);
}

// 53-bit uniform in [0,1)
__device__ __forceinline__ double uniform53(uint64_t idx, uint64_t seed, uint32_t stream) {
    uint32_t o[4];
    philox_at(idx, seed, stream, o);
    const uint64_t b = (static_cast<uint64_t>(o[1]) << 32) | o[0];
    return static_cast<double>(b >> 11) * 1.1102230246251565404236316680908203125e-16;
}

// ------------------------------------------------------------------------------------------
// FP64 Gamma (Lanczos g = 7, 9 terms) and K_nu.
// ------------------------------------------------------------------------------------------
__device__ inline double gamma_fp64(double x) {
    const double c[9] = {0.99999999999980993, 676.5203681218851, -1259.1392167224028,
                         771.32342877765313, -176.61502916214059, 12.507343278686905,
                         -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};
    const double pi = 3.14159265358979323846;
    double refl = 1.0;
    const bool reflect = x < 0.5;
    if (reflect) {
        refl = pi / sin(pi * x);
        x = 1.0 - x;
    }
    x -= 1.0;
    double a = c[0];
    const double t = x + 7.5;
    for (int i = 1; i < 9; ++i) a += c[i] / (x + (double)i);
    const double g = sqrt(2.0 * pi) * pow(t, x + 0.5) * exp(-t) * a;
    return reflect ? refl / g : g;
}

__device__ inline double envj_fp64(int n, double x) { return 0.5 * log10(6.28 * n) - n * log10(1.36 * x / n); }

__device__ inline int start_order1_fp64(double x, int mp) {
    const double a0 = fabs(x);
    int n0 = (int)(1.1 * a0) + 1;
    double f0 = envj_fp64(n0, a0) - mp;
    int n1 = n0 + 5;
    double f1 = envj_fp64(n1, a0) - mp;
    int nn = n1;
    for (int it = 0; it < 20; ++it) {
        nn = (int)(n1 - (n1 - n0) / (1.0 - f0 / f1));
        const double f = envj_fp64(nn, a0) - mp;
        if (abs(nn - n1) < 1) break;
        n0 = n1; f0 = f1; n1 = nn; f1 = f;
    }
    return nn;
}

__device__ inline int start_order2_fp64(double x, int n, int mp) {
    const double a0 = fabs(x), hmp = 0.5 * mp, ejn = envj_fp64(n, a0);
    double obj;
    int n0;
    if (ejn <= hmp) { obj = mp; n0 = (int)(1.1 * a0) + 1; }
    else { obj = hmp + ejn; n0 = n; }
    double f0 = envj_fp64(n0, a0) - obj;
    int n1 = n0 + 5;
    double f1 = envj_fp64(n1, a0) - obj;
    int nn = n1;
    for (int it = 0; it < 20; ++it) {
        nn = (int)(n1 - (n1 - n0) / (1.0 - f0 / f1));
        const double f = envj_fp64(nn, a0) - obj;
        if (abs(nn - n1) < 1) break;
        n0 = n1; f0 = f1; n1 = nn; f1 = f;
    }
    return nn + 10;
}

// The scheme below follows the routine IKV of Shanjie Zhang and Jianming Jin, "Computation of
// Special Functions" (Wiley, 1996); IKV is copyrighted by Shanjie Zhang and Jianming Jin, who give
// permission to incorporate it into a user program provided that the copyright is acknowledged
// (SRC_F/bessel.f90:13522-13532).  The start orders above (start_order1/2) follow their MSTA1
// and MSTA2 of the same book.
//
// K_nu(x), nu >= 0, x > 0: K_{v0} by series (x <= 9) or asymptotic expansion, I_{v0} by
// series (x <= 18) or asymptotic, I_{v0+1}/I_{v0} from a backward recurrence, K_{v0+1} from
// the Wronskian I_v K_{v+1} + I_{v+1} K_v = 1/x, then the forward recurrence to K_nu.
__device__ inline double kv_fp64(double nu, double x) {
    const double pi = 3.14159265358979323846;
    if (!(x > 0.0)) return __builtin_inf();
    const int n = (int)nu;
    const double v0 = nu - n, x2 = x * x, piv = pi * v0, vt = 4.0 * v0 * v0;
    const int k0 = (x < 35.0) ? 14 : (x < 50.0 ? 10 : 8);
    const double gap = gamma_fp64(1.0 + v0);
    const double a1 = (v0 == 0.0) ? 1.0 : pow(0.5 * x, v0) / gap;
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
        const double ca = exp(x) / sqrt(2.0 * pi * x);
        double s = 1.0, r = 1.0;
        for (int k = 1; k <= k0; ++k) {
            r = -0.125 * r * (vt - (2.0 * k - 1.0) * (2.0 * k - 1.0)) / (k * x);
            s += r;
        }
        bi0 = ca * s;
    }
    int ntop = 1;
    int mstart = start_order1_fp64(x, 200);
    if (mstart < ntop) ntop = mstart;
    else mstart = start_order2_fp64(x, ntop, 15);
    double f2 = 0.0, f1 = 1e-100, f = 0.0, b0 = 0.0, b1 = 0.0;
    for (int k = mstart; k >= 0; --k) {
        f = 2.0 * (v0 + k + 1.0) / x * f1 + f2;
        if (k == 1) b1 = f;
        if (k == 0) b0 = f;
        f2 = f1;
        f1 = f;
    }
    const double cs = bi0 / f;
    const double I0 = cs * b0, I1 = cs * b1;
    double bk0;
    if (x <= 9.0) {
        if (v0 == 0.0) {
            const double ct = -log(0.5 * x) - 0.5772156649015329;
            double s = 0.0, w0 = 0.0, r = 1.0, ww = 0.0;
            for (int k = 1; k <= 50; ++k) {
                w0 += 1.0 / k;
                r = 0.25 * r / ((double)k * k) * x2;
                s += r * (w0 + ct);
                const double wa = fabs(s);
                if (k > 1 && fabs((wa - ww) / wa) < 1e-15) break;
                ww = wa;
            }
            bk0 = ct + s;
        } else {
            const double gan = gamma_fp64(1.0 - v0);
            const double a2 = 1.0 / (gan * pow(0.5 * x, v0));
            double s = a2 - a1, r1 = 1.0, r2 = 1.0, ww = 0.0;
            for (int k = 1; k <= 120; ++k) {
                r1 = 0.25 * r1 * x2 / (k * (k - v0));
                r2 = 0.25 * r2 * x2 / (k * (k + v0));
                s += a2 * r1 - a1 * r2;
                const double wa = fabs(s);
                if (k > 1 && fabs((wa - ww) / wa) < 1e-15) break;
                ww = wa;
            }
            bk0 = 0.5 * pi * s / sin(piv);
        }
    } else {
        const double cb = exp(-x) * sqrt(0.5 * pi / x);
        double s = 1.0, r = 1.0;
        for (int k = 1; k <= k0; ++k) {
            r = 0.125 * r * (vt - (2.0 * k - 1.0) * (2.0 * k - 1.0)) / (k * x);
            s += r;
        }
        bk0 = cb * s;
    }
    if (n == 0) return bk0;
    double bk1 = (1.0 / x - I1 * bk0) / I0;
    for (int k = 2; k <= n; ++k) {
        const double bk2 = 2.0 * (v0 + k - 1.0) / x * bk1 + bk0;
        bk0 = bk1;
        bk1 = bk2;
    }
    return bk1;
}

}  // namespace scarl
