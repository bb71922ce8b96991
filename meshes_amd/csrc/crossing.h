// crossing.h -- device-side transfer matrices of 2x2 crossings, in double precision.
//
// The row tables that drive the propagation kernels are built once per call from the
// slot-layout parameters by these functions: sin/cos are evaluated in FP64 and each
// matrix element is rounded to FP32 once, so the FP32 layer loop never sees the error
// of a fast single-precision sincos (the reference builds T with __sincosf per block,
// gpu/crossing.cu:40-48).
//
// Formulas:
//   MZI          crossing.py:216-231, gpu/crossing.cu:36-84
//   SYM          crossing.py:377-381, gpu/crossing.cu:86-126 (derivatives: :104-114)
//   ORTH         gpu/crossing.cu:128-148
//   MZI_GENERIC  crossing.py:465-488 (T4 T3 T2 T1; derivatives by the product rule)
#pragma once
#include <hip/hip_runtime.h>

namespace mp {

struct dc {
    double r, i;
};
__device__ __forceinline__ dc mk(double r, double i) { return dc{r, i}; }
__device__ __forceinline__ dc cadd(dc a, dc b) { return dc{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ dc csub(dc a, dc b) { return dc{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ dc cmul(dc a, dc b) { return dc{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ dc cscale(dc a, double s) { return dc{a.r * s, a.i * s}; }
__device__ __forceinline__ dc cconj(dc a) { return dc{a.r, -a.i}; }
__device__ __forceinline__ dc cexpi(double t) {
    double s, c;
    sincos(t, &s, &c);
    return dc{c, s};
}
__device__ __forceinline__ dc cmuli(dc a) { return dc{-a.i, a.r}; }  // i*a
__device__ __forceinline__ double cabs(dc a) { return sqrt(a.r * a.r + a.i * a.i); }

struct M2 {
    dc a00, a01, a10, a11;
};
__device__ __forceinline__ M2 mzero() { return M2{mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)}; }
__device__ __forceinline__ M2 meye() { return M2{mk(1, 0), mk(0, 0), mk(0, 0), mk(1, 0)}; }
__device__ __forceinline__ M2 mmul(const M2& A, const M2& B) {
    return M2{cadd(cmul(A.a00, B.a00), cmul(A.a01, B.a10)), cadd(cmul(A.a00, B.a01), cmul(A.a01, B.a11)),
              cadd(cmul(A.a10, B.a00), cmul(A.a11, B.a10)), cadd(cmul(A.a10, B.a01), cmul(A.a11, B.a11))};
}
__device__ __forceinline__ M2 madd(const M2& A, const M2& B) {
    return M2{cadd(A.a00, B.a00), cadd(A.a01, B.a01), cadd(A.a10, B.a10), cadd(A.a11, B.a11)};
}
__device__ __forceinline__ M2 mscale(const M2& A, double s) {
    return M2{cscale(A.a00, s), cscale(A.a01, s), cscale(A.a10, s), cscale(A.a11, s)};
}
__device__ __forceinline__ M2 mscalec(const M2& A, dc s) {
    return M2{cmul(A.a00, s), cmul(A.a01, s), cmul(A.a10, s), cmul(A.a11, s)};
}
// Conjugate transpose.
__device__ __forceinline__ M2 mherm(const M2& A) {
    return M2{cconj(A.a00), cconj(A.a10), cconj(A.a01), cconj(A.a11)};
}
__device__ __forceinline__ dc mdet(const M2& A) { return csub(cmul(A.a00, A.a11), cmul(A.a01, A.a10)); }
// Inverse via the adjugate.
__device__ __forceinline__ M2 minv(const M2& A) {
    dc det = csub(cmul(A.a00, A.a11), cmul(A.a01, A.a10));
    double n = det.r * det.r + det.i * det.i;
    dc id = dc{det.r / n, -det.i / n};
    return M2{cmul(A.a11, id), cmul(cscale(A.a01, -1.0), id), cmul(cscale(A.a10, -1.0), id), cmul(A.a00, id)};
}

enum { X_MZI = 0, X_SYM = 1, X_ORTH = 2, X_GENERIC = 3 };

__host__ __device__ constexpr int n_phase_of(int X) { return X == X_ORTH ? 1 : 2; }
__host__ __device__ constexpr int n_split_of(int X) {
    return X == X_MZI ? 2 : X == X_SYM ? 1 : X == X_ORTH ? 0 : 14;
}

// Splitter of the generic MZI: [[cos(pi/4+x), i sin(pi/4+x)], [i sin(pi/4+x), cos(pi/4+x)]].
__device__ __forceinline__ M2 generic_splitter(double x) {
    double s, c;
    sincos(0.78539816339744830962 + x, &s, &c);
    return M2{mk(c, 0), mk(0, s), mk(0, s), mk(c, 0)};
}

// T and dT/dp_k (k < n_phase_of(X)) of one crossing.  s may be null (ideal splitters /
// zero imperfections).  dT may be null.
__device__ inline void crossing_eval(int X, const float* p, const float* s, M2* T, M2* dT) {
    if (X == X_MZI) {
        const double th = p[0], ph = p[1];
        const double a = s ? (double)s[0] : 0.0, b = s ? (double)s[1] : 0.0;
        double S, C, Sp, Cp, Sm, Cm;
        sincos(0.5 * th, &S, &C);
        sincos(a + b, &Sp, &Cp);
        sincos(a - b, &Sm, &Cm);
        const dc g = cexpi(0.5 * th), e = cexpi(ph);
        M2 M{cmul(e, mk(-C * Sp, S * Cm)), mk(-S * Sm, C * Cp), cmul(e, mk(S * Sm, C * Cp)), mk(-C * Sp, -S * Cm)};
        *T = mscalec(M, g);
        if (dT) {
            M2 dM{cmul(e, mk(0.5 * S * Sp, 0.5 * C * Cm)), mk(-0.5 * C * Sm, -0.5 * S * Cp),
                  cmul(e, mk(0.5 * C * Sm, -0.5 * S * Cp)), mk(0.5 * S * Sp, -0.5 * C * Cm)};
            M2 hT = mscalec(*T, mk(0.0, 0.5));
            dT[0] = madd(hT, mscalec(dM, g));
            dT[1] = M2{cmuli(T->a00), mk(0, 0), cmuli(T->a10), mk(0, 0)};
        }
    } else if (X == X_SYM) {
        const double th = p[0], ph = p[1];
        const double al = s ? (double)s[0] : 0.0;
        double S, C, S2, C2;
        sincos(0.5 * th, &S, &C);
        sincos(2.0 * al, &S2, &C2);
        const dc e = cexpi(ph), ec = cconj(e);
        *T = M2{cmul(e, mk(S, C * S2)), mk(0, C * C2), mk(0, C * C2), cmul(ec, mk(S, -C * S2))};
        if (dT) {
            dT[0] = M2{cmul(e, mk(0.5 * C, -0.5 * S * S2)), mk(0, -0.5 * S * C2), mk(0, -0.5 * S * C2),
                       cmul(ec, mk(0.5 * C, 0.5 * S * S2))};
            dT[1] = M2{cmuli(T->a00), mk(0, 0), mk(0, 0), cscale(cmuli(T->a11), -1.0)};
        }
    } else if (X == X_ORTH) {
        const double th = p[0];
        double S, C;
        sincos(0.5 * th, &S, &C);
        *T = M2{mk(S, 0), mk(-C, 0), mk(C, 0), mk(S, 0)};
        if (dT) dT[0] = M2{mk(0.5 * C, 0), mk(0.5 * S, 0), mk(-0.5 * S, 0), mk(0.5 * C, 0)};
    } else {  // X_GENERIC: (a, b, p11, p21, p12, p22, q11, q21, q12, q22, g11, g21, g12, g22)
        double q[14];
        for (int k = 0; k < 14; ++k) q[k] = s ? (double)s[k] : 0.0;
        const double th = p[0], ph = p[1];
        const dc t1a = cscale(cexpi(ph + q[2] + q[6]), exp(0.5 * q[10]));
        const dc t1b = cscale(cexpi(q[3] + q[7]), exp(0.5 * q[11]));
        const dc t3a = cscale(cexpi(th + q[4] + q[8]), exp(0.5 * q[12]));
        const dc t3b = cscale(cexpi(q[5] + q[9]), exp(0.5 * q[13]));
        const M2 T1{t1a, mk(0, 0), mk(0, 0), t1b};
        const M2 T3{t3a, mk(0, 0), mk(0, 0), t3b};
        const M2 T2 = generic_splitter(q[0]), T4 = generic_splitter(q[1]);
        const M2 T21 = mmul(T2, T1);
        *T = mmul(T4, mmul(T3, T21));
        if (dT) {
            const M2 dT3{cmuli(t3a), mk(0, 0), mk(0, 0), mk(0, 0)};
            dT[0] = mmul(T4, mmul(dT3, T21));
            dT[1] = M2{cmuli(T->a00), mk(0, 0), cmuli(T->a10), mk(0, 0)};
        }
    }
}

// Second derivatives contracted with a direction q: S_k = sum_m q_m d2T/dp_k dp_m, from T and
// dT/dp_k of the same crossing.  MZI and the generic MZI: d2T/dth2 = i dT/dth,
// d2T/dth dph = dT/dth diag(i, 0), d2T/dph2 = T diag(-1, 0).  SYM: d2T/dth2 = -T/4,
// d2T/dth dph = diag(i dT/dth_00, -i dT/dth_11), d2T/dph2 = diag(-T00, -T11).  ORTH: -T/4.
__device__ inline void crossing_d2q(int X, const M2& T, const M2* dT, const double* q, M2* S) {
    if (X == X_ORTH) {
        S[0] = mscale(T, -0.25 * q[0]);
        return;
    }
    M2 H00, H01, H11;
    if (X == X_SYM) {
        H00 = mscale(T, -0.25);
        H01 = M2{cmuli(dT[0].a00), mk(0, 0), mk(0, 0), cscale(cmuli(dT[0].a11), -1.0)};
        H11 = M2{cscale(T.a00, -1.0), mk(0, 0), mk(0, 0), cscale(T.a11, -1.0)};
    } else {
        H00 = M2{cmuli(dT[0].a00), cmuli(dT[0].a01), cmuli(dT[0].a10), cmuli(dT[0].a11)};
        H01 = M2{cmuli(dT[0].a00), mk(0, 0), cmuli(dT[0].a10), mk(0, 0)};
        H11 = M2{cscale(T.a00, -1.0), mk(0, 0), cscale(T.a10, -1.0), mk(0, 0)};
    }
    S[0] = madd(mscale(H00, q[0]), mscale(H01, q[1]));
    S[1] = madd(mscale(H01, q[0]), mscale(H11, q[1]));
}

}  // namespace mp
