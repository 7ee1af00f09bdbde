// svd3.h — 3x3 singular value decomposition for one particle per lane.
//
// Same algorithm and sign/ordering conventions as the reference's QR/Givens
// SVD (QRSVD.hpp:226-374 with givens.hpp:12-210): Givens upper
// bidiagonalisation, implicit Wilkinson-shift QR sweeps until a diagonal or
// super-diagonal entry drops below 128*eps*max(|B|/2, 1), then a 2x2 polar +
// symmetric eigen step on the remaining block and the reference's sort/sign
// fix-up, so that F = U diag(S) V^T with S0 >= S1 >= |S2| and only S2 < 0.
//
// Written for registers: matrices are plain T[3][3] arrays fully unrolled by
// the compiler; a Givens rotation is (i, k, c, s).  The sweep loop is bounded
// (kMaxSweeps) so that a non-converging input can never stall a wavefront; the
// reference loop is unbounded.  Templated on float and double (the stress
// update runs it in FP64 by default: FP64 is half the FP32 rate on CDNA4 and
// the update is bound by particle-state traffic).
#pragma once
#include <hip/hip_runtime.h>

namespace mpmx {

template <typename T> struct Eps;
template <> struct Eps<float>  { static constexpr float  v = 1.1920928955078125e-07f; };
template <> struct Eps<double> { static constexpr double v = 2.220446049250313080847e-16; };

constexpr int kMaxSweeps = 64;

template <typename T>
struct Rot {
    int i, k;
    T c, s;
};

template <typename T>
__device__ __forceinline__ Rot<T> rot_id(int i, int k) { return Rot<T>{i, k, T(1), T(0)}; }

// 1 / sqrt(d) for d > 0.  FP64: the hardware estimate refined by two Newton
// steps, y += y (1 - d y^2) / 2 (within about an ulp of 1/sqrt(d), a tenth of
// the instructions of a square root and a division); FP32: 1 / sqrtf.
__device__ __forceinline__ float rsqrt_acc(float d) { return 1.0f / sqrtf(d); }
__device__ __forceinline__ double rsqrt_acc(double d) {
    double y = __builtin_amdgcn_rsq(d);
    double e = fma(-d * y, y, 1.0);
    y = fma(y * 0.5, e, y);
    e = fma(-d * y, y, 1.0);
    return fma(y * 0.5, e, y);
}

// [c -s; s c] [a; b] = [*; 0]
template <typename T>
__device__ __forceinline__ void rot_conv(Rot<T>& g, T a, T b) {
    const T d = a * a + b * b;
    g.c = T(1); g.s = T(0);
    if (d != T(0)) { const T t = rsqrt_acc(d); g.c = a * t; g.s = -b * t; }
}
// [c -s; s c] [a; b] = [0; *]
template <typename T>
__device__ __forceinline__ void rot_unconv(Rot<T>& g, T a, T b) {
    const T d = a * a + b * b;
    g.c = T(0); g.s = T(1);
    if (d != T(0)) { const T t = rsqrt_acc(d); g.s = a * t; g.c = b * t; }
}
// rows i,k of A
template <typename T, int NC>
__device__ __forceinline__ void rot_rows(const Rot<T>& g, T (&A)[NC][NC]) {
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const T a = A[g.i][j], b = A[g.k][j];
        A[g.i][j] = g.c * a - g.s * b;
        A[g.k][j] = g.s * a + g.c * b;
    }
}
// columns i,k of A
template <typename T>
__device__ __forceinline__ void rot_cols(const Rot<T>& g, T (&A)[3][3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const T a = A[r][g.i], b = A[r][g.k];
        A[r][g.i] = g.c * a - g.s * b;
        A[r][g.k] = g.s * a + g.c * b;
    }
}

template <typename T>
__device__ __forceinline__ void eye3(T (&A)[3][3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) A[r][c] = (r == c) ? T(1) : T(0);
}

// bring [x x 0; x x x; 0 0 x] to upper bidiagonal, accumulating U and V
template <typename T>
__device__ __forceinline__ void chase(T (&H)[3][3], T (&U)[3][3], T (&V)[3][3]) {
    Rot<T> r1 = rot_id<T>(0, 1);
    rot_conv(r1, H[0][0], H[1][0]);
    Rot<T> r2 = rot_id<T>(1, 2);
    if (H[1][0] != T(0))
        rot_conv(r2, H[0][0] * H[0][1] + H[1][0] * H[1][1], H[0][0] * H[0][2] + H[1][0] * H[1][2]);
    else
        rot_conv(r2, H[0][1], H[0][2]);
    rot_rows(r1, H);
    rot_cols(r2, H);
    rot_cols(r2, V);
    Rot<T> r3 = rot_id<T>(1, 2);
    rot_conv(r3, H[1][1], H[2][1]);
    rot_rows(r3, H);
    rot_cols(r1, U);
    rot_cols(r3, U);
}

// 2x2 block at (t,t): polar rotation, then the symmetric 2x2 eigen rotation
template <typename T>
__device__ __forceinline__ void block2(int t, const T (&B)[3][3], T (&U)[3][3], T (&S)[3], T (&V)[3][3]) {
    const int other = (t == 1) ? 0 : 2;
    S[other] = B[other][other];
    T A2[2][2] = {{B[t][t], B[t][t + 1]}, {B[t + 1][t], B[t + 1][t + 1]}};
    Rot<T> u = rot_id<T>(0, 1), v = rot_id<T>(0, 1);
    // polar part
    const T x0 = A2[0][0] + A2[1][1], x1 = A2[1][0] - A2[0][1];
    const T d2 = x0 * x0 + x1 * x1;
    if (d2 != T(0)) { const T t = rsqrt_acc(d2); u.c = x0 * t; u.s = -x1 * t; }
    rot_rows(u, A2);
    // symmetric part
    const T x = A2[0][0], y = A2[0][1], z = A2[1][1];
    const T y2 = y * y;
    T cs, sn, s0, s1;
    if (y2 == T(0)) {
        cs = T(1); sn = T(0); s0 = x; s1 = z;
    } else {
        const T tau = T(0.5) * (x - z);
        const T w = sqrt(tau * tau + y2);
        const T tt = (tau > T(0)) ? y / (tau + w) : y / (tau - w);
        cs = rsqrt_acc(tt * tt + T(1));
        sn = -tt * cs;
        const T c2 = cs * cs, s2 = sn * sn, csy2 = T(2) * cs * sn * y;
        s0 = c2 * x - csy2 + s2 * z;
        s1 = s2 * x + csy2 + c2 * z;
    }
    if (s0 < s1) { const T tmp = s0; s0 = s1; s1 = tmp; v.c = -sn; v.s = cs; }
    else         { v.c = cs; v.s = sn; }
    // u <- u * v
    const T nc = u.c * v.c - u.s * v.s, ns = u.s * v.c + u.c * v.s;
    u.c = nc; u.s = ns;
    S[t] = s0;
    S[t + 1] = s1;
    u.i += t; u.k += t;
    v.i += t; v.k += t;
    rot_cols(u, U);
    rot_cols(v, V);
}

template <typename T>
__device__ __forceinline__ void swapv(T& a, T& b) { const T t = a; a = b; b = t; }
template <typename T>
__device__ __forceinline__ void swapc(T (&A)[3][3], int a, int b) {
#pragma unroll
    for (int r = 0; r < 3; ++r) swapv(A[r][a], A[r][b]);
}
template <typename T>
__device__ __forceinline__ void negc(T (&A)[3][3], int a) {
#pragma unroll
    for (int r = 0; r < 3; ++r) A[r][a] = -A[r][a];
}
template <typename T>
__device__ __forceinline__ void flip(int j, T (&U)[3][3], T (&S)[3]) {
    S[j] = -S[j];
    negc(U, j);
}

// ordering/sign fix-up after a block at t = 0 (rows/cols 0,1)
template <typename T>
__device__ __forceinline__ void order0(T (&U)[3][3], T (&s)[3], T (&V)[3][3]) {
    if (fabs(s[1]) >= fabs(s[2])) {
        if (s[1] < T(0)) { flip(1, U, s); flip(2, U, s); }
        return;
    }
    if (s[2] < T(0)) { flip(1, U, s); flip(2, U, s); }
    swapv(s[1], s[2]); swapc(U, 1, 2); swapc(V, 1, 2);
    if (s[1] > s[0]) { swapv(s[0], s[1]); swapc(U, 0, 1); swapc(V, 0, 1); }
    else { negc(U, 2); negc(V, 2); }
}
// ordering/sign fix-up after a block at t = 1 (rows/cols 1,2)
template <typename T>
__device__ __forceinline__ void order1(T (&U)[3][3], T (&s)[3], T (&V)[3][3]) {
    if (fabs(s[0]) >= s[1]) {
        if (s[0] < T(0)) { flip(0, U, s); flip(2, U, s); }
        return;
    }
    swapv(s[0], s[1]); swapc(U, 0, 1); swapc(V, 0, 1);
    if (fabs(s[1]) < fabs(s[2])) { swapv(s[1], s[2]); swapc(U, 1, 2); swapc(V, 1, 2); }
    else { negc(U, 1); negc(V, 1); }
    if (s[1] < T(0)) { flip(1, U, s); flip(2, U, s); }
}

// A (row-major 3x3) -> U, S, V with A = U diag(S) V^T.  Returns sweeps used.
template <typename T>
__device__ __forceinline__ int svd3(const T (&A)[3][3], T (&U)[3][3], T (&S)[3], T (&V)[3][3]) {
    T B[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) B[r][c] = A[r][c];
    eye3(U);
    eye3(V);
    {
        Rot<T> r = rot_id<T>(1, 2);
        rot_conv(r, B[1][0], B[2][0]);
        rot_rows(r, B);
        rot_cols(r, U);
        chase(B, U, V);
    }
    T a0 = B[0][0], a1 = B[1][1], a2 = B[2][2], b0 = B[0][1], b1 = B[1][2];
    const T nrm = T(0.5) * sqrt(a0 * a0 + a1 * a1 + a2 * a2 + b0 * b0 + b1 * b1);
    const T tol = Eps<T>::v * T(128) * (nrm > T(1) ? nrm : T(1));
    int sweeps = 0;
    while (fabs(a0) > tol && fabs(a1) > tol && fabs(a2) > tol && fabs(b0) > tol && fabs(b1) > tol &&
           sweeps < kMaxSweeps) {
        // Wilkinson shift of the trailing 2x2 of B^T B
        const T ta = a1 * a1 + b0 * b0, tc = a2 * a2 + b1 * b1, tb = a1 * b1;
        const T dd = T(0.5) * (ta - tc);
        const T bs = tb * tb;
        const T mu = tc - copysign(bs / (fabs(dd) + sqrt(dd * dd + bs)), dd);
        Rot<T> r = rot_id<T>(0, 1);
        rot_conv(r, a0 * a0 - mu, a0 * b0);
        rot_cols(r, B);
        rot_cols(r, V);
        chase(B, U, V);
        a0 = B[0][0]; a1 = B[1][1]; a2 = B[2][2]; b0 = B[0][1]; b1 = B[1][2];
        ++sweeps;
    }
    S[0] = S[1] = S[2] = T(0) / T(0);
    if (fabs(b1) <= tol) {
        block2<T>(0, B, U, S, V);
        order0(U, S, V);
    } else if (fabs(b0) <= tol) {
        block2<T>(1, B, U, S, V);
        order1(U, S, V);
    } else if (fabs(a1) <= tol) {
        Rot<T> r1 = rot_id<T>(1, 2);
        rot_unconv(r1, B[1][2], B[2][2]);
        rot_rows(r1, B);
        rot_cols(r1, U);
        block2<T>(0, B, U, S, V);
        order0(U, S, V);
    } else if (fabs(a2) <= tol) {
        Rot<T> r1 = rot_id<T>(1, 2);
        rot_conv(r1, B[1][1], B[1][2]);
        rot_cols(r1, B);
        rot_cols(r1, V);
        Rot<T> r2 = rot_id<T>(0, 2);
        rot_conv(r2, B[0][0], B[0][2]);
        rot_cols(r2, B);
        rot_cols(r2, V);
        block2<T>(0, B, U, S, V);
        order0(U, S, V);
    } else if (fabs(a0) <= tol) {
        Rot<T> r1 = rot_id<T>(0, 1);
        rot_unconv(r1, B[0][1], B[1][1]);
        rot_rows(r1, B);
        rot_cols(r1, U);
        Rot<T> r2 = rot_id<T>(0, 2);
        rot_unconv(r2, B[0][2], B[2][2]);
        rot_rows(r2, B);
        rot_cols(r2, U);
        block2<T>(1, B, U, S, V);
        order1(U, S, V);
    }
    return sweeps;
}

}  // namespace mpmx
