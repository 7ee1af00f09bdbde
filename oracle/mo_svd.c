/*
 * mo_svd.c — oracle (test infrastructure): FP64 restatement of the reference's
 * 3x3 QR/Givens SVD.
 *
 *   givens.hpp:12-131   GivensRotation (conventional / unconventional, row and
 *                       column rotations, composition)
 *   givens.hpp:145-185  zero_chasing
 *   givens.hpp:194-210  upper_bidiagonalize
 *   QRSVD.hpp:27-45     polar_decomposition (2x2)
 *   QRSVD.hpp:49-100    qr_svd2x2
 *   QRSVD.hpp:104-110   wilkinson_shift
 *   QRSVD.hpp:114-196   flip_sign, sort_sigma<0>, sort_sigma<1>
 *   QRSVD.hpp:200-221   process<t>
 *   QRSVD.hpp:226-322   qr_svd (implicit-shift QR sweeps and deflation cases)
 *
 * The reference's Eigen matrices are replaced by plain arrays; the reference's
 * abs() on device resolves to the floating-point absolute value.  Result:
 * F = U diag(S) V^T, S0 >= S1 >= |S2|, only S2 may be negative.
 */
#include "mpmx_oracle.h"
#include <float.h>
#include <math.h>

typedef struct { int i, k; double c, s; } giv;

static giv giv_id(int i, int k) { giv g = {i, k, 1.0, 0.0}; return g; }

/* givens.hpp:46-57 */
static void giv_conventional(giv* g, double a, double b) {
    const double d = a * a + b * b;
    g->c = 1.0; g->s = 0.0;
    if (d != 0.0) { const double t = 1.0 / sqrt(d); g->c = a * t; g->s = -b * t; }
}
/* givens.hpp:60-71 */
static void giv_unconventional(giv* g, double a, double b) {
    const double d = a * a + b * b;
    g->c = 0.0; g->s = 1.0;
    if (d != 0.0) { const double t = 1.0 / sqrt(d); g->s = a * t; g->c = b * t; }
}
/* givens.hpp:88-100: rows i,k of A (3 columns) */
static void giv_row3(const giv* g, double A[3][3]) {
    for (int j = 0; j < 3; ++j) {
        const double t1 = A[g->i][j], t2 = A[g->k][j];
        A[g->i][j] = g->c * t1 - g->s * t2;
        A[g->k][j] = g->s * t1 + g->c * t2;
    }
}
/* givens.hpp:104-114: columns i,k of A (3 rows) */
static void giv_col3(const giv* g, double A[3][3]) {
    for (int r = 0; r < 3; ++r) {
        const double t1 = A[r][g->i], t2 = A[r][g->k];
        A[r][g->i] = g->c * t1 - g->s * t2;
        A[r][g->k] = g->s * t1 + g->c * t2;
    }
}
static void giv_row2(const giv* g, double A[2][2]) {
    for (int j = 0; j < 2; ++j) {
        const double t1 = A[g->i][j], t2 = A[g->k][j];
        A[g->i][j] = g->c * t1 - g->s * t2;
        A[g->k][j] = g->s * t1 + g->c * t2;
    }
}
/* givens.hpp:117-123 */
static void giv_compose(giv* g, const giv* r) {
    const double nc = g->c * r->c - g->s * r->s;
    const double ns = g->s * r->c + g->c * r->s;
    g->c = nc; g->s = ns;
}

static void set_identity(double A[3][3]) {
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) A[r][c] = (r == c) ? 1.0 : 0.0;
}

/* givens.hpp:145-185 */
static void zero_chasing(double H[3][3], double U[3][3], double V[3][3]) {
    giv r1 = giv_id(0, 1);
    giv_conventional(&r1, H[0][0], H[1][0]);
    giv r2 = giv_id(1, 2);
    if (H[1][0] != 0.0)
        giv_conventional(&r2, H[0][0] * H[0][1] + H[1][0] * H[1][1],
                         H[0][0] * H[0][2] + H[1][0] * H[1][2]);
    else
        giv_conventional(&r2, H[0][1], H[0][2]);
    giv_row3(&r1, H);
    giv_col3(&r2, H);
    giv_col3(&r2, V);
    giv r3 = giv_id(1, 2);
    giv_conventional(&r3, H[1][1], H[2][1]);
    giv_row3(&r3, H);
    giv_col3(&r1, U);
    giv_col3(&r3, U);
}

/* givens.hpp:194-210 */
static void upper_bidiagonalize(double H[3][3], double U[3][3], double V[3][3]) {
    set_identity(U);
    set_identity(V);
    giv r = giv_id(1, 2);
    giv_conventional(&r, H[1][0], H[2][0]);
    giv_row3(&r, H);
    giv_col3(&r, U);
    zero_chasing(H, U, V);
}

/* QRSVD.hpp:27-45 */
static void polar_decomposition(const double A[2][2], double S[2][2], giv* R) {
    S[0][0] = A[0][0]; S[0][1] = A[0][1]; S[1][0] = A[1][0]; S[1][1] = A[1][1];
    const double x0 = A[0][0] + A[1][1], x1 = A[1][0] - A[0][1];
    const double d = sqrt(x0 * x0 + x1 * x1);
    if (d != 0.0) { R->c = x0 / d; R->s = -x1 / d; }
    else          { R->c = 1.0;    R->s = 0.0; }
    giv_row2(R, S);
}

/* QRSVD.hpp:49-100 */
static void qr_svd2x2(const double A[2][2], double S[2], giv* U, giv* V) {
    double Ss[2][2];
    polar_decomposition(A, Ss, U);
    double cosine, sine;
    const double x = Ss[0][0], y = Ss[0][1], z = Ss[1][1];
    const double y2 = y * y;
    if (y2 == 0.0) {
        cosine = 1.0; sine = 0.0;
        S[0] = x; S[1] = z;
    } else {
        const double tau = 0.5 * (x - z);
        const double w = sqrt(tau * tau + y2);
        double t;
        if (tau > 0.0) t = y / (tau + w);
        else           t = y / (tau - w);
        cosine = 1.0 / sqrt(t * t + 1.0);
        sine = -t * cosine;
        const double c2 = cosine * cosine;
        const double csy2 = 2.0 * cosine * sine * y;
        const double s2 = sine * sine;
        S[0] = c2 * x - csy2 + s2 * z;
        S[1] = s2 * x + csy2 + c2 * z;
    }
    if (S[0] < S[1]) {
        const double t = S[0]; S[0] = S[1]; S[1] = t;
        V->c = -sine; V->s = cosine;
    } else {
        V->c = cosine; V->s = sine;
    }
    giv_compose(U, V);
}

/* QRSVD.hpp:104-110 */
static double wilkinson_shift(double a1, double b1, double a2) {
    const double d = 0.5 * (a1 - a2);
    const double bs = b1 * b1;
    return a2 - copysign(bs / (fabs(d) + sqrt(d * d + bs)), d);
}

static void flip_sign(int j, double U[3][3], double S[3]) {
    S[j] = -S[j];
    for (int r = 0; r < 3; ++r) U[r][j] = -U[r][j];
}
static void swap_d(double* a, double* b) { const double t = *a; *a = *b; *b = t; }
static void swap_cols(double A[3][3], int a, int b) {
    for (int r = 0; r < 3; ++r) swap_d(&A[r][a], &A[r][b]);
}
static void neg_col(double A[3][3], int a) {
    for (int r = 0; r < 3; ++r) A[r][a] = -A[r][a];
}

/* QRSVD.hpp:127-160 */
static void sort_sigma0(double U[3][3], double s[3], double V[3][3]) {
    if (fabs(s[1]) >= fabs(s[2])) {
        if (s[1] < 0.0) { flip_sign(1, U, s); flip_sign(2, U, s); }
        return;
    }
    if (s[2] < 0.0) { flip_sign(1, U, s); flip_sign(2, U, s); }
    swap_d(&s[1], &s[2]);
    swap_cols(U, 1, 2);
    swap_cols(V, 1, 2);
    if (s[1] > s[0]) {
        swap_d(&s[0], &s[1]);
        swap_cols(U, 0, 1);
        swap_cols(V, 0, 1);
    } else {
        neg_col(U, 2);
        neg_col(V, 2);
    }
}
/* QRSVD.hpp:162-195 */
static void sort_sigma1(double U[3][3], double s[3], double V[3][3]) {
    if (fabs(s[0]) >= s[1]) {
        if (s[0] < 0.0) { flip_sign(0, U, s); flip_sign(2, U, s); }
        return;
    }
    swap_d(&s[0], &s[1]);
    swap_cols(U, 0, 1);
    swap_cols(V, 0, 1);
    if (fabs(s[1]) < fabs(s[2])) {
        swap_d(&s[1], &s[2]);
        swap_cols(U, 1, 2);
        swap_cols(V, 1, 2);
    } else {
        neg_col(U, 1);
        neg_col(V, 1);
    }
    if (s[1] < 0.0) { flip_sign(1, U, s); flip_sign(2, U, s); }
}

/* QRSVD.hpp:200-221 */
static void process(int t, double B[3][3], double U[3][3], double S[3], double V[3][3]) {
    giv u = giv_id(0, 1), v = giv_id(0, 1);
    const int other = (t == 1) ? 0 : 2;
    S[other] = B[other][other];
    double B2[2][2] = {{B[t][t], B[t][t + 1]}, {B[t + 1][t], B[t + 1][t + 1]}};
    double S2[2];
    qr_svd2x2(B2, S2, &u, &v);
    S[t] = S2[0];
    S[t + 1] = S2[1];
    u.i += t; u.k += t;
    v.i += t; v.k += t;
    giv_col3(&u, U);
    giv_col3(&v, V);
}

static int g_last_sweeps = 0;
int mo_svd3_last_sweeps(void) { return g_last_sweeps; }

/* QRSVD.hpp:226-322 */
static void qr_svd(const double A[3][3], double S[3], double U[3][3], double V[3][3]) {
    set_identity(U);
    set_identity(V);
    double B[3][3];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) B[r][c] = A[r][c];
    upper_bidiagonalize(B, U, V);

    giv r = giv_id(0, 1);
    double alpha[3] = {B[0][0], B[1][1], B[2][2]};
    double beta[2] = {B[0][1], B[1][2]};
    double gamma[2] = {alpha[0] * beta[0], alpha[1] * beta[1]};
    const double eta = DBL_EPSILON * 128.0;
    const double nrm = 0.5 * sqrt(alpha[0] * alpha[0] + alpha[1] * alpha[1] + alpha[2] * alpha[2] +
                                  beta[0] * beta[0] + beta[1] * beta[1]);
    const double tol = eta * (nrm > 1.0 ? nrm : 1.0);

    int sweeps = 0;
    while (fabs(alpha[0]) > tol && fabs(alpha[1]) > tol && fabs(alpha[2]) > tol &&
           fabs(beta[0]) > tol && fabs(beta[1]) > tol) {
        const double mu = wilkinson_shift(alpha[1] * alpha[1] + beta[0] * beta[0], gamma[1],
                                          alpha[2] * alpha[2] + beta[1] * beta[1]);
        giv_conventional(&r, alpha[0] * alpha[0] - mu, gamma[0]);
        giv_col3(&r, B);
        giv_col3(&r, V);
        zero_chasing(B, U, V);
        alpha[0] = B[0][0]; alpha[1] = B[1][1]; alpha[2] = B[2][2];
        beta[0] = B[0][1]; beta[1] = B[1][2];
        gamma[0] = alpha[0] * beta[0];
        gamma[1] = alpha[1] * beta[1];
        ++sweeps;
    }
    g_last_sweeps = sweeps;

    if (fabs(beta[1]) <= tol) {
        process(0, B, U, S, V);
        sort_sigma0(U, S, V);
    } else if (fabs(beta[0]) <= tol) {
        process(1, B, U, S, V);
        sort_sigma1(U, S, V);
    } else if (fabs(alpha[1]) <= tol) {
        giv r1 = giv_id(1, 2);
        giv_unconventional(&r1, B[1][2], B[2][2]);
        giv_row3(&r1, B);
        giv_col3(&r1, U);
        process(0, B, U, S, V);
        sort_sigma0(U, S, V);
    } else if (fabs(alpha[2]) <= tol) {
        giv r1 = giv_id(1, 2);
        giv_conventional(&r1, B[1][1], B[1][2]);
        giv_col3(&r1, B);
        giv_col3(&r1, V);
        giv r2 = giv_id(0, 2);
        giv_conventional(&r2, B[0][0], B[0][2]);
        giv_col3(&r2, B);
        giv_col3(&r2, V);
        process(0, B, U, S, V);
        sort_sigma0(U, S, V);
    } else if (fabs(alpha[0]) <= tol) {
        giv r1 = giv_id(0, 1);
        giv_unconventional(&r1, B[0][1], B[1][1]);
        giv_row3(&r1, B);
        giv_col3(&r1, U);
        giv r2 = giv_id(0, 2);
        giv_unconventional(&r2, B[0][2], B[2][2]);
        giv_row3(&r2, B);
        giv_col3(&r2, U);
        process(1, B, U, S, V);
        sort_sigma1(U, S, V);
    }
}

void mo_svd3(const double F[9], double U[9], double S[3], double V[9]) {
    double A[3][3], Um[3][3], Vm[3][3];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) A[r][c] = F[r + 3 * c];
    S[0] = S[1] = S[2] = NAN;   /* a non-finite input leaves every branch untaken */
    qr_svd(A, S, Um, Vm);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) { U[r + 3 * c] = Um[r][c]; V[r + 3 * c] = Vm[r][c]; }
}
