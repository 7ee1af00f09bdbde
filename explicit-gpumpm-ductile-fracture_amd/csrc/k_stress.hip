// k_stress.hip — per-particle von Mises return mapping with hardening,
// psi+ history, phase-field coefficient and Kirchhoff stress.
//
// Same maths as applyVonMises (ExplicitTimeIntegrator.cu:132-370), including the
// reference's choices: the yield test uses the full norm of the trial Kirchhoff
// stress, dlambda = y / ((2/3) g tr tau) capped at 100, alpha is stored clamped
// to alpha_c while the unclamped value feeds G_c, Hencky strain components
// above 3 are rescaled by 3/|eps|, and psi+ takes the deviatoric energy only
// under compression (J < 1).  C^-1 (the reference's in-kernel Gauss-Jordan
// Inverse, :71-129) is the same for every particle and comes precomputed in
// DevParams.
//
// One lane per particle; F, sigma component-major (coalesced across lanes).
#include "payload.h"
#include "svd3.h"

namespace mpmx {

struct StressIn {
    float F[3][3];
    float c, g, alpha, psi;
};
struct StressOut {
    float F[3][3];
    float g, alpha, psi, FP, vol;
    float sigma[3][3];
};

template <typename T>
__device__ __forceinline__ void usv(const T (&U)[3][3], const T (&d)[3], const T (&V)[3][3], T (&P)[3][3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
            P[r][cc] = d[0] * U[r][0] * V[cc][0] + d[1] * U[r][1] * V[cc][1] + d[2] * U[r][2] * V[cc][2];
}

// the return mapping in precision T (float, or double for the near-degenerate fallback)
template <typename T>
__device__ __forceinline__ void von_mises(const DevParams& dp, const StressIn& in, StressOut& out,
                                          const T (&U)[3][3], T (&S)[3], const T (&V)[3][3]) {
    const T lambda = dp.lambda, mu = dp.mu, kappa = dp.kappa;
    T g = in.g;
    T alpha = in.alpha;
    T Fm[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) Fm[r][cc] = in.F[r][cc];

    const T c0 = T(2) * mu + lambda;
    T eps[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) eps[a] = (S[a] > T(0)) ? log(S[a]) : T(0);
    // ln S of the final singular values: eps, or eps_nn after a return
    T lnS[3] = {eps[0], eps[1], eps[2]};
    T tau_tr[3];
    tau_tr[0] = c0 * eps[0] + lambda * eps[1] + lambda * eps[2];
    tau_tr[1] = lambda * eps[0] + c0 * eps[1] + lambda * eps[2];
    tau_tr[2] = lambda * eps[0] + lambda * eps[1] + c0 * eps[2];
    const T tau_tr_F = sqrt(tau_tr[0] * tau_tr[0] + tau_tr[1] * tau_tr[1] + tau_tr[2] * tau_tr[2]);
    const T tau_tr_sum = tau_tr[0] + tau_tr[1] + tau_tr[2];
    const T divider = tau_tr_sum / T(3);
    T tau_dev_tr[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) tau_dev_tr[a] = g * (tau_tr[a] - divider);

    const T p23 = T(2.0f / 3.0f);
    const T sq23 = sqrt(p23);
    const T Rh = T(dp.sq23f) * T(dp.yield) - (T(dp.sq23f) * T(dp.yield) - T(dp.yield)) * exp(-T(dp.hard) * alpha);
    const T y = tau_tr_F - g * Rh * sq23;
    if (y >= T(1e-12)) {
        T dlambda = T(0);
        const T bss = p23 * g * tau_tr_sum;
        if (bss > T(0)) dlambda = y / bss;
        if (dlambda > T(dp.dlambda_cap)) dlambda = T(dp.dlambda_cap);
        if (alpha < T(dp.alpha_c)) {
            alpha += sq23 * dlambda;
            out.alpha = (float)alpha;
        }
        if (alpha > T(dp.alpha_c)) out.alpha = dp.alpha_c;
        const T tdn = sqrt(tau_dev_tr[0] * tau_dev_tr[0] + tau_dev_tr[1] * tau_dev_tr[1] + tau_dev_tr[2] * tau_dev_tr[2]);
        // no deviatoric trial stress (equal singular values, which FP32
        // rounding produces near F = s R): no flow direction, no flow; the
        // reference's (1/0) * 0 would give NaN
        const T taudiv = tdn > T(0) ? (T(1) / tdn) * p23 * g * tau_tr_sum * dlambda : T(0);
        const T third = (T(1) / T(3)) * tau_tr_sum;
        T tau_nn[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) tau_nn[a] = (tau_dev_tr[a] - tau_dev_tr[a] * taudiv) + third;
        T eps_nn[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            eps_nn[a] = T(dp.Cinv[3 * a]) * tau_nn[0] + T(dp.Cinv[3 * a + 1]) * tau_nn[1] +
                        T(dp.Cinv[3 * a + 2]) * tau_nn[2];
        const T normS = sqrt(eps_nn[0] * eps_nn[0] + eps_nn[1] * eps_nn[1] + eps_nn[2] * eps_nn[2]);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (eps_nn[a] > T(dp.eps_cap)) eps_nn[a] = eps_nn[a] * (T(dp.eps_cap) / normS);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            S[a] = exp(eps_nn[a]);
            lnS[a] = eps_nn[a];
        }
        usv(U, S, V, Fm);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) out.F[r][cc] = (float)Fm[r][cc];

    const T J = S[0] * S[1] * S[2];
    // J^(-2/3) and ln J from the logarithms already at hand where every
    // singular value is positive (one exp instead of pow and log)
    const bool pos = S[0] > T(0) && S[1] > T(0) && S[2] > T(0);
    const T lnJ = lnS[0] + lnS[1] + lnS[2];
    const T J_bar = pos ? exp(-p23 * lnJ) : pow(J, -p23);
    T be[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) be[a] = J_bar * S[a] * S[a];
    const T be_sum = be[0] + be[1] + be[2];
    const T ln_J = pos ? lnJ : ((J > T(0)) ? log(J) : T(0));
    const T Psi_vol = (kappa * T(0.5)) * (((J * J - T(1)) * T(0.5)) - ln_J);
    const T Psi_dev = (mu * T(0.5)) * (be_sum - T(3));
    const T Psi_pos = (J < T(1)) ? Psi_dev : Psi_dev + Psi_vol;
    const T smax = fmax(S[0], fmax(S[1], S[2]));
    T psi = in.psi;
    if (smax >= T(dp.crit_stretch) && Psi_pos > psi && Psi_pos < T(1e6)) psi = Psi_pos;
    out.psi = (float)psi;

    const T pp = alpha / T(2);
    const T G_C = (pp < T(dp.ps)) ? T(dp.Gc0)
                                  : T(dp.Gc0) * (((T(1) - T(dp.residual_c)) * exp(T(dp.ps) - pp)) + T(dp.residual_c));
    const T theta = dp.theta;
    out.FP = (float)(T(4) * T(dp.L0) * (T(1) - theta) * psi * (T(1) / G_C) + T(1));

    const T tau_vol = (kappa * T(0.5)) * (J * J - T(1));
    const T bthird = (T(1) / T(3)) * be_sum;
    g = T(in.c) * T(in.c) * (T(1) - theta) + theta;
    out.g = (float)g;
    T P_hat[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T td = mu * (be[a] - bthird);
        P_hat[a] = (J < T(1)) ? g * td + tau_vol : g * (td + tau_vol);
    }
    T P[3][3];
    usv(U, P_hat, V, P);
    // sigma = V0 P F^T
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
            out.sigma[r][cc] = (float)((P[r][0] * Fm[cc][0] + P[r][1] * Fm[cc][1] + P[r][2] * Fm[cc][2]) * T(dp.V0));
    const T det = Fm[0][0] * (Fm[1][1] * Fm[2][2] - Fm[1][2] * Fm[2][1]) -
                  Fm[0][1] * (Fm[1][0] * Fm[2][2] - Fm[1][2] * Fm[2][0]) +
                  Fm[0][2] * (Fm[1][0] * Fm[2][1] - Fm[1][1] * Fm[2][0]);
    out.vol = (float)(T(dp.V0) * det);
}

// sigma = V0 U diag(P_hat) V^T F^T, F = in.F
template <typename T>
__device__ __forceinline__ void contribution(const DevParams& dp, const StressIn& in, const T (&U)[3][3],
                                             const T (&P_hat)[3], const T (&V)[3][3], StressOut& out) {
    T P[3][3];
    usv(U, P_hat, V, P);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
            out.sigma[r][cc] = (float)((P[r][0] * T(in.F[cc][0]) + P[r][1] * T(in.F[cc][1]) + P[r][2] * T(in.F[cc][2])) *
                                       T(dp.V0));
}

template <typename T>
__device__ __forceinline__ T det3(const float (&F)[3][3]) {
    return T(F[0][0]) * (T(F[1][1]) * T(F[2][2]) - T(F[1][2]) * T(F[2][1])) -
           T(F[0][1]) * (T(F[1][0]) * T(F[2][2]) - T(F[1][2]) * T(F[2][0])) +
           T(F[0][2]) * (T(F[1][0]) * T(F[2][1]) - T(F[1][1]) * T(F[2][0]));
}

// computeContributionNeoHookean (ExplicitTimeIntegrator.cu:372-446): phase-
// field fracture without plasticity.  F, g and alpha are left as they are; as
// in the reference the psi+ history takes kappa/2 ((J - 1)/2 - ln J) while the
// stress uses kappa/2 (J^2 - 1), FP takes G_c = 3 and (1 - 0.001) literally,
// and ln J is unguarded.
template <typename T>
__device__ __forceinline__ void neo_hookean(const DevParams& dp, const StressIn& in, StressOut& out,
                                            const T (&U)[3][3], const T (&S)[3], const T (&V)[3][3]) {
    const T mu = dp.mu, kappa = dp.kappa;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) out.F[r][cc] = in.F[r][cc];
    out.g = in.g;
    out.alpha = in.alpha;
    const T J = S[0] * S[1] * S[2];
    const T J_bar = pow(J, -T(2.0f / 3.0f));
    T be[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) be[a] = J_bar * S[a] * S[a];
    const T be_sum = be[0] + be[1] + be[2];
    const T ln_J = log(J);
    const T Psi_vol = kappa / T(2) * ((J - T(1)) / T(2) - ln_J);
    const T Psi_dev = mu / T(2) * (be_sum - T(3));
    const T Psi_pos = (J < T(1)) ? Psi_dev : Psi_dev + Psi_vol;
    const T smax = fmax(S[0], fmax(S[1], S[2]));
    T psi = in.psi;
    if (smax >= T(dp.crit_stretch) && Psi_pos > psi && Psi_pos < T(1e6)) psi = Psi_pos;
    out.psi = (float)psi;
    out.FP = (float)(T(4) * T(dp.L0) * (T(1) - T(0.001f)) * psi * (T(1) / T(3)) + T(1));
    const T tau_vol = kappa / T(2) * (J * J - T(1));
    const T third = (T(1) / T(3)) * be_sum;
    const T theta = dp.theta;
    const T g = T(in.c) * T(in.c) * (T(1) - theta) + theta;
    T P_hat[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T td = mu * (be[a] - third);
        P_hat[a] = (J < T(1)) ? g * td + tau_vol : g * (td + tau_vol);
    }
    contribution<T>(dp, in, U, P_hat, V, out);
    out.vol = (float)(T(dp.V0) * det3<T>(in.F));
}

// computeContributionFixedCorotated (ExplicitTimeIntegrator.cu:22-47): no
// plasticity and no fracture; only sigma changes (FP = 1 and vol = V0 det F
// for the record; the solve does not run)
template <typename T>
__device__ __forceinline__ void fixed_corotated(const DevParams& dp, const StressIn& in, StressOut& out,
                                                const T (&U)[3][3], const T (&S)[3], const T (&V)[3][3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) out.F[r][cc] = in.F[r][cc];
    out.g = in.g;
    out.alpha = in.alpha;
    out.psi = in.psi;
    const T mu = dp.mu, lambda = dp.lambda;
    const T J = S[0] * S[1] * S[2];
    const T scaled_mu = T(2) * mu, scaled_lambda = lambda * (J - T(1));
    const T P_hat[3] = {scaled_mu * (S[0] - T(1)) + scaled_lambda * (S[1] * S[2]),
                        scaled_mu * (S[1] - T(1)) + scaled_lambda * (S[0] * S[2]),
                        scaled_mu * (S[2] - T(1)) + scaled_lambda * (S[0] * S[1])};
    contribution<T>(dp, in, U, P_hat, V, out);
    out.FP = 1.0f;
    out.vol = (float)(T(dp.V0) * det3<T>(in.F));
}

// the energy of the scene: von Mises (2), Neo-Hookean (1), fixed corotated (0)
template <typename T>
__device__ __forceinline__ void energy(const DevParams& dp, const StressIn& in, StressOut& out,
                                       const T (&U)[3][3], T (&S)[3], const T (&V)[3][3]) {
    if (dp.model == 2) von_mises<T>(dp, in, out, U, S, V);
    else if (dp.model == 1) neo_hookean<T>(dp, in, out, U, S, V);
    else fixed_corotated<T>(dp, in, out, U, S, V);
}

// ALL64: SVD and return mapping in FP64 for every particle.  Otherwise FP32,
// redone in FP64 where the smallest singular-value gap is below svd_gap (the
// directions of near-equal singular values are ill-conditioned in FP32).
// Returns whether the FP64 redo ran.
template <bool ALL64>
__device__ __forceinline__ bool stress_one(const DevParams& dp, const StressIn& in, StressOut& out) {
    out.alpha = in.alpha;
    if (ALL64) {
        double Ad[3][3], Ud[3][3], Sd[3], Vd[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) Ad[r][cc] = in.F[r][cc];
        svd3<double>(Ad, Ud, Sd, Vd);
        energy<double>(dp, in, out, Ud, Sd, Vd);
        return false;
    }
    float A[3][3], U[3][3], S[3], V[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) A[r][cc] = in.F[r][cc];
    svd3<float>(A, U, S, V);
    const float gap = fminf(fabsf(S[0] - S[1]), fabsf(S[1] - S[2]));
    if (gap < dp.svd_gap) {
        double Ad[3][3], Ud[3][3], Sd[3], Vd[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) Ad[r][cc] = A[r][cc];
        svd3<double>(Ad, Ud, Sd, Vd);
        energy<double>(dp, in, out, Ud, Sd, Vd);
        return true;
    }
    energy<float>(dp, in, out, U, S, V);
    return false;
}

// one atomic per wave for the FP64-redo counter
__device__ __forceinline__ void count_fallbacks(bool fb, int* counter) {
    const unsigned long long m = __ballot(fb);
    const unsigned long long act = __ballot(1);
    if (m && (int)(threadIdx.x & 63) == __ffsll((long long)act) - 1) atomicAdd(counter, (int)__popcll(m));
}

// state access of the two storage forms
__device__ __forceinline__ void stress_load(const PState& s, int i, int n, StressIn& in) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) in.F[r][cc] = s.F[(int64_t)(r + 3 * cc) * n + i];
    in.c = s.c[i];
    in.g = s.g[i];
    in.alpha = s.alpha[i];
    in.psi = s.psi[i];
}
__device__ __forceinline__ void stress_load(const MState& s, int i, int n, StressIn& in) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
            in.F[r][cc] = s.F32 ? s.F32[(int64_t)(r + 3 * cc) * n + i]
                                : h2f(s.F[(int64_t)(r + 3 * cc) * n + i]) + (r == cc ? 1.0f : 0.0f);
    in.c = bf2f(s.c[i]);
    in.g = bf2f(s.g[i]);
    in.alpha = bf2f(s.alpha[i]);
    in.psi = bf2f(s.psi[i]);
}
__device__ __forceinline__ void stress_store(const PState& s, int i, int n, const StressOut& out) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) s.F[(int64_t)(r + 3 * cc) * n + i] = out.F[r][cc];
    s.g[i] = out.g;
    s.alpha[i] = out.alpha;
    s.psi[i] = out.psi;
}
__device__ __forceinline__ void stress_store(const MState& s, int i, int n, const StressOut& out) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            if (s.F32) s.F32[(int64_t)(r + 3 * cc) * n + i] = out.F[r][cc];
            else s.F[(int64_t)(r + 3 * cc) * n + i] = f2h(out.F[r][cc] - (r == cc ? 1.0f : 0.0f));
        }
    s.g[i] = f2bf(out.g);
    s.alpha[i] = f2bf(out.alpha);
    s.psi[i] = f2bf(out.psi);
}

// stress of particle i from its F and phase-field state (stress_compute);
// then (stress_write) F, g, alpha, psi written back, V stored and, where
// sigma is not null, sigma (component r + 3 c) and FP (the fused stress pass
// carries both in the P2G payload and stores them only when the caller keeps
// them)
template <bool ALL64, typename ST>
__device__ __forceinline__ void stress_compute(const DevParams& sp, const ST& s, int i, int n, DevCounts* dc,
                                               StressOut& out) {
    StressIn in;
    stress_load(s, i, n, in);
    count_fallbacks(stress_one<ALL64>(sp, in, out), &dc->svd_fallbacks);
}
template <typename ST>
__device__ __forceinline__ void stress_write(const ST& s, int i, int n, float* __restrict__ sigma,
                                             float* __restrict__ FP, float* __restrict__ vol, DevCounts* dc,
                                             const StressOut& out) {
    bool finite = true;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            if (sigma) sigma[(int64_t)(r + 3 * cc) * n + i] = out.sigma[r][cc];
            finite = finite && isfinite(out.sigma[r][cc]);
        }
    stress_store(s, i, n, out);
    if (sigma) FP[i] = out.FP;
    vol[i] = out.vol;
    if (!finite) atomicOr(&dc->err, kErrNonFinite);
}

// one thread per particle; every wave holds particles of one scene, whose
// constants it reads once
template <bool ALL64, typename ST>
__global__ __launch_bounds__(256) void k_stress(DevParams dp, ST s, WaveMap wm, float* __restrict__ sigma,
                                                float* __restrict__ FP, float* __restrict__ vol, int n,
                                                DevCounts* dc) {
    int i, scene;
    if (!wave_particle(wm, n, i, scene)) return;
    StressOut out;
    stress_compute<ALL64>(dp.scenes[scene], s, i, n, dc, out);
    stress_write(s, i, n, sigma, FP, vol, dc, out);
}

// the state fields the stress leaves unchanged, moved from o of s to i of d
// (every load before any store); v and B only with vb set: the G2P of the
// substep writes both afresh without reading them, so a substep that runs it
// leaves them out
template <typename ST>
__device__ __forceinline__ void carry_state(const ST& s, int o, const ST& d, int i, int n, bool vb) {
    decltype(+s.x[0]) x[3];
    decltype(+s.v[0]) v[3];
    decltype(+s.B[0]) B[9];
#pragma unroll
    for (int a = 0; a < 3; ++a) x[a] = s.x[3 * o + a];
    if (vb) {
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = s.v[3 * o + a];
#pragma unroll
        for (int a = 0; a < 9; ++a) B[a] = s.B[(int64_t)a * n + o];
    }
    const auto c = s.c[o];
    const auto mass = s.mass[o];
    const int id = s.id[o];
#pragma unroll
    for (int a = 0; a < 3; ++a) d.x[3 * i + a] = x[a];
    if (vb) {
#pragma unroll
        for (int a = 0; a < 3; ++a) d.v[3 * i + a] = v[a];
#pragma unroll
        for (int a = 0; a < 9; ++a) d.B[(int64_t)a * n + i] = B[a];
    }
    d.c[i] = c;
    d.mass[i] = mass;
    d.id[i] = id;
}

// v and B in the new order, for a substep that stopped between its stress
// pass (which left them out) and its G2P
template <typename ST>
__global__ void k_carry_vb(const uint32_t* __restrict__ perm, ST s, ST d, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = (int)perm[i];
    decltype(+s.v[0]) v[3];
    decltype(+s.B[0]) B[9];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = s.v[3 * o + a];
#pragma unroll
    for (int a = 0; a < 9; ++a) B[a] = s.B[(int64_t)a * n + o];
#pragma unroll
    for (int a = 0; a < 3; ++a) d.v[3 * i + a] = v[a];
#pragma unroll
    for (int a = 0; a < 9; ++a) d.B[(int64_t)a * n + i] = B[a];
}

hipError_t finish_carry_vb(mpmx_ctx* c) {
    const uint32_t* perm = c->vb_perm;
    if (!perm) return hipSuccess;
    c->vb_perm = nullptr;
    const int n = (int)c->n, src = 1 - c->cur;
    if (c->cfg.precision != MPMX_FP32)
        k_carry_vb<<<(n + 255) / 256, 256, 0, c->stream>>>(perm, c->ms[src], c->ms[c->cur], n);
    else
        k_carry_vb<<<(n + 255) / 256, 256, 0, c->stream>>>(perm, c->ps[src], c->ps[c->cur], n);
    return hipGetLastError();
}

// the stress pass of a substep with the gather P2G: after the stress, the
// particle's P2G payload record and its page's channel bounds, from the
// stress still in registers (the same values k_p2g_prep reads back from
// sigma, FP and V, so the records are bit-identical); pbound is zero on entry.
// With perm set the pass also performs the sort's state gather: sorted
// particle i reads its state at perm[i] of s and writes all of it to i of d
// (the stand-alone gather's reads are the ones the stress makes anyway);
// without, s and d are the same sorted state.
// Three waves per SIMD (158 VGPRs): 138 us against 161 us at the two waves an
// unconstrained allocation allows, 175 us at four (1M particles).
template <bool ALL64, int TRANSFER, typename ST>
__global__ __launch_bounds__(256, 3) void k_stress_prep(DevParams dp, ST s, ST d, const uint32_t* __restrict__ perm,
                                                     int carry_vb, WaveMap wm, float* __restrict__ sigma,
                                                     float* __restrict__ FP, float* __restrict__ vol, BlockTables t,
                                                     float* __restrict__ pay, unsigned* __restrict__ pbound, float dt,
                                                     int n, const uint8_t* __restrict__ wclass,
                                                     int* __restrict__ order, DevCounts* dc) {
    dt = substep_dt(dt, dc);
    // the first workgroup orders the blocks for the block kernels that follow
    // (k_block_lists has just listed them); dispatched first, it runs beside
    // the stress waves instead of after them
    if ((int)blockIdx.x < kOrderGroups) {
        __shared__ int ocnt[kOrderClasses * 256];
        __shared__ int olds[4];
        int nb = dc->num_blocks;
        nb = nb < t.cap_blocks ? nb : (int)t.cap_blocks;
        order_blocks(wclass, order, nb, ocnt, olds);
        return;
    }
    int i, scene;
    const bool valid = wave_particle(wm, n, (int)blockIdx.x - kOrderGroups, i, scene);
    float b[kP2GCh];
#pragma unroll
    for (int k = 0; k < kP2GCh; ++k) b[k] = 0.f;
    uint32_t pk = 0xffffffffu;
    if (valid) {
        const DevParams& sp = dp.scenes[scene];
        const int o = perm ? (int)perm[i] : i;
        StressOut out;
        stress_compute<ALL64>(sp, s, o, n, dc, out);
        pk = t.keys[i] >> 6;
        float sig[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) sig[r + 3 * cc] = out.sigma[r][cc];
        // the payload reads x, m, c, v and B, none of which the stress
        // changes: it goes first, so that its loads need not wait behind the
        // stores of the updated state
        prep_particle<TRANSFER>(sp, s, sig, out.FP, out.vol, o, i, n, dt, pay, b);
        if (perm) carry_state(s, o, d, i, n, carry_vb != 0);
        stress_write(d, i, n, sigma, FP, vol, dc, out);
    }
    page_bounds(valid, pk, b, t, pbound, dc);
}

template <bool ALL64>
static void stress_dispatch(mpmx_ctx* c) {
    const int n = (int)c->n;
    if (c->cfg.precision != MPMX_FP32)
        k_stress<ALL64, MState><<<particle_blocks(c), c->particle_wg, 0, c->stream>>>(c->dp, c->ms[c->cur], c->wmap, c->sigma,
                                                                           c->FP, c->vol, n, c->dc);
    else
        k_stress<ALL64, PState><<<particle_blocks(c), c->particle_wg, 0, c->stream>>>(c->dp, c->ps[c->cur], c->wmap, c->sigma,
                                                                           c->FP, c->vol, n, c->dc);
}

hipError_t launch_stress(mpmx_ctx* c) {
    c->stress_fields = true;
    if (c->cfg.stress_fp64) stress_dispatch<true>(c);
    else stress_dispatch<false>(c);
    return hipGetLastError();
}

template <bool ALL64, int TRANSFER, typename ST>
static void stress_prep_dispatch(mpmx_ctx* c, ST st, ST dst, float dt) {
    k_stress_prep<ALL64, TRANSFER, ST><<<particle_blocks(c) + kOrderGroups, c->particle_wg, 0, c->stream>>>(
        c->dp, st, dst, c->gather_perm, c->g2p_follows ? 0 : 1, c->wmap, c->cfg.store_stress ? c->sigma : nullptr, c->FP, c->vol,
        block_tables_of(c), c->payload, c->pbound, dt, (int)c->n, c->block_wclass, c->block_order, c->dc);
}

template <bool ALL64, typename ST>
static void stress_prep_dispatch(mpmx_ctx* c, ST st, ST dst, float dt) {
    if (c->P.transfer == 0) stress_prep_dispatch<ALL64, 0>(c, st, dst, dt);
    else stress_prep_dispatch<ALL64, 1>(c, st, dst, dt);
}

// a sort whose gather is pending (sort_keys) hands the state over here: the
// pass reads buffer cur through the permutation and writes the other one
template <bool ALL64>
static void stress_prep_dispatch(mpmx_ctx* c, float dt) {
    const int to = c->gather_perm ? 1 - c->cur : c->cur;
    if (c->cfg.precision != MPMX_FP32) stress_prep_dispatch<ALL64>(c, c->ms[c->cur], c->ms[to], dt);
    else stress_prep_dispatch<ALL64>(c, c->ps[c->cur], c->ps[to], dt);
    // v and B wait for the G2P when the pass gathered without them
    if (c->gather_perm && c->g2p_follows) c->vb_perm = c->gather_perm;
    c->cur = to;
    c->gather_perm = nullptr;
}

hipError_t launch_stress_prep(mpmx_ctx* c, float dt) {
    c->stress_fields = c->cfg.store_stress != 0;
    if (c->cfg.stress_fp64) stress_prep_dispatch<true>(c, dt);
    else stress_prep_dispatch<false>(c, dt);
    return hipGetLastError();
}

}  // namespace mpmx
