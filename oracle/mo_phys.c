/*
 * mo_phys.c — oracle (test infrastructure): FP64 serial restatement of the
 * reference substep.
 *
 *   ExplicitTimeIntegrator.cu:49-62    computePhase_field_vol (V = V0 det F)
 *   ExplicitTimeIntegrator.cu:64-68    R(alpha) hardening
 *   ExplicitTimeIntegrator.cu:71-129   Inverse (Gauss-Jordan)
 *   ExplicitTimeIntegrator.cu:132-370  applyVonMises
 *   ExplicitTimeIntegrator.cu:485-609  max/"dot" reductions (block-max-128)
 *   ExplicitTimeIntegrator.cu:611-698  grid vector kernels, precondition_subd
 *   ExplicitTimeIntegrator.cu:881-995  integrate (substep order, PCG driver loop)
 *   P2GKernal.cu:6-155                 P2G_MLS (per-particle maths; the
 *                                      reference launch covers 256 of its 512 lanes)
 *   P2GKernal.cu:512-702               P2G_APIC
 *   P2GKernal.cu:897-1030              volP2G_APIC (phase-field right-hand side)
 *   P2GKernal.cu:1032-1157             preConditionP2G_APIC (Jacobi diagonal)
 *   P2GKernal.cu:1162-1326             AxP2G_APIC
 *   G2PKernal.cu:32-200                G2P_MLS
 *   G2PKernal.cu:204-441               G2P_APIC
 *   G2PKernal.cu:1006-1123             AxG2P_APIC
 *   TimeIntegrator.cu:63-111           undateGrid_kernal
 *   TimeIntegrator.cu:117-165          collideWithGround (sticky walls)
 *   Materials.cpp:4-14                 lambda, mu, kappa
 *   Particle_kernal.cu:544-573         reorder (re-key and re-sort)
 *
 * Float literals of the reference (2.f/3.f, 0.0001f, 1e-6f, 11.137f, 0.01f) are
 * kept as the float values the FP64 reference build promotes.
 */
#include "mpmx_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CH_M 0
#define CH_F0 4
#define CH_C 7
#define CH_W 8
#define CH_TAG 9
#define CH_RHS 10
#define CH_DIAG 11
#define NCH 12

static const double F23 = (double)(2.0f / 3.0f);

/* Materials.cpp:4-14 */
static void lame(const mo_params* p, double* lambda, double* mu, double* kappa) {
    *lambda = p->E * p->nu / ((1.0 + p->nu) * (1.0 - 2.0 * p->nu));
    *mu = p->E / (2.0 * (1.0 + p->nu));
    *kappa = p->E / (3.0 * (1.0 - 2.0 * p->nu));
}

/* ExplicitTimeIntegrator.cu:71-129 (row-major in/out) */
void mo_inverse3(const double in[9], double mat[9]) {
    for (int i = 0; i < 9; ++i) mat[i] = in[i];
    int swapR[3], swapC[3];
    int pivot[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        int pr = 0, pc = 0;
        double maxValue = 0;
        for (int j = 0; j < 3; ++j)
            if (pivot[j] != 1)
                for (int k = 0; k < 3; ++k)
                    if (pivot[k] == 0 && mat[j * 3 + k] > maxValue) {
                        maxValue = mat[j * 3 + k];
                        pr = j;
                        pc = k;
                    }
        if (pr != pc)
            for (int j = 0; j < 3; ++j) {
                const double pv = mat[pr * 3 + j];
                mat[pr * 3 + j] = mat[pc * 3 + j];
                mat[pc * 3 + j] = pv;
            }
        swapC[i] = pc;
        swapR[i] = pr;
        ++pivot[i];
        const double inv = 1.0 / mat[pc * 3 + pc];
        mat[pc * 3 + pc] = 1;
        for (int j = 0; j < 3; ++j) mat[pc * 3 + j] *= inv;
        for (int j = 0; j < 3; ++j)
            if (j != pc) {
                const double ratio = mat[j * 3 + pc];
                mat[j * 3 + pc] = 0.0;
                for (int k = 0; k < 3; ++k) mat[j * 3 + k] -= mat[pc * 3 + k] * ratio;
            }
    }
    for (int i = 0; i < 3; ++i)
        if (swapR[i] != swapC[i])
            for (int j = 0; j < 3; ++j) {
                const double pv = mat[j * 3 + swapC[i]];
                mat[j * 3 + swapC[i]] = mat[j * 3 + swapR[i]];
                mat[j * 3 + swapR[i]] = pv;
            }
}

/* ExplicitTimeIntegrator.cu:64-68 */
static double hardening_R(double alpha, double Y, double h) {
    const double s23 = (double)sqrtf(2.0f / 3.0f);
    return s23 * Y - (s23 * Y - Y) / exp(h * alpha);
}

/* P = U diag(d) V^T (column-major), ExplicitTimeIntegrator.cu:276-278, 362-365 */
static void usv(const double U[9], const double d[3], const double V[9], double P[9]) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            P[r + 3 * c] = d[0] * U[r] * V[c] + d[1] * U[r + 3] * V[c + 3] + d[2] * U[r + 6] * V[c + 6];
}

/* ExplicitTimeIntegrator.cu:132-370 */
void mo_von_mises(const mo_params* p, double F[9], double phase, double* g_io, double* alpha_io,
                  double* psi_io, double* FP_out, double sigma[9], double* vol_out) {
    double lambda, mu, kappa;
    lame(p, &lambda, &mu, &kappa);
    double g = *g_io;
    double alpha = *alpha_io;

    double U[9], S[3], V[9];
    mo_svd3(F, U, S, V);

    double C[9] = {2 * mu + lambda, lambda, lambda, lambda, 2 * mu + lambda, lambda,
                   lambda, lambda, 2 * mu + lambda};
    double eps[3];
    for (int a = 0; a < 3; ++a) eps[a] = (S[a] > 0.0) ? log(S[a]) : 0.0;
    double tau_tr[3];
    for (int a = 0; a < 3; ++a) tau_tr[a] = C[3 * a] * eps[0] + C[3 * a + 1] * eps[1] + C[3 * a + 2] * eps[2];
    const double tau_tr_F = sqrt(tau_tr[0] * tau_tr[0] + tau_tr[1] * tau_tr[1] + tau_tr[2] * tau_tr[2]);
    const double tau_tr_sum = tau_tr[0] + tau_tr[1] + tau_tr[2];
    const double divider = tau_tr_sum / 3.0;
    double tau_dev_tr[3];
    for (int a = 0; a < 3; ++a) tau_dev_tr[a] = g * (tau_tr[a] - divider);

    double parameter1 = F23;
    const double y = tau_tr_F - g * hardening_R(alpha, p->yield_stress, p->hardening) * sqrt(parameter1);
    if (y >= 1e-12) {
        double dlambda = 0.0;
        const double bss = parameter1 * g * tau_tr_sum;
        if (bss > 0.0) dlambda = y / bss;
        if (dlambda > p->dlambda_cap) dlambda = p->dlambda_cap;
        if (alpha < p->alpha_c) {
            alpha += sqrt(parameter1) * dlambda;
            *alpha_io = alpha;
        }
        if (alpha > p->alpha_c) *alpha_io = p->alpha_c;

        const double tau_dev_tr_norm = sqrt(tau_dev_tr[0] * tau_dev_tr[0] + tau_dev_tr[1] * tau_dev_tr[1] +
                                            tau_dev_tr[2] * tau_dev_tr[2]);
        /* A trial stress with no deviatoric part (equal singular values) has
         * no flow direction: the reference evaluates (1/0) * ... and, with
         * dlambda = 0 under compression, inf * 0 = NaN.  Here it does not flow. */
        const double taudiv = tau_dev_tr_norm > 0.0
                                  ? (1.0 / tau_dev_tr_norm) * parameter1 * g * tau_tr_sum * dlambda
                                  : 0.0;
        double tau_nn[3];
        const double parameter2 = (1.0 / 3.0) * tau_tr_sum;
        for (int a = 0; a < 3; ++a) tau_nn[a] = (tau_dev_tr[a] - tau_dev_tr[a] * taudiv) + parameter2;
        double D[9];
        mo_inverse3(C, D);
        double eps_nn[3];
        for (int a = 0; a < 3; ++a) eps_nn[a] = D[3 * a] * tau_nn[0] + D[3 * a + 1] * tau_nn[1] + D[3 * a + 2] * tau_nn[2];
        const double normS = sqrt(eps_nn[0] * eps_nn[0] + eps_nn[1] * eps_nn[1] + eps_nn[2] * eps_nn[2]);
        for (int a = 0; a < 3; ++a)
            if (eps_nn[a] > p->eps_cap) eps_nn[a] = eps_nn[a] * (p->eps_cap / normS);
        for (int a = 0; a < 3; ++a) S[a] = exp(eps_nn[a]);
        usv(U, S, V, F);
    }

    const double J = S[0] * S[1] * S[2];
    const double J_bar = pow(J, -parameter1);
    double be_bar[3];
    for (int a = 0; a < 3; ++a) be_bar[a] = J_bar * S[a] * S[a];
    const double be_bar_sum = be_bar[0] + be_bar[1] + be_bar[2];
    const double ln_J = (J > 0.0) ? log(J) : 0.0;
    const double Psi_vol = (kappa * 0.5) * (((J * J - 1.0) * 0.5) - ln_J);
    const double Psi_dev = (mu * 0.5) * (be_bar[0] + be_bar[1] + be_bar[2] - 3.0);
    const double Psi_pos = (J < 1.0) ? Psi_dev : Psi_dev + Psi_vol;
    /* crit_stretch gates the history update on the largest principal stretch
     * (0 keeps the reference behaviour: always eligible) */
    const double smax = fmax(S[0], fmax(S[1], S[2]));
    if (smax >= p->crit_stretch && Psi_pos > *psi_io && Psi_pos < 1e6) *psi_io = Psi_pos;

    const double pp = alpha / 2;
    double G_C;
    if (pp < p->ps) G_C = p->Gc0;
    else G_C = p->Gc0 * (((1.0 - p->residual_c) * exp(p->ps - pp)) + p->residual_c);
    const double theta = p->theta;
    const double L0 = (1.0 / p->N) * 0.5;
    *FP_out = 4 * L0 * (1 - theta) * (*psi_io) * (1.0 / G_C) + 1.0;

    parameter1 = (kappa * 0.5) * (J * J - 1.0);
    const double tau_vol = parameter1;
    parameter1 = (1.0 / 3.0) * be_bar_sum;
    double tau_dev[3];
    for (int a = 0; a < 3; ++a) tau_dev[a] = mu * (be_bar[a] - parameter1);

    g = phase * phase * (1 - theta) + theta;
    *g_io = g;
    double P_hat[3];
    for (int a = 0; a < 3; ++a)
        P_hat[a] = (J < 1.0) ? g * tau_dev[a] + tau_vol : g * (tau_dev[a] + tau_vol);
    double P[9];
    usv(U, P_hat, V, P);
    /* sigma = V0 * P * F^T (ExplicitTimeIntegrator.cu:367-369) */
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            sigma[r + 3 * c] = (P[r] * F[c] + P[r + 3] * F[c + 3] + P[r + 6] * F[c + 6]) * p->V0;
    /* ExplicitTimeIntegrator.cu:58-60 */
    const double det = F[0] * F[4] * F[8] + F[3] * F[7] * F[2] + F[6] * F[1] * F[5] -
                       F[2] * F[4] * F[6] - F[1] * F[3] * F[8] - F[0] * F[5] * F[7];
    *vol_out = p->V0 * det;
}

static double det3(const double F[9]) {
    return F[0] * F[4] * F[8] + F[3] * F[7] * F[2] + F[6] * F[1] * F[5] - F[2] * F[4] * F[6] -
           F[1] * F[3] * F[8] - F[0] * F[5] * F[7];
}

/* sigma = V0 * U diag(P_hat) V^T * F^T */
static void contribution(const mo_params* p, const double U[9], const double P_hat[3], const double V[9],
                         const double F[9], double sigma[9]) {
    double P[9];
    usv(U, P_hat, V, P);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            sigma[r + 3 * c] = (P[r] * F[c] + P[r + 3] * F[c + 3] + P[r + 6] * F[c + 6]) * p->V0;
}

/* computeContributionNeoHookean (ExplicitTimeIntegrator.cu:372-446): phase-field
 * fracture without plasticity; F is not changed.  As in the reference the
 * psi+ history takes the volumetric term kappa/2 ((J - 1)/2 - ln J) (:400)
 * while the stress uses kappa/2 (J^2 - 1) (:421-422), FP takes G_c = 3 and
 * (1 - 0.001) as literals (:416-418), ln J is unguarded (:397), and
 * g = c^2 (1 - theta) + theta is used but not stored (:426). */
void mo_neo_hookean(const mo_params* p, const double F[9], double phase, double* psi_io, double* FP_out,
                    double sigma[9], double* vol_out) {
    double lambda, mu, kappa;
    lame(p, &lambda, &mu, &kappa);
    double U[9], S[3], V[9];
    mo_svd3(F, U, S, V);
    const double J = S[0] * S[1] * S[2];
    const double J_bar = pow(J, -F23);
    double be[3];
    for (int a = 0; a < 3; ++a) be[a] = J_bar * S[a] * S[a];
    const double be_sum = be[0] + be[1] + be[2];
    const double ln_J = log(J);
    const double Psi_vol = kappa / 2.0 * ((J - 1.0) / 2.0 - ln_J);
    const double Psi_dev = mu / 2.0 * (be_sum - 3.0);
    const double Psi_pos = (J < 1.0) ? Psi_dev : Psi_dev + Psi_vol;
    const double smax = fmax(S[0], fmax(S[1], S[2]));
    if (smax >= p->crit_stretch && Psi_pos > *psi_io && Psi_pos < 1e6) *psi_io = Psi_pos;
    const double G_C = 3.0;
    const double L0 = 0.5 * (1.0 / p->N);
    *FP_out = 4.0 * L0 * (1.0 - (double)0.001f) * (*psi_io) * (1.0 / G_C) + 1.0;
    const double tau_vol = kappa / 2.0 * (J * J - 1.0);
    const double third = (1.0 / 3.0) * be_sum;
    const double theta = p->theta;
    const double g = phase * phase * (1.0 - theta) + theta;
    double P_hat[3];
    for (int a = 0; a < 3; ++a) {
        const double td = mu * (be[a] - third);
        P_hat[a] = (J < 1.0) ? g * td + tau_vol : g * (td + tau_vol);
    }
    contribution(p, U, P_hat, V, F, sigma);
    *vol_out = p->V0 * det3(F);
}

/* computeContributionFixedCorotated (ExplicitTimeIntegrator.cu:22-47): no
 * plasticity, no fracture */
void mo_fixed_corotated(const mo_params* p, const double F[9], double sigma[9]) {
    double lambda, mu, kappa;
    lame(p, &lambda, &mu, &kappa);
    double U[9], S[3], V[9];
    mo_svd3(F, U, S, V);
    const double J = S[0] * S[1] * S[2];
    const double scaled_mu = 2.0 * mu, scaled_lambda = lambda * (J - 1.0);
    const double P_hat[3] = {scaled_mu * (S[0] - 1.0) + scaled_lambda * (S[1] * S[2]),
                             scaled_mu * (S[1] - 1.0) + scaled_lambda * (S[0] * S[2]),
                             scaled_mu * (S[2] - 1.0) + scaled_lambda * (S[0] * S[1])};
    contribution(p, U, P_hat, V, F, sigma);
}

/* ------------------------------------------------------------------ simulator */
struct mo_sim {
    mo_params p;
    int n;
    double *x, *v, *F, *B, *m, *c, *g, *alpha, *psi, *FP, *vol, *sigma;
    int* order;
    uint64_t* keys;      /* sorted keys */
    uint64_t* tmpkeys;
    int* tmpperm;
    /* tables */
    int cap_p, cap_tp, cap_vp;
    int gcap;            /* pages the grid arrays currently hold */
    uint64_t* page_key;
    int *page2particle, *adj, *vpo, *target, *p2c, *c2p;
    int counts[4];
    int M;               /* page lookup edge */
    int* lookup;         /* dense page coords -> sparse id */
    double* grid;        /* NCH * cap_tp * 64 */
    double* grid_p2g;    /* copy after P2G (stage 0) */
    double* grid_upd;    /* copy after update (stage 1) */
    double *r, *xx, *mr, *s, *q;
    int cg_iters, cg_breakdown;
    double cg_norm;
};

static void* xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

static int page_coord(uint64_t key, int axis) {
    uint64_t m[3];
    mo_masks(m);
    return mo_bit_pack(m[axis], key << 12) >> 2;
}

static int in_domain(const mo_sim* s) {
    const double lo = 2.0 / s->p.N, hi = 1.0 - 2.0 / s->p.N;
    for (int i = 0; i < 3 * s->n; ++i)
        if (!(s->x[i] >= lo && s->x[i] <= hi)) return 0;
    return 1;
}

/* re-key from positions and stable-sort the current order (Particle_kernal.cu:544-573) */
static void resort(mo_sim* s) {
    for (int k = 0; k < s->n; ++k) {
        const int id = s->order[k];
        s->tmpkeys[k] = mo_linear_offset(mo_base_node(s->x[3 * id], s->p.N),
                                         mo_base_node(s->x[3 * id + 1], s->p.N),
                                         mo_base_node(s->x[3 * id + 2], s->p.N));
    }
    mo_sort_keys(s->n, s->tmpkeys, s->tmpperm, s->keys);
    int* neworder = (int*)malloc(sizeof(int) * (size_t)(s->n ? s->n : 1));
    for (int k = 0; k < s->n; ++k) neworder[k] = s->order[s->tmpperm[k]];
    memcpy(s->order, neworder, sizeof(int) * (size_t)s->n);
    free(neworder);
}

mo_sim* mo_sim_create(const mo_params* p, int n, const double* x, const double* v, const double* mass) {
    mo_sim* s = (mo_sim*)calloc(1, sizeof(mo_sim));
    s->p = *p;
    s->n = n;
    s->x = (double*)xcalloc((size_t)n * 3, sizeof(double));
    s->v = (double*)xcalloc((size_t)n * 3, sizeof(double));
    s->F = (double*)xcalloc((size_t)n * 9, sizeof(double));
    s->B = (double*)xcalloc((size_t)n * 9, sizeof(double));
    s->sigma = (double*)xcalloc((size_t)n * 9, sizeof(double));
    s->m = (double*)xcalloc((size_t)n, sizeof(double));
    s->c = (double*)xcalloc((size_t)n, sizeof(double));
    s->g = (double*)xcalloc((size_t)n, sizeof(double));
    s->alpha = (double*)xcalloc((size_t)n, sizeof(double));
    s->psi = (double*)xcalloc((size_t)n, sizeof(double));
    s->FP = (double*)xcalloc((size_t)n, sizeof(double));
    s->vol = (double*)xcalloc((size_t)n, sizeof(double));
    s->order = (int*)xcalloc((size_t)n, sizeof(int));
    s->keys = (uint64_t*)xcalloc((size_t)n, sizeof(uint64_t));
    s->tmpkeys = (uint64_t*)xcalloc((size_t)n, sizeof(uint64_t));
    s->tmpperm = (int*)xcalloc((size_t)n, sizeof(int));
    memcpy(s->x, x, sizeof(double) * 3 * (size_t)n);
    memcpy(s->v, v, sizeof(double) * 3 * (size_t)n);
    memcpy(s->m, mass, sizeof(double) * (size_t)n);
    /* Particles::initialize_kernal (Particle_kernal.cu:441-481), initMatrix (TimeIntegrator.cu:46-51) */
    for (int k = 0; k < n; ++k) {
        for (int a = 0; a < 9; ++a) s->F[9 * k + a] = (a % 3 == a / 3) ? 1.0 : 0.0;
        s->c[k] = 1.0;
        s->g[k] = 1.0;
        s->order[k] = k;
    }
    s->cap_p = n + 1;
    s->cap_vp = n + 1;
    s->cap_tp = 8 * (n + 1);
    s->page_key = (uint64_t*)xcalloc((size_t)s->cap_tp, sizeof(uint64_t));
    s->page2particle = (int*)xcalloc((size_t)s->cap_p + 1, sizeof(int));
    s->adj = (int*)xcalloc((size_t)7 * s->cap_p, sizeof(int));
    s->vpo = (int*)xcalloc((size_t)s->cap_p + 1, sizeof(int));
    s->target = (int*)xcalloc((size_t)s->cap_vp, sizeof(int));
    s->p2c = (int*)xcalloc((size_t)n, sizeof(int));
    s->c2p = (int*)xcalloc((size_t)n + 1, sizeof(int));
    s->M = p->N / 4 + 2;
    s->lookup = (int*)malloc(sizeof(int) * (size_t)s->M * s->M * s->M);
    s->grid = NULL;
    resort(s);
    return s;
}

void mo_sim_destroy(mo_sim* s) {
    if (!s) return;
    free(s->x); free(s->v); free(s->F); free(s->B); free(s->sigma); free(s->m); free(s->c);
    free(s->g); free(s->alpha); free(s->psi); free(s->FP); free(s->vol); free(s->order);
    free(s->keys); free(s->tmpkeys); free(s->tmpperm); free(s->page_key); free(s->page2particle);
    free(s->adj); free(s->vpo); free(s->target); free(s->p2c); free(s->c2p); free(s->lookup);
    free(s->grid); free(s->grid_p2g); free(s->grid_upd);
    free(s->r); free(s->xx); free(s->mr); free(s->s); free(s->q);
    free(s);
}

/* node (i,j,k) -> index page*64 + cell, or -1 */
static int node_index(const mo_sim* s, int i, int j, int k) {
    const int M = s->M;
    const int pi = i >> 2, pj = j >> 2, pk = k >> 2;
    if (pi < 0 || pj < 0 || pk < 0 || pi >= M || pj >= M || pk >= M) return -1;
    const int id = s->lookup[(pi * M + pj) * M + pk];
    if (id < 0) return -1;
    return id * 64 + ((i & 3) << 4 | (j & 3) << 2 | (k & 3));
}

static double* CH(mo_sim* s, int ch) { return s->grid + (size_t)ch * s->gcap * 64; }

/* A node takes part in the grid update and the damage solve when some
 * particle gives it a positive weight.  The reference tests its tag channel
 * (0.0001 per contribution, TimeIntegrator.cu:67-69), which also marks nodes
 * reached only with weight exactly 0 (a particle coordinate exactly on a cell
 * centre) and then divides by their zero mass; sum w > 0 is the same set
 * without those nodes. */
static int ACTIVE(mo_sim* s, int ni) { return CH(s, CH_W)[ni] > 0.0; }

/* quadratic B-spline weights and gradients (P2GKernal.cu:74-82, 237-258) */
static void weights(const double xp[3], double inv_dx, double w[3][3], double wg[3][3]) {
    for (int a = 0; a < 3; ++a) {
        double d0 = xp[a] * inv_dx;
        double z = 1.5 - d0;
        w[a][0] = 0.5 * z * z;
        wg[a][0] = -z;
        d0 = d0 - 1.0;
        w[a][1] = 0.75 - d0 * d0;
        wg[a][1] = -d0 * 2;
        z = 1.5 - (1.0 - d0);
        w[a][2] = 0.5 * z * z;
        wg[a][2] = z;
    }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) wg[a][b] *= inv_dx;
}

static void base_of(const mo_sim* s, int id, int b[3], double xp[3]) {
    const double dx = 1.0 / s->p.N;
    for (int a = 0; a < 3; ++a) {
        b[a] = mo_base_node(s->x[3 * id + a], s->p.N);
        xp[a] = s->x[3 * id + a] - b[a] * dx;
    }
}

static int build_tables(mo_sim* s) {
    const int rc = mo_page_tables(s->n, s->keys, s->p.vpage, s->cap_p, s->cap_tp, s->cap_vp,
                                  s->page_key, s->page2particle, s->adj, s->vpo, s->target,
                                  s->p2c, s->c2p, s->counts);
    if (rc) return rc;
    const int M = s->M;
    for (int i = 0; i < M * M * M; ++i) s->lookup[i] = -1;
    for (int id = 0; id < s->counts[1]; ++id) {
        const int pi = page_coord(s->page_key[id], 0), pj = page_coord(s->page_key[id], 1),
                  pk = page_coord(s->page_key[id], 2);
        if (pi < M && pj < M && pk < M) s->lookup[(pi * M + pj) * M + pk] = id;
    }
    if (s->counts[1] > s->gcap) {
        const int cap = s->counts[1] + s->counts[1] / 2 + 8;
        free(s->grid); free(s->grid_p2g); free(s->grid_upd);
        free(s->r); free(s->xx); free(s->mr); free(s->s); free(s->q);
        s->grid = (double*)malloc(sizeof(double) * NCH * (size_t)cap * 64);
        s->grid_p2g = (double*)malloc(sizeof(double) * NCH * (size_t)cap * 64);
        s->grid_upd = (double*)malloc(sizeof(double) * NCH * (size_t)cap * 64);
        s->r = (double*)malloc(sizeof(double) * (size_t)cap * 64);
        s->xx = (double*)malloc(sizeof(double) * (size_t)cap * 64);
        s->mr = (double*)malloc(sizeof(double) * (size_t)cap * 64);
        s->s = (double*)malloc(sizeof(double) * (size_t)cap * 64);
        s->q = (double*)malloc(sizeof(double) * (size_t)cap * 64);
        s->gcap = cap;
    }
    for (int ch = 0; ch < NCH; ++ch) memset(CH(s, ch), 0, sizeof(double) * (size_t)s->counts[1] * 64);
    return 0;
}

/* P2G_MLS (P2GKernal.cu:52-136) or P2G_APIC (P2GKernal.cu:565-680) for all particles */
static void p2g(mo_sim* s, double dt) {
    const double dx = 1.0 / s->p.N, inv_dx = (double)s->p.N;
    const double Dinv = 4.0 * inv_dx * inv_dx;
    for (int k = 0; k < s->n; ++k) {
        const int id = s->order[k];
        int b[3];
        double xp[3], w[3][3], wg[3][3];
        base_of(s, id, b, xp);
        weights(xp, inv_dx, w, wg);
        const double mass = s->m[id];
        const double* sig = s->sigma + 9 * id;
        const double* v = s->v + 3 * id;
        double Q[9];
        for (int a = 0; a < 9; ++a)
            Q[a] = s->p.transfer == 0 ? (s->B[9 * id + a] * mass - sig[a] * dt) * Dinv
                                      : s->B[9 * id + a] * Dinv;
        const double C = s->c[id];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int kk = 0; kk < 3; ++kk) {
                    const int ni = node_index(s, b[0] + i, b[1] + j, b[2] + kk);
                    const double weight = w[0][i] * w[1][j] * w[2][kk];
                    const double d[3] = {i * dx - xp[0], j * dx - xp[1], kk * dx - xp[2]};
                    const double val0 = mass * weight;
                    CH(s, CH_M)[ni] += val0;
                    if (s->p.transfer == 0) {
                        for (int a = 0; a < 3; ++a)
                            CH(s, 1 + a)[ni] += val0 * v[a] + (Q[a] * d[0] + Q[a + 3] * d[1] + Q[a + 6] * d[2]) * weight;
                        CH(s, CH_TAG)[ni] += (double)0.0001f;
                    } else {
                        const double g3[3] = {wg[0][i] * w[1][j] * w[2][kk], w[0][i] * wg[1][j] * w[2][kk],
                                              w[0][i] * w[1][j] * wg[2][kk]};
                        for (int a = 0; a < 3; ++a) {
                            CH(s, CH_F0 + a)[ni] += -(sig[a] * g3[0] + sig[a + 3] * g3[1] + sig[a + 6] * g3[2]);
                            CH(s, 1 + a)[ni] += (v[a] + (Q[a] * d[0] + Q[a + 3] * d[1] + Q[a + 6] * d[2])) * (mass * weight);
                        }
                        CH(s, CH_TAG)[ni] += (double)0.01f;
                    }
                    CH(s, CH_C)[ni] += weight * C;
                    CH(s, CH_W)[ni] += weight;
                }
    }
}

/* undateGrid_kernal (TimeIntegrator.cu:63-111) */
static void grid_update(mo_sim* s, double dt) {
    const int nn = s->counts[1] * 64;
    for (int ni = 0; ni < nn; ++ni) {
        if (!ACTIVE(s, ni)) continue;
        double mass = CH(s, CH_M)[ni];
        mass = 1.0 / mass;
        CH(s, CH_C)[ni] = CH(s, CH_C)[ni] / CH(s, CH_W)[ni];
        for (int a = 0; a < 3; ++a) CH(s, 1 + a)[ni] *= mass;
        for (int a = 0; a < 3; ++a) CH(s, 1 + a)[ni] += s->p.gravity[a] * dt;
        mass = dt * mass;
        for (int a = 0; a < 3; ++a) {
            const double test = CH(s, CH_F0 + a)[ni] * mass;
            if (!isnan(test) && !isinf(test)) CH(s, 1 + a)[ni] += test;
        }
    }
}

/* one grid node's velocity after the boundaries: the grip slabs first, as in
 * PullGrid_kernal (TimeIntegrator.cu:169-261, here two x-slabs moving at
 * -+grip_v), then the sticky walls of collideWithGround
 * (TimeIntegrator.cu:117-165): a node within wall_cells of a face whose
 * velocity points out through that face stops entirely */
void mo_node_boundary(const mo_params* p, int i, int j, int k, double v[3]) {
    const int N = p->N, W = p->wall_cells;
    if (p->grip_lo >= 0) {
        if (i <= p->grip_lo) { v[0] = -p->grip_v; v[1] = 0; v[2] = 0; }
        if (i >= p->grip_hi) { v[0] = p->grip_v; v[1] = 0; v[2] = 0; }
    }
    if (i <= W && v[0] < 0.0) { v[0] = 0; v[1] = 0; v[2] = 0; }
    if (i > N - W && v[0] > 0.0) { v[0] = 0; v[1] = 0; v[2] = 0; }
    if (j <= W && v[1] < 0.0) { v[1] = 0; v[0] = 0; v[2] = 0; }
    if (j > N - W && v[1] > 0.0) { v[1] = 0; v[0] = 0; v[2] = 0; }
    if (k <= W && v[2] < 0.0) { v[2] = 0; v[0] = 0; v[1] = 0; }
    if (k > N - W && v[2] > 0.0) { v[2] = 0; v[0] = 0; v[1] = 0; }
}

static void boundary(mo_sim* s) {
    for (int id = 0; id < s->counts[1]; ++id) {
        const int pi = page_coord(s->page_key[id], 0) * 4, pj = page_coord(s->page_key[id], 1) * 4,
                  pk = page_coord(s->page_key[id], 2) * 4;
        for (int cell = 0; cell < 64; ++cell) {
            const int i = pi + (cell >> 4), j = pj + ((cell >> 2) & 3), k = pk + (cell & 3);
            const int ni = id * 64 + cell;
            double v[3] = {CH(s, 1)[ni], CH(s, 2)[ni], CH(s, 3)[ni]};
            mo_node_boundary(&s->p, i, j, k, v);
            CH(s, 1)[ni] = v[0];
            CH(s, 2)[ni] = v[1];
            CH(s, 3)[ni] = v[2];
        }
    }
}

/* the MLS grid-to-particle speed clamp (G2PKernal.cu:134-142): a velocity
 * faster than vmax keeps its direction at speed vmax */
void mo_speed_clamp(double vmax, double v[3]) {
    const double norms = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (norms > vmax)
        for (int a = 0; a < 3; ++a) v[a] = v[a] / norms * vmax;
}

/* the monotone damage update of G2P (G2PKernal.cu:155): c only decreases
 * and stays >= 0 */
double mo_damage_update(double c, double dc) { return fmax(fmin(c, c + dc), 0.0); }

/* MPMSimulator::computeDt (Simulator.cpp:200-225) over
 * Particles::getMaxParticleVel_c (Particle_kernal.cu:421-440): dt =
 * min(dt_default, cfl dx / max|v|).  wave = 1 adds each particle's elastic
 * wave speed sqrt((lambda + 2 mu) V0 / m) to its speed (an extension). */
double mo_sim_cfl_dt(const mo_sim* s, double dt_default, double cfl, int wave, double* max_speed) {
    double lambda, mu, kappa;
    lame(&s->p, &lambda, &mu, &kappa);
    double best = 0.0;
    for (int id = 0; id < s->n; ++id) {
        const double* v = s->v + 3 * id;
        double sp = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (wave) sp += sqrt((lambda + 2.0 * mu) * s->p.V0 / s->m[id]);
        if (sp > best) best = sp;
    }
    if (max_speed) *max_speed = best;
    const double dx = 1.0 / s->p.N;
    double dt = dt_default;
    if (best > 0 && cfl * dx / best < dt_default) dt = cfl * dx / best;
    return dt;
}

/* getDotProduct (ExplicitTimeIntegrator.cu:731-747): myDotSum sums each
   128-entry block (:574-610) and mysumFm reduces the block sums with a
   maximum, not a sum (:528-571), so mode 0 is the largest 128-entry block
   sum; mode 1 is the true dot product. */
double mo_vdot(int n, const double* a, const double* b, int mode) {
    if (mode == 1) {
        double acc = 0;
        for (int i = 0; i < n; ++i) acc += a[i] * b[i];
        return acc;
    }
    double best = 0;
    for (int blk = 0; blk * 128 < n; ++blk) {
        double acc = 0;
        const int e = (blk + 1) * 128 < n ? (blk + 1) * 128 : n;
        for (int i = blk * 128; i < e; ++i) acc += a[i] * b[i];
        if (blk == 0 || acc > best) best = acc;
    }
    return best;
}
static double vdot(const mo_sim* s, const double* a, const double* b) {
    return mo_vdot(s->counts[1] * 64, a, b, s->p.dot_mode);
}
static double vnorm(const mo_sim* s, const double* a) {
    const int n = s->counts[1] * 64;
    double best = 0;
    for (int i = 0; i < n; ++i)
        if (a[i] * a[i] > best) best = a[i] * a[i];
    return sqrt(best);
}

/* preConditionP2G_APIC + precondition_subd (P2GKernal.cu:1032-1157, ExplicitTimeIntegrator.cu:673-698) */
static void precondition(mo_sim* s, double dt) {
    const double inv_dx = (double)s->p.N;
    for (int k = 0; k < s->n; ++k) {
        const int id = s->order[k];
        int b[3];
        double xp[3], w[3][3], wg[3][3];
        base_of(s, id, b, xp);
        weights(xp, inv_dx, w, wg);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int kk = 0; kk < 3; ++kk) {
                    const int ni = node_index(s, b[0] + i, b[1] + j, b[2] + kk);
                    CH(s, CH_DIAG)[ni] += s->vol[id] * w[0][i] * w[1][j] * w[2][kk] * (s->FP[id] * s->p.M + (1.0 / dt));
                }
    }
    const int nn = s->counts[1] * 64;
    for (int ni = 0; ni < nn; ++ni) {
        if (!ACTIVE(s, ni)) continue;
        const double mr = CH(s, CH_DIAG)[ni];
        CH(s, CH_DIAG)[ni] = 0.0;
        if (fabs(mr) < 1e-30) continue;
        s->mr[ni] = s->r[ni] / mr;
    }
}

/* AxG2P_APIC + AxP2G_APIC (G2PKernal.cu:1006-1123, P2GKernal.cu:1162-1326): q = A s */
static void apply_A(mo_sim* s, double dt, double* col) {
    const double inv_dx = (double)s->p.N;
    const double L0 = 0.5 * (1.0 / s->p.N);
    for (int k = 0; k < s->n; ++k) {
        const int id = s->order[k];
        int b[3];
        double xp[3], w[3][3], wg[3][3];
        base_of(s, id, b, xp);
        weights(xp, inv_dx, w, wg);
        double val[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int kk = 0; kk < 3; ++kk) {
                    const int ni = node_index(s, b[0] + i, b[1] + j, b[2] + kk);
                    const double gc = s->s[ni];
                    val[0] += wg[0][i] * w[1][j] * w[2][kk] * gc;
                    val[1] += w[0][i] * wg[1][j] * w[2][kk] * gc;
                    val[2] += w[0][i] * w[1][j] * wg[2][kk] * gc;
                }
        const double PARA = s->vol[id] * 4.0 * L0 * L0;
        for (int a = 0; a < 3; ++a) col[3 * id + a] = val[a] * PARA;
    }
    for (int k = 0; k < s->n; ++k) {
        const int id = s->order[k];
        int b[3];
        double xp[3], w[3][3], wg[3][3];
        base_of(s, id, b, xp);
        weights(xp, inv_dx, w, wg);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int kk = 0; kk < 3; ++kk) {
                    const int ni = node_index(s, b[0] + i, b[1] + j, b[2] + kk);
                    const double g3[3] = {wg[0][i] * w[1][j] * w[2][kk], w[0][i] * wg[1][j] * w[2][kk],
                                          w[0][i] * w[1][j] * wg[2][kk]};
                    const double HW = g3[0] * col[3 * id] + g3[1] * col[3 * id + 1] + g3[2] * col[3 * id + 2];
                    const double gx = s->s[ni];
                    s->q[ni] += gx * s->vol[id] * w[0][i] * w[1][j] * w[2][kk] * (s->FP[id] * s->p.M + 1.0 / dt) + HW * s->p.M;
                }
    }
}

/* ExplicitTimeIntegrator.cu:909-968 */
static int phase_field(mo_sim* s, double dt) {
    const double inv_dx = (double)s->p.N;
    const int nn = s->counts[1] * 64;
    /* computePhase_field_vol (:49-62) */
    for (int id = 0; id < s->n; ++id) {
        const double* F = s->F + 9 * id;
        const double det = F[0] * F[4] * F[8] + F[3] * F[7] * F[2] + F[6] * F[1] * F[5] -
                           F[2] * F[4] * F[6] - F[1] * F[3] * F[8] - F[0] * F[5] * F[7];
        s->vol[id] = s->p.V0 * det;
    }
    /* volP2G_APIC (P2GKernal.cu:1001-1022) */
    for (int k = 0; k < s->n; ++k) {
        const int id = s->order[k];
        int b[3];
        double xp[3], w[3][3], wg[3][3];
        base_of(s, id, b, xp);
        weights(xp, inv_dx, w, wg);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int kk = 0; kk < 3; ++kk) {
                    const int ni = node_index(s, b[0] + i, b[1] + j, b[2] + kk);
                    const double gc = CH(s, CH_C)[ni];
                    CH(s, CH_RHS)[ni] += s->vol[id] * w[0][i] * w[1][j] * w[2][kk] * (gc / dt + s->p.M);
                }
    }
    for (int i = 0; i < nn; ++i) { s->r[i] = 0; s->xx[i] = 0; s->mr[i] = 0; s->q[i] = 0; }
    for (int i = 0; i < nn; ++i)
        if (ACTIVE(s, i)) s->r[i] = CH(s, CH_RHS)[i];
    double norm = vnorm(s, s->r);
    precondition(s, dt);
    memcpy(s->s, s->mr, sizeof(double) * (size_t)nn);
    double rho_old = vdot(s, s->r, s->mr);
    double* col = (double*)xcalloc((size_t)s->n * 3, sizeof(double));
    int it = 0, breakdown = 0;
    for (it = 0; it < s->p.cg_max_iter; ++it) {
        for (int i = 0; i < nn; ++i) s->q[i] = 0;
        apply_A(s, dt, col);
        const double sq = vdot(s, s->s, s->q);
        const double alpha = rho_old / sq;
        for (int i = 0; i < nn; ++i)
            if (ACTIVE(s, i)) {
                s->xx[i] += s->s[i] * alpha;
                s->r[i] = s->r[i] - s->q[i] * alpha;
            }
        norm = vnorm(s, s->r);
        if (norm < s->p.cg_tol) { ++it; break; }
        precondition(s, dt);
        const double rho = vdot(s, s->r, s->mr);
        if (rho_old == 0.0) {
            for (int i = 0; i < nn; ++i) s->xx[i] = 1e20;
            breakdown = 1;
            ++it;
            break;
        }
        const double beta = rho / rho_old;
        rho_old = rho;
        for (int i = 0; i < nn; ++i)
            if (ACTIVE(s, i)) s->s[i] = s->s[i] * beta + s->mr[i];
    }
    free(col);
    s->cg_iters = it;
    s->cg_breakdown = breakdown;
    s->cg_norm = norm;
    /* update_grid_phase (:632-642): channel 7 becomes the damage increment */
    for (int i = 0; i < nn; ++i)
        if (ACTIVE(s, i)) CH(s, CH_C)[i] = s->xx[i] - CH(s, CH_C)[i];
    return breakdown ? -3 : 0;
}

/* G2P_MLS (G2PKernal.cu:80-198) or G2P_APIC (G2PKernal.cu:255-440) */
static void g2p(mo_sim* s, double dt) {
    const double dx = 1.0 / s->p.N, inv_dx = (double)s->p.N;
    const double Dinv = 4.0 * inv_dx * inv_dx;
    const double vmax = s->p.vmax;
    for (int k = 0; k < s->n; ++k) {
        const int id = s->order[k];
        int b[3];
        double xp[3], w[3][3], wg[3][3];
        base_of(s, id, b, xp);
        weights(xp, inv_dx, w, wg);
        double val[4] = {0, 0, 0, 0};
        double Bn[9] = {0}, G[9] = {0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int kk = 0; kk < 3; ++kk) {
                    const int ni = node_index(s, b[0] + i, b[1] + j, b[2] + kk);
                    const double weight = w[0][i] * w[1][j] * w[2][kk];
                    const double vi[3] = {CH(s, 1)[ni], CH(s, 2)[ni], CH(s, 3)[ni]};
                    val[0] += weight * vi[0];
                    val[1] += weight * vi[1];
                    val[2] += weight * vi[2];
                    val[3] += weight * CH(s, CH_C)[ni];
                    const double d[3] = {i * dx - xp[0], j * dx - xp[1], kk * dx - xp[2]};
                    const double g3[3] = {wg[0][i] * w[1][j] * w[2][kk], w[0][i] * wg[1][j] * w[2][kk],
                                          w[0][i] * w[1][j] * wg[2][kk]};
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) {
                            Bn[r + 3 * c] += weight * vi[r] * d[c];
                            G[r + 3 * c] += vi[r] * g3[c];
                        }
                }
        if (s->p.transfer == 0) mo_speed_clamp(vmax, val);
        for (int a = 0; a < 3; ++a) s->v[3 * id + a] = val[a];
        s->c[id] = mo_damage_update(s->c[id], val[3]);
        for (int a = 0; a < 3; ++a) s->x[3 * id + a] += val[a] * dt;
        double A[9];
        for (int a = 0; a < 9; ++a)
            A[a] = (s->p.transfer == 0 ? Bn[a] * dt * Dinv : G[a] * dt);
        A[0] += 1.0; A[4] += 1.0; A[8] += 1.0;
        double* F = s->F + 9 * id;
        double R[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                R[r + 3 * c] = A[r] * F[3 * c] + A[r + 3] * F[3 * c + 1] + A[r + 6] * F[3 * c + 2];
        memcpy(F, R, sizeof(R));
        memcpy(s->B + 9 * id, Bn, sizeof(Bn));
    }
}

int mo_sim_substep(mo_sim* s, double dt) {
    if (!in_domain(s)) return -2;
    /* DomainTransformer::rebuild (Transformer.cpp:71-82) */
    if (build_tables(s)) return -1;
    /* computeForceCoefficient (ExplicitTimeIntegrator.cu:845-865), by Energy_Model */
    for (int id = 0; id < s->n; ++id) {
        if (s->p.model == 2)
            mo_von_mises(&s->p, s->F + 9 * id, s->c[id], &s->g[id], &s->alpha[id], &s->psi[id],
                         &s->FP[id], s->sigma + 9 * id, &s->vol[id]);
        else if (s->p.model == 1)
            mo_neo_hookean(&s->p, s->F + 9 * id, s->c[id], &s->psi[id], &s->FP[id], s->sigma + 9 * id,
                           &s->vol[id]);
        else
            mo_fixed_corotated(&s->p, s->F + 9 * id, s->sigma + 9 * id);
    }
    const size_t gbytes = sizeof(double) * NCH * (size_t)s->gcap * 64;
    p2g(s, dt);
    memcpy(s->grid_p2g, s->grid, gbytes);
    grid_update(s, dt);
    memcpy(s->grid_upd, s->grid, gbytes);
    int rc = 0;
    /* Energy_Model 0 has no fracture: integrate skips the solve (:905-907) */
    if (s->p.phase_field && s->p.model != 0) {
        rc = phase_field(s, dt);
    } else {
        /* no damage solve: the damage increment is zero */
        const int nn = s->counts[1] * 64;
        for (int i = 0; i < nn; ++i) CH(s, CH_C)[i] = 0.0;
        s->cg_iters = 0;
    }
    boundary(s);
    g2p(s, dt);
    resort(s);
    return rc;
}

static int field_width(int f) {
    switch (f) {
    case MO_X: case MO_V: return 3;
    case MO_F: case MO_B: case MO_SIGMA: return 9;
    default: return 1;
    }
}
static double* field_ptr(const mo_sim* s, int f) {
    switch (f) {
    case MO_X: return s->x;
    case MO_V: return s->v;
    case MO_F: return s->F;
    case MO_B: return s->B;
    case MO_C: return s->c;
    case MO_ALPHA: return s->alpha;
    case MO_PSI: return s->psi;
    case MO_G: return s->g;
    case MO_MASS: return s->m;
    case MO_FP: return s->FP;
    case MO_VOL: return s->vol;
    case MO_SIGMA: return s->sigma;
    default: return NULL;
    }
}
int mo_sim_get(const mo_sim* s, int f, double* out) {
    double* p = field_ptr(s, f);
    if (!p) return -1;
    memcpy(out, p, sizeof(double) * (size_t)s->n * field_width(f));
    return 0;
}
int mo_sim_set(mo_sim* s, int f, const double* in) {
    double* p = field_ptr(s, f);
    if (!p) return -1;
    memcpy(p, in, sizeof(double) * (size_t)s->n * field_width(f));
    if (f == MO_X) resort(s);
    return 0;
}
/* Round the particle state to float32, as a store in FP32 would: used to
 * measure how far storage rounding alone moves a trajectory. */
void mo_sim_round_f32(mo_sim* s) {
    double* f3[] = {s->x, s->v};
    double* f9[] = {s->F, s->B};
    double* f1[] = {s->c, s->g, s->alpha, s->psi};
    for (int k = 0; k < 2; ++k)
        for (size_t i = 0; i < (size_t)s->n * 3; ++i) f3[k][i] = (double)(float)f3[k][i];
    for (int k = 0; k < 2; ++k)
        for (size_t i = 0; i < (size_t)s->n * 9; ++i) f9[k][i] = (double)(float)f9[k][i];
    for (int k = 0; k < 4; ++k)
        for (size_t i = 0; i < (size_t)s->n; ++i) f1[k][i] = (double)(float)f1[k][i];
}

int mo_sim_get_order(const mo_sim* s, int* order, uint64_t* keys) {
    if (order) memcpy(order, s->order, sizeof(int) * (size_t)s->n);
    if (keys) memcpy(keys, s->keys, sizeof(uint64_t) * (size_t)s->n);
    return 0;
}
void mo_sim_counts(const mo_sim* s, int counts[6]) {
    for (int i = 0; i < 4; ++i) counts[i] = s->counts[i];
    counts[4] = s->cg_iters;
    counts[5] = s->cg_breakdown;
}
double mo_sim_cg_norm(const mo_sim* s) { return s->cg_norm; }
int mo_sim_grid(const mo_sim* s, int stage, int ch, double* out) {
    if (!s->grid || ch < 0 || ch >= NCH) return -1;
    const double* src = stage == 0 ? s->grid_p2g : stage == 1 ? s->grid_upd : s->grid;
    memcpy(out, src + (size_t)ch * s->gcap * 64, sizeof(double) * (size_t)s->counts[1] * 64);
    return 0;
}
int mo_sim_pages(const mo_sim* s, uint64_t* page_key) {
    memcpy(page_key, s->page_key, sizeof(uint64_t) * (size_t)s->counts[1]);
    return s->counts[1];
}
void mo_sim_conservation(const mo_sim* s, double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0;
    for (int id = 0; id < s->n; ++id) {
        out[0] += s->m[id];
        for (int a = 0; a < 3; ++a) out[1 + a] += s->m[id] * s->v[3 * id + a];
    }
}

void mo_g2p_affine(int N, const double x[3], const double A[9], const double b[3], double vp[3],
                   double Bm[9], double G[9]) {
    const double dx = 1.0 / N, inv_dx = (double)N;
    int base[3];
    double xp[3], w[3][3], wg[3][3];
    for (int a = 0; a < 3; ++a) {
        base[a] = mo_base_node(x[a], N);
        xp[a] = x[a] - base[a] * dx;
    }
    weights(xp, inv_dx, w, wg);
    for (int a = 0; a < 3; ++a) vp[a] = 0;
    for (int k = 0; k < 9; ++k) { Bm[k] = 0; G[k] = 0; }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int kk = 0; kk < 3; ++kk) {
                const double xi[3] = {(base[0] + i) * dx, (base[1] + j) * dx, (base[2] + kk) * dx};
                double vi[3];
                for (int r = 0; r < 3; ++r) vi[r] = A[r] * xi[0] + A[r + 3] * xi[1] + A[r + 6] * xi[2] + b[r];
                const double weight = w[0][i] * w[1][j] * w[2][kk];
                const double d[3] = {i * dx - xp[0], j * dx - xp[1], kk * dx - xp[2]};
                const double g3[3] = {wg[0][i] * w[1][j] * w[2][kk], w[0][i] * wg[1][j] * w[2][kk],
                                      w[0][i] * w[1][j] * wg[2][kk]};
                for (int r = 0; r < 3; ++r) {
                    vp[r] += weight * vi[r];
                    for (int c = 0; c < 3; ++c) {
                        Bm[r + 3 * c] += weight * vi[r] * d[c];
                        G[r + 3 * c] += vi[r] * g3[c];
                    }
                }
            }
}

/* single-stage P2G-MLS for the product's P2G parity test */
int mo_p2g_mls(const mo_params* p, int n, const uint64_t* sorted_keys, const double* x,
               const double* v, const double* Bm, const double* sigma, const double* m,
               const double* c, const double* vol, const double* FP, double dt, int tp,
               const uint64_t* page_key, double* out) {
    (void)sorted_keys;
    mo_sim* s = (mo_sim*)calloc(1, sizeof(mo_sim));
    s->p = *p;
    s->n = n;
    s->M = p->N / 4 + 2;
    s->cap_tp = tp;
    s->gcap = tp;
    s->lookup = (int*)malloc(sizeof(int) * (size_t)s->M * s->M * s->M);
    for (int i = 0; i < s->M * s->M * s->M; ++i) s->lookup[i] = -1;
    for (int id = 0; id < tp; ++id) {
        const int pi = page_coord(page_key[id], 0), pj = page_coord(page_key[id], 1), pk = page_coord(page_key[id], 2);
        if (pi < s->M && pj < s->M && pk < s->M) s->lookup[(pi * s->M + pj) * s->M + pk] = id;
    }
    s->grid = (double*)calloc((size_t)NCH * tp * 64, sizeof(double));
    s->x = (double*)x; s->v = (double*)v; s->B = (double*)Bm; s->sigma = (double*)sigma;
    s->m = (double*)m; s->c = (double*)c; s->vol = (double*)vol; s->FP = (double*)FP;
    s->order = (int*)malloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int k = 0; k < n; ++k) s->order[k] = k;
    s->p.transfer = 0;
    p2g(s, dt);
    const int chs[6] = {CH_M, 1, 2, 3, CH_C, CH_W};
    for (int q = 0; q < 6; ++q) memcpy(out + (size_t)q * tp * 64, CH(s, chs[q]), sizeof(double) * (size_t)tp * 64);
    /* the phase-field sums of volP2G / preConditionP2G before their node factors:
     * sum_p V w and sum_p V FP w (P2GKernal.cu:1012, 1123) */
    const double inv_dx = (double)p->N;
    double* vw = out + (size_t)6 * tp * 64;
    double* vfw = out + (size_t)7 * tp * 64;
    memset(vw, 0, sizeof(double) * (size_t)tp * 128);
    for (int k = 0; k < n; ++k) {
        int b[3];
        double xp[3], w[3][3], wg[3][3];
        base_of(s, k, b, xp);
        weights(xp, inv_dx, w, wg);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int kk = 0; kk < 3; ++kk) {
                    const int ni = node_index(s, b[0] + i, b[1] + j, b[2] + kk);
                    const double wt = w[0][i] * w[1][j] * w[2][kk];
                    vw[ni] += vol[k] * wt;
                    vfw[ni] += vol[k] * FP[k] * wt;
                }
    }
    free(s->grid); free(s->lookup); free(s->order); free(s);
    return 0;
}
