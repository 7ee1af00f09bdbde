/*
 * mpmx oracle — serial FP64 C restatement of the reference's explicit MLS-MPM
 * substep with von Mises plasticity and plastic phase-field fracture.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product library (libmpmx_hip.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference checkout, directory GPU_MPM/).
 *
 * Conventions shared with the product (the "canonical" order):
 *  - a particle's cell key is the 64-bit SPGrid offset of its quadratic base
 *    node (Particle_kernal.cu:158-165);
 *  - particles are stably sorted by that key, ties kept in previous sorted
 *    order (the reference's within-cell order depends on atomic scheduling,
 *    Particle_kernal.cu:239, so it has no fixed order to match);
 *  - particle pages get sparse ids in ascending page-key (Morton) order,
 *    neighbour-only pages follow, also in ascending page-key order
 *    (the reference numbers them by atomicAdd order, Transformer_kernal.cu:85).
 * Matrices are 3x3 column-major: M(r,c) = m[r + 3c] (reference d_F layout).
 *
 * Pinned by closed forms (tests/test_oracle_pins.py): the branch quirks of
 * applyVonMises (compressive trial no flow, the dlambda cap, alpha storage
 * against the local alpha of G_c, the log-strain rescale), the walls and
 * grips (mo_node_boundary), the MLS speed clamp, the damage update and the
 * CFL step; the PCG's block-max "dot" (mo_vdot) in tests/test_oracle.py.
 * Parity unpinned (checked only as a faithful transcription, no golden vector
 * or closed form fixes them): the QR-SVD's choice among equivalent sign
 * conventions beyond what the SVD tests fix, and the reference's float-literal
 * constants (their effect is ~1e-8 relative, below every bar).
 */
#ifndef MPMX_ORACLE_H
#define MPMX_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct mo_params {
    int    N;            /* cells per axis, dx = 1/N                              */
    int    transfer;     /* 0 = MLS (P2GKernal.cu:6), 1 = APIC (P2GKernal.cu:512) */
    int    phase_field;  /* 1 = phase-field damage solve on                       */
    int    dot_mode;     /* 0 = reference block-max-128 "dot", 1 = true dot       */
    int    cg_max_iter;  /* 100 (ExplicitTimeIntegrator.cu:923)                   */
    int    vpage;        /* particles per virtual page, 256 (Transformer_kernal.cu:119) */
    int    wall_cells;   /* sticky wall width, 4 (TimeIntegrator.cu:135)          */
    int    grip_lo;      /* nodes with i <= grip_lo get v = (-grip_v,0,0); <0 off */
    int    grip_hi;      /* nodes with i >= grip_hi get v = (+grip_v,0,0)         */
    double grip_v;
    double E, nu;        /* Young's modulus, Poisson ratio (Materials.cpp:4-14)   */
    double V0;           /* particle rest volume                                  */
    double yield_stress; /* 800 in the reference (ExplicitTimeIntegrator.cu:188)  */
    double hardening;    /* 3  (:189)   */
    double alpha_c;      /* 1  (:194)   */
    double dlambda_cap;  /* 100 (:217)  */
    double eps_cap;      /* 3  (:268)   */
    double Gc0;          /* 3  (:320)   */
    double residual_c;   /* 0.1 (:330)  */
    double ps;           /* 0.01 (:323) */
    double theta;        /* float 1e-4 (:334) */
    double M;            /* parabolic_M = 50 (Materials.cpp:13) */
    double gravity[3];   /* (0,-10,0) (TimeIntegrator.cu:100) */
    double vmax;         /* MLS |v| clamp 11.137 (G2PKernal.cu:134-142) */
    double cg_tol;       /* 1e-6 (ExplicitTimeIntegrator.cu:924) */
    double crit_stretch; /* psi+ history only grows while max stretch >= this; 0 = always (reference) */
    int    model;        /* 2 von Mises (default), 1 Neo-Hookean phase field, 0 fixed corotated */
} mo_params;

/* reference defaults (constants listed above), N = 128, E = 1e5, nu = 0.3 */
void mo_default_params(mo_params* p);

/* ---------------- SPGrid masks / keys  (SPGridMask.cpp, Particle_kernal.cu) ---- */
void     mo_masks(uint64_t m[3]);                                  /* SPGridMask.cpp:78-129 */
uint64_t mo_bit_spread(uint64_t mask, int data);                   /* Particle_kernal.cu:31-45 */
int      mo_bit_pack(uint64_t mask, uint64_t data);                /* TimeIntegrator.cu:26-45 */
uint64_t mo_linear_offset(int i, int j, int k);                    /* SPGridMask.cpp:73-76 */
uint64_t mo_packed_add(uint64_t a, uint64_t b);                    /* Transformer_kernal.cu:6-13 */
int      mo_base_node(double x, int N);                            /* Particle_kernal.cu:162 */
/* keys for n particles, positions AoS xyz (double or float) */
void     mo_cell_keys_f64(int n, const double* x, int N, uint64_t* keys);
void     mo_cell_keys_f32(int n, const float* x, int N, uint64_t* keys);

/* canonical stable sort: perm[new] = old, sorted[new] = keys[perm[new]] */
void     mo_sort_keys(int n, const uint64_t* keys, int* perm, uint64_t* sorted);

/* page tables from sorted keys (Transformer_kernal.cu:129-209, canonical order).
 * Capacities: page_key[cap_tp], page2particle[cap_p+1], adj[7*cap_p], vpo[cap_p+1],
 * target[cap_vp], particle2cell[n], cell2particle[n+1].  Returns 0, or -1 on overflow.
 * counts[0..3] = numPage, numTotalPage, numVirtualPage, numCell. */
int mo_page_tables(int n, const uint64_t* sorted, int vpage, int cap_p, int cap_tp, int cap_vp,
                   uint64_t* page_key, int* page2particle, int* adj, int* vpo, int* target,
                   int* particle2cell, int* cell2particle, int counts[4]);

/* ---------------- 3x3 QR/Givens SVD (QRSVD.hpp, givens.hpp) ------------------ */
void mo_svd3(const double F[9], double U[9], double S[3], double V[9]);
/* number of QR sweeps the last mo_svd3 call made (QRSVD.hpp:252-271) */
int  mo_svd3_last_sweeps(void);

/* Gauss-Jordan inverse (ExplicitTimeIntegrator.cu:71-129), row-major 3x3 */
void mo_inverse3(const double in[9], double out[9]);
/* the PCG "dot" of n entries: mode 0 the largest 128-entry block sum (the
   reference's getDotProduct), mode 1 the true dot product */
double mo_vdot(int n, const double* a, const double* b, int mode);

/* von Mises return mapping + psi+ history + phase-field coefficients for one
 * particle (ExplicitTimeIntegrator.cu:132-370).  F is updated in place.
 * c = damage phase (1 intact).  Outputs FP (phase-field coefficient), the
 * stress contribution sigma = V0 * P * F^T and vol = V0 * det F. */
void mo_von_mises(const mo_params* p, double F[9], double c, double* g, double* alpha,
                  double* psi, double* FP, double sigma[9], double* vol);

/* Neo-Hookean phase-field energy (ExplicitTimeIntegrator.cu:372-446): psi+
 * history, FP, sigma = V0 P F^T, vol = V0 det F; F unchanged */
void mo_neo_hookean(const mo_params* p, const double F[9], double c, double* psi, double* FP, double sigma[9],
                    double* vol);
/* fixed corotated energy (ExplicitTimeIntegrator.cu:22-47): sigma = V0 P F^T */
void mo_fixed_corotated(const mo_params* p, const double F[9], double sigma[9]);

/* ---------------- whole-substep simulator ----------------------------------- */
typedef struct mo_sim mo_sim;

mo_sim* mo_sim_create(const mo_params* p, int n, const double* x, const double* v,
                      const double* mass);
void    mo_sim_destroy(mo_sim* s);
/* one substep (ExplicitTimeIntegrator.cu:881-995).  Returns 0, or a negative
 * code: -1 table overflow, -2 particle outside [2dx, 1-2dx], -3 CG breakdown. */
int     mo_sim_substep(mo_sim* s, double dt);

/* field ids for get/set (particle fields in original-id order) */
enum { MO_X = 0, MO_V = 1, MO_F = 2, MO_B = 3, MO_C = 4, MO_ALPHA = 5, MO_PSI = 6,
       MO_G = 7, MO_MASS = 8, MO_FP = 9, MO_VOL = 10, MO_SIGMA = 11 };
/* copies n*width doubles (x, v: 3 per particle AoS; F, B, sigma: 9 per particle,
 * column-major per particle; scalars: 1) */
int     mo_sim_get(const mo_sim* s, int field, double* out);
int     mo_sim_set(mo_sim* s, int field, const double* in);
/* round the particle state (x, v, F, B, c, g, alpha, psi) to float32 */
void    mo_sim_round_f32(mo_sim* s);
/* current sorted order: order[new] = original id; keys sorted */
int     mo_sim_get_order(const mo_sim* s, int* order, uint64_t* keys);
/* counts[0..5] = numPage, numTotalPage, numVirtualPage, numCell, cg_iters, cg_breakdown */
void    mo_sim_counts(const mo_sim* s, int counts[6]);
/* last residual norm of the phase-field solve */
double  mo_sim_cg_norm(const mo_sim* s);
/* grid channel ch (reference channel numbering 0..11) after the last substep's
 * P2G + update, page-major [numTotalPage*64]; plus the page keys. */
int     mo_sim_grid(const mo_sim* s, int stage, int ch, double* out);
int     mo_sim_pages(const mo_sim* s, uint64_t* page_key);
/* FP64 conservation sums over particles: [mass, px, py, pz] */
void    mo_sim_conservation(const mo_sim* s, double out[4]);
/* MPMSimulator::computeDt (Simulator.cpp:200-225) with getMaxParticleVel_c
 * (Particle_kernal.cu:421-440): min(dt_default, cfl dx / max speed); wave = 1
 * adds each particle's elastic wave speed sqrt((lambda + 2 mu) V0 / m).
 * *max_speed (may be NULL) receives the largest speed. */
double  mo_sim_cfl_dt(const mo_sim* s, double dt_default, double cfl, int wave, double* max_speed);

/* single-node and single-particle pieces the substep uses */
/* grips (PullGrid_kernal, TimeIntegrator.cu:169-261) then sticky walls
 * (collideWithGround, TimeIntegrator.cu:117-165) on node (i, j, k)'s velocity */
void    mo_node_boundary(const mo_params* p, int i, int j, int k, double v[3]);
/* MLS speed clamp (G2PKernal.cu:134-142) */
void    mo_speed_clamp(double vmax, double v[3]);
/* monotone damage update max(min(c, c + dc), 0) (G2PKernal.cu:155) */
double  mo_damage_update(double c, double dc);

/* Stage-wise entry points for single-kernel parity (used on the same inputs as
 * the product kernels).  All particle arrays are in the given (sorted) order. */
/* P2G-MLS grid of n particles already in sorted order; grid channels returned
 * for the given page list (canonical order from mo_page_tables). out[ch][tp*64]
 * for ch in {m, p0, p1, p2, sum w c, sum w, sum V w, sum V FP w} (8 channels) */
int mo_p2g_mls(const mo_params* p, int n, const uint64_t* sorted_keys, const double* x,
               const double* v, const double* Bm, const double* sigma, const double* m,
               const double* c, const double* vol, const double* FP, double dt, int tp,
               const uint64_t* page_key, double* out);

/* G2P gather of one particle from an affine grid velocity v_i = A x_i + b,
 * evaluated at the 27 stencil nodes (G2PKernal.cu:94-185): returns the particle
 * velocity v_p = sum w v_i and B = sum w v_i (x_i - x_p)^T (column-major), plus
 * the APIC velocity gradient sum v_i grad w^T.  A is column-major. */
void mo_g2p_affine(int N, const double x[3], const double A[9], const double b[3], double vp[3],
                   double Bm[9], double G[9]);

#ifdef __cplusplus
}
#endif
#endif
