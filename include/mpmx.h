/*
 * mpmx.h — C ABI of libmpmx_hip.so, the MI355X (gfx950) implementation of the
 * explicit MLS-MPM substep with von Mises plasticity and plastic phase-field
 * fracture.
 *
 * The reference (KemengHuang/Explicit-GPUMPM-Ductile-Fracture) has no C ABI:
 * its boundary is the C++ classes MPMSimulator, MPMTimeIntegrator /
 * ExplicitTimeIntegrator, Particles, SPGrid and DomainTransformer with public
 * device pointers.  Each entry point below names the reference method it
 * replaces (paths relative to the reference's GPU_MPM/ directory); the C++
 * classes of the same names in this repository are thin wrappers over these
 * calls.
 *
 * Conventions
 *  - POD structs only; no C++ exception crosses the ABI; every call returns an
 *    mpmx_status.  mpmx_last_error() gives a message for the last failure.
 *  - The context owns all device memory and one HIP stream.  Host buffers are
 *    caller-owned and copied.
 *  - Stage calls are asynchronous on the context stream; mpmx_get_*,
 *    mpmx_sync and mpmx_destroy synchronise.  Device-side errors (particle left
 *    the domain, non-finite state, table capacity exceeded) are latched in a
 *    device error word and reported at the next synchronising call.  A PCG
 *    breakdown is not an error: as in the reference (ExplicitTimeIntegrator.cu:
 *    959-962) the substep goes on with the damage unchanged, and the breakdown
 *    is counted in mpmx_stats.cg_breakdown / mpmx_scene_stats.cg_breakdown.
 *  - A call that fails after it has started to replace the particle state
 *    (mpmx_set_particles*) leaves the context empty: the stage calls then
 *    return MPMX_ERR_ARG until particles are set again.
 *  - One context per host thread.
 *  - Particle fields are returned in the caller's original particle order.
 *    3x3 matrices are column-major per particle: M(r,c) = m[r + 3c].
 */
#ifndef MPMX_H
#define MPMX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mpmx_status {
    MPMX_OK = 0,
    MPMX_ERR_HIP = 1,            /* HIP runtime call failed                       */
    MPMX_ERR_ARG = 2,            /* bad argument / call order                     */
    MPMX_ERR_CAPACITY = 3,       /* page or virtual-page table capacity exceeded  */
    MPMX_ERR_OUT_OF_DOMAIN = 4,  /* a particle left [2dx, 1-2dx]                  */
    MPMX_ERR_NONFINITE = 5,      /* NaN/Inf in particle state                     */
    MPMX_ERR_CG_BREAKDOWN = 6,   /* reserved; a breakdown is counted, not returned */
    MPMX_ERR_NO_DEVICE = 7       /* no usable HIP device                          */
} mpmx_status;

/* Scene constants.  The reference hard-codes these as compile-time constants
 * and kernel literals (Setting.h:3-62, ExplicitTimeIntegrator.cu:188-335,
 * Materials.cpp:4-14, TimeIntegrator.cu:100-165, G2PKernal.cu:134-142); here
 * they are per-scene runtime values.  mpmx_default_params() fills the
 * reference values. */
typedef struct mpmx_params {
    int    N;            /* cells per axis, dx = 1/N (Setting.h:13)               */
    int    transfer;     /* 0 = MLS (P2GKernal.cu:6), 1 = APIC (P2GKernal.cu:512)  */
    int    phase_field;  /* 1 = solve the damage field every substep              */
    int    dot_mode;     /* 0 = reference "dot" = max over 128-entry block sums,
                            1 = true inner product (ExplicitTimeIntegrator.cu:573-609) */
    int    cg_max_iter;  /* 100 */
    int    vpage;        /* particles per virtual page (Transformer_kernal.cu:119): 64, 128 or 256 */
    int    wall_cells;   /* sticky wall width in cells, 4 */
    int    grip_lo;      /* grid nodes with i <= grip_lo move at (-grip_v,0,0); < 0 disables */
    int    grip_hi;      /* grid nodes with i >= grip_hi move at (+grip_v,0,0) */
    double grip_v;
    double E, nu;        /* Young's modulus, Poisson ratio */
    double V0;           /* particle rest volume */
    double yield_stress, hardening, alpha_c, dlambda_cap, eps_cap;
    double Gc0, residual_c, ps, theta, M;
    double gravity[3];
    double vmax;         /* MLS grid-to-particle speed clamp */
    double cg_tol;
    double crit_stretch; /* psi+ history grows only while max principal stretch >= this (0: always) */
    int    model;        /* energy (Setting.h:38 Energy_Model): 2 = von Mises plastic phase-field
                            fracture (default), 1 = Neo-Hookean phase-field fracture
                            (computeContributionNeoHookean), 0 = fixed corotated, no
                            fracture (computeContributionFixedCorotated; phase field off) */
} mpmx_params;

typedef enum mpmx_precision {
    MPMX_FP32 = 0,       /* FP32 particle state and grid                          */
    MPMX_MIXED = 1,      /* x and F FP32, v/C/c/alpha/psi/g BF16, FP32 grid; stress
                            arithmetic as mpmx_config.stress_fp64 says (88 B of
                            state per particle against 120 B in FP32)              */
    MPMX_MIXED_F16 = 2   /* as MPMX_MIXED with F stored as FP16(F - I) (70 B per
                            particle; one store rounds F by up to 2^-11 |F - I|)    */
} mpmx_precision;

typedef struct mpmx_config {
    int64_t max_particles;  /* capacity; 0 = size of the first mpmx_set_particles */
    int64_t max_pages;      /* total (particle + neighbour) page capacity; 0 = auto */
    int     device;         /* HIP device ordinal */
    int     precision;      /* mpmx_precision */
    double  svd_gap;        /* FP32 stress arithmetic: FP64 redo when min singular-value gap < this (1e-4) */
    int     cg_check_every; /* PCG: the first host check of a substep comes after as many
                               iterations as the previous substep took, later checks every
                               this many iterations (1)                                */
    int     p2g_variant;    /* transfer kernels: 0 = default (2); 1 = page tiles over
                               virtual pages with LDS and global float atomics (P2G,
                               G2P, PCG K s); 2 = 2x2x2-page blocks, P2G and K s as
                               single-wave gathers without atomics; 3 = as 2, but the
                               P2G scatters each block's own particles into a 10^3-node
                               tile and adds the neighbours' halos in a second pass
                               (same sums to rounding, slower on every measured scene);
                               other values: MPMX_ERR_ARG                        */
    int     block;          /* threads per transfer block; 0 = auto (= vpage) */
    int     stress_fp64;    /* arithmetic of the SVD and return mapping: 1 = FP64 for every
                               particle (default; state stays in its storage precision),
                               0 = FP32, redone in FP64 where the singular-value gap is
                               below svd_gap (0 = never)                              */
    int     use_graph;      /* 1: mpmx_run / mpmx_time_run replay each substep as two
                               captured HIP graphs (sort .. PCG, then G2P) instead of ~50
                               launches; the PCG runs as many iterations in the graph as
                               recent substeps needed and finishes any remaining ones
                               eagerly.  Results are identical either way.  Default 0:
                               on ROCm 7.2 the replays were 1-2 % slower than the direct
                               launches for this kernel sequence                       */
    int     store_stress;   /* 1: the fused stress pass of mpmx_substep / mpmx_run also
                               stores each particle's sigma and FP for mpmx_get_field
                               (40 B per particle per substep that the substep itself
                               does not need); 0 (default): only mpmx_stress stores
                               them, and mpmx_get_field of MPMX_FIELD_SIGMA / _FP after a
                               fused substep returns MPMX_ERR_ARG                      */
    int     sort_variant;   /* the substep's sort (same order either way: key ascending,
                               ties in the previous order): 1 = LSD radix sort of every
                               key; 2 = the movers only (the particles whose key changed
                               since the last sort are radix-sorted and merged with the
                               rest, which are in order already); 0 = default: 2 while the
                               last published mover count is at most n/4, 1 above (back
                               to 2 below n/8); other values: MPMX_ERR_ARG             */
} mpmx_config;

typedef struct mpmx_ctx mpmx_ctx;

typedef struct mpmx_stats {
    int64_t num_particles;
    int     num_page;          /* particle pages             (Transformer.cuh:39 _numPage)        */
    int     num_total_page;    /* + neighbour-only pages     (_numTotalPage)                      */
    int     num_virtual_page;  /* (_numVirtualPage)                                             */
    int     cg_iters;          /* PCG iterations of the last substep                             */
    int     cg_breakdown;
    int     substeps;          /* substeps taken by this context                                 */
    double  cg_norm;           /* final max|r| of the last PCG                                   */
    double  mass;              /* FP64 sums over particles: mass and momentum                    */
    double  momentum[3];
    double  kinetic_energy;
    int64_t svd_fallbacks;     /* stress_fp64 = 0: particle updates redone in FP64 for a small
                                  singular-value gap, since the particles were set             */
} mpmx_stats;

/* one scene of a batch context (mpmx_get_scene_stats) */
typedef struct mpmx_scene_stats {
    int64_t num_particles;
    int64_t first_particle;    /* the scene's particles are [first, first + num) of the upload order */
    int     num_page;          /* the scene's particle pages                                     */
    int     num_total_page;    /* + its neighbour-only pages                                     */
    int     cg_iters;          /* iterations of its last PCG                                     */
    int     cg_breakdown;
    double  cg_norm;
    double  mass;              /* FP64 sums over the scene's particles                           */
    double  momentum[3];
    double  kinetic_energy;
} mpmx_scene_stats;

/* per-kernel average durations of the last timed region (mpmx_timing_begin/end) */
typedef struct mpmx_kernel_time {
    char   name[32];
    int    launches;
    double total_ms;
} mpmx_kernel_time;

void        mpmx_default_params(mpmx_params* p);
void        mpmx_default_config(mpmx_config* c);
const char* mpmx_status_string(mpmx_status s);
const char* mpmx_last_error(const mpmx_ctx* ctx);
int         mpmx_device_count(void);

/* MPMSimulator::MallocDeviceMEM + Particles::Particles_Malloc_MEM +
 * SPGrid::SPGrid_MEM_Malloc + DomainTransformer::Transformer_Malloc_MEM
 * (Simulator.cpp:16-31, Particles.cpp:54-94, SPGrid.cpp:10-55, Transformer.cpp:5-35) */
mpmx_status mpmx_create(const mpmx_params* p, const mpmx_config* cfg, mpmx_ctx** out);
/* MPMSimulator::FreeDeviceMEM and the destructors (Simulator.cpp:33-43, Particles.cpp:96-125,
 * SPGrid.cpp:63-79, Transformer.cpp:50-69) */
mpmx_status mpmx_destroy(mpmx_ctx* ctx);

/* Particles::initialize / initialize_kernal (Particles.cpp:8-31, Particle_kernal.cu:441-481)
 * and MPMSimulator::buildIntegrator's F = I, B = 0 (Simulator.cpp:67-73, TimeIntegrator.cu:46-51).
 * x, v: n*3 doubles (xyz per particle); mass: n doubles.  Keys and the first
 * canonical sort are computed here. */
mpmx_status mpmx_set_particles(mpmx_ctx* ctx, int64_t n, const double* x, const double* v,
                               const double* mass);

/* Batched scenes: num_scenes independent scenes, each with its own parameters
 * (grid N, material, yield, grips, walls, gravity, energy model), advanced
 * together by every stage and substep call with one dt.  No counterpart in
 * the reference, which runs one scene per process (MPMSimulator::build, Simulator.cpp:76-198).
 * transfer, vpage, dot_mode and cg_max_iter must agree across scenes;
 * p2g_variant 1 is single-scene only.  Each scene's PCG keeps its own
 * scalars, iteration count and stopping test.  Positions stay in each
 * scene's own unit cube.  The key space limits the batch:
 * 6 + 3 ceil(log2(max N / 4)) + ceil(log2(num_scenes)) <= 32 (256 scenes of
 * N <= 256).  A single-scene context is mpmx_create_batch(1, ...). */
mpmx_status mpmx_create_batch(int num_scenes, const mpmx_params* params, const mpmx_config* cfg,
                              mpmx_ctx** out);
/* counts[num_scenes]; x, v (3 per particle) and mass hold the scenes one
 * after another.  Fields come back in this order (mpmx_get_field). */
mpmx_status mpmx_set_particles_batch(mpmx_ctx* ctx, const int64_t* counts, const double* x, const double* v,
                                     const double* mass);
int         mpmx_num_scenes(const mpmx_ctx* ctx);
/* counts, PCG state and FP64 conservation sums of one scene */
mpmx_status mpmx_get_scene_stats(mpmx_ctx* ctx, int scene, mpmx_scene_stats* out);

/* ---- stages of one substep, in order (ExplicitTimeIntegrator.cu:881-995) ---- */
/* Particles::reorder / sort_by_offsets (Particle_kernal.cu:484-573): canonical
 * stable sort by cell key, then every particle field is gathered to the new order */
mpmx_status mpmx_bin_sort(mpmx_ctx* ctx);
/* DomainTransformer::rebuild (Transformer.cpp:71-82) */
mpmx_status mpmx_rebuild(mpmx_ctx* ctx);
/* ExplicitTimeIntegrator::computeForceCoefficient -> applyVonMises
 * (ExplicitTimeIntegrator.cu:846-865, 132-370) */
mpmx_status mpmx_stress(mpmx_ctx* ctx);
/* MPMTimeIntegrator::transferP2G (TimeIntegrator.cu:299-372) */
mpmx_status mpmx_p2g(mpmx_ctx* ctx, double dt);
/* MPMTimeIntegrator::undateGrid + resolveCollision (TimeIntegrator.cu:378-403),
 * also forms the phase-field right-hand side and Jacobi diagonal */
mpmx_status mpmx_grid_update(mpmx_ctx* ctx, double dt);
/* the damage solve inside ExplicitTimeIntegrator::integrate, ExplicitTimeIntegrator.cu:909-968 */
mpmx_status mpmx_phase_field(mpmx_ctx* ctx, double dt);
/* MPMTimeIntegrator::transferG2P (TimeIntegrator.cu:411-474) */
mpmx_status mpmx_g2p(mpmx_ctx* ctx, double dt);

/* ExplicitTimeIntegrator::integrate (ExplicitTimeIntegrator.cu:881-995):
 * bin_sort, rebuild, stress, p2g, grid_update, phase_field, g2p.  With the
 * default gather P2G the stress pass also writes the P2G's per-particle
 * terms (one pass instead of two); the results are bit-identical to calling
 * the stages one by one. */
mpmx_status mpmx_substep(mpmx_ctx* ctx, double dt);
/* nsteps substeps with fixed dt (no host synchronisation in between unless the
 * PCG needs a convergence check) */
mpmx_status mpmx_run(mpmx_ctx* ctx, int nsteps, double dt);
/* Particles::getMaxParticleVel_c (Particle_kernal.cu:421-440): max |v|^2 over
 * particles (one FP32 value; synchronises) */
mpmx_status mpmx_max_velocity_sq(mpmx_ctx* ctx, double* out);
/* MPMSimulator::computeDt (Simulator.cpp:200-225): min(dt_default, cfl/(N max|v|)),
 * max over the particles of every scene with each scene's own N.  The
 * reduction runs on the device (partials in context memory, no allocation per
 * call); the result is also left on the device as the step of the next
 * substep launched with dt <= 0.  Synchronises to return *dt_out. */
mpmx_status mpmx_compute_dt(mpmx_ctx* ctx, double dt_default, double cfl, double* dt_out);
/* MPMSimulator::simulateStick's computeDt + integrate loop (Simulator.cpp:
 * 227-259) without host round trips: before each of nsteps substeps a device
 * reduction sets dt = min(dt_default, cfl / max_p((|v_p| + wave c_p) N_s)),
 * with c_p = sqrt((lambda + 2 mu) V0 / m_p) the particle's elastic wave speed
 * when wave = 1 (0: speed only, the reference's rule), and the substep's
 * kernels read that dt on the device.  *time_out (may be NULL) receives the
 * simulated time of the nsteps substeps (one synchronisation at the end). */
mpmx_status mpmx_run_cfl(mpmx_ctx* ctx, int nsteps, double dt_default, double cfl, int wave, double* time_out);
/* mpmx_compute_dt that allocates nothing (no counterpart in the reference):
 * for callers that cannot allocate device memory, such as code under HIP
 * graph capture or a serving loop that forbids hipMalloc.  Same result, bit
 * for bit.  Its scratch (4096 partial maxima and two FP64 results, 16 KiB)
 * is one static device array per GPU, held until the process ends and shared
 * by every call on that GPU: calls that overlap on one GPU race.  Contexts of
 * more than 33,554,432 particles are refused with MPMX_ERR_CAPACITY. */
mpmx_status mpmx_compute_dt_noalloc(mpmx_ctx* ctx, double dt_default, double cfl, double* dt_out);

mpmx_status mpmx_sync(mpmx_ctx* ctx);

/* field ids for mpmx_get_field / mpmx_set_field */
typedef enum mpmx_field {
    MPMX_FIELD_X = 0,       /* 3 per particle */
    MPMX_FIELD_V = 1,       /* 3 */
    MPMX_FIELD_F = 2,       /* 9, column-major */
    MPMX_FIELD_B = 3,       /* 9, the affine matrix C (MLS/APIC B) */
    MPMX_FIELD_C = 4,       /* damage phase c (1 = intact) */
    MPMX_FIELD_ALPHA = 5,   /* equivalent plastic strain */
    MPMX_FIELD_PSI = 6,     /* psi+ history maximum */
    MPMX_FIELD_G = 7,       /* degradation g(c) */
    MPMX_FIELD_MASS = 8,
    MPMX_FIELD_FP = 9,      /* phase-field coefficient of the last stress update */
    MPMX_FIELD_VOL = 10,    /* V0 det F of the last stress update */
    MPMX_FIELD_SIGMA = 11   /* 9: V0 P F^T of the last stress update */
} mpmx_field;
/* Particles::Attribs_Back_Host (Particles.cpp:33-52): copies n*width doubles in
 * original particle order */
mpmx_status mpmx_get_field(mpmx_ctx* ctx, int field, double* host, size_t count);
mpmx_status mpmx_set_field(mpmx_ctx* ctx, int field, const double* host, size_t count);

/* current canonical order: order[i] = original id of the particle in sorted
 * slot i; keys[i] = its 64-bit SPGrid offset (Particles::d_indices, cell_id) */
mpmx_status mpmx_get_order(mpmx_ctx* ctx, int32_t* order, uint64_t* keys);

/* DomainTransformer tables of the last rebuild (Transformer.cuh:24-50):
 * page_key[num_total_page] (= d_pageOffset >> 12), page2particle[num_page+1],
 * adj[7*num_page] (neighbour j of page p at adj[j*num_page+p]),
 * vpo[num_page+1], target[num_virtual_page] (1-based, as d_targetPage),
 * particle2cell[n] (1-based), cell2particle[num_cell+1]; counts[4] = numPage,
 * numTotalPage, numVirtualPage, numCell.  Any pointer may be NULL. */
mpmx_status mpmx_get_tables(mpmx_ctx* ctx, uint64_t* page_key, int32_t* page2particle, int32_t* adj,
                            int32_t* vpo, int32_t* target, int32_t* particle2cell,
                            int32_t* cell2particle, int32_t counts[4]);

/* grid channel after the last grid update, page-major [num_total_page*64]:
 * ch 0 mass, 1-3 velocity, 4 damage c (or its increment after the PCG),
 * 5 sum of weights, 6 phase-field rhs, 7 Jacobi diagonal; 8 = PCG solution x.
 * SPGrid::d_channels (SPGrid.cpp:34-49) with channels renumbered. */
mpmx_status mpmx_get_grid(mpmx_ctx* ctx, int channel, float* host, size_t count);
/* the grid channels right after P2G (before the update): 0 m, 1-3 momentum,
 * 4 sum w c, 5 sum w, 6 sum V w, 7 sum V FP w */
mpmx_status mpmx_get_grid_p2g(mpmx_ctx* ctx, int channel, float* host, size_t count);

mpmx_status mpmx_get_stats(mpmx_ctx* ctx, mpmx_stats* out);

/* Raw device pointers of the current buffers, for callers that work on the
 * device data directly as the reference's classes expose it (public members of
 * Particles, SPGrid and DomainTransformer: Particles.cuh:24-74, SPGrid.h:15-33,
 * Transformer.cuh:24-50).  Particle buffers are double-buffered and swap on
 * every sort, so re-read these after mpmx_bin_sort / mpmx_substep.
 * Layouts: pos, vel float[3n] xyz; F, B, sigma float[9n] component-major
 * (entry (r,c) of particle i at [(r+3c)*n + i]); scalars float[n]; keys
 * uint32[n] = Morton page key << 6 | cell; grid float[pages][8][64];
 * PCG vectors float[pages*64]; page_key uint32; adj int[7][cap_pages];
 * target int (0-based).  Fields absent in the active precision are NULL. */
typedef struct mpmx_device_ptrs {
    void* pos;
    void* vel;
    void* F;
    void* B;
    void* phase_c;
    void* alpha;
    void* psi;
    void* g;
    void* mass;
    int32_t* ids;
    float* sigma;
    float* FP;
    float* vol;
    uint32_t* keys;
    float* grid;
    float* pcg_x;
    float* pcg_r;
    float* pcg_z;
    float* pcg_s;
    float* pcg_q;
    uint32_t* page_key;
    int32_t* page2particle;
    int32_t* adj;
    int32_t* vpo;
    int32_t* target;
    int64_t cap_pages;      /* stride of adj */
    int64_t cap_total_pages;
    int     precision;
} mpmx_device_ptrs;
mpmx_status mpmx_get_device_pointers(mpmx_ctx* ctx, mpmx_device_ptrs* out);

/* kernel timing with HIP events on the context stream */
mpmx_status mpmx_timing_begin(mpmx_ctx* ctx);
mpmx_status mpmx_timing_end(mpmx_ctx* ctx, mpmx_kernel_time* out, int max_entries, int* n_entries);
/* elapsed milliseconds of `nsteps` substeps measured with events on the
 * context stream (inputs already resident) */
mpmx_status mpmx_time_run(mpmx_ctx* ctx, int nsteps, double dt, double* ms);

/* ---- synthetic scenes (host only, no GPU) ----
 * scene ids: 0 cube (20^3 cells at [54,74)^3 of 128^3, 8 ppc jittered),
 * 1 notched bar (1,048,576 uniform particles, 256^3), 2 notched bar 8 ppc,
 * 3 two-block impact (2 x 80^3 cells x 8 ppc, 512^3), 4 sphere, 5 scaled cube.
 * Query the particle count with x == NULL; returns particle count or -1. */
int64_t mpmx_scene(int scene, uint64_t seed, double scale, mpmx_params* p, double* x, double* v,
                   double* mass, double* dt);
/* batches of independent scenes: config 4 = 64 scenes of 10k-500k particles,
 * 5 = 256 scenes of 5k-200k (log-uniform; cube / notched bar / sphere in turn,
 * grids 64^3-256^3, E log-uniform in [1e4, 1e6]); `scale` scales the counts.
 * *num_scenes: in, at most this many scenes when > 0; out, the count.  Size
 * the arrays with null pointers first; returns the particle total or -1.
 * params[s], counts[s]; x, v, mass hold the scenes one after another. */
int64_t mpmx_scene_batch(int config, uint64_t seed, double scale, int* num_scenes, mpmx_params* params,
                         int64_t* counts, double* x, double* v, double* mass, double* dt);

/* ---- launch shapes and their autotuner ----
 * None of these changes any result: the block kernels stride over the
 * device-side block count, and per-particle work does not depend on how it is
 * grouped.  No reference counterpart (its launch shapes are fixed). */
typedef struct mpmx_tuning {
    int particle_wg;  /* threads per workgroup of the per-particle kernels (stress, P2G payload): 64/128/256 */
    int p2g_pct;      /* workgroups of the block P2G, in % of the 2x2x2-block count (10..1000)      */
    int g2p_pct;      /* ... of the block G2P                                                         */
    int cg_pct;       /* ... of the block K s kernels of the phase-field solve                        */
} mpmx_tuning;
mpmx_status mpmx_get_tuning(mpmx_ctx* ctx, mpmx_tuning* out);
mpmx_status mpmx_set_tuning(mpmx_ctx* ctx, const mpmx_tuning* t);
/* Times trial_steps substeps of the context's current state for each
 * candidate shape (particle_wg 64/128/256 with the block kernels at 50/100/200 %),
 * restoring the state after every trial, so the simulation does not advance;
 * keeps for each kernel family the fastest candidate and returns it in *best
 * (may be NULL). */
mpmx_status mpmx_autotune(mpmx_ctx* ctx, double dt, int trial_steps, mpmx_tuning* best);

/* ---- scene files and particle files (host; checkpoints also touch the GPU) ---- */
/* Sample::readObj (LoadSample.cpp:163-194): a point-list OBJ, one particle per
 * "v x y z" line, y moved by y_shift (the reference's loader uses -0.2).
 * Returns the count (x may be NULL to size it) or -1. */
int64_t mpmx_read_obj_points(const char* path, double y_shift, double* x, int64_t cap);
/* Sample::LoadSDF + GenerateUniformSamples + Initial_data (LoadSample.cpp:3-160):
 * particles uniform inside a level set (text: "ni nj nk", "minx miny minz",
 * "dx", ni*nj*nk values, i fastest), scaled into the box res[3] (cells) at
 * center[3] as the reference places it, samples_per_cell per grid cell (the
 * reference uses 20); a seeded SplitMix64 stream replaces rand().  Positions
 * in units of the N^3 domain.  Returns the count (x may be NULL) or -1. */
int64_t mpmx_sample_sdf(const char* path, uint64_t seed, int N, const unsigned* center, const unsigned* res,
                        double samples_per_cell, double* x, int64_t cap);
/* MPMSimulator::writePartio (declared at Simulator.h:35, never defined in the
 * reference): every particle in upload order as binary PLY: x y z vx vy vz c
 * (float) and the scene index (int). */
mpmx_status mpmx_write_particles(mpmx_ctx* ctx, const char* path);
/* Checkpoint of the whole particle state and the context's parameters; a
 * context loaded from it (on `device`) continues bit-exactly.  No reference
 * counterpart (the reference keeps no checkpoints). */
mpmx_status mpmx_save_checkpoint(mpmx_ctx* ctx, const char* path);
mpmx_status mpmx_load_checkpoint(const char* path, int device, mpmx_ctx** out);

#ifdef __cplusplus
}
#endif
#endif
