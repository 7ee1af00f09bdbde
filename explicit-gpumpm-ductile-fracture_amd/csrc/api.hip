// api.hip — the C ABI of libmpmx_hip.so (include/mpmx.h): context lifecycle,
// substep stages, field access, statistics and kernel timing.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>

#include "access.h"
#include "mpmx_internal.h"
#include "precision.h"

using namespace mpmx;

namespace mpmx {

// ---- device kernels owned by the API --------------------------------------
__global__ void k_upload_fp32(PState s, const double* __restrict__ x, const double* __restrict__ v,
                              const double* __restrict__ m, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s.x[3 * i + a] = (float)x[3 * i + a];
        s.v[3 * i + a] = (float)v[3 * i + a];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        s.B[(int64_t)k * n + i] = 0.f;
        s.F[(int64_t)k * n + i] = (k % 3 == k / 3) ? 1.f : 0.f;
    }
    s.c[i] = 1.f;      // Particle_kernal.cu:451-454
    s.g[i] = 1.f;
    s.alpha[i] = 0.f;
    s.psi[i] = 0.f;
    s.mass[i] = (float)m[i];
    s.id[i] = i;
}
__global__ void k_upload_mixed(MState s, const double* __restrict__ x, const double* __restrict__ v,
                               const double* __restrict__ m, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s.x[3 * i + a] = (float)x[3 * i + a];
        s.v[3 * i + a] = f2bf((float)v[3 * i + a]);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        s.B[(int64_t)k * n + i] = f2bf(0.f);
        if (s.F32) s.F32[(int64_t)k * n + i] = (k % 3 == k / 3) ? 1.f : 0.f;
        else s.F[(int64_t)k * n + i] = f2h(0.f);
    }
    s.c[i] = f2bf(1.f);
    s.g[i] = f2bf(1.f);
    s.alpha[i] = f2bf(0.f);
    s.psi[i] = f2bf(0.f);
    s.mass[i] = (float)m[i];
    s.id[i] = i;
}

// ---- particle reductions: one partial per workgroup, then one workgroup
// combines the partials in a fixed order (no atomics, bit-reproducible) ----
// mpmx_ctx::red_part holds kRedParts * 5 doubles of partials, then results
// from double kRedResult on
constexpr int kRedParts = 1024;
constexpr int kRedResult = 5 * kRedParts;
constexpr int kRedDoubles = kRedResult + 8;

__global__ void k_clear_time(DevCounts* dc) {
    if (threadIdx.x == 0) dc->time = 0.0;
}

// FP64 sums of mass, momentum and kinetic energy of particles [0, n) of the
// arrays given (a range of the current order): part[5 * blockIdx.x + q]
__global__ __launch_bounds__(256) void k_conservation_part(const float* __restrict__ mass,
                                                           const float* __restrict__ v,
                                                           const uint16_t* __restrict__ vb, int n,
                                                           double* __restrict__ part) {
    __shared__ double lds[5][4];
    double acc[5] = {0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double m = mass[i];
        double vv[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) vv[a] = vb ? (double)bf2f(vb[3 * i + a]) : (double)v[3 * i + a];
        acc[0] += m;
        acc[1] += m * vv[0];
        acc[2] += m * vv[1];
        acc[3] += m * vv[2];
        acc[4] += 0.5 * m * (vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        double a = acc[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
        if ((threadIdx.x & 63) == 0) lds[q][threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += lds[threadIdx.x][w];
        part[5 * blockIdx.x + threadIdx.x] = s;
    }
}
// out[q] = sum of the G partials of quantity q, in workgroup order
__global__ void k_conservation_final(const double* __restrict__ part, int G, double* __restrict__ out) {
    if (threadIdx.x < 5) {
        double s = 0;
        for (int g = 0; g < G; ++g) s += part[5 * g + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// Largest CFL speed of the particles: q_p = ((|v_p| + wave c_p) N_s)^2 with
// c_p = sqrt((lambda + 2 mu) V0 / m_p) the elastic wave speed (mode bit 0)
// and N_s the cells per axis of the particle's scene (mode bit 1); mode 0 is
// |v_p|^2 (getMaxParticleVel_c).  Waves hold particles of one scene (the
// per-particle wave map), so a wave reads its scene's constants once.
// part[blockIdx.x] = the workgroup's largest q (floats, exact under max).
__global__ __launch_bounds__(256) void k_speed_part(DevParams dp, const float* __restrict__ v,
                                                    const uint16_t* __restrict__ vb, const float* __restrict__ mass,
                                                    WaveMap wm, int n, int mode, float* __restrict__ part) {
    __shared__ float lds[4];
    const int64_t W = wm.first ? wm.waves : ((int64_t)n + 63) / 64;
    const int lane = threadIdx.x & 63;
    float best = 0.f;
    for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < W;
         w += ((int64_t)gridDim.x * blockDim.x) >> 6) {
        int i, scene;
        bool valid;
        if (wm.first) {
            i = wm.first[w] + lane;
            scene = __builtin_amdgcn_readfirstlane(wm.scene[w]);
            valid = lane < wm.count[w];
        } else {
            i = (int)(w * 64 + lane);
            scene = 0;
            valid = i < n;
        }
        if (!valid) continue;
        float va[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) va[a] = vb ? bf2f(vb[3 * i + a]) : v[3 * i + a];
        float q = va[0] * va[0] + va[1] * va[1] + va[2] * va[2];
        if (mode) {
            const DevParams& sp = dp.scenes[scene];
            float s = sqrtf(q);
            if (mode & 1) s += sqrtf((sp.lambda + 2.0f * sp.mu) * sp.V0 / mass[i]);
            if (mode & 2) s *= (float)sp.N;
            q = s * s;
        }
        best = fmaxf(best, q);
    }
    best = combine_waves<true>(wave_max(best), lds);
    if (threadIdx.x == 0) part[blockIdx.x] = best;
}
// out[0] = the largest q, out[1] = dt = min(dt_default, cfl / sqrt(q)) (FP64,
// Simulator.cpp:213); with set_dt the device keeps dt for the next substep
// launched with dt <= 0 and adds it to the simulated time
__global__ __launch_bounds__(256) void k_speed_final(const float* __restrict__ part, int G, double dt_default,
                                                     double cfl, int set_dt, DevCounts* dc, double* __restrict__ out) {
    __shared__ float lds[4];
    const float q = reduce_parts<true>(part, G, lds);
    if (threadIdx.x == 0) {
        const double vmax = sqrt((double)q);
        double dt = dt_default;
        if (vmax > 0 && cfl / vmax < dt_default) dt = cfl / vmax;
        out[0] = (double)q;
        out[1] = dt;
        if (set_dt) {
            dc->dt = (float)dt;
            dc->time += (double)(float)dt;
        }
    }
}

// cell tables (Transformer_kernal.cu:15-36): head flags, then ids
__global__ void k_cell_flags(const uint32_t* __restrict__ keys, int n, int* flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}
__global__ void k_cell_ids(const uint32_t* __restrict__ keys, int n, const int* __restrict__ excl,
                           int* p2c, int* c2p, const int* num_cell) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool head = (i == 0 || keys[i] != keys[i - 1]);
    p2c[i] = excl[i] + (head ? 1 : 0);
    if (head) c2p[excl[i]] = i;
    if (i == n - 1) c2p[*num_cell] = n;
}

int particle_blocks(const mpmx_ctx* c) {
    const int64_t threads = c->wmap.first ? c->wmap.waves * 64 : c->n;
    const int wg = c->particle_wg;
    return (int)std::max<int64_t>(1, (threads + wg - 1) / wg);
}

// the captured substeps no longer match the context (buffers, capacities or
// launch shapes changed)
void destroy_graphs(mpmx_ctx* c) {
    for (auto& g : c->graphs) {
        if (g.a) (void)hipGraphExecDestroy(g.a);
        if (g.b) (void)hipGraphExecDestroy(g.b);
        g = mpmx_ctx::SubstepGraph{};
    }
}

void note_cg_iters(mpmx_ctx* c, int iters) {
    for (int q = 7; q > 0; --q) c->cg_recent[q] = c->cg_recent[q - 1];
    c->cg_recent[0] = iters;
}

void timer_push(mpmx_ctx* c, const char* name, bool begin) {
    if (!c->timing) return;
    TimerEntry* t = nullptr;
    for (auto& e : c->timers)
        if (e.name == name) t = &e;
    if (!t) {
        c->timers.push_back(TimerEntry{name, {}});
        t = &c->timers.back();
    }
    if (begin) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
        (void)hipEventRecord(a, c->stream);
        t->ev.push_back({a, b});
    } else if (!t->ev.empty()) {
        (void)hipEventRecord(t->ev.back().second, c->stream);
    }
}

}  // namespace mpmx

// ---- helpers ----------------------------------------------------------------
static mpmx_status fail(mpmx_ctx* c, mpmx_status s, const std::string& msg) {
    if (c) c->last_error = msg;
    return s;
}
static mpmx_status hipfail(mpmx_ctx* c, hipError_t e, const char* where) {
    return fail(c, MPMX_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}
#define HIPCHK(c, expr)                                     \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return hipfail(c, _e, #expr); \
    } while (0)

template <typename T>
static hipError_t dalloc(T** p, int64_t count) {
    *p = nullptr;
    if (count <= 0) count = 1;
    return hipMalloc((void**)p, sizeof(T) * (size_t)count);
}
template <typename T>
static void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

static int ceil_log2(int64_t v) {
    int b = 0;
    while ((1ll << b) < v) ++b;
    return b;
}

// C^-1 by Gauss-Jordan with full pivoting, as Inverse (ExplicitTimeIntegrator.cu:71-129)
static void gauss_jordan3(const double in[9], double mat[9]) {
    for (int i = 0; i < 9; ++i) mat[i] = in[i];
    int swapR[3], swapC[3], pivot[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        int pr = 0, pc = 0;
        double maxValue = 0;
        for (int j = 0; j < 3; ++j)
            if (pivot[j] != 1)
                for (int k = 0; k < 3; ++k)
                    if (pivot[k] == 0 && mat[j * 3 + k] > maxValue) { maxValue = mat[j * 3 + k]; pr = j; pc = k; }
        if (pr != pc)
            for (int j = 0; j < 3; ++j) std::swap(mat[pr * 3 + j], mat[pc * 3 + j]);
        swapC[i] = pc;
        swapR[i] = pr;
        ++pivot[i];
        const double inv = 1.0 / mat[pc * 3 + pc];
        mat[pc * 3 + pc] = 1;
        for (int j = 0; j < 3; ++j) mat[pc * 3 + j] *= inv;
        for (int j = 0; j < 3; ++j)
            if (j != pc) {
                const double ratio = mat[j * 3 + pc];
                mat[j * 3 + pc] = 0;
                for (int k = 0; k < 3; ++k) mat[j * 3 + k] -= mat[pc * 3 + k] * ratio;
            }
    }
    for (int i = 0; i < 3; ++i)
        if (swapR[i] != swapC[i])
            for (int j = 0; j < 3; ++j) std::swap(mat[j * 3 + swapC[i]], mat[j * 3 + swapR[i]]);
}

// the constants of one scene (the batch fields are set by set_scene_geometry)
static DevParams scene_dev_params(const mpmx_params& P, const mpmx_config& cfg) {
    DevParams d;
    std::memset(&d, 0, sizeof(d));
    d.N = P.N;
    d.transfer = P.transfer;
    // the fixed corotated energy has no fracture: integrate skips the solve (Energy_Model 0)
    d.phase_field = (P.phase_field && P.model != 0) ? 1 : 0;
    d.model = P.model;
    d.dot_mode = P.dot_mode;
    d.wall = P.wall_cells;
    d.grip_lo = P.grip_lo;
    d.grip_hi = P.grip_hi;
    d.vpage = P.vpage;
    d.dx = (float)(1.0 / P.N);
    d.inv_dx = (float)P.N;
    d.Dinv = (float)(4.0 * P.N * P.N);
    d.grip_v = (float)P.grip_v;
    // Materials.cpp:4-14
    const double lambda = P.E * P.nu / ((1 + P.nu) * (1 - 2 * P.nu));
    const double mu = P.E / (2 * (1 + P.nu));
    const double kappa = P.E / (3 * (1 - 2 * P.nu));
    d.lambda = (float)lambda;
    d.mu = (float)mu;
    d.kappa = (float)kappa;
    d.V0 = (float)P.V0;
    const double C[9] = {2 * mu + lambda, lambda, lambda, lambda, 2 * mu + lambda, lambda,
                         lambda, lambda, 2 * mu + lambda};
    double Ci[9];
    gauss_jordan3(C, Ci);
    for (int k = 0; k < 9; ++k) d.Cinv[k] = (float)Ci[k];
    d.yield = (float)P.yield_stress;
    d.hard = (float)P.hardening;
    d.alpha_c = (float)P.alpha_c;
    d.dlambda_cap = (float)P.dlambda_cap;
    d.eps_cap = (float)P.eps_cap;
    d.Gc0 = (float)P.Gc0;
    d.residual_c = (float)P.residual_c;
    d.ps = (float)P.ps;
    d.theta = (float)P.theta;
    d.M = (float)P.M;
    for (int a = 0; a < 3; ++a) d.gravity[a] = (float)P.gravity[a];
    d.vmax = (float)P.vmax;
    d.cg_tol = (float)P.cg_tol;
    d.crit_stretch = (float)P.crit_stretch;
    d.L0 = (float)(0.5 / P.N);
    d.sq23f = sqrtf(2.0f / 3.0f);
    d.svd_gap = (float)cfg.svd_gap;
    return d;
}

static int64_t morton3_host(int64_t x, int64_t y, int64_t z) {
    int64_t k = 0;
    for (int b = 0; b < 20; ++b)
        k |= (((z >> b) & 1) << (3 * b)) | (((y >> b) & 1) << (3 * b + 1)) | (((x >> b) & 1) << (3 * b + 2));
    return k;
}
static void unmorton3_host(int64_t k, int64_t& x, int64_t& y, int64_t& z) {
    x = y = z = 0;
    for (int b = 0; b < 20; ++b) {
        z |= ((k >> (3 * b)) & 1) << b;
        y |= ((k >> (3 * b + 1)) & 1) << b;
        x |= ((k >> (3 * b + 2)) & 1) << b;
    }
}

// page space of the context: one scene spans the Morton cube of 2^axis_bits
// pages per axis that holds its pages and their +1 neighbours; a batch gives
// scene s slot s, a cube of 2^slot_log2 pages per axis (the largest scene's
// pages), at slot coordinates unmorton(s)
static void set_scene_geometry(mpmx_ctx* c) {
    const int S = c->num_scenes;
    if (S == 1) {
        c->axis_bits = ceil_log2(c->scene_P[0].N / 4 + 1);
        c->slot_log2 = c->axis_bits;
    } else {
        int b = 2;
        for (const auto& P : c->scene_P) b = std::max(b, ceil_log2(P.N / 4));
        c->slot_log2 = b;
        c->axis_bits = b + (ceil_log2(S) + 2) / 3;
    }
    const int shift = 3 * c->slot_log2;
    c->mask_words = std::max<int64_t>(1, ((int64_t)S << shift) / 64);
    std::vector<DevParams> sd((size_t)S);
    bool any_pf = false;
    for (int k = 0; k < S; ++k) {
        DevParams d = scene_dev_params(c->scene_P[k], c->cfg);
        d.num_scenes = S;
        d.slot_shift = shift;
        int64_t ox, oy, oz;
        unmorton3_host(k, ox, oy, oz);
        const int64_t cells = (int64_t)4 << c->slot_log2;
        d.origin[0] = (int)(ox * cells);
        d.origin[1] = (int)(oy * cells);
        d.origin[2] = (int)(oz * cells);
        d.page_limit = std::min(1 << ceil_log2(c->scene_P[k].N / 4 + 1), 1 << c->slot_log2);
        any_pf = any_pf || d.phase_field;
        sd[(size_t)k] = d;
    }
    for (auto& d : sd) d.scenes = c->scene_dp;
    c->dp = sd[0];
    c->dp.phase_field = any_pf ? 1 : 0;   // batch-wide: some scene runs the solve
    (void)hipMemcpy(c->scene_dp, sd.data(), sizeof(DevParams) * S, hipMemcpyHostToDevice);
}

static mpmx_status read_counts(mpmx_ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->hc, c->dc, sizeof(DevCounts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MPMX_OK;
}
static mpmx_status check_err(mpmx_ctx* c) {
    mpmx_status s = read_counts(c);
    if (s != MPMX_OK) return s;
    const unsigned e = c->hc->err;
    if (e & kErrCapacity) return fail(c, MPMX_ERR_CAPACITY, "page table capacity exceeded");
    if (e & kErrNonFinite) return fail(c, MPMX_ERR_NONFINITE, "non-finite particle state");
    if (e & kErrDomain) return fail(c, MPMX_ERR_OUT_OF_DOMAIN, "a particle left [2dx, 1-2dx]");
    if (e & kErrBreakdown) return fail(c, MPMX_ERR_CG_BREAKDOWN, "phase-field PCG breakdown");
    return MPMX_OK;
}

static void free_particles(mpmx_ctx* c) {
    destroy_graphs(c);
    c->old_keys_sorted = false;
    c->tables_match_keys = false;
    for (int b = 0; b < 2; ++b) {
        PState& s = c->ps[b];
        dfree(s.x); dfree(s.v); dfree(s.B); dfree(s.F); dfree(s.c); dfree(s.alpha); dfree(s.psi); dfree(s.g);
        dfree(s.mass); dfree(s.id);
        MState& m = c->ms[b];
        dfree(m.x); dfree(m.v); dfree(m.B); dfree(m.F); dfree(m.F32); dfree(m.c); dfree(m.alpha); dfree(m.psi); dfree(m.g);
        dfree(m.mass); dfree(m.id);
    }
    dfree(c->sigma); dfree(c->FP); dfree(c->vol); dfree(c->cg_col); dfree(c->payload); dfree(c->pbound); dfree(c->cbound);
    dfree(c->keys); dfree(c->keys_alt); dfree(c->keys_next); dfree(c->vals); dfree(c->vals_alt); dfree(c->hist);
    dfree(c->mv_tcount); dfree(c->mv_toff); dfree(c->mv_old); dfree(c->mv_slot);
    dfree(c->mv_tbound); dfree(c->mv_cpre); dfree(c->mv_cmask);
    for (int b = 0; b < 2; ++b) { dfree(c->mv_key[b]); dfree(c->mv_val[b]); }
    dfree(c->pmask); dfree(c->nmask); dfree(c->pflag); dfree(c->wcount); dfree(c->scount);
    dfree(c->page_key); dfree(c->page2particle); dfree(c->adj); dfree(c->nvp); dfree(c->vpo); dfree(c->target); dfree(c->block_key); dfree(c->bcount); dfree(c->cell_start);
    if (c->block_lists) (void)hipFree(c->block_lists);
    c->block_lists = nullptr;
    if (c->block_g2p) (void)hipFree(c->block_g2p);
    c->block_g2p = nullptr;
    dfree(c->block_order);
    dfree(c->block_wclass);
    c->order_ok = false;
    dfree(c->grid); dfree(c->grid_p2g);
    dfree(c->cg_x); dfree(c->cg_r); dfree(c->cg_z); dfree(c->cg_s); dfree(c->cg_q); dfree(c->cg_part); dfree(c->cg_amask);
    dfree(c->scan_tmp);
    if (c->p2g_part) (void)hipFree(c->p2g_part);
    c->p2g_part = nullptr;
    c->cap_part = 0;
}

static mpmx_status alloc_grid(mpmx_ctx* c, int64_t cap_tp) {
    destroy_graphs(c);
    dfree(c->grid); dfree(c->grid_p2g);
    dfree(c->cg_x); dfree(c->cg_r); dfree(c->cg_z); dfree(c->cg_s); dfree(c->cg_q); dfree(c->cg_part); dfree(c->cg_amask);
    c->cap_tp = cap_tp;
    // workgroups of the node-space kernels: about four 128-node blocks each at
    // capacity, at most 4 per CU for one scene and 8 for a batch, a multiple
    // of 8 (one share per XCD).  Every workgroup of a scene reduces all of
    // the scene's partials, so one scene keeps fewer; a batch splits them
    // among its scenes, whose small solves are latency-bound (on the
    // 256-scene batch 8 per CU take the vector kernels from 0.32 to 0.24 ms
    // per substep, on one 8M scene from 0.034 to 0.039 ms).
    int64_t nb = ((cap_tp + 1) / 2 + 3) / 4;
    nb = std::min<int64_t>(nb, (int64_t)c->num_cu * (c->num_scenes > 1 ? 8 : 4));
    nb = std::max<int64_t>(nb, c->num_scenes);   // at least one per scene
    c->cg_nb = (int)std::max<int64_t>(8, (nb + 7) / 8 * 8);
    HIPCHK(c, dalloc(&c->cg_part, 4 * (int64_t)c->cg_nb));
    // PCG workgroups per scene: one each, the rest in proportion to the
    // scenes' particle counts (largest remainders); a lone scene gets all
    {
        const int S = c->num_scenes, W = c->cg_nb;
        std::vector<int> nwg((size_t)S, 1);
        const int64_t ntot = std::accumulate(c->scene_n.begin(), c->scene_n.end(), (int64_t)0);
        const int extra = W - S;
        std::vector<std::pair<double, int>> rem;
        int given = 0;
        for (int k = 0; k < S; ++k) {
            const double share = ntot > 0 ? (double)extra * (double)c->scene_n[(size_t)k] / (double)ntot : 0.0;
            const int whole = (int)share;
            nwg[(size_t)k] += whole;
            given += whole;
            rem.push_back({share - whole, k});
        }
        std::sort(rem.begin(), rem.end(), [](const auto& a, const auto& b) {
            return a.first != b.first ? a.first > b.first : a.second < b.second;
        });
        for (int q = 0; q < extra - given; ++q) nwg[(size_t)rem[(size_t)q % S].second] += 1;
        std::vector<SceneCG> h((size_t)S);
        std::vector<int> wg_scene((size_t)W, 0);
        std::memset(h.data(), 0, sizeof(SceneCG) * h.size());
        int w0 = 0;
        for (int k = 0; k < S; ++k) {
            h[(size_t)k].wg0 = w0;
            h[(size_t)k].nwg = nwg[(size_t)k];
            for (int q = 0; q < nwg[(size_t)k]; ++q) wg_scene[(size_t)(w0 + q)] = k;
            w0 += nwg[(size_t)k];
        }
        dfree(c->wg_scene);
        HIPCHK(c, dalloc(&c->wg_scene, W));
        HIPCHK(c, hipMemcpy(c->wg_scene, wg_scene.data(), sizeof(int) * W, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->scg, h.data(), sizeof(SceneCG) * S, hipMemcpyHostToDevice));
        HIPCHK(c, launch_scene_counts(c));   // page ranges of the current tables
    }
    HIPCHK(c, dalloc(&c->grid, cap_tp * kPageFloats));
    if (c->keep_p2g) HIPCHK(c, dalloc(&c->grid_p2g, cap_tp * kPageFloats));
    HIPCHK(c, dalloc(&c->cg_x, cap_tp * 64));
    HIPCHK(c, dalloc(&c->cg_r, cap_tp * 64));
    HIPCHK(c, dalloc(&c->cg_z, cap_tp * 64));
    HIPCHK(c, dalloc(&c->cg_s, cap_tp * 64));
    HIPCHK(c, dalloc(&c->cg_q, cap_tp * 64));
    HIPCHK(c, dalloc(&c->cg_amask, cap_tp));
    HIPCHK(c, fill_zero(c, c->grid, cap_tp * kPageFloats * 4));
    HIPCHK(c, fill_zero(c, c->cg_x, cap_tp * 64 * 4));
    if (c->cfg.p2g_variant == 3) {
        // owner-block P2G records (50 KiB per block): four times the blocks of
        // the first table build, as the grid pages are sized from its pages
        if (c->p2g_part) (void)hipFree(c->p2g_part);
        c->p2g_part = nullptr;
        c->cap_part = std::min<int64_t>(c->tp_bound, std::max<int64_t>(4 * (int64_t)c->hc->num_blocks, 1024));
        HIPCHK(c, hipMalloc(&c->p2g_part, (size_t)c->cap_part * (size_t)kOwnRecBytes));
    }
    return MPMX_OK;
}

static mpmx_status upload_particles(mpmx_ctx* c, const int64_t* counts, const std::vector<int>& start, int64_t n,
                                    const double* x, const double* v, const double* mass);
static hipError_t noalloc_scratch(int device, float** out);

// ---- public API --------------------------------------------------------------
extern "C" {

void mpmx_default_params(mpmx_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->N = 128;
    p->transfer = 0;
    p->phase_field = 1;
    p->dot_mode = 0;
    p->cg_max_iter = 100;
    p->vpage = 256;
    p->wall_cells = 4;
    p->grip_lo = -1;
    p->grip_hi = 1 << 30;
    p->grip_v = 0;
    p->E = 1e5;
    p->nu = 0.3;
    p->V0 = 1.0 / (128.0 * 128.0 * 128.0 * 8.0);
    p->yield_stress = 800;
    p->hardening = 3;
    p->alpha_c = 1;
    p->dlambda_cap = 100;
    p->eps_cap = 3;
    p->Gc0 = 3;
    p->residual_c = 0.1;
    p->ps = 0.01;
    p->theta = (double)0.0001f;
    p->M = 50;
    p->gravity[0] = 0;
    p->gravity[1] = -10;
    p->gravity[2] = 0;
    p->vmax = (double)11.137f;
    p->cg_tol = (double)1e-6f;
    p->crit_stretch = 0;
    p->model = 2;
}

void mpmx_default_config(mpmx_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->precision = MPMX_FP32;
    c->svd_gap = 1e-4;
    c->cg_check_every = 1;
    c->stress_fp64 = 1;
    c->use_graph = 0;
}

const char* mpmx_status_string(mpmx_status s) {
    switch (s) {
    case MPMX_OK: return "ok";
    case MPMX_ERR_HIP: return "HIP runtime error";
    case MPMX_ERR_ARG: return "invalid argument";
    case MPMX_ERR_CAPACITY: return "capacity exceeded";
    case MPMX_ERR_OUT_OF_DOMAIN: return "particle out of domain";
    case MPMX_ERR_NONFINITE: return "non-finite state";
    case MPMX_ERR_CG_BREAKDOWN: return "PCG breakdown";
    case MPMX_ERR_NO_DEVICE: return "no HIP device";
    }
    return "unknown";
}

const char* mpmx_last_error(const mpmx_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

int mpmx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

mpmx_status mpmx_create(const mpmx_params* p, const mpmx_config* cfg, mpmx_ctx** out) {
    return mpmx_create_batch(1, p, cfg, out);
}

mpmx_status mpmx_create_batch(int num_scenes, const mpmx_params* ps, const mpmx_config* cfg, mpmx_ctx** out) {
    if (!ps || !out || num_scenes < 1) return MPMX_ERR_ARG;
    *out = nullptr;
    int slot_log2 = 2;
    for (int k = 0; k < num_scenes; ++k) {
        const mpmx_params* p = ps + k;
        if (p->N < 16 || p->N > 1024 || (p->N & 3)) return MPMX_ERR_ARG;
        if (p->vpage != 64 && p->vpage != 128 && p->vpage != 256) return MPMX_ERR_ARG;
        if (p->transfer < 0 || p->transfer > 1) return MPMX_ERR_ARG;
        if (p->model < 0 || p->model > 2) return MPMX_ERR_ARG;
        // batch-wide settings
        if (p->transfer != ps[0].transfer || p->vpage != ps[0].vpage || p->dot_mode != ps[0].dot_mode ||
            p->cg_max_iter != ps[0].cg_max_iter)
            return MPMX_ERR_ARG;
        slot_log2 = std::max(slot_log2, ceil_log2(p->N / 4));
    }
    // keys: 6 cell bits + the slot's page bits + the slot index
    if (num_scenes > 1 && 6 + 3 * slot_log2 + ceil_log2(num_scenes) > 32) return MPMX_ERR_ARG;
    if (cfg && (cfg->p2g_variant < 0 || cfg->p2g_variant > 3)) return MPMX_ERR_ARG;
    if (cfg && (cfg->sort_variant < 0 || cfg->sort_variant > 2)) return MPMX_ERR_ARG;
    if (num_scenes > 1 && cfg && cfg->p2g_variant == 1) return MPMX_ERR_ARG;
    if (mpmx_device_count() <= 0) return MPMX_ERR_NO_DEVICE;
    mpmx_ctx* c = new mpmx_ctx();
    c->P = ps[0];
    c->num_scenes = num_scenes;
    c->scene_P.assign(ps, ps + num_scenes);
    if (cfg) c->cfg = *cfg;
    else mpmx_default_config(&c->cfg);
    c->device = c->cfg.device;
    if (hipSetDevice(c->device) != hipSuccess) {
        delete c;
        return MPMX_ERR_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
        c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return MPMX_ERR_HIP;
    }
    if (hipMalloc((void**)&c->dc, sizeof(DevCounts)) != hipSuccess ||
        hipHostMalloc((void**)&c->hc, sizeof(DevCounts), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&c->pub, sizeof(DevCounts), hipHostMallocCoherent) != hipSuccess) {
        if (c->dc) (void)hipFree(c->dc);
        if (c->hc) (void)hipHostFree(c->hc);
        delete c;
        return MPMX_ERR_HIP;
    }
    std::memset(c->hc, 0, sizeof(DevCounts));
    std::memset(c->pub, 0, sizeof(DevCounts));
    (void)hipMemcpy(c->dc, c->hc, sizeof(DevCounts), hipMemcpyHostToDevice);
    if (hipMalloc((void**)&c->scene_dp, sizeof(DevParams) * num_scenes) != hipSuccess ||
        hipMalloc((void**)&c->scg, sizeof(SceneCG) * num_scenes) != hipSuccess ||
        hipMalloc((void**)&c->scene_start, sizeof(int) * (num_scenes + 1)) != hipSuccess ||
        hipMalloc((void**)&c->red_part, sizeof(double) * kRedDoubles) != hipSuccess) {
        mpmx_destroy(c);
        return MPMX_ERR_HIP;
    }
    set_scene_geometry(c);
    {   // the device address of the allocation-free CFL step's scratch, once
        // per device (a failure here is reported by that entry point itself)
        float* s = nullptr;
        (void)noalloc_scratch(c->device, &s);
    }
    *out = c;
    return MPMX_OK;
}

mpmx_status mpmx_destroy(mpmx_ctx* c) {
    if (!c) return MPMX_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& t : c->timers)
        for (auto& e : t.ev) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    free_particles(c);
    dfree(c->scene_dp);
    dfree(c->scg);
    dfree(c->scene_start);
    dfree(c->red_part);
    dfree(c->wg_scene);
    dfree(c->wave_buf);
    if (c->dc) (void)hipFree(c->dc);
    if (c->hc) (void)hipHostFree(c->hc);
    if (c->pub) (void)hipHostFree(c->pub);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MPMX_OK;
}

mpmx_status mpmx_set_particles(mpmx_ctx* c, int64_t n, const double* x, const double* v, const double* mass) {
    if (c && c->num_scenes != 1) return fail(c, MPMX_ERR_ARG, "batch context: use mpmx_set_particles_batch");
    return mpmx_set_particles_batch(c, &n, x, v, mass);
}

mpmx_status mpmx_set_particles_batch(mpmx_ctx* c, const int64_t* counts, const double* x, const double* v,
                                     const double* mass) {
    if (!c || !counts || !x || !v || !mass) return fail(c, MPMX_ERR_ARG, "bad particle arrays");
    int64_t n = 0;
    std::vector<int> start((size_t)c->num_scenes + 1, 0);
    for (int k = 0; k < c->num_scenes; ++k) {
        if (counts[k] <= 0) return fail(c, MPMX_ERR_ARG, "every scene needs particles");
        n += counts[k];
        if (n > (1ll << 30)) return fail(c, MPMX_ERR_ARG, "too many particles");
        start[(size_t)k + 1] = (int)n;
    }
    HIPCHK(c, hipSetDevice(c->device));
    for (int k = 0; k < c->num_scenes; ++k) {
        // the float test the kernels apply to the stored (float) positions
        const float dxf = (float)(1.0 / c->scene_P[(size_t)k].N);
        const float lo = 2.0f * dxf, hi = 1.0f - 2.0f * dxf;
        for (int64_t i = 3 * (int64_t)start[(size_t)k]; i < 3 * (int64_t)start[(size_t)k + 1]; ++i) {
            const float xf = (float)x[i];
            if (!(xf >= lo && xf <= hi)) return fail(c, MPMX_ERR_OUT_OF_DOMAIN, "particle outside [2dx, 1-2dx]");
        }
    }
    // from here on the old state is gone: a failure leaves an empty context
    // (n = 0, every buffer freed), which the stage calls refuse
    free_particles(c);
    mpmx_status s = upload_particles(c, counts, start, n, x, v, mass);
    if (s != MPMX_OK) {
        const std::string msg = c->last_error;
        (void)hipStreamSynchronize(c->stream);
        free_particles(c);
        dfree(c->wave_buf);
        c->wmap = WaveMap{};
        c->n = 0;
        c->scene_n.assign((size_t)c->num_scenes, 0);
        c->sorted = false;
        c->last_error = msg;
    }
    return s;
}

}  // extern "C"

// buffers, upload, first sort and tables of mpmx_set_particles_batch
static mpmx_status upload_particles(mpmx_ctx* c, const int64_t* counts, const std::vector<int>& start, int64_t n,
                                    const double* x, const double* v, const double* mass) {
    c->scene_n.assign(counts, counts + c->num_scenes);
    HIPCHK(c, hipMemcpy(c->scene_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
    // waves of the per-particle kernels: each scene's range (contiguous in
    // every sorted order) padded to whole waves
    dfree(c->wave_buf);
    c->wmap = WaveMap{};
    if (c->num_scenes > 1) {
        std::vector<int> first, count, scene;
        for (int k = 0; k < c->num_scenes; ++k)
            for (int64_t q = 0; q < counts[k]; q += 64) {
                first.push_back(start[(size_t)k] + (int)q);
                count.push_back((int)std::min<int64_t>(64, counts[k] - q));
                scene.push_back(k);
            }
        const int64_t W = (int64_t)first.size();
        HIPCHK(c, dalloc(&c->wave_buf, 3 * W));
        HIPCHK(c, hipMemcpy(c->wave_buf, first.data(), sizeof(int) * W, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->wave_buf + W, count.data(), sizeof(int) * W, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->wave_buf + 2 * W, scene.data(), sizeof(int) * W, hipMemcpyHostToDevice));
        c->wmap = WaveMap{c->wave_buf, c->wave_buf + W, c->wave_buf + 2 * W, W};
    }
    c->n = n;
    c->cap_n = std::max<int64_t>(n, c->cfg.max_particles);
    const bool mixed = c->cfg.precision != MPMX_FP32;
    for (int b = 0; b < 2; ++b) {
        if (mixed) {
            MState& s = c->ms[b];
            HIPCHK(c, dalloc(&s.x, 3 * n));
            HIPCHK(c, dalloc(&s.v, 3 * n));
            HIPCHK(c, dalloc(&s.B, 9 * n));
            if (c->cfg.precision == MPMX_MIXED_F16) HIPCHK(c, dalloc(&s.F, 9 * n));
            else HIPCHK(c, dalloc(&s.F32, 9 * n));
            HIPCHK(c, dalloc(&s.c, n));
            HIPCHK(c, dalloc(&s.alpha, n));
            HIPCHK(c, dalloc(&s.psi, n));
            HIPCHK(c, dalloc(&s.g, n));
            HIPCHK(c, dalloc(&s.mass, n));
            HIPCHK(c, dalloc(&s.id, n));
        } else {
            PState& s = c->ps[b];
            HIPCHK(c, dalloc(&s.x, 3 * n));
            HIPCHK(c, dalloc(&s.v, 3 * n));
            HIPCHK(c, dalloc(&s.B, 9 * n));
            HIPCHK(c, dalloc(&s.F, 9 * n));
            HIPCHK(c, dalloc(&s.c, n));
            HIPCHK(c, dalloc(&s.alpha, n));
            HIPCHK(c, dalloc(&s.psi, n));
            HIPCHK(c, dalloc(&s.g, n));
            HIPCHK(c, dalloc(&s.mass, n));
            HIPCHK(c, dalloc(&s.id, n));
        }
    }
    HIPCHK(c, dalloc(&c->sigma, 9 * n));
    HIPCHK(c, dalloc(&c->FP, n));
    HIPCHK(c, dalloc(&c->vol, n));
    HIPCHK(c, dalloc(&c->cg_col, 3 * n));
    HIPCHK(c, dalloc(&c->payload, 28 * n));
    HIPCHK(c, dalloc(&c->keys, n));
    HIPCHK(c, dalloc(&c->keys_alt, n));
    HIPCHK(c, dalloc(&c->keys_next, n));
    HIPCHK(c, dalloc(&c->vals, n));
    HIPCHK(c, dalloc(&c->vals_alt, n));
    c->sort_tiles = (int)((n + 2047) / 2048);
    HIPCHK(c, dalloc(&c->hist, (int64_t)256 * c->sort_tiles));
    HIPCHK(c, dalloc(&c->mv_tcount, c->sort_tiles));
    HIPCHK(c, dalloc(&c->mv_toff, c->sort_tiles));
    HIPCHK(c, dalloc(&c->mv_old, n));
    HIPCHK(c, dalloc(&c->mv_slot, n));
    HIPCHK(c, dalloc(&c->mv_tbound, c->sort_tiles + 1));
    HIPCHK(c, dalloc(&c->mv_cpre, n / 64 + 1));
    HIPCHK(c, dalloc(&c->mv_cmask, n / 64 + 1));
    for (int b = 0; b < 2; ++b) {
        HIPCHK(c, dalloc(&c->mv_key[b], n));
        HIPCHK(c, dalloc(&c->mv_val[b], n));
    }
    c->sort_by_movers = c->cfg.sort_variant != 1;
    // keys: cell bits, the slot's page bits and (batch) the slot index
    c->key_bits = c->num_scenes == 1 ? 6 + 3 * std::max(1, ceil_log2(c->P.N / 4))
                                     : 6 + 3 * c->slot_log2 + ceil_log2(c->num_scenes);
    // pages (mask_words from the scene geometry)
    HIPCHK(c, dalloc(&c->pmask, c->mask_words));
    HIPCHK(c, dalloc(&c->nmask, c->mask_words));
    HIPCHK(c, dalloc(&c->pflag, 8 * c->mask_words));
    HIPCHK(c, fill_zero(c, c->pflag, 64 * c->mask_words));
    HIPCHK(c, dalloc(&c->wcount, c->mask_words));
    HIPCHK(c, dalloc(&c->scount, c->mask_words));
    int64_t pages_domain = 0, pages_nb = 0;
    for (const auto& P : c->scene_P) {
        const int64_t q = P.N / 4;
        pages_domain += q * q * q;
        pages_nb += (q + 1) * (q + 1) * (q + 1);
    }
    c->cap_p = std::min<int64_t>(n, pages_domain);
    const int64_t tp_bound = std::min<int64_t>(8 * c->cap_p, pages_nb);
    c->cap_vp = n / c->P.vpage + c->cap_p + 1;
    HIPCHK(c, dalloc(&c->page_key, tp_bound));
    c->tp_bound = tp_bound;
    HIPCHK(c, dalloc(&c->block_key, tp_bound));
    HIPCHK(c, dalloc(&c->pbound, 8 * (c->cap_p + 1)));
    HIPCHK(c, dalloc(&c->cbound, c->cap_p + 1));
    HIPCHK(c, dalloc(&c->bcount, c->mask_words));
    HIPCHK(c, dalloc(&c->page2particle, c->cap_p + 1));
    HIPCHK(c, dalloc(&c->cell_start, 65 * c->cap_p));
    HIPCHK(c, hipMalloc(&c->block_lists, sizeof(BlockClasses) * (size_t)tp_bound));
    HIPCHK(c, hipMalloc(&c->block_g2p, sizeof(BlockG2P) * (size_t)tp_bound));
    HIPCHK(c, dalloc(&c->block_order, tp_bound));
    HIPCHK(c, dalloc(&c->block_wclass, tp_bound));
    HIPCHK(c, dalloc(&c->adj, 7 * c->cap_p));
    HIPCHK(c, dalloc(&c->nvp, c->cap_p + 1));
    HIPCHK(c, dalloc(&c->vpo, c->cap_p + 1));
    HIPCHK(c, dalloc(&c->target, c->cap_vp));
    int64_t tiles = (256ll * c->sort_tiles + 1023) / 1024;
    tiles = std::max(tiles, 3 * ((c->mask_words + 1023) / 1024));   // the page tables' three word scans
    tiles = std::max(tiles, (c->cap_p + 1023) / 1024);
    tiles = std::max(tiles, (n + 1023) / 1024);
    c->scan_tmp_cap = tiles + 1;
    HIPCHK(c, dalloc(&c->scan_tmp, c->scan_tmp_cap));
    // upload
    double *dx_ = nullptr, *dv_ = nullptr, *dm_ = nullptr;
    HIPCHK(c, dalloc(&dx_, 3 * n));
    HIPCHK(c, dalloc(&dv_, 3 * n));
    HIPCHK(c, dalloc(&dm_, n));
    HIPCHK(c, hipMemcpyAsync(dx_, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dv_, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dm_, mass, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    c->cur = 0;
    if (mixed) k_upload_mixed<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(c->ms[0], dx_, dv_, dm_, (int)n);
    else k_upload_fp32<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(c->ps[0], dx_, dv_, dm_, (int)n);
    HIPCHK(c, hipGetLastError());
    std::memset(c->hc, 0, sizeof(DevCounts));
    HIPCHK(c, hipMemcpyAsync(c->dc, c->hc, sizeof(DevCounts), hipMemcpyHostToDevice, c->stream));
    // keys, first canonical sort, first tables (with the bound as capacity)
    HIPCHK(c, compute_keys(c));
    HIPCHK(c, sort_and_gather(c));
    c->cap_tp = tp_bound;
    HIPCHK(c, build_tables(c));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(dx_); dfree(dv_); dfree(dm_);
    mpmx_status s = check_err(c);
    if (s != MPMX_OK) return s;
    int64_t cap = c->cfg.max_pages;
    if (cap <= 0) cap = std::min<int64_t>(tp_bound, std::max<int64_t>(4 * (int64_t)c->hc->num_total, 4096));
    cap = std::min(cap, tp_bound);
    s = alloc_grid(c, cap);
    if (s != MPMX_OK) return s;
    c->sorted = true;
    c->substeps = 0;
    c->stress_fields = false;
    return MPMX_OK;
}

extern "C" {

mpmx_status mpmx_bin_sort(mpmx_ctx* c) {
    if (!c || c->n == 0) return fail(c, MPMX_ERR_ARG, "no particles");
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->sorted) {
        MPMX_TIMED(c, "sort", HIPCHK(c, sort_and_gather(c)));
        c->sorted = true;
    }
    return MPMX_OK;
}

mpmx_status mpmx_rebuild(mpmx_ctx* c) {
    if (!c || c->n == 0) return fail(c, MPMX_ERR_ARG, "no particles");
    MPMX_TIMED(c, "tables", HIPCHK(c, build_tables(c)));
    return MPMX_OK;
}

mpmx_status mpmx_stress(mpmx_ctx* c) {
    if (!c || c->n == 0) return fail(c, MPMX_ERR_ARG, "no particles");
    MPMX_TIMED(c, "stress", HIPCHK(c, launch_stress(c)));
    return MPMX_OK;
}

}  // extern "C"

// The stages with their time step: dt > 0 as given, dt <= 0 the step the
// device-side CFL left in the counters (mpmx_run_cfl).  The public entry
// points accept dt > 0 only.
static mpmx_status p2g_dt(mpmx_ctx* c, double dt) {
    MPMX_TIMED(c, "p2g", HIPCHK(c, launch_p2g(c, (float)dt)));
    return MPMX_OK;
}
// stress and P2G of a substep; with the gather P2G one pass computes the
// stress and the P2G payload
static mpmx_status stress_and_p2g(mpmx_ctx* c, double dt) {
    if (c->cfg.p2g_variant == 1) {
        mpmx_status s = mpmx_stress(c);
        return s == MPMX_OK ? p2g_dt(c, dt) : s;
    }
    HIPCHK(c, launch_stress_p2g(c, (float)dt));
    return MPMX_OK;
}
static mpmx_status grid_update_dt(mpmx_ctx* c, double dt) {
    MPMX_TIMED(c, "grid_update", HIPCHK(c, launch_grid_update(c, (float)dt)));
    return MPMX_OK;
}
static mpmx_status g2p_dt(mpmx_ctx* c, double dt) {
    MPMX_TIMED(c, "g2p", HIPCHK(c, launch_g2p(c, (float)dt)));
    c->sorted = false;
    return MPMX_OK;
}
// the sort of a substep: ahead of the fused stress pass the state gather
// waits for that pass (sort_keys), which performs it on the way; the scatter
// P2G takes the state sorted
static mpmx_status substep_sort(mpmx_ctx* c) {
    if (c->sorted || c->cfg.p2g_variant == 1) return mpmx_bin_sort(c);
    MPMX_TIMED(c, "sort", HIPCHK(c, sort_keys(c)));
    c->sorted = true;
    return MPMX_OK;
}
static mpmx_status substep_dt(mpmx_ctx* c, double dt) {
    HIPCHK(c, hipSetDevice(c->device));
    mpmx_status s;
    if ((s = substep_sort(c)) != MPMX_OK) return s;
    s = mpmx_rebuild(c);
    c->g2p_follows = true;
    if (s == MPMX_OK) s = stress_and_p2g(c, dt);
    c->g2p_follows = false;
    if (c->gather_perm) HIPCHK(c, finish_gather(c));   // a stage before the stress pass failed
    if (s == MPMX_OK) s = grid_update_dt(c, dt);
    if (s == MPMX_OK) s = mpmx_phase_field(c, dt);
    if (s == MPMX_OK && c->dp.phase_field) note_cg_iters(c, c->cg_spec);
    if (s == MPMX_OK) s = g2p_dt(c, dt);
    if (c->vb_perm) HIPCHK(c, finish_carry_vb(c));   // a stage before the G2P failed
    if (s != MPMX_OK) return s;
    ++c->substeps;
    return MPMX_OK;
}
static mpmx_status check_ready(mpmx_ctx* c, double dt) {
    if (!c || c->n == 0) return fail(c, MPMX_ERR_ARG, "no particles");
    if (!(dt > 0) || !std::isfinite(dt)) return fail(c, MPMX_ERR_ARG, "dt must be positive and finite");
    return MPMX_OK;
}

extern "C" {

mpmx_status mpmx_p2g(mpmx_ctx* c, double dt) {
    mpmx_status s = check_ready(c, dt);
    return s == MPMX_OK ? p2g_dt(c, dt) : s;
}

mpmx_status mpmx_grid_update(mpmx_ctx* c, double dt) {
    mpmx_status s = check_ready(c, dt);
    return s == MPMX_OK ? grid_update_dt(c, dt) : s;
}

mpmx_status mpmx_phase_field(mpmx_ctx* c, double dt) {
    if (!c || c->n == 0) return fail(c, MPMX_ERR_ARG, "no particles");
    HIPCHK(c, launch_pcg(c, (float)dt));
    return MPMX_OK;
}

mpmx_status mpmx_g2p(mpmx_ctx* c, double dt) {
    mpmx_status s = check_ready(c, dt);
    return s == MPMX_OK ? g2p_dt(c, dt) : s;
}

mpmx_status mpmx_substep(mpmx_ctx* c, double dt) {
    mpmx_status s = check_ready(c, dt);
    return s == MPMX_OK ? substep_dt(c, dt) : s;
}

// ---- captured substeps ---------------------------------------------------
// Captures the substep that starts from the current parity (particles
// unsorted in buffer c->cur) as graph A (sort, tables, stress, P2G + grid
// update, `iters` PCG iterations, the counter copy to the host) and graph B
// (G2P).  The host-side state the stages update during capture is put back
// afterwards; replays set it themselves.
static mpmx_status capture_substep(mpmx_ctx* c, double dt, int iters) {
    const int p = c->cur;
    choose_sort_path(c);
    const bool by_movers = c->sort_by_movers && c->old_keys_sorted;   // the path sort_keys takes
    uint32_t* const keys0 = c->keys;
    uint32_t* const keys1 = c->keys_alt;
    auto& G = c->graphs[p];
    if (G.a) (void)hipGraphExecDestroy(G.a);
    if (G.b) (void)hipGraphExecDestroy(G.b);
    G = mpmx_ctx::SubstepGraph{};
    hipGraph_t ga = nullptr, gb = nullptr;
    mpmx_status s = MPMX_OK;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->g2p_follows = true;   // graph B holds the G2P
    if ((s = substep_sort(c)) == MPMX_OK && (s = mpmx_rebuild(c)) == MPMX_OK &&
        (s = stress_and_p2g(c, dt)) == MPMX_OK && (s = grid_update_dt(c, dt)) == MPMX_OK) {
        hipError_t e = c->dp.phase_field ? queue_pcg_iterations(c, 0, iters) : hipSuccess;
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->hc, c->dc, sizeof(DevCounts), hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) s = hipfail(c, e, "capture of the substep");
    }
    hipError_t ee = hipStreamEndCapture(c->stream, &ga);
    c->g2p_follows = false;
    if (s == MPMX_OK && ee != hipSuccess) s = hipfail(c, ee, "hipStreamEndCapture");
    if (s == MPMX_OK) {
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        s = g2p_dt(c, dt);
        ee = hipStreamEndCapture(c->stream, &gb);
        if (s == MPMX_OK && ee != hipSuccess) s = hipfail(c, ee, "hipStreamEndCapture");
    }
    if (s == MPMX_OK) {
        hipError_t e = hipGraphInstantiate(&G.a, ga, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipGraphInstantiate(&G.b, gb, nullptr, nullptr, 0);
        if (e != hipSuccess) s = hipfail(c, e, "hipGraphInstantiate");
    }
    if (ga) (void)hipGraphDestroy(ga);
    if (gb) (void)hipGraphDestroy(gb);
    c->cur = p;
    c->keys = keys0;
    c->keys_alt = keys1;
    c->sorted = false;
    c->gather_perm = nullptr;
    c->vb_perm = nullptr;
    if (s != MPMX_OK) {
        if (G.a) (void)hipGraphExecDestroy(G.a);
        if (G.b) (void)hipGraphExecDestroy(G.b);
        G = mpmx_ctx::SubstepGraph{};
        return s;
    }
    G.valid = true;
    G.dt = (float)dt;
    G.iters = iters;
    G.by_movers = by_movers;
    G.blocks = c->hc->num_blocks;
    return MPMX_OK;
}

// one substep as two graph replays and one host check of the PCG
static mpmx_status substep_graph(mpmx_ctx* c, double dt) {
    const int p = c->cur;
    int want = 1;
    for (int q : c->cg_recent) want = std::max(want, q);
    want = std::min(want, std::max(1, c->P.cg_max_iter));
    auto& G = c->graphs[p];
    choose_sort_path(c);
    const bool by_movers = c->sort_by_movers && c->old_keys_sorted;
    if (!G.valid || G.dt != (float)dt || G.iters != want || c->hc->num_blocks > G.blocks + G.blocks / 4 + 64 ||
        G.by_movers != by_movers) {
        mpmx_status s = capture_substep(c, dt, want);
        if (s != MPMX_OK) return s;
    }
    HIPCHK(c, hipGraphLaunch(G.a, c->stream));
    c->cur = 1 - p;
    std::swap(c->keys, c->keys_alt);   // what the captured sort did
    c->sorted = true;
    c->grid_fused = c->cfg.p2g_variant != 1;
    if (c->cfg.p2g_variant != 1) c->vb_perm = c->vals_alt;   // graph A's gather left v and B to graph B
    hipError_t e = hipStreamSynchronize(c->stream);          // graph A copied the counters to the host
    if (e == hipSuccess && c->dp.phase_field) {
        if (!c->hc->cg_done && G.iters < c->P.cg_max_iter)
            e = launch_pcg_from(c, G.iters, c->cfg.cg_check_every > 0 ? c->cfg.cg_check_every : 1);
        c->cg_spec = c->hc->cg_iters > 0 ? c->hc->cg_iters : 1;
        note_cg_iters(c, c->hc->cg_iters);
    }
    if (e != hipSuccess) {
        (void)finish_carry_vb(c);
        return hipfail(c, e, "substep between its graphs");
    }
    HIPCHK(c, hipGraphLaunch(G.b, c->stream));
    c->vb_perm = nullptr;
    c->sorted = false;
    ++c->substeps;
    return MPMX_OK;
}

static mpmx_status run_dt(mpmx_ctx* c, int k, double dt) {
    // the first substep after an upload or a sort, and timed runs, go eagerly
    (void)k;
    const bool graph = c->cfg.use_graph && !c->timing && !c->sorted;
    return graph ? substep_graph(c, dt) : substep_dt(c, dt);
}

mpmx_status mpmx_run(mpmx_ctx* c, int nsteps, double dt) {
    mpmx_status s = check_ready(c, dt);
    for (int k = 0; s == MPMX_OK && k < nsteps; ++k) s = run_dt(c, k, dt);
    return s;
}

mpmx_status mpmx_sync(mpmx_ctx* c) {
    if (!c) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return check_err(c);
}

}  // extern "C"

// workgroups of the particle reductions (one partial each)
static int reduction_groups(const mpmx_ctx* c) {
    const int64_t threads = c->wmap.first ? c->wmap.waves * 64 : c->n;
    return (int)std::min<int64_t>(kRedParts, std::max<int64_t>(1, (threads + 1023) / 1024));
}

// the CFL speed reduction (k_speed_part / k_speed_final) with its partials
// in `part` (kRedParts floats) and its two results in `out`
static hipError_t launch_speed(mpmx_ctx* c, int mode, double dt_default, double cfl, int set_dt, float* part,
                               double* out, int groups = 0) {
    const bool mixed = c->cfg.precision != MPMX_FP32;
    const int G = groups > 0 ? groups : reduction_groups(c);
    k_speed_part<<<G, 256, 0, c->stream>>>(c->dp, mixed ? nullptr : c->ps[c->cur].v,
                                           mixed ? c->ms[c->cur].v : nullptr,
                                           mixed ? c->ms[c->cur].mass : c->ps[c->cur].mass, c->wmap, (int)c->n,
                                           mode, part);
    k_speed_final<<<1, 256, 0, c->stream>>>(part, G, dt_default, cfl, set_dt, c->dc, out);
    return hipGetLastError();
}

// ---- the allocation-free CFL step (mpmx_compute_dt_noalloc) -----------------
// Its scratch is this file-scope device array: the speed reduction's partial
// maxima (one float per workgroup, at most kNoAllocParts) followed by the two
// FP64 results.  Each GPU holds its own copy for the life of the process;
// host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address) but the address hipGetSymbolAddress
// returns for the device, resolved once per device by the library's init
// call (mpmx_create_batch) or, failing that, by the first call on the device.
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU race; each call writes every partial it reads, so nothing is cleared.
constexpr int kNoAllocParts = 4096;
constexpr int kNoAllocPerThread = 32;   // particles per thread of the partial pass at the limit
constexpr int64_t kNoAllocMaxParticles = (int64_t)kNoAllocParts * 256 * kNoAllocPerThread;   // 33,554,432
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float scratch[kNoAllocParts + 4];
static float* g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// device address of `scratch` on `device` (the calling thread's current device)
static hipError_t noalloc_scratch(int device, float** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<float*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

extern "C" {

mpmx_status mpmx_max_velocity_sq(mpmx_ctx* c, double* out) {
    if (!c || !out || c->n == 0) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    double* res = c->red_part + kRedResult;
    HIPCHK(c, launch_speed(c, 0, 1.0, 1.0, 0, reinterpret_cast<float*>(c->red_part), res));
    double h[2];
    HIPCHK(c, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = h[0];
    return MPMX_OK;
}

mpmx_status mpmx_compute_dt(mpmx_ctx* c, double dt_default, double cfl, double* dt_out) {
    if (!c || !dt_out || c->n == 0 || !(dt_default > 0) || !(cfl > 0)) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    // Simulator.cpp:213 per scene: max over particles of |v| N_s
    double* res = c->red_part + kRedResult;
    HIPCHK(c, launch_speed(c, 2, dt_default, cfl, 1, reinterpret_cast<float*>(c->red_part), res));
    double h[2];
    HIPCHK(c, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *dt_out = h[1];
    return MPMX_OK;
}

mpmx_status mpmx_compute_dt_noalloc(mpmx_ctx* c, double dt_default, double cfl, double* dt_out) {
    if (!c || !dt_out || c->n == 0 || !(dt_default > 0) || !(cfl > 0)) return MPMX_ERR_ARG;
    const int64_t threads = c->wmap.first ? c->wmap.waves * 64 : c->n;
    if (threads > kNoAllocMaxParticles)
        return fail(c, MPMX_ERR_CAPACITY,
                    "mpmx_compute_dt_noalloc: more than 33,554,432 particles (its static scratch holds 4096 "
                    "partials); use mpmx_compute_dt");
    HIPCHK(c, hipSetDevice(c->device));
    float* s = nullptr;
    HIPCHK(c, noalloc_scratch(c->device, &s));
    const int64_t per = 256 * kNoAllocPerThread;
    const int G = (int)std::max<int64_t>(1, (threads + per - 1) / per);
    double* res = reinterpret_cast<double*>(s + kNoAllocParts);
    HIPCHK(c, launch_speed(c, 2, dt_default, cfl, 1, s, res, G));
    double h[2];
    HIPCHK(c, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *dt_out = h[1];
    return MPMX_OK;
}

mpmx_status mpmx_run_cfl(mpmx_ctx* c, int nsteps, double dt_default, double cfl, int wave, double* time_out) {
    mpmx_status s = check_ready(c, dt_default);
    if (s != MPMX_OK) return s;
    if (!(cfl > 0)) return fail(c, MPMX_ERR_ARG, "cfl must be positive");
    HIPCHK(c, hipSetDevice(c->device));
    k_clear_time<<<1, 64, 0, c->stream>>>(c->dc);
    for (int k = 0; s == MPMX_OK && k < nsteps; ++k) {
        HIPCHK(c, launch_speed(c, 2 | (wave ? 1 : 0), dt_default, cfl, 2, reinterpret_cast<float*>(c->red_part),
                               c->red_part + kRedResult));
        s = run_dt(c, k, -1.0);   // the kernels read the step the CFL kernels left on the device
    }
    if (s != MPMX_OK) return s;
    if (time_out) {
        HIPCHK(c, hipMemcpyAsync(c->hc, c->dc, sizeof(DevCounts), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *time_out = c->hc->time;
    }
    return MPMX_OK;
}

static int field_width(int f) {
    switch (f) {
    case MPMX_FIELD_X: case MPMX_FIELD_V: return 3;
    case MPMX_FIELD_F: case MPMX_FIELD_B: case MPMX_FIELD_SIGMA: return 9;
    default: return 1;
    }
}

// copy a device field of the current (sorted) order to doubles in sorted order
static mpmx_status download_sorted(mpmx_ctx* c, int f, std::vector<double>& out) {
    const int64_t n = c->n;
    const int w = field_width(f);
    out.assign((size_t)(n * w), 0.0);
    const bool mixed = c->cfg.precision != MPMX_FP32;
    const PState& s = c->ps[c->cur];
    const MState& m = c->ms[c->cur];
    auto get_f32 = [&](const float* p, int64_t count, std::vector<float>& h) -> hipError_t {
        h.resize((size_t)count);
        return hipMemcpy(h.data(), p, sizeof(float) * count, hipMemcpyDeviceToHost);
    };
    auto get_u16 = [&](const uint16_t* p, int64_t count, std::vector<uint16_t>& h) -> hipError_t {
        h.resize((size_t)count);
        return hipMemcpy(h.data(), p, sizeof(uint16_t) * count, hipMemcpyDeviceToHost);
    };
    std::vector<float> hf;
    std::vector<uint16_t> hu;
    // AoS (3 per particle) fields
    if (f == MPMX_FIELD_X) {
        HIPCHK(c, get_f32(mixed ? m.x : s.x, 3 * n, hf));
        for (int64_t i = 0; i < 3 * n; ++i) out[i] = hf[i];
        return MPMX_OK;
    }
    if (f == MPMX_FIELD_V) {
        if (mixed) {
            HIPCHK(c, get_u16(m.v, 3 * n, hu));
            for (int64_t i = 0; i < 3 * n; ++i) out[i] = bf2f(hu[i]);
        } else {
            HIPCHK(c, get_f32(s.v, 3 * n, hf));
            for (int64_t i = 0; i < 3 * n; ++i) out[i] = hf[i];
        }
        return MPMX_OK;
    }
    // component-major 9-wide fields -> per particle column-major
    if (f == MPMX_FIELD_F || f == MPMX_FIELD_B || f == MPMX_FIELD_SIGMA) {
        if (f == MPMX_FIELD_SIGMA || !mixed || (f == MPMX_FIELD_F && m.F32)) {
            const float* p = f == MPMX_FIELD_SIGMA ? c->sigma : (f == MPMX_FIELD_F ? (mixed ? m.F32 : s.F) : s.B);
            HIPCHK(c, get_f32(p, 9 * n, hf));
            for (int64_t i = 0; i < n; ++i)
                for (int k = 0; k < 9; ++k) out[9 * i + k] = hf[k * n + i];
        } else {
            HIPCHK(c, get_u16(f == MPMX_FIELD_F ? m.F : m.B, 9 * n, hu));
            for (int64_t i = 0; i < n; ++i)
                for (int k = 0; k < 9; ++k)
                    out[9 * i + k] = f == MPMX_FIELD_F ? (double)h2f(hu[k * n + i]) + ((k % 4) == 0 ? 1.0 : 0.0)
                                                       : (double)bf2f(hu[k * n + i]);
        }
        return MPMX_OK;
    }
    const float* pf = nullptr;
    const uint16_t* pu = nullptr;
    switch (f) {
    case MPMX_FIELD_C: pf = s.c; pu = m.c; break;
    case MPMX_FIELD_ALPHA: pf = s.alpha; pu = m.alpha; break;
    case MPMX_FIELD_PSI: pf = s.psi; pu = m.psi; break;
    case MPMX_FIELD_G: pf = s.g; pu = m.g; break;
    case MPMX_FIELD_MASS: pf = mixed ? m.mass : s.mass; pu = nullptr; break;
    case MPMX_FIELD_FP: pf = c->FP; pu = nullptr; break;
    case MPMX_FIELD_VOL: pf = c->vol; pu = nullptr; break;
    default: return fail(c, MPMX_ERR_ARG, "unknown field");
    }
    if (mixed && pu) {
        HIPCHK(c, get_u16(pu, n, hu));
        for (int64_t i = 0; i < n; ++i) out[i] = bf2f(hu[i]);
    } else {
        HIPCHK(c, get_f32(pf, n, hf));
        for (int64_t i = 0; i < n; ++i) out[i] = hf[i];
    }
    return MPMX_OK;
}

static mpmx_status download_ids(mpmx_ctx* c, std::vector<int>& ids) {
    ids.resize((size_t)c->n);
    const int* p = c->cfg.precision != MPMX_FP32 ? c->ms[c->cur].id : c->ps[c->cur].id;
    HIPCHK(c, hipMemcpy(ids.data(), p, sizeof(int) * c->n, hipMemcpyDeviceToHost));
    return MPMX_OK;
}

mpmx_status mpmx_get_field(mpmx_ctx* c, int f, double* host, size_t count) {
    if (!c || !host || c->n == 0) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const int w = field_width(f);
    if (count < (size_t)(c->n * w)) return fail(c, MPMX_ERR_ARG, "output buffer too small");
    if ((f == MPMX_FIELD_SIGMA || f == MPMX_FIELD_FP) && !c->stress_fields)
        return fail(c, MPMX_ERR_ARG,
                    "sigma and FP of the last stress update are not kept: set mpmx_config.store_stress "
                    "or run the stages (mpmx_stress)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> sorted;
    std::vector<int> ids;
    mpmx_status s = download_sorted(c, f, sorted);
    if (s != MPMX_OK) return s;
    if ((s = download_ids(c, ids)) != MPMX_OK) return s;
    for (int64_t i = 0; i < c->n; ++i) {
        const int o = ids[i];
        if (o < 0 || o >= c->n) return fail(c, MPMX_ERR_HIP, "corrupt particle ids");
        for (int k = 0; k < w; ++k) host[(size_t)o * w + k] = sorted[(size_t)i * w + k];
    }
    return check_err(c);
}

mpmx_status mpmx_set_field(mpmx_ctx* c, int f, const double* host, size_t count) {
    if (!c || !host || c->n == 0) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const int w = field_width(f);
    if (count < (size_t)(c->n * w)) return fail(c, MPMX_ERR_ARG, "input buffer too small");
    if (f == MPMX_FIELD_X || f == MPMX_FIELD_SIGMA || f == MPMX_FIELD_FP || f == MPMX_FIELD_VOL)
        return fail(c, MPMX_ERR_ARG, "field is derived or fixes the order; not settable");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<int> ids;
    mpmx_status s = download_ids(c, ids);
    if (s != MPMX_OK) return s;
    const int64_t n = c->n;
    const bool mixed = c->cfg.precision != MPMX_FP32;
    const PState& ps = c->ps[c->cur];
    const MState& ms = c->ms[c->cur];
    std::vector<float> hf;
    std::vector<uint16_t> hu;
    auto put = [&](void* dst, const void* src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); };
    if (f == MPMX_FIELD_V) {
        hf.resize(3 * n);
        for (int64_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) hf[3 * i + a] = (float)host[3 * ids[i] + a];
        if (mixed) {
            hu.resize(3 * n);
            for (int64_t i = 0; i < 3 * n; ++i) hu[i] = f2bf(hf[i]);
            HIPCHK(c, put(ms.v, hu.data(), 2 * hu.size()));
        } else {
            HIPCHK(c, put(ps.v, hf.data(), 4 * hf.size()));
        }
        return MPMX_OK;
    }
    if (f == MPMX_FIELD_F || f == MPMX_FIELD_B) {
        hf.resize(9 * n);
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < 9; ++k) hf[k * n + i] = (float)host[9 * ids[i] + k];
        if (mixed && f == MPMX_FIELD_F && ms.F32) {
            HIPCHK(c, put(ms.F32, hf.data(), 4 * hf.size()));
        } else if (mixed) {
            hu.resize(9 * n);
            for (int64_t i = 0; i < n; ++i)
                for (int k = 0; k < 9; ++k)
                    hu[k * n + i] = f == MPMX_FIELD_F ? f2h(hf[k * n + i] - ((k % 4) == 0 ? 1.f : 0.f))
                                                      : f2bf(hf[k * n + i]);
            HIPCHK(c, put(f == MPMX_FIELD_F ? ms.F : ms.B, hu.data(), 2 * hu.size()));
        } else {
            HIPCHK(c, put(f == MPMX_FIELD_F ? ps.F : ps.B, hf.data(), 4 * hf.size()));
        }
        return MPMX_OK;
    }
    hf.resize(n);
    for (int64_t i = 0; i < n; ++i) hf[i] = (float)host[ids[i]];
    float* pf = nullptr;
    uint16_t* pu = nullptr;
    switch (f) {
    case MPMX_FIELD_C: pf = ps.c; pu = ms.c; break;
    case MPMX_FIELD_ALPHA: pf = ps.alpha; pu = ms.alpha; break;
    case MPMX_FIELD_PSI: pf = ps.psi; pu = ms.psi; break;
    case MPMX_FIELD_G: pf = ps.g; pu = ms.g; break;
    case MPMX_FIELD_MASS: pf = mixed ? ms.mass : ps.mass; break;
    default: return fail(c, MPMX_ERR_ARG, "unknown field");
    }
    if (mixed && pu) {
        hu.resize(n);
        for (int64_t i = 0; i < n; ++i) hu[i] = f2bf(hf[i]);
        HIPCHK(c, put(pu, hu.data(), 2 * hu.size()));
    } else {
        HIPCHK(c, put(pf, hf.data(), 4 * hf.size()));
    }
    return MPMX_OK;
}

mpmx_status mpmx_get_order(mpmx_ctx* c, int32_t* order, uint64_t* keys) {
    if (!c || c->n == 0) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (order) {
        std::vector<int> ids;
        mpmx_status s = download_ids(c, ids);
        if (s != MPMX_OK) return s;
        std::memcpy(order, ids.data(), sizeof(int) * c->n);
    }
    if (keys) {
        std::vector<uint32_t> k((size_t)c->n);
        HIPCHK(c, hipMemcpy(k.data(), c->slot_keys(), sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost));
        // (Morton page key << 6 | cell) -> 64-bit SPGrid offset (page << 12 | cell << 2)
        for (int64_t i = 0; i < c->n; ++i)
            keys[i] = ((uint64_t)(k[i] >> 6) << 12) | ((uint64_t)(k[i] & 63u) << 2);
    }
    return check_err(c);
}

mpmx_status mpmx_get_tables(mpmx_ctx* c, uint64_t* page_key, int32_t* page2particle, int32_t* adj, int32_t* vpo,
                            int32_t* target, int32_t* particle2cell, int32_t* cell2particle, int32_t counts[4]) {
    if (!c || c->n == 0) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->vpages_valid) HIPCHK(c, build_vpage_tables(c));
    mpmx_status s = read_counts(c);
    if (s != MPMX_OK) return s;
    const int np = std::min<int64_t>(c->hc->num_page, c->cap_p);
    const int nt = std::min<int64_t>(c->hc->num_total, c->cap_tp);
    const int nv = std::min<int64_t>(c->hc->num_vp, c->cap_vp);
    const int n = (int)c->n;
    int num_cell = 0;
    if (particle2cell || cell2particle || counts) {
        int *flags = nullptr, *excl = nullptr, *p2c = nullptr, *c2p = nullptr, *nc = nullptr;
        HIPCHK(c, dalloc(&flags, n));
        HIPCHK(c, dalloc(&excl, n));
        HIPCHK(c, dalloc(&p2c, n));
        HIPCHK(c, dalloc(&c2p, n + 1));
        HIPCHK(c, dalloc(&nc, 1));
        k_cell_flags<<<(n + 255) / 256, 256, 0, c->stream>>>(c->keys, n, flags);
        HIPCHK(c, scan_exclusive(c, flags, excl, n, nullptr, n, nc));
        k_cell_ids<<<(n + 255) / 256, 256, 0, c->stream>>>(c->keys, n, excl, p2c, c2p, nc);
        HIPCHK(c, hipMemcpyAsync(&num_cell, nc, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (particle2cell) HIPCHK(c, hipMemcpy(particle2cell, p2c, sizeof(int) * n, hipMemcpyDeviceToHost));
        if (cell2particle) HIPCHK(c, hipMemcpy(cell2particle, c2p, sizeof(int) * (num_cell + 1), hipMemcpyDeviceToHost));
        dfree(flags); dfree(excl); dfree(p2c); dfree(c2p); dfree(nc);
    }
    if (page_key) {
        std::vector<uint32_t> pk((size_t)std::max(nt, 1));
        HIPCHK(c, hipMemcpy(pk.data(), c->page_key, sizeof(uint32_t) * nt, hipMemcpyDeviceToHost));
        for (int i = 0; i < nt; ++i) page_key[i] = pk[i];
    }
    if (page2particle) HIPCHK(c, hipMemcpy(page2particle, c->page2particle, sizeof(int) * (np + 1), hipMemcpyDeviceToHost));
    if (adj)
        for (int q = 0; q < 7; ++q)
            HIPCHK(c, hipMemcpy(adj + (size_t)q * np, c->adj + (size_t)q * c->cap_p, sizeof(int) * np,
                                hipMemcpyDeviceToHost));
    if (vpo) HIPCHK(c, hipMemcpy(vpo, c->vpo, sizeof(int) * (np + 1), hipMemcpyDeviceToHost));
    if (target) {
        HIPCHK(c, hipMemcpy(target, c->target, sizeof(int) * nv, hipMemcpyDeviceToHost));
        for (int v = 0; v < nv; ++v) target[v] += 1;   // the reference's d_targetPage is 1-based
    }
    if (counts) {
        counts[0] = c->hc->num_page;
        counts[1] = c->hc->num_total;
        counts[2] = c->hc->num_vp;
        counts[3] = num_cell;
    }
    return check_err(c);
}

static mpmx_status get_grid_from(mpmx_ctx* c, const float* src, int ch, float* host, size_t count) {
    if (!c || !host || ch < 0 || ch > 8) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    mpmx_status s = read_counts(c);
    if (s != MPMX_OK) return s;
    const int64_t nt = std::min<int64_t>(c->hc->num_total, c->cap_tp);
    if (count < (size_t)(nt * 64)) return fail(c, MPMX_ERR_ARG, "output buffer too small");
    if (ch == 8) {
        HIPCHK(c, hipMemcpy(host, c->cg_x, sizeof(float) * nt * 64, hipMemcpyDeviceToHost));
        return MPMX_OK;
    }
    std::vector<float> g((size_t)(nt * kPageFloats));
    HIPCHK(c, hipMemcpy(g.data(), src, sizeof(float) * nt * kPageFloats, hipMemcpyDeviceToHost));
    for (int64_t p = 0; p < nt; ++p)
        for (int k = 0; k < 64; ++k) host[p * 64 + k] = g[p * kPageFloats + ch * 64 + k];
    return MPMX_OK;
}

mpmx_status mpmx_get_grid(mpmx_ctx* c, int ch, float* host, size_t count) {
    return get_grid_from(c, c ? c->grid : nullptr, ch, host, count);
}

mpmx_status mpmx_get_grid_p2g(mpmx_ctx* c, int ch, float* host, size_t count) {
    if (!c) return MPMX_ERR_ARG;
    if (!c->keep_p2g) {
        // enable the copy from the next grid update on
        c->keep_p2g = true;
        HIPCHK(c, dalloc(&c->grid_p2g, c->cap_tp * kPageFloats));
        return fail(c, MPMX_ERR_ARG, "P2G copy enabled; call again after the next grid update");
    }
    return get_grid_from(c, c->grid_p2g, ch, host, count);
}

// FP64 mass, momentum and kinetic energy of particles [first, first + count)
// of the current order
static mpmx_status conservation_sums(mpmx_ctx* c, int64_t first, int64_t count, double h[5]) {
    const bool mixed = c->cfg.precision != MPMX_FP32;
    const int G = (int)std::min<int64_t>(kRedParts, std::max<int64_t>(1, (count + 1023) / 1024));
    double* out = c->red_part + kRedResult;
    k_conservation_part<<<G, 256, 0, c->stream>>>((mixed ? c->ms[c->cur].mass : c->ps[c->cur].mass) + first,
                                                  mixed ? nullptr : c->ps[c->cur].v + 3 * first,
                                                  mixed ? c->ms[c->cur].v + 3 * first : nullptr, (int)count,
                                                  c->red_part);
    k_conservation_final<<<1, 64, 0, c->stream>>>(c->red_part, G, out);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(h, out, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MPMX_OK;
}

mpmx_status mpmx_get_stats(mpmx_ctx* c, mpmx_stats* out) {
    if (!c || !out) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    std::memset(out, 0, sizeof(*out));
    out->num_particles = c->n;
    out->substeps = c->substeps;
    if (c->n > 0) {
        double h[5];
        mpmx_status s = conservation_sums(c, 0, c->n, h);
        if (s != MPMX_OK) return s;
        out->mass = h[0];
        out->momentum[0] = h[1];
        out->momentum[1] = h[2];
        out->momentum[2] = h[3];
        out->kinetic_energy = h[4];
    }
    if (c->n > 0 && !c->vpages_valid) HIPCHK(c, build_vpage_tables(c));
    mpmx_status s = read_counts(c);
    if (s != MPMX_OK) return s;
    out->num_page = c->hc->num_page;
    out->num_total_page = c->hc->num_total;
    out->num_virtual_page = c->hc->num_vp;
    out->cg_iters = c->hc->cg_iters;
    out->cg_breakdown = c->hc->cg_breakdown;
    out->svd_fallbacks = (unsigned)c->hc->svd_fallbacks;
    std::vector<SceneCG> g((size_t)c->num_scenes);
    HIPCHK(c, hipMemcpy(g.data(), c->scg, sizeof(SceneCG) * g.size(), hipMemcpyDeviceToHost));
    double norm = 0;
    for (const auto& e : g) norm = std::max(norm, (double)e.norm);
    out->cg_norm = norm;
    return check_err(c);
}

mpmx_status mpmx_get_scene_stats(mpmx_ctx* c, int scene, mpmx_scene_stats* out) {
    if (!c || !out || scene < 0 || scene >= c->num_scenes || c->n == 0) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    std::memset(out, 0, sizeof(*out));
    int64_t first = 0;
    for (int k = 0; k < scene; ++k) first += c->scene_n[(size_t)k];
    out->num_particles = c->scene_n[(size_t)scene];
    out->first_particle = first;
    // the sort keeps each scene's particles contiguous, in scene order
    double h[5];
    mpmx_status s = conservation_sums(c, first, out->num_particles, h);
    if (s != MPMX_OK) return s;
    out->mass = h[0];
    out->momentum[0] = h[1];
    out->momentum[1] = h[2];
    out->momentum[2] = h[3];
    out->kinetic_energy = h[4];
    SceneCG g;
    HIPCHK(c, hipMemcpy(&g, c->scg + scene, sizeof(SceneCG), hipMemcpyDeviceToHost));
    out->num_page = g.np;
    out->num_total_page = g.np + g.nn;
    out->cg_iters = g.iters;
    out->cg_breakdown = g.breakdown;
    out->cg_norm = g.norm;
    return check_err(c);
}

int mpmx_num_scenes(const mpmx_ctx* c) { return c ? c->num_scenes : 0; }

mpmx_status mpmx_get_device_pointers(mpmx_ctx* c, mpmx_device_ptrs* o) {
    if (!c || !o) return MPMX_ERR_ARG;
    std::memset(o, 0, sizeof(*o));
    const bool mixed = c->cfg.precision != MPMX_FP32;
    if (mixed) {
        const MState& s = c->ms[c->cur];
        o->pos = s.x; o->vel = s.v; o->F = s.F32 ? (void*)s.F32 : (void*)s.F; o->B = s.B; o->phase_c = s.c;
        o->alpha = s.alpha; o->psi = s.psi; o->g = s.g; o->mass = s.mass; o->ids = s.id;
    } else {
        const PState& s = c->ps[c->cur];
        o->pos = s.x; o->vel = s.v; o->F = s.F; o->B = s.B; o->phase_c = s.c; o->alpha = s.alpha;
        o->psi = s.psi; o->g = s.g; o->mass = s.mass; o->ids = s.id;
    }
    o->sigma = c->sigma; o->FP = c->FP; o->vol = c->vol; o->keys = c->slot_keys();
    o->grid = c->grid; o->pcg_x = c->cg_x; o->pcg_r = c->cg_r; o->pcg_z = c->cg_z; o->pcg_s = c->cg_s;
    o->pcg_q = c->cg_q; o->page_key = c->page_key; o->page2particle = c->page2particle; o->adj = c->adj;
    o->vpo = c->vpo; o->target = c->target; o->cap_pages = c->cap_p; o->cap_total_pages = c->cap_tp;
    o->precision = c->cfg.precision;
    return MPMX_OK;
}

mpmx_status mpmx_timing_begin(mpmx_ctx* c) {
    if (!c) return MPMX_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto& t : c->timers)
        for (auto& e : t.ev) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    c->timers.clear();
    c->timing = true;
    return MPMX_OK;
}

mpmx_status mpmx_timing_end(mpmx_ctx* c, mpmx_kernel_time* out, int max_entries, int* n_entries) {
    if (!c) return MPMX_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->timing = false;
    int k = 0;
    for (auto& t : c->timers) {
        double total = 0;
        for (auto& e : t.ev) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) total += ms;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (out && k < max_entries) {
            std::memset(out[k].name, 0, sizeof(out[k].name));
            std::strncpy(out[k].name, t.name.c_str(), sizeof(out[k].name) - 1);
            out[k].launches = (int)t.ev.size();
            out[k].total_ms = total;
        }
        ++k;
    }
    c->timers.clear();
    if (n_entries) *n_entries = k;
    return check_err(c);
}

mpmx_status mpmx_time_run(mpmx_ctx* c, int nsteps, double dt, double* ms) {
    if (!c || !ms) return MPMX_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipEventRecord(a, c->stream));
    mpmx_status s = mpmx_run(c, nsteps, dt);
    HIPCHK(c, hipEventRecord(b, c->stream));
    HIPCHK(c, hipEventSynchronize(b));
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t;
    if (s != MPMX_OK) return s;
    return check_err(c);
}

}  // extern "C"
