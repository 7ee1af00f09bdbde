// k_grid.hip — P2G as a gather over grid pages, fused with the grid update and
// the PCG set-up; the stand-alone grid update that follows the scatter P2G.
//
//   P2G   P2G_MLS / P2G_APIC (P2GKernal.cu:6-155, 512-702) plus the phase-field
//         particle passes volP2G_APIC / preConditionP2G_APIC (P2GKernal.cu:
//         897-1157), whose node sums are sum(V w) and sum(V FP w) times node
//         constants and ride along as two more channels;
//   grid  undateGrid_kernal + collideWithGround / PullGrid (TimeIntegrator.cu:
//         63-261) and the PCG initialisation of ExplicitTimeIntegrator.cu:914-935.
//
// Gather form: a workgroup owns one 2x2x2 block of grid pages (8^3 nodes, the
// pages 8b .. 8b+7 in Morton order) at a time.  Particles of page P reach the
// nodes of P and its +1 neighbours, so block b receives from the 27 particle
// pages 2b-1 .. 2b+1 (per axis); of a page below the block along an axis only
// the particles whose cell has index & 3 >= 2 on that axis reach it.  The
// workgroup loads the block's source lists (runs of passing cells, built with
// the page tables by k_block_lists), takes the fixed-point scale of each
// channel from the per-page bounds of its 27 source pages (written with the
// payload), and adds every listed particle's contributions to the block's
// nodes in LDS with 64-bit integer atomics.  It then updates its nodes in
// place and writes them once: no global atomics, no zero fill of the grid,
// and sums that do not depend on the order of the additions.  Each particle
// is read by the one to eight blocks it reaches (about twice on average).
#include "payload.h"

namespace mpmx {


// grid update of one node with the raw P2G sums ch[] (undateGrid_kernal, the
// grips of PullGrid_kernal and collideWithGround), then the node's PCG start
// values r = rhs, z = r / D.  Nodes with sum w = 0 stay inactive.  Returns r z.
__device__ __forceinline__ float node_update(const DevParams& dp, int gi, int gj, int gk, float (&ch)[kP2GCh],
                                             float dt, float& r, float& z) {
    r = 0.f;
    z = 0.f;
    const float W = ch[5];
    if (!(W > 0.f)) return 0.f;
    const float inv = 1.0f / ch[0];
    const float c = ch[4] / W;
    float v0 = ch[1] * inv + dp.gravity[0] * dt;
    float v1 = ch[2] * inv + dp.gravity[1] * dt;
    float v2 = ch[3] * inv + dp.gravity[2] * dt;
    if (dp.grip_lo >= 0) {
        if (gi <= dp.grip_lo) { v0 = -dp.grip_v; v1 = 0.f; v2 = 0.f; }
        if (gi >= dp.grip_hi) { v0 = dp.grip_v; v1 = 0.f; v2 = 0.f; }
    }
    const int lo = dp.wall, hi = dp.N - dp.wall;
    if (gi <= lo && v0 < 0.f) { v0 = 0.f; v1 = 0.f; v2 = 0.f; }
    if (gi > hi && v0 > 0.f) { v0 = 0.f; v1 = 0.f; v2 = 0.f; }
    if (gj <= lo && v1 < 0.f) { v0 = 0.f; v1 = 0.f; v2 = 0.f; }
    if (gj > hi && v1 > 0.f) { v0 = 0.f; v1 = 0.f; v2 = 0.f; }
    if (gk <= lo && v2 < 0.f) { v0 = 0.f; v1 = 0.f; v2 = 0.f; }
    if (gk > hi && v2 > 0.f) { v0 = 0.f; v1 = 0.f; v2 = 0.f; }
    ch[1] = v0;
    ch[2] = v1;
    ch[3] = v2;
    ch[4] = c;
    const float VW = ch[6], VFW = ch[7];
    const float rhs = VW * (c / dt + dp.M);
    const float D = dp.M * VFW + VW / dt;
    ch[6] = rhs;
    ch[7] = D;
    r = rhs;
    z = (fabsf(D) < 1e-30f) ? 0.f : r / D;
    return r * z;
}

struct CGVec {
    float* x;
    float* r;
    float* z;
    float* s;
    float* q;
    uint64_t* amask;   // per grid page: bit c set when node c has sum w > 0 (block P2G only)
};

// PCG scalars start over every substep
// (per scene: a scene whose energy has no fracture starts done)
// (run by the 64 lanes of one wave)
__device__ __forceinline__ void cg_reset(const DevParams& dp, SceneCG* __restrict__ scg, DevCounts* dc) {
    int active = 0;
    for (int s = threadIdx.x; s < dp.num_scenes; s += 64) {
        const int on = dp.scenes[s].phase_field ? 1 : 0;
        scg[s].done = 1 - on;
        scg[s].iters = 0;
        scg[s].breakdown = 0;
        scg[s].norm = 0.f;
        active += on;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) active += __shfl_xor(active, o, 64);
    if (threadIdx.x == 0) {
        dc->cg_active = active;
        dc->cg_done = active ? 0 : 1;
        dc->cg_iters = 0;
        dc->cg_breakdown = 0;
        dc->cg_norm = 0.f;
    }
}
__global__ void k_cg_reset(DevParams dp, SceneCG* __restrict__ scg, DevCounts* dc) { cg_reset(dp, scg, dc); }



// the particle's terms times the channel scales (powers of two, so every
// product below is the unscaled product times the scale, exactly)
template <int TRANSFER>
__device__ __forceinline__ void scale_terms(P2GTerms& T, const float (&sc)[kP2GCh]) {
    T.m *= sc[0];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        T.ar[r] *= sc[1 + r];
        T.Qd[r] *= sc[1 + r];
        T.Qd[r + 3] *= sc[1 + r];
        T.Qd[r + 6] *= sc[1 + r];
        if (TRANSFER == 1) {
            T.sg[r] *= sc[1 + r];
            T.sg[r + 3] *= sc[1 + r];
            T.sg[r + 6] *= sc[1 + r];
        }
    }
    T.cphase *= sc[4];
    T.V *= sc[6];
    T.VF *= sc[7];
}

// add the particle's contributions (terms scaled by scale_terms, sc5 the
// scale of the weight channel) to the block's 8^3 nodes in fixed point.
// Nodes outside the block are skipped; a wave whose lanes all hold particles
// of one source class (see BlockClasses) skips them together.
// The affine term ar + Qd (ii, jj, kk) is accumulated axis by axis with one
// fused multiply-add per nonzero offset, shared by the nodes of a row; a zero
// offset adds nothing (for finite terms it could only change the sign of a
// zero, which converts to the same integer).
template <int TRANSFER>
__device__ __forceinline__ void p2g_fix(const P2GTerms& T, float sc5, unsigned long long* acc) {
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
        const int la = T.l0[0] + ii;
        if (la < 0 || la >= 8) continue;
        float ax[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) ax[r] = ii == 0 ? T.ar[r] : fmaf(T.Qd[r], (float)ii, T.ar[r]);
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const int lb = T.l0[1] + jj;
            if (lb < 0 || lb >= 8) continue;
            const float wxy = T.w[0][ii] * T.w[1][jj];
            float axy[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) axy[r] = jj == 0 ? ax[r] : fmaf(T.Qd[r + 3], (float)jj, ax[r]);
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                const int lc = T.l0[2] + kk;
                if (lc < 0 || lc >= 8) continue;
                const float wt = wxy * T.w[2][kk];
                float pr[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) pr[r] = wt * (kk == 0 ? axy[r] : fmaf(T.Qd[r + 6], (float)kk, axy[r]));
                if (TRANSFER == 1) {
                    const float g0 = T.gw[0][ii] * T.w[1][jj] * T.w[2][kk];
                    const float g1 = T.w[0][ii] * T.gw[1][jj] * T.w[2][kk];
                    const float g2 = T.w[0][ii] * T.w[1][jj] * T.gw[2][kk];
#pragma unroll
                    for (int r = 0; r < 3; ++r) pr[r] -= T.sg[r] * g0 + T.sg[r + 3] * g1 + T.sg[r + 6] * g2;
                }
                unsigned long long* a = acc + acc_node(la, lb, lc);
                const float add[kP2GCh] = {wt * T.m, pr[0], pr[1], pr[2], wt * T.cphase, wt * sc5, wt * T.V, wt * T.VF};
                // the weights are >= 0, so channel 5 needs no sign extension
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k)
                    atomicAdd(a + k * kAccStride, k == 5 ? to_fix_scaled_nonneg(add[k]) : to_fix_scaled(add[k]));
            }
        }
    }
}


// payload of every particle and its page's channel bounds (page_bounds)
template <int TRANSFER, typename ST>
__global__ __launch_bounds__(256) void k_p2g_prep(DevParams dp, ST s, WaveMap wm, const float* __restrict__ sigma,
                                                  const float* __restrict__ FP, const float* __restrict__ vol,
                                                  BlockTables t, float* __restrict__ pay, unsigned* __restrict__ pbound,
                                                  float dt, int n, const DevCounts* __restrict__ dc) {
    dt = substep_dt(dt, dc);
    int i, scene;
    const bool valid = wave_particle(wm, n, i, scene);
    float b[kP2GCh];
#pragma unroll
    for (int k = 0; k < kP2GCh; ++k) b[k] = 0.f;
    const uint32_t pk = valid ? t.keys[i] >> 6 : 0xffffffffu;
    if (valid) {
        float sig[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) sig[k] = sigma[(int64_t)k * n + i];
        prep_particle<TRANSFER>(dp.scenes[scene], s, sig, FP[i], vol[i], i, i, n, dt, pay, b);
    }
    page_bounds(valid, pk, b, t, pbound, dc);
}

// the calling thread's run [jb, je) of the entries of one class of the
// block's sources: the particles of one cell (adjacent after the sort) go
// through one lane in turn instead of hitting the same nodes in one
// instruction
template <int TRANSFER>
__device__ __forceinline__ void p2g_class(const BlockClasses& S, const BlockTables& t, const float* __restrict__ pay,
                                          const int (&org)[3], const float (&sc)[kP2GCh], unsigned long long* acc,
                                          int cls, int jb, int je) {
    for (int j = jb; j < je; ++j) {
        const int i = class_entry(S, cls, j);
        P2GTerms T;
        payload_load<TRANSFER>(pay, i, t.keys[i], org, T);
        scale_terms<TRANSFER>(T, sc);
        p2g_fix<TRANSFER>(T, sc[5], acc);
    }
}

template <int TRANSFER, typename ST>
__global__ __launch_bounds__(256, 4) void k_p2g_block(DevParams dp, const float* __restrict__ pay,
                                                   const unsigned* __restrict__ pbound, BlockTables t,
                                                   const BlockClasses* __restrict__ lists, float* __restrict__ grid,
                                                   float* __restrict__ grid_raw, CGVec cg, float dt, int n,
                                                   const DevCounts* __restrict__ dc) {
    __shared__ unsigned long long acc[kP2GCh * kAccStride];
    __shared__ BlockClasses S;
    __shared__ int sexp[kP2GCh];
    dt = substep_dt(dt, dc);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int nb = dc->num_blocks;
    nb = nb < t.cap_blocks ? nb : (int)t.cap_blocks;
    FOR_LOGICAL_ITEMS(item, nb) {
        const int blk = t.order ? t.order[item] : item;
        const uint32_t bkey = t.block_key[blk];
        int bc[3];
        unmorton(bkey, bc[0], bc[1], bc[2]);
        const DevParams sp = dp.scenes[__builtin_amdgcn_readfirstlane(scene_of_page(dp, bkey << 3))];
        // the block's source lists, built with the page tables
        const BlockClasses* L = lists + blk;
        classes_load(L, S);
        for (int q = tid; q < kP2GCh * kAccStride; q += 256) acc[q] = 0ull;
        const int org[3] = {8 * bc[0], 8 * bc[1], 8 * bc[2]};
        // scales: the largest contribution per channel of the 27 source pages
        if (wave == 1) {
            const int pg = lane < 27 ? L->page[lane] : -1;
            float bmax[kP2GCh];
#pragma unroll
            for (int k = 0; k < kP2GCh; ++k)
                bmax[k] = wave_max(pg >= 0 ? __uint_as_float(pbound[(int64_t)pg * kP2GCh + k]) : 0.f);
            if (lane < kP2GCh) {
                float v = bmax[0];
#pragma unroll
                for (int k = 1; k < kP2GCh; ++k) v = (lane == k) ? bmax[k] : v;
                sexp[lane] = fix_exp(v);
            }
        }
        __syncthreads();
        int e[kP2GCh];
        float sc[kP2GCh];
#pragma unroll
        for (int k = 0; k < kP2GCh; ++k) {
            e[k] = sexp[k];
            sc[k] = ldexpf(1.0f, e[k]);
        }
        int cls, jb, je;
        if (thread_class(S, cls, jb, je)) p2g_class<TRANSFER>(S, t, pay, org, sc, acc, cls, jb, je);
        __syncthreads();
        // grid update of the block's pages, written page-major (coalesced)
        for (int q = tid; q < 512; q += 256) {
            const int o = q >> 6, cell = q & 63;
            const int G = S.out_page[o];
            if (G < 0) continue;
            const int lx = ((o >> 2) & 1) * 4 + (cell >> 4), ly = ((o >> 1) & 1) * 4 + ((cell >> 2) & 3),
                      lz = (o & 1) * 4 + (cell & 3);
            const int node = acc_node(lx, ly, lz);
            float ch[kP2GCh];
#pragma unroll
            for (int k = 0; k < kP2GCh; ++k) ch[k] = from_fix(acc[k * kAccStride + node], e[k]);
            if (grid_raw) {
                float* rp = grid_raw + (int64_t)G * kPageFloats + cell;
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k) rp[k * 64] = ch[k];
            }
            float r, z;
            node_update(sp, org[0] + lx - sp.origin[0], org[1] + ly - sp.origin[1], org[2] + lz - sp.origin[2], ch,
                        dt, r, z);
            float* gp = grid + (int64_t)G * kPageFloats + cell;
#pragma unroll
            for (int k = 0; k < kP2GCh; ++k) gp[k * 64] = ch[k];
            if (sp.phase_field) {
                const int64_t nn = (int64_t)G * 64 + cell;
                cg.x[nn] = 0.f;
                cg.r[nn] = r;
                cg.z[nn] = z;
                cg.s[nn] = z;
                // the wave holds the 64 nodes of page G: their activity bits
                const uint64_t act = __ballot(ch[5] > 0.f);
                if (cell == 0) cg.amask[G] = act;
            }
        }
        __syncthreads();
    }
}


// the grid update after the scatter P2G: one wave per 128-node PCG block, two
// nodes per lane
__global__ __launch_bounds__(256) void k_grid_update(DevParams dp, const uint32_t* __restrict__ page_key,
                                                     float* __restrict__ grid, CGVec cg,
                                                     float* __restrict__ part_rho, float dt, int64_t cap_tp,
                                                     const DevCounts* __restrict__ dc) {
    __shared__ float lds[4];
    dt = substep_dt(dt, dc);
    int nt = dc->num_total;
    nt = nt < cap_tp ? nt : (int)cap_tp;
    const int nblk = (nt + 1) / 2;
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const bool maxdot = dp.dot_mode == 0;
    float wacc = maxdot ? -INFINITY : 0.f;
    for (int blk = wave; blk < nblk; blk += nwaves) {
        float rz = 0.f;
        const int n0 = blk * 128 + lane * 2;
        const int pg = n0 >> 6;
        if (pg < nt) {
            int px, py, pz;
            unmorton(page_key[pg], px, py, pz);
            float* gp = grid + (int64_t)pg * kPageFloats;
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const int cell = (n0 & 63) + hh;
                float ch[kP2GCh];
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k) ch[k] = gp[k * 64 + cell];
                float r, z;
                rz += node_update(dp, px * 4 + (cell >> 4), py * 4 + ((cell >> 2) & 3), pz * 4 + (cell & 3), ch,
                                  dt, r, z);
#pragma unroll
                for (int k = 1; k < kP2GCh; ++k) gp[k * 64 + cell] = ch[k];
                if (dp.phase_field) {
                    const int64_t nn = (int64_t)pg * 64 + cell;
                    cg.x[nn] = 0.f;
                    cg.r[nn] = r;
                    cg.z[nn] = z;
                    cg.s[nn] = z;
                    cg.q[nn] = 0.f;
                }
            }
        }
        const float bs = wave_sum(rz);
        wacc = maxdot ? fmaxf(wacc, bs) : wacc + bs;
    }
    const float r = maxdot ? combine_waves<true>(wacc, lds) : combine_waves<false>(wacc, lds);
    if (threadIdx.x == 0) part_rho[blockIdx.x] = r;
}

// source lists and G2P ranges of every 2x2x2 block (block_lists_wave), one
// wave per block; they depend on the page tables only and serve the block
// P2G, the K s gathers of the substep's PCG and the block G2P
__global__ __launch_bounds__(256) void k_block_lists(DevParams dp, BlockTables t, BlockClasses* __restrict__ lists,
                                                     BlockG2P* __restrict__ g2p, uint8_t* __restrict__ wclass, int n,
                                                     SceneCG* __restrict__ scg, unsigned* __restrict__ pbound,
                                                     DevCounts* dc) {
    // the substep's PCG reset and the clear of the per-page channel bounds
    // ride along (the next kernels, the stress pass and the block P2G, need
    // both)
    if (blockIdx.x == 0 && threadIdx.x < 64) cg_reset(dp, scg, dc);
    {
        int64_t npb = dc->num_page;
        npb = npb < t.cap_p ? npb : t.cap_p;
        const int64_t m = npb * kP2GCh;
        for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x)
            pbound[q] = 0u;
    }
    const int np_raw = dc->num_page;
    const int np = np_raw < t.cap_p ? np_raw : (int)t.cap_p;
    int nt = dc->num_total;
    nt = nt < t.cap_tp ? nt : (int)t.cap_tp;
    int nb = dc->num_blocks;
    nb = nb < t.cap_blocks ? nb : (int)t.cap_blocks;
    // one block per wave
    const int per = (int)(blockDim.x >> 6), wave = (int)(threadIdx.x >> 6);
    const int end = xcd_item_end(nb), step = xcd_item_step(per);
    for (int blk = xcd_item_first(nb, per) + wave; blk < end; blk += step) {
        const int entries = block_lists_wave(t, t.block_key[blk], n, np_raw, np, nt, lists + blk, g2p + blk);
        if ((threadIdx.x & 63) == 0) wclass[blk] = (uint8_t)work_class(entries);
    }
}

hipError_t launch_block_lists(mpmx_ctx* c) {
    int g = (block_grid(c, 100) + 3) / 4;
    g = (g + 7) / 8 * 8;
    MPMX_TIMED(c, "block_lists", (k_block_lists<<<g, 256, 0, c->stream>>>(
                                      c->dp, block_tables_of(c), static_cast<BlockClasses*>(c->block_lists),
                                      static_cast<BlockG2P*>(c->block_g2p), c->block_wclass, (int)c->n, c->scg,
                                      c->pbound, c->dc)));
    return hipGetLastError();
}

hipError_t launch_p2g_scatter(mpmx_ctx* c, float dt);

hipError_t launch_cg_reset(mpmx_ctx* c) {
    k_cg_reset<<<1, 64, 0, c->stream>>>(c->dp, c->scg, c->dc);
    return hipGetLastError();
}

static CGVec cg_vec(mpmx_ctx* c) { return CGVec{c->cg_x, c->cg_r, c->cg_z, c->cg_s, c->cg_q, c->cg_amask}; }

// ---- owner-block P2G (p2g_variant 3) ------------------------------------------
// Every particle is read once, by the block that owns its base node.  A
// workgroup takes one 2x2x2-page block B at a time: B's own particles (a
// contiguous range of the sorted order) reach nodes 0..9 of B per axis, so
// all 27 contributions of every particle land in a 10^3-node fixed-point
// tile in LDS with every lane active (k_p2g_own).  Nodes 2..7 along every
// axis (the core, 216 of B's 512) receive only B's particles and are final:
// the same workgroup runs their grid update and PCG set-up.  The other 784
// nodes of the tile, B's shell (a coordinate below 2) and its halo (a
// coordinate of 8 or 9, the shell of the blocks above), go to B's record in
// HBM with B's channel exponents; a second pass (k_p2g_shell) sums each
// shell node from the records of the up to 8 blocks that reach it, in a
// fixed order, in FP64, then updates it.  Each block's exponents come from
// the bounds of its own particles (the per-page bounds of its 8 pages), so a
// term keeps 30 bits relative to the largest contribution of its own block;
// the integer sums are exact and the FP64 combination has a fixed order, so
// the result is bit-reproducible.  It is slower than the gather (variant 2)
// on every measured scene: a record is 784 x 8 x 8 bytes, three times the
// block's 512 x 8 x 4 bytes of grid output, written and read back, and the
// 64 KiB tile leaves room for two workgroups per CU.
constexpr int kOwnE = 10;                   // tile edge (nodes)
constexpr int kOwn3 = kOwnE * kOwnE * kOwnE;
constexpr int kOwnSlots = kOwn3 - 216;       // nodes outside the core [2,8)^3
// record of one block: a 64-byte header (flag, exponents) and the 784 slots
// of each channel as 64-bit integers
constexpr int64_t kOwnRecU64 = 8 + (int64_t)kP2GCh * kOwnSlots;
struct OwnHeader {
    int flag;            // 1 when the block owns particles (its record holds sums)
    int e[kP2GCh];       // channel exponents of its terms
    int pad[7];
};

__host__ __device__ constexpr bool own_core(int X, int Y, int Z) {
    return X >= 2 && X < 8 && Y >= 2 && Y < 8 && Z >= 2 && Z < 8;
}
// slot of non-core tile node (X, Y, Z): its rank among the non-core nodes in
// X-major, then Y, then Z order (planes X < 2 or >= 8 hold 100 slots, the
// others 64; rows Y < 2 or >= 8 of those hold 10, the others 4: Z 0, 1, 8, 9)
__device__ __forceinline__ int own_slot(int X, int Y, int Z) {
    const int xi = X < 2 ? X : (X < 8 ? X - 2 : X - 8);
    int s = X < 2 ? 100 * X : (X < 8 ? 200 + 64 * xi : 584 + 100 * xi);
    if (X < 2 || X >= 8) return s + 10 * Y + Z;
    s += Y < 2 ? 10 * Y : (Y < 8 ? 20 + 4 * (Y - 2) : 44 + 10 * (Y - 8));
    if (Y < 2 || Y >= 8) return s + Z;
    return s + (Z < 2 ? Z : Z - 6);
}

// the particle's 27 contributions, all inside the tile (terms scaled by
// scale_terms; the affine term accumulated as in p2g_fix)
template <int TRANSFER>
__device__ __forceinline__ void p2g_fix_own(const P2GTerms& T, float sc5, unsigned long long* acc) {
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
        const int la = T.l0[0] + ii;
        float ax[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) ax[r] = ii == 0 ? T.ar[r] : fmaf(T.Qd[r], (float)ii, T.ar[r]);
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const int lb = T.l0[1] + jj;
            const float wxy = T.w[0][ii] * T.w[1][jj];
            float axy[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) axy[r] = jj == 0 ? ax[r] : fmaf(T.Qd[r + 3], (float)jj, ax[r]);
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                const int lc = T.l0[2] + kk;
                const float wt = wxy * T.w[2][kk];
                float pr[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) pr[r] = wt * (kk == 0 ? axy[r] : fmaf(T.Qd[r + 6], (float)kk, axy[r]));
                if (TRANSFER == 1) {
                    const float g0 = T.gw[0][ii] * T.w[1][jj] * T.w[2][kk];
                    const float g1 = T.w[0][ii] * T.gw[1][jj] * T.w[2][kk];
                    const float g2 = T.w[0][ii] * T.w[1][jj] * T.gw[2][kk];
#pragma unroll
                    for (int r = 0; r < 3; ++r) pr[r] -= T.sg[r] * g0 + T.sg[r + 3] * g1 + T.sg[r + 6] * g2;
                }
                unsigned long long* a = acc + (la * kOwnE + lb) * kOwnE + lc;
                const float add[kP2GCh] = {wt * T.m, pr[0], pr[1], pr[2], wt * T.cphase, wt * sc5, wt * T.V, wt * T.VF};
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k)
                    atomicAdd(a + k * kOwn3, k == 5 ? to_fix_scaled_nonneg(add[k]) : to_fix_scaled(add[k]));
            }
        }
    }
}

// the grid update of own node q of a block (q = page o of the block, cell),
// raw sums ch; writes the grid page, the raw copy and the PCG start values,
// returns whether the node is active (sum w > 0)
__device__ __forceinline__ bool own_node_out(const DevParams& sp, const int (&org)[3], int o, int cell, int G,
                                             float (&ch)[kP2GCh], float* __restrict__ grid,
                                             float* __restrict__ grid_raw, const CGVec& cg, float dt) {
    const int lx = ((o >> 2) & 1) * 4 + (cell >> 4), ly = ((o >> 1) & 1) * 4 + ((cell >> 2) & 3),
              lz = (o & 1) * 4 + (cell & 3);
    if (grid_raw) {
        float* rp = grid_raw + (int64_t)G * kPageFloats + cell;
#pragma unroll
        for (int k = 0; k < kP2GCh; ++k) rp[k * 64] = ch[k];
    }
    float r, z;
    node_update(sp, org[0] + lx - sp.origin[0], org[1] + ly - sp.origin[1], org[2] + lz - sp.origin[2], ch, dt, r, z);
    float* gp = grid + (int64_t)G * kPageFloats + cell;
#pragma unroll
    for (int k = 0; k < kP2GCh; ++k) gp[k * 64] = ch[k];
    if (sp.phase_field) {
        const int64_t nn = (int64_t)G * 64 + cell;
        cg.x[nn] = 0.f;
        cg.r[nn] = r;
        cg.z[nn] = z;
        cg.s[nn] = z;
    }
    return ch[5] > 0.f;
}

template <int TRANSFER>
__global__ __launch_bounds__(256, 2) void k_p2g_own(DevParams dp, const float* __restrict__ pay,
                                                 const unsigned* __restrict__ pbound, BlockTables t,
                                                 const BlockClasses* __restrict__ lists,
                                                 const BlockG2P* __restrict__ info, float* __restrict__ grid,
                                                 float* __restrict__ grid_raw, CGVec cg,
                                                 unsigned long long* __restrict__ part, int64_t cap_part, float dt,
                                                 DevCounts* dc) {
    __shared__ unsigned long long acc[kP2GCh * kOwn3];
    __shared__ int sexp[kP2GCh];
    dt = substep_dt(dt, dc);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int nb = dc->num_blocks;
    nb = nb < t.cap_blocks ? nb : (int)t.cap_blocks;
    FOR_LOGICAL_ITEMS(item, nb) {
        const int blk = t.order ? t.order[item] : item;
        if (blk >= cap_part) {
            if (tid == 0) atomicOr(&dc->err, kErrCapacity);
            continue;
        }
        const uint32_t bkey = t.block_key[blk];
        int bc[3];
        unmorton(bkey, bc[0], bc[1], bc[2]);
        const DevParams sp = dp.scenes[__builtin_amdgcn_readfirstlane(scene_of_page(dp, bkey << 3))];
        const BlockClasses* L = lists + blk;
        const BlockG2P* B = info + blk;
        const int pa = __builtin_amdgcn_readfirstlane(B->pa), pb = __builtin_amdgcn_readfirstlane(B->pb);
        const bool has = pa < pb;
        unsigned long long* rec = part + (int64_t)blk * kOwnRecU64;
        // exponents from the bounds of the block's own pages (source slots 0..7)
        if (wave == 0) {
            const int pg = lane < 8 ? L->page[lane] : -1;
            float bmax[kP2GCh];
#pragma unroll
            for (int k = 0; k < kP2GCh; ++k)
                bmax[k] = wave_max(pg >= 0 ? __uint_as_float(pbound[(int64_t)pg * kP2GCh + k]) : 0.f);
            if (lane < kP2GCh) {
                float v = bmax[0];
#pragma unroll
                for (int k = 1; k < kP2GCh; ++k) v = (lane == k) ? bmax[k] : v;
                sexp[lane] = fix_exp(v);
            }
        }
        if (has)
            for (int q = tid; q < kP2GCh * kOwn3; q += 256) acc[q] = 0ull;
        __syncthreads();
        int e[kP2GCh];
        float sc[kP2GCh];
#pragma unroll
        for (int k = 0; k < kP2GCh; ++k) {
            e[k] = sexp[k];
            sc[k] = ldexpf(1.0f, e[k]);
        }
        if (tid < 16) {
            int* h = reinterpret_cast<int*>(rec);
            h[tid] = tid == 0 ? (has ? 1 : 0) : (tid <= kP2GCh ? e[tid - 1] : 0);
        }
        const int org[3] = {8 * bc[0], 8 * bc[1], 8 * bc[2]};
        if (has) {
            // a wave takes 64 consecutive particles, lane l the one at
            // (l mod 8) * 8 + l / 8, so that the up to 8 particles of a cell,
            // adjacent in the sorted order, sit in different atomics
            const int off = 64 * wave + ((lane & 7) << 3) + (lane >> 3);
            for (int c0 = pa; c0 < pb; c0 += 256) {
                const int i = c0 + off;
                if (i < pb) {
                    P2GTerms T;
                    payload_load<TRANSFER>(pay, i, t.keys[i], org, T);
                    // an own particle's base node lies in the block; stale
                    // tables after a capacity overflow (latched) must not
                    // index outside the tile
                    const bool inside = (unsigned)T.l0[0] < 8u && (unsigned)T.l0[1] < 8u && (unsigned)T.l0[2] < 8u;
                    if (inside) {
                        scale_terms<TRANSFER>(T, sc);
                        p2g_fix_own<TRANSFER>(T, sc[5], acc);
                    }
                }
            }
        }
        __syncthreads();
        // the core nodes are final: grid update, written page-major (a wave
        // holds the 64 nodes of one page)
        for (int q = tid; q < 512; q += 256) {
            const int o = q >> 6, cell = q & 63;
            const int G = L->out_page[o];
            const int lx = ((o >> 2) & 1) * 4 + (cell >> 4), ly = ((o >> 1) & 1) * 4 + ((cell >> 2) & 3),
                      lz = (o & 1) * 4 + (cell & 3);
            bool act = false;
            if (G >= 0 && own_core(lx, ly, lz)) {
                const int node = (lx * kOwnE + ly) * kOwnE + lz;
                float ch[kP2GCh];
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k) ch[k] = has ? from_fix(acc[k * kOwn3 + node], e[k]) : 0.f;
                act = own_node_out(sp, org, o, cell, G, ch, grid, grid_raw, cg, dt);
            }
            const uint64_t bits = __ballot(act);
            if (sp.phase_field && cell == 0 && G >= 0) cg.amask[G] = bits;   // the shell pass adds its nodes
        }
        // the shell and halo nodes to the block's record
        if (has) {
            unsigned long long* data = rec + 8;
            for (int q = tid; q < kOwn3; q += 256) {
                const int X = q / 100, Y = (q / 10) % 10, Z = q % 10;
                if (own_core(X, Y, Z)) continue;
                const int s = own_slot(X, Y, Z);
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k) data[(int64_t)k * kOwnSlots + s] = acc[k * kOwn3 + q];
            }
        }
        __syncthreads();
    }
}

// the shell nodes of every block: the sums of the records of the (up to 8)
// blocks B - d, d in {0,1}^3, that reach each node, in the order of d, in
// FP64 with each record's exponents, then the grid update
__global__ __launch_bounds__(256) void k_p2g_shell(DevParams dp, BlockTables t, const BlockClasses* __restrict__ lists,
                                                   float* __restrict__ grid, float* __restrict__ grid_raw, CGVec cg,
                                                   const unsigned long long* __restrict__ part, int64_t cap_part,
                                                   float dt, const DevCounts* __restrict__ dc) {
    __shared__ int src[8];
    __shared__ int sexp[8][kP2GCh];
    dt = substep_dt(dt, dc);
    const int tid = threadIdx.x;
    int nb = dc->num_blocks;
    nb = nb < t.cap_blocks ? nb : (int)t.cap_blocks;
    FOR_LOGICAL_ITEMS(blk, nb) {
        if (blk >= cap_part) continue;   // k_p2g_own has latched the capacity error
        const uint32_t bkey = t.block_key[blk];
        int bc[3];
        unmorton(bkey, bc[0], bc[1], bc[2]);
        const int scene = __builtin_amdgcn_readfirstlane(scene_of_page(dp, bkey << 3));
        const DevParams sp = dp.scenes[scene];
        const BlockClasses* L = lists + blk;
        if (tid < 8) {
            const int dx = (tid >> 2) & 1, dy = (tid >> 1) & 1, dz = tid & 1;
            const int sx = bc[0] - dx, sy = bc[1] - dy, sz = bc[2] - dz;
            int id = -1;
            if (sx >= 0 && sy >= 0 && sz >= 0) {
                const uint32_t sk = morton((uint32_t)sx, (uint32_t)sy, (uint32_t)sz);
                if (scene_of_page(dp, sk << 3) == scene) {
                    id = block_rank(t, sk);
                    if (id >= cap_part) id = -1;
                }
            }
            if (id >= 0) {
                const OwnHeader* h = reinterpret_cast<const OwnHeader*>(part + (int64_t)id * kOwnRecU64);
                if (!h->flag) id = -1;
                else
                    for (int k = 0; k < kP2GCh; ++k) sexp[tid][k] = h->e[k];
            }
            src[tid] = id;
        }
        __syncthreads();
        const int org[3] = {8 * bc[0], 8 * bc[1], 8 * bc[2]};
        for (int q = tid; q < 512; q += 256) {
            const int o = q >> 6, cell = q & 63;
            const int G = L->out_page[o];
            const int lx = ((o >> 2) & 1) * 4 + (cell >> 4), ly = ((o >> 1) & 1) * 4 + ((cell >> 2) & 3),
                      lz = (o & 1) * 4 + (cell & 3);
            bool act = false;
            const bool shell = G >= 0 && !own_core(lx, ly, lz);
            if (shell) {
                double s[kP2GCh];
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k) s[k] = 0.0;
                for (int d = 0; d < 8; ++d) {
                    const int dx = (d >> 2) & 1, dy = (d >> 1) & 1, dz = d & 1;
                    const int id = src[d];
                    if (id < 0 || (dx && lx >= 2) || (dy && ly >= 2) || (dz && lz >= 2)) continue;
                    const unsigned long long* data = part + (int64_t)id * kOwnRecU64 + 8;
                    const int sl = own_slot(lx + 8 * dx, ly + 8 * dy, lz + 8 * dz);
#pragma unroll
                    for (int k = 0; k < kP2GCh; ++k)
                        s[k] += ldexp((double)(long long)data[(int64_t)k * kOwnSlots + sl], -sexp[d][k]);
                }
                float ch[kP2GCh];
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k) ch[k] = (float)s[k];
                act = own_node_out(sp, org, o, cell, G, ch, grid, grid_raw, cg, dt);
            }
            const uint64_t bits = __ballot(act);
            if (sp.phase_field && cell == 0 && G >= 0) cg.amask[G] |= bits;
        }
        __syncthreads();
    }
}

// block P2G (with grid update and PCG set-up) over the payload and page bounds
template <int TRANSFER, typename ST>
static hipError_t gather_blocks_t(mpmx_ctx* c, float dt) {
    const BlockTables t = block_tables_of(c);
    float* raw = (c->keep_p2g && c->grid_p2g) ? c->grid_p2g : nullptr;
    // about one workgroup per block: the hardware balances their uneven work
    const int g = block_grid(c, c->p2g_pct);
    if (c->cfg.p2g_variant == 3) {
        // owner blocks, then the shell nodes (both timed as the P2G stage)
        unsigned long long* part = static_cast<unsigned long long*>(c->p2g_part);
        timer_push(c, "p2g_block", true);
        k_p2g_own<TRANSFER><<<g, 256, 0, c->stream>>>(c->dp, c->payload, c->pbound, t,
                                                       static_cast<BlockClasses*>(c->block_lists),
                                                       static_cast<BlockG2P*>(c->block_g2p), c->grid, raw,
                                                       cg_vec(c), part, c->cap_part, dt, c->dc);
        k_p2g_shell<<<block_grid(c, 100), 256, 0, c->stream>>>(c->dp, t, static_cast<BlockClasses*>(c->block_lists),
                                                               c->grid, raw, cg_vec(c), part, c->cap_part, dt,
                                                               c->dc);
        timer_push(c, "p2g_block", false);
        return hipGetLastError();
    }
    MPMX_TIMED(c, "p2g_block", (k_p2g_block<TRANSFER, ST><<<g, 256, 0, c->stream>>>(
                                    c->dp, c->payload, c->pbound, t, static_cast<BlockClasses*>(c->block_lists),
                                    c->grid, raw, cg_vec(c), dt, (int)c->n, c->dc)));
    // the PCG's starting "dot"(r, z) comes with its first "dot"(s, q) (k_cg_dot)
    return hipGetLastError();
}

template <int TRANSFER, typename ST>
static hipError_t gather_dispatch_t(mpmx_ctx* c, ST st, float dt) {
    c->order_ok = false;   // the stage path leaves the blocks in their own order
    const BlockTables t = block_tables_of(c);
    hipError_t e = fill_zero_dev_count(c, reinterpret_cast<float*>(c->pbound), &c->dc->num_page, kP2GCh, c->cap_p);
    if (e != hipSuccess) return e;
    MPMX_TIMED(c, "p2g_prep", (k_p2g_prep<TRANSFER, ST><<<particle_blocks(c), c->particle_wg, 0, c->stream>>>(
                                   c->dp, st, c->wmap, c->sigma, c->FP, c->vol, t, c->payload, c->pbound, dt,
                                   (int)c->n, c->dc)));
    return gather_blocks_t<TRANSFER, ST>(c, dt);
}

template <typename ST>
static hipError_t gather_dispatch(mpmx_ctx* c, ST st, float dt) {
    return c->P.transfer == 0 ? gather_dispatch_t<0>(c, st, dt) : gather_dispatch_t<1>(c, st, dt);
}

// P2G; the gather variant (default) also runs the grid update and PCG set-up
hipError_t launch_p2g(mpmx_ctx* c, float dt) {
    hipError_t e = launch_cg_reset(c);
    if (e != hipSuccess) return e;
    if (c->cfg.p2g_variant == 1) {
        c->grid_fused = false;
        return launch_p2g_scatter(c, dt);
    }
    e = (c->cfg.precision != MPMX_FP32) ? gather_dispatch(c, c->ms[c->cur], dt) : gather_dispatch(c, c->ps[c->cur], dt);
    c->grid_fused = true;
    return e;
}

// stress and gather P2G of a substep: the stress pass also writes the
// payload and page bounds, so the P2G's own payload pass is skipped.  Same
// results as launch_stress then launch_p2g, bit for bit.
hipError_t launch_stress_p2g(mpmx_ctx* c, float dt) {
    // the PCG reset and the zero page bounds come from k_block_lists, which
    // the substep's table build has just run
    hipError_t e = hipSuccess;
    c->order_ok = true;   // the stress pass orders the blocks
    MPMX_TIMED(c, "stress_prep", e = launch_stress_prep(c, dt));
    if (e != hipSuccess) return e;
    const bool mixed = c->cfg.precision != MPMX_FP32;
    if (c->P.transfer == 0) e = mixed ? gather_blocks_t<0, MState>(c, dt) : gather_blocks_t<0, PState>(c, dt);
    else e = mixed ? gather_blocks_t<1, MState>(c, dt) : gather_blocks_t<1, PState>(c, dt);
    c->grid_fused = true;
    return e;
}

hipError_t launch_grid_update(mpmx_ctx* c, float dt) {
    if (c->grid_fused) return hipSuccess;
    if (c->keep_p2g && c->grid_p2g) {
        hipError_t e = hipMemcpyAsync(c->grid_p2g, c->grid, sizeof(float) * c->cap_tp * kPageFloats,
                                      hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) return e;
    }
    k_grid_update<<<c->cg_nb, 256, 0, c->stream>>>(c->dp, c->page_key, c->grid, cg_vec(c), c->cg_part + c->cg_nb,
                                                    dt, c->cap_tp, c->dc);
    return hipGetLastError();
}

}  // namespace mpmx
