// k_transfer.hip — scatter particle-to-grid transfer and grid-to-particle
// transfer (MLS and APIC).
//
//   P2G  replaces P2G_MLS / P2G_APIC (P2GKernal.cu:6-155, 512-702) and the two
//        phase-field particle passes volP2G_APIC / preConditionP2G_APIC
//        (P2GKernal.cu:897-1157): their per-node sums are sum(V w) and
//        sum(V FP w) times node constants, so they ride along as two more
//        channels of the same pass.
//   G2P  replaces G2P_MLS / G2P_APIC (G2PKernal.cu:32-441) and update_grid_phase
//        (ExplicitTimeIntegrator.cu:632-642), whose increment x - c is formed
//        while the tile is staged.
//
// One workgroup works through a contiguous run of virtual pages (<= vpage
// particles of one 4x4x4-cell page, sorted by cell).  The 6x6x6 node tile that
// the page's particles touch is staged in LDS.  The scatter P2G here (variant
// 1) accumulates that tile with LDS float atomics and flushes non-zero entries
// to HBM with global float atomics (8 pages: the page and its 7 +x/+y/+z
// neighbours); the default gather P2G is in k_grid.hip.
#include "access.h"

namespace mpmx {

constexpr int kP2GCh = 8;   // m, p0, p1, p2, sum w c, sum w, sum V w, sum V FP w

template <int BLOCK, int TRANSFER, typename ST>
__global__ __launch_bounds__(BLOCK) void k_p2g(DevParams dp, ST s, const float* __restrict__ sigma,
                                               const float* __restrict__ FP, const float* __restrict__ vol,
                                               Tables t, float* __restrict__ grid, float dt, int n,
                                               const DevCounts* __restrict__ dc) {
    __shared__ float tile[kP2GCh * kTileN];
    dt = substep_dt(dt, dc);
    int nvp = dc->num_vp;
    nvp = nvp < t.cap_vp ? nvp : (int)t.cap_vp;
    int vb, ve;
    block_range(nvp, vb, ve);
    for (int vp = vb; vp < ve; ++vp) {
        for (int q = threadIdx.x; q < kP2GCh * kTileN; q += BLOCK) tile[q] = 0.f;
        __syncthreads();
        int p, s0, e;
        vpage_range(t, vp, BLOCK, n, p, s0, e);
        const int i = s0 + threadIdx.x;
        if (i < e) {
            int pc[3];
            unmorton(t.page_key[p], pc[0], pc[1], pc[2]);
            float d0[3], w[3][3], gw[3][3];
            int l[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float xa = ld_x(s, i, a);
                const int b = base_node(xa, dp.N);
                d0[a] = xa * dp.inv_dx - (float)b;
                int la = b - 4 * pc[a];
                l[a] = la < 0 ? 0 : (la > 3 ? 3 : la);
                if (TRANSFER == 0) bspline(d0[a], w[a]);
                else bspline_grad(d0[a], w[a], gw[a]);
            }
            const float m = ld_m(s, i);
            const float cphase = ld_c(s, i);
            const float V = vol[i];
            const float VF = V * FP[i];
            float v[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) v[a] = ld_v(s, i, a);
            // affine matrix per node offset in units of dx: p_r = w (a_r + sum_c Qd(r,c) o_c)
            float Qd[9], sg[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const float Bk = ld_B(s, i, k, n);
                const float sk = sigma[(int64_t)k * n + i];
                if (TRANSFER == 0) Qd[k] = (Bk * m - sk * dt) * dp.Dinv * dp.dx;
                else { Qd[k] = Bk * dp.Dinv * dp.dx * m; sg[k] = sk * dt * dp.inv_dx; }
            }
            float ar[3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
                ar[r] = m * v[r] - (Qd[r] * d0[0] + Qd[r + 3] * d0[1] + Qd[r + 6] * d0[2]);
#pragma unroll
            for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) {
                    const float wxy = w[0][ii] * w[1][jj];
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) {
                        const float wt = wxy * w[2][kk];
                        float pr[3];
#pragma unroll
                        for (int r = 0; r < 3; ++r)
                            pr[r] = wt * (ar[r] + Qd[r] * ii + Qd[r + 3] * jj + Qd[r + 6] * kk);
                        if (TRANSFER == 1) {
                            // + dt * (-sigma grad w), gradient in units of 1/dx
                            const float g0 = gw[0][ii] * w[1][jj] * w[2][kk];
                            const float g1 = w[0][ii] * gw[1][jj] * w[2][kk];
                            const float g2 = w[0][ii] * w[1][jj] * gw[2][kk];
#pragma unroll
                            for (int r = 0; r < 3; ++r) pr[r] -= sg[r] * g0 + sg[r + 3] * g1 + sg[r + 6] * g2;
                        }
                        const int node = (l[0] + ii) * 36 + (l[1] + jj) * 6 + (l[2] + kk);
                        atomicAdd(&tile[0 * kTileN + node], wt * m);
                        atomicAdd(&tile[1 * kTileN + node], pr[0]);
                        atomicAdd(&tile[2 * kTileN + node], pr[1]);
                        atomicAdd(&tile[3 * kTileN + node], pr[2]);
                        atomicAdd(&tile[4 * kTileN + node], wt * cphase);
                        atomicAdd(&tile[5 * kTileN + node], wt);
                        atomicAdd(&tile[6 * kTileN + node], wt * V);
                        atomicAdd(&tile[7 * kTileN + node], wt * VF);
                    }
                }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < kP2GCh * kTileN; q += BLOCK) {
            const float val = tile[q];
            if (val != 0.f) {
                const int ch = q / kTileN, node = q - ch * kTileN;
                const int a = node / 36, b = (node / 6) % 6, cc = node % 6;
                const int64_t gn = tile_node(t, p, a, b, cc);
                if (gn >= 0) atomicAdd(&grid[(gn >> 6) * kPageFloats + ch * 64 + (gn & 63)], val);
            }
        }
        __syncthreads();
    }
}

// G2P of particle i from a staged node region of R^3 nodes whose node (0,0,0)
// is node org (scene-local cells): channel ch of region node (a, b, c) at
// tv[ch R^3 + (a R + b) R + c] (v0, v1, v2 and the damage increment).  MLS
// velocity clamp, c update, F = (I + dt grad v) F, B, x, the next cell key
// (with the scene's slot origin) and the error flags.
// position and damage of a particle, which the block G2P loads ahead of the
// node staging (F is read after the node loop, where it is needed: held
// through the loop it would cost a wave per SIMD)
struct G2PIn {
    float x[3], c;
};
template <typename ST>
__device__ __forceinline__ void g2p_load(const ST& s, int i, G2PIn& P) {
#pragma unroll
    for (int a = 0; a < 3; ++a) P.x[a] = ld_x(s, i, a);
    P.c = ld_c(s, i);
}

template <int TRANSFER, int R, typename ST>
__device__ __forceinline__ void g2p_particle(const DevParams& dp, const ST& s, int i, int n, float dt,
                                             const int (&org)[3], const float* tv, const G2PIn& P,
                                             uint32_t* __restrict__ keys, DevCounts* dc) {
    constexpr int R3 = R * R * R;
    float d0[3], w[3][3], gw[3][3];
    int l[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float xa = P.x[a];
        const int b = base_node(xa, dp.N);
        d0[a] = xa * dp.inv_dx - (float)b;
        const int la = b - org[a];
        l[a] = la < 0 ? 0 : (la > R - 3 ? R - 3 : la);
        if (TRANSFER == 0) bspline(d0[a], w[a]);
        else bspline_grad(d0[a], w[a], gw[a]);
    }
    float vp3[3] = {0.f, 0.f, 0.f}, dcs = 0.f;
    float Bn[9], G[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) { Bn[k] = 0.f; G[k] = 0.f; }
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                const int node = ((l[0] + ii) * R + (l[1] + jj)) * R + (l[2] + kk);
                const float wt = w[0][ii] * w[1][jj] * w[2][kk];
                const float vi[3] = {tv[node], tv[R3 + node], tv[2 * R3 + node]};
                dcs += wt * tv[3 * R3 + node];
                const float o[3] = {((float)ii - d0[0]) * dp.dx, ((float)jj - d0[1]) * dp.dx,
                                    ((float)kk - d0[2]) * dp.dx};
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    vp3[r] += wt * vi[r];
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) Bn[r + 3 * cc] += wt * vi[r] * o[cc];
                }
                if (TRANSFER == 1) {
                    const float g3[3] = {gw[0][ii] * w[1][jj] * w[2][kk] * dp.inv_dx,
                                         w[0][ii] * gw[1][jj] * w[2][kk] * dp.inv_dx,
                                         w[0][ii] * w[1][jj] * gw[2][kk] * dp.inv_dx};
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int cc = 0; cc < 3; ++cc) G[r + 3 * cc] += vi[r] * g3[cc];
                }
            }
        // one slab of nine nodes at a time: the scheduler would otherwise
        // hoist all 108 staged-node reads and hold them in registers
        __builtin_amdgcn_sched_barrier(0);
    }
    // B and the damage update first: stored before the velocity clamp's
    // branch, their sums over the 27 nodes are not sunk past it (where their
    // terms would hold over a hundred registers)
#pragma unroll
    for (int k = 0; k < 9; ++k) st_B(s, i, k, n, Bn[k]);
    const float pc0 = P.c;
    const float nc = pc0 + dcs;
    st_c(s, i, fmaxf(fminf(pc0, nc), 0.f));
    if (TRANSFER == 0) {
        const float nv = sqrtf(vp3[0] * vp3[0] + vp3[1] * vp3[1] + vp3[2] * vp3[2]);
        if (nv > dp.vmax) {
#pragma unroll
            for (int r = 0; r < 3; ++r) vp3[r] = vp3[r] / nv * dp.vmax;
        }
    }
    float xn[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        st_v(s, i, r, vp3[r]);
        xn[r] = P.x[r] + vp3[r] * dt;
        st_x(s, i, r, xn[r]);
    }
    float A[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) A[k] = (TRANSFER == 0) ? Bn[k] * dt * dp.Dinv : G[k] * dt;
    A[0] += 1.f;
    A[4] += 1.f;
    A[8] += 1.f;
    float F[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = ld_F(s, i, k, n);
    bool finite = true;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            // (A F)_(r,cc) with the fused multiply-adds spelled out, so that
            // the rounding does not depend on how the compiler pairs them
            const float f = fmaf(A[r + 6], F[3 * cc + 2], fmaf(A[r], F[3 * cc], A[r + 3] * F[3 * cc + 1]));
            finite = finite && isfinite(f);
            st_F(s, i, r + 3 * cc, n, f);
        }
    const float lo = 2.0f * dp.dx, hi = 1.0f - 2.0f * dp.dx;
    if (!(xn[0] >= lo && xn[0] <= hi && xn[1] >= lo && xn[1] <= hi && xn[2] >= lo && xn[2] <= hi))
        atomicOr(&dc->err, isfinite(xn[0] + xn[1] + xn[2]) ? kErrDomain : kErrNonFinite);
    if (!finite) atomicOr(&dc->err, kErrNonFinite);
    keys[i] = key_of(base_node(xn[0], dp.N) + dp.origin[0], base_node(xn[1], dp.N) + dp.origin[1],
                     base_node(xn[2], dp.N) + dp.origin[2]);
}

// the grid values G2P reads at global node gn (page * 64 + cell): v0..v2 and
// the damage increment x - c of update_grid_phase (touched nodes only)
__device__ __forceinline__ void g2p_node(const DevParams& dp, const float* __restrict__ grid,
                                         const float* __restrict__ cgx, bool breakdown, int64_t gn, float& v0,
                                         float& v1, float& v2, float& dcn) {
    const float* gp = grid + (gn >> 6) * kPageFloats + (gn & 63);
    v0 = gp[64];
    v1 = gp[128];
    v2 = gp[192];
    dcn = 0.f;
    if (dp.phase_field && gp[5 * 64] > 0.f) dcn = (breakdown ? 1e20f : cgx[gn]) - gp[4 * 64];
}

// page-tile G2P (variant 1): one virtual page at a time, its 6^3 tile in LDS
template <int BLOCK, int TRANSFER, typename ST>
__global__ __launch_bounds__(BLOCK) void k_g2p(DevParams dp, ST s, Tables t, const float* __restrict__ grid,
                                               const float* __restrict__ cgx, uint32_t* __restrict__ keys,
                                               float dt, int n, DevCounts* dc) {
    __shared__ float tv[4 * kTileN];
    dt = substep_dt(dt, dc);
    int nvp = dc->num_vp;
    nvp = nvp < t.cap_vp ? nvp : (int)t.cap_vp;
    const bool breakdown = dc->cg_breakdown != 0;
    int vb, ve;
    block_range(nvp, vb, ve);
    for (int vp = vb; vp < ve; ++vp) {
        int p, s0, e;
        vpage_range(t, vp, BLOCK, n, p, s0, e);
        for (int q = threadIdx.x; q < kTileN; q += BLOCK) {
            const int a = q / 36, b = (q / 6) % 6, cc = q % 6;
            const int64_t gn = tile_node(t, p, a, b, cc);
            float v0 = 0.f, v1 = 0.f, v2 = 0.f, dcn = 0.f;
            if (gn >= 0) g2p_node(dp, grid, cgx, breakdown, gn, v0, v1, v2, dcn);
            tv[q] = v0;
            tv[kTileN + q] = v1;
            tv[2 * kTileN + q] = v2;
            tv[3 * kTileN + q] = dcn;
        }
        __syncthreads();
        const int i = s0 + threadIdx.x;
        if (i < e) {
            int pc[3];
            unmorton(t.page_key[p], pc[0], pc[1], pc[2]);
            const int org[3] = {4 * pc[0], 4 * pc[1], 4 * pc[2]};
            G2PIn P;
            g2p_load(s, i, P);
            g2p_particle<TRANSFER, kTile>(dp, s, i, n, dt, org, tv, P, keys, dc);
        }
        __syncthreads();
    }
}

// block G2P (variant 2): a workgroup stages the 10^3 nodes that the particles
// of one 2x2x2 page block reach (the block's 8^3 nodes and two layers of the
// +x/+y/+z neighbours) and runs every particle of the block; the block's
// particles are one contiguous range because its 8 pages are consecutive
// Morton codes.  The region's grid pages and the particle range come with the
// page tables (BlockG2P).
constexpr int kBR = 10;                       // staged region edge
constexpr int kBR3 = kBR * kBR * kBR;

template <int TRANSFER, typename ST>
__global__ __launch_bounds__(256) void k_g2p_block(DevParams dp, ST s, BlockTables t,
                                                   const BlockG2P* __restrict__ info, const float* __restrict__ grid,
                                                   const float* __restrict__ cgx, uint32_t* __restrict__ keys,
                                                   float dt, int n, const SceneCG* __restrict__ scg, DevCounts* dc) {
    __shared__ float tv[4 * kBR3];
    dt = substep_dt(dt, dc);
    int nb = dc->num_blocks;
    nb = nb < t.cap_blocks ? nb : (int)t.cap_blocks;
    FOR_LOGICAL_ITEMS(item, nb) {
        const int blk = t.order ? t.order[item] : item;
        const uint32_t bkey = t.block_key[blk];
        const int scene = __builtin_amdgcn_readfirstlane(scene_of_page(dp, bkey << 3));
        const DevParams sp = dp.scenes[scene];   // loaded once per block, before the particle loop
        const bool breakdown = scg[scene].breakdown != 0;
        int bc[3];
        unmorton(bkey, bc[0], bc[1], bc[2]);
        const int tid = threadIdx.x;
        const BlockG2P* B = info + blk;
        const int pa = __builtin_amdgcn_readfirstlane(B->pa), pb = __builtin_amdgcn_readfirstlane(B->pb);
        if (pa < pb) {
            // the thread's first particle, loaded while the region is staged
            G2PIn P;
            if (pa + tid < pb) g2p_load(s, pa + tid, P);
            for (int q = tid; q < kBR3; q += 256) {
                const int X = q / 100, Y = (q / 10) % 10, Z = q % 10;
                const int G = B->page[(X >> 2) * 9 + (Y >> 2) * 3 + (Z >> 2)];
                float v0 = 0.f, v1 = 0.f, v2 = 0.f, dcn = 0.f;
                if (G >= 0)
                    g2p_node(sp, grid, cgx, breakdown, (int64_t)G * 64 + (((X & 3) << 4) | ((Y & 3) << 2) | (Z & 3)),
                             v0, v1, v2, dcn);
                tv[q] = v0;
                tv[kBR3 + q] = v1;
                tv[2 * kBR3 + q] = v2;
                tv[3 * kBR3 + q] = dcn;
            }
            __syncthreads();
            const int org[3] = {8 * bc[0] - sp.origin[0], 8 * bc[1] - sp.origin[1], 8 * bc[2] - sp.origin[2]};
            for (int i = pa + tid; i < pb; i += 256) {
                if (i != pa + tid) g2p_load(s, i, P);
                g2p_particle<TRANSFER, kBR>(sp, s, i, n, dt, org, tv, P, keys, dc);
            }
            __syncthreads();
        }
    }
}

static Tables tables_of(mpmx_ctx* c) {
    return Tables{c->page_key, c->page2particle, c->adj, c->vpo, c->target, c->cap_p, c->cap_tp, c->cap_vp};
}

// workgroups for the 2x2x2-block kernels: pct % of the last block count the
// host saw (refreshed at every PCG check) plus a margin, a multiple of 8.  The
// kernels stride over the device-side count, so a stale hint or any pct
// changes only the balance, never the results.
int block_grid(mpmx_ctx* c, int pct) {
    int64_t nb = c->hc->num_blocks;
    // no count seen yet (the first build): enough workgroups to fill the GPU,
    // which stride over the device-side count
    if (nb <= 0) nb = std::min<int64_t>(c->tp_bound / 8, (int64_t)c->num_cu * 64);
    nb = nb + nb / 8 + 64;
    nb = nb * (pct > 0 ? pct : 100) / 100;
    const int64_t cap = c->tp_bound / 8 + 8;
    nb = nb < cap ? nb : cap;
    nb = (nb + 7) / 8 * 8;
    return nb < 8 ? 8 : (int)nb;
}

int transfer_grid_blocks(mpmx_ctx* c) {
    // enough workgroups to fill every CU several times over; a multiple of 8 so
    // that the XCD-aware remap of block ids applies
    int g = c->num_cu * 8;
    const int64_t need = c->cap_vp;
    if (need < g) g = (int)((need + 7) / 8 * 8);
    return g < 8 ? 8 : g;
}

template <int BLOCK, typename ST>
static void p2g_dispatch(mpmx_ctx* c, ST st, float dt) {
    const int G = transfer_grid_blocks(c);
    const Tables t = tables_of(c);
    if (c->P.transfer == 0)
        k_p2g<BLOCK, 0, ST><<<G, BLOCK, 0, c->stream>>>(c->dp, st, c->sigma, c->FP, c->vol, t, c->grid, dt,
                                                        (int)c->n, c->dc);
    else
        k_p2g<BLOCK, 1, ST><<<G, BLOCK, 0, c->stream>>>(c->dp, st, c->sigma, c->FP, c->vol, t, c->grid, dt,
                                                        (int)c->n, c->dc);
}

template <typename ST>
static void p2g_block(mpmx_ctx* c, ST st, float dt) {
    switch (c->P.vpage) {
    case 64: p2g_dispatch<64>(c, st, dt); break;
    case 128: p2g_dispatch<128>(c, st, dt); break;
    default: p2g_dispatch<256>(c, st, dt); break;
    }
}

// scatter variant (p2g_variant 1): zero the grid, then every workgroup adds its
// pages' tiles with global atomics; the grid update runs as a separate pass
hipError_t launch_p2g_scatter(mpmx_ctx* c, float dt) {
    hipError_t e = fill_zero_dev_count(c, c->grid, &c->dc->num_total, kPageFloats, c->cap_tp);
    if (e != hipSuccess) return e;
    if (c->cfg.precision != MPMX_FP32) p2g_block(c, c->ms[c->cur], dt);
    else p2g_block(c, c->ps[c->cur], dt);
    return hipGetLastError();
}

template <int BLOCK, typename ST>
static void g2p_dispatch(mpmx_ctx* c, ST st, float dt) {
    const int G = transfer_grid_blocks(c);
    const Tables t = tables_of(c);
    if (c->P.transfer == 0)
        k_g2p<BLOCK, 0, ST><<<G, BLOCK, 0, c->stream>>>(c->dp, st, t, c->grid, c->cg_x, c->keys_next, dt, (int)c->n,
                                                        c->dc);
    else
        k_g2p<BLOCK, 1, ST><<<G, BLOCK, 0, c->stream>>>(c->dp, st, t, c->grid, c->cg_x, c->keys_next, dt, (int)c->n,
                                                        c->dc);
}
template <typename ST>
static void g2p_block(mpmx_ctx* c, ST st, float dt) {
    switch (c->P.vpage) {
    case 64: g2p_dispatch<64>(c, st, dt); break;
    case 128: g2p_dispatch<128>(c, st, dt); break;
    default: g2p_dispatch<256>(c, st, dt); break;
    }
}

template <typename ST>
static void g2p_block_kernel(mpmx_ctx* c, ST st, float dt) {
    // the blocks in the gathers' order (by gather entries; a block's own
    // particles, the G2P's work, go with them): 0.34 -> 0.29 ms on the 8M
    // impact against block order
    const BlockTables t = block_tables_of(c);
    const int g = block_grid(c, c->g2p_pct);
    MPMX_TIMED(c, "g2p_block", ({
                   if (c->P.transfer == 0)
                       k_g2p_block<0, ST><<<(int)g, 256, 0, c->stream>>>(c->dp, st, t, static_cast<const BlockG2P*>(c->block_g2p), c->grid, c->cg_x, c->keys_next, dt,
                                                                         (int)c->n, c->scg, c->dc);
                   else
                       k_g2p_block<1, ST><<<(int)g, 256, 0, c->stream>>>(c->dp, st, t, static_cast<const BlockG2P*>(c->block_g2p), c->grid, c->cg_x, c->keys_next, dt,
                                                                         (int)c->n, c->scg, c->dc);
               }));
}

hipError_t launch_g2p(mpmx_ctx* c, float dt) {
    c->vb_perm = nullptr;   // the G2P writes v and B of every particle
    if (c->cfg.p2g_variant == 1) {
        if (c->cfg.precision != MPMX_FP32) g2p_block(c, c->ms[c->cur], dt);
        else g2p_block(c, c->ps[c->cur], dt);
    } else {
        if (c->cfg.precision != MPMX_FP32) g2p_block_kernel(c, c->ms[c->cur], dt);
        else g2p_block_kernel(c, c->ps[c->cur], dt);
    }
    return hipGetLastError();
}

}  // namespace mpmx
