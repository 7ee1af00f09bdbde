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
// Gather form: a workgroup of one wavefront owns one 2x2x2 block of grid
// pages (8^3 nodes, the pages 8b .. 8b+7 in Morton order) at a time.
// Particles of page P reach the nodes of P and its +1 neighbours, so block b
// receives from the 27 particle pages 2b-1 .. 2b+1 (per axis); of a page below
// the block along an axis only the particles whose cell has index & 3 >= 2 on
// that axis reach it.  The wave walks the concatenated particle ranges of
// those pages and adds each particle's 8 channels to the block's nodes in LDS
// with plain read-add-write: LDS float atomics cost ~190 CU cycles per wave
// instruction on gfx950 against ~18 for a read-add-write, and with one wave
// per block no other wave touches the accumulator.  Lanes of one wave never
// hit the same node in the same instruction: particles of one cell (adjacent
// after the sort) are spread over rounds, and a compiler barrier between
// stencil slots keeps the in-order LDS semantics across lanes.  The wave
// then updates its nodes in place and writes them once: no atomics at all,
// no zero fill of the grid.  Each particle is read by the one to eight blocks
// it reaches (about twice on average).
#include "access.h"

namespace mpmx {

constexpr int kP2GCh = 8;   // m, p0, p1, p2, sum w c, sum w, sum V w, sum V FP w

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
};

// PCG scalars start over every substep
__global__ void k_cg_reset(DevParams dp, DevCounts* dc) {
    if (threadIdx.x != 0) return;
    dc->cg_done = dp.phase_field ? 0 : 1;
    dc->cg_iters = 0;
    dc->cg_breakdown = 0;
    dc->cg_norm = 0.f;
}

// per-particle P2G terms, computed once and kept in registers across rounds
struct P2GTerms {
    float w[3][3], gw[3][3];
    int l0[3];
    float m, cphase, V, VF;
    float ar[3], Qd[9], sg[9];
};

template <int TRANSFER, typename ST>
__device__ __forceinline__ void p2g_terms(const DevParams& dp, const ST& s, const float* __restrict__ sigma,
                                          const float* __restrict__ FP, const float* __restrict__ vol, int i,
                                          int n, float dt, const int (&org)[3], P2GTerms& T) {
    float d0[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float xa = ld_x(s, i, a);
        const int b = base_node(xa, dp.N);
        d0[a] = xa * dp.inv_dx - (float)b;
        T.l0[a] = b - org[a];
        if (TRANSFER == 0) bspline(d0[a], T.w[a]);
        else bspline_grad(d0[a], T.w[a], T.gw[a]);
    }
    T.m = ld_m(s, i);
    T.cphase = ld_c(s, i);
    T.V = vol[i];
    T.VF = T.V * FP[i];
    float v[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = ld_v(s, i, a);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float Bk = ld_B(s, i, k, n);
        const float sk = sigma[(int64_t)k * n + i];
        if (TRANSFER == 0) T.Qd[k] = (Bk * T.m - sk * dt) * dp.Dinv * dp.dx;
        else { T.Qd[k] = Bk * dp.Dinv * dp.dx * T.m; T.sg[k] = sk * dt * dp.inv_dx; }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) T.ar[r] = T.m * v[r] - (T.Qd[r] * d0[0] + T.Qd[r + 3] * d0[1] + T.Qd[r + 6] * d0[2]);
}

// add the particle's contributions to the 8^3 block accumulator (read-add-
// write; the caller guarantees that no two active lanes share a cell)
template <int TRANSFER>
__device__ __forceinline__ void p2g_rmw(const P2GTerms& T, float* acc) {
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
        const int la = T.l0[0] + ii;
        if (la < 0 || la >= 8) continue;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const int lb = T.l0[1] + jj;
            if (lb < 0 || lb >= 8) continue;
            const float wxy = T.w[0][ii] * T.w[1][jj];
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                const int lc = T.l0[2] + kk;
                if (lc < 0 || lc >= 8) continue;
                const float wt = wxy * T.w[2][kk];
                float pr[3];
#pragma unroll
                for (int r = 0; r < 3; ++r)
                    pr[r] = wt * (T.ar[r] + T.Qd[r] * ii + T.Qd[r + 3] * jj + T.Qd[r + 6] * kk);
                if (TRANSFER == 1) {
                    const float g0 = T.gw[0][ii] * T.w[1][jj] * T.w[2][kk];
                    const float g1 = T.w[0][ii] * T.gw[1][jj] * T.w[2][kk];
                    const float g2 = T.w[0][ii] * T.w[1][jj] * T.gw[2][kk];
#pragma unroll
                    for (int r = 0; r < 3; ++r) pr[r] -= T.sg[r] * g0 + T.sg[r + 3] * g1 + T.sg[r + 6] * g2;
                }
                float* a = acc + (la * 8 + lb) * kAccZ + lc;
                // all eight reads, then all eight writes: one LDS round trip
                float o[kP2GCh];
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k) o[k] = a[k * kAccStride];
                const float add[kP2GCh] = {wt * T.m, pr[0], pr[1], pr[2], wt * T.cphase, wt, wt * T.V, wt * T.VF};
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k) a[k * kAccStride] = o[k] + add[k];
                // another lane may add to this node in a later slot
                __asm__ volatile("" ::: "memory");
            }
        }
    }
}


template <int TRANSFER, typename ST>
__global__ __launch_bounds__(64) void k_p2g_block(DevParams dp, ST s, const float* __restrict__ sigma,
                                                  const float* __restrict__ FP, const float* __restrict__ vol,
                                                  BlockTables t, float* __restrict__ grid,
                                                  float* __restrict__ grid_raw, CGVec cg, float dt, int n,
                                                  const DevCounts* __restrict__ dc) {
    __shared__ float acc[kP2GCh * kAccStride];
    __shared__ BlockSources S;
    const int lane = threadIdx.x;
    const int np_raw = dc->num_page;
    const int np = np_raw < t.cap_p ? np_raw : (int)t.cap_p;
    int nt = dc->num_total;
    nt = nt < t.cap_tp ? nt : (int)t.cap_tp;
    int nb = dc->num_blocks;
    nb = nb < t.cap_blocks ? nb : (int)t.cap_blocks;
    int bb, be;
    block_range(nb, bb, be);
    for (int blk = bb; blk < be; ++blk) {
        const uint32_t bkey = t.block_key[blk];
        int bc[3];
        unmorton(bkey, bc[0], bc[1], bc[2]);
        block_sources(t, bkey, bc, n, np_raw, np, nt, S);
        for (int q = lane; q < kP2GCh * kAccStride; q += 64) acc[q] = 0.f;
        __syncthreads();
        const int total = S.pre[27];
        const int org[3] = {8 * bc[0], 8 * bc[1], 8 * bc[2]};
        for (int j0 = 0; j0 < total; j0 += 64) {
            int i, rank, rmax;
            const bool valid = block_entry(t, S, j0 + lane, total, i, rank, rmax);
            if (!__any(valid)) continue;
            P2GTerms T;
            if (valid) p2g_terms<TRANSFER>(dp, s, sigma, FP, vol, i, n, dt, org, T);
            for (int r = 0; r <= rmax; ++r) {
                if (valid && rank == r) p2g_rmw<TRANSFER>(T, acc);
                __asm__ volatile("" ::: "memory");
            }
        }
        __syncthreads();
        // grid update of the block's pages, written page-major (coalesced)
        for (int q = lane; q < 512; q += 64) {
            const int o = q >> 6, cell = q & 63;
            const int G = S.out_page[o];
            if (G < 0) continue;
            const int lx = ((o >> 2) & 1) * 4 + (cell >> 4), ly = ((o >> 1) & 1) * 4 + ((cell >> 2) & 3),
                      lz = (o & 1) * 4 + (cell & 3);
            const int node = (lx * 8 + ly) * kAccZ + lz;
            float ch[kP2GCh];
#pragma unroll
            for (int k = 0; k < kP2GCh; ++k) ch[k] = acc[k * kAccStride + node];
            if (grid_raw) {
                float* rp = grid_raw + (int64_t)G * kPageFloats + cell;
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k) rp[k * 64] = ch[k];
            }
            float r, z;
            node_update(dp, org[0] + lx, org[1] + ly, org[2] + lz, ch, dt, r, z);
            float* gp = grid + (int64_t)G * kPageFloats + cell;
#pragma unroll
            for (int k = 0; k < kP2GCh; ++k) gp[k * 64] = ch[k];
            if (dp.phase_field) {
                const int64_t nn = (int64_t)G * 64 + cell;
                cg.x[nn] = 0.f;
                cg.r[nn] = r;
                cg.z[nn] = z;
                cg.s[nn] = z;
                cg.q[nn] = 0.f;
            }
        }
        __syncthreads();
    }
}

// partials of the PCG's first "dot"(r, z): one wave per 128-node block (the
// reference's 128-entry dot blocks follow page order, not the 2x2x2 blocks)
__global__ __launch_bounds__(256) void k_cg_rho0(DevParams dp, const float* __restrict__ cgr,
                                                 const float* __restrict__ cgz, float* __restrict__ part_rho,
                                                 int64_t cap_tp, const DevCounts* __restrict__ dc) {
    __shared__ float lds[4];
    int nt = dc->num_total;
    nt = nt < cap_tp ? nt : (int)cap_tp;
    const int nblk = (nt + 1) / 2;
    const int lane = threadIdx.x & 63;
    const bool maxdot = dp.dot_mode == 0;
    float wacc = maxdot ? -INFINITY : 0.f;
    for (int blk = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; blk < nblk; blk += (gridDim.x * blockDim.x) >> 6) {
        const int n0 = blk * 128 + lane * 2;
        float rz = 0.f;
        if ((n0 >> 6) < nt) {
            const float2 r = *(const float2*)(cgr + n0);
            const float2 z = *(const float2*)(cgz + n0);
            rz = r.x * z.x + r.y * z.y;
        }
        const float bs = wave_sum(rz);
        wacc = maxdot ? fmaxf(wacc, bs) : wacc + bs;
    }
    const float r = maxdot ? combine_waves<true>(wacc, lds) : combine_waves<false>(wacc, lds);
    if (threadIdx.x == 0) part_rho[blockIdx.x] = r;
}

// the grid update after the scatter P2G: one wave per 128-node PCG block, two
// nodes per lane
__global__ __launch_bounds__(256) void k_grid_update(DevParams dp, const uint32_t* __restrict__ page_key,
                                                     float* __restrict__ grid, CGVec cg,
                                                     float* __restrict__ part_rho, float dt, int64_t cap_tp,
                                                     const DevCounts* __restrict__ dc) {
    __shared__ float lds[4];
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

hipError_t launch_p2g_scatter(mpmx_ctx* c, float dt);

hipError_t launch_cg_reset(mpmx_ctx* c) {
    k_cg_reset<<<1, 64, 0, c->stream>>>(c->dp, c->dc);
    return hipGetLastError();
}

static CGVec cg_vec(mpmx_ctx* c) { return CGVec{c->cg_x, c->cg_r, c->cg_z, c->cg_s, c->cg_q}; }

template <typename ST>
static void gather_dispatch(mpmx_ctx* c, ST st, float dt) {
    const BlockTables t = block_tables_of(c);
    float* raw = (c->keep_p2g && c->grid_p2g) ? c->grid_p2g : nullptr;
    // one wave per workgroup; ~8 resident per CU (18.4 KiB of LDS each)
    int64_t g = (c->tp_bound + 7) / 8 * 8;
    g = g < (int64_t)c->num_cu * 16 ? g : (int64_t)c->num_cu * 16;
    MPMX_TIMED(c, "p2g_block", ({
                   if (c->P.transfer == 0)
                       k_p2g_block<0, ST><<<(int)g, 64, 0, c->stream>>>(c->dp, st, c->sigma, c->FP, c->vol, t, c->grid,
                                                                         raw, cg_vec(c), dt, (int)c->n, c->dc);
                   else
                       k_p2g_block<1, ST><<<(int)g, 64, 0, c->stream>>>(c->dp, st, c->sigma, c->FP, c->vol, t, c->grid,
                                                                         raw, cg_vec(c), dt, (int)c->n, c->dc);
               }));
    if (c->dp.phase_field)
        k_cg_rho0<<<c->cg_nb, 256, 0, c->stream>>>(c->dp, c->cg_r, c->cg_z, c->cg_part + c->cg_nb, c->cap_tp,
                                                    c->dc);
}

// P2G; the gather variant (default) also runs the grid update and PCG set-up
hipError_t launch_p2g(mpmx_ctx* c, float dt) {
    hipError_t e = launch_cg_reset(c);
    if (e != hipSuccess) return e;
    if (c->cfg.p2g_variant == 1) {
        c->grid_fused = false;
        return launch_p2g_scatter(c, dt);
    }
    if (c->cfg.precision == MPMX_MIXED) gather_dispatch(c, c->ms[c->cur], dt);
    else gather_dispatch(c, c->ps[c->cur], dt);
    c->grid_fused = true;
    return hipGetLastError();
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
