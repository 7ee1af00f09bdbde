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
// the particles whose cell has index & 3 >= 2 on that axis reach it.  All 256
// threads walk the concatenated particle ranges of those pages, accumulate the
// 8 channels of the 512 nodes in LDS with LDS float atomics, and the
// workgroup then updates its nodes in place and writes them once: no global
// atomics and no zero fill of the grid.  Each particle is read by the one to
// eight blocks it reaches (about twice on average).
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

// contributions of particle i to the R^3 nodes starting at node org (node
// coordinates): the MLS / APIC momentum with the stress term, mass, sum w c,
// sum w, sum V w and sum V FP w, added to acc[ch R^3 + (a R + b) R + c] with
// LDS atomics
template <int R, int TRANSFER, typename ST>
__device__ __forceinline__ void p2g_particle(const DevParams& dp, const ST& s, const float* __restrict__ sigma,
                                             const float* __restrict__ FP, const float* __restrict__ vol, int i,
                                             int n, float dt, const int (&org)[3], float* acc) {
    constexpr int R3 = R * R * R;
    float d0[3], w[3][3], gw[3][3];
    int l0[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float xa = ld_x(s, i, a);
        const int b = base_node(xa, dp.N);
        d0[a] = xa * dp.inv_dx - (float)b;
        l0[a] = b - org[a];
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
    for (int r = 0; r < 3; ++r) ar[r] = m * v[r] - (Qd[r] * d0[0] + Qd[r + 3] * d0[1] + Qd[r + 6] * d0[2]);
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
        const int la = l0[0] + ii;
        if (la < 0 || la >= R) continue;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const int lb = l0[1] + jj;
            if (lb < 0 || lb >= R) continue;
            const float wxy = w[0][ii] * w[1][jj];
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                const int lc = l0[2] + kk;
                if (lc < 0 || lc >= R) continue;
                const float wt = wxy * w[2][kk];
                float pr[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) pr[r] = wt * (ar[r] + Qd[r] * ii + Qd[r + 3] * jj + Qd[r + 6] * kk);
                if (TRANSFER == 1) {
                    // + dt * (-sigma grad w), gradient in units of 1/dx
                    const float g0 = gw[0][ii] * w[1][jj] * w[2][kk];
                    const float g1 = w[0][ii] * gw[1][jj] * w[2][kk];
                    const float g2 = w[0][ii] * w[1][jj] * gw[2][kk];
#pragma unroll
                    for (int r = 0; r < 3; ++r) pr[r] -= sg[r] * g0 + sg[r + 3] * g1 + sg[r + 6] * g2;
                }
                const int node = (la * R + lb) * R + lc;
                atomicAdd(&acc[0 * R3 + node], wt * m);
                atomicAdd(&acc[1 * R3 + node], pr[0]);
                atomicAdd(&acc[2 * R3 + node], pr[1]);
                atomicAdd(&acc[3 * R3 + node], pr[2]);
                atomicAdd(&acc[4 * R3 + node], wt * cphase);
                atomicAdd(&acc[5 * R3 + node], wt);
                atomicAdd(&acc[6 * R3 + node], wt * V);
                atomicAdd(&acc[7 * R3 + node], wt * VF);
            }
        }
    }
}

struct BlockTables {
    const uint64_t* pmask;
    const uint64_t* nmask;
    const int* wpre;           // page rank prefix sums (particle, neighbour-only)
    const int* spre;
    const int* page2particle;
    const uint32_t* keys;      // sorted particle keys
    const uint32_t* block_key;
    int64_t cap_p, cap_tp, cap_blocks;
    int axis_pages;            // page coordinates lie in [0, axis_pages)
};

template <int TRANSFER, typename ST>
__global__ __launch_bounds__(256) void k_p2g_block(DevParams dp, ST s, const float* __restrict__ sigma,
                                                   const float* __restrict__ FP, const float* __restrict__ vol,
                                                   BlockTables t, float* __restrict__ grid,
                                                   float* __restrict__ grid_raw, CGVec cg, float dt, int n,
                                                   const DevCounts* __restrict__ dc) {
    __shared__ float acc[kP2GCh * 512];
    __shared__ int src_start[27], src_pre[28], src_cnt[27];
    __shared__ int src_filter[27];
    __shared__ int out_page[8];
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
        if (threadIdx.x < 27) {
            const int q = threadIdx.x;
            const int dx = q / 9 - 1, dy = (q / 3) % 3 - 1, dz = q % 3 - 1;
            const int px = 2 * bc[0] + dx, py = 2 * bc[1] + dy, pz = 2 * bc[2] + dz;
            int st = 0, cnt = 0;
            if (px >= 0 && py >= 0 && pz >= 0 && px < t.axis_pages && py < t.axis_pages && pz < t.axis_pages) {
                const int id = page_rank(morton((uint32_t)px, (uint32_t)py, (uint32_t)pz), t.pmask, t.nmask, t.wpre,
                                         t.spre, np_raw);
                if (id >= 0 && id < np) {
                    int a = t.page2particle[id], b = t.page2particle[id + 1];
                    a = a < 0 ? 0 : (a > n ? n : a);
                    b = b < a ? a : (b > n ? n : b);
                    st = a;
                    cnt = b - a;
                }
            }
            src_start[q] = st;
            src_cnt[q] = cnt;
            src_filter[q] = (dx < 0 ? 1 : 0) | (dy < 0 ? 2 : 0) | (dz < 0 ? 4 : 0);
        } else if (threadIdx.x >= 64 && threadIdx.x < 72) {
            const int o = threadIdx.x - 64;
            const int id = page_rank((bkey << 3) | (uint32_t)o, t.pmask, t.nmask, t.wpre, t.spre, np_raw);
            out_page[o] = (id >= 0 && id < nt) ? id : -1;
        }
        for (int q = threadIdx.x; q < kP2GCh * 512; q += 256) acc[q] = 0.f;
        __syncthreads();
        if (threadIdx.x == 0) {
            int run = 0;
            for (int q = 0; q < 27; ++q) {
                src_pre[q] = run;
                run += src_cnt[q];
            }
            src_pre[27] = run;
        }
        __syncthreads();
        const int total = src_pre[27];
        const int org[3] = {8 * bc[0], 8 * bc[1], 8 * bc[2]};
        for (int j = threadIdx.x; j < total; j += 256) {
            int lo = 0, hi = 26;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (src_pre[mid] <= j) lo = mid;
                else hi = mid - 1;
            }
            const int i = src_start[lo] + (j - src_pre[lo]);
            const int f = src_filter[lo];
            if (f) {
                const uint32_t cell = t.keys[i] & 63u;
                if (((f & 1) && !((cell >> 5) & 1u)) || ((f & 2) && !((cell >> 3) & 1u)) ||
                    ((f & 4) && !((cell >> 1) & 1u)))
                    continue;
            }
            p2g_particle<8, TRANSFER>(dp, s, sigma, FP, vol, i, n, dt, org, acc);
        }
        __syncthreads();
        // grid update of the block's pages, written page-major (coalesced)
        for (int q = threadIdx.x; q < 512; q += 256) {
            const int o = q >> 6, cell = q & 63;
            const int G = out_page[o];
            if (G < 0) continue;
            const int lx = ((o >> 2) & 1) * 4 + (cell >> 4), ly = ((o >> 1) & 1) * 4 + ((cell >> 2) & 3),
                      lz = (o & 1) * 4 + (cell & 3);
            const int node = (lx * 8 + ly) * 8 + lz;
            float ch[kP2GCh];
#pragma unroll
            for (int k = 0; k < kP2GCh; ++k) ch[k] = acc[k * 512 + node];
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
    const BlockTables t{c->pmask, c->nmask, c->wcount, c->scount, c->page2particle, c->keys, c->block_key,
                        c->cap_p, c->cap_tp, c->tp_bound, 1 << c->axis_bits};
    float* raw = (c->keep_p2g && c->grid_p2g) ? c->grid_p2g : nullptr;
    int64_t g = (c->tp_bound + 7) / 8 * 8;
    g = g < (int64_t)c->num_cu * 8 ? g : (int64_t)c->num_cu * 8;
    if (c->P.transfer == 0)
        k_p2g_block<0, ST><<<(int)g, 256, 0, c->stream>>>(c->dp, st, c->sigma, c->FP, c->vol, t, c->grid, raw,
                                                          cg_vec(c), dt, (int)c->n, c->dc);
    else
        k_p2g_block<1, ST><<<(int)g, 256, 0, c->stream>>>(c->dp, st, c->sigma, c->FP, c->vol, t, c->grid, raw,
                                                          cg_vec(c), dt, (int)c->n, c->dc);
    if (c->P.phase_field)
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
