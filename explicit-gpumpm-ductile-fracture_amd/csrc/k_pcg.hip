// k_pcg.hip — the phase-field solve: Jacobi-preconditioned CG on the grid,
// (D + M K) x = b, with D_i = sum_p V w (FP M + 1/dt) and
// (K s)_i = sum_p grad w_ip . (4 L0^2 V_p sum_j grad w_jp s_j), L0 = dx/2.
//
// Same iteration as the reference driver loop (ExplicitTimeIntegrator.cu:914-968)
// with its reductions: the "dot product" is the maximum over 128-entry blocks
// of the block sums and the norm is max|r| (dot_mode 1 gives true inner
// products).  Differences of form, not of maths:
//  - AxG2P_APIC and AxP2G_APIC (G2PKernal.cu:1006-1123, P2GKernal.cu:1162-1326)
//    are one kernel: a workgroup gathers grad w . s for its particles from the
//    staged s tile and scatters M grad w . col back, with no round trip of the
//    per-particle vector col through HBM;
//  - the diagonal term s_i V w (FP M + 1/dt) summed over particles is s_i D_i,
//    applied per node instead of per particle;
//  - D is computed once per substep (the reference rebuilds it every
//    iteration, preConditionP2G_APIC);
//  - scalars never leave the device: each node kernel writes one partial per
//    workgroup (max or sum of its block sums, max|r|) and the next kernel
//    reduces the partials in a fixed order, so no same-address atomics and a
//    bit-reproducible solve; the convergence test is evaluated identically by
//    every workgroup, and every kernel returns at once after convergence, so a
//    fixed number of iterations can be queued without host round trips;
//  - a breakdown (rho = 0 before convergence) sets x = 1e20 for G2P as the
//    reference does (:959-962), which leaves c unchanged that substep; it is
//    counted in the statistics, not raised as an error.
//
// Partials: cg_part[0] s.q, cg_part[1 + (k & 1)] r.z entering iteration k,
// cg_part[3] max|r|, each cg_nb floats.  Every scene of a batch runs its own
// solve: its own 128-node blocks, partials, scalars and stopping test
// (SceneBlocks); the batch iterates until every scene has stopped.
#include "access.h"
#include <chrono>
#include <cstring>

namespace mpmx {

// the PCG has stopped (converged, broken down): read once per workgroup so
// that every thread of it takes the same branch
__device__ __forceinline__ bool cg_stopped(const DevCounts* dc) {
    __shared__ int sdone;
    if (threadIdx.x == 0) sdone = dc->cg_done;
    __syncthreads();
    return sdone != 0;
}

// "dot" of the partials: max of block sums (dot_mode 0) or their sum
__device__ __forceinline__ float reduce_dot(const DevParams& dp, const float* parts, int cnt, float* lds) {
    return dp.dot_mode == 0 ? reduce_parts<true>(parts, cnt, lds) : reduce_parts<false>(parts, cnt, lds);
}
__device__ __forceinline__ float combine_dot(const DevParams& dp, float v, float* lds) {
    return dp.dot_mode == 0 ? combine_waves<true>(v, lds) : combine_waves<false>(v, lds);
}

// q += M grad w . col for every particle, col = V dx^2 sum_j grad w_j s_j
template <int BLOCK, typename ST>
__global__ __launch_bounds__(BLOCK) void k_cg_apply(DevParams dp, ST s, const float* __restrict__ vol, Tables t,
                                                    const float* __restrict__ cgs, float* __restrict__ cgq,
                                                    int n, DevCounts* dc) {
    if (cg_stopped(dc)) return;
    __shared__ float ts[kTileN];
    __shared__ float tq[kTileN];
    int nvp = dc->num_vp;
    nvp = nvp < t.cap_vp ? nvp : (int)t.cap_vp;
    int vb, ve;
    block_range(nvp, vb, ve);
    const float para_scale = 4.0f * dp.L0 * dp.L0 * dp.inv_dx * dp.inv_dx;  // dx^2 / dx^2 with grads in 1/dx
    for (int vp = vb; vp < ve; ++vp) {
        int p, s0, e;
        vpage_range(t, vp, BLOCK, n, p, s0, e);
        for (int q = threadIdx.x; q < kTileN; q += BLOCK) {
            const int a = q / 36, b = (q / 6) % 6, cc = q % 6;
            const int64_t gn = tile_node(t, p, a, b, cc);
            ts[q] = gn >= 0 ? cgs[gn] : 0.f;
            tq[q] = 0.f;
        }
        __syncthreads();
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
                bspline_grad(d0[a], w[a], gw[a]);
            }
            float gs[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) {
                        const float sv = ts[(l[0] + ii) * 36 + (l[1] + jj) * 6 + (l[2] + kk)];
                        gs[0] += gw[0][ii] * w[1][jj] * w[2][kk] * sv;
                        gs[1] += w[0][ii] * gw[1][jj] * w[2][kk] * sv;
                        gs[2] += w[0][ii] * w[1][jj] * gw[2][kk] * sv;
                    }
            // col in units where grad w is per cell: col = V * 4 L0^2 * grad(s) ; scatter M grad w . col
            const float sc = vol[i] * para_scale * dp.M;
            const float c0 = gs[0] * sc, c1 = gs[1] * sc, c2 = gs[2] * sc;
#pragma unroll
            for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) {
                        const float hw = gw[0][ii] * w[1][jj] * w[2][kk] * c0 + w[0][ii] * gw[1][jj] * w[2][kk] * c1 +
                                         w[0][ii] * w[1][jj] * gw[2][kk] * c2;
                        atomicAdd(&tq[(l[0] + ii) * 36 + (l[1] + jj) * 6 + (l[2] + kk)], hw);
                    }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < kTileN; q += BLOCK) {
            const float val = tq[q];
            if (val != 0.f) {
                const int a = q / 36, b = (q / 6) % 6, cc = q % 6;
                const int64_t gn = tile_node(t, p, a, b, cc);
                if (gn >= 0) atomicAdd(&cgq[gn], val);
            }
        }
        __syncthreads();
    }
}

// ---- block form of (K s) (variant 2) ----------------------------------------
// col_p = M V dx^2 sum_j grad w_jp s_j for every particle of a 2x2x2 page
// block, from the block's 10^3-node s region staged in LDS; with it, the
// largest |col_p|_1 of each of the block's pages (cbound, float bits), from
// which k_cg_q takes its fixed-point scale
template <typename ST>
__global__ __launch_bounds__(256) void k_cg_col(DevParams dp, ST s, const float* __restrict__ vol, BlockTables t,
                                                const float* __restrict__ cgs, float* __restrict__ col,
                                                unsigned* __restrict__ cbound, int n,
                                                const SceneCG* __restrict__ scg, DevCounts* dc) {
    constexpr int R = 10, R3 = R * R * R;
    __shared__ float ts[R3];
    __shared__ int spg[27];
    __shared__ int prange[2];
    __shared__ int opg[8];          // page ids of the block's own pages (-1: none)
    __shared__ unsigned omax[8];    // their largest |col|_1 (float bits)
    if (cg_stopped(dc)) return;
    const int np_raw = dc->num_page;
    const int np = np_raw < t.cap_p ? np_raw : (int)t.cap_p;
    int nt = dc->num_total;
    nt = nt < t.cap_tp ? nt : (int)t.cap_tp;
    int nb = dc->num_blocks;
    nb = nb < t.cap_blocks ? nb : (int)t.cap_blocks;
    const int tid = threadIdx.x;
    FOR_LOGICAL_ITEMS(item, nb) {
        const int blk = t.order ? t.order[item] : item;
        const uint32_t bkey = t.block_key[blk];
        const int scene = __builtin_amdgcn_readfirstlane(scene_of_page(dp, bkey << 3));
        if (scg[scene].done) continue;   // uniform over the workgroup
        const DevParams sp = dp.scenes[scene];
        const float para_scale = 4.0f * sp.L0 * sp.L0 * sp.inv_dx * sp.inv_dx;
        int bc[3];
        unmorton(bkey, bc[0], bc[1], bc[2]);
        if (tid < 27) {
            const int px = 2 * bc[0] + tid / 9, py = 2 * bc[1] + (tid / 3) % 3, pz = 2 * bc[2] + tid % 3;
            int id = -1;
            const uint32_t key = scene_page_key(t, px, py, pz, scene);
            if (key != ~0u) id = page_rank(key, t.pmask, t.nmask, t.wpre, t.spre, np_raw);
            spg[tid] = (id >= 0 && id < nt) ? id : -1;
        }
        if (tid >= 64 && tid < 128) {
            const int o = tid - 64;
            int id = -1;
            if (o < 8) id = page_rank((bkey << 3) | (uint32_t)o, t.pmask, t.nmask, t.wpre, t.spre, np_raw);
            const bool own = id >= 0 && id < np;
            // opg and omax are written and read by these 8 threads only
            if (o < 8) {
                opg[o] = own ? id : -1;
                omax[o] = 0u;
            }
            int lo = own ? id : 0x7fffffff, hi = own ? id : -1;
#pragma unroll
            for (int w = 32; w > 0; w >>= 1) {
                lo = min(lo, __shfl_xor(lo, w, 64));
                hi = max(hi, __shfl_xor(hi, w, 64));
            }
            if (o == 0) {
                int a = 0, b = 0;
                if (hi >= 0) {
                    a = t.page2particle[lo];
                    b = t.page2particle[hi + 1];
                    a = a < 0 ? 0 : (a > n ? n : a);
                    b = b < a ? a : (b > n ? n : b);
                }
                prange[0] = a;
                prange[1] = b;
            }
        }
        __syncthreads();
        const int pa = prange[0], pb = prange[1];
        if (pa < pb) {
            for (int q = tid; q < R3; q += 256) {
                const int X = q / 100, Y = (q / 10) % 10, Z = q % 10;
                const int G = spg[(X >> 2) * 9 + (Y >> 2) * 3 + (Z >> 2)];
                ts[q] = G >= 0 ? cgs[(int64_t)G * 64 + (((X & 3) << 4) | ((Y & 3) << 2) | (Z & 3))] : 0.f;
            }
            __syncthreads();
            // the region's first node in scene-local cells
            const int org[3] = {8 * bc[0] - sp.origin[0], 8 * bc[1] - sp.origin[1], 8 * bc[2] - sp.origin[2]};
            for (int i = pa + tid; i < pb; i += 256) {
                float d0[3], w[3][3], gw[3][3];
                int l[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float xa = ld_x(s, i, a);
                    const int b = base_node(xa, sp.N);
                    d0[a] = xa * sp.inv_dx - (float)b;
                    const int la = b - org[a];
                    l[a] = la < 0 ? 0 : (la > R - 3 ? R - 3 : la);
                    bspline_grad(d0[a], w[a], gw[a]);
                }
                float gs[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                        for (int kk = 0; kk < 3; ++kk) {
                            const float sv = ts[((l[0] + ii) * R + (l[1] + jj)) * R + (l[2] + kk)];
                            gs[0] += gw[0][ii] * w[1][jj] * w[2][kk] * sv;
                            gs[1] += w[0][ii] * gw[1][jj] * w[2][kk] * sv;
                            gs[2] += w[0][ii] * w[1][jj] * gw[2][kk] * sv;
                        }
                const float sc = vol[i] * para_scale * sp.M;
                const float c0 = gs[0] * sc, c1 = gs[1] * sc, c2 = gs[2] * sc;
                col[3 * (int64_t)i + 0] = c0;
                col[3 * (int64_t)i + 1] = c1;
                col[3 * (int64_t)i + 2] = c2;
                // the particle's page is child (key >> 6) & 7 of the block
                atomicMax(&omax[(t.keys[i] >> 6) & 7u], __float_as_uint(fabsf(c0) + fabsf(c1) + fabsf(c2)));
            }
        }
        __syncthreads();
        if (tid >= 64 && tid < 72 && opg[tid - 64] >= 0) cbound[opg[tid - 64]] = omax[tid - 64];
    }
}

// particle of entry e (0 <= e < S.pre[kMaxRuns]) of a block's source runs,
// all classes in order: the last run that starts at or before e
__device__ __forceinline__ int block_entry(const BlockClasses& S, int e) {
    int lo = 0, hi = kMaxRuns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S.pre[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return S.start[lo] + (e - S.pre[lo]);
}

// (K s) contributions of one particle at xp with column cs (col times the
// block's power-of-two scale, exact): grad w_ip . col_p into the block's
// nodes in fixed point; nodes outside the block are skipped.  grad w . col is
// w_z (gw_x w_y c0 + w_x gw_y c1) + gw_z (w_x w_y c2), whose two brackets are
// shared by the three nodes of a row.
__device__ __forceinline__ void cgq_entry(const DevParams& sp, const float (&xp)[3], const float (&cs)[3],
                                          const int (&org)[3], unsigned long long* acc) {
    float w[3][3], gw[3][3];
    int l0[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int b = base_node(xp[a], sp.N);
        const float d0 = xp[a] * sp.inv_dx - (float)b;
        l0[a] = b - org[a];
        bspline_grad(d0, w[a], gw[a]);
    }
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
        const int la = l0[0] + ii;
        if (la < 0 || la >= 8) continue;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const int lb = l0[1] + jj;
            if (lb < 0 || lb >= 8) continue;
            const float rA = fmaf(gw[0][ii] * w[1][jj], cs[0], w[0][ii] * gw[1][jj] * cs[1]);
            const float rB = w[0][ii] * w[1][jj] * cs[2];
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                const int lc = l0[2] + kk;
                if (lc < 0 || lc >= 8) continue;
                const float hw = fmaf(w[2][kk], rA, gw[2][kk] * rB);
                atomicAdd(&acc[acc_node(la, lb, lc)], to_fix_scaled(hw));
            }
        }
    }
}

// entries staged in LDS per round, and the staged words per component (one
// pad word every 32, so that the runs of neighbouring lanes start in
// different banks)
constexpr int kQChunk = 1024;
constexpr int kQSkew = kQChunk + kQChunk / 32;
__device__ __forceinline__ int q_slot(int q) { return q + (q >> 5); }

// q_i = sum_p grad w_ip . col_p + D_i s_i over the 8^3 nodes of a block, as a
// gather with 64-bit fixed-point LDS atomics; q is written whole (no zero
// fill, no global atomics).  The block's source lists are the ones the block
// P2G reads (built with the page tables, same block numbering).  The scale
// comes from the largest |col|_1 of the 27 source pages (cbound, k_cg_col),
// which bounds |grad w . col| for every listed particle.  The listed
// particles are read in rounds of kQChunk: consecutive lanes load
// consecutive entries (runs of adjacent particles, whole cache lines) into
// LDS, then each thread takes a run of adjacent staged entries, so the lanes
// of a wave hold particles of different cells when they add into the nodes.
template <typename ST>
__global__ __launch_bounds__(256) void k_cg_q(DevParams dp, ST s, BlockTables t,
                                              const BlockClasses* __restrict__ lists, const float* __restrict__ col,
                                              const unsigned* __restrict__ cbound, const float* __restrict__ grid,
                                              const float* __restrict__ cgs, float* __restrict__ cgq, int n,
                                              const SceneCG* __restrict__ scg, DevCounts* dc) {
    __shared__ unsigned long long acc[kAccStride];
    __shared__ BlockClasses S;
    __shared__ float stg[6 * kQSkew];   // x (3 components), then col (3)
    __shared__ int sexp;
    if (cg_stopped(dc)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int nb = dc->num_blocks;
    nb = nb < t.cap_blocks ? nb : (int)t.cap_blocks;
    (void)n;
    FOR_LOGICAL_ITEMS(item, nb) {
        const int blk = t.order ? t.order[item] : item;
        const uint32_t bkey = t.block_key[blk];
        const int scene = __builtin_amdgcn_readfirstlane(scene_of_page(dp, bkey << 3));
        if (scg[scene].done) continue;   // uniform over the workgroup
        const DevParams sp = dp.scenes[scene];
        int bc[3];
        unmorton(bkey, bc[0], bc[1], bc[2]);
        classes_load(lists + blk, S);
        for (int q = tid; q < kAccStride; q += 256) acc[q] = 0ull;
        if (wave == 0) {
            const int pg = lane < 27 ? lists[blk].page[lane] : -1;
            const float bm = wave_max(pg >= 0 ? __uint_as_float(cbound[pg]) : 0.f);
            if (lane == 0) sexp = fix_exp(bm);
        }
        __syncthreads();
        const int e = sexp;
        const float sc = ldexpf(1.0f, e);
        // the block's first node in scene-local cells
        const int org[3] = {8 * bc[0] - sp.origin[0], 8 * bc[1] - sp.origin[1], 8 * bc[2] - sp.origin[2]};
        const int total = S.pre[kMaxRuns];
        for (int c0 = 0; c0 < total; c0 += kQChunk) {
            const int cnt = total - c0 < kQChunk ? total - c0 : kQChunk;
            for (int q = tid; q < cnt; q += 256) {
                const int i = block_entry(S, c0 + q);
                const int k = q_slot(q);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    stg[a * kQSkew + k] = ld_x(s, i, a);
                    stg[(3 + a) * kQSkew + k] = col[3 * (int64_t)i + a];
                }
            }
            __syncthreads();
            const int per = (cnt + 255) >> 8;
            const int qb = tid * per, qe = qb + per < cnt ? qb + per : cnt;
            for (int q = qb; q < qe; ++q) {
                const int k = q_slot(q);
                const float xp[3] = {stg[k], stg[kQSkew + k], stg[2 * kQSkew + k]};
                const float cs[3] = {stg[3 * kQSkew + k] * sc, stg[4 * kQSkew + k] * sc, stg[5 * kQSkew + k] * sc};
                cgq_entry(sp, xp, cs, org, acc);
            }
            __syncthreads();
        }
        for (int q = tid; q < 512; q += 256) {
            const int o = q >> 6, cell = q & 63;
            const int G = S.out_page[o];
            if (G < 0) continue;
            const int lx = ((o >> 2) & 1) * 4 + (cell >> 4), ly = ((o >> 1) & 1) * 4 + ((cell >> 2) & 3),
                      lz = (o & 1) * 4 + (cell & 3);
            const int64_t nn = (int64_t)G * 64 + cell;
            const float D = grid[(int64_t)G * kPageFloats + 7 * 64 + cell];
            cgq[nn] = from_fix(acc[acc_node(lx, ly, lz)], e) + cgs[nn] * D;
        }
        __syncthreads();
    }
}

// the workgroup's scene has stopped (converged, broken down): uniform branch
__device__ __forceinline__ bool scene_stopped(const SceneCG* scg, int scene) {
    __shared__ int sdone;
    if (threadIdx.x == 0) sdone = scg[scene].done;
    __syncthreads();
    return sdone != 0;
}

// q = q + s D (page-tile variant; the block variant adds s D in k_cg_q);
// partial of "dot"(s, q).  With rho0 set (first iteration of the block
// variant) also the partial of the starting "dot"(r, z) into part_rho.  The
// fused multiply-adds are spelled out so that the rounding does not depend
// on how the compiler pairs the products.
__global__ __launch_bounds__(256) void k_cg_dot(DevParams dp, const float* __restrict__ grid,
                                                const float* __restrict__ cgs, float* __restrict__ cgq,
                                                float* __restrict__ part_sq, int add_diag,
                                                const float* __restrict__ cgr, const float* __restrict__ cgz,
                                                float* __restrict__ part_rho, int rho0, int64_t cap_tp,
                                                const SceneCG* __restrict__ scg, const int* __restrict__ wg_scene,
                                                const DevCounts* __restrict__ dc) {
    __shared__ float lds[4];
    const SceneBlocks sb = scene_blocks(scg, wg_scene);
    if (scene_stopped(scg, sb.scene)) return;
    const int num_page = dc->num_page;
    const bool maxdot = dp.dot_mode == 0;
    float wacc = maxdot ? -INFINITY : 0.f, racc = wacc;
    for (int blk = sb.wave; blk < sb.nblk; blk += sb.nwaves) {
        const int64_t n0 = scene_block_node(sb, num_page, cap_tp, blk);
        float part = 0.f, rz = 0.f;
        if (n0 >= 0) {
            const float2 sv = *(const float2*)(cgs + n0);
            float2 qv = *(const float2*)(cgq + n0);
            if (add_diag) {
                const float* D = grid + (n0 >> 6) * kPageFloats + 7 * 64 + (n0 & 63);
                qv.x += sv.x * D[0];
                qv.y += sv.y * D[1];
                *(float2*)(cgq + n0) = qv;
            }
            part = fmaf(sv.y, qv.y, sv.x * qv.x);
            if (rho0) {
                const float2 r = *(const float2*)(cgr + n0);
                const float2 z = *(const float2*)(cgz + n0);
                rz = fmaf(r.x, z.x, r.y * z.y);
            }
        }
        const float bs = wave_sum(part);
        wacc = maxdot ? fmaxf(wacc, bs) : wacc + bs;
        if (rho0) {
            const float br = wave_sum(rz);
            racc = maxdot ? fmaxf(racc, br) : racc + br;
        }
    }
    const float r = combine_dot(dp, wacc, lds);
    if (threadIdx.x == 0) part_sq[blockIdx.x] = r;
    if (rho0) {
        const float rr = combine_dot(dp, racc, lds);
        if (threadIdx.x == 0) part_rho[blockIdx.x] = rr;
    }
}

// alpha = rho_old / sq; x += alpha s; r -= alpha q; q = 0; z = r / D where
// |D| >= 1e-30 (else z keeps its value, as precondition_subd); partials of
// max|r| and "dot"(r, z).  Only nodes touched by particles (sum w > 0) change.
__global__ __launch_bounds__(256) void k_cg_update(DevParams dp, const float* __restrict__ grid,
                                                   float* __restrict__ cgx, float* __restrict__ cgr,
                                                   float* __restrict__ cgz, const float* __restrict__ cgs,
                                                   float* __restrict__ cgq, float* __restrict__ part, int nb, int k,
                                                   int64_t cap_tp, const SceneCG* __restrict__ scg,
                                                   const int* __restrict__ wg_scene,
                                                   const uint64_t* __restrict__ amask, int zero_q,
                                                   const DevCounts* __restrict__ dc) {
    __shared__ float lds[4];
    const SceneBlocks sb = scene_blocks(scg, wg_scene);
    if (scene_stopped(scg, sb.scene)) return;
    const float sq = reduce_dot(dp, part + sb.wg0, sb.nwg, lds);
    const float rho_old = reduce_dot(dp, part + (1 + (k & 1)) * nb + sb.wg0, sb.nwg, lds);
    const float alpha = rho_old / sq;
    const int num_page = dc->num_page;
    const bool maxdot = dp.dot_mode == 0;
    float wacc = maxdot ? -INFINITY : 0.f, wmax = 0.f;
    for (int blk = sb.wave; blk < sb.nblk; blk += sb.nwaves) {
        const int64_t n0 = scene_block_node(sb, num_page, cap_tp, blk);
        float rz = 0.f, rmax = 0.f;
        if (n0 >= 0) {
            const float* gp = grid + (n0 >> 6) * kPageFloats + (n0 & 63);
            const uint64_t am = amask ? amask[n0 >> 6] >> (n0 & 63) : 0ull;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int64_t nn = n0 + h;
                if (amask ? ((am >> h) & 1ull) != 0ull : gp[5 * 64 + h] > 0.f) {
                    const float D = gp[7 * 64 + h];
                    cgx[nn] += cgs[nn] * alpha;
                    const float r = cgr[nn] - cgq[nn] * alpha;
                    cgr[nn] = r;
                    float z;
                    if (!(fabsf(D) < 1e-30f)) {
                        z = r / D;
                        cgz[nn] = z;
                    } else {
                        z = cgz[nn];   // read only where it is kept
                    }
                    rmax = fmaxf(rmax, fabsf(r));
                    rz += r * z;
                }
                if (zero_q) cgq[nn] = 0.f;   // the page-tile K s adds into q; the block one writes it whole
            }
        }
        const float bs = wave_sum(rz);
        wacc = maxdot ? fmaxf(wacc, bs) : wacc + bs;
        wmax = fmaxf(wmax, wave_max(rmax));
    }
    const float rho = combine_dot(dp, wacc, lds);
    const float norm = combine_waves<true>(wmax, lds);
    if (threadIdx.x == 0) {
        part[(1 + ((k + 1) & 1)) * nb + blockIdx.x] = rho;
        part[3 * nb + blockIdx.x] = norm;
    }
}

// convergence / breakdown test of the workgroup's scene, then s = s beta + z.
// The scene's first workgroup records the scene's iteration count and norm
// and, when it stops, takes it off the batch's active count; the last scene
// to stop sets the batch flag that turns every later PCG launch into a no-op.
__global__ __launch_bounds__(256) void k_cg_direction(DevParams dp, const float* __restrict__ grid,
                                                      const float* __restrict__ cgz, float* __restrict__ cgs,
                                                      const float* __restrict__ part, int nb, int k,
                                                      int64_t cap_tp, SceneCG* __restrict__ scg,
                                                      const int* __restrict__ wg_scene,
                                                      const uint64_t* __restrict__ amask, DevCounts* dc) {
    __shared__ float lds[4];
    const SceneBlocks sb = scene_blocks(scg, wg_scene);
    if (scene_stopped(scg, sb.scene)) return;
    const float norm = reduce_parts<true>(part + 3 * nb + sb.wg0, sb.nwg, lds);
    const float rho_old = reduce_dot(dp, part + (1 + (k & 1)) * nb + sb.wg0, sb.nwg, lds);
    const float rho = reduce_dot(dp, part + (1 + ((k + 1) & 1)) * nb + sb.wg0, sb.nwg, lds);
    const bool converged = norm < dp.scenes[sb.scene].cg_tol;
    const bool stop = converged || rho_old == 0.f;
    if ((int)blockIdx.x == sb.wg0 && threadIdx.x == 0) {
        scg[sb.scene].iters = k + 1;
        scg[sb.scene].norm = norm;
        atomicMax(&dc->cg_iters, k + 1);
        if (stop) {
            scg[sb.scene].done = 1;
            if (!converged) {
                scg[sb.scene].breakdown = 1;
                atomicAdd(&dc->cg_breakdown, 1);
            }
            if (atomicSub(&dc->cg_active, 1) == 1) dc->cg_done = 1;
        }
    }
    if (stop) return;
    const float beta = rho / rho_old;
    const int num_page = dc->num_page;
    for (int blk = sb.wave; blk < sb.nblk; blk += sb.nwaves) {
        const int64_t n0 = scene_block_node(sb, num_page, cap_tp, blk);
        if (n0 >= 0) {
            const float* gp = grid + (n0 >> 6) * kPageFloats + (n0 & 63);
            const uint64_t am = amask ? amask[n0 >> 6] >> (n0 & 63) : 0ull;
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (amask ? ((am >> h) & 1ull) != 0ull : gp[5 * 64 + h] > 0.f)
                    cgs[n0 + h] = cgs[n0 + h] * beta + cgz[n0 + h];
        }
    }
}

template <int BLOCK, typename ST>
static void apply_dispatch(mpmx_ctx* c, ST st) {
    const int G = transfer_grid_blocks(c);
    const Tables t{c->page_key, c->page2particle, c->adj, c->vpo, c->target, c->cap_p, c->cap_tp, c->cap_vp};
    k_cg_apply<BLOCK, ST><<<G, BLOCK, 0, c->stream>>>(c->dp, st, c->vol, t, c->cg_s, c->cg_q, (int)c->n, c->dc);
}
template <typename ST>
static void apply_block(mpmx_ctx* c, ST st) {
    switch (c->P.vpage) {
    case 64: apply_dispatch<64>(c, st); break;
    case 128: apply_dispatch<128>(c, st); break;
    default: apply_dispatch<256>(c, st); break;
    }
}

// one PCG iteration k: four launches, all no-ops once the device flag says done
template <typename ST>
static void apply_blocks(mpmx_ctx* c, ST st) {
    const BlockTables t = block_tables_of(c);
    const int g1 = block_grid(c, c->cg_pct), g2 = g1;
    MPMX_TIMED(c, "cg_col", (k_cg_col<ST><<<(int)g1, 256, 0, c->stream>>>(
                                c->dp, st, c->vol, t, c->cg_s, c->cg_col, c->cbound, (int)c->n, c->scg, c->dc)));
    MPMX_TIMED(c, "cg_q", (k_cg_q<ST><<<(int)g2, 256, 0, c->stream>>>(
                              c->dp, st, t, static_cast<const BlockClasses*>(c->block_lists), c->cg_col, c->cbound,
                              c->grid, c->cg_s, c->cg_q, (int)c->n, c->scg, c->dc)));
}

static hipError_t pcg_iteration(mpmx_ctx* c, int k) {
    const bool blocks = c->cfg.p2g_variant != 1;
    if (blocks)
        MPMX_TIMED(c, "cg_apply", (c->cfg.precision != MPMX_FP32) ? apply_blocks(c, c->ms[c->cur])
                                                                   : apply_blocks(c, c->ps[c->cur]));
    else
        MPMX_TIMED(c, "cg_apply", (c->cfg.precision != MPMX_FP32) ? apply_block(c, c->ms[c->cur])
                                                                   : apply_block(c, c->ps[c->cur]));
    const int nb = c->cg_nb;
    // the block P2G records which nodes are active; the page-tile variant's
    // grid update does not, and the kernels then test sum w themselves
    const uint64_t* amask = blocks ? c->cg_amask : nullptr;
    MPMX_TIMED(c, "cg_vector", ({
                   k_cg_dot<<<nb, 256, 0, c->stream>>>(c->dp, c->grid, c->cg_s, c->cg_q, c->cg_part,
                                                       blocks ? 0 : 1, c->cg_r, c->cg_z, c->cg_part + nb,
                                                       (blocks && k == 0) ? 1 : 0, c->cap_tp, c->scg,
                                                       c->wg_scene, c->dc);
                   k_cg_update<<<nb, 256, 0, c->stream>>>(c->dp, c->grid, c->cg_x, c->cg_r, c->cg_z, c->cg_s,
                                                          c->cg_q, c->cg_part, nb, k, c->cap_tp, c->scg,
                                                          c->wg_scene, amask, blocks ? 0 : 1, c->dc);
                   k_cg_direction<<<nb, 256, 0, c->stream>>>(c->dp, c->grid, c->cg_z, c->cg_s, c->cg_part, nb, k,
                                                             c->cap_tp, c->scg, c->wg_scene, amask, c->dc);
               }));
    return hipGetLastError();
}

// Iterations are queued without host round trips and turn into no-ops once the
// device flag says done.  The first chunk is as long as the previous substep's
// solve (one host check per substep while the iteration count holds), later
// chunks cg_check_every long.
hipError_t launch_pcg_from(mpmx_ctx* c, int it0, int first_chunk) {
    const int maxit = c->P.cg_max_iter;
    const int every = c->cfg.cg_check_every > 0 ? c->cfg.cg_check_every : 1;
    int chunk = first_chunk > 0 ? first_chunk : 1;
    int it = it0;
    while (it < maxit) {
        chunk = (maxit - it) < chunk ? (maxit - it) : chunk;
        for (int k = 0; k < chunk; ++k) {
            hipError_t e = pcg_iteration(c, it + k);
            if (e != hipSuccess) return e;
        }
        it += chunk;
        // read the convergence flag and the counters (the block count sizes
        // the next launches of the block kernels)
        hipError_t e = counts_to_host(c);
        if (e != hipSuccess) return e;
        if (c->hc->cg_done || it >= maxit) break;
        chunk = every;
    }
    c->cg_spec = c->hc->cg_iters > 0 ? c->hc->cg_iters : 1;
    return hipSuccess;
}

hipError_t launch_pcg(mpmx_ctx* c, float dt) {
    (void)dt;
    if (!c->dp.phase_field) return hipSuccess;
    return launch_pcg_from(c, 0, c->cg_spec);
}

// iterations k0 .. k1-1 queued without a host check (for graph capture)
hipError_t queue_pcg_iterations(mpmx_ctx* c, int k0, int k1) {
    for (int k = k0; k < k1; ++k) {
        hipError_t e = pcg_iteration(c, k);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---- counters to the host without a stream synchronisation -----------------
// One wave copies the 31 counter words to coherent host memory, makes them
// visible system-wide, then writes the sequence number the host waits for.
__global__ void k_publish_counts(const DevCounts* __restrict__ dc, DevCounts* pub, unsigned seq) {
    constexpr int kWords = (int)(sizeof(DevCounts) / sizeof(unsigned));
    const unsigned* src = reinterpret_cast<const unsigned*>(dc);
    unsigned* dst = reinterpret_cast<unsigned*>(pub);
    const int lane = threadIdx.x;
    if (lane < kWords - 1) __hip_atomic_store(&dst[lane], src[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
    if (lane == 0) __hip_atomic_store(&pub->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

hipError_t counts_to_host(mpmx_ctx* c) {
    const unsigned want = ++c->pub_seq;
    k_publish_counts<<<1, 64, 0, c->stream>>>(c->dc, c->pub, want);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the copy lands when every earlier kernel of the stream has finished;
    // past two seconds (a stalled stream, or a launch that failed late) the
    // stream synchronisation below reports what happened
    const auto t0 = std::chrono::steady_clock::now();
    bool seen = false;
    for (int k = 0;; ++k) {
        if (__atomic_load_n(&c->pub->seq, __ATOMIC_ACQUIRE) == want) {
            seen = true;
            break;
        }
        if ((k & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) break;
    }
    if (!seen) {
        if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return e;
        if (__atomic_load_n(&c->pub->seq, __ATOMIC_ACQUIRE) != want) return hipErrorUnknown;
    }
    std::memcpy(c->hc, c->pub, sizeof(DevCounts));
    return hipSuccess;
}

}  // namespace mpmx
