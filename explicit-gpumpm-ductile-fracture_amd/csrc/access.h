// access.h — particle-field accessors for the two storage forms (FP32 and the
// mixed-precision codecs), so that the transfer kernels are written once.
#pragma once
#include "mpmx_internal.h"
#include "precision.h"

namespace mpmx {

// page tables passed to the transfer kernels
struct Tables {
    const uint32_t* page_key;
    const int* page2particle;
    const int* adj;      // [7][cap_p]
    const int* vpo;
    const int* target;   // 0-based page of each virtual page
    int64_t cap_p, cap_tp, cap_vp;
};

// tables of the 2x2x2-page block kernels
struct BlockTables {
    const uint64_t* pmask;
    const uint64_t* nmask;
    const int* wpre;           // page rank prefix sums (particle, neighbour-only)
    const int* spre;
    const int* page2particle;
    const uint32_t* keys;      // sorted particle keys
    const uint32_t* block_key;
    int64_t cap_p, cap_tp, cap_blocks;
    int slot_shift;            // scene of a page key = key >> slot_shift
};

inline BlockTables block_tables_of(const mpmx_ctx* c) {
    return BlockTables{c->pmask, c->nmask, c->wcount, c->scount, c->page2particle, c->keys, c->block_key,
                       c->cap_p, c->cap_tp, c->tp_bound, c->dp.slot_shift};
}

// page key of page (px, py, pz) (non-negative) when that page lies in the
// slot of `scene` (the page space of one scene), else ~0u
__device__ __forceinline__ uint32_t scene_page_key(const BlockTables& t, int px, int py, int pz, int scene) {
    const uint32_t k = morton((uint32_t)px, (uint32_t)py, (uint32_t)pz);
    return (int)(k >> t.slot_shift) == scene ? k : ~0u;
}

// ---- shared pieces of the block gathers ---------------------------------------
// The particle pages 2b-1 .. 2b+1 (per axis) of block b, their particle ranges
// and the cell filter that keeps the particles reaching the block; lanes 32..39
// resolve the block's 8 grid pages.  Pages of another scene's slot are not
// sources.  Called by one wavefront (the first of the workgroup); LDS arrays
// owned by the caller.
struct BlockSources {
    int start[27], pre[28], filter[27];
    int page[27];          // particle page id of each source, or -1
    int out_page[8];
};

__device__ __forceinline__ void block_sources(const BlockTables& t, uint32_t bkey, const int (&bc)[3], int n,
                                              int np_raw, int np, int nt, BlockSources& S) {
    const int lane = threadIdx.x;
    int cnt = 0;
    const int scene = (int)((bkey << 3) >> t.slot_shift);
    if (lane < 27) {
        const int dx = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dz = lane % 3 - 1;
        const int px = 2 * bc[0] + dx, py = 2 * bc[1] + dy, pz = 2 * bc[2] + dz;
        int st = 0, page_of_source = -1;
        const uint32_t key = (px >= 0 && py >= 0 && pz >= 0) ? scene_page_key(t, px, py, pz, scene) : ~0u;
        if (key != ~0u) {
            const int id = page_rank(key, t.pmask, t.nmask, t.wpre, t.spre, np_raw);
            if (id >= 0 && id < np) {
                page_of_source = id;
                int a = t.page2particle[id], b = t.page2particle[id + 1];
                a = a < 0 ? 0 : (a > n ? n : a);
                b = b < a ? a : (b > n ? n : b);
                st = a;
                cnt = b - a;
            }
        }
        S.start[lane] = st;
        S.page[lane] = cnt > 0 ? page_of_source : -1;
        S.filter[lane] = (dx < 0 ? 1 : 0) | (dy < 0 ? 2 : 0) | (dz < 0 ? 4 : 0);
    } else if (lane >= 32 && lane < 40) {
        const int o = lane - 32;
        const int id = page_rank((bkey << 3) | (uint32_t)o, t.pmask, t.nmask, t.wpre, t.spre, np_raw);
        S.out_page[o] = (id >= 0 && id < nt) ? id : -1;
    }
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const int up = __shfl_up(incl, o, 64);
        if (lane >= o) incl += up;
    }
    if (lane < 27) S.pre[lane] = incl - cnt;
    if (lane == 26) S.pre[27] = incl;
}

// entry j of the concatenated source ranges: particle index and whether it
// reaches the block (any thread count)
__device__ __forceinline__ bool block_entry_simple(const BlockTables& t, const BlockSources& S, int j, int total,
                                                   int& i) {
    i = 0;
    if (j >= total) return false;
    int lo = 0, hi = 26;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S.pre[mid] <= j) lo = mid;
        else hi = mid - 1;
    }
    i = S.start[lo] + (j - S.pre[lo]);
    const int f = S.filter[lo];
    if (!f) return true;
    const uint32_t cell = t.keys[i] & 63u;
    return !(((f & 1) && !((cell >> 5) & 1u)) || ((f & 2) && !((cell >> 3) & 1u)) ||
             ((f & 4) && !((cell >> 1) & 1u)));
}

// ---- fixed-point accumulation -----------------------------------------------
// LDS float atomics cost ~190 CU cycles per wave instruction on gfx950 and
// 64-bit integer atomics ~7, so the block gathers add in fixed point: a
// channel whose largest single contribution in the block is b uses the scale
// 2^e with b 2^e <= 2^30, each contribution is truncated to a 32-bit integer
// (three VALU operations: scale, convert, sign extension) and added to a
// 64-bit accumulator, which holds 2^33 such terms.  Integer sums are exact and
// order-free: the result is bit-reproducible, and each term carries an error
// below 2^-30 of the block's largest contribution (finer than an FP32 sum).
constexpr int kFixBits = 30;

// exponent e for a channel bound b (>= 0): b 2^e <= 2^kFixBits
__device__ __forceinline__ int fix_exp(float b) {
    if (!(b > 0.f) || !isfinite(b)) return 0;
    int ex;
    frexpf(b, &ex);   // b < 2^ex
    const int e = kFixBits - ex;
    return e < -120 ? -120 : (e > 120 ? 120 : e);
}
// x scaled by sc = 2^e (|x sc| <= 2^30), truncated, sign-extended to 64 bits
__device__ __forceinline__ unsigned long long to_fix(float x, float sc) {
    return (unsigned long long)(long long)(int)(x * sc);
}
__device__ __forceinline__ float from_fix(unsigned long long v, int e) {
    return (float)ldexp((double)(long long)v, -e);
}

// node index inside an 8^3 block accumulator; z padded to 9 so that the nodes
// of one page spread over the LDS banks
constexpr int kAccZ = 9;
constexpr int kAccStride = 8 * 8 * kAccZ;   // floats per channel

__device__ __forceinline__ float ld_x(const PState& s, int i, int a) { return s.x[3 * i + a]; }
__device__ __forceinline__ float ld_x(const MState& s, int i, int a) { return s.x[3 * i + a]; }
__device__ __forceinline__ void st_x(const PState& s, int i, int a, float f) { s.x[3 * i + a] = f; }
__device__ __forceinline__ void st_x(const MState& s, int i, int a, float f) { s.x[3 * i + a] = f; }

__device__ __forceinline__ float ld_v(const PState& s, int i, int a) { return s.v[3 * i + a]; }
__device__ __forceinline__ float ld_v(const MState& s, int i, int a) { return bf2f(s.v[3 * i + a]); }
__device__ __forceinline__ void st_v(const PState& s, int i, int a, float f) { s.v[3 * i + a] = f; }
__device__ __forceinline__ void st_v(const MState& s, int i, int a, float f) { s.v[3 * i + a] = f2bf(f); }

// k = r + 3c, component-major storage
__device__ __forceinline__ float ld_B(const PState& s, int i, int k, int n) { return s.B[(int64_t)k * n + i]; }
__device__ __forceinline__ float ld_B(const MState& s, int i, int k, int n) { return bf2f(s.B[(int64_t)k * n + i]); }
__device__ __forceinline__ void st_B(const PState& s, int i, int k, int n, float f) { s.B[(int64_t)k * n + i] = f; }
__device__ __forceinline__ void st_B(const MState& s, int i, int k, int n, float f) {
    s.B[(int64_t)k * n + i] = f2bf(f);
}

__device__ __forceinline__ float ld_F(const PState& s, int i, int k, int n) { return s.F[(int64_t)k * n + i]; }
__device__ __forceinline__ float ld_F(const MState& s, int i, int k, int n) {
    return h2f(s.F[(int64_t)k * n + i]) + ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f);
}
__device__ __forceinline__ void st_F(const PState& s, int i, int k, int n, float f) { s.F[(int64_t)k * n + i] = f; }
__device__ __forceinline__ void st_F(const MState& s, int i, int k, int n, float f) {
    s.F[(int64_t)k * n + i] = f2h(f - ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f));
}

__device__ __forceinline__ float ld_c(const PState& s, int i) { return s.c[i]; }
__device__ __forceinline__ float ld_c(const MState& s, int i) { return bf2f(s.c[i]); }
__device__ __forceinline__ void st_c(const PState& s, int i, float f) { s.c[i] = f; }
__device__ __forceinline__ void st_c(const MState& s, int i, float f) { s.c[i] = f2bf(f); }

__device__ __forceinline__ float ld_m(const PState& s, int i) { return s.mass[i]; }
__device__ __forceinline__ float ld_m(const MState& s, int i) { return s.mass[i]; }

// tile node (a, b, c) in [0,6)^3 of page p -> global grid node index page*64 + cell,
// or -1 when the neighbour page lies beyond the grid capacity (the table build
// has latched a capacity error then; the node is skipped instead of written)
__device__ __forceinline__ int64_t tile_node(const Tables& t, int p, int a, int b, int c) {
    const int nb = ((a >> 2) << 2) | ((b >> 2) << 1) | (c >> 2);
    const int pg = nb ? t.adj[(int64_t)(nb - 1) * t.cap_p + p] : p;
    if (pg < 0 || pg >= t.cap_tp) return -1;
    return (int64_t)pg * 64 + (((a & 3) << 4) | ((b & 3) << 2) | (c & 3));
}

// range [s, e) of the particles of virtual page vp and its page p, clamped so
// that stale tables after a capacity overflow still index inside the arrays
__device__ __forceinline__ void vpage_range(const Tables& t, int vp, int vpage, int n, int& p, int& s, int& e) {
    p = t.target[vp];
    p = p < 0 ? 0 : (p >= t.cap_p ? (int)t.cap_p - 1 : p);
    s = t.page2particle[p] + (vp - t.vpo[p]) * vpage;
    const int pe = t.page2particle[p + 1];
    e = s + vpage < pe ? s + vpage : pe;
    s = s < 0 ? 0 : (s > n ? n : s);
    e = e < s ? s : (e > n ? n : e);
}

}  // namespace mpmx
