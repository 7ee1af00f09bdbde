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
    const int* cell_start;     // [cap_p][65] first particle of each cell of each particle page
    int64_t cap_p, cap_tp, cap_blocks;
    int slot_shift;            // scene of a page key = key >> slot_shift
    const int* order;          // block of each item of the XCD shares, or null (item = block)
    const int* bpre;           // per bitmap word: blocks numbered before it (scan of the block counts)
};

inline BlockTables block_tables_of(const mpmx_ctx* c) {
    return BlockTables{c->pmask, c->nmask, c->wcount, c->scount, c->page2particle, c->keys, c->block_key,
                       c->cell_start, c->cap_p, c->cap_tp, c->tp_bound, c->dp.slot_shift,
                       c->order_ok ? c->block_order : nullptr, c->bcount};
}

// id of 2x2x2 block bk (Morton block key = page key >> 3), or -1 when none of
// its 8 pages is a particle or neighbour page: blocks are numbered in Morton
// order, a bitmap word holds 8 blocks (one byte each)
__device__ __forceinline__ int block_rank(const BlockTables& t, uint32_t bk) {
    const uint32_t w = bk >> 3, j = bk & 7u;
    const uint64_t m = t.pmask[w] | t.nmask[w];
    if (!((m >> (8 * j)) & 0xffull)) return -1;
    int id = t.bpre[w];
    for (uint32_t k = 0; k < j; ++k) id += ((m >> (8 * k)) & 0xffull) != 0ull;
    return id;
}

// page key of page (px, py, pz) (non-negative) when that page lies in the
// slot of `scene` (the page space of one scene), else ~0u
__device__ __forceinline__ uint32_t scene_page_key(const BlockTables& t, int px, int py, int pz, int scene) {
    const uint32_t k = morton((uint32_t)px, (uint32_t)py, (uint32_t)pz);
    return (int)(k >> t.slot_shift) == scene ? k : ~0u;
}

// ---- shared pieces of the block gathers ---------------------------------------
// A block gathers from the particle pages 2b-1 .. 2b+1 (per axis) of block b:
// its own 8 pages, whose particles all reach it, and the 19 pages just below
// it, whose particles reach it only from cells with index 2 or 3 along each
// axis on which the page is below.

// ---- class-ordered sources -----------------------------------------------------
// The same 27 source pages, grouped by class: bit 2 / 1 / 0 of the class set
// when the page lies below the block along x / y / z.  Only the particles in
// cells with index 2 or 3 along each such axis reach the block, and their
// base node lies one or two nodes below it, so a class knows at compile time
// which node offsets can fall outside the block.  Each source page is listed
// as runs of its passing cells (cell = i<<4 | j<<2 | k, i, j, k in 0..3), with
// the particle ranges from the page's cell starts, so every listed particle
// reaches the block; a workgroup's threads are split among the classes, an
// equal number of entries each, so that most waves hold particles of one
// class and take the same branches.
// class:             0  1(z) 2(y) 3(yz) 4(x) 5(xz) 6(xy) 7(xyz)
// pages:             8  4    4    2     4    2     2     1
// runs per page:     1  16   4    8     1    8     2     4
constexpr int kClassFirstSlot[9] = {0, 8, 12, 16, 18, 22, 24, 26, 27};
constexpr int kClassRuns[8] = {1, 16, 4, 8, 1, 8, 2, 4};
constexpr int kMaxRuns = 132;

__host__ __device__ constexpr int class_first_run(int cls) {
    int r = 0;
    for (int c = 0; c < cls; ++c) r += (kClassFirstSlot[c + 1] - kClassFirstSlot[c]) * kClassRuns[c];
    return r;
}

struct BlockClasses {
    int start[kMaxRuns];     // first particle of run r
    int pre[kMaxRuns + 1];   // particles listed before run r (all classes, in order)
    int page[27];            // particle page id of source slot s, or -1 (slot order)
    int out_page[8];
    int thr0[9];             // threads [thr0[c], thr0[c + 1]) work on class c
    int len[8];              // entries per thread of class c
};

// what the block G2P of a 2x2x2 block needs besides the grid: the grid pages
// of its staged node region (pages 2b .. 2b+2 per axis, -1 where absent) and
// the contiguous range [pa, pb) of its own particles
struct BlockG2P {
    int page[27];
    int pa, pb;
    int pad[3];
};

// copy of a block's lists from HBM into LDS (every thread of the workgroup)
constexpr int kClassWords = (int)(sizeof(BlockClasses) / sizeof(int));
__device__ __forceinline__ void classes_load(const BlockClasses* __restrict__ src, BlockClasses& S) {
    const int* a = reinterpret_cast<const int*>(src);
    int* b = reinterpret_cast<int*>(&S);
    for (int q = threadIdx.x; q < kClassWords; q += blockDim.x) b[q] = a[q];
}

// class and page offset (per axis, -1 below / 0 / 1) of source slot s
__device__ __forceinline__ void source_slot(int s, int& cls, int (&d)[3]) {
    cls = 0;
#pragma unroll
    for (int c = 1; c < 8; ++c) cls += s >= kClassFirstSlot[c] ? 1 : 0;
    const int q = s - kClassFirstSlot[cls];
    int bit = 0;
#pragma unroll
    for (int a = 2; a >= 0; --a) {
        if ((cls >> (2 - a)) & 1) {
            d[a] = -1;
        } else {
            d[a] = (q >> bit) & 1;
            ++bit;
        }
    }
}

// run r of a page of class cls: cells [c0, c0 + len)
__device__ __forceinline__ void class_run(int cls, int r, int& c0, int& len) {
    const int fm = ((cls & 4) ? 32 : 0) | ((cls & 2) ? 8 : 0) | ((cls & 1) ? 2 : 0);
    if (!fm) {
        c0 = 0;
        len = 64;
        return;
    }
    const int low = __builtin_ctz(fm);
    len = 1 << low;
    const int free_bits = (0x3f & ~((2 << low) - 1)) & ~fm;
    int dep = 0, src = 0;
    for (int b = 0; b < 6; ++b)
        if ((free_bits >> b) & 1) {
            dep |= ((r >> src) & 1) << b;
            ++src;
        }
    c0 = fm | dep;
}

// The lists of block bkey, built by one wave (lane = threadIdx.x & 63) and
// written straight to dst: the page ids from the page tables (one rank per
// lane, all at once: the 27 source slots, the G2P region's 27 grid pages and
// the block's own 8 grid pages), the 132 runs
// of passing cells (three per lane) with their particle ranges read from the
// cell starts, the exclusive prefix over the runs, and the threads of each
// class of the 256-thread gathers.  Returns the block's entries.
__device__ __forceinline__ int block_lists_wave(const BlockTables& t, uint32_t bkey, int n, int np_raw, int np,
                                                 int nt, BlockClasses* __restrict__ dst, BlockG2P* __restrict__ g2p) {
    const int lane = threadIdx.x & 63;
    int bc[3];
    unmorton(bkey, bc[0], bc[1], bc[2]);
    const int scene = (int)((bkey << 3) >> t.slot_shift);
    // one page rank per lane, all issued at once: lanes 0..26 the source
    // slots, 27..53 the G2P region's grid pages 2b + (0, 1, 2) per axis,
    // 54..61 the block's own 8 pages
    uint32_t key = ~0u;
    if (lane < 27) {
        int cls, d[3];
        source_slot(lane, cls, d);
        const int px = 2 * bc[0] + d[0], py = 2 * bc[1] + d[1], pz = 2 * bc[2] + d[2];
        key = (px >= 0 && py >= 0 && pz >= 0) ? scene_page_key(t, px, py, pz, scene) : ~0u;
    } else if (lane < 54) {
        const int q = lane - 27;
        key = scene_page_key(t, 2 * bc[0] + q / 9, 2 * bc[1] + (q / 3) % 3, 2 * bc[2] + q % 3, scene);
    } else if (lane < 62) {
        key = (bkey << 3) | (uint32_t)(lane - 54);
    }
    const int id = key != ~0u ? page_rank(key, t.pmask, t.nmask, t.wpre, t.spre, np_raw) : -1;
    // source pages that hold particles
    int pg = -1;
    if (lane < 27) {
        if (id >= 0 && id < np) {
            int a = t.page2particle[id], b = t.page2particle[id + 1];
            a = a < 0 ? 0 : (a > n ? n : a);
            b = b < a ? a : (b > n ? n : b);
            if (b > a) pg = id;
        }
        dst->page[lane] = pg;
    } else if (lane < 54) {
        g2p->page[lane - 27] = (id >= 0 && id < nt) ? id : -1;
    }
    // the block's own 8 pages: grid pages, and the particle range of those
    // that hold particles (consecutive page ids, so one range)
    {
        int lo = 0x7fffffff, hi = -1;
        if (lane >= 54 && lane < 62) {
            dst->out_page[lane - 54] = (id >= 0 && id < nt) ? id : -1;
            if (id >= 0 && id < np) lo = hi = id;
        }
#pragma unroll
        for (int w = 32; w > 0; w >>= 1) {
            lo = min(lo, __shfl_xor(lo, w, 64));
            hi = max(hi, __shfl_xor(hi, w, 64));
        }
        if (lane == 0) {
            int a = 0, b = 0;
            if (hi >= 0) {
                a = t.page2particle[lo];
                b = t.page2particle[hi + 1];
                a = a < 0 ? 0 : (a > n ? n : a);
                b = b < a ? a : (b > n ? n : b);
            }
            g2p->pa = a;
            g2p->pb = b;
        }
    }
    // runs q = lane + 64 k in class order; an absent page's runs are empty
    int excl[3], carry = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int q = lane + 64 * k;
        int cls = 0;
#pragma unroll
        for (int c = 1; c < 8; ++c) cls += q >= class_first_run(c) ? 1 : 0;
        const int nr = kClassRuns[cls];
        const int rel = q - class_first_run(cls);
        const int slot = kClassFirstSlot[cls] + rel / nr;
        const int ps = __shfl(pg, q < kMaxRuns ? slot : 0, 64);
        int st = 0, cnt = 0;
        if (q < kMaxRuns && ps >= 0) {
            int c0, len;
            class_run(cls, rel - (rel / nr) * nr, c0, len);
            int a = t.cell_start[(int64_t)ps * 65 + c0], b = t.cell_start[(int64_t)ps * 65 + c0 + len];
            a = a < 0 ? 0 : (a > n ? n : a);
            b = b < 0 ? 0 : (b > n ? n : b);
            st = a;
            cnt = b > a ? b - a : 0;
        }
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int up = __shfl_up(incl, o, 64);
            if (lane >= o) incl += up;
        }
        excl[k] = carry + incl - cnt;
        if (q < kMaxRuns) {
            dst->start[q] = st;
            dst->pre[q] = excl[k];
        }
        carry += __shfl(incl, 63, 64);
    }
    const int all = carry;
    if (lane == 0) dst->pre[kMaxRuns] = all;
    // every thread takes L = ceil(all / 248) entries of a single class (the
    // last thread of a class fewer) and a class gets ceil(n_c / L) threads
    // (at most 255 in all): waves stay within one class except where two
    // classes meet, their lanes run L entries each, and the threads left
    // over form idle trailing waves rather than short runs in busy ones
    const int c = lane < 8 ? lane : 0;
    int pre_c[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int b = class_first_run(c + h);
        const int v0 = __shfl(excl[0], b & 63, 64), v1 = __shfl(excl[1], b & 63, 64),
                  v2 = __shfl(excl[2], b & 63, 64);
        pre_c[h] = b >= kMaxRuns ? all : ((b >> 6) == 0 ? v0 : ((b >> 6) == 1 ? v1 : v2));
    }
    const int n_c = pre_c[1] - pre_c[0];
    const int L = all > 0 ? (all + 247) / 248 : 1;
    int thr = 0;
    if (lane < 8 && n_c > 0) thr = (n_c + L - 1) / L;
    int tin = thr;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        const int up = __shfl_up(tin, o, 64);
        if (lane >= o) tin += up;
    }
    if (lane < 8) {
        dst->thr0[lane] = tin - thr;
        dst->len[lane] = thr > 0 ? L : 0;
    }
    if (lane == 7) dst->thr0[8] = tin;
    return all;
}

// particles listed for class cls, and entry j of them
__device__ __forceinline__ int class_count(const BlockClasses& S, int cls) {
    return S.pre[cls == 7 ? kMaxRuns : class_first_run(cls + 1)] - S.pre[class_first_run(cls)];
}
__device__ __forceinline__ int class_entry(const BlockClasses& S, int cls, int j) {
    const int lo0 = class_first_run(cls);
    const int e = S.pre[lo0] + j;
    int lo = lo0, hi = lo0 + (kClassFirstSlot[cls + 1] - kClassFirstSlot[cls]) * kClassRuns[cls] - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S.pre[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return S.start[lo] + (e - S.pre[lo]);
}

// The order in which the block kernels take the blocks: by work, longest
// first, in half-octave classes of the block's gather entries (2^l <= n <
// 1.5 2^l and 1.5 2^l <= n < 2^(l+1), highest first, each class in block
// order), dealt
// round-robin over the 8 XCD shares (sorted rank g to share g mod 8 at
// position g / 8, the last nb mod 8 ranks to the shares that hold one item
// more), so that every XCD gets the same mix of long and short blocks and
// takes its long blocks first.  Built by one workgroup (any size that is a
// multiple of 64, at most 256 threads); deterministic.
constexpr int kOrderGroups = 1;
constexpr int kOrderClasses = 32;
__device__ __forceinline__ int wave_incl_scan_i(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}
__device__ __forceinline__ int work_class(int entries) {
    if (entries <= 1) return 0;
    const int l = 31 - __clz(entries);
    const int c = 2 * l + (((entries >> (l - 1)) & 1) ? 1 : 0);   // half octaves: 2^l, 1.5 2^l
    return c < kOrderClasses - 1 ? c : kOrderClasses - 1;
}
// item of sorted rank g among nb
__device__ __forceinline__ int dealt_item(int g, int nb) {
    const int K = nb >> 3;
    if (g < 8 * K) return xcd_share_lo(g & 7, nb) + (g >> 3);
    int rr = g - 8 * K;
    for (int x = 0; x < 8; ++x) {
        const int lo = xcd_share_lo(x, nb);
        if (xcd_share_lo(x + 1, nb) - lo > K) {
            if (rr == 0) return lo + K;
            --rr;
        }
    }
    return g;
}
// cnt: kOrderClasses x blockDim.x ints of LDS, lds4: 4 ints
// wclass: the work class of every block (k_block_lists writes it, one byte
// per block, so that each thread reads a contiguous stretch)
__device__ __forceinline__ void order_blocks(const uint8_t* __restrict__ wclass, int* __restrict__ order, int nb,
                                             int* cnt, int* lds4) {
    const int T = (int)blockDim.x, t = (int)threadIdx.x;
    const int chunk = (nb + T - 1) / T;
    const int b0 = t * chunk, b1 = b0 + chunk < nb ? b0 + chunk : nb;
    // counts per (class, thread), laid out highest class first, then by thread
    for (int k = 0; k < kOrderClasses; ++k) cnt[k * T + t] = 0;
    for (int b = b0; b < b1; ++b) ++cnt[(kOrderClasses - 1 - (int)wclass[b]) * T + t];
    __syncthreads();
    // exclusive scan of the kOrderClasses x T counts: kOrderClasses consecutive
    // entries per thread, then across threads
    int part = 0;
    for (int q = 0; q < kOrderClasses; ++q) part += cnt[t * kOrderClasses + q];
    const int lane = t & 63, w = t >> 6, nw = T >> 6;
    const int inc = wave_incl_scan_i(part);
    if (lane == 63) lds4[w] = inc;
    __syncthreads();
    int off = inc - part;
    for (int q = 0; q < w && q < nw; ++q) off += lds4[q];
    for (int q = 0; q < kOrderClasses; ++q) {
        const int v = cnt[t * kOrderClasses + q];
        cnt[t * kOrderClasses + q] = off;
        off += v;
    }
    __syncthreads();
    // each thread places its blocks at their ranks (its own column: no races)
    for (int b = b0; b < b1; ++b) {
        const int k = kOrderClasses - 1 - (int)wclass[b];
        const int g = cnt[k * T + t]++;
        order[dealt_item(g, nb)] = b;
    }
}

// the class of the calling thread and its entries [jb, je) of that class;
// false for threads without entries
__device__ __forceinline__ bool thread_class(const BlockClasses& S, int& cls, int& jb, int& je) {
    const int t = (int)threadIdx.x;
    cls = -1;
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (t >= S.thr0[c] && t < S.thr0[c + 1]) cls = c;
    if (cls < 0) return false;
    const int len = S.len[cls];
    jb = (t - S.thr0[cls]) * len;
    je = jb + len;
    const int n = class_count(S, cls);
    je = je < n ? je : n;
    return jb < je;
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
// a value already scaled, truncated, sign-extended
__device__ __forceinline__ unsigned long long to_fix_scaled(float xs) {
    return (unsigned long long)(long long)(int)xs;
}
// the same for a value known to be >= 0 (zero- and sign-extension agree)
__device__ __forceinline__ unsigned long long to_fix_scaled_nonneg(float xs) {
    return (unsigned long long)(unsigned)(int)xs;
}
__device__ __forceinline__ float from_fix(unsigned long long v, int e) {
    return (float)ldexp((double)(long long)v, -e);
}

// node index inside an 8^3 block accumulator: z padded to 9 and x strided by
// 76 words so that the lanes of a wave, which hold particles of neighbouring
// cells, spread their 64-bit atomics over the LDS banks (a model of the
// sorted particle order puts this layout at ~25 % fewer conflict cycles than
// the dense 9 x 72 one)
constexpr int kAccZ = 9;
constexpr int kAccX = 76;
constexpr int kAccStride = 8 * kAccX;   // words per channel
__host__ __device__ constexpr int acc_node(int lx, int ly, int lz) { return lx * kAccX + ly * kAccZ + lz; }

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
    if (s.F32) return s.F32[(int64_t)k * n + i];
    return h2f(s.F[(int64_t)k * n + i]) + ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f);
}
__device__ __forceinline__ void st_F(const PState& s, int i, int k, int n, float f) { s.F[(int64_t)k * n + i] = f; }
__device__ __forceinline__ void st_F(const MState& s, int i, int k, int n, float f) {
    if (s.F32) s.F32[(int64_t)k * n + i] = f;
    else s.F[(int64_t)k * n + i] = f2h(f - ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f));
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
