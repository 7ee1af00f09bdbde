// k_sort.hip — cell keys, device-wide exclusive scan, stable LSD radix sort of
// the particle keys, and the gather of every particle field into sorted order.
//
// Replaces Particles::sort_by_offsets / reorder (Particle_kernal.cu:167-249,
// 484-573) and the recursive block scan (cuda_kernal_tools.cu:15-203).  The
// reference buckets particles by (hash-numbered page, cell) with atomics, so
// its page order and within-cell order change from run to run; here the sort
// is a stable radix sort of the Morton key, giving one canonical order: key
// ascending, ties in previous sorted order.
#include "mpmx_internal.h"
#include <utility>

namespace mpmx {

// ---------------------------------------------------------------- fills
__global__ void k_fill_zero_u32(uint32_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}
// zero count_dev[0] * per_unit floats (count clamped to cap_units)
__global__ void k_fill_zero_count(float* p, const int* count_dev, int64_t per_unit, int64_t cap_units) {
    int64_t units = *count_dev;
    units = units < cap_units ? units : cap_units;
    const int64_t n = units * per_unit;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

hipError_t fill_zero(mpmx_ctx* c, void* p, int64_t bytes) {
    const int64_t words = (bytes + 3) / 4;
    int64_t blocks = (words + 255) / 256;
    blocks = blocks < 8192 ? blocks : 8192;
    if (blocks < 1) blocks = 1;
    k_fill_zero_u32<<<(int)blocks, 256, 0, c->stream>>>((uint32_t*)p, words);
    return hipGetLastError();
}
hipError_t fill_zero_dev_count(mpmx_ctx* c, float* p, const int* count_dev, int64_t per_unit,
                               int64_t cap_units) {
    int64_t blocks = (cap_units * per_unit + 255) / 256;
    blocks = blocks < 8192 ? blocks : 8192;
    if (blocks < 1) blocks = 1;
    k_fill_zero_count<<<(int)blocks, 256, 0, c->stream>>>(p, count_dev, per_unit, cap_units);
    return hipGetLastError();
}

// ---------------------------------------------------------------- scan
// Exclusive scan of int32 in two launches: tile sums, then every tile adds up
// the sums of the tiles before it (at most a few hundred) and scans its own
// items from that offset; the last tile writes the total.  The item count
// comes either from the host (n_dev == nullptr) or from device memory, so the
// same launches serve tables whose sizes are only known on the device.
// In-place scans (in == out) are allowed.
constexpr int kScanPer = 4;
constexpr int kScanTile = 256 * kScanPer;

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// block of 256 threads: exclusive scan of one value per thread, returns total
__device__ __forceinline__ int block_excl_scan256(int v, int& total, int* lds4) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int inc = wave_incl_scan(v);
    if (lane == 63) lds4[w] = inc;
    __syncthreads();
    int off = 0;
    total = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int s = lds4[q];
        if (q < w) off += s;
        total += s;
    }
    __syncthreads();
    return off + inc - v;
}

__device__ __forceinline__ int64_t scan_count(const int* n_dev, int64_t n_host) {
    int64_t n = n_dev ? (int64_t)*n_dev : n_host;
    return n < n_host ? n : n_host;   // n_host is the capacity when n_dev is given
}

// tile `tile` of a two-launch exclusive scan: its sum
__device__ __forceinline__ void scan_reduce_tile(const int* in, int* sums, int64_t n, int tile, int* lds4) {
    const int64_t base = (int64_t)tile * kScanTile;
    int local = 0;
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) {
        const int64_t i = base + q * 256 + threadIdx.x;   // coalesced
        local += (i < n) ? in[i] : 0;
    }
    int total;
    (void)block_excl_scan256(local, total, lds4);
    if (threadIdx.x == 0) sums[tile] = total;
}

// tile `tile`: its prefix from the sums of the tiles before it, then the
// scan of its entries; the last tile writes the total
__device__ __forceinline__ void scan_apply_tile(const int* in, int* out, const int* __restrict__ sums, int64_t n,
                                                int tile, int tiles, int* total_dev, int* lds4) {
    int pre = 0;
    for (int t = threadIdx.x; t < tile; t += 256) pre += sums[t];
    int offset;
    (void)block_excl_scan256(pre, offset, lds4);
    const int64_t base = (int64_t)tile * kScanTile + (int64_t)threadIdx.x * kScanPer;
    int v[kScanPer];
    int local = 0;
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) {
        const int64_t i = base + q;
        v[q] = (i < n) ? in[i] : 0;
        local += v[q];
    }
    int total;
    int off = block_excl_scan256(local, total, lds4) + offset;
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) {
        const int64_t i = base + q;
        if (i < n) out[i] = off;
        off += v[q];
    }
    if (tile == tiles - 1 && threadIdx.x == 0 && total_dev) *total_dev = offset + total;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const int* in, int* sums, const int* n_dev, int64_t n_host) {
    __shared__ int lds4[4];
    scan_reduce_tile(in, sums, scan_count(n_dev, n_host), blockIdx.x, lds4);
}

__global__ __launch_bounds__(256) void k_scan_apply(const int* in, int* out, const int* __restrict__ sums,
                                                    const int* n_dev, int64_t n_host, int* total_dev) {
    __shared__ int lds4[4];
    scan_apply_tile(in, out, sums, scan_count(n_dev, n_host), blockIdx.x, gridDim.x, total_dev, lds4);
}

hipError_t scan_exclusive(mpmx_ctx* c, const int* in, int* out, int64_t n_cap, const int* n_dev,
                          int64_t n_host, int* total_dev) {
    const int64_t tiles = (n_cap + kScanTile - 1) / kScanTile;
    if (tiles < 1) return hipSuccess;
    if (tiles > c->scan_tmp_cap) return hipErrorInvalidValue;
    // with a device count, n_host carries the capacity bound
    const int64_t bound = n_dev ? n_cap : n_host;
    k_scan_reduce<<<(int)tiles, 256, 0, c->stream>>>(in, c->scan_tmp, n_dev, bound);
    k_scan_apply<<<(int)tiles, 256, 0, c->stream>>>(in, out, c->scan_tmp, n_dev, bound, total_dev);
    return hipGetLastError();
}

// up to three scans of n entries each in the same two launches (in place
// allowed); array y uses the tile sums at sums + y * tiles
__global__ __launch_bounds__(256) void k_scan_reduce3(ScanSet3 a, int* sums, int64_t n) {
    __shared__ int lds4[4];
    const int y = blockIdx.y;
    scan_reduce_tile(a.in[y], sums + (int64_t)y * gridDim.x, n, blockIdx.x, lds4);
}
__global__ __launch_bounds__(256) void k_scan_apply3(ScanSet3 a, const int* sums, int64_t n) {
    __shared__ int lds4[4];
    const int y = blockIdx.y;
    scan_apply_tile(a.in[y], a.out[y], sums + (int64_t)y * gridDim.x, n, blockIdx.x, gridDim.x, a.total[y], lds4);
}

hipError_t scan_exclusive3(mpmx_ctx* c, const ScanSet3& a, int k, int64_t n) {
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles < 1) return hipSuccess;
    if (k * tiles > c->scan_tmp_cap) return hipErrorInvalidValue;
    const dim3 grid((unsigned)tiles, (unsigned)k);
    k_scan_reduce3<<<grid, 256, 0, c->stream>>>(a, c->scan_tmp, n);
    k_scan_apply3<<<grid, 256, 0, c->stream>>>(a, c->scan_tmp, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- keys
// cal_cellId (Particle_kernal.cu:158-165): key of the quadratic base node.
// Particles are in upload order here: scene s holds [scene_start[s],
// scene_start[s + 1]); the key adds the scene's slot origin.
__global__ void k_keys(DevParams dp, const float* __restrict__ x, uint32_t* __restrict__ keys, int n,
                       const int* __restrict__ scene_start, DevCounts* dc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = dp.num_scenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (scene_start[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const DevParams& sp = dp.scenes[lo];
    const int N = sp.N;
    const float x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
    // the same float bounds as the G2P's domain test and the upload's host check
    const float lo_x = 2.0f * sp.dx, hi_x = 1.0f - 2.0f * sp.dx;
    if (!(x0 >= lo_x && x0 <= hi_x && x1 >= lo_x && x1 <= hi_x && x2 >= lo_x && x2 <= hi_x))
        atomicOr(&dc->err, kErrDomain);
    keys[i] = key_of(base_node(x0, N) + sp.origin[0], base_node(x1, N) + sp.origin[1],
                     base_node(x2, N) + sp.origin[2]);
}

hipError_t compute_keys(mpmx_ctx* c) {
    const int n = (int)c->n;
    const float* x = c->cfg.precision != MPMX_FP32 ? c->ms[c->cur].x : c->ps[c->cur].x;
    k_keys<<<(n + 255) / 256, 256, 0, c->stream>>>(c->dp, x, c->keys_next, n, c->scene_start, c->dc);
    return hipGetLastError();
}

// ---------------------------------------------------------------- radix sort
constexpr int kRsThreads = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsThreads * kRsItems;

// Item count and tile count of a radix launch: from the host, or (n_dev set)
// from device memory, capped at n.  A device count sorts a prefix whose
// length only the device knows (the movers), with the digit-major histogram
// packed for that many tiles; the launch then walks the tiles it needs with
// a stride of its grid.
struct RsSize {
    int n;
    const int* n_dev;
    int tiles;   // host tile count (n_dev == nullptr)
};
__device__ __forceinline__ void rs_size(const RsSize& z, int& n, int& tiles) {
    n = z.n;
    tiles = z.tiles;
    if (z.n_dev) {
        const int d = *z.n_dev;
        n = d < n ? d : n;
        n = n > 0 ? n : 0;
        tiles = (n + kRsTile - 1) / kRsTile;
    }
}

// per-tile digit histogram, digit-major: hist[d * tiles + t]; with a device
// count, workgroup 0 also stores the histogram's length for the scan
__global__ __launch_bounds__(kRsThreads) void k_radix_hist(const uint32_t* __restrict__ keys, RsSize z, int shift,
                                                           uint32_t* __restrict__ hist, int* __restrict__ hist_n) {
    __shared__ uint32_t cnt[256];
    int n, tiles;
    rs_size(z, n, tiles);
    if (hist_n && blockIdx.x == 0 && threadIdx.x == 0) *hist_n = 256 * tiles;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        cnt[threadIdx.x] = 0;
        __syncthreads();
        const int base = tile * kRsTile;
#pragma unroll
        for (int j = 0; j < kRsItems; ++j) {
            const int i = base + j * kRsThreads + threadIdx.x;
            if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
        }
        __syncthreads();
        hist[threadIdx.x * tiles + tile] = cnt[threadIdx.x];
    }
}

// stable scatter: items are ranked in (round, wave, lane) order, which is their
// index order inside the tile (same-digit peers inside a wave are found with
// eight 64-lane ballots), and put in their digit's run of an LDS copy of the
// tile; the tile then leaves in that order, so that consecutive lanes write
// consecutive positions of a digit's run instead of 256 scattered words per
// round.  The tile's count of digit d is the scanned histogram's next
// digit-major entry minus its own (n after the last).  All items of the tile
// are loaded before the ranking starts.  vals == nullptr: the values are the
// items' indices.
__global__ __launch_bounds__(kRsThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t* __restrict__ okeys,
    uint32_t* __restrict__ ovals, RsSize z, int shift, const uint32_t* __restrict__ offs) {
    __shared__ uint32_t gstart[256];   // output position of the tile's first item of digit d
    __shared__ uint32_t lstart[256];   // its slot in the tile copy
    __shared__ uint32_t lnext[256];    // next free slot of digit d
    __shared__ uint32_t wcnt[4][256];
    __shared__ uint32_t skey[kRsTile], sval[kRsTile];
    __shared__ int lds4[4];
    int n, tiles;
    rs_size(z, n, tiles);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int base = tile * kRsTile;
        uint32_t k[kRsItems], v[kRsItems];
#pragma unroll
        for (int j = 0; j < kRsItems; ++j) {
            const int i = base + j * kRsThreads + tid;
            k[j] = i < n ? keys[i] : 0u;
            v[j] = i < n ? (vals ? vals[i] : (uint32_t)i) : 0u;
        }
        {
            const int64_t e = (int64_t)tid * tiles + tile;
            const uint32_t g0 = offs[e];
            const uint32_t g1 = e + 1 < (int64_t)256 * tiles ? offs[e + 1] : (uint32_t)n;
            int total;
            const int ls = block_excl_scan256((int)(g1 - g0), total, lds4);
            gstart[tid] = g0;
            lstart[tid] = (uint32_t)ls;
            lnext[tid] = (uint32_t)ls;
        }
#pragma unroll
        for (int j = 0; j < kRsItems; ++j) {
#pragma unroll
            for (int q = 0; q < 4; ++q) wcnt[q][tid] = 0;
            __syncthreads();
            const bool valid = base + j * kRsThreads + tid < n;
            const uint32_t d = (k[j] >> shift) & 255u;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int bit = (d >> b) & 1;
                const uint64_t bal = __ballot(bit);
                peers &= bit ? bal : ~bal;
            }
            const uint32_t rank = __popcll(peers & lt);
            if (valid && (peers & lt) == 0ull) wcnt[w][d] = __popcll(peers);
            __syncthreads();
            if (valid) {
                uint32_t pos = lnext[d] + rank;
                for (int q = 0; q < w; ++q) pos += wcnt[q][d];
                skey[pos] = k[j];
                sval[pos] = v[j];
            }
            __syncthreads();
            lnext[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        }
        __syncthreads();
        const int cnt = n - base < kRsTile ? n - base : kRsTile;
        for (int q = tid; q < cnt; q += kRsThreads) {
            const uint32_t kq = skey[q];
            const uint32_t d = (kq >> shift) & 255u;
            const uint32_t g = gstart[d] + ((uint32_t)q - lstart[d]);
            okeys[g] = kq;
            ovals[g] = sval[q];
        }
        __syncthreads();   // the tile copy and the digit tables are reused by the next tile
    }
}

// ---------------------------------------------------------------- mover sort
// Between two sorts most particles keep their key (their cell).  Those that
// keep it (the stayers) are already in key order, since the slots hold the
// previous sort's order; only the movers need sorting.  The movers are listed
// in slot order, radix-sorted by their new keys, and both groups are merged
// by position, which gives the same order as a stable sort of all keys (key
// ascending, ties by slot).  Slots in tiles of kRsTile, 8 per thread (two
// runs of 4, below).  old: the previous sort's keys (sorted), cur: the keys
// now.

// A tile's 2048 slots in two halves of 1024; thread t holds slots
// base + 1024 h + 4t + q (item j = 4h + q), so that every load of a wave
// reads 1 KiB in one piece.  Returns the moved flags as bits (bit j).
__device__ __forceinline__ int item_slot(int base, int j) {
    return base + (j >> 2) * 1024 + 4 * (int)threadIdx.x + (j & 3);
}
__device__ __forceinline__ unsigned mover_bits(const uint32_t* __restrict__ old, const uint32_t* __restrict__ cur,
                                               int base, int n, uint32_t (&ko)[kRsItems], uint32_t (&kn)[kRsItems]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int s0 = base + h * 1024 + 4 * (int)threadIdx.x;
        if (s0 + 4 <= n) {
            const uint4 a = *reinterpret_cast<const uint4*>(old + s0);
            const uint4 b = *reinterpret_cast<const uint4*>(cur + s0);
            ko[4 * h] = a.x; ko[4 * h + 1] = a.y; ko[4 * h + 2] = a.z; ko[4 * h + 3] = a.w;
            kn[4 * h] = b.x; kn[4 * h + 1] = b.y; kn[4 * h + 2] = b.z; kn[4 * h + 3] = b.w;
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                ko[4 * h + q] = s0 + q < n ? old[s0 + q] : 0u;
                kn[4 * h + q] = s0 + q < n ? cur[s0 + q] : 0u;
            }
        }
    }
    unsigned bits = 0;
#pragma unroll
    for (int j = 0; j < kRsItems; ++j)
        if (item_slot(base, j) < n && ko[j] != kn[j]) bits |= 1u << j;
    return bits;
}

// movers before each half's first item of the thread, from one block scan
// of both halves' counts packed in one word (at most 1024 each): mr[0] for
// the first half, mr[1] for the second (which follows the whole first half)
__device__ __forceinline__ void half_ranks(unsigned bits, int t0, int (&mr)[2], int& total, int* lds4) {
    const int packed = __popc(bits & 15u) | (__popc(bits >> 4) << 16);
    int tot;
    const int ex = block_excl_scan256(packed, tot, lds4);
    mr[0] = t0 + (ex & 0xffff);
    mr[1] = t0 + (tot & 0xffff) + (ex >> 16);
    total = (tot & 0xffff) + (tot >> 16);
}

// movers per tile
__global__ __launch_bounds__(kRsThreads) void k_mover_count(const uint32_t* __restrict__ old,
                                                            const uint32_t* __restrict__ cur, int n,
                                                            int* __restrict__ tcount) {
    __shared__ int lds4[4];
    uint32_t ko[kRsItems], kn[kRsItems];
    const int c = __popc(mover_bits(old, cur, blockIdx.x * kRsTile, n, ko, kn));
    int total;
    (void)block_excl_scan256(c, total, lds4);
    if (threadIdx.x == 0) tcount[blockIdx.x] = total;
}

// bit k of the low 16 bits of x to bit 4k
__device__ __forceinline__ uint64_t spread4(uint64_t x) {
    x &= 0xffffull;
    x = (x | (x << 24)) & 0x000000ff000000ffull;
    x = (x | (x << 12)) & 0x000f000f000f000full;
    x = (x | (x << 6)) & 0x0303030303030303ull;
    x = (x | (x << 3)) & 0x1111111111111111ull;
    return x;
}

// the movers in slot order: new key, previous key, slot (toff: the tiles'
// exclusive mover offsets); per 64 slots, the movers before them (cpre) and
// which of them move (cmask), so that the movers before any slot s are
// cpre[s / 64] + popcount(cmask[s / 64] below bit s % 64)
__global__ __launch_bounds__(kRsThreads) void k_mover_list(const uint32_t* __restrict__ old,
                                                           const uint32_t* __restrict__ cur, int n,
                                                           const int* __restrict__ toff, uint32_t* __restrict__ mkey,
                                                           uint32_t* __restrict__ mold, uint32_t* __restrict__ mslot,
                                                           int* __restrict__ cpre, uint64_t* __restrict__ cmask) {
    __shared__ int lds4[4];
    uint32_t ko[kRsItems], kn[kRsItems];
    const int base = blockIdx.x * kRsTile;
    const unsigned bits = mover_bits(old, cur, base, n, ko, kn);
    int mr[2], total;
    half_ranks(bits, toff[blockIdx.x], mr, total, lds4);
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        // 16 lanes hold a chunk of 64 slots: lane L's slot q is bit 4 (L % 16) + q
        uint64_t cm = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) cm |= spread4(__ballot((bits >> (4 * h + q)) & 1u) >> (lane & 48)) << q;
        const int s0 = base + h * 1024 + 4 * (int)threadIdx.x;
        if ((lane & 15) == 0 && s0 < n) {
            cpre[s0 >> 6] = mr[h];
            cmask[s0 >> 6] = cm;
        }
        int r = mr[h];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if ((bits >> (4 * h + q)) & 1u) {
                mkey[r] = kn[4 * h + q];
                mold[r] = ko[4 * h + q];
                mslot[r] = (uint32_t)(s0 + q);
                ++r;
            }
    }
}

// Searches over sorted arrays.  An entry q is "below" the target when
// key[q] < k (kLower: the first index at or above k is found), key[q] <= k
// (kUpper: the first index above k), or (key[q], val[q]) < (k, v) (kPair).
enum SearchMode { kLower, kUpper, kPair };
template <int MODE>
__device__ __forceinline__ bool below(const uint32_t* key, const uint32_t* val, int q, uint32_t k, uint32_t v) {
    const uint32_t kq = key[q];
    if (MODE == kLower) return kq < k;
    if (MODE == kUpper) return kq <= k;
    return kq < k || (kq == k && val[q] < v);
}
// the first index in [0, n) not below the target, searched outward from s
// with doubling steps, then by bisection inside the last step: a mover's new
// key lies within a cell or two of the keys around its previous place, so
// the answer is mostly a few entries from s
template <int MODE>
__device__ __forceinline__ int search_near(const uint32_t* key, const uint32_t* val, int n, uint32_t k, uint32_t v,
                                           int s) {
    if (n <= 0) return 0;
    s = s < 0 ? 0 : (s >= n ? n - 1 : s);
    int lo, hi;
    if (below<MODE>(key, val, s, k, v)) {   // forward: entries before lo are below
        int step = 1;
        lo = s + 1;
        hi = s + 1;
        while (hi < n && below<MODE>(key, val, hi, k, v)) {
            lo = hi + 1;
            step <<= 1;
            hi = s + step;
        }
        hi = hi < n ? hi : n;
    } else {   // backward: entry hi is not below
        int step = 1;
        hi = s;
        lo = s - 1;
        while (lo >= 0 && !below<MODE>(key, val, lo, k, v)) {
            hi = lo;
            step <<= 1;
            lo = s - step;
        }
        lo = lo < 0 ? 0 : lo + 1;
    }
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (below<MODE>(key, val, mid, k, v)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// movers in the slots before slot s (s in [0, n]; m in all)
__device__ __forceinline__ int movers_before(const int* __restrict__ cpre, const uint64_t* __restrict__ cmask, int s,
                                             int n, int m) {
    if (s >= n) return m;
    const uint64_t below_s = (s & 63) ? ((~0ull) >> (64 - (s & 63))) : 0ull;
    return cpre[s >> 6] + __popcll(cmask[s >> 6] & below_s);
}

// The previous sort's page tables (consistent with its keys): the first slot
// whose key is >= k, or > k, in two dependent steps: the page's rank among
// the particle pages from the bitmap, then its cell start, or for a page
// without particles the first slot of the next particle page.
struct OldTables {
    const uint64_t* pmask;
    const int* wpre;
    const int* page2particle;
    const int* cell_start;
    const int* num_page;   // device count
    int64_t cap_p;
};
__device__ __forceinline__ int table_bound(const OldTables& t, uint32_t k, bool upper) {
    const uint32_t pk = k >> 6, w = pk >> 6, b = pk & 63u;
    const uint64_t P = t.pmask[w];
    const int rank = t.wpre[w] + __popcll(P & (b ? ((~0ull) >> (64 - b)) : 0ull));
    if ((P >> b) & 1ull) return t.cell_start[(int64_t)rank * 65 + (k & 63u) + (upper ? 1 : 0)];
    return t.page2particle[rank];
}

// The movers' places: mover r of the sorted movers (new key K, list index
// r0, slot i, previous key P) comes after the r movers before it and after
// every stayer with (key, slot) < (K, i): all slots with (previous key, slot)
// < (K, i), less the movers among them.  Since K != P that count of slots is
// a plain bound: for K > P the slots before the first previous key >= K (all
// of them after i), for K < P the slots up to the last previous key <= K
// (all before i); with the tables of the previous sort (tb set) it comes
// from the cell starts, else from a search of the previous keys.  Threads
// 0 .. tiles also give each tile of slots the sorted movers that precede its
// first slot (tbound, for k_place_stayers).
__global__ __launch_bounds__(kRsThreads) void k_place_movers(const uint32_t* __restrict__ old, int n,
                                                             const uint32_t* __restrict__ skey,
                                                             const uint32_t* __restrict__ sr0,
                                                             const uint32_t* __restrict__ mold,
                                                             const uint32_t* __restrict__ mslot, const int* m_dev,
                                                             const int* __restrict__ cpre,
                                                             const uint64_t* __restrict__ cmask,
                                                             const int* __restrict__ toff, int tiles,
                                                             int* __restrict__ tbound, OldTables tb, bool use_tables,
                                                             uint32_t* __restrict__ okeys, uint32_t* __restrict__ perm) {
    int m = *m_dev;
    m = m < n ? m : n;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    // the tile bounds go to the last threads of the grid, which mostly hold
    // no mover: their searches then run beside the movers' own
    const int tq = gid - (int)(gridDim.x * blockDim.x - (tiles + 1));
    if (tq >= 0 && tq < tiles) {
        const int t0 = toff[tq];
        tbound[tq] = search_near<kPair>(skey, sr0, m, old[tq * kRsTile], (uint32_t)t0, t0);
    } else if (tq == tiles) {
        tbound[tiles] = m;
    }
    if (use_tables) {
        const int np = *tb.num_page;
        use_tables = np <= tb.cap_p;   // truncated tables (a capacity error) are not used
    }
    for (int r = gid; r < m; r += gridDim.x * blockDim.x) {
        const uint32_t K = skey[r];
        const int r0 = (int)sr0[r];
        const int i = (int)mslot[r0];
        const bool up = K > mold[r0];
        int all;
        if (use_tables) all = table_bound(tb, K, !up);
        else all = up ? search_near<kLower>(old, nullptr, n, K, 0u, i + 1) : search_near<kUpper>(old, nullptr, n, K, 0u, i - 1);
        all = all < 0 ? 0 : (all > n ? n : all);
        const int pos = r + all - movers_before(cpre, cmask, all, n, m);
        if (pos < 0 || pos >= n) continue;   // only if the previous keys were not sorted
        okeys[pos] = K;
        perm[pos] = (uint32_t)i;
    }
}

// The stayers' places: stayer i (key k, m_i movers in slots before it) comes
// after the i - m_i stayers before it and after the c movers with (key, list
// index) < (k, m_i).  A tile's stayers fall between the bounds of its first
// slot and of the next tile's, and the movers in between are staged in LDS
// when they fit; in each run of 4 slots of a thread the stayers come in
// order, so each one's c continues from the last.
constexpr int kPlaceStage = 1024;
__global__ __launch_bounds__(kRsThreads) void k_place_stayers(const uint32_t* __restrict__ old,
                                                              const uint32_t* __restrict__ cur, int n,
                                                              const int* __restrict__ toff,
                                                              const int* __restrict__ tbound,
                                                              const uint32_t* __restrict__ skey,
                                                              const uint32_t* __restrict__ sr0, const int* m_dev,
                                                              uint32_t* __restrict__ okeys, uint32_t* __restrict__ perm) {
    __shared__ int lds4[4];
    __shared__ uint32_t stk[kPlaceStage], stv[kPlaceStage];
    int m = *m_dev;
    m = m < n ? m : n;
    uint32_t ko[kRsItems], kn[kRsItems];
    const int base = blockIdx.x * kRsTile;
    const unsigned bits = mover_bits(old, cur, base, n, ko, kn);
    int mr[2], total;
    half_ranks(bits, toff[blockIdx.x], mr, total, lds4);
    // the tile's stayers lie between the bounds of its first slot and the
    // next tile's (k_place_movers found them)
    const int lo = tbound[blockIdx.x];
    int hi = tbound[blockIdx.x + 1];
    hi = hi > lo ? (hi < m ? hi : m) : lo;
    const bool staged = hi - lo <= kPlaceStage;
    if (staged)
        for (int q = threadIdx.x; q < hi - lo; q += kRsThreads) {
            stk[q] = skey[lo + q];
            stv[q] = sr0[lo + q];
        }
    __syncthreads();
    const uint32_t* ck = staged ? stk : skey + lo;
    const uint32_t* cv = staged ? stv : sr0 + lo;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        int mrh = mr[h];
        int c = -1;   // the last stayer's count, relative to lo
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = 4 * h + q;
            const int i = item_slot(base, j);
            if (i >= n) break;
            if ((bits >> j) & 1u) {
                ++mrh;
                continue;
            }
            if (c < 0) {
                int a = 0, b = hi - lo;
                while (a < b) {
                    const int mid = (a + b) >> 1;
                    if (below<kPair>(ck, cv, mid, ko[j], (uint32_t)mrh)) a = mid + 1;
                    else b = mid;
                }
                c = a;
            } else {
                while (c < hi - lo && below<kPair>(ck, cv, c, ko[j], (uint32_t)mrh)) ++c;
            }
            const int pos = i - mrh + lo + c;
            if (pos < 0 || pos >= n) continue;   // only if the previous keys were not sorted
            okeys[pos] = ko[j];
            perm[pos] = (uint32_t)i;
        }
    }
}

// ---------------------------------------------------------------- gather
// Every field of particle perm[i] is loaded before any is stored: the source
// and destination state are separate buffers, but the compiler cannot know
// that of plain pointers, and with stores between them it would keep one
// load in flight per thread.
__global__ void k_gather_fp32(const uint32_t* __restrict__ perm, PState s, PState d, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = (int)perm[i];
    float x[3], v[3], B[9], F[9];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        x[a] = s.x[3 * o + a];
        v[a] = s.v[3 * o + a];
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        B[a] = s.B[(int64_t)a * n + o];
        F[a] = s.F[(int64_t)a * n + o];
    }
    const float c = s.c[o], alpha = s.alpha[o], psi = s.psi[o], g = s.g[o], mass = s.mass[o];
    const int id = s.id[o];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        d.x[3 * i + a] = x[a];
        d.v[3 * i + a] = v[a];
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        d.B[(int64_t)a * n + i] = B[a];
        d.F[(int64_t)a * n + i] = F[a];
    }
    d.c[i] = c;
    d.alpha[i] = alpha;
    d.psi[i] = psi;
    d.g[i] = g;
    d.mass[i] = mass;
    d.id[i] = id;
}

__global__ void k_gather_mixed(const uint32_t* __restrict__ perm, MState s, MState d, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = (int)perm[i];
    float x[3], F32[9];
    uint16_t v[3], B[9], F[9];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        x[a] = s.x[3 * o + a];
        v[a] = s.v[3 * o + a];
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        B[a] = s.B[(int64_t)a * n + o];
        if (s.F32) F32[a] = s.F32[(int64_t)a * n + o];
        else F[a] = s.F[(int64_t)a * n + o];
    }
    const uint16_t c = s.c[o], alpha = s.alpha[o], psi = s.psi[o], g = s.g[o];
    const float mass = s.mass[o];
    const int id = s.id[o];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        d.x[3 * i + a] = x[a];
        d.v[3 * i + a] = v[a];
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        d.B[(int64_t)a * n + i] = B[a];
        if (s.F32) d.F32[(int64_t)a * n + i] = F32[a];
        else d.F[(int64_t)a * n + i] = F[a];
    }
    d.c[i] = c;
    d.alpha[i] = alpha;
    d.psi[i] = psi;
    d.g[i] = g;
    d.mass[i] = mass;
    d.id[i] = id;
}

// LSD radix passes over (keys, values): pass p sorts by bits 8p .. 8p+7.
// The keys start in k_in, the values are the items' indices (vals_in ==
// nullptr); the passes alternate between the two buffers of kb and vb so
// that the last one writes kb[1] and vb[1].
static hipError_t radix_passes(mpmx_ctx* c, const uint32_t* k_in, const uint32_t* vals_in, uint32_t* const (&kb)[2],
                               uint32_t* const (&vb)[2], RsSize z, int grid, int* hist_n) {
    const int passes = (c->key_bits + 7) / 8;
    const int64_t hist_cap = (int64_t)256 * c->sort_tiles;
    const uint32_t* ki = k_in;
    const uint32_t* vi = vals_in;
    for (int p = 0; p < passes; ++p) {
        const int o = (passes - 1 - p) % 2 == 0 ? 1 : 0;
        k_radix_hist<<<grid, kRsThreads, 0, c->stream>>>(ki, z, 8 * p, c->hist, hist_n);
        hipError_t e = z.n_dev ? scan_exclusive(c, (const int*)c->hist, (int*)c->hist, hist_cap, hist_n, hist_cap, nullptr)
                               : scan_exclusive(c, (const int*)c->hist, (int*)c->hist, (int64_t)256 * z.tiles,
                                                nullptr, (int64_t)256 * z.tiles, nullptr);
        if (e != hipSuccess) return e;
        k_radix_scatter<<<grid, kRsThreads, 0, c->stream>>>(ki, vi, kb[o], vb[o], z, 8 * p, c->hist);
        ki = kb[o];
        vi = vb[o];
    }
    return hipGetLastError();
}

// every key: c->keys_next radix-sorted into c->keys_alt, the permutation into
// c->vals_alt (c->keys serves as the other key buffer); the movers are
// counted on the way, so that the choice of path sees them
static hipError_t sort_radix(mpmx_ctx* c) {
    const int n = (int)c->n;
    if (c->old_keys_sorted) {
        k_mover_count<<<c->sort_tiles, kRsThreads, 0, c->stream>>>(c->keys, c->keys_next, n, c->mv_tcount);
        hipError_t e = scan_exclusive(c, c->mv_tcount, c->mv_toff, c->sort_tiles, nullptr, c->sort_tiles,
                                      &c->dc->movers);
        if (e != hipSuccess) return e;
    }
    uint32_t* const kb[2] = {c->keys, c->keys_alt};
    uint32_t* const vb[2] = {c->vals, c->vals_alt};
    hipError_t e = radix_passes(c, c->keys_next, nullptr, kb, vb, RsSize{n, nullptr, c->sort_tiles}, c->sort_tiles,
                                nullptr);
    c->gather_perm = c->vals_alt;
    return e;
}

// the movers only (see "mover sort" above): the same order as sort_radix,
// from c->keys (the previous sort's keys) and c->keys_next, into c->keys_alt
// and c->vals_alt
static hipError_t sort_movers(mpmx_ctx* c) {
    const int n = (int)c->n;
    const int tiles = c->sort_tiles;
    k_mover_count<<<tiles, kRsThreads, 0, c->stream>>>(c->keys, c->keys_next, n, c->mv_tcount);
    hipError_t e = scan_exclusive(c, c->mv_tcount, c->mv_toff, tiles, nullptr, tiles, &c->dc->movers);
    if (e != hipSuccess) return e;
    // the movers' radix passes end in mv_key[1] / mv_val[1]; the list goes
    // to the buffer that the first pass does not write
    const int passes = (c->key_bits + 7) / 8;
    uint32_t* const list = c->mv_key[(passes - 1) % 2 == 0 ? 0 : 1];
    k_mover_list<<<tiles, kRsThreads, 0, c->stream>>>(c->keys, c->keys_next, n, c->mv_toff, list, c->mv_old,
                                                      c->mv_slot, c->mv_cpre, c->mv_cmask);
    const int grid = std::min(tiles, 4 * c->num_cu);
    uint32_t* const kb[2] = {c->mv_key[0], c->mv_key[1]};
    uint32_t* const vb[2] = {c->mv_val[0], c->mv_val[1]};
    e = radix_passes(c, list, nullptr, kb, vb, RsSize{n, &c->dc->movers, tiles}, grid, &c->dc->mv_hist_n);
    if (e != hipSuccess) return e;
    // one thread per mover up to n / 8 movers (the movers' searches are
    // chains of dependent loads: latency, hidden by many waves)
    const OldTables tb{c->pmask, c->wcount, c->page2particle, c->cell_start, &c->dc->num_page, c->cap_p};
    k_place_movers<<<tiles, kRsThreads, 0, c->stream>>>(c->keys, n, c->mv_key[1], c->mv_val[1], c->mv_old,
                                                       c->mv_slot, &c->dc->movers, c->mv_cpre, c->mv_cmask,
                                                       c->mv_toff, tiles, c->mv_tbound, tb, c->tables_match_keys,
                                                       c->keys_alt, c->vals_alt);
    k_place_stayers<<<tiles, kRsThreads, 0, c->stream>>>(c->keys, c->keys_next, n, c->mv_toff, c->mv_tbound,
                                                         c->mv_key[1], c->mv_val[1], &c->dc->movers, c->keys_alt,
                                                         c->vals_alt);
    c->gather_perm = c->vals_alt;
    return hipGetLastError();
}

void choose_sort_path(mpmx_ctx* c) {
    if (c->cfg.sort_variant != 0) {
        c->sort_by_movers = c->cfg.sort_variant == 2;
        return;
    }
    const int64_t m = c->hc->movers;
    if (m > c->n / 4) c->sort_by_movers = false;
    else if (m < c->n / 8) c->sort_by_movers = true;
}

// the sort alone: the keys of the particles' slots (c->keys_next) in key
// order become c->keys, the permutation (new -> old) c->gather_perm; the
// particle state still waits in c->cur.  The mover path needs the previous
// sort's keys, so the first sort after an upload takes the radix path.
hipError_t sort_keys(mpmx_ctx* c) {
    choose_sort_path(c);
    const bool movers = c->sort_by_movers && c->old_keys_sorted;
    hipError_t e = movers ? sort_movers(c) : sort_radix(c);
    if (e != hipSuccess) return e;
    std::swap(c->keys, c->keys_alt);
    c->old_keys_sorted = true;
    c->tables_match_keys = false;
    return hipSuccess;
}

// the state gather that sort_keys left pending, on its own
hipError_t finish_gather(mpmx_ctx* c) {
    const uint32_t* perm = c->gather_perm;
    if (!perm) return hipSuccess;
    const int n = (int)c->n;
    const int nxt = 1 - c->cur;
    if (c->cfg.precision != MPMX_FP32)
        k_gather_mixed<<<(n + 255) / 256, 256, 0, c->stream>>>(perm, c->ms[c->cur], c->ms[nxt], n);
    else
        k_gather_fp32<<<(n + 255) / 256, 256, 0, c->stream>>>(perm, c->ps[c->cur], c->ps[nxt], n);
    c->cur = nxt;
    c->gather_perm = nullptr;
    return hipGetLastError();
}

hipError_t sort_and_gather(mpmx_ctx* c) {
    hipError_t e = sort_keys(c);
    return e == hipSuccess ? finish_gather(c) : e;
}

}  // namespace mpmx
