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
    k_keys<<<(n + 255) / 256, 256, 0, c->stream>>>(c->dp, x, c->keys, n, c->scene_start, c->dc);
    return hipGetLastError();
}

// ---------------------------------------------------------------- radix sort
constexpr int kRsThreads = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsThreads * kRsItems;

// per-tile digit histogram, digit-major: hist[d * tiles + t]
__global__ __launch_bounds__(kRsThreads) void k_radix_hist(const uint32_t* __restrict__ keys, int n,
                                                           int shift, uint32_t* __restrict__ hist,
                                                           int tiles) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRsTile;
#pragma unroll
    for (int j = 0; j < kRsItems; ++j) {
        const int i = base + j * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// stable scatter: items are ranked in (round, wave, lane) order, which is their
// index order inside the tile (same-digit peers inside a wave are found with
// eight 64-lane ballots), and put in their digit's run of an LDS copy of the
// tile; the tile then leaves in that order, so that consecutive lanes write
// consecutive positions of a digit's run instead of 256 scattered words per
// round.  The tile's count of digit d is the scanned histogram's next
// digit-major entry minus its own (n after the last).  All items of the tile
// are loaded before the ranking starts.
__global__ __launch_bounds__(kRsThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t* __restrict__ okeys,
    uint32_t* __restrict__ ovals, int n, int shift, const uint32_t* __restrict__ offs, int tiles) {
    __shared__ uint32_t gstart[256];   // output position of the tile's first item of digit d
    __shared__ uint32_t lstart[256];   // its slot in the tile copy
    __shared__ uint32_t lnext[256];    // next free slot of digit d
    __shared__ uint32_t wcnt[4][256];
    __shared__ uint32_t skey[kRsTile], sval[kRsTile];
    __shared__ int lds4[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int base = blockIdx.x * kRsTile;
    uint32_t k[kRsItems], v[kRsItems];
#pragma unroll
    for (int j = 0; j < kRsItems; ++j) {
        const int i = base + j * kRsThreads + tid;
        k[j] = i < n ? keys[i] : 0u;
        v[j] = i < n ? (vals ? vals[i] : (uint32_t)i) : 0u;
    }
    {
        const int64_t e = (int64_t)tid * tiles + blockIdx.x;
        const uint32_t g0 = offs[e];
        const uint32_t g1 = e + 1 < (int64_t)256 * tiles ? offs[e + 1] : (uint32_t)n;
        int total;
        const int ls = block_excl_scan256((int)(g1 - g0), total, lds4);
        gstart[tid] = g0;
        lstart[tid] = (uint32_t)ls;
        lnext[tid] = (uint32_t)ls;
    }
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
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

// the radix passes alone: sorted keys in c->keys and the permutation
// (new -> old) in c->gather_perm; the particle state still waits in c->cur
hipError_t sort_keys(mpmx_ctx* c) {
    const int n = (int)c->n;
    const int tiles = c->sort_tiles;
    uint32_t* kb[2] = {c->keys, c->keys_alt};
    uint32_t* vb[2] = {c->vals, c->vals_alt};
    int ki = 0;    // buffer holding the current keys
    int vi = -1;   // buffer holding the current permutation (-1: identity)
    const int passes = (c->key_bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        const int ko = 1 - ki;
        const int vo = (vi == 1) ? 0 : 1;
        k_radix_hist<<<tiles, kRsThreads, 0, c->stream>>>(kb[ki], n, shift, c->hist, tiles);
        hipError_t e = scan_exclusive(c, (const int*)c->hist, (int*)c->hist, (int64_t)256 * tiles,
                                      nullptr, (int64_t)256 * tiles, nullptr);
        if (e != hipSuccess) return e;
        k_radix_scatter<<<tiles, kRsThreads, 0, c->stream>>>(kb[ki], vi < 0 ? nullptr : vb[vi], kb[ko],
                                                             vb[vo], n, shift, c->hist, tiles);
        ki = ko;
        vi = vo;
    }
    // the sorted keys become c->keys (an odd pass count swaps the two buffers,
    // as sort_swaps_keys says); the permutation (new -> old) feeds the gather
    // from whichever buffer holds it
    if (ki != 0) std::swap(c->keys, c->keys_alt);
    const uint32_t* perm = vi < 0 ? nullptr : vb[vi];
    if (!perm) return hipErrorInvalidValue;   // no pass ran: key_bits is at least 1
    c->gather_perm = perm;
    return hipGetLastError();
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

bool sort_swaps_keys(const mpmx_ctx* c) { return ((c->key_bits + 7) / 8) % 2 == 1; }

}  // namespace mpmx
