// k_tables.hip — sparse page tables from the sorted particle keys.
//
// Replaces DomainTransformer::rebuild (Transformer.cpp:71-82,
// Transformer_kernal.cu:15-209).  The reference numbers pages through a hash
// table with atomicAdd, so ids depend on scheduling.  Here the occupied pages
// and their +x/+y/+z neighbours are set in two dense bitmasks over the Morton
// page space (one bit per page: 2 MiB at 512^3), and a page's id is its rank
// in the mask: particle pages first, in ascending key order, then
// neighbour-only pages, also ascending.  Ranks are popcounts plus a scanned
// per-word prefix, so the adjacency lookup needs no hash probing.  All sizes
// stay on the device (DevCounts), so the whole build can be graph-captured.
#include "mpmx_internal.h"

namespace mpmx {

// neighbour q (0..6) of a page = offset (i,j,k) with q+1 = i*4 + j*2 + k
// (Transformer.cpp:27-34)
__device__ __forceinline__ uint32_t neighbor_key(uint32_t pk, int q) {
    int x, y, z;
    unmorton(pk, x, y, z);
    const int o = q + 1;
    return morton((uint32_t)(x + ((o >> 2) & 1)), (uint32_t)(y + ((o >> 1) & 1)), (uint32_t)(z + (o & 1)));
}

// markPageBoundary + buildHashMapFromPage + supplementAdjacentPages, as bits:
// the first particle of each page sets the page's byte flag (a plain store,
// one writer per byte), then one thread per word packs its 64 flags into the
// bitmap and clears them for the next build.  Setting the bits directly takes
// an atomic per page on shared words, which costs more than both passes.
__global__ void k_mark_pages(const uint32_t* __restrict__ keys, int n, uint8_t* __restrict__ pflag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t pk = keys[i] >> 6;
    if (i > 0 && (keys[i - 1] >> 6) == pk) return;
    pflag[pk] = 1;
}
__global__ void k_pack_pages(uint64_t* __restrict__ pflag, uint64_t* __restrict__ pmask, int64_t words) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    uint64_t* f = pflag + 8 * w;
    uint64_t m = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const uint64_t b = f[q];   // flags of pages 8q .. 8q + 7 of the word, one byte each
#pragma unroll
        for (int j = 0; j < 8; ++j) m |= ((b >> (8 * j)) & 1ull) << (8 * q + j);
        if (b) f[q] = 0ull;   // most of a batch's page space is empty: clear only what was set
    }
    pmask[w] = m;
}

// supplementAdjacentPages as a bit dilation: nmask = the particle pages P plus
// P + e for e in {0,1}^3 (the pages a particle page's nodes reach).  A 64-bit
// word holds the 4x4x4 page cube of one coarse Morton code; inside it the
// local index has z at bits 0 and 3, y at 1 and 4, x at 2 and 5.  +1 along an
// axis moves a page with low bit 0 up by 2^a, one with low bit 1 and high bit
// 0 up by 7 2^a, and the top layer (both bits set) leaves the word for the
// next word along that axis, landing 9 2^a lower in its local index.
__host__ __device__ constexpr uint64_t axis_bits_where(int a, int lo, int hi) {
    uint64_t m = 0;
    for (int i = 0; i < 64; ++i)
        if (((i >> a) & 1) == lo && (hi < 0 || ((i >> (a + 3)) & 1) == hi)) m |= 1ull << i;
    return m;
}
// m dilated by +1 along axis a, prev = the word below along that axis
template <int A>
__device__ __forceinline__ uint64_t dilate(uint64_t m, uint64_t prev) {
    constexpr uint64_t lo0 = axis_bits_where(A, 0, -1), lo1hi0 = axis_bits_where(A, 1, 0),
                       top = axis_bits_where(A, 1, 1);
    return m | ((m & lo0) << (1 << A)) | ((m & lo1hi0) << (7 << A)) | ((prev & top) >> (9 << A));
}

// Each scene's dilation stays inside its slot and below its own page limit
// (the page coordinates its domain alone would hold), so a scene of a batch
// gets the neighbour pages it would get alone, less the ones past the slot,
// which no particle of the domain reaches.
// per word: particle pages, neighbour-only pages and occupied 2x2x2 blocks
// (one byte of the bitmaps each: the 8 pages 8b .. 8b+7 of block b)
__device__ __forceinline__ void word_counts_of(int64_t w, uint64_t P, uint64_t N, int* __restrict__ wcount,
                                               int* __restrict__ scount, int* __restrict__ bcount) {
    wcount[w] = __popcll(P);
    scount[w] = __popcll(N & ~P);
    const uint64_t m = P | N;
    int c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) c += ((m >> (8 * k)) & 0xffull) != 0ull;
    bcount[w] = c;
}

// the neighbour bitmap word and the word's counts
__global__ void k_neighbour_mask(DevParams dp, const uint64_t* __restrict__ pmask, uint64_t* __restrict__ nmask,
                                 int* __restrict__ wcount, int* __restrict__ scount, int* __restrict__ bcount,
                                 int64_t words) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    const int wshift = dp.slot_shift - 6;   // 2^wshift words per slot
    const int64_t local = w & ((1ll << wshift) - 1);
    const int lim = dp.scenes[(int)(w >> wshift)].page_limit >> 2;   // words per axis
    int cx, cy, cz;
    unmorton((uint32_t)local, cx, cy, cz);
    if (cx >= lim || cy >= lim || cz >= lim) {
        nmask[w] = 0ull;
        word_counts_of(w, pmask[w], 0ull, wcount, scount, bcount);
        return;
    }
    const int64_t slot_base = w - local;
    auto P = [&](int dx, int dy, int dz) -> uint64_t {
        const int x = cx - dx, y = cy - dy, z = cz - dz;
        if (x < 0 || y < 0 || z < 0) return 0ull;
        return pmask[slot_base + morton((uint32_t)x, (uint32_t)y, (uint32_t)z)];
    };
    // dilate along z the four words the y and x steps need, then along y, then x
    const uint64_t a00 = dilate<0>(P(0, 0, 0), P(0, 0, 1)), a01 = dilate<0>(P(0, 1, 0), P(0, 1, 1));
    const uint64_t a10 = dilate<0>(P(1, 0, 0), P(1, 0, 1)), a11 = dilate<0>(P(1, 1, 0), P(1, 1, 1));
    const uint64_t b0 = dilate<1>(a00, a01), b1 = dilate<1>(a10, a11);
    const uint64_t N = dilate<2>(b0, b1);
    nmask[w] = N;
    word_counts_of(w, pmask[w], N, wcount, scount, bcount);
}


__device__ __forceinline__ void scene_counts_of(const DevParams& dp, int s, const int* __restrict__ wpre,
                                                const int* __restrict__ spre, SceneCG* __restrict__ scg,
                                                const DevCounts* __restrict__ dc);

// page keys in id order, the block keys (2x2x2 blocks in Morton order) and
// each scene's page ranges, one thread per bitmap word
__global__ void k_page_keys(DevParams dp, const uint64_t* __restrict__ pmask, const uint64_t* __restrict__ nmask,
                            const int* __restrict__ wpre, const int* __restrict__ spre,
                            const int* __restrict__ bpre, int64_t words, uint32_t* page_key, int64_t cap_p,
                            int64_t cap_tp, uint32_t* __restrict__ block_key, int64_t cap_blocks,
                            SceneCG* __restrict__ scg, DevCounts* dc) {
    // a wave per 64 consecutive words: one lane per word finds the words that
    // hold a page (a batch's page space is mostly empty), then the wave takes
    // those words one at a time with one lane per bit: a set bit's id is the
    // word's prefix plus the set bits below it.  The 2x2x2-page blocks of the
    // gather transfer: the 8 pages of block b are the Morton codes 8b .. 8b+7,
    // one byte of the bitmaps, so a block is occupied when that byte of
    // pmask | nmask is non-zero, and blocks are numbered in ascending Morton
    // order by the scan of the per-word block counts.
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int b = (int)(t & 63);
    const int num_page = dc->num_page;
    if (t == 0) {
        const int total = num_page + dc->num_supp;
        dc->num_total = total;
        if (num_page > cap_p || total > cap_tp) atomicOr(&dc->err, kErrCapacity);
    }
    if (t < dp.num_scenes) scene_counts_of(dp, (int)t, wpre, spre, scg, dc);
    const int64_t w0 = (t >> 6) * 64;   // the wave's first word
    uint64_t Pl = 0ull, Nl = 0ull;
    int wl = 0, sl = 0, bl = 0;         // the word's prefixes, loaded with it (the loop below only shuffles)
    if (w0 + b < words) {
        Pl = pmask[w0 + b];
        Nl = nmask[w0 + b];
        wl = wpre[w0 + b];
        sl = spre[w0 + b];
        bl = bpre[w0 + b];
    }
    uint64_t busy = __ballot((Pl | Nl) != 0ull);
    const uint64_t below = b == 0 ? 0ull : (~0ull >> (64 - b));
    while (busy) {
        const int j = __ffsll((long long)busy) - 1;
        busy &= busy - 1ull;
        const int64_t w = w0 + j;
        const uint64_t P = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(Pl >> 32), j, 64) << 32) |
                           (uint64_t)(uint32_t)__shfl((int)(uint32_t)Pl, j, 64);
        const uint64_t N = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(Nl >> 32), j, 64) << 32) |
                           (uint64_t)(uint32_t)__shfl((int)(uint32_t)Nl, j, 64);
        const int wp = __shfl(wl, j, 64), sp = __shfl(sl, j, 64), bp = __shfl(bl, j, 64);
        if ((P >> b) & 1ull) {
            const int id = wp + __popcll(P & below);
            if (id < cap_p && id < cap_tp) page_key[id] = (uint32_t)(w * 64 + b);
        }
        const uint64_t S = N & ~P;
        if ((S >> b) & 1ull) {
            const int id = num_page + sp + __popcll(S & below);
            if (id < cap_tp) page_key[id] = (uint32_t)(w * 64 + b);
        }
        if (b < 8) {
            const uint64_t m = P | N;
            if ((m >> (8 * b)) & 0xffull) {
                int id = bp;
                for (int k = 0; k < b; ++k) id += ((m >> (8 * k)) & 0xffull) != 0ull;
                if (id < cap_blocks) block_key[id] = (uint32_t)(w * 8 + b);
            }
        }
    }
}

// markBlockOffset on page boundaries (Transformer_kernal.cu:22-36): first particle of each page
// first particle of every cell of every particle page (cell2particle of
// Transformer_kernal.cu:15-36, per page): cell_start[p * 65 + c] for c in
// [0, 64], the first particle of the page with cell >= c (the page's end for
// c = 64).  Particles of a page are sorted by cell, so particle i is that
// first particle for the cells after its predecessor's cell up to its own;
// the page's last particle also fills the cells after its own with the end.
// Every entry is written once, without atomics.
// The same pass writes page2particle (the first particle of each page,
// markPageBoundary + the page head scan of Transformer_kernal.cu:38-47).
// Each thread takes kHeadItems particles (a block's items consecutive per
// round, so loads stay coalesced) and issues the key loads of all of them,
// then the bitmap and prefix loads of those that start a cell, before it
// uses any: the two dependent steps are paid once per thread, not per item.
constexpr int kHeadItems = 4;
__global__ void k_page_heads(const uint32_t* __restrict__ keys, int n, const uint64_t* __restrict__ pmask,
                             const uint64_t* __restrict__ nmask, const int* __restrict__ wpre,
                             const int* __restrict__ spre, int* __restrict__ page2particle,
                             int* __restrict__ cell_start, int64_t cap_p, const DevCounts* __restrict__ dc) {
    const int num_page = dc->num_page;
    const int np = num_page < cap_p ? num_page : (int)cap_p;
    const int base = blockIdx.x * blockDim.x * kHeadItems + threadIdx.x;
    uint32_t k[kHeadItems], kp[kHeadItems], kn[kHeadItems];
#pragma unroll
    for (int j = 0; j < kHeadItems; ++j) {
        const int i = base + j * (int)blockDim.x;
        k[j] = i < n ? keys[i] : 0u;
        kp[j] = (i > 0 && i < n) ? keys[i - 1] : ~0u;
        kn[j] = i < n - 1 ? keys[i + 1] : ~0u;
    }
    bool need[kHeadItems];
    uint64_t P[kHeadItems], N[kHeadItems];
    int wp[kHeadItems], sp[kHeadItems];
#pragma unroll
    for (int j = 0; j < kHeadItems; ++j) {
        const int i = base + j * (int)blockDim.x;
        const uint32_t pk = k[j] >> 6;
        const bool first = i == 0 || (kp[j] >> 6) != pk;
        const bool last = i == n - 1 || (kn[j] >> 6) != pk;
        const int prev = first ? -1 : (int)(kp[j] & 63u);
        need[j] = i < n && (prev != (int)(k[j] & 63u) || last);
        P[j] = N[j] = 0ull;
        wp[j] = sp[j] = 0;
        if (need[j]) {
            const uint32_t w = pk >> 6;
            P[j] = pmask[w];
            N[j] = nmask[w];
            wp[j] = wpre[w];
            sp[j] = spre[w];
        }
    }
#pragma unroll
    for (int j = 0; j < kHeadItems; ++j) {
        const int i = base + j * (int)blockDim.x;
        if (i == n - 1 && num_page <= cap_p) page2particle[num_page] = n;
        if (!need[j]) continue;
        const uint32_t pk = k[j] >> 6, b = pk & 63u;
        const int cell = (int)(k[j] & 63u);
        const bool first = i == 0 || (kp[j] >> 6) != pk;
        const bool last = i == n - 1 || (kn[j] >> 6) != pk;
        const int prev = first ? -1 : (int)(kp[j] & 63u);
        // page_rank from the preloaded words
        const uint64_t below = b ? ((~0ull) >> (64 - b)) : 0ull;
        int id = -1;
        if ((P[j] >> b) & 1ull) id = wp[j] + __popcll(P[j] & below);
        else if (((N[j] & ~P[j]) >> b) & 1ull) id = num_page + sp[j] + __popcll(N[j] & ~P[j] & below);
        if (id < 0 || id >= cap_p) continue;
        if (first) page2particle[id] = i;
        if (id >= np) continue;
        int* cs = cell_start + (int64_t)id * 65;
        for (int c = prev + 1; c <= cell; ++c) cs[c] = i;
        if (last)
            for (int c = cell + 1; c <= 64; ++c) cs[c] = i + 1;
    }
}

// establishPageTopology + markPageSize (Transformer_kernal.cu:98-120)
__global__ void k_page_topology(const uint32_t* __restrict__ page_key, const int* __restrict__ page2particle,
                                const uint64_t* __restrict__ pmask, const uint64_t* __restrict__ nmask,
                                const int* __restrict__ wpre, const int* __restrict__ spre, int* adj,
                                int* nvp, int64_t cap_p, int vpage, DevCounts* dc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int num_page = dc->num_page;
    num_page = num_page < cap_p ? num_page : (int)cap_p;
    if (p >= num_page) return;
    const uint32_t pk = page_key[p];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        // -1 only past the edge of a batched scene's page space, which no
        // node of the scene reaches
        adj[(int64_t)q * cap_p + p] = page_rank(neighbor_key(pk, q), pmask, nmask, wpre, spre, dc->num_page);
    }
    const int cnt = page2particle[p + 1] - page2particle[p];
    nvp[p] = (cnt + vpage - 1) / vpage;
}

// markVirtualPageOffset + inclusive scan (Transformer_kernal.cu:122-126, 196-208)
__global__ void k_vtargets(int* vpo, int* target, int64_t cap_p, int64_t cap_vp, DevCounts* dc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int num_page = dc->num_page;
    num_page = num_page < cap_p ? num_page : (int)cap_p;
    if (p >= num_page) return;
    const int nv = dc->num_vp;
    if (p == 0 && nv > cap_vp) atomicOr(&dc->err, kErrCapacity);
    const int b = vpo[p];
    const int e = (p + 1 < num_page) ? vpo[p + 1] : nv;
    if (p == num_page - 1) vpo[num_page] = nv;
    for (int v = b; v < e && v < cap_vp; ++v) target[v] = p;
}

// page ranges of every scene: its particle pages and neighbour-only pages are
// the ranks of its slot's words in the two prefix scans
__device__ __forceinline__ void scene_counts_of(const DevParams& dp, int s, const int* __restrict__ wpre,
                                                const int* __restrict__ spre, SceneCG* __restrict__ scg,
                                                const DevCounts* __restrict__ dc) {
    const int wshift = dp.slot_shift - 6;
    const int64_t w0 = (int64_t)s << wshift, w1 = (int64_t)(s + 1) << wshift;
    const bool last = s + 1 == dp.num_scenes;
    const int p0 = wpre[w0], p1 = last ? dc->num_page : wpre[w1];
    const int n0 = spre[w0], n1 = last ? dc->num_supp : spre[w1];
    scg[s].pfirst = p0;
    scg[s].np = p1 - p0;
    scg[s].nfirst = n0;
    scg[s].nn = n1 - n0;
}
__global__ void k_scene_counts(DevParams dp, const int* __restrict__ wpre, const int* __restrict__ spre,
                               SceneCG* __restrict__ scg, const DevCounts* __restrict__ dc) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < dp.num_scenes) scene_counts_of(dp, s, wpre, spre, scg, dc);
}

hipError_t launch_scene_counts(mpmx_ctx* c) {
    k_scene_counts<<<(c->num_scenes + 255) / 256, 256, 0, c->stream>>>(c->dp, c->wcount, c->scount, c->scg, c->dc);
    return hipGetLastError();
}

// page adjacency and virtual pages (establishPageTopology, markPageSize,
// markVirtualPageOffset and the target pages, Transformer_kernal.cu:98-126):
// the page-tile transfer kernels (p2g_variant 1) use them; the block kernels
// do not, so for those they are built only when asked for
hipError_t build_vpage_tables(mpmx_ctx* c) {
    hipError_t e;
    const int pblocks = (int)((c->cap_p + 255) / 256);
    k_page_topology<<<pblocks, 256, 0, c->stream>>>(c->page_key, c->page2particle, c->pmask, c->nmask,
                                                     c->wcount, c->scount, c->adj, c->nvp, c->cap_p,
                                                     c->P.vpage, c->dc);
    if ((e = scan_exclusive(c, c->nvp, c->vpo, c->cap_p, &c->dc->num_page, 0, &c->dc->num_vp)) != hipSuccess)
        return e;
    k_vtargets<<<pblocks, 256, 0, c->stream>>>(c->vpo, c->target, c->cap_p, c->cap_vp, c->dc);
    c->vpages_valid = true;
    return hipGetLastError();
}

hipError_t build_tables(mpmx_ctx* c) {
    const int n = (int)c->n;
    const int64_t words = c->mask_words;
    hipError_t e;
    c->order_ok = false;   // new blocks: identity order until the fused stress pass orders them
    k_mark_pages<<<(n + 255) / 256, 256, 0, c->stream>>>(c->keys, n, reinterpret_cast<uint8_t*>(c->pflag));
    k_pack_pages<<<(int)((words + 255) / 256), 256, 0, c->stream>>>(c->pflag, c->pmask, words);
    k_neighbour_mask<<<(int)((words + 255) / 256), 256, 0, c->stream>>>(c->dp, c->pmask, c->nmask, c->wcount,
                                                                         c->scount, c->bcount, words);
    ScanSet3 sc{{c->wcount, c->scount, c->bcount}, {c->wcount, c->scount, c->bcount},
                {&c->dc->num_page, &c->dc->num_supp, &c->dc->num_blocks}};
    if ((e = scan_exclusive3(c, sc, 3, words)) != hipSuccess) return e;
    // one thread per word (at least one workgroup: its first threads also
    // write the per-scene page ranges)
    k_page_keys<<<(int)std::max<int64_t>(1, (words + 255) / 256), 256, 0, c->stream>>>(
        c->dp, c->pmask, c->nmask, c->wcount, c->scount, c->bcount, words, c->page_key, c->cap_p, c->cap_tp,
        c->block_key, c->tp_bound, c->scg, c->dc);
    k_page_heads<<<(n + 256 * kHeadItems - 1) / (256 * kHeadItems), 256, 0, c->stream>>>(c->keys, n, c->pmask, c->nmask, c->wcount, c->scount,
                                                          c->page2particle, c->cell_start, c->cap_p, c->dc);
    c->vpages_valid = false;
    c->tables_match_keys = true;
    if (c->cfg.p2g_variant == 1 && (e = build_vpage_tables(c)) != hipSuccess) return e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return c->cfg.p2g_variant == 1 ? hipSuccess : launch_block_lists(c);
}

}  // namespace mpmx
