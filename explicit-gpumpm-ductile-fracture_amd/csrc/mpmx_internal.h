// mpmx_internal.h — context layout and device helpers shared by the kernels of
// libmpmx_hip.so.  Not part of the public ABI (include/mpmx.h is).
//
// HBM layout (FP32 plain path), all buffers owned by the context:
//  particles  (sorted "current" order, double-buffered for the gather):
//      x, v       float[3n] AoS (reference vector3T layout)
//      B, F       float[9n] component-major: M(r,c) of particle i at m[(r+3c)*n + i]
//                 (reference d_F/d_B layout, Particles.cpp:92-93)
//      c, alpha, psi, g, mass  float[n];  id int[n] (original particle id)
//  per-substep: sigma float[9n], FP float[n], vol float[n]
//  keys       uint32[n]: (Morton page key << 6) | cell, order-equivalent to the
//             reference's 64-bit SPGrid offset (Particle_kernal.cu:158-165)
//  pages      page_key uint32[cap_tp] (particle pages ascending, then
//             neighbour-only pages ascending), page2particle int[cap_p+1],
//             adj int[7*cap_p], vpo int[cap_p+1], target int[cap_vp] (0-based)
//  grid       float[cap_tp][8][64]: per page, 8 channels of 64 cells (256 B each)
//             0 m | 1-3 momentum -> velocity | 4 sum w c -> c | 5 sum w |
//             6 sum V w -> PCG rhs | 7 sum V FP w -> Jacobi diagonal
//  PCG        x, r, z, s, q float[cap_tp*64] page-major (the reference's d_grid_*);
//             partials float[4][cg_nb]: s.q, r.z (two, by iteration parity), max|r|
//  blocks     block_key uint32[tp_bound]: occupied 2x2x2-page blocks (page key >> 3)
//  scenes     a context holds one scene or a batch of independent scenes.  Scene
//             s occupies slot s of the Morton page space: a cube of 2^slot_log2
//             pages per axis whose page keys are [s << slot_shift, (s+1) <<
//             slot_shift), so a page's scene is its key >> slot_shift, the
//             sorted particles of a scene are contiguous, and every 2x2x2
//             block belongs to one scene.  Positions stay in the scene's own
//             units; keys add the slot origin.  Per-scene constants: scenes[s]
//             (DevParams); per-scene PCG state and page ranges: SceneCG[s].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/mpmx.h"

namespace mpmx {

constexpr int kGridCh = 8;       // channels per page in the grid slab
constexpr int kPageFloats = kGridCh * 64;
constexpr int kTile = 6;         // 6^3 node tile of one page's particles
// bytes of one block's record of the owner-block P2G (k_grid.hip): a 64-byte
// header and 784 nodes x 8 channels of 64-bit sums
constexpr int64_t kOwnRecBytes = 64 + 8ll * 8 * 784;
constexpr int kTileN = kTile * kTile * kTile;

enum ErrBits : unsigned {
    kErrDomain = 1u,
    kErrNonFinite = 2u,
    kErrCapacity = 4u,
    kErrBreakdown = 8u,
};

// per-scene constants, passed by value to every kernel
struct DevParams {
    int N;
    int transfer;
    int phase_field;   // the solve runs (phase_field set and the energy has fracture)
    int model;         // 2 von Mises, 1 Neo-Hookean phase field, 0 fixed corotated
    int dot_mode;
    int wall;
    int grip_lo, grip_hi;
    int vpage;
    float dx, inv_dx, Dinv;
    float grip_v;
    float lambda, mu, kappa, V0;
    float Cinv[9];
    float yield, hard, alpha_c, dlambda_cap, eps_cap;
    float Gc0, residual_c, ps, theta, M;
    float gravity[3];
    float vmax, cg_tol, crit_stretch;
    float L0;
    float sq23f;   // sqrt(2/3) as float
    float svd_gap; // FP32 stress arithmetic: FP64 redo below this singular-value gap
    // scenes: constants of scene s at scenes[s] (the context's own copy of
    // this struct is scene 0 and carries the batch-wide fields)
    const DevParams* scenes;
    int num_scenes;
    int slot_shift;     // scene of a page key = key >> slot_shift
    int origin[3];      // this scene's cell (0,0,0) in the page space, per axis
    int page_limit;     // neighbour pages exist below this page coordinate (per axis, scene-local)
};

// per-scene PCG state and page ranges (device array, one per scene).  The
// scene's pages are its particle pages [pfirst, pfirst + np) and its
// neighbour-only pages [num_page + nfirst, num_page + nfirst + nn) of the
// context's page numbering; scene-local page l is the l-th of that list, as
// the scene alone would number it.  Workgroups [wg0, wg0 + nwg) of every
// node-space PCG kernel serve the scene.
struct SceneCG {
    int done;
    int iters;
    int breakdown;
    float norm;
    int pfirst, np, nfirst, nn;
    int wg0, nwg;
    int pad[6];
};

// device-resident counters and PCG state.  The PCG scalars themselves are not
// here: every node kernel writes one partial per workgroup (cg_part) and the
// kernels that need a scalar reduce those partials in a fixed order, so the
// solve has no same-address atomics and is bit-reproducible.
struct DevCounts {
    int num_page;
    int num_total;
    int num_vp;
    int num_supp;
    unsigned err;
    int cg_done;
    int cg_iters;
    int cg_breakdown;
    float cg_norm;      // max|r| at the last convergence test
    int num_cell;
    int svd_fallbacks;
    int num_blocks;     // occupied 2x2x2-page blocks
    int cg_active;      // scenes whose PCG is still running
    float dt;           // time step of the substeps launched with dt <= 0 (set by the CFL kernels)
    int movers;         // particles whose key changed since the last sort (mover sort)
    int mv_hist_n;      // digit-histogram entries of the movers' radix passes
    int pad[13];
    double time;        // simulated time summed by the CFL kernels
    unsigned seq;       // host-published copies: written last, names the copy
};

// time step of a substep kernel: the launch argument, or (argument <= 0) the
// step the device-side CFL computed for this substep
__device__ __forceinline__ float substep_dt(float dt, const DevCounts* dc) { return dt > 0.f ? dt : dc->dt; }

struct PState {
    float* x = nullptr;
    float* v = nullptr;
    float* B = nullptr;
    float* F = nullptr;
    float* c = nullptr;
    float* alpha = nullptr;
    float* psi = nullptr;
    float* g = nullptr;
    float* mass = nullptr;
    int* id = nullptr;
};

// mixed-precision particle storage: x float AoS, F float (MPMX_MIXED) or
// half(F - I) (MPMX_MIXED_F16), the rest bf16.  Exactly one of F32 and F is
// set; kernels branch on F32 (a kernel argument, so the branch is uniform).
struct MState {
    float* x = nullptr;
    uint16_t* v = nullptr;   // bf16[3n] AoS
    uint16_t* B = nullptr;   // bf16[9n]
    uint16_t* F = nullptr;   // fp16[9n] of F - I (MPMX_MIXED_F16)
    float* F32 = nullptr;    // float[9n] (MPMX_MIXED)
    uint16_t* c = nullptr;   // bf16
    uint16_t* alpha = nullptr;
    uint16_t* psi = nullptr;
    uint16_t* g = nullptr;
    float* mass = nullptr;
    int* id = nullptr;
};

// Lanes -> particles of the per-particle kernels.  Every wave holds
// particles of one scene: a batch pads each scene's (contiguous) particle
// range to whole waves, so a kernel reads its scene's constants once per wave,
// through the scalar cache.  One scene: wave w holds particles 64w .. 64w+63.
struct WaveMap {
    const int* first;   // per wave: first particle; nullptr for one scene
    const int* count;   // per wave: particles in the wave (<= 64)
    const int* scene;   // per wave: scene
    int64_t waves;
};
struct TimerEntry {
    std::string name;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};

}  // namespace mpmx

struct mpmx_ctx {
    mpmx_params P;
    mpmx_config cfg;
    mpmx::DevParams dp;
    // scenes (a single-scene context is a batch of one)
    int num_scenes = 1;
    std::vector<mpmx_params> scene_P;        // per-scene parameters as given
    std::vector<int64_t> scene_n;            // particles per scene
    mpmx::DevParams* scene_dp = nullptr;     // device [num_scenes]
    mpmx::SceneCG* scg = nullptr;            // device [num_scenes]
    int* wg_scene = nullptr;                 // device [cg_nb]: scene of each PCG workgroup
    int* scene_start = nullptr;              // device [num_scenes + 1]: upload-order particle ranges
    mpmx::WaveMap wmap{};                    // per-particle kernels' waves (batches only)
    int* wave_buf = nullptr;                 // device storage of wmap
    int slot_log2 = 0;                       // pages per slot axis = 2^slot_log2
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0, cap_n = 0;
    int cur = 0;
    mpmx::PState ps[2];
    mpmx::MState ms[2];
    float* sigma = nullptr;
    float* FP = nullptr;
    float* vol = nullptr;
    bool stress_fields = false;      // sigma and FP hold the last stress update
    double* red_part = nullptr;      // [kRedParts * 5 + 8] partials and results of the particle reductions
                                     // (CFL speed, FP64 conservation sums), for the context's lifetime
    // sort: keys holds the sorted keys of the last sort (the tables' keys),
    // keys_next the keys of the particles in their current slots (the G2P
    // writes them), keys_alt receives the next sort, after which keys and
    // keys_alt trade places
    uint32_t* keys = nullptr;
    uint32_t* keys_alt = nullptr;
    uint32_t* keys_next = nullptr;
    uint32_t* vals = nullptr;
    uint32_t* vals_alt = nullptr;
    uint32_t* hist = nullptr;
    int sort_tiles = 0;
    int key_bits = 0;
    bool sorted = false;     // particles gathered to the order of `keys`
    // mover sort (sort_movers): per-tile mover counts and their offsets, the
    // movers' new keys and slots (two buffers each for the radix passes),
    // their previous keys and slots in slot order
    int* mv_tcount = nullptr;
    int* mv_toff = nullptr;
    uint32_t* mv_key[2] = {nullptr, nullptr};
    uint32_t* mv_val[2] = {nullptr, nullptr};
    uint32_t* mv_old = nullptr;
    uint32_t* mv_slot = nullptr;
    bool sort_by_movers = true;   // the path the next sort takes (mpmx_config.sort_variant 0: from
                                  // the last published mover count)
    bool old_keys_sorted = false; // keys holds a sort's output (the mover path's precondition)
    bool tables_match_keys = false;   // the page tables were built from keys (the mover path uses them)
    int* mv_tbound = nullptr;     // [sort_tiles + 1] sorted movers before each tile's first slot
    int* mv_cpre = nullptr;       // [n / 64 + 1] movers before each 64-slot chunk
    uint64_t* mv_cmask = nullptr; // [n / 64 + 1] movers of each chunk
    // keys of the particles in their current slots
    uint32_t* slot_keys() const { return sorted ? keys : keys_next; }
    // pages
    int axis_bits = 0;       // bits per axis of the Morton page space
    int64_t mask_words = 0;
    uint64_t* pmask = nullptr;
    uint64_t* nmask = nullptr;
    uint64_t* pflag = nullptr;   // one byte per page, set by the page's first particle; zero between builds
    int* wcount = nullptr;   // per-word popcounts, then their exclusive scan
    int* scount = nullptr;
    uint32_t* page_key = nullptr;
    int* page2particle = nullptr;
    int* adj = nullptr;
    int* nvp = nullptr;
    int* vpo = nullptr;
    int* target = nullptr;
    int64_t cap_p = 0, cap_tp = 0, cap_vp = 0;
    int64_t tp_bound = 0;    // total pages the domain can hold
    uint32_t* block_key = nullptr;   // [tp_bound] occupied 2x2x2-page blocks, Morton order
    int* cell_start = nullptr;       // [cap_p][65] first particle of each cell of each particle page
    void* block_lists = nullptr;     // [tp_bound] per-block source lists of the block P2G (BlockClasses)
    void* block_g2p = nullptr;       // [tp_bound] per-block G2P pages and particle range (BlockG2P)
    int* block_order = nullptr;      // [tp_bound] blocks of each XCD share, heavy ones first
    uint8_t* block_wclass = nullptr; // [tp_bound] work class of each block (its gather entries)
    // set by sort_keys: the sorted order (new -> old) whose state gather is
    // still pending; the fused stress pass or finish_gather performs it
    const uint32_t* gather_perm = nullptr;
    // set by substep_dt and the substep capture around their fused stress
    // pass: the G2P follows in the same substep, so the pass's gather leaves
    // out v and B (the G2P writes both without reading them); vb_perm then
    // holds the permutation until the G2P has run (finish_carry_vb moves
    // them if it does not)
    bool g2p_follows = false;
    const uint32_t* vb_perm = nullptr;
    void* p2g_part = nullptr;        // [cap_part] per-block records of the owner-block P2G (p2g_variant 3)
    int64_t cap_part = 0;
    bool vpages_valid = false;       // adjacency and virtual pages match the current tables
    // launch shapes (mpmx_tuning; mpmx_autotune picks them): none changes results
    int particle_wg = 256;           // threads per workgroup of the per-particle kernels
    int p2g_pct = 100, g2p_pct = 100, cg_pct = 100;   // block-kernel workgroups, % of the block count
    // captured substeps (mpmx_config.use_graph), one pair per parity of the
    // double-buffered particle state: A = sort .. PCG (k_graph iterations) +
    // the counter copy, B = G2P
    struct SubstepGraph {
        hipGraphExec_t a = nullptr, b = nullptr;
        bool valid = false;
        float dt = 0.f;
        int iters = 0;
        int64_t blocks = 0;
        bool by_movers = false;   // the sort path captured
    } graphs[2];
    int cg_recent[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // PCG iterations of the last substeps
    int* bcount = nullptr;           // [mask_words] per-word block counts, then their scan
    // grid + PCG
    float* grid = nullptr;
    float* grid_p2g = nullptr;   // copy of the P2G result (diagnostics only)
    bool keep_p2g = false;
    float* cg_x = nullptr;
    float* cg_r = nullptr;
    float* cg_z = nullptr;
    float* cg_s = nullptr;
    float* cg_q = nullptr;
    float* cg_part = nullptr;    // [4][cg_nb] per-workgroup partials
    uint64_t* cg_amask = nullptr;   // [cap_tp] active-node bits per grid page (block P2G)
    float* cg_col = nullptr;     // [3n] per-particle M V dx^2 grad s (block K s)
    float* payload = nullptr;    // [28n] per-particle P2G terms (block P2G)
    unsigned* pbound = nullptr;  // [cap_p][8] per-page largest contribution per channel
    unsigned* cbound = nullptr;  // [cap_p] per-page largest |col|_1 of the K s column (float bits)
    int cg_nb = 0;               // workgroups of every node-space kernel
    int cg_spec = 1;             // PCG iterations queued before the first host check
    bool grid_fused = false;     // last P2G also ran the grid update (gather variant)
    bool order_ok = false;       // block_order holds this substep's order (fused stress pass)
    // counters
    mpmx::DevCounts* dc = nullptr;
    mpmx::DevCounts* hc = nullptr;   // pinned host mirror
    mpmx::DevCounts* pub = nullptr;  // coherent pinned copy the device publishes to (counts_to_host)
    unsigned pub_seq = 0;
    // scans
    int* scan_tmp = nullptr;
    int64_t scan_tmp_cap = 0;
    // misc
    int num_cu = 256;
    int substeps = 0;
    bool timing = false;
    std::vector<mpmx::TimerEntry> timers;
    std::string last_error;
};

namespace mpmx {

// ---- device helpers -------------------------------------------------------
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t compact3(uint32_t v) {
    v &= 0x09249249u;
    v = (v ^ (v >> 2)) & 0x030C30C3u;
    v = (v ^ (v >> 4)) & 0x0300F00Fu;
    v = (v ^ (v >> 8)) & 0x030000FFu;
    v = (v ^ (v >> 16)) & 0x000003FFu;
    return v;
}
// Morton page key with the SPGrid bit order: z at bit 0, y at bit 1, x at bit 2
// of every triple (SPGridMask.cpp:96-100)
__device__ __forceinline__ uint32_t morton(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(z) | (spread3(y) << 1) | (spread3(x) << 2);
}
__device__ __forceinline__ void unmorton(uint32_t k, int& x, int& y, int& z) {
    z = (int)compact3(k);
    y = (int)compact3(k >> 1);
    x = (int)compact3(k >> 2);
}

// id of page key k: rank among particle pages, or num_page + rank among the
// neighbour-only pages, or -1 (popcount rank in the page bitmaps)
__device__ __forceinline__ int page_rank(uint32_t k, const uint64_t* __restrict__ pmask,
                                         const uint64_t* __restrict__ nmask, const int* __restrict__ wpre,
                                         const int* __restrict__ spre, int num_page) {
    const uint32_t w = k >> 6, b = k & 63u;
    const uint64_t below = b ? ((~0ull) >> (64 - b)) : 0ull;
    const uint64_t P = pmask[w];
    if ((P >> b) & 1ull) return wpre[w] + __popcll(P & below);
    const uint64_t S = nmask[w] & ~P;
    if ((S >> b) & 1ull) return num_page + spre[w] + __popcll(S & below);
    return -1;
}

// base node floor(x*N + 1/2) - 1 evaluated exactly (the reference's
// (int)(x * one_over_dx + 0.5f) - 1, Particle_kernal.cu:162), clamped to the
// valid range so that a particle that left the domain still yields in-range
// indices (the error itself is latched by the caller)
__device__ __forceinline__ int base_node(float x, int N) {
    double t = (double)x * (double)N + 0.5;
    if (!(t == t)) t = 0.0;
    int b = (int)floor(t) - 1;
    b = b < 0 ? 0 : b;
    b = b > N - 3 ? N - 3 : b;
    return b;
}
__device__ __forceinline__ uint32_t key_of(int bi, int bj, int bk) {
    return (morton((uint32_t)bi >> 2, (uint32_t)bj >> 2, (uint32_t)bk >> 2) << 6) |
           (uint32_t)(((bi & 3) << 4) | ((bj & 3) << 2) | (bk & 3));
}

// quadratic B-spline weights of the 3 nodes base+0..2, d0 = x/dx - base in [0.5,1.5)
// (P2GKernal.cu:74-82), and their derivatives in units of 1/dx
__device__ __forceinline__ void bspline(float d0, float (&w)[3]) {
    float z = 1.5f - d0;
    w[0] = 0.5f * z * z;
    const float d1 = d0 - 1.0f;
    w[1] = 0.75f - d1 * d1;
    z = 1.5f - (1.0f - d1);
    w[2] = 0.5f * z * z;
}
__device__ __forceinline__ void bspline_grad(float d0, float (&w)[3], float (&g)[3]) {
    float z = 1.5f - d0;
    w[0] = 0.5f * z * z;
    g[0] = -z;
    const float d1 = d0 - 1.0f;
    w[1] = 0.75f - d1 * d1;
    g[1] = -d1 * 2.0f;
    z = 1.5f - (1.0f - d1);
    w[2] = 0.5f * z * z;
    g[2] = z;
}

// order-preserving float <-> uint keys for atomicMax over signed floats
__device__ __forceinline__ unsigned fkey(float f) {
    const unsigned b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ __forceinline__ float fkey_dec(unsigned k) {
    const unsigned b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float(b);
#else
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
#endif
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// max (MAX) or sum of parts[0..cnt) for every thread of the workgroup, in a
// fixed order (thread-strided, wave butterfly, waves in order); lds holds one
// float per wave
template <bool MAX>
__device__ __forceinline__ float reduce_parts(const float* __restrict__ parts, int cnt, float* lds) {
    float v = MAX ? -INFINITY : 0.f;
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) v = MAX ? fmaxf(v, parts[k]) : v + parts[k];
    v = MAX ? wave_max(v) : wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = lds[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = MAX ? fmaxf(r, lds[w]) : r + lds[w];
    return r;
}
// combine one value per wave into thread 0 of the workgroup, waves in order
template <bool MAX>
__device__ __forceinline__ float combine_waves(float v, float* lds) {
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = lds[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = MAX ? fmaxf(r, lds[w]) : r + lds[w];
    return r;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Items of `count` for the calling workgroup, `per_wg` consecutive items per
// workgroup step (1, or one per wave): XCD x = blockIdx % 8 owns the
// contiguous share [floor(x count / 8), floor((x + 1) count / 8)) of the
// items, so that every XCD gets an equal share of the items that exist (the
// grids are sized with slack for growth, and the workgroups beyond a share
// exit at once); its workgroups j = blockIdx / 8 take the share's items
// j per_wg, + per_wg G / 8, ...  A grid that is not a multiple of 8 strides
// over all items instead.
// first item of XCD share x of count items: shares differ by at most one item
__device__ __forceinline__ int xcd_share_lo(int x, int count) { return (int)(((int64_t)x * count) >> 3); }
__device__ __forceinline__ int xcd_item_first(int count, int per_wg) {
    const int G = gridDim.x, b = blockIdx.x;
    if (G & 7) return b * per_wg;
    return xcd_share_lo(b & 7, count) + (b >> 3) * per_wg;
}
__device__ __forceinline__ int xcd_item_end(int count) {
    const int G = gridDim.x, b = blockIdx.x;
    if (G & 7) return count;
    return xcd_share_lo((b & 7) + 1, count);
}
__device__ __forceinline__ int xcd_item_step(int per_wg) {
    const int G = gridDim.x;
    return ((G & 7) ? G : (G >> 3)) * per_wg;
}
#define FOR_LOGICAL_ITEMS(var, count)                                                             \
    for (int var = xcd_item_first((count), 1), var##_end = xcd_item_end(count); var < var##_end; \
         var += xcd_item_step(1))

// contiguous [begin, end) chunk of `count` items for the calling workgroup:
// XCD x = blockIdx % 8 owns the share [x s, (x + 1) s), s = ceil(count / 8),
// and its workgroups j = blockIdx / 8 split it into equal chunks (a grid that
// is not a multiple of 8 splits all items by blockIdx)
__device__ __forceinline__ void block_range(int count, int& begin, int& end) {
    const int G = gridDim.x, b = blockIdx.x;
    int lo = 0, hi = count, g = G, j = b;
    if (!(G & 7)) {
        const int s = (count + 7) >> 3;
        lo = (b & 7) * s;
        hi = lo + s < count ? lo + s : count;
        hi = hi > lo ? hi : lo;
        g = G >> 3;
        j = b >> 3;
    }
    const int per = (hi - lo + g - 1) / g;
    begin = lo + j * per;
    end = begin + per < hi ? begin + per : hi;
}

// scene of page key pk
__device__ __forceinline__ int scene_of_page(const DevParams& dp, uint32_t pk) { return (int)(pk >> dp.slot_shift); }

// the calling lane's particle i and the wave's scene, for workgroup wg of the
// particle grid; false for idle lanes
__device__ __forceinline__ bool wave_particle(const WaveMap& m, int n, int wg, int& i, int& scene) {
    const int64_t w = ((int64_t)wg * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (!m.first) {
        i = (int)(w * 64 + lane);
        scene = 0;
        return i < n;
    }
    if (w >= m.waves) {
        i = 0;
        scene = 0;
        return false;
    }
    i = m.first[w] + lane;
    scene = __builtin_amdgcn_readfirstlane(m.scene[w]);
    return lane < m.count[w];
}
__device__ __forceinline__ bool wave_particle(const WaveMap& m, int n, int& i, int& scene) {
    return wave_particle(m, n, (int)blockIdx.x, i, scene);
}

// The 128-node blocks of the PCG's "dot" products, per scene: block b of a
// scene is its scene-local pages 2b and 2b+1 (as the scene alone numbers its
// pages); lane l takes nodes 2 (l & 31) and 2 (l & 31) + 1 of page
// 2b + (l >> 5).  Workgroup blockIdx.x serves scene wg_scene[blockIdx.x]; the
// waves of that scene's workgroups stride over its blocks, and each workgroup
// writes its partial at blockIdx.x, so a scene's partials are [wg0, wg0 + nwg).
struct SceneBlocks {
    int scene, wg0, nwg, pfirst, np, nfirst, nn, nblk, wave, nwaves;
};
__device__ __forceinline__ SceneBlocks scene_blocks(const SceneCG* __restrict__ scg,
                                                    const int* __restrict__ wg_scene) {
    SceneBlocks b;
    b.scene = wg_scene[blockIdx.x];
    const SceneCG& g = scg[b.scene];
    b.wg0 = g.wg0;
    b.nwg = g.nwg;
    b.pfirst = g.pfirst;
    b.np = g.np;
    b.nfirst = g.nfirst;
    b.nn = g.nn;
    b.nblk = (g.np + g.nn + 1) / 2;
    const int wpb = (int)(blockDim.x >> 6);
    b.wave = (int)(blockIdx.x - b.wg0) * wpb + (int)(threadIdx.x >> 6);
    b.nwaves = b.nwg * wpb;
    return b;
}
// first of the lane's two nodes in block blk of the scene, or -1
__device__ __forceinline__ int64_t scene_block_node(const SceneBlocks& b, int num_page, int64_t cap_tp, int blk) {
    const int lane = threadIdx.x & 63;
    const int l = 2 * blk + (lane >> 5);
    if (l >= b.np + b.nn) return -1;
    const int pg = l < b.np ? b.pfirst + l : num_page + b.nfirst + (l - b.np);
    if (pg < 0 || pg >= cap_tp) return -1;
    return (int64_t)pg * 64 + 2 * (lane & 31);
}

}  // namespace mpmx

// ---- host-side launch helpers (defined in the .hip files) -------------------
namespace mpmx {
// up to three exclusive scans of n entries at once; total[y] (may be null)
// receives the sum of array y
struct ScanSet3 {
    const int* in[3];
    int* out[3];
    int* total[3];
};
hipError_t scan_exclusive3(mpmx_ctx* c, const ScanSet3& a, int k, int64_t n);
hipError_t scan_exclusive(mpmx_ctx* c, const int* in, int* out, int64_t n_cap, const int* n_dev,
                          int64_t n_host, int* total_dev);
hipError_t sort_and_gather(mpmx_ctx* c);
hipError_t sort_keys(mpmx_ctx* c);
void choose_sort_path(mpmx_ctx* c);
hipError_t finish_gather(mpmx_ctx* c);
hipError_t finish_carry_vb(mpmx_ctx* c);
hipError_t compute_keys(mpmx_ctx* c);
hipError_t build_tables(mpmx_ctx* c);
hipError_t build_vpage_tables(mpmx_ctx* c);
hipError_t launch_scene_counts(mpmx_ctx* c);
int particle_blocks(const mpmx_ctx* c);   // 256-thread workgroups of the per-particle kernels
hipError_t launch_stress(mpmx_ctx* c);
hipError_t launch_p2g(mpmx_ctx* c, float dt);
hipError_t launch_stress_prep(mpmx_ctx* c, float dt);
hipError_t launch_stress_p2g(mpmx_ctx* c, float dt);
hipError_t launch_block_lists(mpmx_ctx* c);
// the device counters into c->hc once the work queued so far has finished:
// a one-wave kernel copies them into coherent host memory, the host waits
// for its sequence number (no stream synchronisation)
hipError_t counts_to_host(mpmx_ctx* c);
hipError_t launch_grid_update(mpmx_ctx* c, float dt);
hipError_t launch_pcg(mpmx_ctx* c, float dt);
hipError_t launch_pcg_from(mpmx_ctx* c, int it0, int first_chunk);
hipError_t queue_pcg_iterations(mpmx_ctx* c, int k0, int k1);
hipError_t launch_g2p(mpmx_ctx* c, float dt);
hipError_t launch_cg_reset(mpmx_ctx* c);
hipError_t fill_zero(mpmx_ctx* c, void* p, int64_t bytes);
hipError_t fill_zero_dev_count(mpmx_ctx* c, float* p, const int* count_dev, int64_t per_unit,
                               int64_t cap_units);
void timer_push(mpmx_ctx* c, const char* name, bool begin);
void destroy_graphs(mpmx_ctx* c);
int transfer_grid_blocks(mpmx_ctx* c);
int block_grid(mpmx_ctx* c, int pct);
}  // namespace mpmx

#define MPMX_TIMED(ctx, name, stmt)          \
    do {                                     \
        mpmx::timer_push(ctx, name, true);   \
        stmt;                                \
        mpmx::timer_push(ctx, name, false);  \
    } while (0)
