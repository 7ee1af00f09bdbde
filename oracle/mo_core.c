/*
 * mo_core.c — oracle (test infrastructure): SPGrid masks, cell keys, the
 * canonical particle sort and the sparse page tables, restated serially.
 *
 *   SPGridMask.cpp:4-71, 73-76, 78-129   bit masks, Bit_Spread, Linear_Offset
 *   Particle_kernal.cu:31-45             Bit_Spread_Mine (device pdep)
 *   TimeIntegrator.cu:26-45              Bit_Pack_Mine (device pext)
 *   Transformer_kernal.cu:6-13           Packed_Add (masked Morton addition)
 *   Particle_kernal.cu:158-165           cal_cellId (key of the quadratic base node)
 *   Particle_kernal.cu:484-542           sort_by_offsets (bucket sort by page/cell)
 *   Transformer_kernal.cu:15-126,129-209 page/cell boundaries, hash of pages,
 *                                        7 neighbour pages, topology, virtual pages
 *   Transformer.cpp:27-34                neighbour offsets (i*4, j*4, k*4) order
 */
#include "mpmx_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

void mo_default_params(mo_params* p) {
    memset(p, 0, sizeof(*p));
    p->N = 128;
    p->transfer = 0;
    p->phase_field = 1;
    p->dot_mode = 0;
    p->cg_max_iter = 100;
    p->vpage = 256;
    p->wall_cells = 4;
    p->grip_lo = -1;
    p->grip_hi = 1 << 30;
    p->grip_v = 0.0;
    p->E = 1e5;
    p->nu = 0.3;
    p->V0 = 1.0 / (128.0 * 128.0 * 128.0 * 8.0);
    p->yield_stress = 800.0;
    p->hardening = 3.0;
    p->alpha_c = 1.0;
    p->dlambda_cap = 100.0;
    p->eps_cap = 3.0;
    p->Gc0 = 3.0;
    p->residual_c = 0.1;
    p->ps = 0.01;
    p->theta = (double)0.0001f;
    p->M = 50.0;
    p->gravity[0] = 0.0; p->gravity[1] = -10.0; p->gravity[2] = 0.0;
    p->vmax = (double)11.137f;
    p->cg_tol = (double)1e-6f;
    p->crit_stretch = 0.0;
    p->model = 2;
}

/* SPGridMask.cpp:78-88, 96-100, 119-124 with struct_size = 6, element_size = 2
 * (SPGridMask.h:4-7): 4x4x4 cells per page, in-page bits 2..7, page bits 12.. */
static void masks_compute(uint64_t m[3]) {
    const unsigned block_bits = 6, element_size = 2, struct_size = 6;
    const unsigned zb = block_bits / 3 + (block_bits % 3 > 0);
    const unsigned yb = block_bits / 3 + (block_bits % 3 > 1);
    const unsigned xb = block_bits / 3;
    const uint64_t ez = ((1ull << zb) - 1) << element_size;
    const uint64_t ey = ((1ull << yb) - 1) << (element_size + zb);
    const uint64_t ex = ((1ull << xb) - 1) << (element_size + zb + yb);
    const uint64_t pz = (0x9249249249249249ull << (3 - struct_size % 3)) & 0xfffffffffffff000ull;
    const uint64_t py = (0x2492492492492492ull << (3 - struct_size % 3)) & 0xfffffffffffff000ull;
    const uint64_t px = (0x4924924924924924ull << (3 - struct_size % 3)) & 0xfffffffffffff000ull;
    m[0] = px | ex;
    m[1] = py | ey;
    m[2] = pz | ez;
}

void mo_masks(uint64_t m[3]) {
    static uint64_t cached[3];
    static int have = 0;
    if (!have) { masks_compute(cached); have = 1; }
    m[0] = cached[0]; m[1] = cached[1]; m[2] = cached[2];
}

static uint64_t brev64(uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 64; ++b) { r = (r << 1) | (v & 1ull); v >>= 1; }
    return r;
}
static int clz64(uint64_t v) {
    if (!v) return 64;
    int n = 0;
    while (!(v & 0x8000000000000000ull)) { v <<= 1; ++n; }
    return n;
}

/* Particle_kernal.cu:31-45 */
uint64_t mo_bit_spread(uint64_t mask, int data) {
    uint64_t rmask = brev64(mask);
    int dat = data;
    uint64_t result = 0;
    unsigned count = 0;
    while (rmask && dat) {
        const unsigned lz = (unsigned)clz64(rmask) + 1;
        result = (lz >= 64 ? 0 : (result << lz)) | (uint64_t)(dat & 1);
        dat >>= 1;
        rmask = lz >= 64 ? 0 : (rmask << lz);
        count += lz;
    }
    if (count == 0) return 0;
    return brev64(result) >> (64 - count);
}

/* TimeIntegrator.cu:26-45 */
int mo_bit_pack(uint64_t mask, uint64_t data) {
    uint64_t uldata = data, ulresult = 0, rmask = brev64(mask);
    int count = 0;
    while (rmask && uldata) {
        const int lz = clz64(rmask);
        uldata >>= lz;
        ulresult <<= 1;
        ulresult |= (uldata & 1ull);
        ++count;
        uldata >>= 1;
        rmask = (lz + 1 >= 64) ? 0 : (rmask << (lz + 1));
    }
    if (count == 0) return 0;
    return (int)(brev64(ulresult) >> (64 - count));
}

/* SPGridMask.cpp:73-76 */
uint64_t mo_linear_offset(int i, int j, int k) {
    uint64_t m[3];
    mo_masks(m);
    return mo_bit_spread(m[0], i) | mo_bit_spread(m[1], j) | mo_bit_spread(m[2], k);
}

/* Transformer_kernal.cu:6-13 */
uint64_t mo_packed_add(uint64_t i, uint64_t j) {
    uint64_t m[3];
    mo_masks(m);
    const uint64_t all = m[0] | m[1] | m[2];
    const uint64_t x = ((i | ~m[0]) + (j & m[0])) & m[0];
    const uint64_t y = ((i | ~m[1]) + (j & m[1])) & m[1];
    const uint64_t z = ((i | ~m[2]) + (j & m[2])) & m[2];
    const uint64_t w = ((i | all) + (j & ~all)) & ~all;
    return x | y | z | w;
}

/* Particle_kernal.cu:162: (int)(x * one_over_dx + 0.5f) - 1, evaluated exactly
 * (for FP64 inputs of the size used here the product and sum are exact). */
int mo_base_node(double x, int N) {
    return (int)(x * (double)N + 0.5) - 1;
}

void mo_cell_keys_f64(int n, const double* x, int N, uint64_t* keys) {
    for (int p = 0; p < n; ++p)
        keys[p] = mo_linear_offset(mo_base_node(x[3 * p + 0], N), mo_base_node(x[3 * p + 1], N),
                                   mo_base_node(x[3 * p + 2], N));
}
void mo_cell_keys_f32(int n, const float* x, int N, uint64_t* keys) {
    for (int p = 0; p < n; ++p)
        keys[p] = mo_linear_offset(mo_base_node((double)x[3 * p + 0], N),
                                   mo_base_node((double)x[3 * p + 1], N),
                                   mo_base_node((double)x[3 * p + 2], N));
}

typedef struct { uint64_t key; int idx; } kv;
static int kv_cmp(const void* a, const void* b) {
    const kv* x = (const kv*)a;
    const kv* y = (const kv*)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

/* Particle_kernal.cu:484-542 buckets particles by (page, cell); here the
 * order inside one cell is the previous order (stable). */
void mo_sort_keys(int n, const uint64_t* keys, int* perm, uint64_t* sorted) {
    kv* a = (kv*)malloc(sizeof(kv) * (size_t)(n > 0 ? n : 1));
    for (int p = 0; p < n; ++p) { a[p].key = keys[p]; a[p].idx = p; }
    qsort(a, (size_t)n, sizeof(kv), kv_cmp);
    for (int p = 0; p < n; ++p) { perm[p] = a[p].idx; sorted[p] = a[p].key; }
    free(a);
}

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}
/* index of key in ascending array a[0..n), or -1 */
static int find_u64(const uint64_t* a, int n, uint64_t key) {
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] == key) return mid;
        if (a[mid] < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* Transformer.cpp:27-34: offsets of pages (i,j,k) in {0,1}^3 \ {0}, order i*4+j*2+k */
static void neighbor_offsets(uint64_t nb[7]) {
    int q = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                if (i == 0 && j == 0 && k == 0) continue;
                nb[q++] = mo_linear_offset(i * 4, j * 4, k * 4);
            }
}

int mo_page_tables(int n, const uint64_t* sorted, int vpage, int cap_p, int cap_tp, int cap_vp,
                   uint64_t* page_key, int* page2particle, int* adj, int* vpo, int* target,
                   int* particle2cell, int* cell2particle, int counts[4]) {
    /* markCellBoundary + inclusive scan + markBlockOffset (Transformer_kernal.cu:15-36, 129-146) */
    int numCell = 0;
    for (int p = 0; p < n; ++p) {
        if (p == 0 || sorted[p] != sorted[p - 1]) {
            if (cell2particle) cell2particle[numCell] = p;
            ++numCell;
        }
        if (particle2cell) particle2cell[p] = numCell;
    }
    if (cell2particle) cell2particle[numCell] = n;

    /* markPageBoundary + scan + markBlockOffset (Transformer_kernal.cu:26-47, 148-161) */
    int numPage = 0;
    for (int p = 0; p < n; ++p) {
        if (p == 0 || (sorted[p] >> 12) != (sorted[p - 1] >> 12)) {
            if (numPage >= cap_p || numPage >= cap_tp) return -1;
            page2particle[numPage] = p;
            page_key[numPage] = sorted[p] >> 12;
            ++numPage;
        }
    }
    page2particle[numPage] = n;

    /* supplementAdjacentPages (Transformer_kernal.cu:70-96): neighbour pages not
     * holding particles, numbered after the particle pages in ascending key order */
    uint64_t nb[7];
    neighbor_offsets(nb);
    uint64_t* cand = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(7 * numPage + 1));
    int nc = 0;
    for (int p = 0; p < numPage; ++p) {
        const uint64_t okey = page_key[p] << 12;
        for (int q = 0; q < 7; ++q) {
            const uint64_t key = mo_packed_add(okey, nb[q]) >> 12;
            if (find_u64(page_key, numPage, key) < 0) cand[nc++] = key;
        }
    }
    qsort(cand, (size_t)nc, sizeof(uint64_t), cmp_u64);
    int numTotal = numPage;
    for (int c = 0; c < nc; ++c) {
        if (c > 0 && cand[c] == cand[c - 1]) continue;
        if (numTotal >= cap_tp) { free(cand); return -1; }
        page_key[numTotal++] = cand[c];
    }
    free(cand);

    /* establishPageTopology (Transformer_kernal.cu:98-114) */
    const int nsupp = numTotal - numPage;
    for (int p = 0; p < numPage; ++p) {
        const uint64_t okey = page_key[p] << 12;
        for (int q = 0; q < 7; ++q) {
            const uint64_t key = mo_packed_add(okey, nb[q]) >> 12;
            int id = find_u64(page_key, numPage, key);
            if (id < 0) id = numPage + find_u64(page_key + numPage, nsupp, key);
            adj[q * numPage + p] = id;
        }
    }

    /* markPageSize + exclusive scan + markVirtualPageOffset + inclusive scan
     * (Transformer_kernal.cu:116-126, 184-209) */
    int nv = 0;
    for (int p = 0; p < numPage; ++p) {
        vpo[p] = nv;
        const int cnt = page2particle[p + 1] - page2particle[p];
        nv += (cnt + vpage - 1) / vpage;
    }
    vpo[numPage] = nv;
    if (nv > cap_vp) return -1;
    for (int p = 0; p < numPage; ++p)
        for (int v = vpo[p]; v < vpo[p + 1]; ++v) target[v] = p + 1;

    counts[0] = numPage;
    counts[1] = numTotal;
    counts[2] = nv;
    counts[3] = numCell;
    return 0;
}
