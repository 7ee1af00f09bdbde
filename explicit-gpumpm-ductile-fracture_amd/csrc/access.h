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
