// payload.h — per-particle P2G terms, the payload records the block P2G reads
// and the per-page channel bounds that fix its fixed-point scales.  Shared by
// the payload pass of the P2G (k_grid.hip) and the stress pass that writes the
// same records while the particle's stress is still in registers
// (k_stress.hip).
//
// Terms of P2G_MLS / P2G_APIC (P2GKernal.cu:6-155, 512-702): MLS folds the
// stress into the affine term, Q = (B m - dt sigma) D^-1 dx; APIC keeps the
// stress term dt sigma / dx for the weight gradients.  The phase-field
// particle passes (P2GKernal.cu:897-1157) add V and V FP.
#pragma once
#include "access.h"

namespace mpmx {

constexpr int kP2GCh = 8;   // m, p0, p1, p2, sum w c, sum w, sum V w, sum V FP w

// per-particle P2G terms, computed once and kept in registers across rounds
struct P2GTerms {
    float w[3][3], gw[3][3];
    int l0[3];
    float m, cphase, V, VF;
    float ar[3], Qd[9], sg[9];
};

// terms of particle i relative to the block whose first node is org, from the
// state, its Kirchhoff stress sig (component r + 3 c), FP and volume V; d0 is
// its offset from its base node in cells
template <int TRANSFER, typename ST>
__device__ __forceinline__ void p2g_terms(const DevParams& dp, const ST& s, const float (&sig)[9], float FPv,
                                          float V, int i, int n, float dt, const int (&org)[3], float (&d0)[3],
                                          P2GTerms& T) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float xa = ld_x(s, i, a);
        const int b = base_node(xa, dp.N);
        d0[a] = xa * dp.inv_dx - (float)b;
        T.l0[a] = b - org[a];
        if (TRANSFER == 0) bspline(d0[a], T.w[a]);
        else bspline_grad(d0[a], T.w[a], T.gw[a]);
    }
    T.m = ld_m(s, i);
    T.cphase = ld_c(s, i);
    T.V = V;
    T.VF = T.V * FPv;
    float v[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = ld_v(s, i, a);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float Bk = ld_B(s, i, k, n);
        const float sk = sig[k];
        if (TRANSFER == 0) T.Qd[k] = (Bk * T.m - sk * dt) * dp.Dinv * dp.dx;
        else { T.Qd[k] = Bk * dp.Dinv * dp.dx * T.m; T.sg[k] = sk * dt * dp.inv_dx; }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) T.ar[r] = T.m * v[r] - (T.Qd[r] * d0[0] + T.Qd[r + 3] * d0[1] + T.Qd[r + 6] * d0[2]);
}

// largest |contribution| of the particle to any node, per channel (weights <= 1,
// gradients of the quadratic B-spline <= 1 in magnitude per axis)
template <int TRANSFER>
__device__ __forceinline__ void p2g_bounds(const P2GTerms& T, float (&b)[kP2GCh]) {
    b[0] = fabsf(T.m);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float q = fabsf(T.ar[r]) + 2.f * (fabsf(T.Qd[r]) + fabsf(T.Qd[r + 3]) + fabsf(T.Qd[r + 6]));
        if (TRANSFER == 1) q += fabsf(T.sg[r]) + fabsf(T.sg[r + 3]) + fabsf(T.sg[r + 6]);
        b[1 + r] = q;
    }
    b[4] = fabsf(T.cphase);
    b[5] = 1.f;
    b[6] = fabsf(T.V);
    b[7] = fabsf(T.VF);
}

// ---- per-particle P2G payload ------------------------------------------------
// The terms every node contribution of a particle needs, computed once per
// substep in sorted order and stored as float4 records: d0 (the particle's
// offset from its base node in cells), m, ar, Qd, c, V, V FP and, for APIC,
// the stress term sg (MLS: 5 records of 4 floats; APIC: 7, sg[8] in the pad
// slot of the fifth).  The block P2G reads one record per visit (a particle
// is visited by the one to eight blocks it reaches) instead of recomputing
// them from 26 scattered fields; the base node itself comes from the key.
template <int TRANSFER> struct Payload { static constexpr int kFloats = TRANSFER ? 28 : 20; };

template <int TRANSFER>
__device__ __forceinline__ void payload_store(float* __restrict__ pay, int64_t i, const float (&d0)[3],
                                              const P2GTerms& T) {
    float4* q = reinterpret_cast<float4*>(pay + i * Payload<TRANSFER>::kFloats);
    q[0] = make_float4(d0[0], d0[1], d0[2], T.m);
    q[1] = make_float4(T.ar[0], T.ar[1], T.ar[2], T.cphase);
    q[2] = make_float4(T.Qd[0], T.Qd[1], T.Qd[2], T.Qd[3]);
    q[3] = make_float4(T.Qd[4], T.Qd[5], T.Qd[6], T.Qd[7]);
    q[4] = make_float4(T.Qd[8], T.V, T.VF, TRANSFER == 1 ? T.sg[8] : 0.f);
    if (TRANSFER == 1) {
        q[5] = make_float4(T.sg[0], T.sg[1], T.sg[2], T.sg[3]);
        q[6] = make_float4(T.sg[4], T.sg[5], T.sg[6], T.sg[7]);
    }
}

// the terms of particle i (key k) relative to the block whose first node is org
template <int TRANSFER>
__device__ __forceinline__ void payload_load(const float* __restrict__ pay, int64_t i, uint32_t key,
                                             const int (&org)[3], P2GTerms& T) {
    const float4* q = reinterpret_cast<const float4*>(pay + i * Payload<TRANSFER>::kFloats);
    const float4 a = q[0], b = q[1], c0 = q[2], c1 = q[3], c2 = q[4];
    const float d0[3] = {a.x, a.y, a.z};
    T.m = a.w;
    T.ar[0] = b.x; T.ar[1] = b.y; T.ar[2] = b.z;
    T.cphase = b.w;
    T.Qd[0] = c0.x; T.Qd[1] = c0.y; T.Qd[2] = c0.z; T.Qd[3] = c0.w;
    T.Qd[4] = c1.x; T.Qd[5] = c1.y; T.Qd[6] = c1.z; T.Qd[7] = c1.w;
    T.Qd[8] = c2.x; T.V = c2.y; T.VF = c2.z;
    if (TRANSFER == 1) {
        const float4 s0 = q[5], s1 = q[6];
        T.sg[0] = s0.x; T.sg[1] = s0.y; T.sg[2] = s0.z; T.sg[3] = s0.w;
        T.sg[4] = s1.x; T.sg[5] = s1.y; T.sg[6] = s1.z; T.sg[7] = s1.w;
        T.sg[8] = c2.w;
    }
    int px, py, pz;
    unmorton(key >> 6, px, py, pz);
    const int cell = (int)(key & 63u);
    T.l0[0] = 4 * px + (cell >> 4) - org[0];
    T.l0[1] = 4 * py + ((cell >> 2) & 3) - org[1];
    T.l0[2] = 4 * pz + (cell & 3) - org[2];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        if (TRANSFER == 0) bspline(d0[ax], T.w[ax]);
        else bspline_grad(d0[ax], T.w[ax], T.gw[ax]);
    }
}

// payload record i of the particle whose state sits at o of s, and its
// per-channel bounds b
template <int TRANSFER, typename ST>
__device__ __forceinline__ void prep_particle(const DevParams& dp, const ST& s, const float (&sig)[9], float FPv,
                                              float V, int o, int i, int n, float dt, float* __restrict__ pay,
                                              float (&b)[kP2GCh]) {
    const int org[3] = {0, 0, 0};
    P2GTerms T;
    float d0[3];
    p2g_terms<TRANSFER>(dp, s, sig, FPv, V, o, n, dt, org, d0, T);
    payload_store<TRANSFER>(pay, i, d0, T);
    p2g_bounds<TRANSFER>(T, b);
}

// per particle page, the largest contribution per channel: a wave-segmented
// max over the page's contiguous particles (page key pk, bounds b; lanes
// without a particle take part with valid false), then one integer atomicMax
// per page and channel on the float bits, which order like the floats for
// non-negative values.  pbound is zero before the first call of a substep.
__device__ __forceinline__ void page_bounds(bool valid, uint32_t pk, const float (&b)[kP2GCh], const BlockTables& t,
                                            unsigned* __restrict__ pbound, const DevCounts* __restrict__ dc) {
    const int lane = threadIdx.x & 63;
    const int np_raw = dc->num_page;
    const int np = np_raw < t.cap_p ? np_raw : (int)t.cap_p;
    unsigned long long todo = __ballot(valid);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const uint32_t lpk = __shfl(pk, leader, 64);
        const bool mine = valid && pk == lpk;
        todo &= ~__ballot(mine);
        float m[kP2GCh];
#pragma unroll
        for (int k = 0; k < kP2GCh; ++k) m[k] = wave_max(mine ? b[k] : 0.f);
        if (lane == leader) {
            const int id = page_rank(lpk, t.pmask, t.nmask, t.wpre, t.spre, np_raw);
            if (id >= 0 && id < np) {
#pragma unroll
                for (int k = 0; k < kP2GCh; ++k) atomicMax(&pbound[(int64_t)id * kP2GCh + k], __float_as_uint(m[k]));
            }
        }
    }
}

}  // namespace mpmx
