"""ctypes binding of the serial FP64 oracle (test infrastructure only).

The oracle is the checker for the product library: only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import
this module.  The product path (``libmpmx_hip.so``) never loads it.

Functions mirror ``oracle/mpmx_oracle.h``; each C function cites the reference
file:line it restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmpmx_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_spgrid.so")


class Params(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("transfer", C.c_int), ("phase_field", C.c_int),
        ("dot_mode", C.c_int), ("cg_max_iter", C.c_int), ("vpage", C.c_int),
        ("wall_cells", C.c_int), ("grip_lo", C.c_int), ("grip_hi", C.c_int),
        ("grip_v", C.c_double), ("E", C.c_double), ("nu", C.c_double), ("V0", C.c_double),
        ("yield_stress", C.c_double), ("hardening", C.c_double), ("alpha_c", C.c_double),
        ("dlambda_cap", C.c_double), ("eps_cap", C.c_double), ("Gc0", C.c_double),
        ("residual_c", C.c_double), ("ps", C.c_double), ("theta", C.c_double),
        ("M", C.c_double), ("gravity", C.c_double * 3), ("vmax", C.c_double),
        ("cg_tol", C.c_double), ("crit_stretch", C.c_double), ("model", C.c_int),
    ]


FIELDS = {"x": (0, 3), "v": (1, 3), "F": (2, 9), "B": (3, 9), "c": (4, 1), "alpha": (5, 1),
          "psi": (6, 1), "g": (7, 1), "mass": (8, 1), "FP": (9, 1), "vol": (10, 1),
          "sigma": (11, 9)}

_lib = None
_ref = None


def build():
    """Compile the oracle (and oracle/_ref when the reference checkout exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
        L.mo_default_params.argtypes = [C.POINTER(Params)]
        L.mo_masks.argtypes = [u64p]
        L.mo_bit_spread.argtypes = [C.c_uint64, C.c_int]
        L.mo_bit_spread.restype = C.c_uint64
        L.mo_bit_pack.argtypes = [C.c_uint64, C.c_uint64]
        L.mo_bit_pack.restype = C.c_int
        L.mo_linear_offset.argtypes = [C.c_int, C.c_int, C.c_int]
        L.mo_linear_offset.restype = C.c_uint64
        L.mo_packed_add.argtypes = [C.c_uint64, C.c_uint64]
        L.mo_packed_add.restype = C.c_uint64
        L.mo_base_node.argtypes = [C.c_double, C.c_int]
        L.mo_base_node.restype = C.c_int
        L.mo_cell_keys_f64.argtypes = [C.c_int, f64p, C.c_int, u64p]
        L.mo_cell_keys_f32.argtypes = [C.c_int, f32p, C.c_int, u64p]
        L.mo_sort_keys.argtypes = [C.c_int, u64p, i32p, u64p]
        L.mo_page_tables.argtypes = [C.c_int, u64p, C.c_int, C.c_int, C.c_int, C.c_int,
                                     u64p, i32p, i32p, i32p, i32p, i32p, i32p, i32p]
        L.mo_page_tables.restype = C.c_int
        L.mo_svd3.argtypes = [f64p, f64p, f64p, f64p]
        L.mo_svd3_last_sweeps.restype = C.c_int
        L.mo_inverse3.argtypes = [f64p, f64p]
        L.mo_vdot.argtypes = [C.c_int, f64p, f64p, C.c_int]
        L.mo_vdot.restype = C.c_double
        L.mo_neo_hookean.argtypes = [C.POINTER(Params), f64p, C.c_double, f64p, f64p, f64p, f64p]
        L.mo_neo_hookean.restype = None
        L.mo_fixed_corotated.argtypes = [C.POINTER(Params), f64p, f64p]
        L.mo_fixed_corotated.restype = None
        L.mo_von_mises.argtypes = [C.POINTER(Params), f64p, C.c_double, f64p, f64p, f64p,
                                   f64p, f64p, f64p]
        L.mo_sim_create.argtypes = [C.POINTER(Params), C.c_int, f64p, f64p, f64p]
        L.mo_sim_create.restype = C.c_void_p
        L.mo_sim_destroy.argtypes = [C.c_void_p]
        L.mo_sim_substep.argtypes = [C.c_void_p, C.c_double]
        L.mo_sim_substep.restype = C.c_int
        L.mo_sim_get.argtypes = [C.c_void_p, C.c_int, f64p]
        L.mo_sim_get.restype = C.c_int
        L.mo_sim_set.argtypes = [C.c_void_p, C.c_int, f64p]
        L.mo_sim_set.restype = C.c_int
        L.mo_sim_get_order.argtypes = [C.c_void_p, i32p, u64p]
        L.mo_sim_counts.argtypes = [C.c_void_p, i32p]
        L.mo_sim_round_f32.argtypes = [C.c_void_p]
        L.mo_sim_round_f32.restype = None
        L.mo_sim_cg_norm.argtypes = [C.c_void_p]
        L.mo_sim_cg_norm.restype = C.c_double
        L.mo_sim_grid.argtypes = [C.c_void_p, C.c_int, C.c_int, f64p]
        L.mo_sim_grid.restype = C.c_int
        L.mo_sim_pages.argtypes = [C.c_void_p, u64p]
        L.mo_sim_pages.restype = C.c_int
        L.mo_sim_conservation.argtypes = [C.c_void_p, f64p]
        L.mo_sim_cfl_dt.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)]
        L.mo_sim_cfl_dt.restype = C.c_double
        L.mo_node_boundary.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.c_int, f64p]
        L.mo_node_boundary.restype = None
        L.mo_speed_clamp.argtypes = [C.c_double, f64p]
        L.mo_speed_clamp.restype = None
        L.mo_damage_update.argtypes = [C.c_double, C.c_double]
        L.mo_damage_update.restype = C.c_double
        L.mo_p2g_mls.argtypes = [C.POINTER(Params), C.c_int, u64p, f64p, f64p, f64p, f64p,
                                 f64p, f64p, f64p, f64p, C.c_double, C.c_int, u64p, f64p]
        L.mo_p2g_mls.restype = C.c_int
        L.mo_g2p_affine.argtypes = [C.c_int, f64p, f64p, f64p, f64p, f64p, f64p]
        _lib = L
    return _lib


def ref():
    """The reference's own SPGridMask/Materials build, or None if absent."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.ref_masks.argtypes = [np.ctypeslib.ndpointer(np.uint64, flags="C")]
        R.ref_linear_offset.argtypes = [C.c_int, C.c_int, C.c_int]
        R.ref_linear_offset.restype = C.c_uint64
        R.ref_material.argtypes = [C.c_double, C.c_double,
                                   np.ctypeslib.ndpointer(np.float64, flags="C")]
        _ref = R
    return _ref


def default_params(**kw):
    p = Params()
    lib().mo_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "gravity":
            p.gravity[:] = list(v)
        else:
            setattr(p, k, v)
    return p


def masks():
    m = np.zeros(3, np.uint64)
    lib().mo_masks(m)
    return m


def cell_keys(x, N):
    x = np.ascontiguousarray(x)
    n = x.shape[0]
    keys = np.zeros(n, np.uint64)
    if x.dtype == np.float32:
        lib().mo_cell_keys_f32(n, x.reshape(-1), N, keys)
    else:
        lib().mo_cell_keys_f64(n, np.ascontiguousarray(x, np.float64).reshape(-1), N, keys)
    return keys


def sort_keys(keys):
    keys = np.ascontiguousarray(keys, np.uint64)
    n = keys.shape[0]
    perm = np.zeros(n, np.int32)
    out = np.zeros(n, np.uint64)
    lib().mo_sort_keys(n, keys, perm, out)
    return perm, out


def page_tables(sorted_keys, vpage=256):
    """Canonical page tables.  Returns a dict of numpy arrays and counts."""
    s = np.ascontiguousarray(sorted_keys, np.uint64)
    n = s.shape[0]
    cap_p, cap_tp, cap_vp = n + 1, 8 * n + 8, n + 1
    page_key = np.zeros(cap_tp, np.uint64)
    p2p = np.zeros(cap_p + 1, np.int32)
    adj = np.zeros(7 * cap_p, np.int32)
    vpo = np.zeros(cap_p + 1, np.int32)
    target = np.zeros(cap_vp, np.int32)
    p2c = np.zeros(max(n, 1), np.int32)
    c2p = np.zeros(n + 1, np.int32)
    counts = np.zeros(4, np.int32)
    rc = lib().mo_page_tables(n, s, vpage, cap_p, cap_tp, cap_vp, page_key, p2p, adj, vpo,
                              target, p2c, c2p, counts)
    if rc != 0:
        raise RuntimeError("oracle page-table overflow")
    npg, ntot, nv, ncell = (int(c) for c in counts)
    return {
        "numPage": npg, "numTotalPage": ntot, "numVirtualPage": nv, "numCell": ncell,
        "page_key": page_key[:ntot].copy(), "page2particle": p2p[:npg + 1].copy(),
        "adj": adj[:7 * npg].reshape(7, npg).copy(), "vpo": vpo[:npg + 1].copy(),
        "target": target[:nv].copy(), "particle2cell": p2c[:n].copy(),
        "cell2particle": c2p[:ncell + 1].copy(),
    }


def svd3(F):
    """F: 3x3 array (row/col as a matrix).  Returns U, S, V with F = U diag(S) V^T."""
    Fc = np.ascontiguousarray(np.asarray(F, np.float64).T).reshape(-1)  # column-major
    U = np.zeros(9)
    S = np.zeros(3)
    V = np.zeros(9)
    lib().mo_svd3(Fc, U, S, V)
    return U.reshape(3, 3).T.copy(), S, V.reshape(3, 3).T.copy()


def svd3_sweeps():
    return int(lib().mo_svd3_last_sweeps())


def inverse3(A):
    A = np.ascontiguousarray(A, np.float64).reshape(-1)
    out = np.zeros(9)
    lib().mo_inverse3(A, out)
    return out.reshape(3, 3)


def vdot(a, b, mode):
    """The PCG's "dot" of the oracle: mode 0 the reference's largest
    128-entry block sum, mode 1 the true dot product."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return float(lib().mo_vdot(len(a), a, b, mode))


def von_mises(params, F, c, g, alpha, psi):
    """One particle's return mapping.  F as a 3x3 matrix.  Returns dict."""
    Fc = np.ascontiguousarray(np.asarray(F, np.float64).T).reshape(-1).copy()
    gg = np.array([g], np.float64)
    al = np.array([alpha], np.float64)
    ps = np.array([psi], np.float64)
    fp = np.zeros(1)
    sig = np.zeros(9)
    vol = np.zeros(1)
    lib().mo_von_mises(C.byref(params), Fc, float(c), gg, al, ps, fp, sig, vol)
    return {"F": Fc.reshape(3, 3).T.copy(), "g": gg[0], "alpha": al[0], "psi": ps[0],
            "FP": fp[0], "sigma": sig.reshape(3, 3).T.copy(), "vol": vol[0]}


def neo_hookean(params, F, c, psi):
    """One particle's Neo-Hookean phase-field update.  Returns dict."""
    Fc = np.ascontiguousarray(np.asarray(F, np.float64).T).reshape(-1).copy()
    ps = np.array([psi], np.float64)
    fp = np.zeros(1)
    sig = np.zeros(9)
    vol = np.zeros(1)
    lib().mo_neo_hookean(C.byref(params), Fc, float(c), ps, fp, sig, vol)
    return {"psi": ps[0], "FP": fp[0], "sigma": sig.reshape(3, 3).T.copy(), "vol": vol[0]}


def fixed_corotated(params, F):
    """One particle's fixed corotated stress contribution V0 P F^T."""
    Fc = np.ascontiguousarray(np.asarray(F, np.float64).T).reshape(-1).copy()
    sig = np.zeros(9)
    lib().mo_fixed_corotated(C.byref(params), Fc, sig)
    return sig.reshape(3, 3).T.copy()


class Sim:
    """Serial FP64 substep simulator (ExplicitTimeIntegrator::integrate)."""

    def __init__(self, params, x, v, mass):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        v = np.ascontiguousarray(v, np.float64).reshape(-1)
        mass = np.ascontiguousarray(mass, np.float64).reshape(-1)
        self.n = mass.shape[0]
        self.params = params
        self.h = lib().mo_sim_create(C.byref(params), self.n, x, v, mass)

    def __del__(self):
        if getattr(self, "h", None):
            lib().mo_sim_destroy(self.h)
            self.h = None

    def substep(self, dt):
        rc = lib().mo_sim_substep(self.h, float(dt))
        if rc != 0:
            raise RuntimeError("oracle substep failed: %d" % rc)

    def get(self, name):
        fid, w = FIELDS[name]
        out = np.zeros(self.n * w)
        lib().mo_sim_get(self.h, fid, out)
        if w == 3:
            return out.reshape(self.n, 3)
        if w == 9:  # per particle column-major -> (n,3,3) matrices
            return out.reshape(self.n, 3, 3).transpose(0, 2, 1).copy()
        return out

    def set(self, name, arr):
        fid, w = FIELDS[name]
        a = np.asarray(arr, np.float64)
        if w == 9:
            a = a.reshape(self.n, 3, 3).transpose(0, 2, 1)
        lib().mo_sim_set(self.h, fid, np.ascontiguousarray(a).reshape(-1))

    def order(self):
        o = np.zeros(self.n, np.int32)
        k = np.zeros(self.n, np.uint64)
        lib().mo_sim_get_order(self.h, o, k)
        return o, k

    def round_f32(self):
        """Round the particle state to float32 (an FP32 store)."""
        lib().mo_sim_round_f32(self.h)

    def counts(self):
        c = np.zeros(6, np.int32)
        lib().mo_sim_counts(self.h, c)
        return {"numPage": int(c[0]), "numTotalPage": int(c[1]), "numVirtualPage": int(c[2]),
                "numCell": int(c[3]), "cg_iters": int(c[4]), "cg_breakdown": int(c[5])}

    def cg_norm(self):
        return lib().mo_sim_cg_norm(self.h)

    def grid(self, ch, stage=2):
        tp = self.counts()["numTotalPage"]
        out = np.zeros(tp * 64)
        lib().mo_sim_grid(self.h, stage, ch, out)
        return out

    def pages(self):
        tp = self.counts()["numTotalPage"]
        out = np.zeros(max(tp, 1), np.uint64)
        lib().mo_sim_pages(self.h, out)
        return out[:tp]

    def conservation(self):
        out = np.zeros(4)
        lib().mo_sim_conservation(self.h, out)
        return out

    def cfl_dt(self, dt_default, cfl=0.3, wave=False):
        """(dt, largest speed) of MPMSimulator::computeDt on the current state."""
        sp = C.c_double(0.0)
        dt = lib().mo_sim_cfl_dt(self.h, float(dt_default), float(cfl), 1 if wave else 0, C.byref(sp))
        return float(dt), float(sp.value)


def node_boundary(params, i, j, k, v):
    """Grid node (i, j, k)'s velocity after the grips and the sticky walls."""
    out = np.ascontiguousarray(v, np.float64).copy()
    lib().mo_node_boundary(C.byref(params), int(i), int(j), int(k), out)
    return out


def speed_clamp(vmax, v):
    out = np.ascontiguousarray(v, np.float64).copy()
    lib().mo_speed_clamp(float(vmax), out)
    return out


def damage_update(c, dc):
    return float(lib().mo_damage_update(float(c), float(dc)))


def g2p_affine(N, x, A, b):
    """G2P gather from the grid field v = A x + b.  Returns (v_p, B, G) with B, G as 3x3."""
    cm = np.ascontiguousarray(np.asarray(A, np.float64).T).reshape(-1)
    vp = np.zeros(3)
    Bm = np.zeros(9)
    G = np.zeros(9)
    lib().mo_g2p_affine(N, np.ascontiguousarray(x, np.float64), cm, np.ascontiguousarray(b, np.float64),
                        vp, Bm, G)
    return vp, Bm.reshape(3, 3).T.copy(), G.reshape(3, 3).T.copy()


def p2g_mls(params, sorted_keys, x, v, B, sigma, m, c, vol, FP, dt, page_key):
    """Single-stage P2G-MLS on particles already in sorted order.
    B, sigma: (n,3,3) matrices.  Returns (8, tp*64) channels m, p0, p1, p2,
    sum w c, sum w, sum V w, sum V FP w."""
    n = x.shape[0]
    tp = page_key.shape[0]
    out = np.zeros(8 * tp * 64)
    cm = lambda a: np.ascontiguousarray(np.asarray(a, np.float64).transpose(0, 2, 1)).reshape(-1)
    f64 = lambda a: np.ascontiguousarray(a, np.float64).reshape(-1)
    lib().mo_p2g_mls(C.byref(params), n, np.ascontiguousarray(sorted_keys, np.uint64), f64(x), f64(v),
                     cm(B), cm(sigma), f64(m), f64(c), f64(vol), f64(FP), float(dt), tp,
                     np.ascontiguousarray(page_key, np.uint64), out)
    return out.reshape(8, tp * 64)
