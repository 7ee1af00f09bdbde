"""ctypes binding of libmpmx_hip.so (the C ABI declared in include/mpmx.h).

This is the Python face of the product library, used by the tests and the
benchmark.  It never falls back to anything: if the HIP library is missing the
import of the library raises.  The C++ drop-in classes (MPMSimulator,
ExplicitTimeIntegrator, Particles, SPGrid, DomainTransformer) in ``host/`` sit
on the same C ABI.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmpmx_hip.so")

OK, ERR_HIP, ERR_ARG, ERR_CAPACITY, ERR_OUT_OF_DOMAIN, ERR_NONFINITE, ERR_CG_BREAKDOWN, ERR_NO_DEVICE = range(8)
FP32, MIXED, MIXED_F16 = 0, 1, 2
FIELDS = {"x": (0, 3), "v": (1, 3), "F": (2, 9), "B": (3, 9), "c": (4, 1), "alpha": (5, 1),
          "psi": (6, 1), "g": (7, 1), "mass": (8, 1), "FP": (9, 1), "vol": (10, 1), "sigma": (11, 9)}

EXPORTS = [
    "mpmx_default_params", "mpmx_default_config", "mpmx_status_string", "mpmx_last_error",
    "mpmx_device_count", "mpmx_create", "mpmx_destroy", "mpmx_set_particles", "mpmx_bin_sort",
    "mpmx_rebuild", "mpmx_stress", "mpmx_p2g", "mpmx_grid_update", "mpmx_phase_field", "mpmx_g2p",
    "mpmx_substep", "mpmx_run", "mpmx_compute_dt", "mpmx_sync", "mpmx_get_field", "mpmx_set_field",
    "mpmx_get_order", "mpmx_get_tables", "mpmx_get_grid", "mpmx_get_grid_p2g", "mpmx_get_stats",
    "mpmx_timing_begin", "mpmx_timing_end", "mpmx_time_run", "mpmx_scene", "mpmx_get_device_pointers",
    "mpmx_max_velocity_sq", "mpmx_create_batch", "mpmx_set_particles_batch", "mpmx_num_scenes",
    "mpmx_get_scene_stats", "mpmx_scene_batch", "mpmx_read_obj_points", "mpmx_sample_sdf",
    "mpmx_write_particles", "mpmx_save_checkpoint", "mpmx_load_checkpoint", "mpmx_get_tuning",
    "mpmx_set_tuning", "mpmx_autotune", "mpmx_run_cfl", "mpmx_compute_dt_noalloc",
]


class Params(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("transfer", C.c_int), ("phase_field", C.c_int), ("dot_mode", C.c_int),
        ("cg_max_iter", C.c_int), ("vpage", C.c_int), ("wall_cells", C.c_int), ("grip_lo", C.c_int),
        ("grip_hi", C.c_int), ("grip_v", C.c_double), ("E", C.c_double), ("nu", C.c_double),
        ("V0", C.c_double), ("yield_stress", C.c_double), ("hardening", C.c_double),
        ("alpha_c", C.c_double), ("dlambda_cap", C.c_double), ("eps_cap", C.c_double),
        ("Gc0", C.c_double), ("residual_c", C.c_double), ("ps", C.c_double), ("theta", C.c_double),
        ("M", C.c_double), ("gravity", C.c_double * 3), ("vmax", C.c_double), ("cg_tol", C.c_double),
        ("crit_stretch", C.c_double), ("model", C.c_int),
    ]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            val = getattr(self, name)
            d[name] = list(val) if name == "gravity" else val
        return d


class Config(C.Structure):
    _fields_ = [("max_particles", C.c_int64), ("max_pages", C.c_int64), ("device", C.c_int),
                ("precision", C.c_int), ("svd_gap", C.c_double), ("cg_check_every", C.c_int),
                ("p2g_variant", C.c_int), ("block", C.c_int), ("stress_fp64", C.c_int), ("use_graph", C.c_int),
                ("store_stress", C.c_int), ("sort_variant", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("num_particles", C.c_int64), ("num_page", C.c_int), ("num_total_page", C.c_int),
                ("num_virtual_page", C.c_int), ("cg_iters", C.c_int), ("cg_breakdown", C.c_int),
                ("substeps", C.c_int), ("cg_norm", C.c_double), ("mass", C.c_double),
                ("momentum", C.c_double * 3), ("kinetic_energy", C.c_double), ("svd_fallbacks", C.c_int64)]


class SceneStats(C.Structure):
    _fields_ = [("num_particles", C.c_int64), ("first_particle", C.c_int64), ("num_page", C.c_int),
                ("num_total_page", C.c_int), ("cg_iters", C.c_int), ("cg_breakdown", C.c_int),
                ("cg_norm", C.c_double), ("mass", C.c_double), ("momentum", C.c_double * 3),
                ("kinetic_energy", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("particle_wg", C.c_int), ("p2g_pct", C.c_int), ("g2p_pct", C.c_int), ("cg_pct", C.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int), ("total_ms", C.c_double)]


class MpmxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("mpmx status %d: %s" % (status, msg))
        self.status = status


_lib = None


def build():
    """Compile libmpmx_hip.so for gfx950 (hipcc cross-compiles without a GPU)."""
    jobs = str(min(16, os.cpu_count() or 4))
    subprocess.run(["make", "-s", "-j", jobs, "-C", os.path.join(HERE, "csrc")], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libmpmx_hip.so is not built (%s); run build()" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
    sig = {
        "mpmx_default_params": ([C.POINTER(Params)], None),
        "mpmx_default_config": ([C.POINTER(Config)], None),
        "mpmx_status_string": ([C.c_int], C.c_char_p),
        "mpmx_last_error": ([vp], C.c_char_p),
        "mpmx_device_count": ([], C.c_int),
        "mpmx_create": ([C.POINTER(Params), C.POINTER(Config), C.POINTER(vp)], C.c_int),
        "mpmx_destroy": ([vp], C.c_int),
        "mpmx_set_particles": ([vp, C.c_int64, f64p, f64p, f64p], C.c_int),
        "mpmx_bin_sort": ([vp], C.c_int),
        "mpmx_rebuild": ([vp], C.c_int),
        "mpmx_stress": ([vp], C.c_int),
        "mpmx_p2g": ([vp, C.c_double], C.c_int),
        "mpmx_grid_update": ([vp, C.c_double], C.c_int),
        "mpmx_phase_field": ([vp, C.c_double], C.c_int),
        "mpmx_g2p": ([vp, C.c_double], C.c_int),
        "mpmx_substep": ([vp, C.c_double], C.c_int),
        "mpmx_run": ([vp, C.c_int, C.c_double], C.c_int),
        "mpmx_compute_dt": ([vp, C.c_double, C.c_double, C.POINTER(C.c_double)], C.c_int),
        "mpmx_run_cfl": ([vp, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "mpmx_compute_dt_noalloc": ([vp, C.c_double, C.c_double, C.POINTER(C.c_double)], C.c_int),
        "mpmx_sync": ([vp], C.c_int),
        "mpmx_get_field": ([vp, C.c_int, f64p, C.c_size_t], C.c_int),
        "mpmx_set_field": ([vp, C.c_int, f64p, C.c_size_t], C.c_int),
        "mpmx_get_order": ([vp, i32p, u64p], C.c_int),
        "mpmx_get_tables": ([vp, u64p, i32p, i32p, i32p, i32p, i32p, i32p, i32p], C.c_int),
        "mpmx_get_grid": ([vp, C.c_int, f32p, C.c_size_t], C.c_int),
        "mpmx_get_grid_p2g": ([vp, C.c_int, f32p, C.c_size_t], C.c_int),
        "mpmx_get_stats": ([vp, C.POINTER(Stats)], C.c_int),
        "mpmx_timing_begin": ([vp], C.c_int),
        "mpmx_timing_end": ([vp, C.POINTER(KernelTime), C.c_int, C.POINTER(C.c_int)], C.c_int),
        "mpmx_time_run": ([vp, C.c_int, C.c_double, C.POINTER(C.c_double)], C.c_int),
        "mpmx_scene": ([C.c_int, C.c_uint64, C.c_double, C.POINTER(Params), vp, vp, vp,
                        C.POINTER(C.c_double)], C.c_int64),
        "mpmx_get_device_pointers": ([vp, vp], C.c_int),
        "mpmx_max_velocity_sq": ([vp, C.POINTER(C.c_double)], C.c_int),
        "mpmx_create_batch": ([C.c_int, C.POINTER(Params), C.POINTER(Config), C.POINTER(vp)], C.c_int),
        "mpmx_set_particles_batch": ([vp, np.ctypeslib.ndpointer(np.int64, flags="C"), f64p, f64p, f64p],
                                     C.c_int),
        "mpmx_num_scenes": ([vp], C.c_int),
        "mpmx_get_scene_stats": ([vp, C.c_int, C.POINTER(SceneStats)], C.c_int),
        "mpmx_scene_batch": ([C.c_int, C.c_uint64, C.c_double, C.POINTER(C.c_int), vp, vp, vp, vp, vp,
                              C.POINTER(C.c_double)], C.c_int64),
        "mpmx_read_obj_points": ([C.c_char_p, C.c_double, vp, C.c_int64], C.c_int64),
        "mpmx_sample_sdf": ([C.c_char_p, C.c_uint64, C.c_int, C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.c_double,
                             vp, C.c_int64], C.c_int64),
        "mpmx_write_particles": ([vp, C.c_char_p], C.c_int),
        "mpmx_save_checkpoint": ([vp, C.c_char_p], C.c_int),
        "mpmx_load_checkpoint": ([C.c_char_p, C.c_int, C.POINTER(vp)], C.c_int),
        "mpmx_get_tuning": ([vp, C.POINTER(Tuning)], C.c_int),
        "mpmx_set_tuning": ([vp, C.POINTER(Tuning)], C.c_int),
        "mpmx_autotune": ([vp, C.c_double, C.c_int, C.POINTER(Tuning)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def default_params(**kw):
    p = Params()
    lib().mpmx_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "gravity":
            p.gravity[:] = list(v)
        else:
            setattr(p, k, v)
    return p


def default_config(**kw):
    c = Config()
    lib().mpmx_default_config(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def device_count():
    return int(lib().mpmx_device_count())


def scene(scene_id, seed=1, scale=1.0):
    """Seeded synthetic scene: returns (params, x[n,3], v[n,3], mass[n], dt)."""
    L = lib()
    p = Params()
    dt = C.c_double()
    n = L.mpmx_scene(scene_id, seed, scale, C.byref(p), None, None, None, C.byref(dt))
    if n < 0:
        raise ValueError("unknown scene %d" % scene_id)
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    m = np.zeros(n)
    L.mpmx_scene(scene_id, seed, scale, C.byref(p), x.ctypes.data, v.ctypes.data, m.ctypes.data,
                 C.byref(dt))
    return p, x, v, m, dt.value


def scene_batch(config, seed=1, scale=1.0, num_scenes=0):
    """Seeded batch of scenes: returns (params list, counts[S], x[n,3], v[n,3], mass[n], dt)."""
    L = lib()
    S = C.c_int(num_scenes)
    dt = C.c_double()
    n = L.mpmx_scene_batch(config, seed, scale, C.byref(S), None, None, None, None, None, C.byref(dt))
    if n < 0:
        raise ValueError("unknown batch config %d" % config)
    ps = (Params * S.value)()
    counts = np.zeros(S.value, np.int64)
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    m = np.zeros(n)
    L.mpmx_scene_batch(config, seed, scale, C.byref(S), C.cast(ps, C.c_void_p), counts.ctypes.data,
                       x.ctypes.data, v.ctypes.data, m.ctypes.data, C.byref(dt))
    return list(ps), counts, x, v, m, dt.value


def read_obj_points(path, y_shift=-0.2):
    """Point-list OBJ as the reference's Sample::readObj loads it: x[n, 3]."""
    L = lib()
    n = L.mpmx_read_obj_points(path.encode(), y_shift, None, 0)
    if n < 0:
        raise OSError("cannot read %s" % path)
    x = np.zeros((n, 3))
    L.mpmx_read_obj_points(path.encode(), y_shift, x.ctypes.data, n)
    return x


def sample_sdf(path, N, center, res, samples_per_cell=20.0, seed=1):
    """Particles inside a level-set file placed as Sample::Initial_data does: x[n, 3]."""
    L = lib()
    cen = (C.c_uint * 3)(*center)
    rs = (C.c_uint * 3)(*res)
    n = L.mpmx_sample_sdf(path.encode(), seed, N, cen, rs, samples_per_cell, None, 0)
    if n < 0:
        raise OSError("cannot read %s" % path)
    x = np.zeros((n, 3))
    L.mpmx_sample_sdf(path.encode(), seed, N, cen, rs, samples_per_cell, x.ctypes.data, n)
    return x


class Context:
    """One simulation context (device memory + HIP stream) over the C ABI.
    `params` is one Params (one scene) or a list of them (a batch of scenes)."""

    def __init__(self, params, config=None):
        self.L = lib()
        self.h = C.c_void_p()
        cfg = config if config is not None else default_config()
        if isinstance(params, (list, tuple)):
            arr = (Params * len(params))(*params)
            st = self.L.mpmx_create_batch(len(params), arr, C.byref(cfg), C.byref(self.h))
            self.num_scenes = len(params)
        else:
            st = self.L.mpmx_create(C.byref(params), C.byref(cfg), C.byref(self.h))
            self.num_scenes = 1
        if st != OK:
            raise MpmxError(st, self.L.mpmx_status_string(st).decode())
        self.params = params
        self.n = 0

    def close(self):
        if self.h:
            self.L.mpmx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st != OK:
            msg = self.L.mpmx_last_error(self.h)
            raise MpmxError(st, msg.decode() if msg else "")

    def set_particles(self, x, v, mass):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        v = np.ascontiguousarray(v, np.float64).reshape(-1)
        mass = np.ascontiguousarray(mass, np.float64).reshape(-1)
        self.n = mass.shape[0]
        self._chk(self.L.mpmx_set_particles(self.h, self.n, x, v, mass))

    @classmethod
    def load_checkpoint(cls, path, device=0):
        """A context restored from mpmx_save_checkpoint (continues bit-exactly)."""
        self = cls.__new__(cls)
        self.L = lib()
        self.h = C.c_void_p()
        st = self.L.mpmx_load_checkpoint(path.encode(), device, C.byref(self.h))
        if st != OK:
            raise MpmxError(st, self.L.mpmx_status_string(st).decode())
        self.num_scenes = int(self.L.mpmx_num_scenes(self.h))
        self.params = None
        self.n = int(self.stats()["num_particles"])
        return self

    def autotune(self, dt, trial_steps=5):
        """Pick the fastest launch shapes on the current state (which it restores)."""
        t = Tuning()
        self._chk(self.L.mpmx_autotune(self.h, dt, trial_steps, C.byref(t)))
        return t.as_dict()

    def tuning(self):
        t = Tuning()
        self._chk(self.L.mpmx_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning()
        self._chk(self.L.mpmx_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        self._chk(self.L.mpmx_set_tuning(self.h, C.byref(t)))

    def save_checkpoint(self, path):
        self._chk(self.L.mpmx_save_checkpoint(self.h, path.encode()))

    def write_particles(self, path):
        self._chk(self.L.mpmx_write_particles(self.h, path.encode()))

    def set_particles_batch(self, counts, x, v, mass):
        counts = np.ascontiguousarray(counts, np.int64)
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        v = np.ascontiguousarray(v, np.float64).reshape(-1)
        mass = np.ascontiguousarray(mass, np.float64).reshape(-1)
        self.n = mass.shape[0]
        self._chk(self.L.mpmx_set_particles_batch(self.h, counts, x, v, mass))

    def scene_stats(self, s):
        st = SceneStats()
        self._chk(self.L.mpmx_get_scene_stats(self.h, s, C.byref(st)))
        return {"num_particles": st.num_particles, "first_particle": st.first_particle,
                "numPage": st.num_page, "numTotalPage": st.num_total_page, "cg_iters": st.cg_iters,
                "cg_breakdown": st.cg_breakdown, "cg_norm": st.cg_norm, "mass": st.mass,
                "momentum": list(st.momentum), "kinetic_energy": st.kinetic_energy}

    def bin_sort(self): self._chk(self.L.mpmx_bin_sort(self.h))
    def rebuild(self): self._chk(self.L.mpmx_rebuild(self.h))
    def stress(self): self._chk(self.L.mpmx_stress(self.h))
    def p2g(self, dt): self._chk(self.L.mpmx_p2g(self.h, dt))
    def grid_update(self, dt): self._chk(self.L.mpmx_grid_update(self.h, dt))
    def phase_field(self, dt): self._chk(self.L.mpmx_phase_field(self.h, dt))
    def g2p(self, dt): self._chk(self.L.mpmx_g2p(self.h, dt))
    def substep(self, dt): self._chk(self.L.mpmx_substep(self.h, dt))
    def run(self, nsteps, dt): self._chk(self.L.mpmx_run(self.h, nsteps, dt))
    def sync(self): self._chk(self.L.mpmx_sync(self.h))

    def compute_dt(self, dt_default, cfl=0.3):
        out = C.c_double()
        self._chk(self.L.mpmx_compute_dt(self.h, dt_default, cfl, C.byref(out)))
        return out.value

    def max_velocity_sq(self):
        out = C.c_double()
        self._chk(self.L.mpmx_max_velocity_sq(self.h, C.byref(out)))
        return out.value

    def run_cfl(self, nsteps, dt_default, cfl=0.3, wave=False):
        """nsteps substeps, each with the device-side CFL step; returns the simulated time."""
        t = C.c_double()
        self._chk(self.L.mpmx_run_cfl(self.h, nsteps, dt_default, cfl, 1 if wave else 0, C.byref(t)))
        return t.value

    def get(self, name):
        fid, w = FIELDS[name]
        out = np.zeros(self.n * w)
        self._chk(self.L.mpmx_get_field(self.h, fid, out, out.size))
        if w == 3:
            return out.reshape(self.n, 3)
        if w == 9:
            return out.reshape(self.n, 3, 3).transpose(0, 2, 1).copy()
        return out

    def set(self, name, arr):
        fid, w = FIELDS[name]
        a = np.asarray(arr, np.float64)
        if w == 9:
            a = a.reshape(self.n, 3, 3).transpose(0, 2, 1)
        a = np.ascontiguousarray(a).reshape(-1)
        self._chk(self.L.mpmx_set_field(self.h, fid, a, a.size))

    def order(self):
        o = np.zeros(self.n, np.int32)
        k = np.zeros(self.n, np.uint64)
        self._chk(self.L.mpmx_get_order(self.h, o, k))
        return o, k

    def stats(self):
        s = Stats()
        self._chk(self.L.mpmx_get_stats(self.h, C.byref(s)))
        return {"num_particles": s.num_particles, "numPage": s.num_page,
                "numTotalPage": s.num_total_page, "numVirtualPage": s.num_virtual_page,
                "cg_iters": s.cg_iters, "cg_breakdown": s.cg_breakdown, "substeps": s.substeps,
                "cg_norm": s.cg_norm, "mass": s.mass, "momentum": list(s.momentum),
                "kinetic_energy": s.kinetic_energy, "svd_fallbacks": s.svd_fallbacks}

    def svd_fallbacks(self):
        return int(self.stats()["svd_fallbacks"])

    def tables(self):
        st = self.stats()
        n = self.n
        npg, nt, nv = st["numPage"], st["numTotalPage"], st["numVirtualPage"]
        pk = np.zeros(max(nt, 1), np.uint64)
        p2p = np.zeros(npg + 1, np.int32)
        adj = np.zeros(max(7 * npg, 1), np.int32)
        vpo = np.zeros(npg + 1, np.int32)
        tgt = np.zeros(max(nv, 1), np.int32)
        p2c = np.zeros(n, np.int32)
        c2p = np.zeros(n + 1, np.int32)
        counts = np.zeros(4, np.int32)
        self._chk(self.L.mpmx_get_tables(self.h, pk, p2p, adj, vpo, tgt, p2c, c2p, counts))
        ncell = int(counts[3])
        return {"numPage": npg, "numTotalPage": nt, "numVirtualPage": nv, "numCell": ncell,
                "page_key": pk[:nt], "page2particle": p2p, "adj": adj[:7 * npg].reshape(7, npg),
                "vpo": vpo, "target": tgt[:nv], "particle2cell": p2c, "cell2particle": c2p[:ncell + 1]}

    def grid(self, ch, p2g=False):
        nt = self.stats()["numTotalPage"]
        out = np.zeros(nt * 64, np.float32)
        fn = self.L.mpmx_get_grid_p2g if p2g else self.L.mpmx_get_grid
        self._chk(fn(self.h, ch, out, out.size))
        return out

    def enable_p2g_copy(self):
        """Keep a copy of the raw P2G grid from the next grid update on."""
        out = np.zeros(1, np.float32)
        st = self.L.mpmx_get_grid_p2g(self.h, 0, out, 0)
        if st not in (OK, ERR_ARG):
            self._chk(st)

    def timing_begin(self):
        self._chk(self.L.mpmx_timing_begin(self.h))

    def timing_end(self):
        arr = (KernelTime * 64)()
        n = C.c_int()
        self._chk(self.L.mpmx_timing_end(self.h, arr, 64, C.byref(n)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(min(n.value, 64))}

    def time_run(self, nsteps, dt):
        ms = C.c_double()
        self._chk(self.L.mpmx_time_run(self.h, nsteps, dt, C.byref(ms)))
        return ms.value
