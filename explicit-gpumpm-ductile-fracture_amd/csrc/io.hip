// io.hip — particle files of a context: a PLY dump (the reference declares
// MPMSimulator::writePartio, Simulator.h:35, but never defines it), a
// checkpoint that resumes bit-exactly, and the launch-shape autotuner, which
// runs its trials on an in-memory snapshot of the same state.  Host code over
// the context's buffers.
//
// Checkpoint layout (little-endian):
//   "MPMXCKP1", int32 version 2, int32 num_scenes, int32 precision,
//   int32 sorted, int64 n, int64 substeps, int64 cap_total_pages, int64 cg_spec,
//   mpmx_params[num_scenes], mpmx_config, int64 counts[num_scenes],
//   then the particle buffers in the context's current order, as stored on
//   the device: x, v, B, F, c, alpha, psi, g, mass, id, keys (FP32: float
//   fields; mixed: v, B, c, alpha, psi, g as 16-bit codes, F float, or as
//   16-bit codes for MPMX_MIXED_F16).  A loader checks the counts against n
//   and the buffer bytes against the file size before it allocates.
// Resuming restores every buffer the next substep reads, so the run continues
// exactly as it would have without the interruption.
#include <cstdio>
#include <cstring>
#include <vector>

#include "mpmx_internal.h"

using namespace mpmx;

namespace {

constexpr char kMagic[8] = {'M', 'P', 'M', 'X', 'C', 'K', 'P', '1'};

struct Header {
    int32_t version, num_scenes, precision, sorted;
    int64_t n, substeps, cap_tp, cg_spec;
};

// (device pointer, bytes) of every particle buffer, current order
std::vector<std::pair<void*, size_t>> particle_buffers(mpmx_ctx* c) {
    const size_t n = (size_t)c->n;
    std::vector<std::pair<void*, size_t>> b;
    if (c->cfg.precision != MPMX_FP32) {
        const MState& s = c->ms[c->cur];
        b = {{s.x, 12 * n}, {s.v, 6 * n}, {s.B, 18 * n}};
        if (s.F32) b.push_back({s.F32, 36 * n});
        else b.push_back({s.F, 18 * n});
        b.insert(b.end(), {{s.c, 2 * n}, {s.alpha, 2 * n}, {s.psi, 2 * n}, {s.g, 2 * n}, {s.mass, 4 * n},
                           {s.id, 4 * n}});
    } else {
        const PState& s = c->ps[c->cur];
        b = {{s.x, 12 * n}, {s.v, 12 * n},   {s.B, 36 * n}, {s.F, 36 * n}, {s.c, 4 * n},   {s.alpha, 4 * n},
             {s.psi, 4 * n}, {s.g, 4 * n}, {s.mass, 4 * n}, {s.id, 4 * n}};
    }
    b.push_back({c->slot_keys(), 4 * n});
    return b;
}

mpmx_status io_fail(mpmx_ctx* c, mpmx_status s, const char* msg) {
    if (c) c->last_error = msg;
    return s;
}

// device copy of the particle state and the host flags that go with it
struct Snapshot {
    std::vector<std::pair<void*, size_t>> bufs;
    char* mem = nullptr;
    int cur = 0, substeps = 0, cg_spec = 1;
    bool sorted = false;
    uint32_t* keys = nullptr;       // which key buffer is current (a sort may swap them)
    uint32_t* keys_alt = nullptr;
};

hipError_t snapshot_take(mpmx_ctx* c, Snapshot& sn) {
    sn.bufs = particle_buffers(c);
    size_t total = 0;
    for (auto& b : sn.bufs) total += (b.second + 255) / 256 * 256;
    hipError_t e = hipMalloc((void**)&sn.mem, total);
    if (e != hipSuccess) return e;
    size_t off = 0;
    for (auto& b : sn.bufs) {
        if ((e = hipMemcpyAsync(sn.mem + off, b.first, b.second, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess)
            return e;
        off += (b.second + 255) / 256 * 256;
    }
    sn.cur = c->cur;
    sn.keys = c->keys;
    sn.keys_alt = c->keys_alt;
    sn.sorted = c->sorted;
    sn.substeps = c->substeps;
    sn.cg_spec = c->cg_spec;
    return hipStreamSynchronize(c->stream);
}

hipError_t snapshot_restore(mpmx_ctx* c, const Snapshot& sn) {
    size_t off = 0;
    hipError_t e = hipSuccess;
    for (auto& b : sn.bufs) {
        if ((e = hipMemcpyAsync(b.first, sn.mem + off, b.second, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess)
            return e;
        off += (b.second + 255) / 256 * 256;
    }
    c->cur = sn.cur;
    c->keys = sn.keys;
    c->keys_alt = sn.keys_alt;
    c->sorted = sn.sorted;
    // keys (the last sort's output) was restored only if the particles sat
    // in its order; otherwise the sorts in between may have left it unsorted.
    // The page tables are the last run's either way.
    if (!sn.sorted) c->old_keys_sorted = false;
    c->tables_match_keys = false;
    c->substeps = sn.substeps;
    c->cg_spec = sn.cg_spec;
    return hipStreamSynchronize(c->stream);
}

}  // namespace

extern "C" {

mpmx_status mpmx_save_checkpoint(mpmx_ctx* c, const char* path) {
    if (!c || !path || c->n == 0) return MPMX_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
        return io_fail(c, MPMX_ERR_HIP, "device sync failed");
    FILE* f = std::fopen(path, "wb");
    if (!f) return io_fail(c, MPMX_ERR_ARG, "cannot open checkpoint for writing");
    Header h{2, c->num_scenes, c->cfg.precision, c->sorted ? 1 : 0, c->n, c->substeps, c->cap_tp, c->cg_spec};
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(&h, sizeof(h), 1, f) == 1 &&
              std::fwrite(c->scene_P.data(), sizeof(mpmx_params), c->scene_P.size(), f) == c->scene_P.size() &&
              std::fwrite(&c->cfg, sizeof(mpmx_config), 1, f) == 1 &&
              std::fwrite(c->scene_n.data(), sizeof(int64_t), c->scene_n.size(), f) == c->scene_n.size();
    std::vector<char> host;
    for (auto& [ptr, bytes] : particle_buffers(c)) {
        if (!ok) break;
        host.resize(bytes);
        ok = hipMemcpy(host.data(), ptr, bytes, hipMemcpyDeviceToHost) == hipSuccess &&
             std::fwrite(host.data(), 1, bytes, f) == bytes;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? MPMX_OK : io_fail(c, MPMX_ERR_HIP, "checkpoint write failed");
}

static mpmx_status load_checkpoint(const char* path, int device, mpmx_ctx** out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return MPMX_ERR_ARG;
    // the file's size bounds everything read from it
    long fsize = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) fsize = std::ftell(f);
    std::rewind(f);
    char magic[8];
    Header h;
    bool ok = fsize > 0 && std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
              std::fread(&h, sizeof(h), 1, f) == 1 && h.version == 2 && h.num_scenes >= 1 &&
              h.num_scenes <= 4096 && h.n > 0 && h.n <= (1ll << 30) && h.precision >= MPMX_FP32 &&
              h.precision <= MPMX_MIXED_F16;
    std::vector<mpmx_params> ps;
    mpmx_config cfg;
    std::vector<int64_t> counts;
    if (ok) {
        ps.resize((size_t)h.num_scenes);
        counts.resize((size_t)h.num_scenes);
        ok = std::fread(ps.data(), sizeof(mpmx_params), ps.size(), f) == ps.size() &&
             std::fread(&cfg, sizeof(cfg), 1, f) == 1 &&
             std::fread(counts.data(), sizeof(int64_t), counts.size(), f) == counts.size();
    }
    // every scene holds particles and the scenes hold all of them
    int64_t total = 0;
    for (size_t k = 0; ok && k < counts.size(); ++k) {
        ok = counts[k] > 0 && counts[k] <= h.n;
        total += ok ? counts[k] : 0;
    }
    ok = ok && total == h.n && cfg.precision == h.precision;
    // the raw buffers, then a context sized by them
    const size_t n = ok ? (size_t)h.n : 0;
    const bool mixed = h.precision != MPMX_FP32;
    const size_t fbytes = h.precision == MPMX_MIXED_F16 ? 18 : 36;
    const size_t sizes[11] = {12 * n, (mixed ? 6 : 12) * n, (mixed ? 18 : 36) * n, fbytes * n,
                              (mixed ? 2 : 4) * n, (mixed ? 2 : 4) * n, (mixed ? 2 : 4) * n, (mixed ? 2 : 4) * n,
                              4 * n, 4 * n, 4 * n};
    size_t need = 0;
    for (size_t s : sizes) need += s;
    if (ok) ok = (size_t)std::ftell(f) + need == (size_t)fsize;
    if (!ok) {
        std::fclose(f);
        return MPMX_ERR_ARG;
    }
    cfg.device = device;
    cfg.max_pages = h.cap_tp;   // the capacity the run had
    cfg.max_particles = 0;
    std::vector<std::vector<char>> raw(11);
    for (int k = 0; k < 11 && ok; ++k) {
        raw[(size_t)k].resize(sizes[k]);
        ok = std::fread(raw[(size_t)k].data(), 1, sizes[k], f) == sizes[k];
    }
    std::fclose(f);
    if (!ok) return MPMX_ERR_ARG;
    mpmx_ctx* c = nullptr;
    mpmx_status s = mpmx_create_batch(h.num_scenes, ps.data(), &cfg, &c);
    if (s != MPMX_OK) return s;
    // upload order = the checkpoint's order with its ids: allocate through the
    // ordinary path from the stored positions, then overwrite every buffer
    std::vector<double> x(3 * n), v(3 * n, 0.0), m(n);
    const float* xf = reinterpret_cast<const float*>(raw[0].data());
    const float* mf = reinterpret_cast<const float*>(raw[8].data());
    const int* id = reinterpret_cast<const int*>(raw[9].data());
    for (size_t i = 0; i < n; ++i) {
        const size_t o = (size_t)id[i];
        if (o >= n) {
            mpmx_destroy(c);
            return MPMX_ERR_ARG;
        }
        for (int a = 0; a < 3; ++a) x[3 * o + a] = xf[3 * i + a];
        m[o] = mf[i];
    }
    s = mpmx_set_particles_batch(c, counts.data(), x.data(), v.data(), m.data());
    if (s != MPMX_OK) {
        mpmx_destroy(c);
        return s;
    }
    c->sorted = h.sorted != 0;   // which key buffer the saved keys go to
    auto bufs = particle_buffers(c);
    for (size_t k = 0; k < bufs.size(); ++k)
        if (bufs[k].second != raw[k].size() ||
            hipMemcpy(bufs[k].first, raw[k].data(), raw[k].size(), hipMemcpyHostToDevice) != hipSuccess) {
            mpmx_destroy(c);
            return MPMX_ERR_HIP;
        }
    c->sorted = h.sorted != 0;
    c->substeps = (int)h.substeps;
    c->cg_spec = (int)(h.cg_spec > 0 ? h.cg_spec : 1);
    *out = c;
    return MPMX_OK;
}

mpmx_status mpmx_load_checkpoint(const char* path, int device, mpmx_ctx** out) {
    if (!path || !out) return MPMX_ERR_ARG;
    *out = nullptr;
    try {
        return load_checkpoint(path, device, out);
    } catch (...) {   // host allocation failures stay on this side of the ABI
        return MPMX_ERR_ARG;
    }
}

// PLY (binary, little-endian) of every particle in upload order: position
// (each scene's own units), velocity, damage c and scene index
mpmx_status mpmx_write_particles(mpmx_ctx* c, const char* path) {
    if (!c || !path || c->n == 0) return MPMX_ERR_ARG;
    const size_t n = (size_t)c->n;
    std::vector<double> x(3 * n), v(3 * n), cc(n);
    mpmx_status s;
    if ((s = mpmx_get_field(c, MPMX_FIELD_X, x.data(), x.size())) != MPMX_OK) return s;
    if ((s = mpmx_get_field(c, MPMX_FIELD_V, v.data(), v.size())) != MPMX_OK) return s;
    if ((s = mpmx_get_field(c, MPMX_FIELD_C, cc.data(), cc.size())) != MPMX_OK) return s;
    FILE* f = std::fopen(path, "wb");
    if (!f) return io_fail(c, MPMX_ERR_ARG, "cannot open particle file for writing");
    std::fprintf(f,
                 "ply\nformat binary_little_endian 1.0\ncomment mpmx particles: position and velocity in scene "
                 "units, damage c, scene index\nelement vertex %zu\nproperty float x\nproperty float y\n"
                 "property float z\nproperty float vx\nproperty float vy\nproperty float vz\nproperty float c\n"
                 "property int scene\nend_header\n",
                 n);
    bool ok = true;
    int scene = 0;
    int64_t scene_end = c->scene_n[0];
    for (size_t i = 0; i < n && ok; ++i) {
        while ((int64_t)i >= scene_end && scene + 1 < c->num_scenes) scene_end += c->scene_n[(size_t)++scene];
        float rec[7] = {(float)x[3 * i], (float)x[3 * i + 1], (float)x[3 * i + 2], (float)v[3 * i],
                        (float)v[3 * i + 1], (float)v[3 * i + 2], (float)cc[i]};
        ok = std::fwrite(rec, sizeof(float), 7, f) == 7 && std::fwrite(&scene, sizeof(int), 1, f) == 1;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? MPMX_OK : io_fail(c, MPMX_ERR_HIP, "particle file write failed");
}

mpmx_status mpmx_get_tuning(mpmx_ctx* c, mpmx_tuning* out) {
    if (!c || !out) return MPMX_ERR_ARG;
    *out = mpmx_tuning{c->particle_wg, c->p2g_pct, c->g2p_pct, c->cg_pct};
    return MPMX_OK;
}

mpmx_status mpmx_set_tuning(mpmx_ctx* c, const mpmx_tuning* t) {
    if (!c || !t) return MPMX_ERR_ARG;
    if (t->particle_wg != 64 && t->particle_wg != 128 && t->particle_wg != 256) return MPMX_ERR_ARG;
    for (int pct : {t->p2g_pct, t->g2p_pct, t->cg_pct})
        if (pct < 10 || pct > 1000) return MPMX_ERR_ARG;
    destroy_graphs(c);
    c->particle_wg = t->particle_wg;
    c->p2g_pct = t->p2g_pct;
    c->g2p_pct = t->g2p_pct;
    c->cg_pct = t->cg_pct;
    return MPMX_OK;
}

mpmx_status mpmx_autotune(mpmx_ctx* c, double dt, int trial_steps, mpmx_tuning* best) {
    if (!c || c->n == 0 || trial_steps < 1) return MPMX_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return io_fail(c, MPMX_ERR_HIP, "hipSetDevice failed");
    Snapshot sn;
    if (snapshot_take(c, sn) != hipSuccess) {
        if (sn.mem) (void)hipFree(sn.mem);
        return io_fail(c, MPMX_ERR_HIP, "snapshot failed");
    }
    destroy_graphs(c);
    const int wgs[3] = {64, 128, 256}, pcts[3] = {50, 100, 200};
    // per candidate: ms of the kernel families over the trial
    double t_part[3], t_p2g[3], t_g2p[3], t_cg[3];
    mpmx_status st = MPMX_OK;
    std::vector<mpmx_kernel_time> kt(64);
    for (int k = 0; k < 3 && st == MPMX_OK; ++k) {
        c->particle_wg = wgs[k];
        c->p2g_pct = c->g2p_pct = c->cg_pct = pcts[k];
        if (snapshot_restore(c, sn) != hipSuccess) {
            st = io_fail(c, MPMX_ERR_HIP, "snapshot restore failed");
            break;
        }
        int ne = 0;
        if ((st = mpmx_timing_begin(c)) != MPMX_OK) break;
        if ((st = mpmx_run(c, trial_steps, dt)) != MPMX_OK) break;
        if ((st = mpmx_timing_end(c, kt.data(), (int)kt.size(), &ne)) != MPMX_OK) break;
        auto ms = [&](const char* name) {
            for (int q = 0; q < ne; ++q)
                if (std::strcmp(kt[(size_t)q].name, name) == 0) return kt[(size_t)q].total_ms;
            return 0.0;
        };
        // per-particle kernels: the fused stress + payload pass of a substep
        // (the stage-by-stage names cover contexts run stage by stage)
        t_part[k] = ms("stress_prep") + ms("stress") + ms("p2g_prep");
        t_p2g[k] = ms("p2g_block");
        t_g2p[k] = ms("g2p_block");
        t_cg[k] = ms("cg_col") + ms("cg_q");
    }
    // a family none of whose timers ran keeps its default
    auto argmin = [](const double* v, int dflt) {
        if (!(v[0] > 0.0) && !(v[1] > 0.0) && !(v[2] > 0.0)) return dflt;
        int b = 0;
        for (int k = 1; k < 3; ++k)
            if (v[k] < v[b]) b = k;
        return b;
    };
    if (st == MPMX_OK) {
        c->particle_wg = wgs[argmin(t_part, 2)];   // default 256 threads
        c->p2g_pct = pcts[argmin(t_p2g, 1)];       // default 100 %
        c->g2p_pct = pcts[argmin(t_g2p, 1)];
        c->cg_pct = pcts[argmin(t_cg, 1)];
    } else {
        c->particle_wg = 256;
        c->p2g_pct = c->g2p_pct = c->cg_pct = 100;
    }
    destroy_graphs(c);
    if (snapshot_restore(c, sn) != hipSuccess && st == MPMX_OK) st = io_fail(c, MPMX_ERR_HIP, "snapshot restore failed");
    (void)hipFree(sn.mem);
    if (st == MPMX_OK && best) mpmx_get_tuning(c, best);
    return st;
}

}  // extern "C"
