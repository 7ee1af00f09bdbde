// scene.cpp — seeded synthetic scenes (host only).
//
// The reference builds its scenes from an SDF of two dragons or an OBJ point
// list with rand() rejection sampling (MPMSimulator::build, Simulator.cpp:76-198,
// LoadSample.cpp:109-214); those assets are not part of the checkout.  These
// generators produce the benchmark configurations from a SplitMix64 stream:
//   0  cube: 20^3 cells at [54,74)^3 of a 128^3 grid, 8 particles per cell
//   1  notched bar: 1,048,576 uniform particles in a 192x96x96-cell bar of a
//      256^3 grid, V-notch at mid-length 1/4 of the height deep, grips 8 cells
//      from each end pulled at +-0.5 m/s, psi+ gated at stretch 1.05
//   2  notched bar, 8 particles per cell, 128x32x32 cells (1,048,576 particles)
//   3  two-block impact: two 80^3-cell blocks x 8 ppc in a 512^3 grid, 8-cell
//      gap along z, v_z = +-2.5 m/s
//   4  sphere of radius 16 cells at the centre of a 128^3 grid, 8 ppc
// and the batches (mpmx_scene_batch):
//   4  64 scenes, particle counts log-uniform in [10k, 500k]
//   5  256 scenes, particle counts log-uniform in [5k, 200k]
// A batch scene is a cube, a notched bar (4:1:1, grips at +-0.5 m/s) or a
// sphere (falling at 2 m/s), in turn, with E log-uniform in [1e4, 1e6] and
// yield stress E / 100; particles are uniform-random in the shape at a
// density of 1 to 8 per cell, and the grid (64^3, 128^3 or 256^3, uniform
// among those in which the shape spans at most 60 % of an axis) follows.
// Lattice scenes place 2x2x2 particles per cell at (0.25, 0.75) dx with a
// uniform jitter of +-0.125 dx.  `scale` shrinks the geometry (lattice
// scenes) or the particle count (scene 1).  Density follows
// rho = max(1000, (lambda + 2 mu) (dt / (0.3 dx))^2) so the elastic wave
// crosses at most 0.3 cells per substep.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/mpmx.h"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Scene {
    mpmx_params p;
    double dt;
    std::vector<double> x, v;
    double V0;
};

double density_rule(const mpmx_params& p, double dt) {
    const double lambda = p.E * p.nu / ((1 + p.nu) * (1 - 2 * p.nu));
    const double mu = p.E / (2 * (1 + p.nu));
    const double dx = 1.0 / p.N;
    const double r = dt / (0.3 * dx);
    const double rho = (lambda + 2 * mu) * r * r;
    return rho > 1000.0 ? rho : 1000.0;
}

// 8 jittered particles per cell for every cell accepted by `inside`
template <typename F>
void lattice(Scene& s, Rng& rng, int i0, int i1, int j0, int j1, int k0, int k1, F inside,
             double vx = 0, double vy = 0, double vz = 0) {
    const double dx = 1.0 / s.p.N;
    for (int i = i0; i < i1; ++i)
        for (int j = j0; j < j1; ++j)
            for (int k = k0; k < k1; ++k)
                for (int q = 0; q < 8; ++q) {
                    double pos[3];
                    const int cell[3] = {i, j, k};
                    for (int a = 0; a < 3; ++a) {
                        const double sub = ((q >> (2 - a)) & 1) ? 0.75 : 0.25;
                        pos[a] = (cell[a] + sub + (rng.u01() - 0.5) * 0.25) * dx;
                    }
                    if (!inside(pos)) continue;
                    s.x.insert(s.x.end(), pos, pos + 3);
                    s.v.push_back(vx);
                    s.v.push_back(vy);
                    s.v.push_back(vz);
                }
}

bool build(int scene, uint64_t seed, double scale, Scene& s) {
    Rng rng{seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};
    mpmx_default_params(&s.p);
    s.p.E = 1e5;
    s.p.nu = 0.3;
    s.p.yield_stress = 1e3;
    s.dt = 1e-4;
    if (scale <= 0) scale = 1;
    auto all = [](const double*) { return true; };
    switch (scene) {
    case 0: {
        s.p.N = 128;
        int e = (int)std::lround(20 * scale);
        if (e < 1) e = 1;
        const int a = 64 - e / 2;   // [54,74) for e = 20
        lattice(s, rng, a, a + e, a, a + e, a, a + e, all);
        s.V0 = std::pow(1.0 / s.p.N, 3) / 8;
        break;
    }
    case 1: {
        s.p.N = 256;
        const double x0 = 0.125, x1 = 0.875, y0 = 0.3125, y1 = 0.6875, z0 = 0.3125, z1 = 0.6875;
        const double depth = 0.25 * (y1 - y0);
        auto in_notch = [&](double x, double y) { return y > y1 - (depth - std::fabs(x - 0.5)); };
        int64_t n = (int64_t)std::llround(1048576.0 * scale);
        if (n < 1) n = 1;
        s.x.reserve(3 * n);
        while ((int64_t)s.x.size() < 3 * n) {
            const double px = x0 + (x1 - x0) * rng.u01();
            const double py = y0 + (y1 - y0) * rng.u01();
            const double pz = z0 + (z1 - z0) * rng.u01();
            if (in_notch(px, py)) continue;
            s.x.push_back(px);
            s.x.push_back(py);
            s.x.push_back(pz);
        }
        s.v.assign(s.x.size(), 0.0);
        const double vol = (x1 - x0) * (y1 - y0) * (z1 - z0) - depth * depth * (z1 - z0);
        s.V0 = vol / (double)n;
        s.p.grip_lo = 32 + 8;
        s.p.grip_hi = 224 - 8;
        s.p.grip_v = 0.5;
        s.p.crit_stretch = 1.05;
        break;
    }
    case 2: {
        s.p.N = 256;
        int L = (int)std::lround(128 * scale), H = (int)std::lround(32 * scale);
        if (L < 8) L = 8;
        if (H < 4) H = 4;
        const int i0 = 128 - L / 2, j0 = 128 - H / 2, k0 = 128 - H / 2;
        const double dx = 1.0 / s.p.N;
        const double ytop = (j0 + H) * dx, depth = 0.25 * H * dx, xm = 0.5;
        auto outside_notch = [&](const double* q) { return !(q[1] > ytop - (depth - std::fabs(q[0] - xm))); };
        lattice(s, rng, i0, i0 + L, j0, j0 + H, k0, k0 + H, outside_notch);
        s.V0 = std::pow(dx, 3) / 8;
        const int grip = L >= 32 ? 8 : L / 4;
        s.p.grip_lo = i0 + grip;
        s.p.grip_hi = i0 + L - grip;
        s.p.grip_v = 0.5;
        s.p.crit_stretch = 1.05;
        break;
    }
    case 3: {
        s.p.N = 512;
        int b = (int)std::lround(80 * scale);
        if (b < 1) b = 1;
        const int c0 = 256 - b / 2;
        lattice(s, rng, c0, c0 + b, c0, c0 + b, 256 - 4 - b, 256 - 4, all, 0, 0, 2.5);
        lattice(s, rng, c0, c0 + b, c0, c0 + b, 256 + 4, 256 + 4 + b, all, 0, 0, -2.5);
        s.V0 = std::pow(1.0 / s.p.N, 3) / 8;
        break;
    }
    case 4: {
        s.p.N = 128;
        const double r = 16 * scale / s.p.N;
        auto ball = [&](const double* q) {
            const double dx0 = q[0] - 0.5, dy = q[1] - 0.5, dz = q[2] - 0.5;
            return dx0 * dx0 + dy * dy + dz * dz <= r * r;
        };
        const int lo = 64 - (int)std::ceil(16 * scale) - 1, hi = 64 + (int)std::ceil(16 * scale) + 1;
        lattice(s, rng, lo, hi, lo, hi, lo, hi, ball);
        s.V0 = std::pow(1.0 / s.p.N, 3) / 8;
        break;
    }
    default:
        return false;
    }
    s.p.V0 = s.V0;
    return true;
}

// scene k of a batch: parameters and particles appended to x, v
void batch_scene(int config, uint64_t seed, int k, double scale, Scene& s) {
    Rng rng{(seed * 0x2545F4914F6CDD1Dull) ^ (0x9E3779B97F4A7C15ull * (uint64_t)(k + 1))};
    for (int q = 0; q < 4; ++q) rng.next();
    mpmx_default_params(&s.p);
    s.dt = 1e-4;
    const double pmin = config == 4 ? 1e4 : 5e3, pmax = config == 4 ? 5e5 : 2e5;
    int64_t P = (int64_t)std::llround(std::exp(std::log(pmin) + rng.u01() * (std::log(pmax) - std::log(pmin))) * scale);
    if (P < 16) P = 16;
    s.p.E = std::exp(std::log(1e4) + rng.u01() * (std::log(1e6) - std::log(1e4)));
    s.p.nu = 0.3;
    s.p.yield_stress = s.p.E / 100;
    const int shape = k % 3;
    const double ppc = 1.0 + 7.0 * rng.u01();
    const double vol_cells = (double)P / ppc;
    // linear size in cells: cube side, bar length (4:1:1), sphere diameter
    const double L = shape == 0 ? std::cbrt(vol_cells) : shape == 1 ? std::cbrt(16.0 * vol_cells)
                                                                    : std::cbrt(6.0 * vol_cells / M_PI);
    const int grids[3] = {64, 128, 256};
    int first = 0;
    while (first < 2 && L > 0.6 * grids[first]) ++first;
    const int N = grids[first + (int)(rng.next() % (uint64_t)(3 - first))];
    s.p.N = N;
    const double dx = 1.0 / N, c = 0.5;
    const double h = std::max(L, 4.0) * dx;   // size in domain units
    double lo[3], hi[3];
    std::function<bool(const double*)> inside;
    double v0[3] = {0, 0, 0};
    if (shape == 0) {
        for (int a = 0; a < 3; ++a) { lo[a] = c - h / 2; hi[a] = c + h / 2; }
        inside = [](const double*) { return true; };
    } else if (shape == 1) {
        const double w = h / 4;
        lo[0] = c - h / 2; hi[0] = c + h / 2;
        for (int a = 1; a < 3; ++a) { lo[a] = c - w / 2; hi[a] = c + w / 2; }
        const double ytop = hi[1], depth = 0.25 * w;
        inside = [=](const double* q) { return !(q[1] > ytop - (depth - std::fabs(q[0] - c))); };
        const int i0 = (int)std::floor(lo[0] * N), len = (int)std::ceil(h * N);
        const int grip = std::max(1, len / 8);
        s.p.grip_lo = i0 + grip;
        s.p.grip_hi = i0 + len - grip;
        s.p.grip_v = 0.5;
        s.p.crit_stretch = 1.05;
    } else {
        for (int a = 0; a < 3; ++a) { lo[a] = c - h / 2; hi[a] = c + h / 2; }
        const double r2 = h * h / 4;
        inside = [=](const double* q) {
            const double a0 = q[0] - c, a1 = q[1] - c, a2 = q[2] - c;
            return a0 * a0 + a1 * a1 + a2 * a2 <= r2;
        };
        v0[1] = -2.0;
    }
    int64_t got = 0;
    while (got < P) {
        double q[3];
        for (int a = 0; a < 3; ++a) q[a] = lo[a] + (hi[a] - lo[a]) * rng.u01();
        if (!inside(q)) continue;
        s.x.insert(s.x.end(), q, q + 3);
        s.v.insert(s.v.end(), v0, v0 + 3);
        ++got;
    }
    // shape volume by the acceptance of a fixed sample of the bounding box
    double box = 1;
    for (int a = 0; a < 3; ++a) box *= hi[a] - lo[a];
    int acc = 0;
    Rng vr{0x51ED270B27A5F3ull + (uint64_t)k};
    for (int t = 0; t < 20000; ++t) {
        double q[3];
        for (int a = 0; a < 3; ++a) q[a] = lo[a] + (hi[a] - lo[a]) * vr.u01();
        acc += inside(q) ? 1 : 0;
    }
    s.V0 = box * acc / 20000.0 / (double)P;
    s.p.V0 = s.V0;
}

}  // namespace

/* Batches of independent scenes (configs 4 and 5, described above).  Call with
 * null arrays to size them: *num_scenes (in: at most this many scenes when
 * > 0; out: the count) and the particle total (returned).  params[s],
 * counts[s]; x, v, mass hold the scenes one after another. */
extern "C" int64_t mpmx_scene_batch(int config, uint64_t seed, double scale, int* num_scenes, mpmx_params* params,
                                    int64_t* counts, double* x, double* v, double* mass, double* dt) {
    if ((config != 4 && config != 5) || !num_scenes) return -1;
    if (scale <= 0) scale = 1;
    int S = config == 4 ? 64 : 256;
    if (*num_scenes > 0 && *num_scenes < S) S = *num_scenes;
    *num_scenes = S;
    int64_t off = 0;
    for (int k = 0; k < S; ++k) {
        Scene s;
        batch_scene(config, seed, k, scale, s);
        const int64_t n = (int64_t)(s.x.size() / 3);
        if (params) params[k] = s.p;
        if (counts) counts[k] = n;
        if (x) std::memcpy(x + 3 * off, s.x.data(), sizeof(double) * 3 * n);
        if (v) std::memcpy(v + 3 * off, s.v.data(), sizeof(double) * 3 * n);
        if (mass) {
            const double rho = density_rule(s.p, s.dt);
            for (int64_t i = 0; i < n; ++i) mass[off + i] = rho * s.V0;
        }
        if (dt) *dt = s.dt;
        off += n;
    }
    return off;
}

extern "C" int64_t mpmx_scene(int scene, uint64_t seed, double scale, mpmx_params* p, double* x, double* v,
                              double* mass, double* dt) {
    Scene s;
    if (!build(scene, seed, scale, s)) return -1;
    const int64_t n = (int64_t)(s.x.size() / 3);
    const double rho = density_rule(s.p, s.dt);
    if (p) *p = s.p;
    if (dt) *dt = s.dt;
    if (x) std::memcpy(x, s.x.data(), sizeof(double) * 3 * n);
    if (v) std::memcpy(v, s.v.data(), sizeof(double) * 3 * n);
    if (mass)
        for (int64_t i = 0; i < n; ++i) mass[i] = rho * s.V0;
    return n;
}

// ---- scene files ----------------------------------------------------------
// Point-list OBJ (Sample::readObj, LoadSample.cpp:163-194): every line
// "v x y z" is one particle; the reference's loader moves it down by 0.2 in
// y.  Other lines (comments, faces) are skipped.
extern "C" int64_t mpmx_read_obj_points(const char* path, double y_shift, double* x, int64_t cap) {
    if (!path) return -1;
    FILE* f = std::fopen(path, "r");
    if (!f) return -1;
    char line[512];
    int64_t n = 0;
    while (std::fgets(line, sizeof(line), f)) {
        double a, b, c;
        if (line[0] != 'v' || (line[1] != ' ' && line[1] != '\t')) continue;
        if (std::sscanf(line + 1, "%lf %lf %lf", &a, &b, &c) != 3) continue;
        if (x && n < cap) {
            x[3 * n] = a;
            x[3 * n + 1] = b + y_shift;
            x[3 * n + 2] = c;
        }
        ++n;
    }
    std::fclose(f);
    return n;
}

namespace {

// level set on an ni x nj x nk lattice, i fastest (Sample::LoadSDF,
// LoadSample.cpp:3-32: "ni nj nk", "minx miny minz", "dx", then the values)
struct LevelSet {
    int ni = 0, nj = 0, nk = 0;
    double min[3] = {0, 0, 0}, dx = 0;
    std::vector<double> phi;
    double at(int i, int j, int k) const { return phi[(size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * k)]; }
    // trilinear value at lattice coordinates (Sample::fetchGridTrilinear, :39-60)
    double trilinear(double x, double y, double z) const {
        const int x0 = (int)std::floor(x), y0 = (int)std::floor(y), z0 = (int)std::floor(z);
        const int x1 = (int)std::ceil(x), y1 = (int)std::ceil(y), z1 = (int)std::ceil(z);
        const double fx = x - x0, fy = y - y0, fz = z - z0;
        const double c00 = at(x0, y0, z0) * (1 - fx) + at(x1, y0, z0) * fx;
        const double c01 = at(x0, y0, z1) * (1 - fx) + at(x1, y0, z1) * fx;
        const double c10 = at(x0, y1, z0) * (1 - fx) + at(x1, y1, z0) * fx;
        const double c11 = at(x0, y1, z1) * (1 - fx) + at(x1, y1, z1) * fx;
        const double c0 = c00 * (1 - fy) + c10 * fy, c1 = c01 * (1 - fy) + c11 * fy;
        return c0 * (1 - fz) + c1 * fz;
    }
};

bool load_level_set(const char* path, LevelSet& ls) {
    FILE* f = std::fopen(path, "r");
    if (!f) return false;
    bool ok = std::fscanf(f, "%d %d %d", &ls.ni, &ls.nj, &ls.nk) == 3 &&
              std::fscanf(f, "%lf %lf %lf", &ls.min[0], &ls.min[1], &ls.min[2]) == 3 &&
              std::fscanf(f, "%lf", &ls.dx) == 1 && ls.ni >= 4 && ls.nj >= 4 && ls.nk >= 4;
    if (ok) {
        ls.phi.resize((size_t)ls.ni * ls.nj * ls.nk);
        for (auto& v : ls.phi)
            if (std::fscanf(f, "%lf", &v) != 1) {
                ok = false;
                break;
            }
    }
    std::fclose(f);
    return ok;
}

}  // namespace

// Sample::Initial_data + GenerateUniformSamples (LoadSample.cpp:62-160): the
// level set's lattice (less one layer each side) is scaled uniformly into the
// box res (cells) placed with its x and z centred on `center` and its low y at
// center_y + res_y / 4, four cells in; samples_per_cell particles per grid cell
// (the reference uses 20), i.e. validCells x samples_per_cell scale^3 in all,
// where a valid lattice cell has a corner inside (phi < 0).  Samples are drawn
// uniformly over the lattice and kept where the trilinear phi < 0, from a
// SplitMix64 stream instead of rand().  Positions in domain units (dx = 1/N).
// Returns the count (x may be null to size it) or -1.
extern "C" int64_t mpmx_sample_sdf(const char* path, uint64_t seed, int N, const unsigned* center,
                                   const unsigned* res, double samples_per_cell, double* x, int64_t cap) {
    if (!path || !center || !res || N <= 0) return -1;
    LevelSet ls;
    if (!load_level_set(path, ls)) return -1;
    int offset[3];
    for (int a = 0; a < 3; ++a) offset[a] = (int)(center[a] - 0.5 * res[a]);
    offset[1] = (int)(center[1] + 0.25 * res[1]);
    const int lo = 1, hi[3] = {ls.ni - 2, ls.nj - 2, ls.nk - 2};
    double scale = 1e300;
    for (int a = 0; a < 3; ++a) scale = std::min(scale, (double)res[a] / (double)(hi[a] - lo));
    const double per_lattice_cell = samples_per_cell * scale * scale * scale;
    int64_t valid = 0;
    for (int i = 0; i < ls.ni - 1; ++i)
        for (int j = 0; j < ls.nj - 1; ++j)
            for (int k = 0; k < ls.nk - 1; ++k) {
                bool in = false;
                for (int q = 0; q < 8 && !in; ++q) in = ls.at(i + (q >> 2), j + ((q >> 1) & 1), k + (q & 1)) < 0;
                valid += in ? 1 : 0;
            }
    const int64_t count = (int64_t)((double)valid * per_lattice_cell);
    if (!x) return count;
    Rng rng{seed * 0x2545F4914F6CDD1Dull + 0x5DF5ull};
    const double dxN = 1.0 / N;
    for (int64_t s = 0; s < count && s < cap; ++s) {
        double p[3];
        do {
            p[0] = rng.u01() * (ls.ni - 1);
            p[1] = rng.u01() * (ls.nj - 1);
            p[2] = rng.u01() * (ls.nk - 1);
        } while (ls.trilinear(p[0], p[1], p[2]) >= 0);
        for (int a = 0; a < 3; ++a) x[3 * s + a] = ((p[a] - lo) * scale + offset[a] + 4) * dxN;
    }
    return count;
}
