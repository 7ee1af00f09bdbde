// mpmx_classes.cpp — the reference's C++ classes over the C ABI of
// libmpmx_hip.so.  See mpmx_classes.h for the mapping.
#include "mpmx_classes.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

void mpmx_check(mpmx_ctx* ctx, mpmx_status s, const char* where) {
    if (s == MPMX_OK) return;
    std::string msg = std::string(where) + ": " + mpmx_status_string(s);
    if (ctx) msg += std::string(" (") + mpmx_last_error(ctx) + ")";
    throw mpmx_error(s, msg);
}
#define CHK(ctx, call) mpmx_check(ctx, (call), #call)

// ---------------------------------------------------------------- SPGridMask
static unsigned long long pdep64(unsigned long long v, unsigned long long mask) {
    unsigned long long out = 0;
    for (int b = 0; b < 64; ++b)
        if ((mask >> b) & 1ull) {
            out |= (v & 1ull) << b;
            v >>= 1;
        }
    return out;
}
static const unsigned long long kMasks[3] = {0x49249249249240c0ull, 0x2492492492492030ull, 0x924924924924900cull};

unsigned long long SPGridMask::Linear_Offset(unsigned long long* mask, const int i, const int j, const int k) {
    return pdep64((unsigned)i, mask[0]) | pdep64((unsigned)j, mask[1]) | pdep64((unsigned)k, mask[2]);
}
void SPGridMask::getMasks(unsigned long long* x, unsigned long long* y, unsigned long long* z) {
    *x = kMasks[0];
    *y = kMasks[1];
    *z = kMasks[2];
}
void SPGridMask::getMasks(unsigned long long* a) { getMasks(a, a + 1, a + 2); }

// ---------------------------------------------------------------- materials
ElasticMaterial::ElasticMaterial(int materialType, T ym, T pr, T d, T vol) {
    _materialType = materialType;
    _youngsModulus = ym;
    _poissonRatio = pr;
    _density = d;
    _volume = vol;
    _lambda = ym * pr / ((1 + pr) * (1 - 2 * pr));
    _mu = ym / (2 * (1 + pr));
    _kappa = ym / (3 * (1 - 2 * pr));
    parabolic_M = 50;
}

// ---------------------------------------------------------------- Particles
void Particles::initialize(mpmx_ctx* ctx, unsigned int numParticle, const T* h_mass, const vector3T* h_pos,
                           const vector3T* h_vel) {
    _ctx = ctx;
    _numParticle = numParticle;
    CHK(ctx, mpmx_set_particles(ctx, numParticle, &h_pos[0].x, &h_vel[0].x, h_mass));
    refresh();
}

void Particles::refresh() {
    mpmx_device_ptrs p;
    CHK(_ctx, mpmx_get_device_pointers(_ctx, &p));
    // the float members point at float buffers only: in the mixed forms the
    // BF16 / FP16 fields are reached through mpmx_get_device_pointers
    const bool f32 = p.precision == MPMX_FP32;
    d_pos = (float*)p.pos;
    d_orderedVel = f32 ? (float*)p.vel : nullptr;
    d_F = p.precision != MPMX_MIXED_F16 ? (float*)p.F : nullptr;
    d_B = f32 ? (float*)p.B : nullptr;
    d_phase_C = f32 ? (float*)p.phase_c : nullptr;
    d_maxPsi = f32 ? (float*)p.psi : nullptr;
    d_g = f32 ? (float*)p.g : nullptr;
    d_alpha = f32 ? (float*)p.alpha : nullptr;
    d_FP = p.FP;
    d_vol = p.vol;
    d_mass = (float*)p.mass;
    d_indices = p.ids;
    cell_id = p.keys;
}

void Particles::reorder() {
    CHK(_ctx, mpmx_bin_sort(_ctx));
    refresh();
}

void Particles::getMaxParticleVel_c(T* maxVelSquared) { CHK(_ctx, mpmx_max_velocity_sq(_ctx, maxVelSquared)); }

void Particles::Attribs_Back_Host(std::vector<vector3T>& h_pos, std::vector<vector3T>& h_vel,
                                  std::vector<unsigned int>& h_indices) {
    const size_t n = _numParticle;
    h_pos.resize(n);
    h_vel.resize(n);
    h_indices.resize(n);
    CHK(_ctx, mpmx_get_field(_ctx, MPMX_FIELD_X, &h_pos[0].x, 3 * n));
    CHK(_ctx, mpmx_get_field(_ctx, MPMX_FIELD_V, &h_vel[0].x, 3 * n));
    h_phase.resize(n);
    CHK(_ctx, mpmx_get_field(_ctx, MPMX_FIELD_C, h_phase.data(), n));
    std::vector<int32_t> order(n);
    CHK(_ctx, mpmx_get_order(_ctx, order.data(), nullptr));
    for (size_t i = 0; i < n; ++i) h_indices[i] = (unsigned)order[i];
}

// ---------------------------------------------------------------- SPGrid
void SPGrid::initialize(mpmx_ctx* ctx, int width, int height, int depth, T dc) {
    _ctx = ctx;
    _width = width;
    _height = height;
    _depth = depth;
    _dc = dc;
    refresh();
}
void SPGrid::refresh() {
    mpmx_device_ptrs p;
    CHK(_ctx, mpmx_get_device_pointers(_ctx, &p));
    d_grid = p.grid;
    d_grid_r = p.pcg_r;
    d_grid_mr = p.pcg_z;
    d_grid_s = p.pcg_s;
    d_grid_q = p.pcg_q;
    d_grid_x = p.pcg_x;
}

// ---------------------------------------------------------------- DomainTransformer
void DomainTransformer::initialize(mpmx_ctx* ctx, const int numParticle, const int gridVolume) {
    _ctx = ctx;
    _numParticle = numParticle;
    _gridVolume = gridVolume;
    refresh();
}
void DomainTransformer::rebuild() {
    CHK(_ctx, mpmx_rebuild(_ctx));
    refresh();
}
void DomainTransformer::refresh() {
    mpmx_device_ptrs p;
    CHK(_ctx, mpmx_get_device_pointers(_ctx, &p));
    d_pageOffset = p.page_key;
    d_adjPage = p.adj;
    d_page2particle = p.page2particle;
    d_virtualPageOffset = p.vpo;
    d_targetPage = p.target;
    _adjStride = p.cap_pages;
    mpmx_stats st;
    CHK(_ctx, mpmx_get_stats(_ctx, &st));
    _numPage = st.num_page;
    _numTotalPage = st.num_total_page;
    _numVirtualPage = st.num_virtual_page;
}

// ---------------------------------------------------------------- integrators
void MPMTimeIntegrator::transferP2G(const T dt, std::unique_ptr<Particles>&, std::unique_ptr<SPGrid>&,
                                    std::unique_ptr<DomainTransformer>&) {
    CHK(_ctx, mpmx_p2g(_ctx, dt));
}
void MPMTimeIntegrator::undateGrid(const T dt, std::unique_ptr<SPGrid>&, std::unique_ptr<DomainTransformer>&) {
    CHK(_ctx, mpmx_grid_update(_ctx, dt));
}
void MPMTimeIntegrator::transferG2P(const T dt, std::unique_ptr<Particles>&, std::unique_ptr<SPGrid>&,
                                    std::unique_ptr<DomainTransformer>&) {
    CHK(_ctx, mpmx_g2p(_ctx, dt));
}
void ExplicitTimeIntegrator::computeForceCoefficient(Model&) { CHK(_ctx, mpmx_stress(_ctx)); }
void ExplicitTimeIntegrator::solvePhaseField(const T dt) { CHK(_ctx, mpmx_phase_field(_ctx, dt)); }

void ExplicitTimeIntegrator::integrate(int, float* simulationTime, float* preprocessTime, const T dt,
                                       Model& model, std::unique_ptr<SPGrid>& grid,
                                       std::unique_ptr<DomainTransformer>& trans) {
    using clock = std::chrono::steady_clock;
    const bool timed = simulationTime || preprocessTime;
    auto& particles = model.getParticlesPtr();
    if (timed) CHK(_ctx, mpmx_sync(_ctx));
    const auto t0 = clock::now();
    particles->reorder();
    trans->rebuild();
    if (timed) CHK(_ctx, mpmx_sync(_ctx));
    const auto t1 = clock::now();
    computeForceCoefficient(model);
    transferP2G(dt, particles, grid, trans);
    undateGrid(dt, grid, trans);
    solvePhaseField(dt);
    transferG2P(dt, particles, grid, trans);
    if (timed) CHK(_ctx, mpmx_sync(_ctx));
    const auto t2 = clock::now();
    if (preprocessTime) *preprocessTime = std::chrono::duration<float, std::milli>(t1 - t0).count();
    if (simulationTime) *simulationTime = std::chrono::duration<float, std::milli>(t2 - t1).count();
}

// ---------------------------------------------------------------- MPMSimulator
MPMSimulator::MPMSimulator() { mpmx_default_params(&_params); }

MPMSimulator::~MPMSimulator() {
    h_models.clear();
    h_pGrid.reset();
    h_pTrans.reset();
    h_pTimeIntegrator.reset();
    if (_ctx) mpmx_destroy(_ctx);
}

bool MPMSimulator::build(unsigned int buildtype, uint64_t seed, double scale) {
    mpmx_params p;
    double dt = 0;
    const int64_t n = mpmx_scene((int)buildtype, seed, scale, &p, nullptr, nullptr, nullptr, &dt);
    if (n <= 0) return false;
    std::vector<vector3T> pos((size_t)n), vel((size_t)n);
    std::vector<T> mass((size_t)n);
    mpmx_scene((int)buildtype, seed, scale, &p, &pos[0].x, &vel[0].x, mass.data(), &dt);
    return build(p, pos, vel, mass, dt);
}

bool MPMSimulator::build(const mpmx_params& params, const std::vector<vector3T>& pos,
                         const std::vector<vector3T>& vel, const std::vector<T>& mass, T dtDefault,
                         const mpmx_config* cfg) {
    _params = params;
    h_pos = pos;
    h_vel = vel;
    h_mass = mass;
    numParticle = (unsigned)pos.size();
    CHK(nullptr, mpmx_create(&_params, cfg, &_ctx));
    const double density = mass.empty() ? 0.0 : mass[0] / params.V0;
    buildModel(1, params.E, params.nu, density, params.V0, 0);
    buildGrid(params.N, params.N, params.N, 1.0 / params.N);
    buildTransformer(params.N * params.N * params.N);
    buildIntegrator(0, params.transfer, dtDefault);
    return true;
}

void MPMSimulator::buildModel(const int materialType, const T youngsModulus, const T poissonRatio, const T density,
                              const T volume, int) {
    auto particles = std::make_unique<Particles>();
    particles->initialize(_ctx, numParticle, h_mass.data(), h_pos.data(), h_vel.data());
    auto material = std::make_unique<ElasticMaterial>(materialType, youngsModulus, poissonRatio, density, volume);
    h_models.emplace_back(std::move(particles), std::move(material));
}

void MPMSimulator::buildGrid(const int width, const int height, const int depth, const T dc) {
    h_pGrid = std::make_unique<SPGrid>();
    h_pGrid->initialize(_ctx, width, height, depth, dc);
}

void MPMSimulator::buildTransformer(const int gridVolume) {
    h_pTrans = std::make_unique<DomainTransformer>();
    h_pTrans->initialize(_ctx, (int)numParticle, gridVolume);
}

void MPMSimulator::buildIntegrator(const int, const int transferScheme, const T dtDefault) {
    h_pTimeIntegrator = std::make_unique<ExplicitTimeIntegrator>(_ctx, transferScheme, (int)numParticle);
    _dtDefault = dtDefault;
}

T MPMSimulator::computeDt(const T, const T) {
    if (!_useCFL) return _dtDefault;
    double dt = _dtDefault;
    CHK(_ctx, mpmx_compute_dt(_ctx, _dtDefault, 0.3, &dt));
    return dt;
}

void MPMSimulator::simulateStick(float* cflTime, float* preTime, float* simuTime, int type) {
    const auto t0 = std::chrono::steady_clock::now();
    const T dt = computeDt(_currentTime, _currentFrame / _frameRate);
    if (cflTime) *cflTime = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (dt > 0) getTimeIntegratorPtr()->integrate(type, simuTime, preTime, dt, getModel(0), h_pGrid, h_pTrans);
    _currentTime += dt;
    ++_currentFrame;
    h_pGrid->refresh();
    // the reference copies the particles back every frame for its viewer
    // (Simulator.cpp:253); a caller without a viewer turns that off
    if (_readback) getModelParticlePtr()->Attribs_Back_Host(h_pos, h_vel, h_indices);
}

void MPMSimulator::writePartio(const std::string& filename) {
    const size_t n = numParticle;
    std::vector<double> x(3 * n), v(3 * n), c(n), a(n);
    CHK(_ctx, mpmx_get_field(_ctx, MPMX_FIELD_X, x.data(), x.size()));
    CHK(_ctx, mpmx_get_field(_ctx, MPMX_FIELD_V, v.data(), v.size()));
    CHK(_ctx, mpmx_get_field(_ctx, MPMX_FIELD_C, c.data(), c.size()));
    CHK(_ctx, mpmx_get_field(_ctx, MPMX_FIELD_ALPHA, a.data(), a.size()));
    FILE* f = std::fopen(filename.c_str(), "wb");
    if (!f) throw mpmx_error(MPMX_ERR_ARG, "cannot open " + filename);
    const char magic[8] = {'M', 'P', 'M', 'X', 'P', 'R', 'T', '1'};
    const uint64_t count = n;
    std::fwrite(magic, 1, 8, f);
    std::fwrite(&count, 8, 1, f);
    std::fwrite(x.data(), 8, x.size(), f);
    std::fwrite(v.data(), 8, v.size(), f);
    std::fwrite(c.data(), 8, c.size(), f);
    std::fwrite(a.data(), 8, a.size(), f);
    std::fclose(f);
}
