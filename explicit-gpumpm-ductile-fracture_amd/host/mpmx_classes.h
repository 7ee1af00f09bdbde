// mpmx_classes.h — the reference's host-side C++ interface for the substep,
// kept by name and method so that a caller of the reference can switch over:
//   MPMSimulator            (Simulator.h:12-69)
//   MPMTimeIntegrator,
//   ExplicitTimeIntegrator  (TimeIntegrator.cuh:11-61)
//   Particles               (Particles.cuh:7-75)
//   SPGrid                  (SPGrid.h:3-36)
//   DomainTransformer       (Transformer.cuh:9-51)
//   Model                   (Models.h:9-31)
//   Material, ElasticMaterial (Materials.h:6-22)
//   SPGridMask              (SPGridMask.h:2-30)
// Every method forwards to the C ABI of libmpmx_hip.so (include/mpmx.h) on the
// one context the simulator owns.  Public members hold the device pointers and
// counts the reference exposes; the layouts are those of mpmx_device_ptrs
// (FP32 storage, so `T` of the device buffers is float), refreshed after every
// call that can move them.  Compile-time settings of the reference (Setting.h:
// N, dx, scheme, energy model) are runtime fields of mpmx_params here.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpmx.h"

using T = double;   // host-side scalar of the interface (Setting.h:4)

struct vector3T {   // Setting.h:40-44
    T x, y, z;
};

class mpmx_error : public std::runtime_error {
public:
    mpmx_error(mpmx_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    mpmx_status status;
};

// SPGridMask.h / SPGridMask.cpp: 64-bit SPGrid offsets of (i,j,k)
class SPGridMask {
public:
    SPGridMask(unsigned long long, unsigned long long) {}
    static unsigned long long Linear_Offset(unsigned long long* mask, const int i, const int j, const int k);
    void getMasks(unsigned long long* mk_x, unsigned long long* mk_y, unsigned long long* mk_z);
    void getMasks(unsigned long long* mk_a);
    void Cal_Masks() {}
    unsigned int getElement_per_Block() { return 64u; }
};

class Material {
public:
    Material() {}
    virtual ~Material() {}
    int getMaterialType() { return _materialType; }
protected:
    int _materialType = 0;
};

// Materials.cpp:4-14
class ElasticMaterial : public Material {
public:
    ElasticMaterial() = delete;
    ElasticMaterial(int materialType, T ym, T pr, T d, T vol);
    T _youngsModulus, _poissonRatio, _density;
    T _lambda, _mu, _volume;
    T _kappa;
    T parabolic_M;
};

class MPMSimulator;

class Particles {
public:
    Particles() {}
    // Particles.cpp:8-31: uploads mass/pos/vel, builds keys and the first sort
    void initialize(mpmx_ctx* ctx, unsigned int numParticle, const T* h_mass, const vector3T* h_pos,
                    const vector3T* h_vel);
    void reorder();                                   // Particle_kernal.cu:544-573
    void getMaxParticleVel_c(T* maxVelSquared);       // Particle_kernal.cu:421-440
    // Particles.cpp:33-52: positions, damage and original ids back to the host
    void Attribs_Back_Host(std::vector<vector3T>& h_pos, std::vector<vector3T>& h_vel,
                           std::vector<unsigned int>& h_indices);
    void refresh();                                   // re-read the device pointers

    unsigned int _numParticle = 0;
    // device buffers (current sorted order)
    float* d_pos = nullptr;
    float* d_orderedVel = nullptr;
    float* d_F = nullptr;
    float* d_B = nullptr;
    float* d_phase_C = nullptr;
    float* d_maxPsi = nullptr;
    float* d_g = nullptr;
    float* d_alpha = nullptr;
    float* d_FP = nullptr;
    float* d_vol = nullptr;
    float* d_mass = nullptr;
    int* d_indices = nullptr;        // original particle id of each slot
    uint32_t* cell_id = nullptr;     // Morton page key << 6 | cell
    std::vector<T> h_phase;          // damage of the last Attribs_Back_Host, original order
private:
    mpmx_ctx* _ctx = nullptr;
};

class Model {
public:
    Model() {}
    Model(std::unique_ptr<Particles> particle, std::unique_ptr<Material> material)
        : _particle(std::move(particle)), _material(std::move(material)) {}
    auto& getParticlesPtr() { return _particle; }
    auto& getMaterialPtr() { return _material; }
private:
    std::unique_ptr<Particles> _particle;
    std::unique_ptr<Material> _material;
};

class SPGrid {
public:
    SPGrid() {}
    void initialize(mpmx_ctx* ctx, int width, int height, int depth, T dc);
    void clear() {}   // the grid pages are zeroed inside transferP2G
    void refresh();
    int _width = 0, _height = 0, _depth = 0;
    T _dc = 0;
    float* d_grid = nullptr;     // [pages][8][64]
    float* d_grid_r = nullptr;
    float* d_grid_mr = nullptr;
    float* d_grid_s = nullptr;
    float* d_grid_q = nullptr;
    float* d_grid_x = nullptr;
private:
    mpmx_ctx* _ctx = nullptr;
};

class DomainTransformer {
public:
    DomainTransformer() {}
    void initialize(mpmx_ctx* ctx, const int numParticle, const int gridVolume);
    void rebuild();                // Transformer.cpp:71-82
    void refresh();
    int _numParticle = 0, _gridVolume = 0;
    int _numPage = 0, _numCell = 0, _numTotalPage = 0, _numVirtualPage = 0;
    uint32_t* d_pageOffset = nullptr;   // Morton page key per sparse page id
    int* d_adjPage = nullptr;           // [7][cap] neighbour page ids
    int* d_page2particle = nullptr;
    int* d_virtualPageOffset = nullptr;
    int* d_targetPage = nullptr;        // 0-based
    int64_t _adjStride = 0;
private:
    mpmx_ctx* _ctx = nullptr;
};

class MPMTimeIntegrator {
public:
    MPMTimeIntegrator(mpmx_ctx* ctx, int transferScheme, int numParticle)
        : _ctx(ctx), _transferScheme(transferScheme), _numParticle(numParticle) {}
    virtual ~MPMTimeIntegrator() {}
    virtual void integrate(int type, float* time1, float* time2, const T dt, Model& model,
                           std::unique_ptr<SPGrid>& grid, std::unique_ptr<DomainTransformer>& trans) = 0;
    void initMatrix(unsigned int, T*) {}   // F = I, B = 0 happen at upload
    void computeCellIndex(std::unique_ptr<DomainTransformer>&, std::unique_ptr<Particles>&) {}
    void transferP2G(const T dt, std::unique_ptr<Particles>& particles, std::unique_ptr<SPGrid>& grid,
                     std::unique_ptr<DomainTransformer>& trans);
    void undateGrid(const T dt, std::unique_ptr<SPGrid>& grid, std::unique_ptr<DomainTransformer>& trans);
    void resolveCollision(std::unique_ptr<SPGrid>&, std::unique_ptr<DomainTransformer>&) {}  // fused into undateGrid
    void PullGrid(std::unique_ptr<SPGrid>&, std::unique_ptr<DomainTransformer>&, T) {}        // grips: mpmx_params
    void transferG2P(const T dt, std::unique_ptr<Particles>& particles, std::unique_ptr<SPGrid>& grid,
                     std::unique_ptr<DomainTransformer>& trans);
protected:
    mpmx_ctx* _ctx;
    int _transferScheme;
    int _numParticle;
};

class ExplicitTimeIntegrator : public MPMTimeIntegrator {
public:
    ExplicitTimeIntegrator(mpmx_ctx* ctx, int transferScheme, int numParticle)
        : MPMTimeIntegrator(ctx, transferScheme, numParticle) {}
    // ExplicitTimeIntegrator.cu:881-995; time1 = simulation ms, time2 = preprocess ms
    void integrate(int type, float* time1, float* time2, const T dt, Model& model, std::unique_ptr<SPGrid>& grid,
                   std::unique_ptr<DomainTransformer>& trans) override;
    void computeForceCoefficient(Model& model);      // applyVonMises
    void computeParticlePhase_vol(Model&) {}         // V0 det F comes out of computeForceCoefficient
    void solvePhaseField(const T dt);                // the PCG block of integrate
};

class MPMSimulator {
public:
    MPMSimulator();
    ~MPMSimulator();
    // buildtype selects a seeded scene (mpmx_scene ids); the reference's
    // presets need SDF/OBJ assets absent from its checkout
    bool build(unsigned int buildtype, uint64_t seed = 1, double scale = 1.0);
    bool build(const mpmx_params& params, const std::vector<vector3T>& pos, const std::vector<vector3T>& vel,
               const std::vector<T>& mass, T dtDefault, const mpmx_config* cfg = nullptr);
    void buildModel(const int materialType, const T youngsModulus, const T poissonRatio, const T density,
                    const T volume, int idx = 0);
    void buildGrid(const int width, const int height, const int depth, const T dc);
    void buildTransformer(const int gridVolume);
    void buildIntegrator(const int integratorType, const int transferScheme, const T dtDefault);

    auto& getTimeIntegratorPtr() { return h_pTimeIntegrator; }
    auto& getGridPtr() { return h_pGrid; }
    auto& getTransformerPtr() { return h_pTrans; }
    Model& getModel(const unsigned int modelId = 0) { return h_models[modelId]; }
    auto& getModelParticlePtr(const unsigned int modelId = 0) { return getModel(modelId).getParticlesPtr(); }
    auto& getModelMaterialPtr(const unsigned int modelId = 0) { return getModel(modelId).getMaterialPtr(); }

    T computeDt(const T cur, const T next);                                       // Simulator.cpp:200-225
    void simulateStick(float* cflTime, float* preTime, float* simuTime, int type);  // Simulator.cpp:227-259
    void writePartio(const std::string& filename);   // declared in the reference (Simulator.h:35): binary dump
    // no counterpart in the reference: simulateStick's per-frame copy of the
    // particles to h_pos (on by default, as the reference does) and its CFL
    // step (computeDt on the device, off by default: the reference's viewer
    // loop uses the fixed dtDefault unless asked)
    void setReadback(bool on) { _readback = on; }
    void setUseCFL(bool on) { _useCFL = on; }

    std::vector<vector3T> h_pos;
    unsigned int numParticle = 0;
    mpmx_ctx* context() { return _ctx; }

private:
    unsigned int _currentFrame = 0;
    T _currentTime = 0;
    T _frameRate = 48;
    T _dtDefault = 1e-4;
    bool _useCFL = false;
    bool _readback = true;
    mpmx_params _params;
    mpmx_ctx* _ctx = nullptr;
    std::unique_ptr<SPGrid> h_pGrid;
    std::unique_ptr<DomainTransformer> h_pTrans;
    std::unique_ptr<MPMTimeIntegrator> h_pTimeIntegrator;
    std::vector<Model> h_models;
    std::vector<T> h_mass;
    std::vector<vector3T> h_vel;
    std::vector<unsigned int> h_indices;
};

void mpmx_check(mpmx_ctx* ctx, mpmx_status s, const char* where);
