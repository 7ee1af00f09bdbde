// demo_main.cpp — headless driver of the C++ drop-in classes, the way the
// reference's viewer drives them (gl_main.cpp:87-94 init, :304-327 one
// simulateStick per frame), minus the OpenGL window.
//
//   mpmx_demo <scene> <frames> <seed> <scale> <out.bin> [cfl]
// writes a writePartio dump of the final state.  With "cfl" every frame takes
// the CFL step of computeDt (0.3 dx / max|v|, at most the scene's dt) and the
// per-frame copy of the particles to the host is off.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "mpmx_classes.h"

int main(int argc, char** argv) {
    const int scene = argc > 1 ? std::atoi(argv[1]) : 0;
    const int frames = argc > 2 ? std::atoi(argv[2]) : 10;
    const unsigned long long seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 1ull;
    const double scale = argc > 4 ? std::atof(argv[4]) : 1.0;
    const char* out = argc > 5 ? argv[5] : "mpmx_demo.bin";
    const bool cfl = argc > 6 && std::string(argv[6]) == "cfl";
    try {
        MPMSimulator sim;
        if (!sim.build((unsigned)scene, seed, scale)) {
            std::fprintf(stderr, "unknown scene %d\n", scene);
            return 2;
        }
        if (cfl) {
            sim.setUseCFL(true);
            sim.setReadback(false);
        }
        float cfl = 0, pre = 0, simu = 0;
        for (int f = 0; f < frames; ++f) {
            sim.simulateStick(&cfl, &pre, &simu, 0);
            if (f == 0 || f == frames - 1)
                std::printf("frame %d: %u particles, pages %d/%d, preprocess %.3f ms, simulation %.3f ms\n", f,
                            sim.numParticle, sim.getTransformerPtr()->_numPage,
                            sim.getTransformerPtr()->_numTotalPage, pre, simu);
        }
        sim.writePartio(out);
    } catch (const mpmx_error& e) {
        std::fprintf(stderr, "mpmx error: %s\n", e.what());
        return 1;
    }
    return 0;
}
