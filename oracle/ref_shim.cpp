// oracle/_ref shim (test infrastructure): exposes the reference's own
// CPU-compilable SPGridMask (SPGridMask.cpp) and ElasticMaterial
// (Materials.cpp) through a C ABI so the oracle's masks, Morton offsets and
// Lame parameters can be checked against the reference code itself.
#include "SPGridMask.h"
#include "Materials.h"

extern "C" {

void ref_masks(unsigned long long m[3]) {
    SPGridMask g(6, 3);
    g.Cal_Masks();
    g.getMasks(m);
}

unsigned long long ref_linear_offset(int i, int j, int k) {
    unsigned long long m[3];
    SPGridMask g(6, 3);
    g.Cal_Masks();
    g.getMasks(m);
    return SPGridMask::Linear_Offset(m, i, j, k);
}

void ref_material(double E, double nu, double out[4]) {
    ElasticMaterial mat(2, E, nu, 1.0, 1.0);
    out[0] = mat._lambda;
    out[1] = mat._mu;
    out[2] = mat._kappa;
    out[3] = mat.parabolic_M;
}

}
