// ref_materials.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-ABI harness around the reference's own constitutive sources
// (src/material/{NeoHookean,MooneyRivlin,Yeoh}.cpp), compiled where they lie
// by oracle/Makefile into oracle/_ref/.  Tests use it to pin or_piola and
// or_elasticity_coeff of the oracle; nothing in the product loads it.
#include "material/Material.h"

namespace {
EFEM::Material* make(int law, const double* p)
{
    Json j;
    j["id"] = 1;
    j["density"] = 1000.0;
    j["C10"] = p[0];
    j["D1"] = p[3];
    if (law == 0) {
        auto* m = new EFEM::NeoHookean();
        m->initFromJson(j);
        return m;
    }
    if (law == 1) {
        j["C01"] = p[1];
        auto* m = new EFEM::MooneyRivlin();
        m->initFromJson(j);
        return m;
    }
    j["C20"] = p[1];
    j["C30"] = p[2];
    auto* m = new EFEM::Yeoh();
    m->initFromJson(j);
    return m;
}
}

extern "C" void ref_piola(int law, const double* p, const double* F, double* P)
{
    EFEM::Material* m = make(law, p);
    Eigen::Matrix3d Fm;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++)
            Fm(i, k) = F[3 * i + k];
    Eigen::Matrix3d Pm = m->getFirstPiola(Fm);
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++)
            P[3 * i + k] = Pm(i, k);
    delete m;
}

extern "C" double ref_elasticity_coeff(int law, const double* p)
{
    EFEM::Material* m = make(law, p);
    double c = m->elasticity_coeff;
    delete m;
    return c;
}
