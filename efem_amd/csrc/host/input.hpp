// Reference input formats: an EFEM run directory (JSON config + mesh and
// constraint files) read into an efem_problem and solver parameters.
//   mesh file     MeshParser::parse (src/io/MeshParser.h:22-100)
//   element rows  Body::buildElement (src/model/Body.cpp:49-90)
//   fix / force / pressure files   ExternalFixConfig / ExternalForceConfig /
//                 ExternalPressureConfig::configConstraint and IO::readFile
//                 (src/fe/ExternalConstraint.cpp:50-247, src/io/IO.cpp:67-96)
//   materials     ElementConstraintConfig::loadMaterials / configConstraint
//                 (src/fe/ElementConstraint.cpp:11-67), Material::initFromJson
//   analysis      Config::Config (src/io/Config.cpp:23-81)
#pragma once
#include <string>
#include <vector>

#include "../../../include/efem_gpu.h"

namespace efem {

struct InputDeck {
    std::vector<double> X;           // [n][3]
    std::vector<int> conn;           // [E][nen] 0-based (elements with a material)
    int nen = 0;
    int law = 0;
    std::vector<double> params;      // [4] or [E][4]
    int paramsPerElem = 0;
    std::vector<unsigned char> fixed;
    std::vector<int> faces;          // [F][fnn] 0-based
    int fnn = 0;
    double pressure = 0.0;
    int directed = 0;
    double pdir[3] = {0, 0, 0};
    std::vector<int> fnodes;         // constant-force nodes
    double fvec[3] = {0, 0, 0};
    int skipped = 0;                 // elements without a defined material
    // analysis block
    int usePrecond = 1;              // nl_solver "lbfgs_ah" (default) / "lbfgs"
    int maxSolverIters = 0;
    int maxLineSearch = 20;
    double dxCriterion = 1e-8;
    double residualCriterion = 1e-6;
    // gravity (analysis.gravity "non" / "x" / "-x" / ... -> 0 .. 6) and the
    // mass matrix it uses (analysis.consistent_inertia); material densities
    // indexed as params
    int gravity = 0;
    int consistentInertia = 0;
    std::vector<double> density;
};

// false with msg on any error
bool loadInput(const std::string& jsonPath, InputDeck& d, std::string& msg);

}  // namespace efem
