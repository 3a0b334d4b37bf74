// Host-side preprocessing of a problem into the arrays the device kernels
// read.  Stands in for the reference's setup path: Constraint::Constraint
// (src/fe/Constraint.cpp:22-91), IntegralElement::initial
// (src/fe/element/IntegralElement.cpp:3-19), ElementConstraint ctor
// (src/fe/ElementConstraint.cpp:69-83), ExternalPressureFace::initial
// (src/fe/ExternalConstraint.cpp:255-287) and the constraint configs.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/efem_gpu.h"
#include "sparse.hpp"

namespace efem {

struct Model {
    int nNodes = 0, nElems = 0, nen = 0, ng = 0, law = 0, nFree = 0;
    int reduced = 0;
    int paramsPerElem = 0;
    std::vector<double> X;            // [n][3]
    std::vector<int> conn;            // [E][nen] node ids
    std::vector<int> fid;             // node -> free id or -1
    std::vector<int> connDof;         // [nen][E] free id, or nFree for a fixed node
    std::vector<double> vol;          // [E]
    std::vector<double> gw;           // [ng] normalised weights
    std::vector<double> gdN;          // [ng][nen][3] dN/dxi at the Gauss points
    std::vector<double> ginv;         // [ng][9][E] J0^-1 per Gauss point (SoA)
    std::vector<double> params;       // [4] or [E][4]
    std::vector<double> coeff;        // [E] elasticity_coeff
    // pressure faces
    int nFaces = 0, fnn = 0, fq = 0;  // fq: Gauss points per face
    double pressure = 0.0;
    int directed = 0;
    double pdir[3] = {0, 0, 0};
    std::vector<int> faceElem;        // [F]
    std::vector<double> fA0;          // [F][3]
    std::vector<double> fWN;          // [F][fq][nen] w_q N_a(xi_q)
    std::vector<double> fG;           // [F][fq][nen][3] grad_X N_a at xi_q
    // constant nodal forces per free DOF (ExternalForceConstraint)
    std::vector<double> fext;         // [3 nFree]
    // deterministic gather: free node -> slots (e*nen + a, then E*nen + f*nen + a)
    std::vector<int> gptr, gslot;

    int ndof() const { return 3 * nFree; }
};

bool buildModel(const efem_problem& pb, Model& m, std::string& msg);

// C3D4 element patches for a residual assembled in LDS.  Recursive
// coordinate bisection of the element centroids (longest axis, split at a
// multiple of P elements; ties by element id) groups the elements into
// patches of P (the last may hold fewer); within a patch the elements keep
// ascending id.  A patch's records are the free nodes its elements touch,
// by ascending free id; a record's entries are the patch's (lane, local
// node) pairs of that node as 4 lane + a, lane-major.  The entries of each
// record start on a multiple of 8 (padding: 0xffff).  Records are stored
// node-major: a free node's records, in patch order, at [rptr[i],
// rptr[i + 1]); rec[k] = {first entry, entry count, stored position, 0}.
// The face slots (f nen + a, faces ascending) of a free node are listed
// apart (fptr / fslot).
struct Tet4Patches {
    int P = 0, nPatch = 0, nRec = 0;
    std::vector<int> order;          // [E] patch position -> element
    std::vector<int> conn;           // [4][E] free ids, patch order
    std::vector<double> vol;         // [E] patch order
    std::vector<int> roff;           // [nPatch + 1] first record of a patch
    std::vector<int> rec;            // [nRec][4]
    std::vector<uint16_t> ent;       // entries
    std::vector<int> rptr;           // [nFree + 1] stored records of a node
    std::vector<int> fptr, fslot;    // free node -> face slots
};
void buildTet4Patches(const Model& m, int P, Tet4Patches& out);

// material parameters (layout of efem_problem.params) and the per-element
// approximate-Jacobian coefficient derived from them
bool setParams(Model& m, const double* params, std::string& msg);

// Concatenation of independent instances into one model (batched solves):
// nodes, free DOFs, elements and faces of part i follow those of part i-1;
// element slots keep their (element, node) order, face slots follow all
// element slots.  Parts must share the element kind, the integration rule
// and the law.  Per-face pressures go to facePressure; per-part ranges to
// ranges[i] = {elem0, elem1, face0, face1, free0, free1}.
struct PartRange {
    int e0, e1, f0, f1, n0, n1;
};
bool concatModels(const std::vector<Model>& parts, Model& out, std::vector<double>& facePressure,
                  std::vector<PartRange>& ranges, std::string& msg);

// scalar approximate Jacobian over free nodes (full symmetric CSC)
void scalarLaplacian(const Model& m, Csc& L);
// The pattern of scalarLaplacian (L.x zeroed) and, per entry k, its
// contributions src[ptr[k] .. ptr[k+1]) = e nen^2 + a nen + b (element e's
// row slot a, column slot b) in the order scalarLaplacian sums them: a
// device kernel that forms the element matrices and sums them through this
// map reproduces scalarLaplacian's values.
void laplacianMap(const Model& m, Csc& L, std::vector<int>& ptr, std::vector<int>& src);

// free-node coordinates [nFree][3]
std::vector<double> freeCoords(const Model& m);

// approximate Jacobian -> nested dissection -> supernodal factor (false: not
// SPD or no free DOF, msg says which)
bool factorModel(const Model& m, SupernodalFactor& F, std::string& msg,
                 int denseSubtree = kDefaultDenseSubtree, HostWorkspace* ws = nullptr,
                 SupernodalSymbolic* symbolic = nullptr);
// the same from an assembled approximate Jacobian A (ordering + factor).
// With `symbolic`: the ordering and symbolic analysis are computed into it
// when it is empty (n == 0) and reused as they are otherwise (A's pattern
// must be the one they were computed from)
bool factorLaplacian(const Model& m, const Csc& A, SupernodalFactor& F, std::string& msg,
                     int denseSubtree = kDefaultDenseSubtree, HostWorkspace* ws = nullptr,
                     SupernodalSymbolic* symbolic = nullptr);
// the same for independent parts, in parallel over the parts
bool factorModels(const std::vector<Model>& parts, std::vector<SupernodalFactor>& F, std::string& msg,
                  int denseSubtree = kDefaultDenseSubtree);

}  // namespace efem
