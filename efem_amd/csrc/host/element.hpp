// Element library: shape functions, local gradients, Gauss rules and
// pressure-face rules for C3D4, C3D10, C3D8 and C3D20 (full and reduced
// integration).  Mirrors the IntegralElement interface of
// src/fe/element/IntegralElement.h:35-61 (getPointShapeFunc,
// getPointShapeFuncLocalGradient, getGaussPointNum, getGaussPointWeight,
// getGaussPointLocalPosition).
#pragma once
#include <memory>
#include <string>
#include <vector>

namespace efem {

enum class ElementType { C3D4, C3D4R, C3D10, C3D10R, C3D8, C3D8R, C3D20, C3D20R };

class IntegralElement {
public:
    virtual ~IntegralElement() = default;
    ElementType type;
    virtual int nodeNum() const = 0;
    // N[nen] at a local position (getPointShapeFunc)
    virtual void pointShapeFunc(const double xi[3], double* N) const = 0;
    // dN/dxi [nen][3] at a local position (getPointShapeFuncLocalGradient)
    virtual void pointShapeFuncLocalGradient(const double xi[3], double* dN) const = 0;
    virtual int gaussPointNum() const = 0;
    // normalised weights: they sum to 1 and multiply the element volume
    virtual double gaussPointWeight(int q) const = 0;
    virtual void gaussPointLocalPosition(int q, double xi[3]) const = 0;
    bool isTet() const { return nodeNum() == 4 || nodeNum() == 10; }
};

// element by node count (Constraint.cpp:9-20) plus the integration choice
// the reference cannot select (its *R classes are unreachable)
std::unique_ptr<IntegralElement> makeElement(int nen, bool reduced);

// Element::volume restated (src/model/Element.h:25-89): divergence-theorem
// volume of the cell surface, quadratic faces tessellated through their
// mid-edge nodes.  Xe is [nen][3].
double elementVolume(int nen, const double* Xe);

// Gauss points of a pressure face in the parent element's parametric space
// (ExternalPressureFace::buildGaussPoint, src/fe/ExternalConstraint.cpp:289-827)
struct FaceRule {
    int nq = 0;
    double xi[9][3];
    double w[9];
    double A0[3];  // constant reference area vector, pointing into the parent
};
// face_nodes: fnn node ids; elem_nodes: parent connectivity; X: all nodes.
// Returns false (with msg) if the face is not a face of the parent.
bool buildFaceRule(int nen, const int* elem_nodes, int fnn, const int* face_nodes,
                   const double* X, FaceRule& out, std::string& msg);

}  // namespace efem
