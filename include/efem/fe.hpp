// efem/fe.hpp -- the reference's C++ host interfaces of the FE layer, over
// the C ABI of efem_gpu.h: materials, the Constrainable family, the
// Constraint aggregate and the Integrator's solve.  Host code written
// against the reference's classes keeps their names, virtuals and argument
// meanings; every residual, factorisation and solve runs on the GPU.
//
//   Material, NeoHookean, MooneyRivlin, Yeoh     src/material/Material.h:9-51
//   Constrainable (conInitial, conEvaluteResidual, conGetResidual,
//                  conEvaluteJacobi, conGetJacobi)  src/fe/Constraint.h:23-31
//   ElementConstraint                              src/fe/ElementConstraint.h:25-45
//   ExternalFixConstraint / ExternalForceConstraint /
//   ExternalPressureConstraint                     src/fe/ExternalConstraint.h
//   InertiaConstraint                              src/fe/InertiaConstraint.h
//   Constraint (assignDx, updateDx, getResidual,
//               getJacobi, getDimens)              src/fe/Constraint.h:58-77
//   Constraint::integral                           Integrator::integral
//                                                  (src/simulation/Integrator.cpp:22-42)
//
// Differences forced by the device, each refused with std::invalid_argument
// at Constraint construction rather than approximated:
//   - a material is a device law (deviceLaw() >= 0: NeoHookean, MooneyRivlin,
//     Yeoh); a subclass without one has no kernel to run on;
//   - all elements run one law (per-element parameters are free);
//   - at most one follower-pressure (or directed-traction) constraint and
//     one nodal-force vector (nodes of several force constraints may differ,
//     their vectors may not).
// Vectors are free-DOF vectors as in the reference: entry 3 id + c, id the
// node's rank among free nodes (Constraint.cpp:61-72).  Eigen types become
// std::array / std::vector; Json (nlohmann in the reference) becomes the
// numeric fields of a material object.
#pragma once
#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../efem_gpu.h"

namespace efem {
namespace fe {

using Mat3 = std::array<double, 9>;  // row-major 3 x 3
using Vec3 = std::array<double, 3>;

// SparseMatrixTriplet (src/math/MathType.h): one (row, col, value) entry
struct Triplet {
    int row;
    int col;
    double value;
};

// a material object's numeric fields ("id", "density", "C10", ...)
using Json = std::map<std::string, double>;

enum class MaterialType {
    NEOHOOKEAN,
    MOONEY_RIVLIN,
    YEOH,
    IN_NEOHOOKEAN,
    IN_MOONEY_RIVLIN,
    IN_YEOH,
};

class Material {
public:
    int id = -1;
    MaterialType type = MaterialType::IN_NEOHOOKEAN;
    double density = 0.0;
    double elasticity_coeff = 0.0;

    virtual ~Material() = default;
    // P(F) of the law (host evaluation of the code the element kernels run)
    virtual Mat3 getFirstPiola(const Mat3& F) const;
    virtual void initFromJson(const Json& json) = 0;
    // the device law (EFEM_NEOHOOKEAN / _MOONEY_RIVLIN / _YEOH) and its four
    // parameters in efem_problem order; -1: no device kernel
    virtual int deviceLaw() const { return -1; }
    virtual std::array<double, 4> deviceParams() const { return {0.0, 0.0, 0.0, 0.0}; }
};

class NeoHookean : public Material {
public:
    double C10 = 0, D1 = 0;
    void initFromJson(const Json& json) override;
    int deviceLaw() const override { return EFEM_NEOHOOKEAN; }
    std::array<double, 4> deviceParams() const override { return {C10, 0.0, 0.0, D1}; }
};

class MooneyRivlin : public Material {
public:
    double C10 = 0, C01 = 0, D1 = 0;
    void initFromJson(const Json& json) override;
    int deviceLaw() const override { return EFEM_MOONEY_RIVLIN; }
    std::array<double, 4> deviceParams() const override { return {C10, C01, 0.0, D1}; }
};

class Yeoh : public Material {
public:
    double C10 = 0, C20 = 0, C30 = 0, D1 = 0;
    void initFromJson(const Json& json) override;
    int deviceLaw() const override { return EFEM_YEOH; }
    std::array<double, 4> deviceParams() const override { return {C10, C20, C30, D1}; }
};

// the mesh the constraints act on (Model's nodes and elements)
struct Model {
    std::vector<Vec3> nodes;     // reference positions
    int nen = 4;                 // nodes per element: 4, 8, 10 or 20
    std::vector<int> conn;       // [elements][nen], 0-based, reference node order
    std::vector<int> material;   // [elements] material id (ElementConstraintConfig binds by id)
    int reduced = 0;             // 1: reduced integration
    int elements() const { return nen > 0 ? static_cast<int>(conn.size()) / nen : 0; }
};

enum ConstraintType {
    ELEMENT,
    INERTIA,
    EXTERNL,
};

class Constraint;
struct ProblemParts;

class Constrainable {
public:
    ConstraintType constraint_type = EXTERNL;
    virtual ~Constrainable() = default;
    virtual void conInitial() = 0;
    virtual void conEvaluteResidual() = 0;
    // adds this constraint's part of r = f_int - f_ext to `residual`
    virtual void conGetResidual(std::vector<double>& residual) = 0;
    virtual void conEvaluteJacobi() = 0;
    // appends this constraint's approximate-Jacobian triplets
    virtual void conGetJacobi(std::vector<Triplet>& jacobi) = 0;

protected:
    friend class Constraint;
    Constraint* owner_ = nullptr;
    // what this constraint contributes to the device problem
    virtual void describe(ProblemParts& parts) const = 0;
};

// internal forces of every element of the model with its bound material
// (ElementConstraintConfig::configConstraint + ElementConstraint,
// src/fe/ElementConstraint.cpp:11-67,139-210); the approximate Jacobian is
// elasticity_coeff x the scalar Laplacian, K = L (x) I3 (:164-205)
class ElementConstraint : public Constrainable {
public:
    explicit ElementConstraint(std::vector<std::shared_ptr<Material>> materials);
    void conInitial() override {}
    void conEvaluteResidual() override;
    void conGetResidual(std::vector<double>& residual) override;
    void conEvaluteJacobi() override;
    void conGetJacobi(std::vector<Triplet>& jacobi) override;
    const std::vector<std::shared_ptr<Material>>& materials() const { return materials_; }

protected:
    void describe(ProblemParts& parts) const override;

private:
    std::vector<std::shared_ptr<Material>> materials_;
    std::vector<double> res_;
    std::vector<Triplet> jac_;
};

// fixed nodes (ExternalFixConfig, ExternalConstraint.cpp:50-82): they leave
// the free-DOF numbering; no residual, no Jacobian
class ExternalFixConstraint : public Constrainable {
public:
    explicit ExternalFixConstraint(std::vector<int> nodes);
    void conInitial() override {}
    void conEvaluteResidual() override {}
    void conGetResidual(std::vector<double>&) override {}
    void conEvaluteJacobi() override {}
    void conGetJacobi(std::vector<Triplet>&) override {}

protected:
    void describe(ProblemParts& parts) const override;

private:
    std::vector<int> nodes_;
};

// a constant force on each listed node (ExternalForceConstraint,
// ExternalConstraint.cpp:132-175): residual -= force on the free ones
class ExternalForceConstraint : public Constrainable {
public:
    ExternalForceConstraint(const Vec3& force, std::vector<int> nodes);
    void conInitial() override;
    void conEvaluteResidual() override {}
    void conGetResidual(std::vector<double>& residual) override;
    void conEvaluteJacobi() override {}
    void conGetJacobi(std::vector<Triplet>&) override {}

protected:
    void describe(ProblemParts& parts) const override;

private:
    Vec3 force_;
    std::vector<int> nodes_;
    std::vector<double> res_;
};

// follower pressure on element faces (ExternalPressureConstraint,
// ExternalConstraint.cpp:184-247,903-967); with a direction, the directed
// traction of :937-940.  faces: [count][fnn] node ids, fnn 3, 4, 6 or 8
class ExternalPressureConstraint : public Constrainable {
public:
    ExternalPressureConstraint(double pressure, int fnn, std::vector<int> faces);
    ExternalPressureConstraint(double pressure, int fnn, std::vector<int> faces, const Vec3& direction);
    void conInitial() override {}
    void conEvaluteResidual() override;
    void conGetResidual(std::vector<double>& residual) override;
    void conEvaluteJacobi() override {}
    void conGetJacobi(std::vector<Triplet>&) override {}

protected:
    void describe(ProblemParts& parts) const override;

private:
    double pressure_;
    int fnn_;
    std::vector<int> faces_;
    bool directed_ = false;
    Vec3 dir_{0.0, 0.0, 0.0};
    std::vector<double> res_;
};

// gravity -M g (InertiaConstraint::conEvaluteResidual and buildGravity,
// src/fe/InertiaConstraint.cpp:34-84): axis 1 +x, 2 -x, 3 +y, 4 -y, 5 +z,
// 6 -z; the mass matrix consistent or lumped (ElementConstraint.cpp:85-137)
// with the materials' densities
class InertiaConstraint : public Constrainable {
public:
    InertiaConstraint(int axis, bool consistent);
    void conInitial() override;
    void conEvaluteResidual() override {}
    void conGetResidual(std::vector<double>& residual) override;
    void conEvaluteJacobi() override {}
    void conGetJacobi(std::vector<Triplet>&) override {}

protected:
    void describe(ProblemParts& parts) const override;

private:
    int axis_;
    bool consistent_;
    std::vector<double> res_;
};

// the aggregate: DOF numbering, the device context, residual and Jacobian
// assembly over its constraints (src/fe/Constraint.cpp:22-138)
class Constraint {
public:
    // builds the device problem (efem_create); precision EFEM_FP64 or a mask
    // of EFEM_F32_*.  Calls conInitial on every constraint.
    Constraint(const Model& model, std::vector<std::shared_ptr<Constrainable>> constraints,
               int precision = EFEM_FP64);
    ~Constraint();
    Constraint(const Constraint&) = delete;
    Constraint& operator=(const Constraint&) = delete;

    // the increment dx on top of the committed displacement (:93-101)
    void assignDx(const std::vector<double>& dx);
    // commit: node dx += info dx, info dx = 0 (:103-113)
    void updateDx();
    // r = f_int - f_ext at the committed displacement + dx, all constraints
    // in one device pass (:115-126)
    void getResidual(std::vector<double>& residual);
    // every constraint's approximate-Jacobian triplets (:128-138)
    void getJacobi(std::vector<Triplet>& jacobi);
    int getDimens() const { return dimens_; }

    // Integrator::integral: L-BFGS from dx = 0 (Integrator.cpp:22-42);
    // dx holds the increment afterwards (zeroed on a non-finite result).
    // Returns ||dx||_inf < dx_criterion; stats optional
    bool integral(const efem_lbfgs_params& prm, double dx_criterion, efem_solve_stats* stats = nullptr);

    const std::vector<double>& dx() const { return dx_; }
    const std::vector<double>& committed() const { return uc_; }
    efem_ctx* context() const { return ctx_; }
    const efem_problem& problem() const;

    // r restricted to the residual terms in `mask` (1 element internal
    // forces, 2 pressure faces, 4 constant loads), for the per-constraint
    // conEvaluteResidual
    void termResidual(int mask, std::vector<double>& out);
    // the host problem with only the parts selected (for the constant-load
    // and Jacobian parts of single constraints)
    efem_problem partProblem(bool elements, bool forces, bool gravity) const;

private:
    struct Store;
    std::unique_ptr<Store> s_;
    std::vector<std::shared_ptr<Constrainable>> constraints_;
    efem_ctx* ctx_ = nullptr;
    int dimens_ = 0;
    std::vector<double> dx_, uc_;
};

}  // namespace fe
}  // namespace efem
