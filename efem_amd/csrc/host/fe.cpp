// The reference's C++ FE interfaces (include/efem/fe.hpp) over the C ABI.
// Materials evaluate on the host through efem_material_point (the law code
// the element kernels run); residuals, factorisations and solves run on the
// device through the context the Constraint aggregate creates.
#include "../../../include/efem/fe.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace efem {
namespace fe {

namespace {

double field(const Json& j, const char* key, double keep)
{
    const auto it = j.find(key);
    return it == j.end() ? keep : it->second;
}

void check(int rc, const char* what)
{
    if (rc != EFEM_OK) throw std::runtime_error(std::string(what) + ": " + efem_global_error());
}

}  // namespace

// ---- materials -------------------------------------------------------------

Mat3 Material::getFirstPiola(const Mat3& F) const
{
    const int law = deviceLaw();
    if (law < 0) throw std::logic_error("Material::getFirstPiola: no law to evaluate (override it)");
    const std::array<double, 4> p = deviceParams();
    Mat3 P{};
    if (efem_material_point(law, p.data(), F.data(), P.data(), nullptr, nullptr, nullptr) != EFEM_OK)
        throw std::invalid_argument("Material::getFirstPiola: efem_material_point refused the law");
    return P;
}

// initFromJson of src/material/NeoHookean.cpp:22-46 ("t = 2.0" coefficient)
void NeoHookean::initFromJson(const Json& json)
{
    id = static_cast<int>(field(json, "id", id));
    density = field(json, "density", density);
    C10 = field(json, "C10", C10);
    D1 = field(json, "D1", D1);
    elasticity_coeff = 2.94001 * C10 + 2.0 / D1;
    type = MaterialType::IN_NEOHOOKEAN;
}

// src/material/MooneyRivlin.cpp:26-54
void MooneyRivlin::initFromJson(const Json& json)
{
    id = static_cast<int>(field(json, "id", id));
    density = field(json, "density", density);
    C10 = field(json, "C10", C10);
    C01 = field(json, "C01", C01);
    D1 = field(json, "D1", D1);
    elasticity_coeff = 3.818 * C01 + 2.94001 * C10 + 2.0 / D1;
    type = MaterialType::IN_MOONEY_RIVLIN;
}

// src/material/Yeoh.cpp:22-53
void Yeoh::initFromJson(const Json& json)
{
    id = static_cast<int>(field(json, "id", id));
    density = field(json, "density", density);
    C10 = field(json, "C10", C10);
    C20 = field(json, "C20", C20);
    C30 = field(json, "C30", C30);
    D1 = field(json, "D1", D1);
    elasticity_coeff = 2.94001 * C10 + 3.51194 * C20 + 3.3154 * C30 + 2.0 / D1;
    type = MaterialType::IN_YEOH;
}

// ---- what the constraints contribute to the device problem ----------------

struct ProblemParts {
    const Model* model = nullptr;
    int elementConstraints = 0;
    std::vector<std::shared_ptr<Material>> materials;
    std::vector<int> fixed;
    int forceConstraints = 0;
    Vec3 force{0.0, 0.0, 0.0};
    std::vector<int> forceNodes;
    int pressureConstraints = 0;
    double pressure = 0.0;
    int fnn = 0;
    std::vector<int> faces;
    bool directed = false;
    Vec3 dir{0.0, 0.0, 0.0};
    int inertiaConstraints = 0;
    int gravity = 0;
    bool consistent = false;
};

struct Constraint::Store {
    efem_problem pb{};
    std::vector<double> X, params, density;
    std::vector<int> conn, faces, fnodes;
    std::vector<unsigned char> fixed;
    double fvec[3] = {0, 0, 0};
    double pdir[3] = {0, 0, 0};
    int gravity = 0;
};

// ---- constraints -------------------------------------------------------------

ElementConstraint::ElementConstraint(std::vector<std::shared_ptr<Material>> materials)
    : materials_(std::move(materials))
{
    constraint_type = ELEMENT;
}

void ElementConstraint::describe(ProblemParts& p) const
{
    ++p.elementConstraints;
    p.materials = materials_;
}

void ElementConstraint::conEvaluteResidual()
{
    owner_->termResidual(1, res_);
}

void ElementConstraint::conGetResidual(std::vector<double>& residual)
{
    for (size_t i = 0; i < res_.size() && i < residual.size(); ++i) residual[i] += res_[i];
}

void ElementConstraint::conEvaluteJacobi()
{
    const efem_problem pb = owner_->partProblem(true, false, false);
    long long nnz = 0;
    check(efem_host_jacobi(&pb, &nnz, nullptr, nullptr, nullptr, 0), "efem_host_jacobi");
    std::vector<int> r(nnz), c(nnz);
    std::vector<double> v(nnz);
    check(efem_host_jacobi(&pb, &nnz, r.data(), c.data(), v.data(), nnz), "efem_host_jacobi");
    jac_.resize(nnz);
    for (long long q = 0; q < nnz; ++q) jac_[q] = Triplet{r[q], c[q], v[q]};
}

void ElementConstraint::conGetJacobi(std::vector<Triplet>& jacobi)
{
    jacobi.insert(jacobi.end(), jac_.begin(), jac_.end());
}

ExternalFixConstraint::ExternalFixConstraint(std::vector<int> nodes) : nodes_(std::move(nodes))
{
    constraint_type = EXTERNL;
}

void ExternalFixConstraint::describe(ProblemParts& p) const
{
    p.fixed.insert(p.fixed.end(), nodes_.begin(), nodes_.end());
}

ExternalForceConstraint::ExternalForceConstraint(const Vec3& force, std::vector<int> nodes)
    : force_(force), nodes_(std::move(nodes))
{
    constraint_type = EXTERNL;
}

void ExternalForceConstraint::describe(ProblemParts& p) const
{
    if (p.forceConstraints > 0 && p.force != force_)
        throw std::invalid_argument("ExternalForceConstraint: force constraints with different vectors "
                                    "(the device problem holds one nodal-force vector)");
    ++p.forceConstraints;
    p.force = force_;
    p.forceNodes.insert(p.forceNodes.end(), nodes_.begin(), nodes_.end());
}

// the constant residual part -force on this constraint's free nodes
// (ExternalForceConstraint::conEvaluteResidual, ExternalConstraint.cpp:155-170)
void ExternalForceConstraint::conInitial()
{
    efem_problem pb = owner_->partProblem(false, false, false);
    pb.n_fnodes = static_cast<int>(nodes_.size());
    pb.fnodes = nodes_.data();
    pb.fvec = force_.data();
    res_.assign(owner_->getDimens(), 0.0);
    check(efem_host_constant_force(&pb, res_.data(), owner_->getDimens()), "efem_host_constant_force");
    for (double& v : res_) v = -v;
}

void ExternalForceConstraint::conGetResidual(std::vector<double>& residual)
{
    for (size_t i = 0; i < res_.size() && i < residual.size(); ++i) residual[i] += res_[i];
}

ExternalPressureConstraint::ExternalPressureConstraint(double pressure, int fnn, std::vector<int> faces)
    : pressure_(pressure), fnn_(fnn), faces_(std::move(faces))
{
    constraint_type = EXTERNL;
}

ExternalPressureConstraint::ExternalPressureConstraint(double pressure, int fnn, std::vector<int> faces,
                                                       const Vec3& direction)
    : pressure_(pressure), fnn_(fnn), faces_(std::move(faces)), directed_(true), dir_(direction)
{
    constraint_type = EXTERNL;
}

void ExternalPressureConstraint::describe(ProblemParts& p) const
{
    if (p.pressureConstraints > 0)
        throw std::invalid_argument("ExternalPressureConstraint: more than one pressure constraint "
                                    "(the device problem holds one)");
    ++p.pressureConstraints;
    p.pressure = pressure_;
    p.fnn = fnn_;
    p.faces = faces_;
    p.directed = directed_;
    p.dir = dir_;
}

void ExternalPressureConstraint::conEvaluteResidual()
{
    owner_->termResidual(2, res_);
}

void ExternalPressureConstraint::conGetResidual(std::vector<double>& residual)
{
    for (size_t i = 0; i < res_.size() && i < residual.size(); ++i) residual[i] += res_[i];
}

InertiaConstraint::InertiaConstraint(int axis, bool consistent) : axis_(axis), consistent_(consistent)
{
    constraint_type = INERTIA;
}

void InertiaConstraint::describe(ProblemParts& p) const
{
    if (p.inertiaConstraints > 0) throw std::invalid_argument("InertiaConstraint: more than one");
    if (axis_ < 0 || axis_ > 6) throw std::invalid_argument("InertiaConstraint: gravity axis 0..6");
    ++p.inertiaConstraints;
    p.gravity = axis_;
    p.consistent = consistent_;
}

// -M g over the free DOFs (InertiaConstraint.cpp:34-47), constant
void InertiaConstraint::conInitial()
{
    const efem_problem pb = owner_->partProblem(true, false, true);
    res_.assign(owner_->getDimens(), 0.0);
    check(efem_host_constant_force(&pb, res_.data(), owner_->getDimens()), "efem_host_constant_force");
    for (double& v : res_) v = -v;
}

void InertiaConstraint::conGetResidual(std::vector<double>& residual)
{
    for (size_t i = 0; i < res_.size() && i < residual.size(); ++i) residual[i] += res_[i];
}

// ---- the aggregate -----------------------------------------------------------

Constraint::Constraint(const Model& model, std::vector<std::shared_ptr<Constrainable>> constraints, int precision)
    : s_(new Store), constraints_(std::move(constraints))
{
    ProblemParts parts;
    parts.model = &model;
    for (const auto& c : constraints_) {
        if (!c) throw std::invalid_argument("Constraint: null constraint");
        c->describe(parts);
    }
    if (parts.elementConstraints != 1)
        throw std::invalid_argument("Constraint: exactly one ElementConstraint (the model's elements) is needed");
    const int nn = static_cast<int>(model.nodes.size());
    const int E = model.elements();
    if (model.nen <= 0 || static_cast<int>(model.conn.size()) != E * model.nen)
        throw std::invalid_argument("Constraint: conn is not [elements][nen]");
    if (static_cast<int>(model.material.size()) != E)
        throw std::invalid_argument("Constraint: one material id per element is needed");
    // bind materials by id (ElementConstraintConfig::configConstraint,
    // ElementConstraint.cpp:11-37): elements whose id has no material are
    // left out, as there
    std::map<int, std::shared_ptr<Material>> byId;
    for (const auto& m : parts.materials)
        if (m) byId[m->id] = m;
    Store& S = *s_;
    int law = -1;
    for (int e = 0; e < E; ++e) {
        const auto it = byId.find(model.material[e]);
        if (it == byId.end()) continue;
        const Material& mat = *it->second;
        const int l = mat.deviceLaw();
        if (l < 0)
            throw std::invalid_argument("Constraint: material " + std::to_string(mat.id) +
                                        " has no device law (NeoHookean, MooneyRivlin or Yeoh)");
        if (law >= 0 && l != law)
            throw std::invalid_argument("Constraint: the elements' materials mix laws (one law per problem)");
        law = l;
        const std::array<double, 4> p = mat.deviceParams();
        S.params.insert(S.params.end(), p.begin(), p.end());
        S.density.push_back(mat.density);
        S.conn.insert(S.conn.end(), model.conn.begin() + static_cast<size_t>(e) * model.nen,
                      model.conn.begin() + static_cast<size_t>(e + 1) * model.nen);
    }
    if (law < 0) throw std::invalid_argument("Constraint: no element has a bound material");
    S.X.resize(3 * static_cast<size_t>(nn));
    for (int i = 0; i < nn; ++i)
        for (int d = 0; d < 3; ++d) S.X[3 * static_cast<size_t>(i) + d] = model.nodes[i][d];
    S.fixed.assign(nn, 0);
    for (int n : parts.fixed) {
        if (n < 0 || n >= nn) throw std::invalid_argument("ExternalFixConstraint: node out of range");
        S.fixed[n] = 1;
    }
    S.fnodes = parts.forceNodes;
    for (int d = 0; d < 3; ++d) S.fvec[d] = parts.force[d];
    S.faces = parts.faces;
    for (int d = 0; d < 3; ++d) S.pdir[d] = parts.dir[d];
    S.gravity = parts.gravity;

    efem_problem& pb = S.pb;
    pb.n_nodes = nn;
    pb.X = S.X.data();
    pb.n_elems = static_cast<int>(S.conn.size()) / model.nen;
    pb.nen = model.nen;
    pb.conn = S.conn.data();
    pb.reduced = model.reduced;
    pb.law = law;
    pb.params = S.params.data();
    pb.params_per_elem = 1;
    pb.fixed = S.fixed.data();
    pb.n_faces = parts.fnn > 0 ? static_cast<int>(S.faces.size()) / parts.fnn : 0;
    pb.fnn = parts.fnn;
    pb.faces = S.faces.empty() ? nullptr : S.faces.data();
    pb.pressure = parts.pressure;
    pb.pdir = parts.directed ? S.pdir : nullptr;
    pb.n_fnodes = static_cast<int>(S.fnodes.size());
    pb.fnodes = S.fnodes.empty() ? nullptr : S.fnodes.data();
    pb.fvec = S.fvec;
    pb.volumes = nullptr;
    pb.precision = precision;
    pb.gravity = parts.gravity;
    pb.consistent_inertia = parts.consistent ? 1 : 0;
    pb.density = S.density.data();

    check(efem_create(&ctx_, &pb), "efem_create");
    dimens_ = efem_ndof(ctx_);
    dx_.assign(dimens_, 0.0);
    uc_.assign(dimens_, 0.0);
    for (const auto& c : constraints_) {
        c->owner_ = this;
        c->conInitial();
    }
}

Constraint::~Constraint()
{
    if (ctx_) efem_destroy(ctx_);
}

const efem_problem& Constraint::problem() const { return s_->pb; }

efem_problem Constraint::partProblem(bool elements, bool forces, bool gravity) const
{
    efem_problem pb = s_->pb;
    (void)elements;  // the elements always define the free DOFs and the Laplacian
    pb.n_faces = 0;
    pb.faces = nullptr;
    pb.pdir = nullptr;
    if (!forces) {
        pb.n_fnodes = 0;
        pb.fnodes = nullptr;
    }
    if (!gravity) pb.gravity = 0;
    return pb;
}

void Constraint::assignDx(const std::vector<double>& dx)
{
    if (static_cast<int>(dx.size()) != dimens_) throw std::invalid_argument("Constraint::assignDx: size");
    dx_ = dx;
}

void Constraint::updateDx()
{
    for (int i = 0; i < dimens_; ++i) uc_[i] += dx_[i];
    std::fill(dx_.begin(), dx_.end(), 0.0);
    if (efem_commit(ctx_, uc_.data()) != EFEM_OK)
        throw std::runtime_error(std::string("efem_commit: ") + efem_last_error(ctx_));
}

void Constraint::termResidual(int mask, std::vector<double>& out)
{
    std::vector<double> u(dimens_);
    for (int i = 0; i < dimens_; ++i) u[i] = uc_[i] + dx_[i];
    out.assign(dimens_, 0.0);
    int rc = efem_set_option(ctx_, "residual_terms", mask);
    if (rc == EFEM_OK) rc = efem_residual(ctx_, u.data(), out.data());
    const int rc2 = efem_set_option(ctx_, "residual_terms", 7);
    if (rc != EFEM_OK || rc2 != EFEM_OK)
        throw std::runtime_error(std::string("efem_residual: ") + efem_last_error(ctx_));
}

void Constraint::getResidual(std::vector<double>& residual)
{
    termResidual(7, residual);
}

void Constraint::getJacobi(std::vector<Triplet>& jacobi)
{
    for (const auto& c : constraints_) c->conEvaluteJacobi();
    jacobi.clear();
    for (const auto& c : constraints_) c->conGetJacobi(jacobi);
}

bool Constraint::integral(const efem_lbfgs_params& prm, double dx_criterion, efem_solve_stats* stats)
{
    // from dx = 0 on top of the committed displacement
    std::vector<double> u = uc_;
    efem_solve_stats st{};
    const int rc = efem_solve(ctx_, u.data(), &prm, &st);
    if (stats) *stats = st;
    if (rc == EFEM_ERR_NONFINITE) {
        std::fill(dx_.begin(), dx_.end(), 0.0);
        return false;
    }
    if (rc != EFEM_OK) throw std::runtime_error(std::string("efem_solve: ") + efem_last_error(ctx_));
    double dxInf = 0.0;
    for (int i = 0; i < dimens_; ++i) {
        dx_[i] = u[i] - uc_[i];
        dxInf = std::max(dxInf, std::fabs(dx_[i]));
    }
    return dxInf < dx_criterion;
}

}  // namespace fe
}  // namespace efem
