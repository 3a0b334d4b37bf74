// Test driver of the C++ FE interfaces (include/efem/fe.hpp): each command
// prints one JSON object on stdout for tests/test_fe_interface.py.
//   materials          P(F) of NeoHookean / MooneyRivlin / Yeoh built by
//                      initFromJson, and their elasticity_coeff (host only)
//   refusals           problems the device cannot hold are refused at
//                      Constraint construction (host only: nothing reaches
//                      the device before the checks)
//   solve W [opts..]   on synthetic workload W (GPU): the aggregate's
//                      residual against efem_residual on the same problem,
//                      the per-constraint parts, the Jacobian triplets,
//                      integral + updateDx
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/efem/fe.hpp"

using namespace efem::fe;

namespace {

void printArray(const char* name, const double* v, int n, bool last = false)
{
    std::printf("\"%s\": [", name);
    for (int i = 0; i < n; ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
    std::printf("]%s", last ? "" : ", ");
}

int materials()
{
    NeoHookean nh;
    nh.initFromJson({{"id", 3}, {"density", 1100}, {"C10", 1.5e5}, {"D1", 2e-6}});
    MooneyRivlin mr;
    mr.initFromJson({{"id", 4}, {"C10", 1.2e5}, {"C01", 3e4}, {"D1", 5e-6}});
    Yeoh ye;
    ye.initFromJson({{"id", 5}, {"C10", 1e5}, {"C20", -2e3}, {"C30", 150}, {"D1", 1e-5}});
    const Mat3 F = {1.1, 0.05, -0.02, 0.03, 0.95, 0.04, -0.01, 0.02, 1.02};
    const Mat3 Pn = nh.getFirstPiola(F), Pm = mr.getFirstPiola(F), Py = ye.getFirstPiola(F);
    std::printf("{");
    printArray("F", F.data(), 9);
    printArray("P_nh", Pn.data(), 9);
    printArray("P_mr", Pm.data(), 9);
    printArray("P_yeoh", Py.data(), 9);
    const double coeff[3] = {nh.elasticity_coeff, mr.elasticity_coeff, ye.elasticity_coeff};
    printArray("coeff", coeff, 3);
    std::printf("\"ids\": [%d, %d, %d], \"density\": %.17g, ", nh.id, mr.id, ye.id, nh.density);
    std::printf("\"types\": [%d, %d, %d]}\n", static_cast<int>(nh.type), static_cast<int>(mr.type),
                static_cast<int>(ye.type));
    return 0;
}

// a user material with no device law
class Custom : public Material {
public:
    void initFromJson(const Json&) override {}
    Mat3 getFirstPiola(const Mat3& F) const override { return F; }
};

Model tinyModel()
{
    Model m;
    m.nen = 4;
    m.nodes = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1}};
    m.conn = {0, 1, 2, 3, 1, 2, 3, 4};
    m.material = {0, 1};
    return m;
}

bool refused(const std::vector<std::shared_ptr<Constrainable>>& cons, const Model& m)
{
    try {
        Constraint c(m, cons);
    } catch (const std::invalid_argument&) {
        return true;
    } catch (const std::exception&) {
        return false;
    }
    return false;
}

int refusals()
{
    const Model m = tinyModel();
    auto nh0 = std::make_shared<NeoHookean>();
    nh0->initFromJson({{"id", 0}, {"C10", 1e5}, {"D1", 1e-5}});
    auto nh1 = std::make_shared<NeoHookean>();
    nh1->initFromJson({{"id", 1}, {"C10", 2e5}, {"D1", 1e-5}});
    auto mr1 = std::make_shared<MooneyRivlin>();
    mr1->initFromJson({{"id", 1}, {"C10", 1e5}, {"C01", 1e4}, {"D1", 1e-5}});
    auto custom = std::make_shared<Custom>();
    custom->id = 1;
    auto fix = std::make_shared<ExternalFixConstraint>(std::vector<int>{0});
    const bool noLaw = refused({std::make_shared<ElementConstraint>(std::vector<std::shared_ptr<Material>>{nh0, custom}), fix}, m);
    const bool mixed = refused({std::make_shared<ElementConstraint>(std::vector<std::shared_ptr<Material>>{nh0, mr1}), fix}, m);
    const bool noElements = refused({fix}, m);
    const bool twoPressures =
        refused({std::make_shared<ElementConstraint>(std::vector<std::shared_ptr<Material>>{nh0, nh1}), fix,
                 std::make_shared<ExternalPressureConstraint>(1.0, 3, std::vector<int>{1, 2, 4}),
                 std::make_shared<ExternalPressureConstraint>(2.0, 3, std::vector<int>{2, 3, 4})},
                m);
    const bool twoForces =
        refused({std::make_shared<ElementConstraint>(std::vector<std::shared_ptr<Material>>{nh0, nh1}), fix,
                 std::make_shared<ExternalForceConstraint>(Vec3{1, 0, 0}, std::vector<int>{4}),
                 std::make_shared<ExternalForceConstraint>(Vec3{0, 1, 0}, std::vector<int>{3})},
                m);
    std::printf("{\"no_device_law\": %d, \"mixed_laws\": %d, \"no_elements\": %d, \"two_pressures\": %d, "
                "\"two_force_vectors\": %d}\n",
                noLaw, mixed, noElements, twoPressures, twoForces);
    return 0;
}

double maxAbsDiff(const std::vector<double>& a, const std::vector<double>& b)
{
    double d = 0.0;
    for (size_t i = 0; i < a.size(); ++i) d = std::max(d, std::fabs(a[i] - b[i]));
    return d;
}

double maxAbs(const std::vector<double>& a)
{
    double d = 0.0;
    for (double v : a) d = std::max(d, std::fabs(v));
    return d;
}

int solve(int which, const std::vector<double>& opts)
{
    efem_mesh* mh = efem_mesh_build(which, opts.empty() ? nullptr : opts.data(), static_cast<int>(opts.size()));
    if (!mh) throw std::runtime_error("efem_mesh_build");
    efem_problem mp{};
    if (efem_mesh_problem(mh, &mp, EFEM_FP64) != EFEM_OK) throw std::runtime_error("efem_mesh_problem");
    if (mp.params_per_elem) throw std::runtime_error("workload with per-element parameters");
    // the same problem through the classes: one material, the fixed nodes,
    // the pressure faces, a nodal force on the last node and gravity -z
    Model m;
    m.nen = mp.nen;
    m.reduced = mp.reduced;
    for (int i = 0; i < mp.n_nodes; ++i) m.nodes.push_back({mp.X[3 * i], mp.X[3 * i + 1], mp.X[3 * i + 2]});
    m.conn.assign(mp.conn, mp.conn + static_cast<size_t>(mp.n_elems) * mp.nen);
    m.material.assign(mp.n_elems, 7);
    std::shared_ptr<Material> mat;
    if (mp.law == EFEM_NEOHOOKEAN) {
        auto p = std::make_shared<NeoHookean>();
        p->initFromJson({{"id", 7}, {"density", 1000}, {"C10", mp.params[0]}, {"D1", mp.params[3]}});
        mat = p;
    } else if (mp.law == EFEM_MOONEY_RIVLIN) {
        auto p = std::make_shared<MooneyRivlin>();
        p->initFromJson({{"id", 7}, {"density", 1000}, {"C10", mp.params[0]}, {"C01", mp.params[1]}, {"D1", mp.params[3]}});
        mat = p;
    } else {
        auto p = std::make_shared<Yeoh>();
        p->initFromJson({{"id", 7}, {"density", 1000}, {"C10", mp.params[0]}, {"C20", mp.params[1]},
                         {"C30", mp.params[2]}, {"D1", mp.params[3]}});
        mat = p;
    }
    std::vector<int> fixedNodes;
    for (int i = 0; i < mp.n_nodes; ++i)
        if (mp.fixed && mp.fixed[i]) fixedNodes.push_back(i);
    const Vec3 force = {0.0, 0.0, -0.05};
    const std::vector<int> forceNodes = {mp.n_nodes - 1};
    auto ec = std::make_shared<ElementConstraint>(std::vector<std::shared_ptr<Material>>{mat});
    auto pc = std::make_shared<ExternalPressureConstraint>(
        mp.pressure, mp.fnn, std::vector<int>(mp.faces, mp.faces + static_cast<size_t>(mp.n_faces) * mp.fnn));
    auto fc = std::make_shared<ExternalForceConstraint>(force, forceNodes);
    auto ic = std::make_shared<InertiaConstraint>(6, false);
    std::vector<std::shared_ptr<Constrainable>> cons = {
        ec, std::make_shared<ExternalFixConstraint>(fixedNodes), pc, fc, ic};
    Constraint con(m, cons, EFEM_FP64);
    const int n = con.getDimens();

    // the same problem straight through the C ABI
    efem_problem pb = con.problem();
    efem_ctx* ctx = nullptr;
    if (efem_create(&ctx, &pb) != EFEM_OK) throw std::runtime_error(efem_global_error());

    std::mt19937_64 rng(11);
    std::normal_distribution<double> g(0.0, 1e-3);
    std::vector<double> dx(n);
    for (double& v : dx) v = g(rng);
    con.assignDx(dx);
    std::vector<double> r, rc(n), parts(n, 0.0);
    con.getResidual(r);
    efem_residual(ctx, dx.data(), rc.data());
    for (const auto& c : cons) c->conEvaluteResidual();
    for (const auto& c : cons) c->conGetResidual(parts);
    std::vector<Triplet> jac;
    con.getJacobi(jac);
    long long nnz = 0;
    efem_host_jacobi(&pb, &nnz, nullptr, nullptr, nullptr, 0);
    std::vector<int> jr(nnz), jc(nnz);
    std::vector<double> jv(nnz);
    efem_host_jacobi(&pb, &nnz, jr.data(), jc.data(), jv.data(), nnz);
    double jdiff = jac.size() == static_cast<size_t>(nnz) ? 0.0 : 1.0;
    for (long long q = 0; q < nnz && jdiff == 0.0; ++q)
        if (jac[q].row != jr[q] || jac[q].col != jc[q] || jac[q].value != jv[q]) jdiff = 1.0;

    // Integrator::integral from dx = 0, then commit and solve again (the
    // second solve starts at equilibrium)
    con.assignDx(std::vector<double>(n, 0.0));
    efem_lbfgs_params prm{};
    prm.epsilon = 1e-9;
    prm.epsilon_rel = 0.0;
    prm.max_iterations = 2000;
    prm.max_linesearch = 20;
    prm.m = 6;
    prm.lns_fea = 1.0;
    prm.use_precond = 1;
    prm.line_search = EFEM_LS_PROJECTION;
    efem_solve_stats st{};
    const bool conv = con.integral(prm, 1e30, &st);
    const std::vector<double> dx1 = con.dx();
    // the C ABI's own solve from 0 on the same problem
    std::vector<double> u(n, 0.0);
    efem_solve_stats st2{};
    efem_solve(ctx, u.data(), &prm, &st2);
    con.updateDx();
    std::vector<double> r2;
    con.getResidual(r2);
    efem_solve_stats st3{};
    con.integral(prm, 1e30, &st3);

    std::printf("{\"ndof\": %d, \"res_scale\": %.17g, \"res_vs_abi\": %.17g, \"parts_vs_total\": %.17g, "
                "\"jacobi_nnz\": %zu, \"jacobi_mismatch\": %g, \"converged\": %d, \"iterations\": %d, "
                "\"abi_iterations\": %d, \"dx_vs_abi\": %.17g, \"dx_scale\": %.17g, "
                "\"committed_residual\": %.17g, \"second_iterations\": %d, \"committed_vs_dx\": %.17g, "
                "\"rinf\": %.17g, \"second_rinf\": %.17g}\n",
                n, maxAbs(r), maxAbsDiff(r, rc), maxAbsDiff(parts, r), jac.size(), jdiff, conv ? 1 : 0,
                st.iterations, st2.iterations, maxAbsDiff(dx1, u), maxAbs(u), maxAbs(r2), st3.iterations,
                maxAbsDiff(con.committed(), u), st.rinf, st3.rinf);
    efem_destroy(ctx);
    efem_mesh_free(mh);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        const std::string cmd = argc > 1 ? argv[1] : "";
        if (cmd == "materials") return materials();
        if (cmd == "refusals") return refusals();
        if (cmd == "solve" && argc > 2) {
            std::vector<double> opts;
            for (int i = 3; i < argc; ++i) opts.push_back(std::atof(argv[i]));
            return solve(std::atoi(argv[2]), opts);
        }
        std::fprintf(stderr, "usage: fe_driver materials | refusals | solve W [opts...]\n");
        return 2;
    } catch (const std::exception& e) {
        std::printf("{\"error\": \"%s\"}\n", e.what());
        return 1;
    }
}
