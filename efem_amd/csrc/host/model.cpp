// Host preprocessing (see model.hpp).
#include "model.hpp"

#include <chrono>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../material_math.hpp"
#include "element.hpp"

namespace efem {
namespace {

double inv3(const double* A, double* B)
{
    const double c00 = A[4] * A[8] - A[5] * A[7];
    const double c01 = A[5] * A[6] - A[3] * A[8];
    const double c02 = A[3] * A[7] - A[4] * A[6];
    const double d = A[0] * c00 + A[1] * c01 + A[2] * c02;
    const double id = 1.0 / d;
    B[0] = c00 * id;
    B[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    B[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    B[3] = c01 * id;
    B[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    B[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    B[6] = c02 * id;
    B[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    B[8] = (A[0] * A[4] - A[1] * A[3]) * id;
    return d;
}

// J0 = sum_a (X_a - X_0) (x) dN_a ; translation invariant because sum dN_a = 0
void jacobian0(int nen, const double* Xe, const double* dN, double* J0)
{
    for (int k = 0; k < 9; ++k) J0[k] = 0.0;
    for (int a = 1; a < nen; ++a)
        for (int i = 0; i < 3; ++i) {
            const double d = Xe[3 * a + i] - Xe[i];
            for (int j = 0; j < 3; ++j) J0[3 * i + j] += d * dN[3 * a + j];
        }
}

}  // namespace

bool setParams(Model& m, const double* params, std::string& msg)
{
    const size_t np = m.paramsPerElem ? 4 * static_cast<size_t>(m.nElems) : 4;
    for (size_t k = 0; k < np; k += 4)
        if (!(params[k + 3] > 0.0)) {
            msg = "material parameters: D1 must be positive";
            return false;
        }
    m.params.assign(params, params + np);
    m.coeff.resize(m.nElems);
    for (int e = 0; e < m.nElems; ++e)
        m.coeff[e] = elasticityCoeff(m.law, &m.params[m.paramsPerElem ? 4 * static_cast<size_t>(e) : 0]);
    return true;
}

bool concatModels(const std::vector<Model>& parts, Model& out, std::vector<double>& facePressure,
                  std::vector<PartRange>& ranges, std::string& msg)
{
    if (parts.empty()) {
        msg = "batch: no instances";
        return false;
    }
    const Model& p0 = parts[0];
    int N = 0, E = 0, F = 0, NF = 0;
    bool anyFaces = false;
    for (const Model& p : parts) {
        if (p.nen != p0.nen || p.ng != p0.ng || p.reduced != p0.reduced || p.law != p0.law) {
            msg = "batch: instances must share the element kind, integration rule and law";
            return false;
        }
        if (p.nFaces > 0) {
            if (anyFaces && (p.fq != out.fq || p.directed != out.directed ||
                             (p.directed && (p.pdir[0] != out.pdir[0] || p.pdir[1] != out.pdir[1] ||
                                             p.pdir[2] != out.pdir[2])))) {
                msg = "batch: pressure faces must share the face rule and load kind (directed: the same vector)";
                return false;
            }
            if (!anyFaces) {
                out.fq = p.fq;
                out.fnn = p.fnn;
                out.directed = p.directed;
                for (int k = 0; k < 3; ++k) out.pdir[k] = p.pdir[k];
            }
            anyFaces = true;
        }
        N += p.nNodes;
        E += p.nElems;
        F += p.nFaces;
        NF += p.nFree;
    }
    const int nen = p0.nen, ng = p0.ng;
    out.nNodes = N;
    out.nElems = E;
    out.nen = nen;
    out.ng = ng;
    out.law = p0.law;
    out.reduced = p0.reduced;
    out.nFree = NF;
    out.paramsPerElem = 1;
    out.gw = p0.gw;
    out.gdN = p0.gdN;
    out.nFaces = F;
    out.pressure = 1.0;
    out.X.reserve(3 * static_cast<size_t>(N));
    out.conn.reserve(static_cast<size_t>(E) * nen);
    out.fid.reserve(N);
    out.connDof.assign(static_cast<size_t>(nen) * E, 0);
    out.ginv.assign(static_cast<size_t>(ng) * 9 * E, 0.0);
    out.params.reserve(4 * static_cast<size_t>(E));
    facePressure.clear();
    ranges.clear();
    int nOff = 0, eOff = 0, fOff = 0, freeOff = 0;
    out.gptr.assign(1, 0);
    for (const Model& p : parts) {
        ranges.push_back({eOff, eOff + p.nElems, fOff, fOff + p.nFaces, freeOff, freeOff + p.nFree});
        out.X.insert(out.X.end(), p.X.begin(), p.X.end());
        for (int c : p.conn) out.conn.push_back(c + nOff);
        for (int f : p.fid) out.fid.push_back(f >= 0 ? f + freeOff : -1);
        for (int a = 0; a < nen; ++a)
            for (int e = 0; e < p.nElems; ++e) {
                const int d = p.connDof[static_cast<size_t>(a) * p.nElems + e];
                out.connDof[static_cast<size_t>(a) * E + eOff + e] = d == p.nFree ? NF : d + freeOff;
            }
        out.vol.insert(out.vol.end(), p.vol.begin(), p.vol.end());
        for (int q = 0; q < ng; ++q)
            for (int k = 0; k < 9; ++k)
                for (int e = 0; e < p.nElems; ++e)
                    out.ginv[(static_cast<size_t>(q) * 9 + k) * E + eOff + e] =
                        p.ginv[(static_cast<size_t>(q) * 9 + k) * p.nElems + e];
        for (int e = 0; e < p.nElems; ++e)
            for (int k = 0; k < 4; ++k) out.params.push_back(p.params[p.paramsPerElem ? 4 * static_cast<size_t>(e) + k : k]);
        out.coeff.insert(out.coeff.end(), p.coeff.begin(), p.coeff.end());
        for (int f = 0; f < p.nFaces; ++f) {
            out.faceElem.push_back(p.faceElem[f] + eOff);
            facePressure.push_back(p.pressure);
        }
        out.fA0.insert(out.fA0.end(), p.fA0.begin(), p.fA0.end());
        out.fWN.insert(out.fWN.end(), p.fWN.begin(), p.fWN.end());
        out.fG.insert(out.fG.end(), p.fG.begin(), p.fG.end());
        out.fext.insert(out.fext.end(), p.fext.begin(), p.fext.end());
        // gather CSR: element slots shift by the element offset, face slots
        // move behind all element slots
        const int es = p.nElems * nen;
        for (int i = 0; i < p.nFree; ++i) {
            for (int q = p.gptr[i]; q < p.gptr[i + 1]; ++q) {
                const int sl = p.gslot[q];
                out.gslot.push_back(sl < es ? sl + eOff * nen : E * nen + fOff * nen + (sl - es));
            }
            out.gptr.push_back(static_cast<int>(out.gslot.size()));
        }
        nOff += p.nNodes;
        eOff += p.nElems;
        fOff += p.nFaces;
        freeOff += p.nFree;
    }
    return true;
}

bool buildModel(const efem_problem& pb, Model& m, std::string& msg)
{
    if (pb.n_nodes <= 0 || !pb.X || pb.n_elems <= 0 || !pb.conn || !pb.params) {
        msg = "efem_problem: missing nodes, elements or material parameters";
        return false;
    }
    auto el = makeElement(pb.nen, pb.reduced != 0);
    if (!el) {
        msg = "efem_problem: nen must be 4, 8, 10 or 20";
        return false;
    }
    if (pb.law < 0 || pb.law > 2) {
        msg = "efem_problem: unknown constitutive law";
        return false;
    }
    m = Model();
    m.nNodes = pb.n_nodes;
    m.nElems = pb.n_elems;
    m.nen = pb.nen;
    m.reduced = pb.reduced;
    m.law = pb.law;
    m.ng = el->gaussPointNum();
    m.paramsPerElem = pb.params_per_elem ? 1 : 0;
    const int nen = m.nen, E = m.nElems, ng = m.ng;
    m.X.assign(pb.X, pb.X + 3 * static_cast<size_t>(m.nNodes));
    m.conn.assign(pb.conn, pb.conn + static_cast<size_t>(nen) * E);
    for (int c : m.conn)
        if (c < 0 || c >= m.nNodes) {
            msg = "efem_problem: connectivity out of range";
            return false;
        }
    // free-DOF numbering (Constraint.cpp:61-72)
    m.fid.assign(m.nNodes, -1);
    int id = 0;
    for (int n = 0; n < m.nNodes; ++n)
        if (!(pb.fixed && pb.fixed[n])) m.fid[n] = id++;
    m.nFree = id;
    m.connDof.resize(static_cast<size_t>(nen) * E);
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < nen; ++a) {
            const int f = m.fid[m.conn[static_cast<size_t>(e) * nen + a]];
            m.connDof[static_cast<size_t>(a) * E + e] = f >= 0 ? f : m.nFree;
        }
    // materials
    if (!setParams(m, pb.params, msg)) return false;
    // Gauss rule and local gradients
    m.gw.resize(ng);
    m.gdN.resize(static_cast<size_t>(ng) * nen * 3);
    for (int q = 0; q < ng; ++q) {
        double xi[3];
        el->gaussPointLocalPosition(q, xi);
        m.gw[q] = el->gaussPointWeight(q);
        el->pointShapeFuncLocalGradient(xi, &m.gdN[static_cast<size_t>(q) * nen * 3]);
    }
    // element geometry: volume and J0^-1 per Gauss point
    m.vol.resize(E);
    m.ginv.resize(static_cast<size_t>(ng) * 9 * E);
    double Xe[60];
    for (int e = 0; e < E; ++e) {
        for (int a = 0; a < nen; ++a)
            for (int d = 0; d < 3; ++d) Xe[3 * a + d] = m.X[3 * static_cast<size_t>(m.conn[static_cast<size_t>(e) * nen + a]) + d];
        m.vol[e] = pb.volumes ? pb.volumes[e] : elementVolume(nen, Xe);
        for (int q = 0; q < ng; ++q) {
            double J0[9], Ji[9];
            jacobian0(nen, Xe, &m.gdN[static_cast<size_t>(q) * nen * 3], J0);
            const double d = inv3(J0, Ji);
            if (!(std::fabs(d) > 0.0) || !std::isfinite(d)) {
                msg = "efem_problem: degenerate element " + std::to_string(e);
                return false;
            }
            for (int k = 0; k < 9; ++k) m.ginv[(static_cast<size_t>(q) * 9 + k) * E + e] = Ji[k];
        }
    }
    // pressure faces (ExternalPressureConfig / ExternalPressureFace)
    if (pb.n_faces > 0 && pb.faces && (pb.pdir || pb.pressure > 0.0)) {
        m.fnn = pb.fnn;
        m.nFaces = pb.n_faces;
        m.directed = pb.pdir ? 1 : 0;
        if (pb.pdir)
            for (int d = 0; d < 3; ++d) m.pdir[d] = pb.pdir[d];
        m.pressure = pb.pdir ? 1.0 : pb.pressure;
        // node -> elements
        std::vector<int> nptr(m.nNodes + 1, 0), nel;
        for (int c : m.conn) nptr[c + 1]++;
        for (int n = 0; n < m.nNodes; ++n) nptr[n + 1] += nptr[n];
        nel.resize(nptr[m.nNodes]);
        {
            std::vector<int> pos(nptr.begin(), nptr.end() - 1);
            for (int e = 0; e < E; ++e)
                for (int a = 0; a < nen; ++a) nel[pos[m.conn[static_cast<size_t>(e) * nen + a]]++] = e;
        }
        m.faceElem.resize(m.nFaces);
        m.fA0.resize(3 * static_cast<size_t>(m.nFaces));
        for (int f = 0; f < m.nFaces; ++f) {
            const int* fn = pb.faces + static_cast<size_t>(f) * m.fnn;
            for (int i = 0; i < m.fnn; ++i)
                if (fn[i] < 0 || fn[i] >= m.nNodes) {
                    msg = "efem_problem: face node out of range";
                    return false;
                }
            // the unique element holding every face node
            int parent = -1, count = 0;
            for (int p = nptr[fn[0]]; p < nptr[fn[0] + 1]; ++p) {
                const int e = nel[p];
                bool all = true;
                for (int i = 1; i < m.fnn && all; ++i) {
                    bool found = false;
                    for (int a = 0; a < nen; ++a) found |= m.conn[static_cast<size_t>(e) * nen + a] == fn[i];
                    all = found;
                }
                if (all) {
                    parent = e;
                    ++count;
                }
            }
            if (count != 1) {
                msg = "pressure face " + std::to_string(f) + ": no unique parent element";
                return false;
            }
            FaceRule rule;
            if (!buildFaceRule(nen, &m.conn[static_cast<size_t>(parent) * nen], m.fnn, fn, m.X.data(), rule, msg))
                return false;
            if (f == 0) {
                m.fq = rule.nq;
                m.fWN.resize(static_cast<size_t>(m.nFaces) * m.fq * nen);
                m.fG.resize(static_cast<size_t>(m.nFaces) * m.fq * nen * 3);
            }
            m.faceElem[f] = parent;
            for (int d = 0; d < 3; ++d) m.fA0[3 * static_cast<size_t>(f) + d] = rule.A0[d];
            for (int a = 0; a < nen; ++a)
                for (int d = 0; d < 3; ++d) Xe[3 * a + d] = m.X[3 * static_cast<size_t>(m.conn[static_cast<size_t>(parent) * nen + a]) + d];
            for (int q = 0; q < rule.nq; ++q) {
                double N[20], dN[60], J0[9], Ji[9];
                el->pointShapeFunc(rule.xi[q], N);
                el->pointShapeFuncLocalGradient(rule.xi[q], dN);
                jacobian0(nen, Xe, dN, J0);
                inv3(J0, Ji);
                const size_t base = (static_cast<size_t>(f) * m.fq + q) * nen;
                for (int a = 0; a < nen; ++a) {
                    m.fWN[base + a] = rule.w[q] * N[a];
                    // grad_X N_a = J0^-T dN_a
                    for (int i = 0; i < 3; ++i)
                        m.fG[(base + a) * 3 + i] = Ji[i] * dN[3 * a] + Ji[3 + i] * dN[3 * a + 1] + Ji[6 + i] * dN[3 * a + 2];
                }
            }
        }
    }
    // constant nodal forces
    m.fext.assign(3 * static_cast<size_t>(m.nFree), 0.0);
    if (pb.n_fnodes > 0 && pb.fnodes && pb.fvec)
        for (int k = 0; k < pb.n_fnodes; ++k) {
            const int n = pb.fnodes[k];
            if (n < 0 || n >= m.nNodes) {
                msg = "efem_problem: force node out of range";
                return false;
            }
            if (m.fid[n] >= 0)
                for (int d = 0; d < 3; ++d) m.fext[3 * static_cast<size_t>(m.fid[n]) + d] += pb.fvec[d];
        }
    // gravity: -M g over the free DOFs becomes the constant force M g
    // (InertiaConstraint::conEvaluteResidual, InertiaConstraint.cpp:34-49;
    // masses of ElementConstraint::conInitial, ElementConstraint.cpp:85-137:
    // consistent M_ab = V rho sum_q w_q N_a N_b, lumped max(V rho, 4e-12) /
    // nen on the diagonal; only free-free pairs carry mass)
    if (pb.gravity < 0 || pb.gravity > 6) {
        msg = "efem_problem: gravity must be 0 (none) or 1..6 (+x, -x, +y, -y, +z, -z)";
        return false;
    }
    if (pb.gravity != 0 && pb.density) {
        double g[3] = {0.0, 0.0, 0.0};
        g[(pb.gravity - 1) / 2] = pb.gravity % 2 == 1 ? 9.8 : -9.8;
        std::vector<double> Nq(static_cast<size_t>(ng) * nen);
        for (int q = 0; q < ng; ++q) {
            double xi[3];
            el->gaussPointLocalPosition(q, xi);
            el->pointShapeFunc(xi, &Nq[static_cast<size_t>(q) * nen]);
        }
        for (int e = 0; e < E; ++e) {
            const double rho = pb.density[pb.params_per_elem ? e : 0];
            const int* ce = &m.conn[static_cast<size_t>(e) * nen];
            if (pb.consistent_inertia) {
                for (int q = 0; q < ng; ++q) {
                    const double* N = &Nq[static_cast<size_t>(q) * nen];
                    double nFree = 0.0;  // sum of N_b over the free nodes
                    for (int b = 0; b < nen; ++b)
                        if (m.fid[ce[b]] >= 0) nFree += N[b];
                    const double wq = m.vol[e] * rho * m.gw[q];
                    for (int a = 0; a < nen; ++a) {
                        const int f = m.fid[ce[a]];
                        if (f < 0) continue;
                        for (int d = 0; d < 3; ++d) m.fext[3 * static_cast<size_t>(f) + d] += wq * N[a] * nFree * g[d];
                    }
                }
            } else {
                double mass = m.vol[e] * rho;
                if (mass < 4e-12) mass = 4e-12;
                mass /= nen;
                for (int a = 0; a < nen; ++a) {
                    const int f = m.fid[ce[a]];
                    if (f < 0) continue;
                    for (int d = 0; d < 3; ++d) m.fext[3 * static_cast<size_t>(f) + d] += mass * g[d];
                }
            }
        }
    }
    // gather CSR: free node -> element slots, then face slots, ascending
    m.gptr.assign(m.nFree + 1, 0);
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < nen; ++a) {
            const int f = m.fid[m.conn[static_cast<size_t>(e) * nen + a]];
            if (f >= 0) m.gptr[f + 1]++;
        }
    for (int fc = 0; fc < m.nFaces; ++fc)
        for (int a = 0; a < nen; ++a) {
            const int f = m.fid[m.conn[static_cast<size_t>(m.faceElem[fc]) * nen + a]];
            if (f >= 0) m.gptr[f + 1]++;
        }
    for (int i = 0; i < m.nFree; ++i) m.gptr[i + 1] += m.gptr[i];
    m.gslot.resize(m.gptr[m.nFree]);
    {
        std::vector<int> pos(m.gptr.begin(), m.gptr.end() - 1);
        for (int e = 0; e < E; ++e)
            for (int a = 0; a < nen; ++a) {
                const int f = m.fid[m.conn[static_cast<size_t>(e) * nen + a]];
                if (f >= 0) m.gslot[pos[f]++] = e * nen + a;
            }
        for (int fc = 0; fc < m.nFaces; ++fc)
            for (int a = 0; a < nen; ++a) {
                const int f = m.fid[m.conn[static_cast<size_t>(m.faceElem[fc]) * nen + a]];
                if (f >= 0) m.gslot[pos[f]++] = E * nen + fc * nen + a;
            }
    }
    return true;
}

// L_ij = sum_e sum_q V_e w_q c_e grad N_i . grad N_j over free nodes
// (ElementConstraint::conEvaluteJacobi, ElementConstraint.cpp:164-205,
// the I3 factor dropped)
void scalarLaplacian(const Model& m, Csc& L)
{
    // column j (a free node) collects, over its incident elements in element
    // order, V w c grad N_a . grad N_j for the element's free nodes a; each
    // column is independent, so columns run in parallel and every entry is
    // still summed in element order
    const int nf = m.nFree, nen = m.nen, E = m.nElems, ng = m.ng;
    std::vector<int> eptr(nf + 1, 0), elist;
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < nen; ++a) {
            const int j = m.connDof[static_cast<size_t>(a) * E + e];
            if (j < nf) eptr[j + 1]++;
        }
    for (int j = 0; j < nf; ++j) eptr[j + 1] += eptr[j];
    elist.resize(eptr[nf]);
    {
        std::vector<int> pos(eptr.begin(), eptr.end() - 1);
        for (int e = 0; e < E; ++e)
            for (int a = 0; a < nen; ++a) {
                const int j = m.connDof[static_cast<size_t>(a) * E + e];
                if (j < nf) elist[pos[j]++] = e * nen + a;  // element, local slot
            }
    }
    std::vector<std::vector<int>> ci(nf);
    std::vector<std::vector<double>> cx(nf);
#pragma omp parallel
    {
        std::vector<double> G(static_cast<size_t>(nen) * 3);
        std::vector<std::pair<int, double>> col;
        std::vector<double> kcol(nen);
#pragma omp for schedule(dynamic, 256)
        for (int j = 0; j < nf; ++j) {
            col.clear();
            for (int p = eptr[j]; p < eptr[j + 1]; ++p) {
                const int e = elist[p] / nen, b = elist[p] % nen;
                std::fill(kcol.begin(), kcol.end(), 0.0);
                const double ks = m.vol[e] * m.coeff[e];
                for (int q = 0; q < ng; ++q) {
                    double Ji[9];
                    for (int k = 0; k < 9; ++k) Ji[k] = m.ginv[(static_cast<size_t>(q) * 9 + k) * E + e];
                    const double* dN = &m.gdN[static_cast<size_t>(q) * nen * 3];
                    for (int a = 0; a < nen; ++a)
                        for (int i = 0; i < 3; ++i)
                            G[3 * a + i] = Ji[i] * dN[3 * a] + Ji[3 + i] * dN[3 * a + 1] + Ji[6 + i] * dN[3 * a + 2];
                    const double w = ks * m.gw[q];
                    for (int a = 0; a < nen; ++a)
                        kcol[a] += w * (G[3 * a] * G[3 * b] + G[3 * a + 1] * G[3 * b + 1] + G[3 * a + 2] * G[3 * b + 2]);
                }
                for (int a = 0; a < nen; ++a) {
                    const int ia = m.connDof[static_cast<size_t>(a) * E + e];
                    if (ia < nf) col.push_back({ia, kcol[a]});
                }
            }
            std::stable_sort(col.begin(), col.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            for (size_t k = 0; k < col.size();) {
                size_t k2 = k;
                double sum = 0.0;
                while (k2 < col.size() && col[k2].first == col[k].first) sum += col[k2++].second;
                ci[j].push_back(col[k].first);
                cx[j].push_back(sum);
                k = k2;
            }
        }
    }
    L.n = nf;
    L.p.assign(nf + 1, 0);
    for (int j = 0; j < nf; ++j) L.p[j + 1] = L.p[j] + static_cast<int>(ci[j].size());
    L.i.resize(L.p[nf]);
    L.x.resize(L.p[nf]);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nf; ++j) {
        std::copy(ci[j].begin(), ci[j].end(), L.i.begin() + L.p[j]);
        std::copy(cx[j].begin(), cx[j].end(), L.x.begin() + L.p[j]);
    }
}

void buildTet4Patches(const Model& m, int P, Tet4Patches& out)
{
    const int E = m.nElems, nen = m.nen;
    out = Tet4Patches{};
    out.P = P;
    if (E <= 0 || nen != 4 || P <= 0) return;
    std::vector<double> cen(3 * static_cast<size_t>(E));
    for (int e = 0; e < E; ++e)
        for (int d = 0; d < 3; ++d) {
            double s = 0.0;
            for (int a = 0; a < 4; ++a) s += m.X[3 * static_cast<size_t>(m.conn[4 * static_cast<size_t>(e) + a]) + d];
            cen[3 * static_cast<size_t>(e) + d] = 0.25 * s;
        }
    std::vector<int> idx(E);
    for (int e = 0; e < E; ++e) idx[e] = e;
    // bisection: [lo, hi) split at lo + P * floor(patches / 2) along the
    // longest extent of its centroids; a range of at most P is one patch
    std::vector<std::pair<int, int>> todo{{0, E}}, leaves;
    while (!todo.empty()) {
        const auto [lo, hi] = todo.back();
        todo.pop_back();
        const int n = hi - lo;
        if (n <= P) {
            leaves.push_back({lo, hi});
            continue;
        }
        double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
        for (int i = lo; i < hi; ++i)
            for (int d = 0; d < 3; ++d) {
                const double v = cen[3 * static_cast<size_t>(idx[i]) + d];
                mn[d] = std::min(mn[d], v);
                mx[d] = std::max(mx[d], v);
            }
        int ax = 0;
        for (int d = 1; d < 3; ++d)
            if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
        const int np = (n + P - 1) / P;
        const int mid = lo + P * (np / 2);
        std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](int a, int b) {
            const double va = cen[3 * static_cast<size_t>(a) + ax], vb = cen[3 * static_cast<size_t>(b) + ax];
            return va < vb || (va == vb && a < b);
        });
        todo.push_back({mid, hi});
        todo.push_back({lo, mid});
    }
    std::sort(leaves.begin(), leaves.end());
    for (const auto& [lo, hi] : leaves) std::sort(idx.begin() + lo, idx.begin() + hi);
    out.nPatch = static_cast<int>(leaves.size());
    out.order = idx;
    out.conn.resize(4 * static_cast<size_t>(E));
    out.vol.resize(E);
    for (int pe = 0; pe < E; ++pe) {
        const int e = idx[pe];
        for (int a = 0; a < 4; ++a) out.conn[static_cast<size_t>(a) * E + pe] = m.connDof[static_cast<size_t>(a) * E + e];
        out.vol[pe] = m.vol[e];
    }
    // records: per patch the free nodes it touches, ascending; entries
    // lane-major
    out.roff.assign(1, 0);
    std::vector<int> recNode;
    std::vector<std::pair<int, int>> pairs;  // (free id, 4 lane + a)
    for (int p = 0; p < out.nPatch; ++p) {
        const int lo = leaves[p].first, hi = leaves[p].second;
        pairs.clear();
        for (int pe = lo; pe < hi; ++pe)
            for (int a = 0; a < 4; ++a) {
                const int f = out.conn[static_cast<size_t>(a) * E + pe];
                if (f < m.nFree) pairs.push_back({f, 4 * (pe - lo) + a});
            }
        std::sort(pairs.begin(), pairs.end());
        for (size_t i = 0; i < pairs.size();) {
            size_t j = i;
            while (j < pairs.size() && pairs[j].first == pairs[i].first) ++j;
            const int first = static_cast<int>(out.ent.size());
            out.rec.push_back(first);
            out.rec.push_back(static_cast<int>(j - i));
            out.rec.push_back(0);
            out.rec.push_back(0);
            for (size_t k = i; k < j; ++k) out.ent.push_back(static_cast<uint16_t>(pairs[k].second));
            while (out.ent.size() % 8) out.ent.push_back(0xffff);
            recNode.push_back(pairs[i].first);
            i = j;
        }
        out.roff.push_back(static_cast<int>(recNode.size()));
    }
    out.nRec = static_cast<int>(recNode.size());
    // node-major positions of the records (patch order within a node)
    out.rptr.assign(m.nFree + 1, 0);
    for (int k = 0; k < out.nRec; ++k) out.rptr[recNode[k] + 1]++;
    for (int i = 0; i < m.nFree; ++i) out.rptr[i + 1] += out.rptr[i];
    {
        std::vector<int> pos(out.rptr.begin(), out.rptr.end() - 1);
        for (int k = 0; k < out.nRec; ++k) out.rec[4 * static_cast<size_t>(k) + 2] = pos[recNode[k]]++;
    }
    // face slots per free node, faces ascending
    out.fptr.assign(m.nFree + 1, 0);
    for (int fc = 0; fc < m.nFaces; ++fc)
        for (int a = 0; a < nen; ++a) {
            const int f = m.fid[m.conn[static_cast<size_t>(m.faceElem[fc]) * nen + a]];
            if (f >= 0) out.fptr[f + 1]++;
        }
    for (int i = 0; i < m.nFree; ++i) out.fptr[i + 1] += out.fptr[i];
    out.fslot.resize(out.fptr[m.nFree]);
    std::vector<int> pos(out.fptr.begin(), out.fptr.end() - 1);
    for (int fc = 0; fc < m.nFaces; ++fc)
        for (int a = 0; a < nen; ++a) {
            const int f = m.fid[m.conn[static_cast<size_t>(m.faceElem[fc]) * nen + a]];
            if (f >= 0) out.fslot[pos[f]++] = fc * nen + a;
        }
}

std::vector<double> freeCoords(const Model& m)
{
    std::vector<double> xyz(3 * static_cast<size_t>(m.nFree));
    for (int n = 0; n < m.nNodes; ++n)
        if (m.fid[n] >= 0)
            for (int d = 0; d < 3; ++d) xyz[3 * static_cast<size_t>(m.fid[n]) + d] = m.X[3 * static_cast<size_t>(n) + d];
    return xyz;
}

void laplacianMap(const Model& m, Csc& L, std::vector<int>& ptr, std::vector<int>& src)
{
    // the same column lists as scalarLaplacian, with each contribution named
    // (element e, row slot a, column slot b) -> e nen^2 + a nen + b instead
    // of valued; stable sorting keeps element order within an entry
    const int nf = m.nFree, nen = m.nen, E = m.nElems;
    std::vector<int> eptr(nf + 1, 0), elist;
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < nen; ++a) {
            const int j = m.connDof[static_cast<size_t>(a) * E + e];
            if (j < nf) eptr[j + 1]++;
        }
    for (int j = 0; j < nf; ++j) eptr[j + 1] += eptr[j];
    elist.resize(eptr[nf]);
    {
        std::vector<int> pos(eptr.begin(), eptr.end() - 1);
        for (int e = 0; e < E; ++e)
            for (int a = 0; a < nen; ++a) {
                const int j = m.connDof[static_cast<size_t>(a) * E + e];
                if (j < nf) elist[pos[j]++] = e * nen + a;
            }
    }
    std::vector<std::vector<int>> ci(nf), cp(nf), cs(nf);
#pragma omp parallel
    {
        std::vector<std::pair<int, int>> col;
#pragma omp for schedule(dynamic, 256)
        for (int j = 0; j < nf; ++j) {
            col.clear();
            for (int p = eptr[j]; p < eptr[j + 1]; ++p) {
                const int e = elist[p] / nen, b = elist[p] % nen;
                for (int a = 0; a < nen; ++a) {
                    const int ia = m.connDof[static_cast<size_t>(a) * E + e];
                    if (ia < nf) col.push_back({ia, (e * nen + a) * nen + b});
                }
            }
            std::stable_sort(col.begin(), col.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            for (size_t k = 0; k < col.size();) {
                size_t k2 = k;
                ci[j].push_back(col[k].first);
                cp[j].push_back(static_cast<int>(cs[j].size()));
                while (k2 < col.size() && col[k2].first == col[k].first) cs[j].push_back(col[k2++].second);
                k = k2;
            }
        }
    }
    L.n = nf;
    L.p.assign(nf + 1, 0);
    for (int j = 0; j < nf; ++j) L.p[j + 1] = L.p[j] + static_cast<int>(ci[j].size());
    L.i.resize(L.p[nf]);
    L.x.assign(L.p[nf], 0.0);
    std::vector<long long> soff(nf + 1, 0);
    for (int j = 0; j < nf; ++j) soff[j + 1] = soff[j] + static_cast<long long>(cs[j].size());
    ptr.resize(static_cast<size_t>(L.p[nf]) + 1);
    src.resize(static_cast<size_t>(soff[nf]));
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nf; ++j) {
        std::copy(ci[j].begin(), ci[j].end(), L.i.begin() + L.p[j]);
        for (size_t k = 0; k < cp[j].size(); ++k) ptr[L.p[j] + k] = static_cast<int>(soff[j] + cp[j][k]);
        std::copy(cs[j].begin(), cs[j].end(), src.begin() + soff[j]);
    }
    ptr[L.p[nf]] = static_cast<int>(soff[nf]);
}

bool factorLaplacian(const Model& m, const Csc& A, SupernodalFactor& F, std::string& msg, int denseSubtree,
                     HostWorkspace* ws, SupernodalSymbolic* symbolic)
{
    if (m.nFree == 0) {
        msg = "no free DOF";
        return false;
    }
    if (!symbolic) {
        const std::vector<int> perm = nestedDissection(A, freeCoords(m));
        return supernodalCholesky(A, perm, F, msg, denseSubtree, ws);
    }
    double symSeconds = 0.0;
    if (symbolic->n == 0) {
        const auto t0 = std::chrono::steady_clock::now();
        const std::vector<int> perm = nestedDissection(A, freeCoords(m));
        if (!supernodalSymbolic(A, perm, *symbolic, msg, denseSubtree)) {
            *symbolic = SupernodalSymbolic();
            return false;
        }
        symSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } else if (symbolic->n != A.n || symbolic->anz != static_cast<long long>(A.i.size()) ||
               A.x.size() != A.i.size()) {
        msg = "symbolic analysis does not match the matrix";
        return false;
    }
    if (!supernodalNumeric(A, *symbolic, F, msg, ws)) return false;
    F.seconds += symSeconds;
    return true;
}

bool factorModel(const Model& m, SupernodalFactor& F, std::string& msg, int denseSubtree, HostWorkspace* ws,
                 SupernodalSymbolic* symbolic)
{
    if (m.nFree == 0) {
        msg = "no free DOF";
        return false;
    }
    Csc A;
    scalarLaplacian(m, A);
    return factorLaplacian(m, A, F, msg, denseSubtree, ws, symbolic);
}

bool factorModels(const std::vector<Model>& parts, std::vector<SupernodalFactor>& F, std::string& msg,
                  int denseSubtree)
{
    const int np = static_cast<int>(parts.size());
    F.clear();
    F.resize(np);
    std::vector<std::string> msgs(np);
    std::vector<char> ok(np, 0);
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < np; ++i) ok[i] = factorModel(parts[i], F[i], msgs[i], denseSubtree) ? 1 : 0;
    for (int i = 0; i < np; ++i)
        if (!ok[i]) {
            msg = "instance " + std::to_string(i) + ": " + msgs[i];
            return false;
        }
    return true;
}

}  // namespace efem
