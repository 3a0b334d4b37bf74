// Host preprocessing (see model.hpp).
#include "model.hpp"

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
    const size_t np = m.paramsPerElem ? 4 * static_cast<size_t>(E) : 4;
    m.params.assign(pb.params, pb.params + np);
    for (size_t k = 0; k < np; k += 4)
        if (!(m.params[k + 3] > 0.0)) {
            msg = "efem_problem: D1 must be positive";
            return false;
        }
    m.coeff.resize(E);
    for (int e = 0; e < E; ++e)
        m.coeff[e] = elasticityCoeff(m.law, &m.params[m.paramsPerElem ? 4 * static_cast<size_t>(e) : 0]);
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
    const int nf = m.nFree, nen = m.nen, E = m.nElems, ng = m.ng;
    std::vector<int> cnt(nf + 1, 0);
    for (int e = 0; e < E; ++e)
        for (int b = 0; b < nen; ++b) {
            const int jb = m.connDof[static_cast<size_t>(b) * E + e];
            if (jb >= nf) continue;
            for (int a = 0; a < nen; ++a)
                if (m.connDof[static_cast<size_t>(a) * E + e] < nf) cnt[jb + 1]++;
        }
    for (int j = 0; j < nf; ++j) cnt[j + 1] += cnt[j];
    std::vector<int> ri(cnt[nf]);
    std::vector<double> rx(cnt[nf]);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    std::vector<double> G(static_cast<size_t>(nen) * 3), Ke(static_cast<size_t>(nen) * nen);
    for (int e = 0; e < E; ++e) {
        std::fill(Ke.begin(), Ke.end(), 0.0);
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
                for (int b = 0; b < nen; ++b)
                    Ke[a * nen + b] += w * (G[3 * a] * G[3 * b] + G[3 * a + 1] * G[3 * b + 1] + G[3 * a + 2] * G[3 * b + 2]);
        }
        for (int b = 0; b < nen; ++b) {
            const int jb = m.connDof[static_cast<size_t>(b) * E + e];
            if (jb >= nf) continue;
            for (int a = 0; a < nen; ++a) {
                const int ia = m.connDof[static_cast<size_t>(a) * E + e];
                if (ia >= nf) continue;
                ri[pos[jb]] = ia;
                rx[pos[jb]++] = Ke[a * nen + b];
            }
        }
    }
    L.n = nf;
    L.p.assign(nf + 1, 0);
    L.i.clear();
    L.x.clear();
    std::vector<std::pair<int, double>> col;
    for (int j = 0; j < nf; ++j) {
        col.clear();
        for (int k = cnt[j]; k < cnt[j + 1]; ++k) col.push_back({ri[k], rx[k]});
        std::stable_sort(col.begin(), col.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < col.size();) {
            size_t k2 = k;
            double s = 0.0;
            while (k2 < col.size() && col[k2].first == col[k].first) s += col[k2++].second;
            L.i.push_back(col[k].first);
            L.x.push_back(s);
            k = k2;
        }
        L.p[j + 1] = static_cast<int>(L.i.size());
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

}  // namespace efem
