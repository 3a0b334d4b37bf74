// Element library (see element.hpp).  Shape functions are written from the
// standard Lagrange / serendipity families over the node layouts the
// reference uses; each class cites the reference file it stands in for.
#include "element.hpp"

#include <cmath>
#include <cstring>

namespace efem {
namespace {

// ---- nodal parametric coordinates (node layouts fixed by the reference) --

// C3D8: VTK hexahedron order (HexahedraElement.cpp:15-58)
constexpr double kHex8[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

// C3D20: the reference's parametrisation, node 0 at (1,-1,1)
// (HexahedraElement.cpp:147-238); in VTK (r,s,t): (xi,eta,zeta) = (-t,r,-s)
constexpr double kHex20[20][3] = {
    {1, -1, 1},  {1, 1, 1},  {1, 1, -1},  {1, -1, -1}, {-1, -1, 1}, {-1, 1, 1}, {-1, 1, -1},
    {-1, -1, -1}, {1, 0, 1}, {1, 1, 0},   {1, 0, -1},  {1, -1, 0},  {-1, 0, 1}, {-1, 1, 0},
    {-1, 0, -1}, {-1, -1, 0}, {0, -1, 1}, {0, 1, 1},   {0, 1, -1},  {0, -1, -1}};

// C3D10 edge nodes as barycentric pairs, l0 = 1-xi-eta-zeta, l1 = xi,
// l2 = eta, l3 = zeta (TetrahedraElement.cpp:62-79)
constexpr int kTet10Edge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

// ---- Gauss rules (IntegralElement.inl:3-223) ------------------------------

void tetRule(int n, int q, double xi[3], double& w)
{
    if (n == 1) {
        xi[0] = xi[1] = xi[2] = 0.25;
        w = 1.0;
        return;
    }
    // the reference's truncated 8-digit constants
    const double a = 0.58541020, b = 0.13819660;
    xi[0] = xi[1] = xi[2] = b;
    if (q < 3) xi[q] = a;
    w = 0.25;
}

void hexRule(int n, int q, double xi[3], double& w)
{
    if (n == 1) {
        xi[0] = xi[1] = xi[2] = 0.0;
        w = 1.0;
        return;
    }
    if (n == 8) {
        // point q takes the sign pattern of C3D8 node q
        const double g = 1.00 / std::sqrt(3.0);
        for (int d = 0; d < 3; ++d) xi[d] = kHex8[q][d] * g;
        w = 0.125;
        return;
    }
    // 3x3x3 with xi varying fastest; weights {5,8,5}^3 / 5832
    const double r = std::sqrt(3.00 / 5.00);
    const double pt[3] = {-r, 0.0, r};
    const int wt[3] = {5, 8, 5};
    const int i = q % 3, j = (q / 3) % 3, k = q / 9;
    xi[0] = pt[i];
    xi[1] = pt[j];
    xi[2] = pt[k];
    w = double(wt[i] * wt[j] * wt[k]) / 5832.0;
}

// ---- element classes -----------------------------------------------------

class Tet4 : public IntegralElement {  // TetElement3D4 (TetrahedraElement.cpp:4-43)
public:
    explicit Tet4(bool reduced) { type = reduced ? ElementType::C3D4R : ElementType::C3D4; }
    int nodeNum() const override { return 4; }
    void pointShapeFunc(const double xi[3], double* N) const override
    {
        N[0] = xi[0];
        N[1] = xi[1];
        N[2] = xi[2];
        N[3] = 1.0 - xi[0] - xi[1] - xi[2];
    }
    void pointShapeFuncLocalGradient(const double*, double* dN) const override
    {
        static const double g[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, -1, -1, -1};
        std::memcpy(dN, g, sizeof(g));
    }
    int gaussPointNum() const override { return 1; }
    double gaussPointWeight(int) const override { return 1.0; }
    void gaussPointLocalPosition(int q, double xi[3]) const override
    {
        double w;
        tetRule(1, q, xi, w);
    }
};

class Tet10 : public IntegralElement {  // TetElement3D10 / 3D10R (TetrahedraElement.cpp:51-215)
public:
    explicit Tet10(bool reduced) : nq_(reduced ? 1 : 4)
    {
        type = reduced ? ElementType::C3D10R : ElementType::C3D10;
    }
    int nodeNum() const override { return 10; }
    void pointShapeFunc(const double xi[3], double* N) const override
    {
        const double l[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
        for (int i = 0; i < 4; ++i) N[i] = l[i] * (2.0 * l[i] - 1.0);
        for (int e = 0; e < 6; ++e) N[4 + e] = 4.0 * l[kTet10Edge[e][0]] * l[kTet10Edge[e][1]];
    }
    void pointShapeFuncLocalGradient(const double xi[3], double* dN) const override
    {
        const double l[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
        static const double dl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int i = 0; i < 4; ++i)
            for (int d = 0; d < 3; ++d) dN[3 * i + d] = (4.0 * l[i] - 1.0) * dl[i][d];
        for (int e = 0; e < 6; ++e) {
            const int a = kTet10Edge[e][0], b = kTet10Edge[e][1];
            for (int d = 0; d < 3; ++d) dN[3 * (4 + e) + d] = 4.0 * (l[a] * dl[b][d] + l[b] * dl[a][d]);
        }
    }
    int gaussPointNum() const override { return nq_; }
    double gaussPointWeight(int q) const override
    {
        double xi[3], w;
        tetRule(nq_, q, xi, w);
        return w;
    }
    void gaussPointLocalPosition(int q, double xi[3]) const override
    {
        double w;
        tetRule(nq_, q, xi, w);
    }

private:
    int nq_;
};

class Hex8 : public IntegralElement {  // HexElement3D8 / 3D8R (HexahedraElement.cpp:4-134)
public:
    explicit Hex8(bool reduced) : nq_(reduced ? 1 : 8)
    {
        type = reduced ? ElementType::C3D8R : ElementType::C3D8;
    }
    int nodeNum() const override { return 8; }
    void pointShapeFunc(const double xi[3], double* N) const override
    {
        for (int a = 0; a < 8; ++a)
            N[a] = 0.125 * (1.0 + kHex8[a][0] * xi[0]) * (1.0 + kHex8[a][1] * xi[1]) *
                   (1.0 + kHex8[a][2] * xi[2]);
    }
    void pointShapeFuncLocalGradient(const double xi[3], double* dN) const override
    {
        for (int a = 0; a < 8; ++a) {
            const double f0 = 1.0 + kHex8[a][0] * xi[0], f1 = 1.0 + kHex8[a][1] * xi[1],
                         f2 = 1.0 + kHex8[a][2] * xi[2];
            dN[3 * a + 0] = 0.125 * kHex8[a][0] * f1 * f2;
            dN[3 * a + 1] = 0.125 * f0 * kHex8[a][1] * f2;
            dN[3 * a + 2] = 0.125 * f0 * f1 * kHex8[a][2];
        }
    }
    int gaussPointNum() const override { return nq_; }
    double gaussPointWeight(int q) const override
    {
        double xi[3], w;
        hexRule(nq_, q, xi, w);
        return w;
    }
    void gaussPointLocalPosition(int q, double xi[3]) const override
    {
        double w;
        hexRule(nq_, q, xi, w);
    }

private:
    int nq_;
};

class Hex20 : public IntegralElement {  // HexElement3D20 / 3D20R (HexahedraElement.cpp:136-361)
public:
    explicit Hex20(bool reduced) : nq_(reduced ? 8 : 27)
    {
        type = reduced ? ElementType::C3D20R : ElementType::C3D20;
    }
    int nodeNum() const override { return 20; }
    void pointShapeFunc(const double xi[3], double* N) const override
    {
        for (int a = 0; a < 20; ++a) {
            const double* s = kHex20[a];
            if (a < 8) {
                N[a] = (1.0 + s[0] * xi[0]) * (1.0 + s[1] * xi[1]) * (1.0 + s[2] * xi[2]) *
                       (s[0] * xi[0] + s[1] * xi[1] + s[2] * xi[2] - 2.0) / 8.0;
            } else {
                const int k = s[0] == 0.0 ? 0 : (s[1] == 0.0 ? 1 : 2);
                const int i = (k + 1) % 3, j = (k + 2) % 3;
                N[a] = (1.0 - xi[k] * xi[k]) * (1.0 + s[i] * xi[i]) * (1.0 + s[j] * xi[j]) / 4.0;
            }
        }
    }
    void pointShapeFuncLocalGradient(const double xi[3], double* dN) const override
    {
        for (int a = 0; a < 20; ++a) {
            const double* s = kHex20[a];
            if (a < 8) {
                const double f[3] = {1.0 + s[0] * xi[0], 1.0 + s[1] * xi[1], 1.0 + s[2] * xi[2]};
                const double g = s[0] * xi[0] + s[1] * xi[1] + s[2] * xi[2] - 2.0;
                for (int d = 0; d < 3; ++d) {
                    const double other = f[(d + 1) % 3] * f[(d + 2) % 3];
                    dN[3 * a + d] = s[d] * other * (g + f[d]) / 8.0;
                }
            } else {
                const int k = s[0] == 0.0 ? 0 : (s[1] == 0.0 ? 1 : 2);
                const int i = (k + 1) % 3, j = (k + 2) % 3;
                const double q = 1.0 - xi[k] * xi[k];
                const double fi = 1.0 + s[i] * xi[i], fj = 1.0 + s[j] * xi[j];
                dN[3 * a + k] = -2.0 * xi[k] * fi * fj / 4.0;
                dN[3 * a + i] = q * s[i] * fj / 4.0;
                dN[3 * a + j] = q * fi * s[j] / 4.0;
            }
        }
    }
    int gaussPointNum() const override { return nq_; }
    double gaussPointWeight(int q) const override
    {
        double xi[3], w;
        hexRule(nq_, q, xi, w);
        return w;
    }
    void gaussPointLocalPosition(int q, double xi[3]) const override
    {
        double w;
        hexRule(nq_, q, xi, w);
    }

private:
    int nq_;
};

// ---- surface tessellation for the element volume -------------------------
constexpr int kTetFaces[4][3] = {{0, 1, 3}, {1, 2, 3}, {2, 0, 3}, {0, 2, 1}};
constexpr int kHexFaces[6][4] = {{0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4},
                                 {3, 7, 6, 2}, {0, 3, 2, 1}, {4, 5, 6, 7}};
constexpr int kHex20Edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                   {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

int midNode(int nen, int a, int b)
{
    if (nen == 10) {
        for (int e = 0; e < 6; ++e)
            if ((kTet10Edge[e][0] == a && kTet10Edge[e][1] == b) ||
                (kTet10Edge[e][0] == b && kTet10Edge[e][1] == a))
                return 4 + e;
    } else if (nen == 20) {
        for (int e = 0; e < 12; ++e)
            if ((kHex20Edge[e][0] == a && kHex20Edge[e][1] == b) ||
                (kHex20Edge[e][0] == b && kHex20Edge[e][1] == a))
                return 8 + e;
    }
    return -1;
}

// signed 6x volume of the tetrahedron (origin, a, b, c)
double triple(const double* X, int a, int b, int c)
{
    const double* pa = X + 3 * a;
    const double* pb = X + 3 * b;
    const double* pc = X + 3 * c;
    return pa[0] * (pb[1] * pc[2] - pb[2] * pc[1]) + pa[1] * (pb[2] * pc[0] - pb[0] * pc[2]) +
           pa[2] * (pb[0] * pc[1] - pb[1] * pc[0]);
}

void cross(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

std::unique_ptr<IntegralElement> makeElement(int nen, bool reduced)
{
    switch (nen) {
        case 4: return std::make_unique<Tet4>(reduced);
        case 10: return std::make_unique<Tet10>(reduced);
        case 8: return std::make_unique<Hex8>(reduced);
        case 20: return std::make_unique<Hex20>(reduced);
    }
    return nullptr;
}

double elementVolume(int nen, const double* Xe)
{
    // the tessellation is translation invariant; shift to node 0 for accuracy
    double Y[60];
    for (int a = 0; a < nen; ++a)
        for (int d = 0; d < 3; ++d) Y[3 * a + d] = Xe[3 * a + d] - Xe[d];
    double s = 0.0;
    if (nen == 4 || nen == 10) {
        for (const auto& f : kTetFaces) {
            const int a = f[0], b = f[1], c = f[2];
            if (nen == 4) {
                s += triple(Y, a, b, c);
            } else {
                const int ab = midNode(nen, a, b), bc = midNode(nen, b, c), ca = midNode(nen, c, a);
                s += triple(Y, a, ab, ca) + triple(Y, ab, b, bc) + triple(Y, ca, bc, c) +
                     triple(Y, ab, bc, ca);
            }
        }
    } else {
        for (const auto& f : kHexFaces) {
            const int a = f[0], b = f[1], c = f[2], d = f[3];
            if (nen == 8) {
                s += triple(Y, a, b, c) + triple(Y, a, c, d);
            } else {
                const int ab = midNode(nen, a, b), bc = midNode(nen, b, c), cd = midNode(nen, c, d),
                          da = midNode(nen, d, a);
                s += triple(Y, a, ab, da) + triple(Y, b, bc, ab) + triple(Y, c, cd, bc) +
                     triple(Y, d, da, cd) + triple(Y, ab, bc, cd) + triple(Y, ab, cd, da);
            }
        }
    }
    return std::fabs(s) / 6.0;
}

bool buildFaceRule(int nen, const int* en, int fnn, const int* fn, const double* X, FaceRule& out,
                   std::string& msg)
{
    bool absent[20];
    for (int a = 0; a < nen; ++a) {
        absent[a] = true;
        for (int i = 0; i < fnn; ++i)
            if (fn[i] == en[a]) absent[a] = false;
    }
    auto P = [&](int node) { return X + 3 * static_cast<size_t>(node); };
    auto orient = [&](double* A, int toward) {
        double t = 0;
        for (int d = 0; d < 3; ++d) t += (P(en[toward])[d] - P(fn[0])[d]) * A[d];
        if (t < 0)
            for (int d = 0; d < 3; ++d) A[d] = -A[d];
    };
    if (fnn == 3 || fnn == 6) {
        if ((fnn == 3 && nen != 4) || (fnn == 6 && nen != 10)) {
            msg = "pressure face node count does not match its parent element";
            return false;
        }
        int k = -1;
        for (int a = 0; a < 4; ++a)
            if (absent[a]) {
                k = a;
                break;
            }
        if (k < 0) {
            msg = "degenerate pressure face";
            return false;
        }
        // A0 = 1/2 (x2 - x0) x (x1 - x0), toward the vertex off the face
        double v1[3], v2[3];
        for (int d = 0; d < 3; ++d) {
            v1[d] = P(fn[2])[d] - P(fn[0])[d];
            v2[d] = P(fn[1])[d] - P(fn[0])[d];
        }
        cross(v1, v2, out.A0);
        for (double& a : out.A0) a *= 0.5;
        orient(out.A0, k);
        if (fnn == 3) {
            // centroid of the face (ExternalConstraint.cpp:292-337)
            out.nq = 1;
            out.w[0] = 1.0;
            for (int d = 0; d < 3; ++d) out.xi[0][d] = 1.0 / 3.0;
            if (k < 3) out.xi[0][k] = 0.0;
        } else {
            // (2/3, 1/6, 1/6) over the face vertices (:338-409): the vertex
            // coordinates are l1..l3 = xi, eta, zeta and l0 = 1 - sum
            static const int order[4][3] = {{1, 2, 3}, {2, 3, 0}, {1, 3, 0}, {1, 2, 0}};
            out.nq = 3;
            for (int q = 0; q < 3; ++q) {
                double l[4] = {0, 0, 0, 0};
                for (int t = 0; t < 3; ++t) l[order[k][t]] = (t == q) ? 2.0 / 3.0 : 1.0 / 6.0;
                out.xi[q][0] = l[1];
                out.xi[q][1] = l[2];
                out.xi[q][2] = l[3];
                out.w[q] = 1.0 / 3.0;
            }
        }
        return true;
    }
    if (fnn == 4 || fnn == 8) {
        if ((fnn == 4 && nen != 8) || (fnn == 8 && nen != 20)) {
            msg = "pressure face node count does not match its parent element";
            return false;
        }
        // case = pair of parent nodes missing from the face -> fixed axis,
        // its value, and the node A0 is oriented toward
        // (:419-563 for C3D8 faces, :580-814 for C3D20 faces)
        static const int q4[6][5] = {{0, 2, 2, 1, 0},  {0, 5, 1, 1, 0},  {0, 7, 0, 1, 0},
                                     {6, 3, 1, -1, 6}, {6, 1, 0, -1, 6}, {6, 4, 2, -1, 6}};
        static const int q8[6][5] = {{0, 2, 0, -1, 0}, {0, 5, 2, -1, 0}, {0, 7, 1, 1, 0},
                                     {6, 3, 2, 1, 6},  {6, 1, 1, -1, 6}, {6, 4, 0, 1, 6}};
        const int(*cs)[5] = fnn == 4 ? q4 : q8;
        int ci = -1;
        for (int i = 0; i < 6; ++i)
            if (absent[cs[i][0]] && absent[cs[i][1]]) {
                ci = i;
                break;
            }
        if (ci < 0) {
            msg = "pressure face is not a face of its parent element";
            return false;
        }
        const int axis = cs[ci][2], toward = cs[ci][4];
        const double val = cs[ci][3];
        const int f0 = axis == 0 ? 1 : 0;
        const int f1 = axis == 2 ? 1 : 2;
        if (fnn == 4) {
            const double s = std::sqrt(1.0 / 3.0);
            static const int a0[4] = {-1, 1, 1, -1}, a1[4] = {-1, -1, 1, 1};
            out.nq = 4;
            for (int q = 0; q < 4; ++q) {
                out.xi[q][axis] = val;
                out.xi[q][f0] = a0[q] * s;
                out.xi[q][f1] = a1[q] * s;
                out.w[q] = 0.25;
            }
        } else {
            const double r = std::sqrt(0.6);
            const double pt[3] = {-r, 0.0, r};
            static const double wf[9] = {25, 40, 25, 40, 64, 40, 25, 40, 25};
            out.nq = 9;
            for (int q = 0; q < 9; ++q) {
                out.xi[q][axis] = val;
                out.xi[q][f0] = pt[q / 3];
                out.xi[q][f1] = pt[q % 3];
                out.w[q] = 0.25 * (wf[q] / 81.0);
            }
        }
        // two triangles over the first four face nodes, each oriented inward
        double A01[3], A02[3], v1[3], v2[3];
        for (int d = 0; d < 3; ++d) {
            v1[d] = P(fn[2])[d] - P(fn[0])[d];
            v2[d] = P(fn[1])[d] - P(fn[0])[d];
        }
        cross(v1, v2, A01);
        for (int d = 0; d < 3; ++d) {
            v1[d] = P(fn[0])[d] - P(fn[3])[d];
            v2[d] = P(fn[2])[d] - P(fn[3])[d];
        }
        cross(v1, v2, A02);
        for (int d = 0; d < 3; ++d) {
            A01[d] *= 0.5;
            A02[d] *= 0.5;
        }
        orient(A01, toward);
        orient(A02, toward);
        for (int d = 0; d < 3; ++d) out.A0[d] = A01[d] + A02[d];
        return true;
    }
    msg = "unsupported pressure face node count";
    return false;
}

}  // namespace efem
