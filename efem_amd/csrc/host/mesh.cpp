// Synthetic workload generators (see mesh.hpp) and their C ABI.
#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>

#include "../../../include/efem_gpu.h"

namespace efem {

double SplitMix64::normal()
{
    // Box-Muller on two 53-bit uniforms; u1 in (0,1]
    const double u1 = 1.0 - uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

namespace {

// local faces in VTK order, outward for a positively oriented cell
constexpr int kTetFaces[4][3] = {{0, 1, 3}, {1, 2, 3}, {2, 0, 3}, {0, 2, 1}};
constexpr int kHexFaces[6][4] = {{0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4},
                                 {3, 7, 6, 2}, {0, 3, 2, 1}, {4, 5, 6, 7}};
constexpr int kTet10Edge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
constexpr int kHex20Edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                   {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
constexpr int kHexCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

int edgeMid(int nen, int a, int b)
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

// Builds a structured mesh on a "doubled" lattice (corner nodes at even
// lattice coordinates, mid-edge nodes at the lattice midpoints) over cells
// (i,j,k); `pos` maps a doubled lattice coordinate to a position, `wrapI`
// makes the first axis periodic (cylinder).  Node ids follow the lattice
// order with the LAST axis fastest and the first slowest.
struct Lattice {
    int ni, nj, nk;  // cells
    bool wrapI = false;
    int li() const { return wrapI ? 2 * ni : 2 * ni + 1; }
    int lj() const { return 2 * nj + 1; }
    int lk() const { return 2 * nk + 1; }
    long idx(int I, int J, int K) const
    {
        if (wrapI) I = ((I % li()) + li()) % li();
        return (static_cast<long>(I) * lj() + J) * lk() + K;
    }
};

Mesh buildLattice(const Lattice& L, int kind, const std::function<void(int, int, int, double*)>& pos)
{
    Mesh m;
    m.nen = kind;
    const long total = static_cast<long>(L.li()) * L.lj() * L.lk();
    std::vector<int> gid(total, -1);
    std::vector<long> elemPts;  // doubled-lattice index per element node
    // Kuhn permutations: path (0,0,0) -> +e_a -> +e_b -> (1,1,1)
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int i = 0; i < L.ni; ++i)
        for (int j = 0; j < L.nj; ++j)
            for (int k = 0; k < L.nk; ++k) {
                if (kind == 8 || kind == 20) {
                    int c[20][3];
                    for (int a = 0; a < 8; ++a) {
                        c[a][0] = 2 * (i + kHexCorner[a][0]);
                        c[a][1] = 2 * (j + kHexCorner[a][1]);
                        c[a][2] = 2 * (k + kHexCorner[a][2]);
                    }
                    int n = 8;
                    if (kind == 20) {
                        for (int e = 0; e < 12; ++e)
                            for (int d = 0; d < 3; ++d)
                                c[8 + e][d] = (c[kHex20Edge[e][0]][d] + c[kHex20Edge[e][1]][d]) / 2;
                        n = 20;
                    }
                    for (int a = 0; a < n; ++a) elemPts.push_back(L.idx(c[a][0], c[a][1], c[a][2]));
                } else {
                    for (const auto& p : perms) {
                        int v[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}};
                        v[1][p[0]] = 1;
                        v[2][p[0]] = 1;
                        v[2][p[1]] = 1;
                        // sign of the permutation decides the orientation
                        const int inv = (p[0] > p[1]) + (p[0] > p[2]) + (p[1] > p[2]);
                        int order[4] = {0, 1, 2, 3};
                        if (inv & 1) std::swap(order[1], order[2]);
                        int c[10][3];
                        for (int a = 0; a < 4; ++a) {
                            c[a][0] = 2 * (i + v[order[a]][0]);
                            c[a][1] = 2 * (j + v[order[a]][1]);
                            c[a][2] = 2 * (k + v[order[a]][2]);
                        }
                        int n = 4;
                        if (kind == 10) {
                            for (int e = 0; e < 6; ++e)
                                for (int d = 0; d < 3; ++d)
                                    c[4 + e][d] = (c[kTet10Edge[e][0]][d] + c[kTet10Edge[e][1]][d]) / 2;
                            n = 10;
                        }
                        for (int a = 0; a < n; ++a) elemPts.push_back(L.idx(c[a][0], c[a][1], c[a][2]));
                    }
                }
            }
    for (long p : elemPts) gid[p] = 0;
    int nn = 0;
    for (long p = 0; p < total; ++p)
        if (gid[p] == 0) gid[p] = nn++ + 1;  // 1-based while marking
    m.X.resize(3 * static_cast<size_t>(nn));
    for (int I = 0; I < L.li(); ++I)
        for (int J = 0; J < L.lj(); ++J)
            for (int K = 0; K < L.lk(); ++K) {
                const long p = L.idx(I, J, K);
                if (gid[p] > 0) pos(I, J, K, &m.X[3 * static_cast<size_t>(gid[p] - 1)]);
            }
    m.conn.resize(elemPts.size());
    for (size_t t = 0; t < elemPts.size(); ++t) m.conn[t] = gid[elemPts[t]] - 1;
    m.fixed.assign(nn, 0);
    return m;
}

// element faces all of whose nodes satisfy `on`
void extractFaces(Mesh& m, const std::function<bool(const double*)>& on, int fnnCheck)
{
    const int nen = m.nen;
    const bool tet = nen == 4 || nen == 10;
    const bool quad = nen == 10 || nen == 20;
    m.fnn = tet ? (quad ? 6 : 3) : (quad ? 8 : 4);
    (void)fnnCheck;
    m.faces.clear();
    for (int e = 0; e < m.nElems(); ++e) {
        const int* c = &m.conn[static_cast<size_t>(e) * nen];
        const int nf = tet ? 4 : 6;
        for (int f = 0; f < nf; ++f) {
            int corners[4];
            const int nc = tet ? 3 : 4;
            for (int t = 0; t < nc; ++t) corners[t] = tet ? kTetFaces[f][t] : kHexFaces[f][t];
            bool all = true;
            for (int t = 0; t < nc && all; ++t) all = on(&m.X[3 * static_cast<size_t>(c[corners[t]])]);
            if (!all) continue;
            for (int t = 0; t < nc; ++t) m.faces.push_back(c[corners[t]]);
            if (quad)
                for (int t = 0; t < nc; ++t) m.faces.push_back(c[edgeMid(nen, corners[t], corners[(t + 1) % nc])]);
        }
    }
}

}  // namespace

Mesh makeBox(const BoxSpec& s)
{
    Lattice L{s.nx, s.ny, s.nz};
    const double hx = s.lx / s.nx, hy = s.ly / s.ny, hz = s.lz / s.nz;
    const int kind = s.kind == 20 ? 20 : s.kind;
    Mesh m = buildLattice(L, kind, [&](int I, int J, int K, double* x) {
        x[0] = 0.5 * I * hx;
        x[1] = 0.5 * J * hy;
        x[2] = 0.5 * K * hz;
    });
    if (s.jitter > 0.0) {
        // interior corner nodes move by U(-j h, j h) per coordinate
        SplitMix64 rng(s.seed);
        const double tol = 1e-12;
        for (int n = 0; n < m.nNodes(); ++n) {
            double* x = &m.X[3 * static_cast<size_t>(n)];
            const bool interior = x[0] > tol && x[0] < s.lx - tol * s.lx && x[1] > tol &&
                                  x[1] < s.ly - tol * s.ly && x[2] > tol && x[2] < s.lz - tol * s.lz;
            const double du = (2.0 * rng.uniform() - 1.0) * s.jitter * hx;
            const double dv = (2.0 * rng.uniform() - 1.0) * s.jitter * hy;
            const double dw = (2.0 * rng.uniform() - 1.0) * s.jitter * hz;
            if (interior) {
                x[0] += du;
                x[1] += dv;
                x[2] += dw;
            }
        }
    }
    return m;
}

static bool near(double a, double b, double scale) { return std::fabs(a - b) <= 1e-9 * scale; }

void fixPlane(Mesh& m, int axis, double value)
{
    double scale = 0;
    for (double v : m.X) scale = std::max(scale, std::fabs(v));
    for (int n = 0; n < m.nNodes(); ++n)
        if (near(m.X[3 * static_cast<size_t>(n) + axis], value, scale)) m.fixed[n] = 1;
}

void pressurePlane(Mesh& m, int axis, double value, double p, int raxis, double lo, double hi)
{
    double scale = 0;
    for (double v : m.X) scale = std::max(scale, std::fabs(v));
    extractFaces(
        m,
        [&](const double* x) {
            if (!near(x[axis], value, scale)) return false;
            if (raxis >= 0 && (x[raxis] < lo - 1e-9 * scale || x[raxis] > hi + 1e-9 * scale)) return false;
            return true;
        },
        0);
    m.pressure = p;
}

Mesh makeWorkload(int which, const double* opts, int n_opts)
{
    auto opt = [&](int i, double d) { return (opts && i < n_opts && opts[i] > 0) ? opts[i] : d; };
    if (which == 1) {
        // cantilever 0.10 x 0.01 x 0.01, 20x4x4 C3D8, NH, 2 kPa on z = max
        BoxSpec s;
        s.nx = static_cast<int>(opt(0, 20));
        s.ny = static_cast<int>(opt(1, 4));
        s.nz = static_cast<int>(opt(2, 4));
        s.lx = 0.005 * s.nx;
        s.ly = 0.0025 * s.ny;
        s.lz = 0.0025 * s.nz;
        s.kind = 8;
        Mesh m = makeBox(s);
        m.law = EFEM_NEOHOOKEAN;
        m.params = {114000.0, 0.0, 0.0, 1e-6};
        fixPlane(m, 0, 0.0);
        pressurePlane(m, 2, s.lz, opt(4, 2000.0));
        return m;
    }
    if (which == 2) {
        // beam 0.2 x 0.02 x 0.02, 200x20x20 hexes -> 6 Kuhn C3D4 each,
        // interior nodes jittered +-10 % of h, MR, 3 kPa on z = max
        BoxSpec s;
        s.nx = static_cast<int>(opt(0, 200));
        s.ny = static_cast<int>(opt(1, 20));
        s.nz = static_cast<int>(opt(2, 20));
        s.lx = 0.001 * s.nx;
        s.ly = 0.001 * s.ny;
        s.lz = 0.001 * s.nz;
        s.kind = 4;
        s.jitter = 0.1;
        s.seed = static_cast<uint64_t>(opt(3, 20240531));
        Mesh m = makeBox(s);
        m.law = EFEM_MOONEY_RIVLIN;
        m.params = {80000.0, 20000.0, 0.0, 1e-6};
        fixPlane(m, 0, 0.0);
        pressurePlane(m, 2, s.lz, opt(4, 3000.0));
        return m;
    }
    if (which == 3) {
        // hollow cylinder r in [10,15] mm, h 20 mm; 48 (theta) x 4 (r) x 40 (z)
        // C3D20R; mid-side nodes on the arcs; Yeoh; 10 kPa inside; z = 0 fixed
        const int nt = static_cast<int>(opt(0, 48)), nr = static_cast<int>(opt(1, 4)),
                  nzc = static_cast<int>(opt(2, 40));
        const double ri = 0.010, ro = 0.015, h = 0.020;
        Lattice L{nr, nt, nzc};
        L.wrapI = false;
        // axes: I -> r (radial, not periodic), J -> theta, K -> z.  The lattice
        // is built with theta as the middle axis and closed by merging below.
        Lattice Lp{nr, nt, nzc};
        Mesh m = buildLattice(Lp, 20, [&](int I, int J, int K, double* x) {
            const double r = ri + 0.5 * I * (ro - ri) / nr;
            const double t = 0.5 * J * 6.283185307179586 / nt;
            x[0] = r * std::cos(t);
            x[1] = r * std::sin(t);
            x[2] = 0.5 * K * h / nzc;
        });
        // close the seam theta = 2 pi onto theta = 0: merge coincident nodes
        {
            const int nn = m.nNodes();
            std::vector<int> remap(nn);
            for (int n = 0; n < nn; ++n) remap[n] = n;
            // nodes at theta = 2pi have lattice J = 2 nt; find partner with J = 0
            // by position (both lie on the same (r, z))
            std::vector<std::pair<long, int>> keyed;
            keyed.reserve(nn);
            auto key = [&](const double* x) {
                const long a = std::lround(x[0] * 1e8), b = std::lround(x[1] * 1e8),
                           c = std::lround(x[2] * 1e8);
                return (a * 1000003L + b) * 1000033L + c;
            };
            for (int n = 0; n < nn; ++n) keyed.push_back({key(&m.X[3 * static_cast<size_t>(n)]), n});
            std::sort(keyed.begin(), keyed.end());
            for (size_t t = 1; t < keyed.size(); ++t)
                if (keyed[t].first == keyed[t - 1].first) {
                    const int a = keyed[t - 1].second, b = keyed[t].second;
                    remap[std::max(a, b)] = std::min(a, b);
                }
            std::vector<int> newid(nn, -1);
            int cnt = 0;
            for (int n = 0; n < nn; ++n)
                if (remap[n] == n) newid[n] = cnt++;
            std::vector<double> X2(3 * static_cast<size_t>(cnt));
            for (int n = 0; n < nn; ++n)
                if (remap[n] == n)
                    for (int d = 0; d < 3; ++d) X2[3 * static_cast<size_t>(newid[n]) + d] = m.X[3 * static_cast<size_t>(n) + d];
            for (int& c : m.conn) c = newid[remap[c]];
            m.X.swap(X2);
            m.fixed.assign(cnt, 0);
        }
        m.reduced = 1;
        m.law = EFEM_YEOH;
        m.params = {500000.0, -10000.0, 1000.0, 1e-7};
        fixPlane(m, 2, 0.0);
        extractFaces(
            m,
            [&](const double* x) { return std::fabs(std::hypot(x[0], x[1]) - ri) <= 1e-9; }, 0);
        m.pressure = opt(4, 10000.0);
        return m;
    }
    if (which == 4) {
        // bar 0.08 x 0.02 x 0.015, 64x16x12 hexes -> 6 Kuhn C3D10 each; NH with
        // C10_e = 114000 exp(0.2 z_e), z_e ~ N(0,1); 20 kPa on y = min for
        // x in [0.02, 0.06]; x = 0 fixed
        BoxSpec s;
        s.nx = static_cast<int>(opt(0, 64));
        s.ny = static_cast<int>(opt(1, 16));
        s.nz = static_cast<int>(opt(2, 12));
        s.lx = 0.00125 * s.nx;
        s.ly = 0.00125 * s.ny;
        s.lz = 0.00125 * s.nz;
        s.kind = 10;
        Mesh m = makeBox(s);
        SplitMix64 rng(static_cast<uint64_t>(opt(3, 7)));
        m.law = EFEM_NEOHOOKEAN;
        m.params_per_elem = 1;
        m.params.resize(4 * static_cast<size_t>(m.nElems()));
        for (int e = 0; e < m.nElems(); ++e) {
            const double z = rng.normal();
            m.params[4 * static_cast<size_t>(e)] = 114000.0 * std::exp(0.2 * z);
            m.params[4 * static_cast<size_t>(e) + 1] = 0.0;
            m.params[4 * static_cast<size_t>(e) + 2] = 0.0;
            m.params[4 * static_cast<size_t>(e) + 3] = 1e-6;
        }
        fixPlane(m, 0, 0.0);
        pressurePlane(m, 1, 0.0, opt(4, 20000.0), 0, 0.25 * s.lx, 0.75 * s.lx);
        return m;
    }
    if (which == 5) {
        // one instance of the batch of cantilevers: 4x4 C3D8 cross-section,
        // h = 5 mm, length L ~ U{10..40} elements, NH C10 ~ U[80, 150] kPa,
        // D1 = 1e-6, x = 0 fixed, pressure ~ U[0.5, 5] kPa on z = max.
        // opts = {instance index, seed, load scale}; the draws depend on the
        // index and the seed only.
        const int idx = (opts && n_opts > 0) ? static_cast<int>(opts[0]) : 0;
        const uint64_t seed = static_cast<uint64_t>(opt(1, 20240531));
        const double scale = opt(2, 1.0);
        SplitMix64 rng(seed + 0x9E3779B97F4A7C15ULL * static_cast<uint64_t>(idx + 1));
        const int len = 10 + std::min(30, static_cast<int>(rng.uniform() * 31.0));
        const double c10 = 80000.0 + 70000.0 * rng.uniform();
        const double pr = 500.0 + 4500.0 * rng.uniform();
        BoxSpec s;
        s.nx = len;
        s.ny = 4;
        s.nz = 4;
        s.lx = 0.005 * len;
        s.ly = 0.02;
        s.lz = 0.02;
        s.kind = 8;
        Mesh m = makeBox(s);
        m.law = EFEM_NEOHOOKEAN;
        m.params = {c10, 0.0, 0.0, 1e-6};
        fixPlane(m, 0, 0.0);
        pressurePlane(m, 2, s.lz, pr * scale);
        return m;
    }
    // custom box: {nx, ny, nz, lx, ly, lz, kind, law, pressure, jitter, seed, reduced}
    BoxSpec s;
    s.nx = static_cast<int>(opt(0, 4));
    s.ny = static_cast<int>(opt(1, 2));
    s.nz = static_cast<int>(opt(2, 2));
    s.lx = opt(3, 0.004);
    s.ly = opt(4, 0.002);
    s.lz = opt(5, 0.002);
    s.kind = static_cast<int>(opt(6, 8));
    s.jitter = (opts && n_opts > 9) ? opts[9] : 0.0;
    s.seed = static_cast<uint64_t>(opt(10, 1));
    Mesh m = makeBox(s);
    m.law = (opts && n_opts > 7) ? static_cast<int>(opts[7]) : 0;
    m.reduced = (opts && n_opts > 11) ? static_cast<int>(opts[11]) : 0;
    m.params = {114000.0, m.law == 1 ? 20000.0 : (m.law == 2 ? -1000.0 : 0.0), m.law == 2 ? 100.0 : 0.0, 1e-6};
    fixPlane(m, 0, 0.0);
    pressurePlane(m, 2, s.lz, opt(8, 2000.0));
    return m;
}

}  // namespace efem

// ---- C ABI -----------------------------------------------------------------
struct efem_mesh {
    efem::Mesh m;
};

extern "C" efem_mesh* efem_mesh_build(int which, const double* opts, int n_opts)
{
    try {
        auto* h = new efem_mesh;
        h->m = efem::makeWorkload(which, opts, n_opts);
        return h;
    } catch (...) {
        return nullptr;
    }
}

extern "C" void efem_mesh_free(efem_mesh* m) { delete m; }

extern "C" int efem_mesh_sizes(const efem_mesh* h, long long* s)
{
    if (!h || !s) return EFEM_ERR_ARG;
    const efem::Mesh& m = h->m;
    s[0] = m.nNodes();
    s[1] = m.nElems();
    s[2] = m.nen;
    s[3] = m.reduced;
    s[4] = m.law;
    s[5] = m.params_per_elem;
    s[6] = m.nFaces();
    s[7] = m.fnn;
    return EFEM_OK;
}

extern "C" int efem_mesh_arrays(const efem_mesh* h, double* X, int* conn, double* params,
                                unsigned char* fixed, int* faces, double* pressure)
{
    if (!h) return EFEM_ERR_ARG;
    const efem::Mesh& m = h->m;
    if (X) std::memcpy(X, m.X.data(), sizeof(double) * m.X.size());
    if (conn) std::memcpy(conn, m.conn.data(), sizeof(int) * m.conn.size());
    if (params) std::memcpy(params, m.params.data(), sizeof(double) * m.params.size());
    if (fixed) std::memcpy(fixed, m.fixed.data(), m.fixed.size());
    if (faces && !m.faces.empty()) std::memcpy(faces, m.faces.data(), sizeof(int) * m.faces.size());
    if (pressure) *pressure = m.pressure;
    return EFEM_OK;
}

extern "C" int efem_mesh_problem(const efem_mesh* h, efem_problem* pb, int precision)
{
    if (!h || !pb) return EFEM_ERR_ARG;
    const efem::Mesh& m = h->m;
    std::memset(pb, 0, sizeof(*pb));
    pb->n_nodes = m.nNodes();
    pb->X = m.X.data();
    pb->n_elems = m.nElems();
    pb->nen = m.nen;
    pb->conn = m.conn.data();
    pb->reduced = m.reduced;
    pb->law = m.law;
    pb->params = m.params.data();
    pb->params_per_elem = m.params_per_elem;
    pb->fixed = m.fixed.data();
    pb->n_faces = m.nFaces();
    pb->fnn = m.fnn;
    pb->faces = m.faces.empty() ? nullptr : m.faces.data();
    pb->pressure = m.pressure;
    pb->precision = precision;
    return EFEM_OK;
}
