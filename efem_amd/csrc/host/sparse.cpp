// Sparse Cholesky of the scalar approximate Jacobian (see sparse.hpp).
// Written from the published algorithms: geometric nested dissection,
// Liu's elimination tree with path compression, row-subtree column counts,
// fundamental + relaxed supernodes, and the multifrontal method of
// Duff & Reid / Liu.
#include "sparse.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

namespace efem {
namespace {

// ---- nested dissection ----------------------------------------------------
struct NdState {
    const Csc* A;
    const std::vector<double>* xyz;
    std::vector<int> side;  // 0 untouched, 1 left, 2 right (per call)
    std::vector<int> out;
};

void ndRecurse(NdState& st, std::vector<int>& nodes)
{
    const int n = static_cast<int>(nodes.size());
    if (n <= 64) {
        st.out.insert(st.out.end(), nodes.begin(), nodes.end());
        return;
    }
    const auto& X = *st.xyz;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v : nodes)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], X[3 * static_cast<size_t>(v) + d]);
            hi[d] = std::max(hi[d], X[3 * static_cast<size_t>(v) + d]);
        }
    int ax = 0;
    for (int d = 1; d < 3; ++d)
        if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    const int h = n / 2;
    std::nth_element(nodes.begin(), nodes.begin() + h, nodes.end(), [&](int a, int b) {
        const double xa = X[3 * static_cast<size_t>(a) + ax], xb = X[3 * static_cast<size_t>(b) + ax];
        return xa < xb || (xa == xb && a < b);
    });
    for (int k = 0; k < n; ++k) st.side[nodes[k]] = k < h ? 1 : 2;
    // one-sided vertex separators from either side; keep the smaller
    std::vector<int> sepR, sepL;
    for (int k = 0; k < n; ++k) {
        const int v = nodes[k];
        const int other = k < h ? 2 : 1;
        bool touch = false;
        for (int p = st.A->p[v]; p < st.A->p[v + 1] && !touch; ++p) touch = st.side[st.A->i[p]] == other;
        if (touch) (k < h ? sepL : sepR).push_back(v);
    }
    const bool useR = sepR.size() <= sepL.size();
    const std::vector<int>& sep = useR ? sepR : sepL;
    std::vector<int> left, right;
    left.reserve(h);
    right.reserve(n - h);
    std::vector<char> inSep;
    for (int v : sep) st.side[v] = 3;
    for (int k = 0; k < n; ++k) {
        const int v = nodes[k];
        if (st.side[v] == 3) continue;
        (k < h ? left : right).push_back(v);
    }
    for (int v : nodes) st.side[v] = 0;
    if (left.empty() || right.empty()) {
        std::sort(nodes.begin(), nodes.end());
        st.out.insert(st.out.end(), nodes.begin(), nodes.end());
        return;
    }
    std::vector<int> sepCopy = sep;
    nodes.clear();
    nodes.shrink_to_fit();
    ndRecurse(st, left);
    ndRecurse(st, right);
    std::sort(sepCopy.begin(), sepCopy.end());
    st.out.insert(st.out.end(), sepCopy.begin(), sepCopy.end());
}

// ---- dense kernels on column-major blocks (leading dimension ld) ----------

// in-place lower Cholesky of the leading n x n block; false if not SPD
bool potrf(double* A, int n, int ld)
{
    for (int j = 0; j < n; ++j) {
        double* cj = A + static_cast<size_t>(j) * ld;
        for (int k = 0; k < j; ++k) {
            const double* ck = A + static_cast<size_t>(k) * ld;
            const double s = ck[j];
            if (s == 0.0) continue;
            for (int i = j; i < n; ++i) cj[i] -= ck[i] * s;
        }
        const double d = cj[j];
        if (!(d > 0.0)) return false;
        const double r = std::sqrt(d);
        cj[j] = r;
        const double ir = 1.0 / r;
        for (int i = j + 1; i < n; ++i) cj[i] *= ir;
    }
    return true;
}

// A21 (m x n at rows [n, n+m)) <- A21 L11^-T, L11 the leading n x n lower block
void trsmRightLT(double* A, int n, int m, int ld)
{
    for (int k = 0; k < n; ++k) {
        double* ck = A + static_cast<size_t>(k) * ld + n;
        for (int p = 0; p < k; ++p) {
            const double s = A[static_cast<size_t>(p) * ld + k];  // L(k,p)
            if (s == 0.0) continue;
            const double* cp = A + static_cast<size_t>(p) * ld + n;
            for (int i = 0; i < m; ++i) ck[i] -= cp[i] * s;
        }
        const double ir = 1.0 / A[static_cast<size_t>(k) * ld + k];
        for (int i = 0; i < m; ++i) ck[i] *= ir;
    }
}

// U (m x m lower, column-major, ldu) -= A21 A21^T, A21 = rows [n, n+m) of A
void syrkLower(const double* A, int n, int m, int ld, double* U, int ldu)
{
    const int KB = 64;
    for (int k0 = 0; k0 < n; k0 += KB) {
        const int k1 = std::min(n, k0 + KB);
        for (int j = 0; j < m; ++j) {
            double* uj = U + static_cast<size_t>(j) * ldu;
            for (int k = k0; k < k1; ++k) {
                const double* ak = A + static_cast<size_t>(k) * ld + n;
                const double s = ak[j];
                if (s == 0.0) continue;
                for (int i = j; i < m; ++i) uj[i] -= ak[i] * s;
            }
        }
    }
}

}  // namespace

std::vector<int> nestedDissection(const Csc& A, const std::vector<double>& xyz)
{
    NdState st;
    st.A = &A;
    st.xyz = &xyz;
    st.side.assign(A.n, 0);
    st.out.reserve(A.n);
    std::vector<int> nodes(A.n);
    std::iota(nodes.begin(), nodes.end(), 0);
    ndRecurse(st, nodes);
    return st.out;
}

bool supernodalCholesky(const Csc& A, const std::vector<int>& permIn, SupernodalFactor& F,
                        std::string& msg)
{
    const auto t0 = std::chrono::steady_clock::now();
    const int n = A.n;
    F = SupernodalFactor();
    F.n = n;
    if (n == 0) {
        msg = "empty matrix";
        return false;
    }
    std::vector<int> perm = permIn;
    std::vector<int> pinv(n), parent(n), Up, Ui;
    auto buildUpper = [&]() {
        // pattern of the upper triangle of P A P^T, column-wise
        for (int k = 0; k < n; ++k) pinv[perm[k]] = k;
        std::vector<int> cnt(n + 1, 0);
        for (int j = 0; j < n; ++j)
            for (int p = A.p[j]; p < A.p[j + 1]; ++p)
                if (pinv[A.i[p]] <= pinv[j]) cnt[pinv[j] + 1]++;
        for (int j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
        Up = cnt;
        Ui.assign(cnt[n], 0);
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int p = A.p[j]; p < A.p[j + 1]; ++p)
                if (pinv[A.i[p]] <= pinv[j]) Ui[pos[pinv[j]]++] = pinv[A.i[p]];
    };
    auto etree = [&]() {
        std::vector<int> anc(n, -1);
        for (int k = 0; k < n; ++k) {
            parent[k] = -1;
            for (int p = Up[k]; p < Up[k + 1]; ++p) {
                int i = Ui[p];
                while (i != -1 && i < k) {
                    const int next = anc[i];
                    anc[i] = k;
                    if (next == -1) {
                        parent[i] = k;
                        break;
                    }
                    i = next;
                }
            }
        }
    };
    buildUpper();
    etree();
    // postorder so that every subtree is a contiguous column range
    {
        std::vector<int> head(n, -1), next(n, -1), post;
        post.reserve(n);
        for (int j = n - 1; j >= 0; --j)
            if (parent[j] != -1) {
                next[j] = head[parent[j]];
                head[parent[j]] = j;
            }
        std::vector<int> stack;
        for (int r = 0; r < n; ++r) {
            if (parent[r] != -1) continue;
            stack.push_back(r);
            while (!stack.empty()) {
                const int p = stack.back();
                const int c = head[p];
                if (c == -1) {
                    stack.pop_back();
                    post.push_back(p);
                } else {
                    head[p] = next[c];
                    stack.push_back(c);
                }
            }
        }
        std::vector<int> p2(n);
        for (int k = 0; k < n; ++k) p2[k] = perm[post[k]];
        perm.swap(p2);
        buildUpper();
        etree();
    }
    // column counts from the row subtrees (ereach of each row)
    std::vector<int> cc(n, 1), flag(n, -1);
    for (int k = 0; k < n; ++k) {
        flag[k] = k;
        for (int p = Up[k]; p < Up[k + 1]; ++p) {
            for (int i = Ui[p]; flag[i] != k; i = parent[i]) {
                cc[i]++;
                flag[i] = k;
            }
        }
    }
    // fundamental supernodes, then relaxed merging along etree chains
    std::vector<int> nchild(n, 0);
    for (int j = 0; j < n; ++j)
        if (parent[j] != -1) nchild[parent[j]]++;
    // (no supernode is wider than kMaxWidth: the device stages one
    // supernode's right-hand sides in a kMaxWidth-wide LDS buffer)
    const int kMaxWidth = 256;
    std::vector<int> fstart;
    for (int j = 0; j < n; ++j) {
        const bool cont = j > 0 && parent[j - 1] == j && cc[j - 1] == cc[j] + 1 && nchild[j] == 1 &&
                          j - fstart.back() < kMaxWidth;
        if (!cont) fstart.push_back(j);
    }
    fstart.push_back(n);
    std::vector<int> first;
    {
        int g0 = fstart[0];
        long long zeros = 0;
        for (size_t f = 0; f + 1 < fstart.size(); ++f) {
            const int s1 = fstart[f + 1] - 1;  // last column of fundamental f
            if (f + 2 < fstart.size()) {
                // candidate: merge fundamental f+1 into the current group
                const int b0 = fstart[f + 1], b1 = fstart[f + 2] - 1;
                if (parent[s1] == b0 && b1 - g0 + 1 <= kMaxWidth) {
                    // extra zeros of the merged group [g0, b1]
                    long long z = 0, tot = 0;
                    for (int j = g0; j <= b1; ++j) {
                        const long long full = (b1 - j) + cc[b1];
                        z += full - cc[j];
                        tot += full;
                    }
                    const int width = b1 - g0 + 1;
                    if (width <= 16 || z <= 0.15 * tot) {
                        zeros = z;
                        continue;
                    }
                }
            }
            first.push_back(g0);
            g0 = fstart[f + 1];
            zeros = 0;
        }
        (void)zeros;
    }
    first.push_back(n);
    const int ns = static_cast<int>(first.size()) - 1;
    F.nsuper = ns;
    F.first = first;
    std::vector<int> snOf(n);
    for (int s = 0; s < ns; ++s)
        for (int j = first[s]; j < first[s + 1]; ++j) snOf[j] = s;
    F.parent.assign(ns, -1);
    for (int s = 0; s < ns; ++s) {
        const int pj = parent[first[s + 1] - 1];
        F.parent[s] = pj == -1 ? -1 : snOf[pj];
    }
    std::vector<std::vector<int>> children(ns);
    for (int s = 0; s < ns; ++s)
        if (F.parent[s] >= 0) children[F.parent[s]].push_back(s);
    // lower pattern + values of P A P^T by columns
    std::vector<int> Lp(n + 1, 0), Li;
    std::vector<double> Lx;
    {
        for (int j = 0; j < n; ++j)
            for (int p = A.p[j]; p < A.p[j + 1]; ++p)
                if (pinv[A.i[p]] >= pinv[j]) Lp[pinv[j] + 1]++;
        for (int j = 0; j < n; ++j) Lp[j + 1] += Lp[j];
        Li.resize(Lp[n]);
        Lx.resize(Lp[n]);
        std::vector<int> pos(Lp.begin(), Lp.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int p = A.p[j]; p < A.p[j + 1]; ++p)
                if (pinv[A.i[p]] >= pinv[j]) {
                    Li[pos[pinv[j]]] = pinv[A.i[p]];
                    Lx[pos[pinv[j]]++] = A.x[p];
                }
    }
    // symbolic: row sets of the supernodes
    F.rowPtr.assign(ns + 1, 0);
    std::vector<int> mark(n, -1);
    for (int s = 0; s < ns; ++s) {
        const int j0 = first[s], j1 = first[s + 1] - 1;
        std::vector<int> off;
        for (int j = j0; j <= j1; ++j)
            for (int p = Lp[j]; p < Lp[j + 1]; ++p) {
                const int i = Li[p];
                if (i > j1 && mark[i] != s) {
                    mark[i] = s;
                    off.push_back(i);
                }
            }
        for (int c : children[s]) {
            const int m0 = F.rowPtr[c], w = first[c + 1] - first[c];
            const int mc = F.rowPtr[c + 1] - m0;
            for (int r = w; r < mc; ++r) {
                const int i = F.rows[m0 + r];
                if (i > j1 && mark[i] != s) {
                    mark[i] = s;
                    off.push_back(i);
                }
            }
        }
        std::sort(off.begin(), off.end());
        for (int j = j0; j <= j1; ++j) F.rows.push_back(j);
        F.rows.insert(F.rows.end(), off.begin(), off.end());
        F.rowPtr[s + 1] = static_cast<int>(F.rows.size());
    }
    F.valPtr.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
        const long long w = first[s + 1] - first[s], m = F.rowPtr[s + 1] - F.rowPtr[s];
        F.valPtr[s + 1] = F.valPtr[s] + w * m;
        F.nnzL += w * m - w * (w - 1) / 2;
        F.flops += static_cast<double>(w) * w * w / 3.0 + static_cast<double>(m - w) * w * w +
                   static_cast<double>(m - w) * (m - w) * w;
    }
    F.L.assign(F.valPtr[ns], 0.0);
    // numeric multifrontal factorisation; update matrices kept on a stack
    std::vector<std::vector<double>> stack;
    std::vector<int> loc(n, -1);
    for (int s = 0; s < ns; ++s) {
        const int j0 = first[s], w = first[s + 1] - j0;
        const int r0 = F.rowPtr[s], m = F.rowPtr[s + 1] - r0;
        for (int r = 0; r < m; ++r) loc[F.rows[r0 + r]] = r;
        std::vector<double> front(static_cast<size_t>(m) * m, 0.0);
        for (int j = j0; j < j0 + w; ++j)
            for (int p = Lp[j]; p < Lp[j + 1]; ++p)
                front[static_cast<size_t>(loc[j]) * m + loc[Li[p]]] += Lx[p];
        // children were processed last-first; their updates are on top
        for (size_t t = children[s].size(); t-- > 0;) {
            const int c = children[s][t];
            std::vector<double> U = std::move(stack.back());
            stack.pop_back();
            const int cw = first[c + 1] - first[c];
            const int cm = F.rowPtr[c + 1] - F.rowPtr[c] - cw;
            const int* cr = &F.rows[F.rowPtr[c] + cw];
            for (int b = 0; b < cm; ++b) {
                double* dst = &front[static_cast<size_t>(loc[cr[b]]) * m];
                const double* src = &U[static_cast<size_t>(b) * cm];
                for (int a = b; a < cm; ++a) dst[loc[cr[a]]] += src[a];
            }
        }
        if (!potrf(front.data(), w, m)) {
            msg = "approximate Jacobian is not positive definite";
            return false;
        }
        const int nr = m - w;
        trsmRightLT(front.data(), w, nr, m);
        if (nr > 0) {
            std::vector<double> U(static_cast<size_t>(nr) * nr, 0.0);
            for (int b = 0; b < nr; ++b)
                for (int a = b; a < nr; ++a)
                    U[static_cast<size_t>(b) * nr + a] = front[static_cast<size_t>(w + b) * m + w + a];
            syrkLower(front.data(), w, nr, m, U.data(), nr);
            if (F.parent[s] >= 0) stack.push_back(std::move(U));
        }
        double* Ls = &F.L[F.valPtr[s]];
        for (int k = 0; k < w; ++k) {
            for (int i = 0; i < k; ++i) Ls[static_cast<size_t>(k) * m + i] = 0.0;
            for (int i = k; i < m; ++i) Ls[static_cast<size_t>(k) * m + i] = front[static_cast<size_t>(k) * m + i];
        }
    }
    // device blocks B_s = [L11^-1 ; L21 L11^-1]
    F.B.assign(F.valPtr[ns], 0.0);
    for (int s = 0; s < ns; ++s) {
        const int w = first[s + 1] - first[s], m = F.rowPtr[s + 1] - F.rowPtr[s];
        const double* Ls = &F.L[F.valPtr[s]];
        double* Bs = &F.B[F.valPtr[s]];
        // L11^-1, column by column: solve L11 x = e_k
        for (int k = 0; k < w; ++k) {
            double* x = Bs + static_cast<size_t>(k) * m;
            x[k] = 1.0 / Ls[static_cast<size_t>(k) * m + k];
            for (int j = k + 1; j < w; ++j) {
                double acc = 0.0;
                for (int p = k; p < j; ++p) acc += Ls[static_cast<size_t>(p) * m + j] * x[p];
                x[j] = -acc / Ls[static_cast<size_t>(j) * m + j];
            }
        }
        // M = L21 L11^-1  (rows w..m-1)
        for (int k = 0; k < w; ++k) {
            double* mk = Bs + static_cast<size_t>(k) * m + w;
            for (int p = k; p < w; ++p) {
                const double s2 = Bs[static_cast<size_t>(k) * m + p];  // Linv(p,k)
                if (s2 == 0.0) continue;
                const double* lp = Ls + static_cast<size_t>(p) * m + w;
                for (int i = 0; i < m - w; ++i) mk[i] += lp[i] * s2;
            }
        }
    }
    F.level.assign(ns, 0);
    F.nlevels = 0;
    for (int s = 0; s < ns; ++s) {
        if (F.parent[s] >= 0) F.level[F.parent[s]] = std::max(F.level[F.parent[s]], F.level[s] + 1);
        F.nlevels = std::max(F.nlevels, F.level[s] + 1);
    }
    F.perm = perm;
    F.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

void factorSolve(const SupernodalFactor& F, const double* b, double* x)
{
    const int n = F.n;
    std::vector<double> y(n);
    for (int k = 0; k < n; ++k) y[k] = b[F.perm[k]];
    for (int s = 0; s < F.nsuper; ++s) {
        const int j0 = F.first[s], w = F.first[s + 1] - j0;
        const int m = F.rowPtr[s + 1] - F.rowPtr[s];
        const int* rows = &F.rows[F.rowPtr[s]];
        const double* Ls = &F.L[F.valPtr[s]];
        for (int k = 0; k < w; ++k) {
            const double* c = Ls + static_cast<size_t>(k) * m;
            const double v = y[j0 + k] / c[k];
            y[j0 + k] = v;
            for (int i = k + 1; i < m; ++i) y[rows[i]] -= c[i] * v;
        }
    }
    for (int s = F.nsuper - 1; s >= 0; --s) {
        const int j0 = F.first[s], w = F.first[s + 1] - j0;
        const int m = F.rowPtr[s + 1] - F.rowPtr[s];
        const int* rows = &F.rows[F.rowPtr[s]];
        const double* Ls = &F.L[F.valPtr[s]];
        for (int k = w - 1; k >= 0; --k) {
            const double* c = Ls + static_cast<size_t>(k) * m;
            double v = y[j0 + k];
            for (int i = k + 1; i < m; ++i) v -= c[i] * y[rows[i]];
            y[j0 + k] = v / c[k];
        }
    }
    for (int k = 0; k < n; ++k) x[F.perm[k]] = y[k];
}

void blockInverseSolve3(const SupernodalFactor& F, const double* b, double* x)
{
    const int n = F.n;
    std::vector<int> uOff(F.nsuper + 1, 0);
    for (int s = 0; s < F.nsuper; ++s) {
        const int w = F.first[s + 1] - F.first[s], m = F.rowPtr[s + 1] - F.rowPtr[s];
        uOff[s + 1] = uOff[s] + (m - w);
    }
    std::vector<int> fptr, fidx;
    frontGatherMap(F, uOff, fptr, fidx);
    std::vector<double> u(3 * static_cast<size_t>(uOff[F.nsuper])), y(3 * static_cast<size_t>(n)),
        z(3 * static_cast<size_t>(n)), f;
    for (int s = 0; s < F.nsuper; ++s) {
        const int j0 = F.first[s], w = F.first[s + 1] - j0, m = F.rowPtr[s + 1] - F.rowPtr[s];
        const int fr0 = F.rowPtr[s];
        const double* B = &F.B[F.valPtr[s]];
        f.assign(3 * static_cast<size_t>(m), 0.0);
        for (int r = 0; r < m; ++r)
            for (int c = 0; c < 3; ++c) {
                double v = r < w ? b[3 * static_cast<size_t>(F.perm[j0 + r]) + c] : 0.0;
                for (int p = fptr[fr0 + r]; p < fptr[fr0 + r + 1]; ++p) v += u[3 * static_cast<size_t>(fidx[p]) + c];
                f[3 * r + c] = v;
            }
        for (int r = 0; r < m; ++r)
            for (int c = 0; c < 3; ++c) {
                double acc = 0.0;
                for (int k = 0; k < w; ++k) acc += B[static_cast<size_t>(k) * m + r] * f[3 * k + c];
                if (r < w)
                    y[3 * static_cast<size_t>(j0 + r) + c] = acc;
                else
                    u[3 * static_cast<size_t>(uOff[s] + r - w) + c] = f[3 * r + c] - acc;
            }
    }
    for (int s = F.nsuper - 1; s >= 0; --s) {
        const int j0 = F.first[s], w = F.first[s + 1] - j0, m = F.rowPtr[s + 1] - F.rowPtr[s];
        const int* rows = &F.rows[F.rowPtr[s]];
        const double* B = &F.B[F.valPtr[s]];
        for (int k = 0; k < w; ++k)
            for (int c = 0; c < 3; ++c) {
                double acc = 0.0;
                for (int r = k; r < m; ++r) {
                    const double wv = r < w ? y[3 * static_cast<size_t>(j0 + r) + c] : -z[3 * static_cast<size_t>(rows[r]) + c];
                    acc += B[static_cast<size_t>(k) * m + r] * wv;
                }
                z[3 * static_cast<size_t>(j0 + k) + c] = acc;
            }
    }
    for (int j = 0; j < n; ++j)
        for (int c = 0; c < 3; ++c) x[3 * static_cast<size_t>(F.perm[j]) + c] = z[3 * static_cast<size_t>(j) + c];
}

void frontGatherMap(const SupernodalFactor& F, const std::vector<int>& uOff, std::vector<int>& fptr,
                    std::vector<int>& fidx)
{
    // front row (rowPtr[s] + r) <- update rows of the children of s, children
    // in ascending order, each child's update rows in order
    const int ns = F.nsuper;
    std::vector<std::vector<int>> children(ns);
    for (int s = 0; s < ns; ++s)
        if (F.parent[s] >= 0) children[F.parent[s]].push_back(s);
    const int total = F.rowPtr[ns];
    std::vector<int> cnt(total + 1, 0);
    std::vector<int> loc(F.n, -1);
    auto forEach = [&](auto&& fn) {
        for (int s = 0; s < ns; ++s) {
            const int r0 = F.rowPtr[s], m = F.rowPtr[s + 1] - r0;
            for (int r = 0; r < m; ++r) loc[F.rows[r0 + r]] = r;
            for (int c : children[s]) {
                const int cw = F.first[c + 1] - F.first[c];
                const int cm = F.rowPtr[c + 1] - F.rowPtr[c];
                for (int rr = cw; rr < cm; ++rr) fn(r0 + loc[F.rows[F.rowPtr[c] + rr]], uOff[c] + rr - cw);
            }
        }
    };
    forEach([&](int frow, int) { cnt[frow + 1]++; });
    for (int i = 0; i < total; ++i) cnt[i + 1] += cnt[i];
    fptr = cnt;
    fidx.assign(cnt[total], 0);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    forEach([&](int frow, int upos) { fidx[pos[frow]++] = upos; });
}

}  // namespace efem
