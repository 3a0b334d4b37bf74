// Sparse Cholesky of the scalar approximate Jacobian (see sparse.hpp).
// Written from the published algorithms: geometric nested dissection,
// Liu's elimination tree with path compression, row-subtree column counts,
// fundamental + relaxed supernodes, and the multifrontal method of
// Duff & Reid / Liu.
#include "sparse.hpp"
#include "../limits.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <atomic>
#include <functional>
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

// A21 (m x n at rows [n, n+m)) <- A21 L11^-T, L11 the leading n x n lower
// block.  Rows are independent: row blocks of 128 run as tasks when `par`.
__attribute__((target_clones("arch=x86-64-v3", "default")))
void trsmRows(double* A, int n, int ld, int i0, int i1)
{
    for (int k = 0; k < n; ++k) {
        double* ck = A + static_cast<size_t>(k) * ld + n;
        for (int p = 0; p < k; ++p) {
            const double s = A[static_cast<size_t>(p) * ld + k];  // L(k,p)
            if (s == 0.0) continue;
            const double* cp = A + static_cast<size_t>(p) * ld + n;
            for (int i = i0; i < i1; ++i) ck[i] -= cp[i] * s;
        }
        const double ir = 1.0 / A[static_cast<size_t>(k) * ld + k];
        for (int i = i0; i < i1; ++i) ck[i] *= ir;
    }
}

void trsmRightLT(double* A, int n, int m, int ld, bool par)
{
    if (!par) {
        trsmRows(A, n, ld, 0, m);
        return;
    }
#pragma omp taskloop grainsize(1)
    for (int i0 = 0; i0 < m; i0 += 128) trsmRows(A, n, ld, i0, std::min(m, i0 + 128));
}

// C[i][j] -= sum_k A[i][k] B[j][k] for a 64 x 64 (or smaller) tile, A and B
// rows of column-major panels with leading dimension ld; 8 x 4 register tiles
__attribute__((target_clones("arch=x86-64-v3", "default")))
void gemmNTTile(double* C, int ldc, const double* A, const double* B, int ld, int mi, int nj, int kk,
                bool lowerDiag)
{
    for (int j = 0; j < nj; j += 4) {
        const int jn = std::min(4, nj - j);
        int i = lowerDiag ? (j / 8) * 8 : 0;
        for (; i < mi; i += 8) {
            const int in = std::min(8, mi - i);
            double acc[4][8] = {};
            if (in == 8 && jn == 4) {
                for (int k = 0; k < kk; ++k) {
                    const double* a = A + static_cast<size_t>(k) * ld + i;
                    const double* b = B + static_cast<size_t>(k) * ld + j;
                    for (int jj = 0; jj < 4; ++jj)
                        for (int ii = 0; ii < 8; ++ii) acc[jj][ii] += a[ii] * b[jj];
                }
            } else {
                for (int k = 0; k < kk; ++k) {
                    const double* a = A + static_cast<size_t>(k) * ld + i;
                    const double* b = B + static_cast<size_t>(k) * ld + j;
                    for (int jj = 0; jj < jn; ++jj)
                        for (int ii = 0; ii < in; ++ii) acc[jj][ii] += a[ii] * b[jj];
                }
            }
            for (int jj = 0; jj < jn; ++jj)
                for (int ii = 0; ii < in; ++ii) C[static_cast<size_t>(j + jj) * ldc + i + ii] -= acc[jj][ii];
        }
    }
}

// U (m x m lower, column-major, ldu) -= A21 A21^T, A21 = rows [n, n+m) of A;
// 64 x 64 tiles of the lower triangle, as tasks when `par`
void syrkLower(const double* A, int n, int m, int ld, double* U, int ldu, bool par)
{
    const int T = 64;
    const int nt = (m + T - 1) / T;
    auto tile = [&](int bj, int bi) {
        const int j0 = bj * T, i0 = bi * T;
        const int nj = std::min(T, m - j0), mi = std::min(T, m - i0);
        // the strict upper part of diagonal tiles is computed and ignored
        gemmNTTile(U + static_cast<size_t>(j0) * ldu + i0, ldu, A + n + i0, A + n + j0, ld, mi, nj, n, bi == bj);
    };
    if (!par) {
        for (int bj = 0; bj < nt; ++bj)
            for (int bi = bj; bi < nt; ++bi) tile(bj, bi);
        return;
    }
#pragma omp taskloop grainsize(1) collapse(2)
    for (int bj = 0; bj < nt; ++bj)
        for (int bi = 0; bi < nt; ++bi)
            if (bi >= bj) tile(bj, bi);
}

// B = [L11^-1 ; L21 L11^-1] from the supernode's column-major L block (m x w)
__attribute__((target_clones("arch=x86-64-v3", "default")))
void mRows(const double* L, double* B, int w, int m, int i0, int i1)
{
    // M(i, k) = sum_{p >= k} L21(i, p) Linv(p, k)
    for (int k = 0; k < w; ++k) {
        double* mk = B + static_cast<size_t>(k) * m;
        for (int p = k; p < w; ++p) {
            const double s = B[static_cast<size_t>(k) * m + p];  // Linv(p, k)
            if (s == 0.0) continue;
            const double* lp = L + static_cast<size_t>(p) * m;
            for (int i = i0; i < i1; ++i) mk[i] += lp[i] * s;
        }
    }
}

void blockInverse(const double* L, double* B, int w, int m, bool par)
{
    // L11^-1 column by column: solve L11 x = e_k
    for (int k = 0; k < w; ++k) {
        double* x = B + static_cast<size_t>(k) * m;
        x[k] = 1.0 / L[static_cast<size_t>(k) * m + k];
        for (int j = k + 1; j < w; ++j) {
            double acc = 0.0;
            for (int p = k; p < j; ++p) acc += L[static_cast<size_t>(p) * m + j] * x[p];
            x[j] = -acc / L[static_cast<size_t>(j) * m + j];
        }
    }
    if (!par) {
        mRows(L, B, w, m, w, m);
        return;
    }
#pragma omp taskloop grainsize(1)
    for (int i0 = w; i0 < m; i0 += 128) mRows(L, B, w, m, i0, std::min(m, i0 + 128));
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

void mergeFactors(std::vector<SupernodalFactor>& parts, SupernodalFactor& out, std::vector<int>& snOffset)
{
    out = SupernodalFactor();
    long long bs = 0;
    for (const auto& P : parts) {
        out.n += P.n;
        out.nsuper += P.nsuper;
        bs += P.Bsize;
        out.nlevels = std::max(out.nlevels, P.nlevels);
        out.nnzL += P.nnzL;
        out.flops += P.flops;
        out.seconds = std::max(out.seconds, P.seconds);
        out.numericSeconds = std::max(out.numericSeconds, P.numericSeconds);
    }
    out.Bsize = bs;
    out.B.reset(new double[static_cast<size_t>(std::max(bs, 1LL))]);
    snOffset.assign(1, 0);
    int col = 0, sn = 0, row = 0;
    long long val = 0;
    for (auto& P : parts) {
        for (int v : P.perm) out.perm.push_back(v + col);
        for (int s = 0; s < P.nsuper; ++s) {
            out.first.push_back(P.first[s] + col);
            out.rowPtr.push_back(P.rowPtr[s] + row);
            out.valPtr.push_back(P.valPtr[s] + val);
            out.parent.push_back(P.parent[s] >= 0 ? P.parent[s] + sn : -1);
            out.level.push_back(P.level[s]);
        }
        for (int r : P.rows) out.rows.push_back(r + col);
        std::memcpy(out.B.get() + val, P.B.get(), static_cast<size_t>(P.Bsize) * sizeof(double));
        col += P.n;
        sn += P.nsuper;
        row += static_cast<int>(P.rows.size());
        val += P.Bsize;
        snOffset.push_back(sn);
        P.B.reset();  // the merged copy holds the blocks
    }
    out.first.push_back(col);
    out.rowPtr.push_back(row);
    out.valPtr.push_back(val);
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
    const int kMaxWidth = kMaxSupernodeWidth;
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
    // numeric multifrontal factorisation, tree-parallel (OpenMP tasks over
    // subtrees; big fronts split their dense kernels into tasks too).  Each
    // supernode's update matrix lives in upd[s] until its parent absorbs it.
    // B is filled supernode by supernode inside the tasks (no serial zero-fill)
    F.B.reset(new double[static_cast<size_t>(F.valPtr[ns])]);
    F.Bsize = F.valPtr[ns];
    std::vector<std::vector<double>> upd(ns);
    std::vector<long long> subtreeCols(ns, 0);
    for (int s = 0; s < ns; ++s) {
        subtreeCols[s] += first[s + 1] - first[s];
        if (F.parent[s] >= 0) subtreeCols[F.parent[s]] += subtreeCols[s];
    }
    std::atomic<int> failed{0};
    const auto tn0 = std::chrono::steady_clock::now();
    auto factorFront = [&](int s) {
        if (failed.load(std::memory_order_relaxed)) return;
        const int j0 = first[s], w = first[s + 1] - j0;
        const int r0 = F.rowPtr[s], m = F.rowPtr[s + 1] - r0;
        const int* rows = &F.rows[r0];
        // row -> front position by binary search in the sorted row set
        auto locOf = [&](int gi) { return static_cast<int>(std::lower_bound(rows, rows + m, gi) - rows); };
        std::vector<double> front(static_cast<size_t>(m) * m, 0.0);
        for (int j = j0; j < j0 + w; ++j)
            for (int p = Lp[j]; p < Lp[j + 1]; ++p) front[static_cast<size_t>(j - j0) * m + locOf(Li[p])] += Lx[p];
        for (int c : children[s]) {
            std::vector<double>& U = upd[c];
            const int cw = first[c + 1] - first[c];
            const int cm = F.rowPtr[c + 1] - F.rowPtr[c] - cw;
            const int* cr = &F.rows[F.rowPtr[c] + cw];
            std::vector<int> cl(cm);
            for (int b = 0; b < cm; ++b) cl[b] = locOf(cr[b]);
            for (int b = 0; b < cm; ++b) {
                double* dst = &front[static_cast<size_t>(cl[b]) * m];
                const double* src = &U[static_cast<size_t>(b) * cm];
                for (int a = b; a < cm; ++a) dst[cl[a]] += src[a];
            }
            std::vector<double>().swap(U);
        }
        if (!potrf(front.data(), w, m)) {
            failed = 1;
            return;
        }
        const int nr = m - w;
        const bool big = static_cast<double>(nr) * nr * w > 2e7;
        trsmRightLT(front.data(), w, nr, m, big);
        if (nr > 0 && F.parent[s] >= 0) {
            std::vector<double> U(static_cast<size_t>(nr) * nr, 0.0);
            for (int b = 0; b < nr; ++b)
                for (int a = b; a < nr; ++a)
                    U[static_cast<size_t>(b) * nr + a] = front[static_cast<size_t>(w + b) * m + w + a];
            syrkLower(front.data(), w, nr, m, U.data(), nr, big);
            upd[s] = std::move(U);
        }
        // the first w columns of the front now hold [L11; L21] (ld = m)
        double* Bs = &F.B[F.valPtr[s]];
        std::fill(Bs, Bs + static_cast<size_t>(w) * m, 0.0);
        blockInverse(front.data(), Bs, w, m, big);
    };
    // postorder recursion with tasks for large subtrees
    std::function<void(int)> process = [&](int s) {
        for (int c : children[s]) {
            if (subtreeCols[c] > 512) {
#pragma omp task firstprivate(c) shared(process)
                process(c);
            } else {
                process(c);
            }
        }
#pragma omp taskwait
        factorFront(s);
    };
#pragma omp parallel
#pragma omp single
    {
        for (int s = 0; s < ns; ++s)
            if (F.parent[s] < 0) {
#pragma omp task firstprivate(s) shared(process)
                process(s);
            }
#pragma omp taskwait
    }
    if (failed) {
        msg = "approximate Jacobian is not positive definite";
        return false;
    }
    F.numericSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - tn0).count();
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
