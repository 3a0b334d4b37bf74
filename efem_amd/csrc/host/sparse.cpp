// Sparse Cholesky of the scalar approximate Jacobian (see sparse.hpp).
// Written from the published algorithms: geometric nested dissection,
// Liu's elimination tree with path compression, row-subtree column counts,
// fundamental + relaxed supernodes, and the multifrontal method of
// Duff & Reid / Liu.
#include "sparse.hpp"
#include "../limits.hpp"
#include "dense.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <atomic>
#include <functional>
#include <memory>
#include <numeric>

namespace efem {
namespace {

// ---- nested dissection ----------------------------------------------------
struct NdState {
    const Csc* A;
    const std::vector<double>* xyz;
    // per node: 0, or 4 (base + 1) + {1 left, 2 right, 3 separator} while
    // the call whose output starts at `base` is splitting it.  Calls that
    // run concurrently (sibling subtrees) have different bases, so a node's
    // mark is never mistaken for another call's.
    std::vector<int> side;
    std::vector<int> out;  // the ordering; each call fills its own range
};

// Orders `nodes` into out[base, base + nodes.size()): left part, right
// part, separator.  Subproblems above 4096 nodes run as OpenMP tasks.
void ndRecurse(NdState& st, std::vector<int>& nodes, int base)
{
    const int n = static_cast<int>(nodes.size());
    if (n <= 64) {
        std::copy(nodes.begin(), nodes.end(), st.out.begin() + base);
        return;
    }
    const int key = 4 * (base + 1);
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
    for (int k = 0; k < n; ++k) st.side[nodes[k]] = key + (k < h ? 1 : 2);
    // one-sided vertex separators from either side; keep the smaller
    std::vector<int> sepR, sepL;
    for (int k = 0; k < n; ++k) {
        const int v = nodes[k];
        const int other = key + (k < h ? 2 : 1);
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
    for (int v : sep) st.side[v] = key + 3;
    for (int k = 0; k < n; ++k) {
        const int v = nodes[k];
        if (st.side[v] == key + 3) continue;
        (k < h ? left : right).push_back(v);
    }
    for (int v : nodes) st.side[v] = 0;
    if (left.empty() || right.empty()) {
        std::sort(nodes.begin(), nodes.end());
        std::copy(nodes.begin(), nodes.end(), st.out.begin() + base);
        return;
    }
    std::vector<int> sepCopy = sep;
    nodes.clear();
    nodes.shrink_to_fit();
    const int nl = static_cast<int>(left.size()), nr = static_cast<int>(right.size());
    std::sort(sepCopy.begin(), sepCopy.end());
    std::copy(sepCopy.begin(), sepCopy.end(), st.out.begin() + base + nl + nr);
    if (n > 4096) {
#pragma omp task default(shared)
        ndRecurse(st, left, base);
#pragma omp task default(shared)
        ndRecurse(st, right, base + nl);
#pragma omp taskwait
    } else {
        ndRecurse(st, left, base);
        ndRecurse(st, right, base + nl);
    }
}

// ---- dense front operations (column-major, leading dimension ld) -----------
// The O(n^3) work of every operation is dense::gemm (host/dense.cpp); big
// fronts split their rows (or tiles) into OpenMP tasks.

// A21 (m rows below the n x n pivot block) <- A21 L11^-T
void trsmRightLT(double* A, int n, int m, int ld, bool par)
{
    if (!par) {
        dense::trsmRightLT(A + n, ld, A, ld, n, 0, m);
        return;
    }
#pragma omp taskloop grainsize(1)
    for (int i0 = 0; i0 < m; i0 += 128) dense::trsmRightLT(A + n, ld, A, ld, n, i0, std::min(m, i0 + 128));
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
        dense::gemm(U + static_cast<size_t>(j0) * ldu + i0, ldu, A + n + i0, ld, A + n + j0, ld, true, mi, nj, n,
                    -1.0);
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

// in place: the supernode's column-major m x w block P (ld m) holds
// [L11; L21] on entry and [L11^-1; L21 L11^-1] on return
void blockInverse(double* P, int w, int m, bool par)
{
    std::vector<double> Linv(static_cast<size_t>(w) * w);
    dense::trtri(P, m, Linv.data(), w, w, par && w >= 256);
    auto chunk = [&](int i0, int i1) {
        const int r = i1 - i0;
        std::vector<double> t(static_cast<size_t>(r) * w, 0.0);
        dense::trmmRightLower(t.data(), r, P + i0, m, Linv.data(), w, w, r);
        for (int k = 0; k < w; ++k)
            std::memcpy(P + i0 + static_cast<size_t>(k) * m, t.data() + static_cast<size_t>(k) * r, sizeof(double) * r);
    };
    if (m > w) {
        if (par) {
#pragma omp taskloop grainsize(1)
            for (int i0 = w; i0 < m; i0 += 128) chunk(i0, std::min(m, i0 + 128));
        } else {
            chunk(w, m);
        }
    }
    for (int k = 0; k < w; ++k)
        std::memcpy(P + static_cast<size_t>(k) * m, Linv.data() + static_cast<size_t>(k) * w, sizeof(double) * w);
}

}  // namespace

std::vector<int> nestedDissection(const Csc& A, const std::vector<double>& xyz)
{
    NdState st;
    st.A = &A;
    st.xyz = &xyz;
    st.side.assign(A.n, 0);
    st.out.assign(A.n, -1);
    std::vector<int> nodes(A.n);
    std::iota(nodes.begin(), nodes.end(), 0);
#pragma omp parallel
#pragma omp single
    ndRecurse(st, nodes, 0);
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

HostWorkspace::~HostWorkspace()
{
    for (auto& kv : capacity_) delete[] kv.first;
}

std::unique_ptr<double[]> HostWorkspace::takeBlocks(size_t count)
{
    if (blocks_ && blockCount_ >= count) {
        blockCount_ = 0;
        return std::move(blocks_);
    }
    blocks_.reset();
    blockCount_ = 0;
    return std::unique_ptr<double[]>(new double[std::max<size_t>(count, 1)]);
}

void HostWorkspace::giveBlocks(std::unique_ptr<double[]> b, size_t count)
{
    blocks_ = std::move(b);
    blockCount_ = blocks_ ? count : 0;
}

double* HostWorkspace::takeUpdate(size_t count)
{
    std::lock_guard<std::mutex> lock(mu_);
    auto it = free_.lower_bound(count);
    if (it != free_.end()) {
        double* p = it->second;
        free_.erase(it);
        return p;
    }
    double* p = new double[std::max<size_t>(count, 1)];
    capacity_[p] = count;
    return p;
}

void HostWorkspace::giveUpdate(double* p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lock(mu_);
    free_.emplace(capacity_.at(p), p);
}

bool supernodalSymbolic(const Csc& A, const std::vector<int>& permIn, SupernodalSymbolic& F, std::string& msg,
                        int denseSubtree)
{
    const auto t0 = std::chrono::steady_clock::now();
    const int n = A.n;
    F = SupernodalSymbolic();
    F.n = n;
    F.anz = static_cast<long long>(A.i.size());
    if (n == 0) {
        msg = "empty matrix";
        return false;
    }
    std::vector<int> perm = permIn;
    std::vector<int> pinv(n), parent(n);
    // elimination tree of P A P^T (Liu's algorithm with path compression)
    // straight from A's full (symmetric) pattern: column k of P A P^T is
    // column perm[k] of A renumbered by pinv, its upper entries those < k
    for (int k = 0; k < n; ++k) pinv[perm[k]] = k;
    auto etree = [&]() {
        std::vector<int> anc(n, -1);
        for (int k = 0; k < n; ++k) {
            parent[k] = -1;
            const int col = perm[k];
            for (int p = A.p[col]; p < A.p[col + 1]; ++p) {
                int i = pinv[A.i[p]];
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
        // relabel: the tree is unchanged, only its node numbers
        std::vector<int> ipost(n), p2(n), par2(n);
        for (int k = 0; k < n; ++k) ipost[post[k]] = k;
        for (int k = 0; k < n; ++k) {
            p2[k] = perm[post[k]];
            par2[k] = parent[post[k]] == -1 ? -1 : ipost[parent[post[k]]];
        }
        perm.swap(p2);
        parent.swap(par2);
        for (int k = 0; k < n; ++k) pinv[perm[k]] = k;
    }
    // lower pattern + values of P A P^T by columns
    std::vector<int>& Lp = F.Lp;
    std::vector<int>& Li = F.Li;
    std::vector<int>& Lsrc = F.Lsrc;
    Lp.assign(n + 1, 0);
    {
        // column k of P A P^T is column perm[k] of A: counted and filled in
        // permuted order, columns in parallel (each writes its own range;
        // entries in A's order within the column)
#pragma omp parallel for schedule(static, 1024)
        for (int k = 0; k < n; ++k) {
            const int col = perm[k];
            int cnt = 0;
            for (int p = A.p[col]; p < A.p[col + 1]; ++p) cnt += pinv[A.i[p]] >= k;
            Lp[k + 1] = cnt;
        }
        for (int k = 0; k < n; ++k) Lp[k + 1] += Lp[k];
        Li.resize(Lp[n]);
        Lsrc.resize(Lp[n]);
#pragma omp parallel for schedule(static, 1024)
        for (int k = 0; k < n; ++k) {
            const int col = perm[k];
            int q = Lp[k];
            for (int p = A.p[col]; p < A.p[col + 1]; ++p) {
                const int i = pinv[A.i[p]];
                if (i >= k) {
                    Li[q] = i;
                    Lsrc[q++] = p;
                }
            }
        }
    }
    // column counts of L (diagonal included) by the row-subtree skeleton
    // method of Gilbert, Ng & Peyton (SIAM J. Matrix Anal. Appl. 15(4),
    // 1994) on the postordered tree: entry A(i, j), i > j, starts a new
    // branch of row subtree i only when j is a leaf of it (first descendant
    // of j beyond the last one seen for i); the branch from j joins the
    // previous leaf's path at their least common ancestor q, found by path
    // compression.  delta[] holds per-node differences, summed up the tree.
    std::vector<int> cc(n, 0);
    {
        std::vector<int> fdesc(n, -1), maxFirst(n, -1), prevLeaf(n, -1), anc(n);
        for (int k = 0; k < n; ++k) {
            cc[k] = fdesc[k] == -1 ? 1 : 0;
            for (int j = k; j != -1 && fdesc[j] == -1; j = parent[j]) fdesc[j] = k;
        }
        std::iota(anc.begin(), anc.end(), 0);
        for (int j = 0; j < n; ++j) {
            if (parent[j] != -1) cc[parent[j]]--;
            for (int p = Lp[j]; p < Lp[j + 1]; ++p) {
                const int i = Li[p];
                if (i <= j || fdesc[j] <= maxFirst[i]) continue;
                maxFirst[i] = fdesc[j];
                const int jprev = prevLeaf[i];
                prevLeaf[i] = j;
                cc[j]++;
                if (jprev != -1) {
                    int q = jprev;
                    while (q != anc[q]) q = anc[q];
                    for (int t = jprev; t != q;) {
                        const int up = anc[t];
                        anc[t] = q;
                        t = up;
                    }
                    cc[q]--;
                }
            }
            if (parent[j] != -1) anc[j] = parent[j];
        }
        for (int j = 0; j < n; ++j)
            if (parent[j] != -1) cc[parent[j]] += cc[j];
    }
    // fundamental supernodes, then relaxed merging along etree chains
    std::vector<int> nchild(n, 0);
    for (int j = 0; j < n; ++j)
        if (parent[j] != -1) nchild[parent[j]]++;
    // (no supernode is wider than kMaxWidth: the device stages one
    // supernode's right-hand sides in a kMaxWidth-wide LDS buffer)
    const int kMaxWidth = kMaxSupernodeWidth;
    // Every maximal subtree of at most denseSubtree columns (contiguous in
    // postorder) becomes one dense supernode: the bottom of the tree loses
    // several levels -- each level is one launch per sweep of the device
    // triangular solves -- for a few percent more stored entries (on the
    // 265k-DOF beam with 64: 17 -> 12 levels, B +7 %, factor flops +2 %).
    const int kDenseSubtree = std::max(1, std::min(denseSubtree, kMaxSupernodeWidth));
    std::vector<int> sz(n, 1), denseRoot(n, -1);
    for (int j = 0; j < n; ++j)
        if (parent[j] != -1) sz[parent[j]] += sz[j];
    for (int j = n - 1; j >= 0; --j)
        if (denseRoot[j] < 0 && sz[j] <= kDenseSubtree && (parent[j] == -1 || sz[parent[j]] > kDenseSubtree))
            for (int i = j - sz[j] + 1; i <= j; ++i) denseRoot[i] = j;
    std::vector<int> fstart;
    for (int j = 0; j < n; ++j) {
        bool cont = j > 0 && parent[j - 1] == j && cc[j - 1] == cc[j] + 1 && nchild[j] == 1 &&
                    j - fstart.back() < kMaxWidth;
        if (denseRoot[j] >= 0)
            cont = j > 0 && denseRoot[j - 1] == denseRoot[j];
        else if (j > 0 && denseRoot[j - 1] >= 0)
            cont = false;
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
    std::vector<std::vector<int>>& children = F.children;
    children.assign(ns, {});
    for (int s = 0; s < ns; ++s)
        if (F.parent[s] >= 0) children[F.parent[s]].push_back(s);
    // symbolic: row sets of the supernodes -- the rows of its columns below
    // the supernode and its children's update rows, sorted without
    // duplicates -- children before parents, subtrees of more than 512
    // columns as OpenMP tasks
    std::vector<long long>& subtreeCols = F.subtreeCols;
    subtreeCols.assign(ns, 0);
    for (int s = 0; s < ns; ++s) {
        subtreeCols[s] += first[s + 1] - first[s];
        if (F.parent[s] >= 0) subtreeCols[F.parent[s]] += subtreeCols[s];
    }
    std::vector<std::vector<int>> offRows(ns);
    auto rowSet = [&](int s) {
        const int j0 = first[s], j1 = first[s + 1] - 1;
        std::vector<int> off;
        for (int j = j0; j <= j1; ++j)
            for (int p = Lp[j]; p < Lp[j + 1]; ++p)
                if (Li[p] > j1) off.push_back(Li[p]);
        for (int c : children[s])
            for (int i : offRows[c])
                if (i > j1) off.push_back(i);
        std::sort(off.begin(), off.end());
        off.erase(std::unique(off.begin(), off.end()), off.end());
        offRows[s] = std::move(off);
    };
    std::function<void(int)> rowsBelow = [&](int s) {
        for (int c : children[s]) {
            if (subtreeCols[c] > 512) {
#pragma omp task firstprivate(c) shared(rowsBelow)
                rowsBelow(c);
            } else {
                rowsBelow(c);
            }
        }
#pragma omp taskwait
        rowSet(s);
    };
#pragma omp parallel
#pragma omp single
    {
        for (int s = 0; s < ns; ++s)
            if (F.parent[s] < 0) {
#pragma omp task firstprivate(s) shared(rowsBelow)
                rowsBelow(s);
            }
#pragma omp taskwait
    }
    F.rowPtr.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s)
        F.rowPtr[s + 1] = F.rowPtr[s] + (first[s + 1] - first[s]) + static_cast<int>(offRows[s].size());
    F.rows.resize(F.rowPtr[ns]);
#pragma omp parallel for schedule(dynamic, 64)
    for (int s = 0; s < ns; ++s) {
        int* r = &F.rows[F.rowPtr[s]];
        for (int j = first[s]; j < first[s + 1]; ++j) *r++ = j;
        std::copy(offRows[s].begin(), offRows[s].end(), r);
    }
    std::vector<std::vector<int>>().swap(offRows);
    F.valPtr.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
        const long long w = first[s + 1] - first[s], m = F.rowPtr[s + 1] - F.rowPtr[s];
        F.valPtr[s + 1] = F.valPtr[s] + w * m;
        F.nnzL += w * m - w * (w - 1) / 2;
        F.flops += static_cast<double>(w) * w * w / 3.0 + static_cast<double>(m - w) * w * w +
                   static_cast<double>(m - w) * (m - w) * w;
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

bool supernodalNumeric(const Csc& A, const SupernodalSymbolic& S, SupernodalFactor& F, std::string& msg,
                       HostWorkspace* ws)
{
    const auto t0 = std::chrono::steady_clock::now();
    const int n = S.n, ns = S.nsuper;
    F = SupernodalFactor();
    F.n = n;
    F.perm = S.perm;
    F.nsuper = ns;
    F.first = S.first;
    F.rowPtr = S.rowPtr;
    F.rows = S.rows;
    F.valPtr = S.valPtr;
    F.parent = S.parent;
    F.level = S.level;
    F.nlevels = S.nlevels;
    F.nnzL = S.nnzL;
    F.flops = S.flops;
    const std::vector<int>& first = S.first;
    const std::vector<int>& Lp = S.Lp;
    const std::vector<int>& Li = S.Li;
    const std::vector<int>& Lsrc = S.Lsrc;
    const std::vector<std::vector<int>>& children = S.children;
    const std::vector<long long>& subtreeCols = S.subtreeCols;
    // numeric multifrontal factorisation, tree-parallel (OpenMP tasks over
    // subtrees; big fronts split their dense kernels into tasks too).  Each
    // supernode's update matrix lives in upd[s] until its parent absorbs it.
    // B is filled supernode by supernode inside the tasks (no serial zero-fill)
    F.B = ws ? ws->takeBlocks(static_cast<size_t>(F.valPtr[ns]))
             : std::unique_ptr<double[]>(new double[static_cast<size_t>(std::max(F.valPtr[ns], 1LL))]);
    F.Bsize = F.valPtr[ns];
    // update matrices: from the workspace's pool (pages already mapped by
    // an earlier factorisation) or the heap
    struct FreeUpdate {
        HostWorkspace* ws;
        void operator()(double* p) const
        {
            if (ws) ws->giveUpdate(p);
            else delete[] p;
        }
    };
    const FreeUpdate freeUpdate{ws};
    using UpdPtr = std::unique_ptr<double[], FreeUpdate>;
    std::vector<UpdPtr> upd;
    upd.reserve(ns);
    for (int s = 0; s < ns; ++s) upd.emplace_back(nullptr, freeUpdate);
    std::atomic<int> failed{0};
    const auto tn0 = std::chrono::steady_clock::now();
    // One supernode: the front's first w columns [F11; F21] are assembled,
    // factored and inverted in place in the supernode's slot P of B (m x w,
    // ld m, the layout B keeps); its trailing (m - w) x (m - w) lower part is
    // the update matrix U handed to the parent.
    auto factorFront = [&](int s) {
        if (failed.load(std::memory_order_relaxed)) return;
        const int j0 = first[s], w = first[s + 1] - j0;
        const int r0 = F.rowPtr[s], m = F.rowPtr[s + 1] - r0;
        const int nr = m - w;
        const int* rows = &F.rows[r0];
        double* P = &F.B[F.valPtr[s]];
        // global row -> front position (no OpenMP scheduling point while a
        // task uses it, so the map is per thread)
        thread_local std::vector<int> loc;
        if (static_cast<int>(loc.size()) < n) loc.resize(n);
        for (int k = 0; k < m; ++k) loc[rows[k]] = k;
        std::fill(P, P + static_cast<size_t>(m) * w, 0.0);
        UpdPtr U(nullptr, freeUpdate);
        if (nr > 0 && F.parent[s] >= 0) {
            const size_t cnt = static_cast<size_t>(nr) * nr;
            U.reset(ws ? ws->takeUpdate(cnt) : new double[cnt]);
            // lower triangle plus the upper part of the 64 x 64 diagonal
            // tiles the update computes in full
            for (int b = 0; b < nr; ++b)
                std::fill(U.get() + static_cast<size_t>(b) * nr + (b / 64) * 64, U.get() + static_cast<size_t>(b + 1) * nr,
                          0.0);
        }
        for (int j = j0; j < j0 + w; ++j) {
            double* col = P + static_cast<size_t>(j - j0) * m;
            for (int p = Lp[j]; p < Lp[j + 1]; ++p) col[loc[Li[p]]] += A.x[Lsrc[p]];
        }
        for (int c : children[s]) {
            const double* Uc = upd[c].get();
            const int cw = first[c + 1] - first[c];
            const int cm = F.rowPtr[c + 1] - F.rowPtr[c] - cw;
            const int* cr = &F.rows[F.rowPtr[c] + cw];
            // child rows are sorted, so a >= b keeps loc[cr[a]] >= loc[cr[b]]
            for (int b = 0; b < cm; ++b) {
                const int cb = loc[cr[b]];
                const double* src = Uc + static_cast<size_t>(b) * cm;
                if (cb < w) {
                    double* dst = P + static_cast<size_t>(cb) * m;
                    for (int a = b; a < cm; ++a) dst[loc[cr[a]]] += src[a];
                } else {
                    double* dst = U.get() + static_cast<size_t>(cb - w) * nr;
                    for (int a = b; a < cm; ++a) dst[loc[cr[a]] - w] += src[a];
                }
            }
            upd[c].reset();
        }
        const bool bigDiag = w >= 256;
        if (!dense::potrf(P, w, m, bigDiag)) {
            failed = 1;
            return;
        }
        const bool big = static_cast<double>(nr) * nr * w > 2e7;
        trsmRightLT(P, w, nr, m, big);
        if (U) {
            syrkLower(P, w, nr, m, U.get(), nr, big);
            upd[s] = std::move(U);
        }
        blockInverse(P, w, m, big || w >= 256);
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
    F.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

bool supernodalCholesky(const Csc& A, const std::vector<int>& perm, SupernodalFactor& F, std::string& msg,
                        int denseSubtree, HostWorkspace* ws)
{
    SupernodalSymbolic S;
    if (!supernodalSymbolic(A, perm, S, msg, denseSubtree)) return false;
    if (!supernodalNumeric(A, S, F, msg, ws)) return false;
    F.seconds += S.seconds;
    return true;
}

void packBlocks(const SupernodalFactor& F, const std::vector<long long>& off, int align, bool f32, void* dst)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int s = 0; s < F.nsuper; ++s) {
        const int w = F.first[s + 1] - F.first[s], m = F.rowPtr[s + 1] - F.rowPtr[s];
        const int ld = (m + align - 1) / align * align;
        for (int k = 0; k < w; ++k) {
            const double* col = F.B.get() + F.valPtr[s] + static_cast<long long>(k) * m;
            const long long o = off[s] + static_cast<long long>(k) * ld;
            if (f32) {
                float* d = static_cast<float*>(dst) + o;
                for (int r = 0; r < m; ++r) d[r] = static_cast<float>(col[r]);
                for (int r = m; r < ld; ++r) d[r] = 0.0f;
            } else {
                double* d = static_cast<double*>(dst) + o;
                std::memcpy(d, col, sizeof(double) * m);
                for (int r = m; r < ld; ++r) d[r] = 0.0;
            }
        }
    }
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

void fillForwardTables(const SupernodalFactor& F, const std::vector<int>& fptr, const std::vector<int>& fidx,
                       const std::vector<FwdTableItem>& items, long long size, int zeroSlot,
                       std::vector<int>& tab)
{
    tab.resize(static_cast<size_t>(size));
    const int ni = static_cast<int>(items.size());
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < ni; ++i) {
        const FwdTableItem& t = items[i];
        const int fr0 = F.rowPtr[t.s], m = F.rowPtr[t.s + 1] - fr0, w = F.first[t.s + 1] - F.first[t.s];
        int* q = tab.data() + t.base;
        std::fill(q, q + static_cast<size_t>(t.K) * (64 + t.W), zeroSlot);
        for (int lane = 0; lane < 64; ++lane) {
            const int r = t.r0 + lane;
            if (r < w || r >= m) continue;
            const int p0 = fptr[fr0 + r], cnt = fptr[fr0 + r + 1] - p0;
            for (int k = 0; k < cnt && k < t.K; ++k) q[static_cast<size_t>(k) * 64 + lane] = fidx[p0 + k];
        }
        int* own = q + static_cast<size_t>(t.K) * 64;
        for (int c = 0; c < w; ++c) {
            const int p0 = fptr[fr0 + c], cnt = fptr[fr0 + c + 1] - p0;
            for (int k = 0; k < cnt && k < t.K; ++k) own[static_cast<size_t>(k) * t.W + c] = fidx[p0 + k];
        }
    }
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
