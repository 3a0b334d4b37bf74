// Plan of the device numeric factorisation (see devplan.hpp).
#include "devplan.hpp"

#include <algorithm>

#include "../limits.hpp"

namespace efem {

namespace {
constexpr int kTile = 64;        // tile edge of the device dense kernels
constexpr int kTrsmRows = 256;   // rows of one triangular-solve task
inline int tiles(int v) { return (v + kTile - 1) / kTile; }
}  // namespace

void buildDevFactorPlan(const SupernodalSymbolic& S, DevFactorPlan& P)
{
    P = DevFactorPlan();
    const int ns = S.nsuper;
    P.n = S.n;
    P.nsuper = ns;
    P.w.resize(ns);
    P.m.resize(ns);
    P.dval.assign(ns + 1, 0);
    P.uOff.assign(ns, -1);
    long long uo = 0;
    for (int s = 0; s < ns; ++s) {
        P.w[s] = S.first[s + 1] - S.first[s];
        P.m[s] = S.rowPtr[s + 1] - S.rowPtr[s];
        P.dval[s + 1] = P.dval[s] + static_cast<long long>(P.w[s]) * blockLd(P.m[s]);
        const int nr = P.m[s] - P.w[s];
        if (nr > 0 && S.parent[s] >= 0) {
            P.uOff[s] = uo;
            uo += static_cast<long long>(nr) * nr;
        }
    }
    P.panelSize = P.dval[ns];
    P.updSize = uo;

    std::vector<std::vector<int>> byLevel(S.nlevels);
    for (int s = 0; s < ns; ++s) byLevel[S.level[s]].push_back(s);

    // child update rows -> positions in the parent's front
    P.relOff.assign(ns, -1);
    std::vector<int> loc(S.n, -1);
    for (int s = 0; s < ns; ++s) {
        if (S.children[s].empty()) continue;
        const int r0 = S.rowPtr[s];
        for (int k = 0; k < P.m[s]; ++k) loc[S.rows[r0 + k]] = k;
        for (int c : S.children[s]) {
            const int cw = P.w[c], cm = P.m[c] - cw;
            P.relOff[c] = static_cast<int>(P.rel.size());
            for (int a = 0; a < cm; ++a) P.rel.push_back(loc[S.rows[S.rowPtr[c] + cw + a]]);
        }
    }

    auto launch = [&](int kind, int first, int count) {
        if (count > 0) P.launches.push_back({kind, first, count});
    };
    auto task = [&](int a, int b, int c, int d) {
        P.tasks.push_back(a);
        P.tasks.push_back(b);
        P.tasks.push_back(c);
        P.tasks.push_back(d);
    };
    auto ntask = [&]() { return static_cast<int>(P.tasks.size() / 4); };

    for (int l = 0; l < S.nlevels; ++l) {
        const std::vector<int>& sl = byLevel[l];
        // A's lower entries of the level's columns -> L arena
        {
            const int a0 = static_cast<int>(P.asmSrc.size());
            for (int s : sl) {
                const int j0 = S.first[s], r0 = S.rowPtr[s], ld = blockLd(P.m[s]);
                for (int k = 0; k < P.m[s]; ++k) loc[S.rows[r0 + k]] = k;
                for (int j = j0; j < j0 + P.w[s]; ++j)
                    for (int p = S.Lp[j]; p < S.Lp[j + 1]; ++p) {
                        P.asmDst.push_back(P.dval[s] + loc[S.Li[p]] + static_cast<long long>(j - j0) * ld);
                        P.asmSrc.push_back(S.Lsrc[p]);
                    }
            }
            launch(DF_ASM, a0, static_cast<int>(P.asmSrc.size()) - a0);
        }
        // extend-add: one task per front column receiving contributions,
        // its (child, child column) pairs in children order
        {
            const int t0 = ntask();
            for (int s : sl) {
                std::vector<std::vector<int>> byCol(P.m[s]);
                for (int c : S.children[s]) {
                    const int cm = P.m[c] - P.w[c];
                    for (int b = 0; b < cm; ++b) byCol[P.rel[P.relOff[c] + b]].push_back(c);
                }
                // the child column b with rel[b] == t, recovered per child
                for (int t = 0; t < P.m[s]; ++t) {
                    if (byCol[t].empty()) continue;
                    const int pb = static_cast<int>(P.pairs.size() / 2);
                    for (int c : byCol[t]) {
                        const int cm = P.m[c] - P.w[c];
                        const int* r = &P.rel[P.relOff[c]];
                        const int b = static_cast<int>(std::lower_bound(r, r + cm, t) - r);
                        P.pairs.push_back(c);
                        P.pairs.push_back(b);
                    }
                    task(s, t, pb, static_cast<int>(P.pairs.size() / 2));
                }
            }
            launch(DF_EXTADD, t0, ntask() - t0);
        }
        // panel factorisation in 64-column steps
        int nbwMax = 0;
        for (int s : sl) nbwMax = std::max(nbwMax, tiles(P.w[s]));
        for (int k = 0; k < nbwMax; ++k) {
            int t0 = ntask();
            for (int s : sl)
                if (tiles(P.w[s]) > k) task(s, k, 0, 0);
            launch(DF_POTRF, t0, ntask() - t0);
            t0 = ntask();
            for (int s : sl) {
                if (tiles(P.w[s]) <= k) continue;
                const int kw = std::min(kTile, P.w[s] - kTile * k);
                for (int r = kTile * k + kw; r < P.m[s]; r += kTrsmRows) task(s, k, r, 0);
            }
            launch(DF_TRSM, t0, ntask() - t0);
            t0 = ntask();
            for (int s : sl) {
                const int nbw = tiles(P.w[s]);
                for (int jt = k + 1; jt < nbw; ++jt)
                    for (int it = jt; it < tiles(P.m[s]); ++it) task(s, k, it, jt);
            }
            launch(DF_UPDATE, t0, ntask() - t0);
        }
        // update matrices
        {
            const int t0 = ntask();
            for (int s : sl) {
                if (P.uOff[s] < 0) continue;
                const int nt = tiles(P.m[s] - P.w[s]);
                for (int it = 0; it < nt; ++it)
                    for (int jt = 0; jt <= it; ++jt) task(s, it, jt, 0);
            }
            launch(DF_SYRK, t0, ntask() - t0);
        }
        // L11^-1, last column tile first, then L21 L11^-1
        for (int J = nbwMax - 1; J >= 0; --J) {
            const int t0 = ntask();
            for (int s : sl) {
                if (tiles(P.w[s]) <= J) continue;
                for (int it = J; it < tiles(P.w[s]); ++it) task(s, J, it, 0);
            }
            launch(DF_BSOLVE, t0, ntask() - t0);
        }
        {
            const int t0 = ntask();
            for (int s : sl)
                for (int it = 0; it < tiles(P.m[s] - P.w[s]); ++it)
                    for (int jt = 0; jt < tiles(P.w[s]); ++jt) task(s, it, jt, 0);
            launch(DF_GEMM21, t0, ntask() - t0);
        }
    }
}

}  // namespace efem

namespace efem {

bool planTop(const SupernodalFactor& F, int zeroSlot, int maxCols, int maxK, int align, TopPlan& P)
{
    P = TopPlan();
    const int ns = F.nsuper;
    if (ns == 0 || maxCols <= 0) return false;
    std::vector<int> w(ns), m(ns), tOff(ns);
    int nt = 0;
    for (int s = 0; s < ns; ++s) {
        w[s] = F.first[s + 1] - F.first[s];
        m[s] = F.rowPtr[s + 1] - F.rowPtr[s];
        tOff[s] = nt;
        nt += m[s] - w[s];
    }
    // supernodes are numbered in postorder: a parent follows its children
    std::vector<int> depth(ns, 0);
    int maxDepth = 0;
    for (int s = ns - 1; s >= 0; --s) {
        depth[s] = F.parent[s] < 0 ? 0 : depth[F.parent[s]] + 1;
        maxDepth = std::max(maxDepth, depth[s]);
    }
    std::vector<long long> colsAt(maxDepth + 1, 0);
    for (int s = 0; s < ns; ++s) colsAt[depth[s]] += w[s];
    long long cum = 0;
    for (int d = 0; d <= maxDepth && cum + colsAt[d] <= maxCols; ++d) {
        cum += colsAt[d];
        P.depth = d;
    }
    if (P.depth < 0) return false;
    P.nT = static_cast<int>(cum);
    P.ld = (P.nT + align - 1) / align * align;
    P.inT.assign(ns, 0);
    std::vector<int> tlOf(F.n, -1);
    std::vector<int> depthStart(P.depth + 2, 0);  // first T-local column of each depth
    for (int d = 0; d <= P.depth; ++d) {
        depthStart[d] = static_cast<int>(P.tcol.size());
        for (int s = 0; s < ns; ++s) {
            if (depth[s] != d) continue;
            P.inT[s] = 1;
            P.super.push_back({s, static_cast<int>(P.tcol.size()), w[s], m[s], 0});
            for (int j = F.first[s]; j < F.first[s + 1]; ++j) {
                tlOf[j] = static_cast<int>(P.tcol.size());
                P.tcol.push_back(j);
            }
        }
    }
    depthStart[P.depth + 1] = P.nT;
    // rows below each T supernode's columns: its ancestors' columns, in T
    for (TopSuperInfo& t : P.super) {
        t.rOff = static_cast<int>(P.rtl.size());
        for (int r = F.rowPtr[t.s] + t.w; r < F.rowPtr[t.s + 1]; ++r) {
            const int q = tlOf[F.rows[r]];
            if (q < 0) return false;  // not closed under ancestors (not a supernodal tree)
            P.rtl.push_back(q);
        }
    }
    // f_T: the update rows of T's children outside T, children in order
    // (the extend-add of the forward sweep)
    std::vector<int> cnt(P.nT, 0);
    int K = 0;
    for (int s = 0; s < ns; ++s) {
        if (P.inT[s] || F.parent[s] < 0 || !P.inT[F.parent[s]]) continue;
        for (int r = F.rowPtr[s] + w[s]; r < F.rowPtr[s + 1]; ++r) {
            const int q = tlOf[F.rows[r]];
            if (q < 0) return false;
            K = std::max(K, ++cnt[q]);
        }
    }
    if (K > maxK) {
        P = TopPlan();
        return false;
    }
    P.K = std::max(K, 1);
    P.tab.assign(static_cast<size_t>(P.K) * P.nT, zeroSlot);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int s = 0; s < ns; ++s) {
        if (P.inT[s] || F.parent[s] < 0 || !P.inT[F.parent[s]]) continue;
        for (int i = 0; i < m[s] - w[s]; ++i) {
            const int q = tlOf[F.rows[F.rowPtr[s] + w[s] + i]];
            P.tab[static_cast<size_t>(cnt[q]++) * P.nT + q] = tOff[s] + i;
        }
    }
    // formation of Z, depth by depth from the roots (64 x 64 tiles)
    constexpr int TB = 64;
    auto task = [&](int a, int b, int c, int d) {
        P.tasks.push_back(a);
        P.tasks.push_back(b);
        P.tasks.push_back(c);
        P.tasks.push_back(d);
    };
    auto ntasks = [&]() { return static_cast<int>(P.tasks.size() / 4); };
    for (int d = 0; d <= P.depth; ++d) {
        const int p0 = depthStart[d], p1 = depthStart[d + 1];
        auto addGemm = [&](int r0, int r1) {
            TopPlan::Step st{0, ntasks(), 0, 0, 0};
            for (size_t q = 0; q < P.super.size(); ++q) {
                const TopSuperInfo& t = P.super[q];
                if (t.cs < p0 || t.cs >= p1 || t.m == t.w) continue;
                for (int r = r0; r < r1; r += TB)
                    for (int j = 0; j < t.w; j += TB) task(static_cast<int>(q), r, j, r1);
            }
            st.count = ntasks() - st.first;
            if (st.count > 0) P.steps.push_back(st);
        };
        if (d > 0) {
            addGemm(0, p0);
            P.steps.push_back({1, 0, 0, p0, p1});
            addGemm(p0, p1);
        }
        TopPlan::Step st{2, ntasks(), 0, 0, 0};
        for (size_t q = 0; q < P.super.size(); ++q) {
            const TopSuperInfo& t = P.super[q];
            if (t.cs < p0 || t.cs >= p1) continue;
            for (int i = 0; i < t.w; i += TB)
                for (int j = 0; j < t.w; j += TB) task(static_cast<int>(q), i, j, 0);
        }
        st.count = ntasks() - st.first;
        if (st.count > 0) P.steps.push_back(st);
    }
    return true;
}

}  // namespace efem
