// Persistent batched solver: many independent instances (load cases or
// unevenly sized meshes) in ONE launch, one 256-lane workgroup per instance.
// Each workgroup runs the reference's L-BFGS loop (LBFGSSolver::solve,
// LBFGS.h:96-174) with the projection line search (LineSearchFEA.h:37-68)
// and the two-loop recursion in LBFGSMat::apply_Hv order (LBFGSMat.h:
// 103-131) on its instance's ranges of the merged model, with no host
// round trip: the residual (element tiles, pressure faces, gather), the
// preconditioner (the instance's supernodal factor, level by level), and
// all dot products are workgroup-local, separated by barriers.  Reductions
// use fixed wave/block trees, so an instance's result does not depend on
// the other instances in the batch.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include <type_traits>
#include "kernels.hpp"

// This file is compiled once per workgroup width (256, 512 and 1024 lanes,
// one object each, -DEFEM_BATCH_BT=...): the autotuner picks the width.
#ifndef EFEM_BATCH_BT
#define EFEM_BATCH_BT 256
#endif

namespace efem {
namespace dev {
namespace {

constexpr int BT = EFEM_BATCH_BT;  // lanes per workgroup (one instance)
constexpr int BW = BT / 64;        // waves per workgroup

// K simultaneous block reductions over the BT lanes (wave trees, then the
// waves in order); op 0 = sum, 1 = max
template <int K>
__device__ __forceinline__ void blockReduceK(double (&v)[K], const int (&op)[K], double (*sh)[BW])
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double t = op[k] ? waveMax(v[k]) : waveSum(v[k]);
        if (lane == 0) sh[k][wave] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double t = sh[k][0];
#pragma unroll
        for (int w = 1; w < BW; ++w) t = op[k] ? maxNan(t, sh[k][w]) : t + sh[k][w];
        v[k] = t;
    }
    __syncthreads();
}

template <int NEN, int NG, typename TF>
union BatchLds {
    ElemTileLds<float, NEN, NG, MODE_FORCE, BT> ef;
    ElemTileLds<double, NEN, NG, MODE_FORCE, BT> ed;
    TrsvFwdLds<TF, BW> tf;
    TF zf[kBatchDenseMaxN * 3];  // the dense apply's right-hand side
};

template <typename T, int NEN, int NG, int LAW>
__device__ __forceinline__ void blockElements(const BatchArgs& a, const BatchRange& I,
                                              ElemTileLds<T, NEN, NG, MODE_FORCE, BT>& L)
{
    const ElemArgs& ea = a.elem;
    T* slots = static_cast<T*>(sizeof(T) == 4 ? a.slots32 : ea.slots);
    const double* ginv = static_cast<const double*>(ea.ginv);
    if (NEN == 4) {
        for (int e = I.e0 + threadIdx.x; e < I.e1; e += BT)
            tet4Lane<T, LAW, MODE_FORCE>(e, ea.nElems, ea.connDof, ginv, ea.vol, 1.0, ea.params, ea.paramsPerElem,
                                          a.x, nullptr, slots);
    } else {
        const double* gdN = static_cast<const double*>(ea.gdN);
        for (int k = threadIdx.x; k < NG * NEN * 3; k += BT) L.sdN[k] = gdN[k];
        for (int e0 = I.e0; e0 < I.e1; e0 += Tile<NG, BT>::epb)
            elemTile<T, NEN, NG, LAW, MODE_FORCE, BT>(ea.nElems, e0, I.e1, ea.connDof, ginv, ea.vol, ea.gw,
                                                       ea.params, ea.paramsPerElem, a.x, nullptr, slots, L);
    }
}

// r = f_int(x) - f_ext(x) on the instance's free nodes (x = a.x, r = a.r)
template <typename T, int NEN, int NG, typename TF>
__device__ void blockResidual(const BatchArgs& a, const BatchRange& I, BatchLds<NEN, NG, TF>& U)
{
    ElemTileLds<T, NEN, NG, MODE_FORCE, BT>& L =
        *reinterpret_cast<ElemTileLds<T, NEN, NG, MODE_FORCE, BT>*>(&U);
    switch (a.elem.law) {
        case LAW_NH: blockElements<T, NEN, NG, LAW_NH>(a, I, L); break;
        case LAW_MR: blockElements<T, NEN, NG, LAW_MR>(a, I, L); break;
        default: blockElements<T, NEN, NG, LAW_YEOH>(a, I, L);
    }
    for (int f = I.f0 + threadIdx.x; f < I.f1; f += BT) pressureFace<NEN, PMODE_FORCE>(a.face, f);
    __syncthreads();
    GatherArgs g = a.gather;
    g.mixed = sizeof(T) == 4;
    if (sizeof(T) == 4) g.eslots = a.slots32;
    for (int i = I.n0 + threadIdx.x; i < I.n1; i += BT) gatherNode<T>(g, i);
    __syncthreads();
}

// out = K^-1 in on the instance (its supernodes of the merged forest): the
// forward items split their columns over the BW waves, the backward items
// give each of their four columns BW / 4 waves
template <typename TF, int NEN, int NG>
__device__ void blockH0(const BatchArgs& a, int inst, const BatchRange& I, const double* in, double* out,
                        BatchLds<NEN, NG, TF>& U, TF (*bred)[3])
{
    const TrsvArgs& t = a.trsv;
    for (int j = I.n0 + threadIdx.x; j < I.n1; j += BT) {
        const int o = t.perm[j];
        a.bp[3 * j] = in[3 * o];
        a.bp[3 * j + 1] = in[3 * o + 1];
        a.bp[3 * j + 2] = in[3 * o + 2];
    }
    __syncthreads();
    const int* li = a.instItems + 4 * static_cast<size_t>(inst) * a.nlev;
    for (int l = 0; l < a.nlev; ++l) {
        const int off = li[4 * l], cnt = li[4 * l + 1];
        for (int it = 0; it < cnt; ++it)
            trsvFwdItem<TF, BW>(t, a.bp, a.items[2 * (off + it)], a.items[2 * (off + it) + 1], U.tf);
        __syncthreads();
    }
    for (int l = a.nlev - 1; l >= 0; --l) {
        const int off = li[4 * l + 2], cnt = li[4 * l + 3];
        for (int it = 0; it < cnt; ++it) {
            trsvBwdItem<TF, BW>(t, a.items[2 * (off + it)], a.items[2 * (off + it) + 1], bred);
            // with several waves per column the item's partial sums sit in
            // bred: the next item may overwrite them only after this one
            // combined them
            if (BW > 4) __syncthreads();
        }
        __syncthreads();
    }
    const TF* z = static_cast<const TF*>(t.z);
    for (int j = I.n0 + threadIdx.x; j < I.n1; j += BT) {
        const int o = t.perm[j];
        out[3 * o] = static_cast<double>(z[3 * j]);
        out[3 * o + 1] = static_cast<double>(z[3 * j + 1]);
        out[3 * o + 2] = static_cast<double>(z[3 * j + 2]);
    }
    __syncthreads();
}

// out = K^-1 in on the instance through its dense Z_i: f_t = in at the
// instance's factor column t (staged in LDS), then z_k = sum_t Z_i[k, t] f_t
// (Z_i symmetric: row k is read as column k's transpose, so a wave's loads
// of one t are one contiguous segment).  Each lane owns RPT consecutive
// rows (one 16-byte load per t) and a wave 64 RPT rows; waves take the row
// groups in turn; the sum over t runs in t order.
template <typename TF>
__device__ __noinline__ void blockH0Dense(const void* zArena, long long zoff, int ld, const int* __restrict__ tc,
                                             const int* __restrict__ perm, int n, const double* in, double* out,
                                             TF* f)
{
    constexpr int RPT = 16 / sizeof(TF);
    using V = typename std::conditional<sizeof(TF) == 4, float4, double2>::type;
    const TF* Z = static_cast<const TF*>(zArena) + zoff;
    for (int t = threadIdx.x; t < n; t += BT) {
        const int o = perm[tc[t]];
        f[3 * t] = static_cast<TF>(in[3 * o]);
        f[3 * t + 1] = static_cast<TF>(in[3 * o + 1]);
        f[3 * t + 2] = static_cast<TF>(in[3 * o + 2]);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int UN = 16;  // loads in flight per lane
    for (int k0 = wave * 64 * RPT; k0 < n; k0 += BW * 64 * RPT) {
        const int k = k0 + lane * RPT;  // this lane's first row (< ld: rows n .. ld - 1 are zero)
        TF acc[RPT][3];
#pragma unroll
        for (int q = 0; q < RPT; ++q) acc[q][0] = acc[q][1] = acc[q][2] = TF(0);
        const bool on = k < n;
        const TF* zk = Z + k;
        int t = 0;
        for (; t + UN <= n; t += UN) {
            V zv[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u)
                zv[u] = on ? *reinterpret_cast<const V*>(zk + static_cast<size_t>(t + u) * ld) : V{};
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const TF* zz = reinterpret_cast<const TF*>(&zv[u]);
                const TF f0 = f[3 * (t + u)], f1 = f[3 * (t + u) + 1], f2 = f[3 * (t + u) + 2];
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    acc[q][0] += zz[q] * f0;
                    acc[q][1] += zz[q] * f1;
                    acc[q][2] += zz[q] * f2;
                }
            }
        }
        for (; t < n; ++t) {
            const V zv = on ? *reinterpret_cast<const V*>(zk + static_cast<size_t>(t) * ld) : V{};
            const TF* zz = reinterpret_cast<const TF*>(&zv);
            const TF f0 = f[3 * t], f1 = f[3 * t + 1], f2 = f[3 * t + 2];
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                acc[q][0] += zz[q] * f0;
                acc[q][1] += zz[q] * f1;
                acc[q][2] += zz[q] * f2;
            }
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            if (k + q < n) {
                const int o = perm[tc[k + q]];
                out[3 * o] = static_cast<double>(acc[q][0]);
                out[3 * o + 1] = static_cast<double>(acc[q][1]);
                out[3 * o + 2] = static_cast<double>(acc[q][2]);
            }
        }
    }
    __syncthreads();
}

template <int NEN, int NG, typename TF>
__device__ __forceinline__ void residualAt(const BatchArgs& a, const BatchRange& I, bool f32,
                                           BatchLds<NEN, NG, TF>& U)
{
    if (f32)
        blockResidual<float, NEN, NG, TF>(a, I, U);
    else
        blockResidual<double, NEN, NG, TF>(a, I, U);
}

// histories of at most kBC pairs take the compact two-loop (one reduction
// for the new pair's Gram row / column and the first loop's dots, one for
// the second loop's dots); longer ones the reference's order
constexpr int kBC = 8;

// The compact two-loop of one instance (the reference's LBFGSMat
// add_correction + apply_Hv recursion regrouped; pairs by age t: slot
// (loc - t) mod m, age 0 the new pair).  First half: the new pair s = x -
// xp, y = r - rp is stored in slot loc and, in the same pass, s.y, y.y, its
// Gram row / column s_loc.y_k, s_k.y_loc and s_k.r of every pair are
// summed; then the first loop newest -> oldest (q_0 = -r): alpha_t =
// (s_t.q_0 - sum_{u newer} alpha_u s_t.y_u) / ys_t, and q = -r - sum_t
// alpha_t y_t into qOut.  gram (kBC x kBC, s_i.y_k by slot), ys and alpha
// live in LDS for the whole solve.  Returns theta = y.y / s.y.  Called by
// every lane of the block (contains barriers).  Its own function (with
// plain arguments): inlined into the persistent kernel its arrays made
// the kernel spill.
__device__ __noinline__ double compactFirst(const double* __restrict__ x, const double* __restrict__ xp,
                                            const double* __restrict__ r, const double* __restrict__ rp, double* S,
                                            double* Y, size_t nt, int m, int loc, int ncn, int d0, int d1,
                                            double* ys, double* alpha, double* gramp, double* redp,
                                            double* __restrict__ qOut)
{
    double(*gram)[kBC] = reinterpret_cast<double(*)[kBC]>(gramp);
    double(*red)[BW] = reinterpret_cast<double(*)[BW]>(redp);
    const int tid = threadIdx.x;
    int slot[kBC];
#pragma unroll
    for (int t = 0; t < kBC; ++t) slot[t] = (loc - t + m) % m;
    double* sl = S + loc * nt;
    double* yl = Y + loc * nt;
    constexpr int K = 3 * kBC + 2;
    double v[K];
    int op[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        v[c] = 0.0;
        op[c] = 0;
    }
    for (int i = d0 + tid; i < d1; i += BT) {
        const double ri = r[i];
        const double si = x[i] - xp[i], yi = ri - rp[i];
        double sk[kBC], yk[kBC];
#pragma unroll
        for (int t = 1; t < kBC; ++t) {
            sk[t] = t < ncn ? S[slot[t] * nt + i] : 0.0;
            yk[t] = t < ncn ? Y[slot[t] * nt + i] : 0.0;
        }
        sl[i] = si;
        yl[i] = yi;
        v[0] += si * yi;
        v[1] += yi * yi;
        v[2 + 2 * kBC] += si * ri;
#pragma unroll
        for (int t = 1; t < kBC; ++t) {
            v[2 + t] += si * yk[t];
            v[2 + kBC + t] += sk[t] * yi;
            v[2 + 2 * kBC + t] += sk[t] * ri;
        }
    }
    blockReduceK<K>(v, op, red);
    if (tid == 0) {
        ys[loc] = v[0];
        gram[loc][loc] = v[0];
        for (int t = 1; t < ncn; ++t) {
            gram[loc][slot[t]] = v[2 + t];
            gram[slot[t]][loc] = v[2 + kBC + t];
        }
    }
    __syncthreads();
    double al[kBC];
#pragma unroll
    for (int t = 0; t < kBC; ++t) {
        al[t] = 0.0;
        if (t < ncn) {
            double acc = -v[2 + 2 * kBC + t];
#pragma unroll
            for (int u = 0; u < t; ++u) acc -= al[u] * gram[slot[t]][slot[u]];
            al[t] = acc / ys[slot[t]];
        }
    }
    if (tid == 0) {
#pragma unroll
        for (int t = 0; t < kBC; ++t)
            if (t < ncn) alpha[slot[t]] = al[t];
    }
    for (int i = d0 + tid; i < d1; i += BT) {
        double yv[kBC];
#pragma unroll
        for (int t = 0; t < kBC; ++t) yv[t] = t < ncn ? Y[slot[t] * nt + i] : 0.0;
        double qi = -r[i];
#pragma unroll
        for (int t = 0; t < kBC; ++t)
            if (t < ncn) qi -= al[t] * yv[t];
        qOut[i] = qi;
    }
    __syncthreads();
    return v[1] / v[0];
}

// Second half of the compact two-loop: with h = H0 q, the dots y_t.h in one
// pass, then oldest -> newest beta_t = (y_t.h + sum_{u older} coef_u
// y_t.s_u) / ys_t, coef_t = alpha_t - beta_t, and d = h + sum_t coef_t s_t.
__device__ __noinline__ void compactSecond(const double* S, const double* Y, size_t nt, int m, int loc, int ncn,
                                           int d0, int d1, const double* ys, const double* alpha,
                                           const double* gramp, double* redp, const double* __restrict__ h,
                                           double* __restrict__ dOut)
{
    const double(*gram)[kBC] = reinterpret_cast<const double(*)[kBC]>(gramp);
    double(*red)[BW] = reinterpret_cast<double(*)[BW]>(redp);
    const int tid = threadIdx.x;
    int slot[kBC];
#pragma unroll
    for (int t = 0; t < kBC; ++t) slot[t] = (loc - t + m) % m;
    double w[kBC];
    int op[kBC];
#pragma unroll
    for (int c = 0; c < kBC; ++c) {
        w[c] = 0.0;
        op[c] = 0;
    }
    for (int i = d0 + tid; i < d1; i += BT) {
        const double hi = h[i];
        double yv[kBC];
#pragma unroll
        for (int t = 0; t < kBC; ++t) yv[t] = t < ncn ? Y[slot[t] * nt + i] : 0.0;
#pragma unroll
        for (int t = 0; t < kBC; ++t) w[t] += yv[t] * hi;
    }
    blockReduceK<kBC>(w, op, red);
    double cf[kBC];
#pragma unroll
    for (int t = kBC - 1; t >= 0; --t) {
        cf[t] = 0.0;
        if (t < ncn) {
            double acc = w[t];
#pragma unroll
            for (int u = t + 1; u < kBC; ++u)
                if (u < ncn) acc += cf[u] * gram[slot[u]][slot[t]];
            cf[t] = alpha[slot[t]] - acc / ys[slot[t]];
        }
    }
    for (int i = d0 + tid; i < d1; i += BT) {
        double sv[kBC];
#pragma unroll
        for (int t = 0; t < kBC; ++t) sv[t] = t < ncn ? S[slot[t] * nt + i] : 0.0;
        double di = h[i];
#pragma unroll
        for (int t = kBC - 1; t >= 0; --t)
            if (t < ncn) di += cf[t] * sv[t];
        dOut[i] = di;
    }
    __syncthreads();
}

template <int NEN, int NG, typename TF>
__global__ __launch_bounds__(BT) void kBatchLbfgs(BatchArgs a)
{
    __shared__ BatchLds<NEN, NG, TF> U;
    // the compact two-loop's reductions carry up to 3 kBC + 2 components
    __shared__ double red[3 * kBC + 2][BW];
    __shared__ double gram[kBC][kBC];  // s_i . y_k of the stored pairs (compact two-loop)
    __shared__ TF bred[BW][3];
    __shared__ double ys[kMaxCompact], alpha[kMaxCompact];
    const int inst = blockIdx.x;
    const BatchRange I = a.ranges[inst];
    const int d0 = 3 * I.n0, d1 = 3 * I.n1;
    const int tid = threadIdx.x;
    const int m = a.m;
    const size_t nt = static_cast<size_t>(a.ntot);
    double* x = a.x;
    double* r = a.r;
    double* d = a.d;
    double* q = a.q;
    const double* xt = a.xt;
    bool f32 = a.elemF32 != 0;
    int rEvals = 0, lsTrials = 0, iters = 0, switchIt = 0, nonfinite = 0, reached = 0;
    double tnorm2 = 0.0, terr = 0.0;

    residualAt<NEN, NG, TF>(a, I, f32, U);
    ++rEvals;
    double rinf, xinf;
    {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        const int op[4] = {0, 1, 1, 0};
        for (int i = d0 + tid; i < d1; i += BT) {
            v[0] += x[i] * x[i];
            v[1] = maxNan(v[1], fabs(x[i]));
            v[2] = maxNan(v[2], fabs(r[i]));
            if (xt) v[3] += xt[i] * xt[i];
        }
        blockReduceK<4>(v, op, red);
        rinf = v[2];
        xinf = v[1];
        tnorm2 = v[3];
        const double xnorm2 = sqrt(v[0]);
        if (rinf <= a.eps || rinf <= a.epsRel * xnorm2) {
            iters = 1;
            goto done;
        }
    }
    {
        const double rinf0 = rinf;
        const bool h0 = a.usePrecond != 0;
        // d = -H0 r
        if (h0) {
            if (a.zArena)
                blockH0Dense<TF>(a.zArena, a.zoff[inst], a.zld[inst], a.ztcol + I.n0, a.trsv.perm, I.n1 - I.n0, r, d,
                                 U.zf);
            else blockH0<TF, NEN, NG>(a, inst, I, r, d, U, bred);
            for (int i = d0 + tid; i < d1; i += BT) d[i] = -d[i];
        } else {
            for (int i = d0 + tid; i < d1; i += BT) d[i] = -r[i];
        }
        __syncthreads();
        int ncorr = 0, ptr = m;
        double theta = 1.0;
        for (int k = 1;; ++k) {
            for (int i = d0 + tid; i < d1; i += BT) {
                a.xp[i] = x[i];
                a.rp[i] = r[i];
            }
            double dgInit;
            {
                double v[1] = {0.0};
                const int op[1] = {0};
                for (int i = d0 + tid; i < d1; i += BT) v[0] += r[i] * d[i];
                blockReduceK<1>(v, op, red);
                dgInit = v[0];
            }
            double cur = 1.0;
            int trials = 0;
            double dist2 = 0.0;
            for (int it = 0; it < a.maxLs; ++it) {
                for (int i = d0 + tid; i < d1; i += BT) x[i] = a.xp[i] + cur * d[i];
                __syncthreads();
                residualAt<NEN, NG, TF>(a, I, f32, U);
                ++rEvals;
                ++trials;
                double v[4] = {0.0, 0.0, 0.0, 0.0};
                const int op[4] = {0, 1, 1, 0};
                for (int i = d0 + tid; i < d1; i += BT) {
                    v[0] += r[i] * d[i];
                    v[1] = maxNan(v[1], fabs(r[i]));
                    v[2] = maxNan(v[2], fabs(x[i]));
                    if (xt) v[3] += (x[i] - xt[i]) * (x[i] - xt[i]);
                }
                blockReduceK<4>(v, op, red);
                rinf = v[1];
                xinf = v[2];
                dist2 = v[3];
                if (fabs(v[0]) <= a.lnsFea * fabs(dgInit)) break;
                cur *= v[0] * dgInit > 0.0 ? 1.5 : 0.5;
            }
            lsTrials += trials;
            iters = k;
            if (!isfinite(rinf) || !isfinite(xinf)) {
                nonfinite = 1;
                break;
            }
            if (xt) {
                // the displacement target (a measurement hook, as in the
                // single-instance solve): distance of the accepted iterate
                terr = sqrt(dist2 / tnorm2);
                if (terr <= a.targetTol) {
                    reached = 1;
                    break;
                }
            }
            if (rinf <= a.eps || rinf <= a.epsRel * xinf) break;
            if (k >= a.maxIt) break;
            if (f32 && a.fp64Switch > 0.0 && rinf <= a.fp64Switch * rinf0) {
                f32 = false;
                switchIt = k;
                residualAt<NEN, NG, TF>(a, I, f32, U);
                ++rEvals;
            }
            // add_correction (LBFGSMat.h:81-98)
            const int loc = ptr % m;
            double* sl = a.S + loc * nt;
            double* yl = a.Y + loc * nt;
            if (m <= kBC) {
                // compact two-loop (the reference's recursion regrouped)
                const int ncn = ncorr < m ? ncorr + 1 : m;
                theta = compactFirst(x, a.xp, r, a.rp, a.S, a.Y, nt, m, loc, ncn, d0, d1, ys, alpha, &gram[0][0],
                                     &red[0][0], d);
                ncorr = ncn;
                ptr = loc + 1;
                if (h0) {
                    if (a.zArena)
                        blockH0Dense<TF>(a.zArena, a.zoff[inst], a.zld[inst], a.ztcol + I.n0, a.trsv.perm,
                                         I.n1 - I.n0, d, q, U.zf);
                    else blockH0<TF, NEN, NG>(a, inst, I, d, q, U, bred);
                } else {
                    for (int i = d0 + tid; i < d1; i += BT) q[i] = d[i] / theta;
                    __syncthreads();
                }
                compactSecond(a.S, a.Y, nt, m, loc, ncn, d0, d1, ys, alpha, &gram[0][0], &red[0][0], q, d);
                continue;
            }
            // reference order (histories longer than kBC)
            {
                double v[2] = {0.0, 0.0};
                const int op[2] = {0, 0};
                for (int i = d0 + tid; i < d1; i += BT) {
                    const double si = x[i] - a.xp[i], yi = r[i] - a.rp[i];
                    sl[i] = si;
                    yl[i] = yi;
                    v[0] += si * yi;
                    v[1] += yi * yi;
                }
                blockReduceK<2>(v, op, red);
                if (tid == 0) ys[loc] = v[0];
                theta = v[1] / v[0];
            }
            if (ncorr < m) ++ncorr;
            ptr = loc + 1;
            // two-loop recursion d = -H r (LBFGSMat::apply_Hv order)
            for (int i = d0 + tid; i < d1; i += BT) d[i] = -r[i];
            __syncthreads();
            int j = ptr % m;
            for (int c = 0; c < ncorr; ++c) {
                j = (j + m - 1) % m;
                const double* sj = a.S + j * nt;
                const double* yj = a.Y + j * nt;
                double v[1] = {0.0};
                const int op[1] = {0};
                for (int i = d0 + tid; i < d1; i += BT) v[0] += sj[i] * d[i];
                blockReduceK<1>(v, op, red);
                const double aj = v[0] / ys[j];
                if (tid == 0) alpha[j] = aj;
                for (int i = d0 + tid; i < d1; i += BT) d[i] -= aj * yj[i];
                __syncthreads();
            }
            if (h0) {
                if (a.zArena)
                    blockH0Dense<TF>(a.zArena, a.zoff[inst], a.zld[inst], a.ztcol + I.n0, a.trsv.perm, I.n1 - I.n0, d,
                                     q, U.zf);
                else blockH0<TF, NEN, NG>(a, inst, I, d, q, U, bred);
                for (int i = d0 + tid; i < d1; i += BT) d[i] = q[i];
            } else {
                for (int i = d0 + tid; i < d1; i += BT) d[i] = d[i] / theta;
            }
            __syncthreads();
            for (int c = 0; c < ncorr; ++c) {
                const double* sj = a.S + j * nt;
                const double* yj = a.Y + j * nt;
                double v[1] = {0.0};
                const int op[1] = {0};
                for (int i = d0 + tid; i < d1; i += BT) v[0] += yj[i] * d[i];
                blockReduceK<1>(v, op, red);
                const double coef = alpha[j] - v[0] / ys[j];
                for (int i = d0 + tid; i < d1; i += BT) d[i] += coef * sj[i];
                __syncthreads();
                j = (j + 1) % m;
            }
        }
    }
done:
    if (tid == 0) {
        double* st = a.stats + kBatchStats * static_cast<size_t>(inst);
        st[0] = iters;
        st[1] = rEvals;
        st[2] = lsTrials;
        st[3] = rinf;
        st[4] = xinf;
        st[5] = nonfinite;
        st[6] = switchIt;
        st[7] = a.usePrecond;
        st[8] = reached;
        st[9] = terr;
    }
}

template <typename TF>
void launchBatchKind(const BatchArgs& a, int count, hipStream_t s)
{
    const int nen = a.elem.nen, ng = a.elem.ng;
    if (nen == 4) kBatchLbfgs<4, 1, TF><<<count, BT, 0, s>>>(a);
    else if (nen == 8 && ng == 8) kBatchLbfgs<8, 8, TF><<<count, BT, 0, s>>>(a);
    else if (nen == 8 && ng == 1) kBatchLbfgs<8, 1, TF><<<count, BT, 0, s>>>(a);
    else if (nen == 10 && ng == 4) kBatchLbfgs<10, 4, TF><<<count, BT, 0, s>>>(a);
    else if (nen == 10 && ng == 1) kBatchLbfgs<10, 1, TF><<<count, BT, 0, s>>>(a);
    else if (nen == 20 && ng == 27) kBatchLbfgs<20, 27, TF><<<count, BT, 0, s>>>(a);
    else if (nen == 20 && ng == 8) kBatchLbfgs<20, 8, TF><<<count, BT, 0, s>>>(a);
}

}  // namespace

#define EFEM_CAT2(a, b) a##b
#define EFEM_CAT(a, b) EFEM_CAT2(a, b)
void EFEM_CAT(launchBatchSolveW, EFEM_BATCH_BT)(const BatchArgs& a, int count, hipStream_t s)
{
    if (count <= 0) return;
    if (a.trsv.mixed)
        launchBatchKind<float>(a, count, s);
    else
        launchBatchKind<double>(a, count, s);
}

}  // namespace dev
}  // namespace efem
