// HIP kernels of the solve loop for gfx950 (CDNA4, wave64).
//
//  * element force: one lane per (element, Gauss point); the element's
//    relative nodal displacements and the local-gradient table are staged in
//    LDS; per-Gauss-point stress moments Q_q = w_q V P J0^-T go through LDS
//    to a second phase with one lane per (element, node) that writes the
//    nodal forces f_a = sum_q Q_q dN_a(xi_q) to a per-(element, slot)
//    buffer.  Linear tets (one Gauss point) use a one-lane-per-element kernel.
//  * assembly: deterministic gather over a node -> slot CSR, FP64 sums, no
//    float atomics (replaces the scatter of ElementConstraint.cpp:154-162).
//  * reductions: fixed-shape block trees + an in-order final pass by the
//    last block to arrive (integer arrival counter), so every dot product is
//    bitwise reproducible.
//  * preconditioner: level-scheduled supernodal triangular solves on the
//    block-inverse form B_s = [L11^-1 ; L21 L11^-1] with three right-hand
//    sides (the interleaved x, y, z components); the forward sweep is
//    multifrontal (each supernode extend-adds only its children's update
//    vectors), the backward sweep reads the ancestors' solved values.
#include <hip/hip_runtime.h>

#include "../material_math.hpp"
#include "kernels.hpp"

namespace efem {
namespace dev {
namespace {

constexpr int kThreads = 256;

// H = Gu J0^-1 is formed in FP64 in both modes: J - 1 (and with it the
// near-incompressible volumetric stress (2/D1)(J-1) cof F, 2/D1 = 2e6 here)
// inherits the rounding of H, so H must carry more than FP32's 24 bits.
// The deviatoric stress and the force moments then run in T.
template <typename T, typename GT>
__device__ __forceinline__ void stressMoment(int law, const double* p, const GT* Gu, const double* Jd,
                                             double wv, T* Q)
{
    double Hd[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Hd[3 * i + j] = static_cast<double>(Gu[3 * i]) * Jd[j] + static_cast<double>(Gu[3 * i + 1]) * Jd[3 + j] +
                            static_cast<double>(Gu[3 * i + 2]) * Jd[6 + j];
    T P[9];
    piolaFromHd<T>(law, p, Hd, P);
    T Ji[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Ji[k] = static_cast<T>(Jd[k]);
    const T w = static_cast<T>(wv);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            Q[3 * i + k] = w * (P[3 * i] * Ji[3 * k] + P[3 * i + 1] * Ji[3 * k + 1] + P[3 * i + 2] * Ji[3 * k + 2]);
}

// ---- linear tetrahedra: one lane per element ------------------------------
template <typename T, int LAW>
__global__ __launch_bounds__(kThreads) void kElemTet4(int E, const int* __restrict__ connDof,
                                                      const double* __restrict__ ginv,
                                                      const double* __restrict__ vol, double w0,
                                                      const double* __restrict__ params, int ppe,
                                                      const double* __restrict__ x,
                                                      T* __restrict__ slots)
{
    const int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= E) return;
    int id[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) id[a] = connDof[a * E + e];
    // dN = (e1, e2, e3, -1 -1 -1): Gu[:,k] = u_k - u_3
    double u3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) u3[c] = x[3 * id[3] + c];
    double Gu[9];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i) Gu[3 * i + k] = x[3 * id[k] + i] - u3[i];
    double Jd[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Jd[k] = ginv[k * E + e];
    T Q[9];
    stressMoment<T, double>(LAW, params + (ppe ? 4 * e : 0), Gu, Jd, vol[e] * w0, Q);
    T* out = slots + 12 * static_cast<size_t>(e);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        out[i] = Q[3 * i];
        out[3 + i] = Q[3 * i + 1];
        out[6 + i] = Q[3 * i + 2];
        out[9 + i] = -(Q[3 * i] + Q[3 * i + 1] + Q[3 * i + 2]);
    }
}

// ---- general elements: LDS-staged (element, Gauss point) lanes ------------
template <int NG>
struct Tile {
    static constexpr int epb = NG >= 4 ? kThreads / NG : 64;  // elements per block
};

template <typename T, int NEN, int NG, int LAW>
__global__ __launch_bounds__(kThreads) void kElemMulti(int E, const int* __restrict__ connDof,
                                                       const double* __restrict__ ginv,
                                                       const double* __restrict__ vol,
                                                       const double* __restrict__ gdN,
                                                       const double* __restrict__ gw,
                                                       const double* __restrict__ params, int ppe,
                                                       const double* __restrict__ x,
                                                       T* __restrict__ slots)
{
    constexpr int EPB = Tile<NG>::epb;
    // the kinematics (relative displacements, local gradients, H) are FP64 in
    // both modes; the stress moments Q and the nodal forces use T
    __shared__ double sdN[NG * NEN * 3];
    __shared__ double su[EPB * NEN * 3];
    __shared__ T sQ[EPB * NG * 9];
    const int tid = threadIdx.x;
    const int e0 = blockIdx.x * EPB;
    for (int k = tid; k < NG * NEN * 3; k += kThreads) sdN[k] = gdN[k];
    // relative displacements u_a - u_0
    for (int k = tid; k < EPB * NEN; k += kThreads) {
        const int el = k / NEN, a = k - el * NEN, e = e0 + el;
        if (e < E) {
            const int id = connDof[a * E + e], id0 = connDof[e];
#pragma unroll
            for (int c = 0; c < 3; ++c) su[3 * k + c] = x[3 * id + c] - x[3 * id0 + c];
        }
    }
    __syncthreads();
    {
        const int el = tid / NG, q = tid - el * NG, e = e0 + el;
        if (el < EPB && e < E) {
            double Gu[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) Gu[k] = 0.0;
            const double* ue = su + el * NEN * 3;
            const double* dq = sdN + q * NEN * 3;
            for (int a = 1; a < NEN; ++a) {
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int k = 0; k < 3; ++k) Gu[3 * i + k] += ue[3 * a + i] * dq[3 * a + k];
            }
            double Jd[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) Jd[k] = ginv[(static_cast<size_t>(q) * 9 + k) * E + e];
            stressMoment<T, double>(LAW, params + (ppe ? 4 * e : 0), Gu, Jd, vol[e] * gw[q], sQ + (el * NG + q) * 9);
        }
    }
    __syncthreads();
    for (int k = tid; k < EPB * NEN; k += kThreads) {
        const int el = k / NEN, a = k - el * NEN, e = e0 + el;
        if (e >= E) continue;
        T f[3] = {T(0), T(0), T(0)};
        for (int q = 0; q < NG; ++q) {
            const T* Q = sQ + (el * NG + q) * 9;
            const double* d = sdN + (q * NEN + a) * 3;
            const T d0 = static_cast<T>(d[0]), d1 = static_cast<T>(d[1]), d2 = static_cast<T>(d[2]);
#pragma unroll
            for (int i = 0; i < 3; ++i) f[i] += Q[3 * i] * d0 + Q[3 * i + 1] * d1 + Q[3 * i + 2] * d2;
        }
        T* out = slots + (static_cast<size_t>(e) * NEN + a) * 3;
        out[0] = f[0];
        out[1] = f[1];
        out[2] = f[2];
    }
}

template <typename T, int NEN, int NG>
void launchMultiLaw(const ElemArgs& a, hipStream_t s)
{
    const int EPB = Tile<NG>::epb;
    const int grid = (a.nElems + EPB - 1) / EPB;
    const double* ginv = static_cast<const double*>(a.ginv);
    const double* gdN = static_cast<const double*>(a.gdN);
    T* slots = static_cast<T*>(a.slots);
    switch (a.law) {
        case LAW_NH:
            kElemMulti<T, NEN, NG, LAW_NH><<<grid, kThreads, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, gdN, a.gw,
                                                                      a.params, a.paramsPerElem, a.x, slots);
            break;
        case LAW_MR:
            kElemMulti<T, NEN, NG, LAW_MR><<<grid, kThreads, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, gdN, a.gw,
                                                                      a.params, a.paramsPerElem, a.x, slots);
            break;
        default:
            kElemMulti<T, NEN, NG, LAW_YEOH><<<grid, kThreads, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, gdN, a.gw,
                                                                        a.params, a.paramsPerElem, a.x, slots);
    }
}

template <typename T>
void launchTet4Law(const ElemArgs& a, hipStream_t s)
{
    const int grid = (a.nElems + kThreads - 1) / kThreads;
    const double* ginv = static_cast<const double*>(a.ginv);
    T* slots = static_cast<T*>(a.slots);
    switch (a.law) {
        case LAW_NH:
            kElemTet4<T, LAW_NH><<<grid, kThreads, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, 1.0, a.params,
                                                            a.paramsPerElem, a.x, slots);
            break;
        case LAW_MR:
            kElemTet4<T, LAW_MR><<<grid, kThreads, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, 1.0, a.params,
                                                            a.paramsPerElem, a.x, slots);
            break;
        default:
            kElemTet4<T, LAW_YEOH><<<grid, kThreads, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, 1.0, a.params,
                                                              a.paramsPerElem, a.x, slots);
    }
}

template <typename T>
void launchElem(const ElemArgs& a, hipStream_t s)
{
    if (a.nen == 4) return launchTet4Law<T>(a, s);
    if (a.nen == 8 && a.ng == 8) return launchMultiLaw<T, 8, 8>(a, s);
    if (a.nen == 8 && a.ng == 1) return launchMultiLaw<T, 8, 1>(a, s);
    if (a.nen == 10 && a.ng == 4) return launchMultiLaw<T, 10, 4>(a, s);
    if (a.nen == 10 && a.ng == 1) return launchMultiLaw<T, 10, 1>(a, s);
    if (a.nen == 20 && a.ng == 27) return launchMultiLaw<T, 20, 27>(a, s);
    if (a.nen == 20 && a.ng == 8) return launchMultiLaw<T, 20, 8>(a, s);
}

// ---- follower pressure: one lane per face (FP64) --------------------------
// f_a = p sum_q w_q N_a(xi_q) J F^-T A0 = p sum_q w_q N_a cof(F_q) A0
// (ExternalConstraint.cpp:829-836); directed: p_vec sum_q w_q N_a |cof F A0|
template <int NEN>
__global__ __launch_bounds__(kThreads) void kPressure(FaceArgs a)
{
    const int f = blockIdx.x * kThreads + threadIdx.x;
    if (f >= a.nFaces) return;
    const int F = a.nFaces;
    double du[NEN][3];
    const int id0 = a.faceDof[f];
#pragma unroll
    for (int n = 0; n < NEN; ++n) {
        const int id = a.faceDof[n * F + f];
#pragma unroll
        for (int c = 0; c < 3; ++c) du[n][c] = a.x[3 * id + c] - a.x[3 * id0 + c];
    }
    double out[NEN][3];
#pragma unroll
    for (int n = 0; n < NEN; ++n) out[n][0] = out[n][1] = out[n][2] = 0.0;
    const double A0[3] = {a.A0[3 * f], a.A0[3 * f + 1], a.A0[3 * f + 2]};
    for (int q = 0; q < a.fq; ++q) {
        const double* G = a.G + (static_cast<size_t>(f) * a.fq + q) * NEN * 3;
        double Fm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
        for (int n = 1; n < NEN; ++n)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) Fm[3 * i + j] += du[n][i] * G[3 * n + j];
        double cf[9];
        cf[0] = Fm[4] * Fm[8] - Fm[5] * Fm[7];
        cf[1] = Fm[5] * Fm[6] - Fm[3] * Fm[8];
        cf[2] = Fm[3] * Fm[7] - Fm[4] * Fm[6];
        cf[3] = Fm[2] * Fm[7] - Fm[1] * Fm[8];
        cf[4] = Fm[0] * Fm[8] - Fm[2] * Fm[6];
        cf[5] = Fm[1] * Fm[6] - Fm[0] * Fm[7];
        cf[6] = Fm[1] * Fm[5] - Fm[2] * Fm[4];
        cf[7] = Fm[2] * Fm[3] - Fm[0] * Fm[5];
        cf[8] = Fm[0] * Fm[4] - Fm[1] * Fm[3];
        double v[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) v[i] = cf[3 * i] * A0[0] + cf[3 * i + 1] * A0[1] + cf[3 * i + 2] * A0[2];
        if (a.directed) {
            const double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            v[0] = nv * a.pdir[0];
            v[1] = nv * a.pdir[1];
            v[2] = nv * a.pdir[2];
        }
        const double* WN = a.WN + (static_cast<size_t>(f) * a.fq + q) * NEN;
#pragma unroll
        for (int n = 0; n < NEN; ++n) {
            const double wn = WN[n];
            out[n][0] += wn * v[0];
            out[n][1] += wn * v[1];
            out[n][2] += wn * v[2];
        }
    }
    double* dst = a.slots + static_cast<size_t>(f) * NEN * 3;
#pragma unroll
    for (int n = 0; n < NEN; ++n)
#pragma unroll
        for (int c = 0; c < 3; ++c) dst[3 * n + c] = a.pressure * out[n][c];
}

// ---- gather assembly ------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void kGather(GatherArgs a)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= a.nFree) return;
    const T* es = static_cast<const T*>(a.eslots);
    double se[3] = {0, 0, 0}, sf[3] = {0, 0, 0};
    const int p1 = a.gptr[i + 1];
    for (int p = a.gptr[i]; p < p1; ++p) {
        const int s = a.gslot[p];
        if (s < a.nElemSlots) {
            const T* v = es + 3 * static_cast<size_t>(s);
            se[0] += static_cast<double>(v[0]);
            se[1] += static_cast<double>(v[1]);
            se[2] += static_cast<double>(v[2]);
        } else {
            const double* v = a.fslots + 3 * static_cast<size_t>(s - a.nElemSlots);
            sf[0] += v[0];
            sf[1] += v[1];
            sf[2] += v[2];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) a.r[3 * i + c] = se[c] - sf[c] - a.fext[3 * i + c];
}

// ---- reductions -------------------------------------------------------------
template <int K>
__device__ __forceinline__ void blockReduce(double (&v)[K], double* sh, bool isMax)
{
    // fixed-shape tree over 256 lanes
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k * kThreads + tid] = v[k];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (tid < w) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double o = sh[k * kThreads + tid + w];
                double& m = sh[k * kThreads + tid];
                m = (isMax && k > 0) ? fmax(m, o) : m + o;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = sh[k * kThreads];
    __syncthreads();
}

template <int K>
__device__ __forceinline__ bool lastBlockIn(double (&v)[K], double* work, unsigned* counter,
                                            double* sh, bool isMax)
{
    __shared__ int amLast;
    blockReduce<K>(v, sh, isMax);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) work[k * kReduceBlocks + blockIdx.x] = v[k];
        __threadfence();
        const unsigned t = atomicAdd(counter, 1u);
        amLast = (t == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!amLast) return false;
    __threadfence();
    // final pass over the block partials in block order
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    for (int b = threadIdx.x; b < static_cast<int>(gridDim.x); b += kThreads)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double o = work[k * kReduceBlocks + b];
            v[k] = (isMax && k > 0) ? fmax(v[k], o) : v[k] + o;
        }
    blockReduce<K>(v, sh, isMax);
    if (threadIdx.x == 0) *counter = 0u;
    return true;
}

int reduceGrid(int n) { return max(1, min(kReduceBlocks, (n + 4 * kThreads - 1) / (4 * kThreads))); }

__global__ __launch_bounds__(kThreads) void kDotMax(const double* __restrict__ a, const double* __restrict__ b,
                                                    const double* __restrict__ c, int n, double* work,
                                                    unsigned* counter, double* out)
{
    __shared__ double sh[3 * kThreads];
    double v[3] = {0, 0, 0};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const double ai = a[i];
        v[0] += ai * b[i];
        v[1] = fmax(v[1], fabs(ai));
        if (c) v[2] = fmax(v[2], fabs(c[i]));
    }
    if (lastBlockIn<3>(v, work, counter, sh, true) && threadIdx.x == 0) {
        out[0] = v[0];
        out[1] = v[1];
        out[2] = v[2];
    }
}

__global__ __launch_bounds__(kThreads) void kDot2(const double* __restrict__ a, const double* __restrict__ b, int n,
                                                  double* work, unsigned* counter, double* ab, double* bb)
{
    __shared__ double sh[2 * kThreads];
    double v[2] = {0, 0};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const double bi = b[i];
        v[0] += a[i] * bi;
        v[1] += bi * bi;
    }
    if (lastBlockIn<2>(v, work, counter, sh, false) && threadIdx.x == 0) {
        *ab = v[0];
        *bb = v[1];
    }
}

__global__ __launch_bounds__(kThreads) void kTwoLoop(double* __restrict__ q, const double* __restrict__ vprev,
                                                     const double* cprev, double sgn,
                                                     const double* __restrict__ w, const double* den,
                                                     const double* aj, double* cout, int mode, int n,
                                                     double* work, unsigned* counter)
{
    __shared__ double sh[kThreads];
    double v[1] = {0};
    const double c = vprev ? sgn * (*cprev) : 0.0;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        double qi = q[i];
        if (vprev) {
            qi += c * vprev[i];
            q[i] = qi;
        }
        v[0] += w[i] * qi;
    }
    if (lastBlockIn<1>(v, work, counter, sh, false) && threadIdx.x == 0) {
        const double r = v[0] / (*den);
        *cout = mode ? (*aj - r) : r;
    }
}

// ---- compact two-loop ------------------------------------------------------
constexpr int KC = kMaxCompact;

// wave64 butterfly then a fixed pass over the block's 4 waves: the sum is
// written to sh[k * 4 + wave]; caller syncs
__device__ __forceinline__ double waveSum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// per-thread sums v[0..cnt) -> block partials -> last block sums the
// partials in block order into red[0..cnt) (shared); returns true in the
// last block (all threads), after a __syncthreads
template <int CNT>
__device__ __forceinline__ bool multiReduce(double (&v)[CNT], int cnt, double* work, unsigned* counter,
                                            double* red)
{
    __shared__ double sh[CNT * 4];
    __shared__ int amLast;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < CNT; ++k) {
        if (k >= cnt) break;
        const double t = waveSum(v[k]);
        if (lane == 0) sh[k * 4 + wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < cnt) {
        const int k = threadIdx.x;
        work[k * kReduceBlocks + blockIdx.x] = ((sh[4 * k] + sh[4 * k + 1]) + sh[4 * k + 2]) + sh[4 * k + 3];
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned t = atomicAdd(counter, 1u);
        amLast = (t == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!amLast) return false;
    __threadfence();
    if (threadIdx.x < cnt) {
        const int k = threadIdx.x;
        double acc = 0.0;
        for (int b = 0; b < static_cast<int>(gridDim.x); ++b) acc += work[k * kReduceBlocks + b];
        red[k] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) *counter = 0u;
    return true;
}

__global__ __launch_bounds__(kThreads) void kGramUpdate(const double* __restrict__ S, const double* __restrict__ Y,
                                                        int n, int nc, int loc, double* G, double* ys, double* yy,
                                                        double* work, unsigned* counter)
{
    __shared__ double red[2 * KC + 2];
    double v[2 * KC + 1];
#pragma unroll
    for (int k = 0; k < 2 * KC + 1; ++k) v[k] = 0.0;
    const double* sl = S + static_cast<size_t>(loc) * n;
    const double* yl = Y + static_cast<size_t>(loc) * n;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const double si = sl[i], yi = yl[i];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            if (k >= nc) break;
            v[k] += si * Y[static_cast<size_t>(k) * n + i];       // s_loc . y_k
            v[KC + k] += S[static_cast<size_t>(k) * n + i] * yi;  // s_k . y_loc
        }
        v[2 * KC] += yi * yi;
    }
    // compact the live sums: [0,nc) s_loc.y_k, [nc,2nc) s_k.y_loc, 2nc y.y
    double c[2 * KC + 1];
#pragma unroll
    for (int k = 0; k < 2 * KC + 1; ++k) c[k] = 0.0;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        if (k >= nc) break;
        c[k] = v[k];
        c[nc + k] = v[KC + k];
    }
    c[2 * nc] = v[2 * KC];
    if (!multiReduce<2 * KC + 1>(c, 2 * nc + 1, work, counter, red)) return;
    if (threadIdx.x == 0) {
        for (int k = 0; k < nc; ++k) {
            G[loc * KC + k] = red[k];
            G[k * KC + loc] = red[nc + k];
        }
        ys[loc] = red[loc];
        *yy = red[2 * nc];
    }
}

__global__ __launch_bounds__(kThreads) void kTwoLoopDots(const double* __restrict__ q, const double* __restrict__ R,
                                                         int n, int nc, int first, int newest, int mcap,
                                                         const double* G, const double* ys, const double* alphaIn,
                                                         double* outCoef, double* work, unsigned* counter)
{
    // first: dots of R_k (= S) with q (= r), alpha recurrence newest->oldest
    // else: dots of R_k (= Y) with q (= H0 output), beta recurrence
    __shared__ double red[KC];
    double v[KC];
#pragma unroll
    for (int k = 0; k < KC; ++k) v[k] = 0.0;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const double qi = q[i];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            if (k >= nc) break;
            v[k] += R[static_cast<size_t>(k) * n + i] * qi;
        }
    }
    if (!multiReduce<KC>(v, nc, work, counter, red)) return;
    if (threadIdx.x == 0) {
        if (first) {
            // q_0 = -r: alpha_i = (s_i.q_0 - sum_{k newer} alpha_k s_i.y_k) / ys_i
            int j = newest;
            for (int t = 0; t < nc; ++t) {
                double acc = -red[j];
                int kk = newest;
                for (int u = 0; u < t; ++u) {
                    acc -= outCoef[kk] * G[j * KC + kk];
                    kk = (kk + mcap - 1) % mcap;
                }
                outCoef[j] = acc / ys[j];
                j = (j + mcap - 1) % mcap;
            }
        } else {
            // beta_i = (y_i.h + sum_{k older} coef_k y_i.s_k) / ys_i ; coef = alpha - beta
            int oldest = (newest + mcap - nc + 1) % mcap;
            int j = oldest;
            for (int t = 0; t < nc; ++t) {
                double acc = red[j];
                int kk = oldest;
                for (int u = 0; u < t; ++u) {
                    acc += outCoef[kk] * G[kk * KC + j];
                    kk = (kk + 1) % mcap;
                }
                outCoef[j] = alphaIn[j] - acc / ys[j];
                j = (j + 1) % mcap;
            }
        }
    }
}

// out = base * sgn_b + sum_k coef_k R_k * sgn_c
__global__ __launch_bounds__(kThreads) void kMultiAxpy(const double* __restrict__ base, double sgnB,
                                                       const double* __restrict__ R, const double* coef,
                                                       double sgnC, int n, int nc, double* __restrict__ out)
{
    __shared__ double c[KC];
    if (threadIdx.x < nc) c[threadIdx.x] = sgnC * coef[threadIdx.x];
    __syncthreads();
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    double v = sgnB * base[i];
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        if (k >= nc) break;
        v += c[k] * R[static_cast<size_t>(k) * n + i];
    }
    out[i] = v;
}

__global__ void kStep(double* x, const double* xp, const double* d, double alpha, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) x[i] = xp[i] + alpha * d[i];
}

__global__ void kDiff2(double* s, double* y, const double* x, const double* xp, const double* r,
                       const double* rp, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        s[i] = x[i] - xp[i];
        y[i] = r[i] - rp[i];
    }
}

__global__ void kScale(double* y, const double* x, double a, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

__global__ void kAxpyDev(double* q, const double* v, const double* c, int neg, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        const double cc = neg ? -(*c) : *c;
        q[i] += cc * v[i];
    }
}

__global__ void kDivDev(double* q, const double* th, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) q[i] /= *th;
}

__global__ void kTheta(const double* ys, const double* yy, double* th)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *th = *yy / *ys;
}

// ---- supernodal triangular solves -----------------------------------------
constexpr int kMaxW = 256;   // widest supernode (host caps it)

template <typename T>
__global__ __launch_bounds__(kThreads) void kTrsvFwd(TrsvArgs a, const double* __restrict__ b,
                                                     const int2* __restrict__ items)
{
    __shared__ T sb[kMaxW * 3];
    __shared__ T red[4][64][3];
    const int2 it = items[blockIdx.x];
    const int s = it.x, r0 = it.y;
    const int w = a.snW[s], m = a.snM[s], j0 = a.snFirst[s], fr0 = a.snRow[s];
    T* u = static_cast<T*>(a.t);
    // front entries of the supernode's own columns: b plus the children's
    // update rows (short lists: at most one entry per child)
    for (int k = threadIdx.x; k < w; k += kThreads) {
        const int old = a.perm[j0 + k];
        double v0 = b[3 * old], v1 = b[3 * old + 1], v2 = b[3 * old + 2];
        for (int p = a.fptr[fr0 + k]; p < a.fptr[fr0 + k + 1]; ++p) {
            const T* uv = u + 3 * static_cast<size_t>(a.fidx[p]);
            v0 += static_cast<double>(uv[0]);
            v1 += static_cast<double>(uv[1]);
            v2 += static_cast<double>(uv[2]);
        }
        sb[3 * k] = static_cast<T>(v0);
        sb[3 * k + 1] = static_cast<T>(v1);
        sb[3 * k + 2] = static_cast<T>(v2);
    }
    __syncthreads();
    // 64 rows per item; the four waves split the k range, then reduce in order
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = r0 + lane;
    const int kc = (w + 3) >> 2;
    const int kb = wave * kc;
    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    if (r < m) {
        int ke = min(w, kb + kc);
        if (r < w) ke = min(ke, r + 1);  // L11^-1 is lower triangular
        const T* Bc = static_cast<const T*>(a.B) + a.snVal[s] + r;
        int k = kb;
        for (; k + 4 <= ke; k += 4) {
            const T b0 = Bc[static_cast<size_t>(k) * m], b1 = Bc[static_cast<size_t>(k + 1) * m];
            const T b2 = Bc[static_cast<size_t>(k + 2) * m], b3 = Bc[static_cast<size_t>(k + 3) * m];
            acc0 += b0 * sb[3 * k] + b1 * sb[3 * k + 3] + b2 * sb[3 * k + 6] + b3 * sb[3 * k + 9];
            acc1 += b0 * sb[3 * k + 1] + b1 * sb[3 * k + 4] + b2 * sb[3 * k + 7] + b3 * sb[3 * k + 10];
            acc2 += b0 * sb[3 * k + 2] + b1 * sb[3 * k + 5] + b2 * sb[3 * k + 8] + b3 * sb[3 * k + 11];
        }
        for (; k < ke; ++k) {
            const T bk = Bc[static_cast<size_t>(k) * m];
            acc0 += bk * sb[3 * k];
            acc1 += bk * sb[3 * k + 1];
            acc2 += bk * sb[3 * k + 2];
        }
    }
    red[wave][lane][0] = acc0;
    red[wave][lane][1] = acc1;
    red[wave][lane][2] = acc2;
    __syncthreads();
    if (wave != 0 || r >= m) return;
    acc0 = ((red[0][lane][0] + red[1][lane][0]) + red[2][lane][0]) + red[3][lane][0];
    acc1 = ((red[0][lane][1] + red[1][lane][1]) + red[2][lane][1]) + red[3][lane][1];
    acc2 = ((red[0][lane][2] + red[1][lane][2]) + red[2][lane][2]) + red[3][lane][2];
    if (r < w) {
        T* out = static_cast<T*>(a.y) + 3 * static_cast<size_t>(j0 + r);
        out[0] = acc0;
        out[1] = acc1;
        out[2] = acc2;
        return;
    }
    // update row: children's contributions to this row minus (L21 L11^-1) f
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int p = a.fptr[fr0 + r]; p < a.fptr[fr0 + r + 1]; ++p) {
        const T* uv = u + 3 * static_cast<size_t>(a.fidx[p]);
        f0 += static_cast<double>(uv[0]);
        f1 += static_cast<double>(uv[1]);
        f2 += static_cast<double>(uv[2]);
    }
    T* out = u + 3 * static_cast<size_t>(a.snT[s] + r - w);
    out[0] = static_cast<T>(f0) - acc0;
    out[1] = static_cast<T>(f1) - acc1;
    out[2] = static_cast<T>(f2) - acc2;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void kTrsvBwd(TrsvArgs a, const int2* __restrict__ items)
{
    // four columns per item, one per wave; lanes stride the rows
    const int2 it = items[blockIdx.x];
    const int s = it.x;
    const int w = a.snW[s], m = a.snM[s], j0 = a.snFirst[s];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = it.y + wave;
    if (k >= w) return;
    const int* rows = a.rows + a.snRow[s];
    const T* y = static_cast<const T*>(a.y);
    T* z = static_cast<T*>(a.z);
    const T* col = static_cast<const T*>(a.B) + a.snVal[s] + static_cast<size_t>(k) * m;
    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    for (int r = k + lane; r < m; r += 64) {
        const T bk = col[r];
        const T* src = r < w ? y + 3 * static_cast<size_t>(j0 + r) : z + 3 * static_cast<size_t>(rows[r]);
        const T sg = r < w ? T(1) : T(-1);
        acc0 += bk * (sg * src[0]);
        acc1 += bk * (sg * src[1]);
        acc2 += bk * (sg * src[2]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        acc0 += __shfl_xor(acc0, off, 64);
        acc1 += __shfl_xor(acc1, off, 64);
        acc2 += __shfl_xor(acc2, off, 64);
    }
    if (lane == 0) {
        T* zo = z + 3 * static_cast<size_t>(j0 + k);
        zo[0] = acc0;
        zo[1] = acc1;
        zo[2] = acc2;
    }
}

template <typename T>
__global__ void kUnpermute(const int* __restrict__ perm, const T* __restrict__ z, double* __restrict__ out, int n)
{
    const int j = blockIdx.x * kThreads + threadIdx.x;
    if (j < n) {
        const int o = perm[j];
        out[3 * o] = static_cast<double>(z[3 * j]);
        out[3 * o + 1] = static_cast<double>(z[3 * j + 1]);
        out[3 * o + 2] = static_cast<double>(z[3 * j + 2]);
    }
}

inline int blocks(int n) { return (n + kThreads - 1) / kThreads; }

}  // namespace

void launchElementForce(const ElemArgs& a, hipStream_t s)
{
    if (a.mixed)
        launchElem<float>(a, s);
    else
        launchElem<double>(a, s);
}

void launchPressure(const FaceArgs& a, hipStream_t s)
{
    if (a.nFaces <= 0) return;
    const int g = blocks(a.nFaces);
    switch (a.nen) {
        case 4: kPressure<4><<<g, kThreads, 0, s>>>(a); break;
        case 8: kPressure<8><<<g, kThreads, 0, s>>>(a); break;
        case 10: kPressure<10><<<g, kThreads, 0, s>>>(a); break;
        case 20: kPressure<20><<<g, kThreads, 0, s>>>(a); break;
    }
}

void launchGather(const GatherArgs& a, hipStream_t s)
{
    if (a.nFree <= 0) return;
    if (a.mixed)
        kGather<float><<<blocks(a.nFree), kThreads, 0, s>>>(a);
    else
        kGather<double><<<blocks(a.nFree), kThreads, 0, s>>>(a);
}

void launchDotMax(const double* a, const double* b, const double* c, int n, double* work, unsigned* counter,
                  double* out, hipStream_t s)
{
    kDotMax<<<reduceGrid(n), kThreads, 0, s>>>(a, b, c, n, work, counter, out);
}

void launchDot2(const double* a, const double* b, int n, double* work, unsigned* counter, double* ab, double* bb,
                hipStream_t s)
{
    kDot2<<<reduceGrid(n), kThreads, 0, s>>>(a, b, n, work, counter, ab, bb);
}

void launchStep(double* x, const double* xp, const double* d, double alpha, int n, hipStream_t s)
{
    if (n > 0) kStep<<<blocks(n), kThreads, 0, s>>>(x, xp, d, alpha, n);
}

void launchDiff2(double* sv, double* yv, const double* x, const double* xp, const double* r, const double* rp,
                 int n, hipStream_t s)
{
    if (n > 0) kDiff2<<<blocks(n), kThreads, 0, s>>>(sv, yv, x, xp, r, rp, n);
}

void launchScale(double* y, const double* x, double a, int n, hipStream_t s)
{
    if (n > 0) kScale<<<blocks(n), kThreads, 0, s>>>(y, x, a, n);
}

void launchTwoLoopStep(double* q, const double* vprev, const double* cprev, const double* w, const double* den,
                       const double* aj, double* cout, int mode, int n, double* work, unsigned* counter,
                       hipStream_t s)
{
    kTwoLoop<<<reduceGrid(n), kThreads, 0, s>>>(q, vprev, cprev, mode ? 1.0 : -1.0, w, den, aj, cout, mode, n,
                                                  work, counter);
}

void launchGramUpdate(const double* S, const double* Y, int n, int nc, int loc, double* G, double* ys, double* yy,
                      double* work, unsigned* counter, hipStream_t s)
{
    kGramUpdate<<<reduceGrid(n), kThreads, 0, s>>>(S, Y, n, nc, loc, G, ys, yy, work, counter);
}

void launchTwoLoopFirst(const double* r, const double* S, const double* Y, int n, int nc, int newest, int mcap,
                        const double* G, const double* ys, double* alpha, double* out, double* work,
                        unsigned* counter, hipStream_t s)
{
    kTwoLoopDots<<<reduceGrid(n), kThreads, 0, s>>>(r, S, n, nc, 1, newest, mcap, G, ys, nullptr, alpha, work,
                                                     counter);
    if (n > 0) kMultiAxpy<<<blocks(n), kThreads, 0, s>>>(r, -1.0, Y, alpha, -1.0, n, nc, out);
}

void launchTwoLoopSecond(const double* h, const double* S, const double* Y, int n, int nc, int newest, int mcap,
                         const double* G, const double* ys, const double* alpha, double* coef, double* out,
                         double* work, unsigned* counter, hipStream_t s)
{
    kTwoLoopDots<<<reduceGrid(n), kThreads, 0, s>>>(h, Y, n, nc, 0, newest, mcap, G, ys, alpha, coef, work,
                                                     counter);
    if (n > 0) kMultiAxpy<<<blocks(n), kThreads, 0, s>>>(h, 1.0, S, coef, 1.0, n, nc, out);
}

void launchAxpyDev(double* q, const double* v, const double* c, int neg, int n, hipStream_t s)
{
    if (n > 0) kAxpyDev<<<blocks(n), kThreads, 0, s>>>(q, v, c, neg, n);
}

void launchDivDev(double* q, const double* theta, int n, hipStream_t s)
{
    if (n > 0) kDivDev<<<blocks(n), kThreads, 0, s>>>(q, theta, n);
}

void launchTheta(const double* ys, const double* yy, double* theta, hipStream_t s)
{
    kTheta<<<1, 64, 0, s>>>(ys, yy, theta);
}

void launchTrsvForward(const TrsvArgs& a, const double* b, const int2* items, int nitems, hipStream_t s)
{
    if (nitems <= 0) return;
    if (a.mixed)
        kTrsvFwd<float><<<nitems, kThreads, 0, s>>>(a, b, items);
    else
        kTrsvFwd<double><<<nitems, kThreads, 0, s>>>(a, b, items);
}

void launchTrsvBackward(const TrsvArgs& a, const int2* items, int nitems, hipStream_t s)
{
    if (nitems <= 0) return;
    if (a.mixed)
        kTrsvBwd<float><<<nitems, kThreads, 0, s>>>(a, items);
    else
        kTrsvBwd<double><<<nitems, kThreads, 0, s>>>(a, items);
}

void launchUnpermute(const TrsvArgs& a, double* out, hipStream_t s)
{
    if (a.n <= 0) return;
    if (a.mixed)
        kUnpermute<float><<<blocks(a.n), kThreads, 0, s>>>(a.perm, static_cast<const float*>(a.z), out, a.n);
    else
        kUnpermute<double><<<blocks(a.n), kThreads, 0, s>>>(a.perm, static_cast<const double*>(a.z), out, a.n);
}

}  // namespace dev
}  // namespace efem
