// Device building blocks shared by the grid kernels (kernels.hip) and the
// persistent batched solver (batch.hip): element stress moments and tiles,
// follower-pressure faces, the deterministic gather, and the supernodal
// triangular-solve work items.  All are __device__ functions for blocks of
// kThreads = 256 lanes (four wave64s).
#pragma once
#include <hip/hip_runtime.h>

#include "../material_math.hpp"
#include "../limits.hpp"
#include "kernels.hpp"

namespace efem {
namespace dev {
namespace {

constexpr int kThreads = 256;

// Workgroups are dispatched to the 8 XCDs round-robin by workgroup id, and
// each XCD has its own L2.  For gather-heavy kernels whose neighbouring
// work items share data (nodes sharing elements), give each XCD one
// contiguous range of logical blocks instead of every 8th block.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcdBlockOf(int b, int g)
{
    const int x = b % kXcds, k = b / kXcds, per = g / kXcds, rem = g % kXcds;
    return x * per + min(x, rem) + k;
}
__device__ __forceinline__ int xcdBlock() { return xcdBlockOf(blockIdx.x, gridDim.x); }

// Cross-lane moves within rows of 16 lanes (DPP; every lane of the wave
// must be active): CTRL 0xB1 quad_perm [1,0,3,2], 0x4E quad_perm
// [2,3,0,1], 0x141 row_half_mirror, 0x140 row_mirror.
template <int CTRL>
__device__ __forceinline__ float dppMove(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dppMove(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float laneValue(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double laneValue(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// op over each row of 16 lanes, in every lane of the row (pairs, quads,
// eights, sixteens; op commutative, so every lane of a row holds the same
// value).  Every lane must be active.
template <typename T, typename Op>
__device__ __forceinline__ T rowReduce16(T v, Op op)
{
    v = op(v, dppMove<0xB1>(v));
    v = op(v, dppMove<0x4E>(v));
    v = op(v, dppMove<0x141>(v));
    v = op(v, dppMove<0x140>(v));
    return v;
}
// op over the wave, in every lane: the rows, then the four row values in
// order.  Every lane must be active.  (A shuffle butterfly costs six
// dependent LDS-crossbar round trips per value.)
template <typename T, typename Op>
__device__ __forceinline__ T waveReduce(T v, Op op)
{
    v = rowReduce16(v, op);
    return op(op(laneValue(v, 0), laneValue(v, 16)), op(laneValue(v, 32), laneValue(v, 48)));
}
struct OpAdd {
    template <typename T>
    __device__ T operator()(T a, T b) const { return a + b; }
};
template <typename T>
__device__ __forceinline__ T waveSumT(T v) { return waveReduce(v, OpAdd{}); }
template <typename T>
__device__ __forceinline__ T rowSum16(T v) { return rowReduce16(v, OpAdd{}); }

// wave64 sum (fixed order), every lane active
__device__ __forceinline__ double waveSum(double v) { return waveSumT(v); }

// max that propagates NaN (fmax drops it): a norm over a vector holding a
// NaN is NaN, so the non-finite checks of the solve see it
__device__ __forceinline__ double maxNan(double a, double b)
{
    return (a != a) ? a : ((b != b) ? b : fmax(a, b));
}

struct OpMaxNan {
    __device__ double operator()(double a, double b) const { return maxNan(a, b); }
};
__device__ __forceinline__ double waveMax(double v) { return waveReduce(v, OpMaxNan{}); }

// H = Gu J0^-1 is formed in FP64 in both modes: J - 1 (and with it the
// near-incompressible volumetric stress (2/D1)(J-1) cof F, 2/D1 = 2e6 here)
// inherits the rounding of H, so H must carry more than FP32's 24 bits.
// The deviatoric stress and the force moments then run in T.
template <typename GT>
__device__ __forceinline__ void gradToH(const GT* Gu, const double* Jd, double* Hd)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Hd[3 * i + j] = static_cast<double>(Gu[3 * i]) * Jd[j] + static_cast<double>(Gu[3 * i + 1]) * Jd[3 + j] +
                            static_cast<double>(Gu[3 * i + 2]) * Jd[6 + j];
}

// MODE_FORCE: Q = w V P(H) J0^-T.  MODE_TANGENT: Q = w V dP J0^-T with
// dP = (dP/dF)[dH] the exact material tangent applied to the direction's
// gradient dH (forward-mode dual numbers through the same stress code).
enum { MODE_FORCE = 0, MODE_TANGENT = 1 };

template <typename T, int MODE>
__device__ __forceinline__ void stressMoment(int law, const double* p, const double* Gu, const double* Gv,
                                             const double* Jd, double wv, T* Q)
{
    double Hd[9];
    gradToH(Gu, Jd, Hd);
    T P[9];
    if (MODE == MODE_TANGENT) {
        double dHd[9];
        gradToH(Gv, Jd, dHd);
        piolaTangent<T>(law, p, Hd, dHd, P);
    } else {
        piolaFromHd<T>(law, p, Hd, P);
    }
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
template <int MODE>
__device__ __forceinline__ void tet4Grad(const double* __restrict__ x, const int* id, double* Gu)
{
    // dN = (e1, e2, e3, -1 -1 -1): Gu[:,k] = u_k - u_3
    double u3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) u3[c] = x[3 * id[3] + c];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i) Gu[3 * i + k] = x[3 * id[k] + i] - u3[i];
}

// J0^-1 of a C3D4 element from its nodes' reference coordinates (Xd: per
// free DOF, the node's coordinate), by the host's operations in the host's
// order without contraction -- J0 = sum_{a>=1} (X_a - X_0) (x) dN_a with the
// local gradients dN_0..3 = e_1, e_2, e_3, (-1, -1, -1), then the cofactor
// inverse -- so it equals the stored J0^-1 bit for bit
__device__ __forceinline__ void tet4Ginv(const double* __restrict__ Xd, const int* id, double* B)
{
#pragma clang fp contract(off)
    double X0[3], d[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) X0[i] = Xd[3 * id[0] + i];
#pragma unroll
    for (int a = 1; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) d[a - 1][i] = Xd[3 * id[a] + i] - X0[i];
    double A[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        A[3 * i] = -d[2][i];
        A[3 * i + 1] = d[0][i] - d[2][i];
        A[3 * i + 2] = d[1][i] - d[2][i];
    }
    const double c00 = A[4] * A[8] - A[5] * A[7];
    const double c01 = A[5] * A[6] - A[3] * A[8];
    const double c02 = A[3] * A[7] - A[4] * A[6];
    const double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
    const double id0 = 1.0 / det;
    B[0] = c00 * id0;
    B[1] = (A[2] * A[7] - A[1] * A[8]) * id0;
    B[2] = (A[1] * A[5] - A[2] * A[4]) * id0;
    B[3] = c01 * id0;
    B[4] = (A[0] * A[8] - A[2] * A[6]) * id0;
    B[5] = (A[2] * A[3] - A[0] * A[5]) * id0;
    B[6] = c02 * id0;
    B[7] = (A[1] * A[6] - A[0] * A[7]) * id0;
    B[8] = (A[0] * A[4] - A[1] * A[3]) * id0;
}

// Xd (may be null): J0^-1 recomputed from the nodes' coordinates when all
// four nodes are free (nF free nodes), else read from ginv (every fixed
// node shares one DOF slot, so its coordinates are not in Xd)
// The four nodal forces (or tangent products) of one C3D4 element with
// free ids id[]: f[3 a + c] for node a, component c.  eg indexes ginv
// (SoA, stride E) and the per-element parameters; volw = V_e w.
template <typename T, int LAW, int MODE>
__device__ __forceinline__ void tet4Forces(const int (&id)[4], int eg, int E, const double* __restrict__ ginv,
                                           double volw, const double* __restrict__ params, int ppe,
                                           const double* __restrict__ x, const double* __restrict__ v,
                                           const double* __restrict__ Xd, int nF, T (&f)[12])
{
    double Gu[9], Gv[9];
    tet4Grad<MODE>(x, id, Gu);
    if (MODE == MODE_TANGENT) tet4Grad<MODE>(v, id, Gv);
    double Jd[9];
    if (Xd && id[0] < nF && id[1] < nF && id[2] < nF && id[3] < nF) {
        tet4Ginv(Xd, id, Jd);
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) Jd[k] = ginv[k * E + eg];
    }
    T Q[9];
    stressMoment<T, MODE>(LAW, params + (ppe ? 4 * eg : 0), Gu, Gv, Jd, volw, Q);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        f[i] = Q[3 * i];
        f[3 + i] = Q[3 * i + 1];
        f[6 + i] = Q[3 * i + 2];
        f[9 + i] = -(Q[3 * i] + Q[3 * i + 1] + Q[3 * i + 2]);
    }
}

template <typename T, int LAW, int MODE>
__device__ __forceinline__ void tet4Lane(int e, int E, const int* __restrict__ connDof,
                                         const double* __restrict__ ginv, const double* __restrict__ vol, double w0,
                                         const double* __restrict__ params, int ppe, const double* __restrict__ x,
                                         const double* __restrict__ v, T* __restrict__ slots,
                                         const double* __restrict__ Xd = nullptr, int nF = 0)
{
    int id[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) id[a] = connDof[a * E + e];
    T f[12];
    tet4Forces<T, LAW, MODE>(id, e, E, ginv, vol[e] * w0, params, ppe, x, v, Xd, nF, f);
    T* out = slots + 12 * static_cast<size_t>(e);
#pragma unroll
    for (int k = 0; k < 12; ++k) out[k] = f[k];
}


// ---- general elements: LDS-staged (element, Gauss point) lanes ------------
template <int NG, int BT = kThreads>
struct Tile {
    static constexpr int epb = NG >= 4 ? BT / NG : BT / 4;  // elements per block of BT lanes
};

// LDS of one element tile (EPB elements x NG Gauss points over BT lanes)
template <typename T, int NEN, int NG, int MODE, int BT = kThreads>
struct ElemTileLds {
    static constexpr int EPB = Tile<NG, BT>::epb;
    static constexpr int NV = MODE == MODE_TANGENT ? 2 : 1;
    double sdN[NG * NEN * 3];
    double su[NV][EPB * NEN * 3];
    T sQ[EPB * NG * 9];
};

// One tile of elements [e0, min(e0 + EPB, eEnd)), called by all BT lanes of
// a block (contains barriers).  The local-gradient table sdN must be written
// before the first call (the tile's first barrier publishes it).
template <typename T, int NEN, int NG, int LAW, int MODE, int BT = kThreads>
__device__ __forceinline__ void elemTile(int E, int e0, int eEnd, const int* __restrict__ connDof,
                                         const double* __restrict__ ginv, const double* __restrict__ vol,
                                         const double* __restrict__ gw, const double* __restrict__ params, int ppe,
                                         const double* __restrict__ x, const double* __restrict__ v,
                                         T* __restrict__ slots, ElemTileLds<T, NEN, NG, MODE, BT>& L)
{
    constexpr int EPB = Tile<NG, BT>::epb;
    constexpr int NV = ElemTileLds<T, NEN, NG, MODE, BT>::NV;
    const int tid = threadIdx.x;
    // relative displacements u_a - u_0 (and the direction's v_a - v_0)
    for (int k = tid; k < EPB * NEN; k += BT) {
        const int el = k / NEN, a = k - el * NEN, e = e0 + el;
        if (e < eEnd) {
            const int id = connDof[a * E + e], id0 = connDof[e];
#pragma unroll
            for (int c = 0; c < 3; ++c) L.su[0][3 * k + c] = x[3 * id + c] - x[3 * id0 + c];
            if (MODE == MODE_TANGENT)
#pragma unroll
                for (int c = 0; c < 3; ++c) L.su[NV - 1][3 * k + c] = v[3 * id + c] - v[3 * id0 + c];
        }
    }
    __syncthreads();
    {
        const int el = tid / NG, q = tid - el * NG, e = e0 + el;
        if (el < EPB && e < eEnd) {
            double G[NV][9];
            const double* dq = L.sdN + q * NEN * 3;
#pragma unroll
            for (int h = 0; h < NV; ++h) {
#pragma unroll
                for (int k = 0; k < 9; ++k) G[h][k] = 0.0;
                const double* ue = L.su[h] + el * NEN * 3;
                for (int a = 1; a < NEN; ++a) {
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int k = 0; k < 3; ++k) G[h][3 * i + k] += ue[3 * a + i] * dq[3 * a + k];
                }
            }
            double Jd[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) Jd[k] = ginv[(static_cast<size_t>(q) * 9 + k) * E + e];
            stressMoment<T, MODE>(LAW, params + (ppe ? 4 * e : 0), G[0], G[NV - 1], Jd, vol[e] * gw[q],
                                  L.sQ + (el * NG + q) * 9);
        }
    }
    __syncthreads();
    for (int k = tid; k < EPB * NEN; k += BT) {
        const int el = k / NEN, a = k - el * NEN, e = e0 + el;
        if (e >= eEnd) continue;
        T f[3] = {T(0), T(0), T(0)};
        for (int q = 0; q < NG; ++q) {
            const T* Q = L.sQ + (el * NG + q) * 9;
            const double* d = L.sdN + (q * NEN + a) * 3;
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


// ---- follower pressure: one lane per face (FP64) --------------------------
// f_a = p sum_q w_q N_a(xi_q) J F^-T A0 = p sum_q w_q N_a cof(F_q) A0
// (ExternalConstraint.cpp:829-836); directed: p_vec sum_q w_q N_a |cof F A0|
// MODE 0: the force.  MODE 1: its tangent applied to a direction v (the
// pressure load stiffness; cof F is quadratic in F, so dual numbers give the
// exact derivative).  MODE 2: the transpose of that tangent applied to v,
// from g(F) = mu . cof(F) A0 with mu = sum_a w N_a v_a:
//   dg/dF[:,K] = (F (A0 x e_K)) x mu.
enum { PMODE_FORCE = 0, PMODE_TANGENT = 1, PMODE_TRANSPOSE = 2 };

template <typename S>
__device__ __forceinline__ void cofactor(const S* Fm, S* cf)
{
    cf[0] = Fm[4] * Fm[8] - Fm[5] * Fm[7];
    cf[1] = Fm[5] * Fm[6] - Fm[3] * Fm[8];
    cf[2] = Fm[3] * Fm[7] - Fm[4] * Fm[6];
    cf[3] = Fm[2] * Fm[7] - Fm[1] * Fm[8];
    cf[4] = Fm[0] * Fm[8] - Fm[2] * Fm[6];
    cf[5] = Fm[1] * Fm[6] - Fm[0] * Fm[7];
    cf[6] = Fm[1] * Fm[5] - Fm[2] * Fm[4];
    cf[7] = Fm[2] * Fm[3] - Fm[0] * Fm[5];
    cf[8] = Fm[0] * Fm[4] - Fm[1] * Fm[3];
}

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

template <int NEN, int MODE>
__device__ __forceinline__ void pressureFace(const FaceArgs& a, int f)
{
    const int F = a.nFaces;
    double du[NEN][3];
    double dv[NEN][3];  // MODE 1: direction relative to node 0; MODE 2: v itself
    const int id0 = a.faceDof[f];
#pragma unroll
    for (int n = 0; n < NEN; ++n) {
        const int id = a.faceDof[n * F + f];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            du[n][c] = a.x[3 * id + c] - a.x[3 * id0 + c];
            if (MODE == PMODE_TANGENT) dv[n][c] = a.v[3 * id + c] - a.v[3 * id0 + c];
            if (MODE == PMODE_TRANSPOSE) dv[n][c] = a.v[3 * id + c];
        }
    }
    double out[NEN][3];
#pragma unroll
    for (int n = 0; n < NEN; ++n) out[n][0] = out[n][1] = out[n][2] = 0.0;
    const double A0[3] = {a.A0[3 * f], a.A0[3 * f + 1], a.A0[3 * f + 2]};
    for (int q = 0; q < a.fq; ++q) {
        const double* G = a.G + (static_cast<size_t>(f) * a.fq + q) * NEN * 3;
        const double* WN = a.WN + (static_cast<size_t>(f) * a.fq + q) * NEN;
        double Fm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
        for (int n = 1; n < NEN; ++n)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) Fm[3 * i + j] += du[n][i] * G[3 * n + j];
        if (MODE == PMODE_FORCE) {
            double cf[9];
            cofactor(Fm, cf);
            double v[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) v[i] = cf[3 * i] * A0[0] + cf[3 * i + 1] * A0[1] + cf[3 * i + 2] * A0[2];
            if (a.directed) {
                const double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                v[0] = nv * a.pdir[0];
                v[1] = nv * a.pdir[1];
                v[2] = nv * a.pdir[2];
            }
#pragma unroll
            for (int n = 0; n < NEN; ++n) {
                const double wn = WN[n];
                out[n][0] += wn * v[0];
                out[n][1] += wn * v[1];
                out[n][2] += wn * v[2];
            }
        } else if (MODE == PMODE_TANGENT) {
            Dual<double> Fd[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) Fd[k] = Dual<double>(Fm[k], 0.0);
#pragma unroll
            for (int n = 1; n < NEN; ++n)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) Fd[3 * i + j].d += dv[n][i] * G[3 * n + j];
            Dual<double> cf[9];
            cofactor(Fd, cf);
            Dual<double> v[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) v[i] = cf[3 * i] * A0[0] + cf[3 * i + 1] * A0[1] + cf[3 * i + 2] * A0[2];
            double dvq[3] = {v[0].d, v[1].d, v[2].d};
            if (a.directed) {
                const Dual<double> nv = sqrt_(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                dvq[0] = nv.d * a.pdir[0];
                dvq[1] = nv.d * a.pdir[1];
                dvq[2] = nv.d * a.pdir[2];
            }
#pragma unroll
            for (int n = 0; n < NEN; ++n) {
                const double wn = WN[n];
                out[n][0] += wn * dvq[0];
                out[n][1] += wn * dvq[1];
                out[n][2] += wn * dvq[2];
            }
        } else {
            double mu[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int n = 0; n < NEN; ++n) {
                mu[0] += WN[n] * dv[n][0];
                mu[1] += WN[n] * dv[n][1];
                mu[2] += WN[n] * dv[n][2];
            }
            if (a.directed) {
                // g = (mu . pdir) |cof F A0|: dg/dF as above with mu' = (mu . pdir) v / |v|
                double cf[9];
                cofactor(Fm, cf);
                double v[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) v[i] = cf[3 * i] * A0[0] + cf[3 * i + 1] * A0[1] + cf[3 * i + 2] * A0[2];
                const double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                const double sc = (mu[0] * a.pdir[0] + mu[1] * a.pdir[1] + mu[2] * a.pdir[2]) / nv;
                mu[0] = sc * v[0];
                mu[1] = sc * v[1];
                mu[2] = sc * v[2];
            }
            double dg[3][3];  // dg[K][b] = dg/dF[b][K]
#pragma unroll
            for (int K = 0; K < 3; ++K) {
                double ak[3] = {0.0, 0.0, 0.0}, w[3];
                // A0 x e_K
                ak[(K + 1) % 3] = A0[(K + 2) % 3];
                ak[(K + 2) % 3] = -A0[(K + 1) % 3];
#pragma unroll
                for (int i = 0; i < 3; ++i) w[i] = Fm[3 * i] * ak[0] + Fm[3 * i + 1] * ak[1] + Fm[3 * i + 2] * ak[2];
                cross3(w, mu, dg[K]);
            }
#pragma unroll
            for (int n = 1; n < NEN; ++n)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const double t = dg[0][b] * G[3 * n] + dg[1][b] * G[3 * n + 1] + dg[2][b] * G[3 * n + 2];
                    out[n][b] += t;
                    out[0][b] -= t;
                }
        }
    }
    double* dst = a.slots + static_cast<size_t>(f) * NEN * 3;
    const double pr = a.fpress ? a.fpress[f] : a.pressure;
#pragma unroll
    for (int n = 0; n < NEN; ++n)
#pragma unroll
        for (int c = 0; c < 3; ++c) dst[3 * n + c] = pr * out[n][c];
}


// ---- gather assembly ------------------------------------------------------
template <typename T>
__device__ __forceinline__ void gatherNode(const GatherArgs& a, int i)
{
    const T* es = static_cast<const T*>(a.eslots);
    double se[3] = {0, 0, 0}, sf[3] = {0, 0, 0};
    const int p1 = a.gptr[i + 1];
    for (int p = a.gptr[i]; p < p1; ++p) {
        const int s = a.gslot[p];
        if (s < a.nElemSlots) {
            if (a.ce == 0.0) continue;
            const T* v = es + 3 * static_cast<size_t>(s);
            se[0] += static_cast<double>(v[0]);
            se[1] += static_cast<double>(v[1]);
            se[2] += static_cast<double>(v[2]);
        } else {
            if (a.cf == 0.0) continue;
            const double* v = a.fslots + 3 * static_cast<size_t>(s - a.nElemSlots);
            sf[0] += v[0];
            sf[1] += v[1];
            sf[2] += v[2];
        }
    }
    // residual: ce = 1, cf = -1 (r = f_int - f_pressure - f_const)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double r = a.ce * se[c] + a.cf * sf[c];
        if (a.fext) r -= a.fext[3 * i + c];
        a.r[3 * i + c] = r;
    }
}


constexpr int kMaxW = kMaxSupernodeWidth;

template <typename T, int WAVES = 4>
struct TrsvFwdLds {
    T sb[kMaxW * 3];
    T red[WAVES][64][3];
};

__device__ __forceinline__ long long itemVal(const int4& d1)
{
    return static_cast<long long>(static_cast<unsigned>(d1.z)) | (static_cast<long long>(d1.w) << 32);
}

// One forward item -- 64 rows [r0, r0 + 64) of supernode s -- by the
// 64 WAVES lanes of a block (contains barriers; every lane must call it).
// The WAVES waves split the column range and are summed in wave order.
// Descriptor d0 = {r0, w, m, j0}, d1 = {front row offset, update-vector
// offset, B offset lo, hi}.  bp is the right-hand side already permuted
// (bp[3 j + c] = b[3 perm[j] + c]).  Multifrontal: the front of s is its
// rows of bp plus its children's update vectors (extend-add through
// fptr/fidx); y_s = L11^-1 f[:w], u_s = f[w:] - (L21 L11^-1) f[:w].
// Extend-add table of one forward item (KT > 0): KT x 64 update-vector
// positions for the item's 64 rows, then KT x W for the supernode's own
// columns (k-major); absent entries point to a zero slot.  The table's
// address depends on the item index only, so its loads do not wait for the
// descriptor.
struct FwdTable {
    const int* tab;  // this item's entries
    int W;           // own-column stride (the level's widest supernode)
};

template <typename T, int WAVES = 4, int KT = 0>
__device__ __forceinline__ void trsvFwdItem(const TrsvArgs& a, const double* __restrict__ bp, int4 d0, int4 d1,
                                            TrsvFwdLds<T, WAVES>& L, FwdTable ft = FwdTable{nullptr, 0})
{
    constexpr int NT = 64 * WAVES;
    const int r0 = d0.x, w = d0.y, m = d0.z, j0 = d0.w, fr0 = d1.x, tb = d1.y;
    const int ld = ldOf(m);
    const long long val = itemVal(d1);
    T* sb = L.sb;
    T* u = static_cast<T*>(a.t);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = r0 + lane;
    // GEMV shape: lane = row, wave q takes columns [kb, ke); B is read in
    // batches of NBT columns, each batch loaded before the previous one is
    // consumed, and the first batch before the right-hand side is gathered
    // (its latency overlaps the gather's dependent index loads)
    constexpr int NBT = 8;
    const int kc = (w + WAVES - 1) / WAVES;
    const int kb = wave * kc;
    int ke = 0;
    if (r < m) {
        ke = min(w, kb + kc);
        if (r < w) ke = min(ke, r + 1);  // L11^-1 is lower triangular
    }
    const T* Bc = static_cast<const T*>(a.B) + val + r;
    T nb[NBT];
#pragma unroll
    for (int q = 0; q < NBT; ++q) nb[q] = kb + q < ke ? Bc[static_cast<size_t>(kb + q) * ld] : T(0);
    // update rows: the children's contributions (wave 0, lane = row)
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    if constexpr (KT > 0) {
        // every position loaded, then every value: two dependent loads, none
        // waiting for the descriptor (pivot rows and absent entries read
        // the zero slot; sums in the order of the row lists)
        if (wave == 0) {
            int q[KT];
#pragma unroll
            for (int k = 0; k < KT; ++k) q[k] = ft.tab[k * 64 + lane];
            T uv[KT][3];
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                const T* up = u + 3 * static_cast<size_t>(q[k]);
                uv[k][0] = up[0];
                uv[k][1] = up[1];
                uv[k][2] = up[2];
            }
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                f0 += static_cast<double>(uv[k][0]);
                f1 += static_cast<double>(uv[k][1]);
                f2 += static_cast<double>(uv[k][2]);
            }
        }
        const int* own = ft.tab + KT * 64;
        for (int k = threadIdx.x; k < w; k += NT) {
            int q[KT];
#pragma unroll
            for (int j = 0; j < KT; ++j) q[j] = own[j * ft.W + k];
            const double* bk = bp + 3 * static_cast<size_t>(j0 + k);
            double v0 = bk[0], v1 = bk[1], v2 = bk[2];
            T uv[KT][3];
#pragma unroll
            for (int j = 0; j < KT; ++j) {
                const T* up = u + 3 * static_cast<size_t>(q[j]);
                uv[j][0] = up[0];
                uv[j][1] = up[1];
                uv[j][2] = up[2];
            }
#pragma unroll
            for (int j = 0; j < KT; ++j) {
                v0 += static_cast<double>(uv[j][0]);
                v1 += static_cast<double>(uv[j][1]);
                v2 += static_cast<double>(uv[j][2]);
            }
            sb[3 * k] = static_cast<T>(v0);
            sb[3 * k + 1] = static_cast<T>(v1);
            sb[3 * k + 2] = static_cast<T>(v2);
        }
    } else {
        if (wave == 0 && r >= w && r < m) {
            const int p1 = a.fptr[fr0 + r + 1];
            for (int p = a.fptr[fr0 + r]; p < p1; ++p) {
                const T* uv = u + 3 * static_cast<size_t>(a.fidx[p]);
                f0 += static_cast<double>(uv[0]);
                f1 += static_cast<double>(uv[1]);
                f2 += static_cast<double>(uv[2]);
            }
        }
        // front entries of the supernode's own columns (short lists: at most
        // one entry per child)
        for (int k = threadIdx.x; k < w; k += NT) {
            const double* bk = bp + 3 * static_cast<size_t>(j0 + k);
            double v0 = bk[0], v1 = bk[1], v2 = bk[2];
            const int p1 = a.fptr[fr0 + k + 1];
            for (int p = a.fptr[fr0 + k]; p < p1; ++p) {
                const T* uv = u + 3 * static_cast<size_t>(a.fidx[p]);
                v0 += static_cast<double>(uv[0]);
                v1 += static_cast<double>(uv[1]);
                v2 += static_cast<double>(uv[2]);
            }
            sb[3 * k] = static_cast<T>(v0);
            sb[3 * k + 1] = static_cast<T>(v1);
            sb[3 * k + 2] = static_cast<T>(v2);
        }
    }
    __syncthreads();
    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    for (int k = kb; k < ke; k += NBT) {
        T b[NBT];
#pragma unroll
        for (int q = 0; q < NBT; ++q) b[q] = nb[q];
        if (k + NBT < ke) {
#pragma unroll
            for (int q = 0; q < NBT; ++q)
                nb[q] = k + NBT + q < ke ? Bc[static_cast<size_t>(k + NBT + q) * ld] : T(0);
        }
#pragma unroll
        for (int q = 0; q < NBT; ++q) {
            if (k + q < ke) {
                acc0 += b[q] * sb[3 * (k + q)];
                acc1 += b[q] * sb[3 * (k + q) + 1];
                acc2 += b[q] * sb[3 * (k + q) + 2];
            }
        }
    }
    L.red[wave][lane][0] = acc0;
    L.red[wave][lane][1] = acc1;
    L.red[wave][lane][2] = acc2;
    __syncthreads();
    if (wave != 0 || r >= m) return;
    acc0 = L.red[0][lane][0];
    acc1 = L.red[0][lane][1];
    acc2 = L.red[0][lane][2];
#pragma unroll
    for (int q = 1; q < WAVES; ++q) {
        acc0 += L.red[q][lane][0];
        acc1 += L.red[q][lane][1];
        acc2 += L.red[q][lane][2];
    }
    if (r < w) {
        T* out = static_cast<T*>(a.y) + 3 * static_cast<size_t>(j0 + r);
        out[0] = acc0;
        out[1] = acc1;
        out[2] = acc2;
        return;
    }
    T* out = u + 3 * static_cast<size_t>(tb + r - w);
    out[0] = static_cast<T>(f0) - acc0;
    out[1] = static_cast<T>(f1) - acc1;
    out[2] = static_cast<T>(f2) - acc2;
}

// One backward item: four columns k0 .. k0 + 3 of supernode s.  With
// WAVES = 4 one wave per column (no barrier); with WAVES = 16 four waves
// per column split its rows and are summed in order through LDS (contains a
// barrier; every lane must call it).  Descriptor d0 = {k0, w, m, j0},
// d1 = {row offset, 0, B offset lo, hi}.
// z_s = L11^-T (y_s - L21^T z_ancestors) = B11^T y_s - B21^T z[rows].
template <typename T, int WAVES = 4>
__device__ __forceinline__ void trsvBwdItem(const TrsvArgs& a, int4 d0, int4 d1, T (*red)[3] = nullptr)
{
    constexpr int G = WAVES / 4;  // waves per column
    const int w = d0.y, m = d0.z, j0 = d0.w;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = wave / G, part = wave - col * G;
    const int k = d0.x + col;
    const int* rows = a.rows + d1.x;
    const T* y = static_cast<const T*>(a.y);
    T* z = static_cast<T*>(a.z);
    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    if (k < w) {
        const T* cl = static_cast<const T*>(a.B) + itemVal(d1) + static_cast<size_t>(k) * ldOf(m);
        constexpr int step = 64 * G;
        // from the line below row k: rows k0 .. k - 1 of the column hold the
        // stored zeros of L11^-1's upper triangle, so a sweep aligned to 32
        // rows reads whole lines at no cost to the result
        int r = (k & ~31) + part * 64 + lane;
        // four rows at a time, their B values and row indices loaded before
        // the dependent gathers (same summation order as one by one).  FP32:
        // software-pipelined -- the next four rows' values and indices load
        // while this four's gathers are in flight (measured: FP32 apply
        // -1.7 %, FP64 +2 % from the extra registers, so FP64 keeps the plain
        // batches)
        if constexpr (sizeof(T) == 8) {
            for (; r + 3 * step < m; r += 4 * step) {
                int rr[4];
                T bk[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    rr[u] = r + u * step;
                    bk[u] = cl[rr[u]];
                }
                int gi[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) gi[u] = rr[u] < w ? -1 : rows[rr[u]];
                T sv[4][3];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const T* src =
                        gi[u] < 0 ? y + 3 * static_cast<size_t>(j0 + rr[u]) : z + 3 * static_cast<size_t>(gi[u]);
                    const T sg = gi[u] < 0 ? T(1) : T(-1);
                    sv[u][0] = sg * src[0];
                    sv[u][1] = sg * src[1];
                    sv[u][2] = sg * src[2];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    acc0 += bk[u] * sv[u][0];
                    acc1 += bk[u] * sv[u][1];
                    acc2 += bk[u] * sv[u][2];
                }
            }
        } else if (r + 3 * step < m) {
            int rr[4], gi[4];
            T bk[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                rr[u] = r + u * step;
                bk[u] = cl[rr[u]];
                gi[u] = rr[u] < w ? -1 : rows[rr[u]];
            }
            for (;;) {
                T sv[4][3];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const T* src =
                        gi[u] < 0 ? y + 3 * static_cast<size_t>(j0 + rr[u]) : z + 3 * static_cast<size_t>(gi[u]);
                    const T sg = gi[u] < 0 ? T(1) : T(-1);
                    sv[u][0] = sg * src[0];
                    sv[u][1] = sg * src[1];
                    sv[u][2] = sg * src[2];
                }
                const int rn = r + 4 * step;
                const bool more = rn + 3 * step < m;
                int rr2[4], gi2[4];
                T bk2[4];
                if (more) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        rr2[u] = rn + u * step;
                        bk2[u] = cl[rr2[u]];
                        gi2[u] = rr2[u] < w ? -1 : rows[rr2[u]];
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    acc0 += bk[u] * sv[u][0];
                    acc1 += bk[u] * sv[u][1];
                    acc2 += bk[u] * sv[u][2];
                }
                r = rn;
                if (!more) break;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    rr[u] = rr2[u];
                    bk[u] = bk2[u];
                    gi[u] = gi2[u];
                }
            }
        }
        for (; r < m; r += step) {
            const T bk = cl[r];
            const T* src = r < w ? y + 3 * static_cast<size_t>(j0 + r) : z + 3 * static_cast<size_t>(rows[r]);
            const T sg = r < w ? T(1) : T(-1);
            acc0 += bk * (sg * src[0]);
            acc1 += bk * (sg * src[1]);
            acc2 += bk * (sg * src[2]);
        }
    }
    acc0 = waveSumT(acc0);
    acc1 = waveSumT(acc1);
    acc2 = waveSumT(acc2);
    if (G == 1) {
        if (lane == 0 && k < w) {
            T* zo = z + 3 * static_cast<size_t>(j0 + k);
            zo[0] = acc0;
            zo[1] = acc1;
            zo[2] = acc2;
        }
        return;
    }
    if (lane == 0) {  // wave = col * G + part
        red[wave][0] = acc0;
        red[wave][1] = acc1;
        red[wave][2] = acc2;
    }
    __syncthreads();
    if (part == 0 && lane == 0 && k < w) {
        T s0 = red[col * G][0], s1 = red[col * G][1], s2 = red[col * G][2];
        for (int q = 1; q < G; ++q) {
            s0 += red[col * G + q][0];
            s1 += red[col * G + q][1];
            s2 += red[col * G + q][2];
        }
        T* zo = z + 3 * static_cast<size_t>(j0 + k);
        zo[0] = s0;
        zo[1] = s1;
        zo[2] = s2;
    }
}

// Staged backward item: columns k0 .. k0 + W - 1 of supernode s by W waves,
// one column each.  The right-hand side v = [y_s; -z[rows]] is staged in
// LDS once per item, CH rows at a time from the 64-row line holding k0
// (rows above a column meet the zero upper triangle of L11^-1), and shared
// by the W columns; each wave issues its column's loads of a chunk before
// the block gathers the chunk's v (every load of the chunk in flight at
// once).  Row r goes to lane r mod 64 in increasing order, so a column's
// sum does not depend on how columns are grouped into items.  Descriptor
// d0 = {k0, w, m, j0} (k0 a multiple of W), d1 = {row offset, 0, B offset
// lo, hi}.  Contains barriers; every lane must call it.
template <typename T>
struct BwdStage {
    static constexpr int NB = sizeof(T) == 4 ? 32 : 16;  // column loads per lane per chunk
    static constexpr int CH = 64 * NB;                    // rows per chunk (LDS 24 KB)
};
template <typename T>
struct TrsvBwdStageLds {
    T v[3][BwdStage<T>::CH];
};

template <typename T, int W>
__device__ __forceinline__ void trsvBwdStageItem(const TrsvArgs& a, int4 d0, int4 d1, TrsvBwdStageLds<T>& L)
{
    constexpr int NT = 64 * W, NB = BwdStage<T>::NB, CH = BwdStage<T>::CH;
    constexpr int RPT = (CH + NT - 1) / NT;  // rows of v each lane stages per chunk
    const int k0 = d0.x, w = d0.y, m = d0.z, j0 = d0.w;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = k0 + wave;
    const bool own = k < w;
    const int* rows = a.rows + d1.x;
    const T* y = static_cast<const T*>(a.y);
    T* z = static_cast<T*>(a.z);
    const T* cl = static_cast<const T*>(a.B) + itemVal(d1) + static_cast<size_t>(min(k, w - 1)) * ldOf(m);
    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    const int base = k0 & ~63;
    for (int c0 = base; c0 < m; c0 += CH) {
        const int cn = min(CH, m - c0);
        T bk[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int r = c0 + lane + 64 * i;
            bk[i] = own && r < m ? cl[r] : T(0);
        }
        if (c0 != base) __syncthreads();  // the previous chunk is done with L.v
        const int cr = (cn + 63) & ~63;   // rows staged (zeros past cn): the rows the columns read
        int gi[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int i = threadIdx.x + j * NT, r = c0 + i;
            gi[j] = i < cn && r >= w ? rows[r] : -1;
        }
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int i = threadIdx.x + j * NT, r = c0 + i;
            if (i >= cr) break;
            T v0 = T(0), v1 = T(0), v2 = T(0);
            if (i < cn) {
                if (r < w) {
                    const T* src = y + 3 * static_cast<size_t>(j0 + r);
                    v0 = src[0];
                    v1 = src[1];
                    v2 = src[2];
                } else {
                    const T* src = z + 3 * static_cast<size_t>(gi[j]);
                    v0 = -src[0];
                    v1 = -src[1];
                    v2 = -src[2];
                }
            }
            L.v[0][i] = v0;
            L.v[1][i] = v1;
            L.v[2][i] = v2;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            if (64 * i >= cn) break;
            const int q = lane + 64 * i;
            acc0 += bk[i] * L.v[0][q];
            acc1 += bk[i] * L.v[1][q];
            acc2 += bk[i] * L.v[2][q];
        }
    }
    acc0 = waveSumT(acc0);
    acc1 = waveSumT(acc1);
    acc2 = waveSumT(acc2);
    if (lane == 0 && own) {
        T* zo = z + 3 * static_cast<size_t>(j0 + k);
        zo[0] = acc0;
        zo[1] = acc1;
        zo[2] = acc2;
    }
}

// Backward item for levels of many narrow supernodes: up to 16 columns
// [k0, k0 + 16) of supernode s by a 256-lane block.  The block stages the
// right-hand side v = [y_s; -z[rows]] of 256 rows at a time in LDS; lane
// t = 16 kk + p accumulates column k0 + kk over the rows p, p + 16, ... of
// the chunk (16 independent loads in flight), and the 16 lanes of a column
// -- adjacent lanes of one wave -- meet in a fixed butterfly.  (4-column
// items give each wave one column, which multiplies the waves of such a
// level beyond what the chip holds at once.)  Descriptor as trsvBwdItem
// with k0 = 0 mod 16.  Contains barriers; every lane must call it.
template <typename T>
struct TrsvBwdNarrowLds {
    T v[256 * 3];
};

template <typename T>
__device__ __forceinline__ void trsvBwdNarrowItem(const TrsvArgs& a, int4 d0, int4 d1, TrsvBwdNarrowLds<T>& L)
{
    const int k0 = d0.x, w = d0.y, m = d0.z, j0 = d0.w;
    const int ld = ldOf(m);
    const int t = threadIdx.x, kk = t >> 4, ph = t & 15;
    const int k = k0 + kk;
    const int* rows = a.rows + d1.x;
    const T* y = static_cast<const T*>(a.y);
    T* z = static_cast<T*>(a.z);
    const T* cl = static_cast<const T*>(a.B) + itemVal(d1) + static_cast<size_t>(min(k, w - 1)) * ld;
    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    // rows below k0 meet only the zero upper triangle of L11^-1
    for (int c0 = k0; c0 < m; c0 += 256) {
        const int c1 = min(m, c0 + 256);
        T bk[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = c0 + ph + 16 * i;
            bk[i] = (k < w && r < c1) ? cl[r] : T(0);
        }
        {
            const int r = c0 + t;
            T v0 = T(0), v1 = T(0), v2 = T(0);
            if (r < c1) {
                const bool own = r < w;
                const T* src = own ? y + 3 * static_cast<size_t>(j0 + r) : z + 3 * static_cast<size_t>(rows[r]);
                const T sg = own ? T(1) : T(-1);
                v0 = sg * src[0];
                v1 = sg * src[1];
                v2 = sg * src[2];
            }
            L.v[3 * t] = v0;
            L.v[3 * t + 1] = v1;
            L.v[3 * t + 2] = v2;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int q = 3 * (ph + 16 * i);
            acc0 += bk[i] * L.v[q];
            acc1 += bk[i] * L.v[q + 1];
            acc2 += bk[i] * L.v[q + 2];
        }
        __syncthreads();
    }
    acc0 = rowSum16(acc0);
    acc1 = rowSum16(acc1);
    acc2 = rowSum16(acc2);
    if (ph == 0 && k < w) {
        T* zo = z + 3 * static_cast<size_t>(j0 + k);
        zo[0] = acc0;
        zo[1] = acc1;
        zo[2] = acc2;
    }
}

}  // namespace
}  // namespace dev
}  // namespace efem
