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
//  * reductions: two launches, block partials in a fixed tree, then one
//    block that combines them in block order and runs the epilogue, so
//    every dot product is bitwise reproducible and no block pays a
//    device-scope fence.
//  * preconditioner: level-scheduled supernodal triangular solves on the
//    block-inverse form B_s = [L11^-1 ; L21 L11^-1] with three right-hand
//    sides (the interleaved x, y, z components); the forward sweep is
//    multifrontal (each supernode extend-adds only its children's update
//    vectors), the backward sweep reads the ancestors' solved values.
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>

#include "device_common.hpp"

namespace efem {
namespace dev {
namespace {

// K block reductions over W waves (defined with the reductions below)
template <int K, int W, unsigned MM>
__device__ __forceinline__ void blockReduce(double (&v)[K]);

template <typename T, int LAW, int MODE, int TPB>
__global__ __launch_bounds__(TPB) void kElemTet4(int E, const int* __restrict__ connDof,
                                                 const double* __restrict__ ginv,
                                                 const double* __restrict__ vol, double w0,
                                                 const double* __restrict__ params, int ppe,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ v,
                                                 T* __restrict__ slots, const double* __restrict__ Xd, int nF)
{
    const int e = xcdBlock() * TPB + threadIdx.x;
    if (e >= E) return;
    tet4Lane<T, LAW, MODE>(e, E, connDof, ginv, vol, w0, params, ppe, x, v, slots, Xd, nF);
}

template <typename T, int NEN, int NG, int LAW, int MODE>
__global__ __launch_bounds__(kThreads) void kElemMulti(int E, const int* __restrict__ connDof,
                                                       const double* __restrict__ ginv,
                                                       const double* __restrict__ vol,
                                                       const double* __restrict__ gdN,
                                                       const double* __restrict__ gw,
                                                       const double* __restrict__ params, int ppe,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ v,
                                                       T* __restrict__ slots)
{
    // the kinematics (relative displacements, local gradients, H) are FP64 in
    // both modes; the stress moments Q and the nodal forces use T
    __shared__ ElemTileLds<T, NEN, NG, MODE> L;
    for (int k = threadIdx.x; k < NG * NEN * 3; k += kThreads) L.sdN[k] = gdN[k];
    elemTile<T, NEN, NG, LAW, MODE>(E, xcdBlock() * Tile<NG>::epb, E, connDof, ginv, vol, gw, params, ppe, x, v,
                                    slots, L);
}

template <typename T, int NEN, int NG, int MODE>
void launchMultiLaw(const ElemArgs& a, const double* v, hipStream_t s)
{
    const int EPB = Tile<NG>::epb;
    const int grid = (a.nElems + EPB - 1) / EPB;
    const double* ginv = static_cast<const double*>(a.ginv);
    const double* gdN = static_cast<const double*>(a.gdN);
    T* slots = static_cast<T*>(a.slots);
    switch (a.law) {
        case LAW_NH:
            kElemMulti<T, NEN, NG, LAW_NH, MODE><<<grid, kThreads, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, gdN,
                                                                            a.gw, a.params, a.paramsPerElem, a.x, v,
                                                                            slots);
            break;
        case LAW_MR:
            kElemMulti<T, NEN, NG, LAW_MR, MODE><<<grid, kThreads, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, gdN,
                                                                            a.gw, a.params, a.paramsPerElem, a.x, v,
                                                                            slots);
            break;
        default:
            kElemMulti<T, NEN, NG, LAW_YEOH, MODE><<<grid, kThreads, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, gdN,
                                                                              a.gw, a.params, a.paramsPerElem, a.x,
                                                                              v, slots);
    }
}

template <typename T, int MODE, int TPB>
void launchTet4Tpb(const ElemArgs& a, const double* v, hipStream_t s)
{
    const int grid = (a.nElems + TPB - 1) / TPB;
    const double* ginv = static_cast<const double*>(a.ginv);
    T* slots = static_cast<T*>(a.slots);
    switch (a.law) {
        case LAW_NH:
            kElemTet4<T, LAW_NH, MODE, TPB><<<grid, TPB, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, 1.0, a.params,
                                                                 a.paramsPerElem, a.x, v, slots, a.Xd, a.nFree);
            break;
        case LAW_MR:
            kElemTet4<T, LAW_MR, MODE, TPB><<<grid, TPB, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, 1.0, a.params,
                                                                 a.paramsPerElem, a.x, v, slots, a.Xd, a.nFree);
            break;
        default:
            kElemTet4<T, LAW_YEOH, MODE, TPB><<<grid, TPB, 0, s>>>(a.nElems, a.connDof, ginv, a.vol, 1.0,
                                                                   a.params, a.paramsPerElem, a.x, v, slots, a.Xd,
                                                                   a.nFree);
    }
}

template <typename T, int MODE>
void launchTet4Law(const ElemArgs& a, const double* v, hipStream_t s)
{
    switch (tuning().tet4Threads) {
        case 64: launchTet4Tpb<T, MODE, 64>(a, v, s); break;
        case 128: launchTet4Tpb<T, MODE, 128>(a, v, s); break;
        default: launchTet4Tpb<T, MODE, 256>(a, v, s);
    }
}

template <typename T, int MODE>
void launchElem(const ElemArgs& a, const double* v, hipStream_t s)
{
    if (a.nen == 4) return launchTet4Law<T, MODE>(a, v, s);
    if (a.nen == 8 && a.ng == 8) return launchMultiLaw<T, 8, 8, MODE>(a, v, s);
    if (a.nen == 8 && a.ng == 1) return launchMultiLaw<T, 8, 1, MODE>(a, v, s);
    if (a.nen == 10 && a.ng == 4) return launchMultiLaw<T, 10, 4, MODE>(a, v, s);
    if (a.nen == 10 && a.ng == 1) return launchMultiLaw<T, 10, 1, MODE>(a, v, s);
    if (a.nen == 20 && a.ng == 27) return launchMultiLaw<T, 20, 27, MODE>(a, v, s);
    if (a.nen == 20 && a.ng == 8) return launchMultiLaw<T, 20, 8, MODE>(a, v, s);
}

template <int NEN, int MODE>
__global__ __launch_bounds__(kThreads) void kPressure(FaceArgs a)
{
    const int f = blockIdx.x * kThreads + threadIdx.x;
    if (f < a.nFaces) pressureFace<NEN, MODE>(a, f);
}

// Follower-pressure force, tiled: FPB faces per block.  Phase 1, one lane
// per (face, Gauss point): F from the LDS-staged relative displacements,
// v = cof(F) A0 (or p_dir |cof F A0|) into LDS; phase 2, one lane per
// (face, node): f_a = p sum_q w_q N_a v_q.  Same arithmetic per face as
// pressureFace<NEN, PMODE_FORCE> (which one lane per face would run
// serially over up to 9 Gauss points and 20 nodes).
template <int NEN>
struct PressureTile {
    static constexpr int kMaxQ = 9;
    static constexpr int fpb = 28;  // 28 x 9 Gauss points <= 256 lanes
};

// One tile of FPB = BT / kMaxQ faces by the BT lanes of a block (contains
// barriers; every lane must call it): relative nodal displacements staged
// in LDS, one lane per (face, Gauss point) forms cof(F) A0, then one lane
// per (face, node) its nodal forces p sum_q WN_q v_q into the face slots.
template <int NEN, int BT>
__device__ __forceinline__ void pressureTileBlock(const FaceArgs& a, int blk)
{
    constexpr int FPB = BT / PressureTile<NEN>::kMaxQ;
    __shared__ double sdu[FPB][NEN][3];
    __shared__ double sv[FPB][PressureTile<NEN>::kMaxQ][3];
    const int F = a.nFaces, fq = a.fq, tid = threadIdx.x;
    const int f0 = blk * FPB;
    for (int k = tid; k < FPB * NEN; k += BT) {
        const int fl = k / NEN, n = k - fl * NEN, f = f0 + fl;
        if (f < F) {
            const int id = a.faceDof[n * F + f], id0 = a.faceDof[f];
#pragma unroll
            for (int c = 0; c < 3; ++c) sdu[fl][n][c] = a.x[3 * id + c] - a.x[3 * id0 + c];
        }
    }
    __syncthreads();
    {
        const int fl = tid / fq, q = tid - fl * fq, f = f0 + fl;
        if (fl < FPB && f < F) {
            const double* G = a.G + (static_cast<size_t>(f) * fq + q) * NEN * 3;
            double Fm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            for (int n = 1; n < NEN; ++n)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) Fm[3 * i + j] += sdu[fl][n][i] * G[3 * n + j];
            double cf[9];
            cofactor(Fm, cf);
            const double A0[3] = {a.A0[3 * f], a.A0[3 * f + 1], a.A0[3 * f + 2]};
            double v[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) v[i] = cf[3 * i] * A0[0] + cf[3 * i + 1] * A0[1] + cf[3 * i + 2] * A0[2];
            if (a.directed) {
                const double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                v[0] = nv * a.pdir[0];
                v[1] = nv * a.pdir[1];
                v[2] = nv * a.pdir[2];
            }
            sv[fl][q][0] = v[0];
            sv[fl][q][1] = v[1];
            sv[fl][q][2] = v[2];
        }
    }
    __syncthreads();
    for (int k = tid; k < FPB * NEN; k += BT) {
        const int fl = k / NEN, n = k - fl * NEN, f = f0 + fl;
        if (f >= F) continue;
        double o0 = 0.0, o1 = 0.0, o2 = 0.0;
        for (int q = 0; q < fq; ++q) {
            const double wn = a.WN[(static_cast<size_t>(f) * fq + q) * NEN + n];
            o0 += wn * sv[fl][q][0];
            o1 += wn * sv[fl][q][1];
            o2 += wn * sv[fl][q][2];
        }
        const double pr = a.fpress ? a.fpress[f] : a.pressure;
        double* dst = a.slots + (static_cast<size_t>(f) * NEN + n) * 3;
        dst[0] = pr * o0;
        dst[1] = pr * o1;
        dst[2] = pr * o2;
    }
}

template <int NEN>
__global__ __launch_bounds__(kThreads) void kPressureTile(FaceArgs a)
{
    pressureTileBlock<NEN, kThreads>(a, blockIdx.x);
}

// C3D4 element forces and the pressure faces in one launch: blocks
// [0, nEB) are element lanes (remapped per XCD over that range), the rest
// face tiles of TPB / kMaxQ faces
template <typename T, int LAW, int TPB>
__global__ __launch_bounds__(TPB) void kElemTet4Faces(ElemArgs ea, FaceArgs fa, int nEB)
{
    if (static_cast<int>(blockIdx.x) < nEB) {
        const int e = xcdBlockOf(blockIdx.x, nEB) * TPB + threadIdx.x;
        if (e < ea.nElems)
            tet4Lane<T, LAW, MODE_FORCE>(e, ea.nElems, ea.connDof, static_cast<const double*>(ea.ginv), ea.vol,
                                         1.0, ea.params, ea.paramsPerElem, ea.x, nullptr,
                                         static_cast<T*>(ea.slots), ea.Xd, ea.nFree);
        return;
    }
    pressureTileBlock<4, TPB>(fa, blockIdx.x - nEB);
}

// C3D4 element forces summed per patch in LDS (see PatchArgs): blocks
// [0, nEB) are patches (remapped per XCD), the rest face tiles
template <typename T, int LAW>
__global__ __launch_bounds__(kPatch) void kElemTet4Patch(ElemArgs ea, PatchArgs pa, FaceArgs fa, int nEB)
{
    if (static_cast<int>(blockIdx.x) >= nEB) {
        pressureTileBlock<4, kPatch>(fa, blockIdx.x - nEB);
        return;
    }
    __shared__ T sq[kPatch * 12];
    const int p = xcdBlockOf(blockIdx.x, nEB);
    const int t = threadIdx.x;
    const int pe = p * kPatch + t;
    const int r0 = pa.roff[p], nr = pa.roff[p + 1] - r0;
    // the lane's first record and its first 16 entries, loaded before the
    // element's work
    int rb = 0, rn = 0, rp = 0;
    uint4 c0 = make_uint4(0, 0, 0, 0), c1 = make_uint4(0, 0, 0, 0);
    if (t < nr) {
        const int4 rc = reinterpret_cast<const int4*>(pa.rec)[r0 + t];
        rb = rc.x;
        rn = rc.y;
        rp = rc.z;
        const uint4* ep = reinterpret_cast<const uint4*>(pa.ent + rb);
        c0 = ep[0];
        if (rn > 8) c1 = ep[1];
    }
    if (pe < pa.E) {
        int id[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) id[a] = pa.conn[a * pa.E + pe];
        T f[12];
        tet4Forces<T, LAW, MODE_FORCE>(id, pa.order[pe], ea.nElems, static_cast<const double*>(ea.ginv),
                                       pa.vol[pe] * 1.0, ea.params, ea.paramsPerElem, ea.x, nullptr, ea.Xd,
                                       ea.nFree, f);
#pragma unroll
        for (int k = 0; k < 12; ++k) sq[12 * t + k] = f[k];
    }
    __syncthreads();
    for (int j = t; j < nr; j += kPatch) {
        if (j != t) {
            const int4 rc = reinterpret_cast<const int4*>(pa.rec)[r0 + j];
            rb = rc.x;
            rn = rc.y;
            rp = rc.z;
            const uint4* ep = reinterpret_cast<const uint4*>(pa.ent + rb);
            c0 = ep[0];
            c1 = rn > 8 ? ep[1] : make_uint4(0, 0, 0, 0);
        }
        const unsigned w[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i < rn) {
                const unsigned q = (w[i >> 1] >> (16 * (i & 1))) & 0xffffu;
                s0 += static_cast<double>(sq[3 * q]);
                s1 += static_cast<double>(sq[3 * q + 1]);
                s2 += static_cast<double>(sq[3 * q + 2]);
            }
        }
        for (int i = 16; i < rn; ++i) {
            const unsigned q = pa.ent[rb + i];
            s0 += static_cast<double>(sq[3 * q]);
            s1 += static_cast<double>(sq[3 * q + 1]);
            s2 += static_cast<double>(sq[3 * q + 2]);
        }
        double* o = pa.part + 3 * static_cast<size_t>(rp);
        o[0] = s0;
        o[1] = s1;
        o[2] = s2;
    }
}

template <typename T>
void launchPatchesT(const ElemArgs& ea, const PatchArgs& pa, const FaceArgs& fa, hipStream_t s)
{
    constexpr int FPB = kPatch / PressureTile<4>::kMaxQ;
    const int nEB = pa.nPatch;
    const int nFB = fa.nFaces > 0 ? (fa.nFaces + FPB - 1) / FPB : 0;
    if (nEB + nFB == 0) return;
    switch (ea.law) {
        case LAW_NH: kElemTet4Patch<T, LAW_NH><<<nEB + nFB, kPatch, 0, s>>>(ea, pa, fa, nEB); break;
        case LAW_MR: kElemTet4Patch<T, LAW_MR><<<nEB + nFB, kPatch, 0, s>>>(ea, pa, fa, nEB); break;
        default: kElemTet4Patch<T, LAW_YEOH><<<nEB + nFB, kPatch, 0, s>>>(ea, pa, fa, nEB);
    }
}

template <typename T, int LAW, int TPB>
void launchTet4FacesTpb(const ElemArgs& ea, const FaceArgs& fa, hipStream_t s)
{
    constexpr int FPB = TPB / PressureTile<4>::kMaxQ;
    const int nEB = (ea.nElems + TPB - 1) / TPB;
    const int nFB = (fa.nFaces + FPB - 1) / FPB;
    kElemTet4Faces<T, LAW, TPB><<<nEB + nFB, TPB, 0, s>>>(ea, fa, nEB);
}

template <typename T, int TPB>
void launchTet4FacesLaw(const ElemArgs& ea, const FaceArgs& fa, hipStream_t s)
{
    switch (ea.law) {
        case LAW_NH: launchTet4FacesTpb<T, LAW_NH, TPB>(ea, fa, s); break;
        case LAW_MR: launchTet4FacesTpb<T, LAW_MR, TPB>(ea, fa, s); break;
        default: launchTet4FacesTpb<T, LAW_YEOH, TPB>(ea, fa, s);
    }
}

template <typename T>
void launchTet4Faces(const ElemArgs& ea, const FaceArgs& fa, hipStream_t s)
{
    switch (tuning().tet4Threads) {
        case 64: launchTet4FacesLaw<T, 64>(ea, fa, s); break;
        case 128: launchTet4FacesLaw<T, 128>(ea, fa, s); break;
        default: launchTet4FacesLaw<T, 256>(ea, fa, s);
    }
}

// ---- per-element scalars (FP64, one lane per element) ---------------------
// MODE 0: energy  out[e] = sum_q w_q V W(F_q)
// MODE 1: parameter sensitivities of lam . f_int,
//         out[4e + k] = sum_q w_q V (dP/dp_k)(F_q) : H_lam,q
// (H_lam = grad lam at the Gauss point).  Used once per adjoint or energy
// call, so the simple layout (global reads, no LDS) is enough.
template <int NEN, int MODE>
__global__ __launch_bounds__(kThreads) void kElemScalar(ElemArgs a, const double* __restrict__ lam,
                                                        double* __restrict__ out)
{
    const int e = blockIdx.x * kThreads + threadIdx.x;
    const int E = a.nElems;
    if (e >= E) return;
    const double* x = a.x;
    const double* ginv = static_cast<const double*>(a.ginv);
    const double* gdN = static_cast<const double*>(a.gdN);
    const double* p = a.params + (a.paramsPerElem ? 4 * e : 0);
    int id[NEN];
#pragma unroll
    for (int n = 0; n < NEN; ++n) id[n] = a.connDof[n * E + e];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < a.ng; ++q) {
        const double* dq = gdN + static_cast<size_t>(q) * NEN * 3;
        double Gu[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Gl[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int n = 1; n < NEN; ++n)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double du = x[3 * id[n] + i] - x[3 * id[0] + i];
                const double dl = MODE == 1 ? lam[3 * id[n] + i] - lam[3 * id[0] + i] : 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    Gu[3 * i + k] += du * dq[3 * n + k];
                    if (MODE == 1) Gl[3 * i + k] += dl * dq[3 * n + k];
                }
            }
        double Jd[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Jd[k] = ginv[(static_cast<size_t>(q) * 9 + k) * E + e];
        double Hd[9];
        gradToH(Gu, Jd, Hd);
        const double wv = a.vol[e] * a.gw[q];
        if (MODE == 0) {
            acc[0] += wv * energyFromH(a.law, p, Hd);
        } else {
            double Hl[9];
            gradToH(Gl, Jd, Hl);
            for (int k = 0; k < 4; ++k) {
                double dP[9];
                piolaParamDerivative<double>(a.law, p, Hd, k, dP);
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < 9; ++j) t += dP[j] * Hl[j];
                acc[k] += wv * t;
            }
        }
    }
    if (MODE == 0) {
        out[e] = acc[0];
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[4 * static_cast<size_t>(e) + k] = acc[k];
    }
}

// LPN lanes per free node (they split its slot list, then combine in a fixed
// butterfly), so LPN times as many independent loads are in flight.  DOT:
// the block's partials of the line-search scalars of the new r as well
// (K = 3, or 4 with the distance to a target), at dotWork[k nb + block]
template <typename T, int LPN, int DOT = 0>
__global__ __launch_bounds__(kThreads) void kGather(GatherArgs a)
{
    const int blk = xcdBlock();
    const int t = blk * kThreads + threadIdx.x;
    const int i = t / LPN, sub = t % LPN;
    const T* es = static_cast<const T*>(a.eslots);
    double se[3] = {0, 0, 0}, sf[3] = {0, 0, 0};
    if (i < a.nFree) {
        if (a.rptr && a.ce != 0.0) {
            const double* er = static_cast<const double*>(a.eslots);
            const int q1 = a.rptr[i + 1];
            for (int q = a.rptr[i] + sub; q < q1; q += LPN) {
                const double* v = er + 3 * static_cast<size_t>(q);
                se[0] += v[0];
                se[1] += v[1];
                se[2] += v[2];
            }
        }
        const int p1 = a.gptr[i + 1];
        for (int p = a.gptr[i] + sub; p < p1; p += LPN) {
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
    }
#pragma unroll
    for (int off = 1; off < LPN; off <<= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            se[c] += __shfl_xor(se[c], off, 64);
            sf[c] += __shfl_xor(sf[c], off, 64);
        }
    if constexpr (DOT == 0) {
        if (i >= a.nFree || sub != 0) return;
        // residual: ce = 1, cf = -1 (r = f_int - f_pressure - f_const)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double r = a.ce * se[c] + a.cf * sf[c];
            if (a.fext) r -= a.fext[3 * i + c];
            a.r[3 * i + c] = r;
        }
    } else {
        double v[DOT];
#pragma unroll
        for (int k = 0; k < DOT; ++k) v[k] = 0.0;
        if (i < a.nFree && sub == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double r = a.ce * se[c] + a.cf * sf[c];
                if (a.fext) r -= a.fext[3 * i + c];
                a.r[3 * i + c] = r;
                const double xi = a.dotX[3 * i + c];
                v[0] += r * a.dotD[3 * i + c];
                v[1] = maxNan(v[1], fabs(r));
                v[2] = maxNan(v[2], fabs(a.dotCb ? xi - a.dotCb[3 * i + c] : xi));
                if constexpr (DOT == 4) {
                    const double e = xi - a.dotXt[3 * i + c];
                    v[3] += e * e;
                }
            }
        }
        blockReduce<DOT, kThreads / 64, 0x6u>(v);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < DOT; ++k) a.dotWork[static_cast<size_t>(k) * gridDim.x + blk] = v[k];
        }
    }
}

// ---- approximate Jacobian ----------------------------------------------------
// One lane per (element, column slot b): column b of the element matrix,
// accumulated over the Gauss points as the host does (the gradients G are
// formed per Gauss point; no multiply-add contraction anywhere).
template <int NEN>
__global__ __launch_bounds__(kThreads) void kLapElem(LapArgs a)
{
#pragma clang fp contract(off)
    const long long t = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    if (t >= static_cast<long long>(a.E) * NEN) return;
    const int e = static_cast<int>(t / NEN), b = static_cast<int>(t - static_cast<long long>(e) * NEN);
    const int E = a.E;
    const double ks = a.vol[e] * elasticityCoeff(a.law, a.params + (a.ppe ? 4 * static_cast<size_t>(e) : 0));
    double kcol[NEN];
#pragma unroll
    for (int i = 0; i < NEN; ++i) kcol[i] = 0.0;
    for (int q = 0; q < a.ng; ++q) {
        double Ji[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Ji[k] = a.ginv[(static_cast<size_t>(q) * 9 + k) * E + e];
        const double* dN = a.gdN + static_cast<size_t>(q) * NEN * 3;
        const double gb0 = Ji[0] * dN[3 * b] + Ji[3] * dN[3 * b + 1] + Ji[6] * dN[3 * b + 2];
        const double gb1 = Ji[1] * dN[3 * b] + Ji[4] * dN[3 * b + 1] + Ji[7] * dN[3 * b + 2];
        const double gb2 = Ji[2] * dN[3 * b] + Ji[5] * dN[3 * b + 1] + Ji[8] * dN[3 * b + 2];
        const double w = ks * a.gw[q];
#pragma unroll
        for (int i = 0; i < NEN; ++i) {
            const double g0 = Ji[0] * dN[3 * i] + Ji[3] * dN[3 * i + 1] + Ji[6] * dN[3 * i + 2];
            const double g1 = Ji[1] * dN[3 * i] + Ji[4] * dN[3 * i + 1] + Ji[7] * dN[3 * i + 2];
            const double g2 = Ji[2] * dN[3 * i] + Ji[5] * dN[3 * i + 1] + Ji[8] * dN[3 * i + 2];
            kcol[i] += w * (g0 * gb0 + g1 * gb1 + g2 * gb2);
        }
    }
    double* out = a.kel + static_cast<size_t>(e) * NEN * NEN;
#pragma unroll
    for (int i = 0; i < NEN; ++i) out[static_cast<size_t>(i) * NEN + b] = kcol[i];
}

// one lane per CSR entry: its contributions summed in the listed order
__global__ __launch_bounds__(kThreads) void kLapSum(LapArgs a)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= a.nnz) return;
    double sum = 0.0;
    for (int p = a.ptr[k]; p < a.ptr[k + 1]; ++p) sum += a.kel[a.src[p]];
    a.vals[k] = sum;
}

// ---- reductions -------------------------------------------------------------
// Two launches per reduction.  Stage 1 streams the vectors with kRB-lane
// blocks and writes each block's partials (wave butterflies, then the wave
// results in wave order) to work[k * kReduceBlocks + block]; stage 2, one
// kRF-lane block, combines the partials in block order with the same tree
// and hands the results to the reduction's epilogue.  The kernel boundary
// orders the partials before stage 2: a single-launch form (each block
// fences, counts its arrival, the last one finishes) pays a device-scope
// fence per block, which on this part writes back the XCD's L2 and grows
// with the grid.  Every reduction shape depends on n and the tuned grid
// only (bitwise reproducible).
constexpr int kRB = 256;  // stage-1 block
constexpr int kRBW = kRB / 64;
constexpr int kRF = 256;  // stage-2 block (one lane per block partial)
constexpr int kRFW = kRF / 64;

// K simultaneous block reductions over the block's 64 W lanes, through
// LDS in chunks of CH components: every lane stores its values; 16 lanes
// per component each combine a strip of 4 W values (lanes j, j + 16, ...,
// in that order); one lane per component combines the 16 strips in order.
// Bit k of the mask MM (a template argument): component k is a maximum
// that propagates NaN (else a sum; components from 32 on are sums).  All
// lanes return the block result.  (Wave butterflies of shuffles took 5.9
// against 2.2 us for 19 components in one block, 7.4 against 2.9 us as the
// last step of a 256-block sweep.)
__host__ __device__ constexpr bool isMax(unsigned maxMask, int k) { return k < 32 && ((maxMask >> k) & 1u); }

template <int K, int W, unsigned MM>
__device__ __forceinline__ void blockReduce(double (&v)[K])
{
    constexpr int NT = 64 * W, G = NT / 16;
    constexpr int CH0 = 4096 / NT;
    constexpr int CH = K < CH0 ? K : CH0;
    __shared__ double buf[CH][NT];
    __shared__ double part[CH][16];
    __shared__ double fin[K];
    const int tid = threadIdx.x;
#pragma unroll
    for (int c0 = 0; c0 < K; c0 += CH) {
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (c0 + c < K) buf[c][tid] = v[c0 + c];
        __syncthreads();
        if (tid < CH * 16) {
            const int c = tid >> 4, j = tid & 15;
            if (c0 + c < K) {
                const bool mx = isMax(MM, c0 + c);
                double x[G];
#pragma unroll
                for (int i = 0; i < G; ++i) x[i] = buf[c][i * 16 + j];
                double t = x[0];
#pragma unroll
                for (int i = 1; i < G; ++i) t = mx ? maxNan(t, x[i]) : t + x[i];
                part[c][j] = t;
            }
        }
        __syncthreads();
        if (tid < CH && c0 + tid < K) {
            const bool mx = isMax(MM, c0 + tid);
            double x[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) x[jj] = part[tid][jj];
            double t = x[0];
#pragma unroll
            for (int jj = 1; jj < 16; ++jj) t = mx ? maxNan(t, x[jj]) : t + x[jj];
            fin[c0 + tid] = t;
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = fin[k];
}

// stage 1: this block's partials of v[0..K) to work
template <int K, unsigned MM>
__device__ __forceinline__ void writePartials(double (&v)[K], double* work)
{
    blockReduce<K, kRBW, MM>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) work[k * kReduceBlocks + blockIdx.x] = v[k];
    }
}

// stage 2: the nb block partials of each component combined (lane t takes
// blocks t, t + kRF, ...; then the block tree); every lane calls epi(v, p),
// p = epi.prefetch() -- what the epilogue reads besides the sums, loaded
// before the partials so that the two loads overlap
template <int K, unsigned MM, typename Epi>
__device__ __forceinline__ void reduceFinishBody(const double* __restrict__ work, int nb, Epi epi, int stride)
{
    const typename Epi::Pre pre = epi.prefetch();
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    // a few components: each lane loads NB of its partials at once (the
    // same order of combination as one by one)
    constexpr int NB = K <= 4 ? 8 : 1;
    for (int b0 = threadIdx.x; b0 < nb; b0 += NB * kRF) {
        double o[NB][K];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int b = b0 + q * kRF;
#pragma unroll
            for (int k = 0; k < K; ++k) o[q][k] = b < nb ? work[static_cast<size_t>(k) * stride + b] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            if (b0 + q * kRF >= nb) break;
#pragma unroll
            for (int k = 0; k < K; ++k) v[k] = isMax(MM, k) ? maxNan(v[k], o[q][k]) : v[k] + o[q][k];
        }
    }
    blockReduce<K, kRFW, MM>(v);
    epi(v, pre);
}

template <int K, unsigned MM, typename Epi>
__global__ __launch_bounds__(kRF) void kReduceFinish(const double* __restrict__ work, int nb, Epi epi,
                                                     int stride = kReduceBlocks)
{
    reduceFinishBody<K, MM, Epi>(work, nb, epi, stride);
}

struct NoPrefetch {
    struct Pre {};
    __device__ Pre prefetch() const { return {}; }
};

int reduceGrid(int n) { return max(1, min(min(kReduceBlocks, tuning().reduceBlocks), (n + kRB - 1) / kRB)); }
// reductions that stream the whole history (2 m + 4 vectors) read several
// times the bytes of a line-search dot, so their grid is tuned on its own
int historyGrid(int n) { return max(1, min(min(kReduceBlocks, tuning().historyBlocks), (n + kRB - 1) / kRB)); }

// epilogue: out[k] = v[k]
template <int K>
struct EpiStore : NoPrefetch {
    double* out;
    __device__ void operator()(const double (&v)[K], Pre) const
    {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] = v[k];
        }
    }
};

// epilogue: out[k] = v[k] and out[4] = *extra (a device scalar carried
// along: the trial's slope), out in pinned host memory
template <int K>
struct EpiStoreCarry : NoPrefetch {
    double* out;
    const double* extra;
    __device__ void operator()(const double (&v)[K], Pre) const
    {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] = v[k];
            out[4] = *extra;
        }
    }
};

// epilogue: *p0 = v[0], *p1 = v[1]
struct EpiStore2 : NoPrefetch {
    double *p0, *p1;
    __device__ void operator()(const double (&v)[2], Pre) const
    {
        if (threadIdx.x == 0) {
            *p0 = v[0];
            *p1 = v[1];
        }
    }
};

// out = {a.b, max|a|, max|c - cb|} (HASC; cb may be null) and
// out[3] = |c - xt|^2 (HASXT)
template <bool HASC, bool HASXT>
__global__ __launch_bounds__(kRB) void kDotMax(const double* __restrict__ a, const double* __restrict__ b,
                                               const double* __restrict__ c, const double* __restrict__ cb,
                                               const double* __restrict__ xt, int n, double* work)
{
    constexpr int K = HASXT ? 4 : 3;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
#pragma unroll 4
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += gridDim.x * kRB) {
        const double ai = a[i];
        v[0] += ai * b[i];
        v[1] = maxNan(v[1], fabs(ai));
        if (HASC) {
            const double ci = c[i];
            v[2] = maxNan(v[2], fabs(cb ? ci - cb[i] : ci));
            if (HASXT) {
                const double e = ci - xt[i];
                v[K - 1] += e * e;
            }
        }
    }
    writePartials<K, 0x6u>(v, work);
}

__global__ __launch_bounds__(kRB) void kDot2(const double* __restrict__ a, const double* __restrict__ b, int n,
                                             double* work)
{
    double v[2] = {0, 0};
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += gridDim.x * kRB) {
        const double bi = b[i];
        v[0] += a[i] * bi;
        v[1] += bi * bi;
    }
    writePartials<2, 0u>(v, work);
}

__global__ __launch_bounds__(kRB) void kTwoLoop(double* __restrict__ q, const double* __restrict__ vprev,
                                                const double* cprev, double sgn,
                                                const double* __restrict__ w, int n, double* work)
{
    double v[1] = {0};
    const double c = vprev ? sgn * (*cprev) : 0.0;
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += gridDim.x * kRB) {
        double qi = q[i];
        if (vprev) {
            qi += c * vprev[i];
            q[i] = qi;
        }
        v[0] += w[i] * qi;
    }
    writePartials<1, 0u>(v, work);
}

// epilogue of a two-loop step: r = v / *den; *cout = mode ? *aj - r : r
struct EpiTwoLoopStep : NoPrefetch {
    const double *den, *aj;
    double* cout;
    int mode;
    __device__ void operator()(const double (&v)[1], Pre) const
    {
        if (threadIdx.x == 0) {
            const double r = v[0] / (*den);
            *cout = mode ? (*aj - r) : r;
        }
    }
};

// ---- compact two-loop ------------------------------------------------------
// Kernels are templated on the history capacity KC (8 covers the reference's
// m = 6; 16 is the maximum) and on the number of stored pairs NC, so that
// every per-thread array has compile-time indexing (no scratch) and every
// load of an element is issued before the first is used (a run-time trip
// count makes each pair's loads wait for the previous pair's).

// layout of the Gram matrix in the scalar area: G[i * kMaxCompact + k]
constexpr int KG = kMaxCompact;

// One pass for the new correction pair and the first loop: s_loc = x - xp
// and y_loc = r - rp are formed (and stored) on the fly, and the same sweep
// over the history takes the Gram row/column of the new pair, y_loc.y_loc
// and the first loop's s_k . r.  Pairs are taken by age t (slot (loc - t)
// mod mcap: the ring order of the stored pairs; age 0 is the new pair);
// the epilogue maps them back to slots.  Partial sums by age:
// v[t] = s_loc . y_t, v[NC + t] = s_t . y_loc, v[2 NC] = y_loc . y_loc,
// v[2 NC + 1 + t] = s_t . r.
template <int NC>
__device__ __forceinline__ void gramStep(double (&v)[3 * NC + 1], double ri, double xi, double xpi, double rpi,
                                         const double* sk, const double* yk, double& si, double& yi)
{
    si = xi - xpi;
    yi = ri - rpi;
    v[0] += si * yi;
    v[NC] += si * yi;
    v[2 * NC] += yi * yi;
    v[2 * NC + 1] += si * ri;
#pragma unroll
    for (int t = 1; t < NC; ++t) {
        v[t] += si * yk[t - 1];
        v[NC + t] += sk[t - 1] * yi;
        v[2 * NC + 1 + t] += sk[t - 1] * ri;
    }
}

__device__ __forceinline__ double2 ld2(const double* p, int i) { return *reinterpret_cast<const double2*>(p + i); }

// P2: each lane takes element pairs (2 p, 2 p + 1) with 16-byte loads and
// stores (n even, every stream 16-byte aligned), the pair's elements summed
// in index order
template <int NC, bool P2>
__global__ __launch_bounds__(kRB) void kGramFirst(double* __restrict__ S, double* __restrict__ Y,
                                                      const double* __restrict__ x, const double* __restrict__ xp,
                                                      const double* __restrict__ r, const double* __restrict__ rp,
                                                      int n, int loc, int mcap, double* work, int g1,
                                                      const double* __restrict__ tfWork, int tfNb, int tfK,
                                                      double* tfOut, const double* tfCarry)
{
    // block g1 (when tfK > 0): a trial's finishing reduction, riding along
    if (tfK && static_cast<int>(blockIdx.x) == g1) {
        if (tfK == 4)
            reduceFinishBody<4, 0x6u>(tfWork, tfNb, EpiStoreCarry<4>{{}, tfOut, tfCarry}, tfNb);
        else
            reduceFinishBody<3, 0x6u>(tfWork, tfNb, EpiStoreCarry<3>{{}, tfOut, tfCarry}, tfNb);
        return;
    }
    static_assert(NC >= 1 && NC <= kMaxCompact, "pair count");
    constexpr int N = 3 * NC + 1;
    double v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = 0.0;
    const double* sAge[NC > 1 ? NC - 1 : 1];
    const double* yAge[NC > 1 ? NC - 1 : 1];
#pragma unroll
    for (int t = 1; t < NC; ++t) {
        const int k = (loc - t + mcap) % mcap;
        sAge[t - 1] = S + static_cast<size_t>(k) * n;
        yAge[t - 1] = Y + static_cast<size_t>(k) * n;
    }
    double* sl = S + static_cast<size_t>(loc) * n;
    double* yl = Y + static_cast<size_t>(loc) * n;
    constexpr int NA = NC > 1 ? NC - 1 : 1;
    if constexpr (P2) {
        const int np = n >> 1;
        for (int p = blockIdx.x * kRB + threadIdx.x; p < np; p += g1 * kRB) {
            const int i = 2 * p;
            const double2 r2 = ld2(r, i), x2 = ld2(x, i), xp2 = ld2(xp, i), rp2 = ld2(rp, i);
            double2 sk2[NA], yk2[NA];
#pragma unroll
            for (int t = 1; t < NC; ++t) {
                sk2[t - 1] = ld2(sAge[t - 1], i);
                yk2[t - 1] = ld2(yAge[t - 1], i);
            }
            double sk[NA], yk[NA], s0, y0, s1, y1;
#pragma unroll
            for (int t = 0; t < NC - 1; ++t) {
                sk[t] = sk2[t].x;
                yk[t] = yk2[t].x;
            }
            gramStep<NC>(v, r2.x, x2.x, xp2.x, rp2.x, sk, yk, s0, y0);
#pragma unroll
            for (int t = 0; t < NC - 1; ++t) {
                sk[t] = sk2[t].y;
                yk[t] = yk2[t].y;
            }
            gramStep<NC>(v, r2.y, x2.y, xp2.y, rp2.y, sk, yk, s1, y1);
            *reinterpret_cast<double2*>(sl + i) = make_double2(s0, s1);
            *reinterpret_cast<double2*>(yl + i) = make_double2(y0, y1);
        }
    } else {
        for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += g1 * kRB) {
            const double ri = r[i], xi = x[i], xpi = xp[i], rpi = rp[i];
            double sk[NA], yk[NA];
#pragma unroll
            for (int t = 1; t < NC; ++t) {
                sk[t - 1] = sAge[t - 1][i];
                yk[t - 1] = yAge[t - 1][i];
            }
            double si, yi;
            gramStep<NC>(v, ri, xi, xpi, rpi, sk, yk, si, yi);
            sl[i] = si;
            yl[i] = yi;
        }
    }
    writePartials<N, 0u>(v, work);
}

// Epilogue of kGramFirst: writes the Gram row / column of loc, ys_loc, yy,
// theta = yy / ys_loc, rs_k = s_k . r, and runs the alpha recurrence newest
// -> oldest (q_0 = -r): alpha_i = (s_i.q_0 - sum_{k newer} alpha_k s_i.y_k)
// / ys_i.
template <int KC, int NC>
struct EpiGramFirst {
    static_assert(KC * KC <= kRF, "one Gram entry per lane");
    double *G, *ys, *yy, *theta, *alpha, *rs;
    int loc, mcap;
    // lane t: stored Gram entry t and ys_t
    struct Pre {
        double g, y;
    };
    __device__ Pre prefetch() const
    {
        const int t = threadIdx.x, i = t / KC, k = t - i * KC;
        Pre p{0.0, 1.0};
        if (t < KC * KC && i < mcap && k < mcap) p.g = G[i * KG + k];
        if (t < mcap) p.y = ys[t];
        return p;
    }
    __device__ void operator()(const double (&v)[3 * NC + 1], Pre pre) const
    {
        constexpr int N = 3 * NC + 1;
        __shared__ double red[N];
        __shared__ double sG[KC][KC], sys[KC], sA[KC], sB[KC], sC[KC];
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < N; ++k) red[k] = v[k];
        }
        __syncthreads();
        // by slot: sA[k] = s_loc . y_k, sB[k] = s_k . y_loc, sC[k] = s_k . r
        if (threadIdx.x < mcap) {
            const int k = threadIdx.x;
            const int t = (loc - k + mcap) % mcap;
            sA[k] = t < NC ? red[t] : 0.0;
            sB[k] = t < NC ? red[NC + t] : 0.0;
            sC[k] = t < NC ? red[2 * NC + 1 + t] : 0.0;
        }
        __syncthreads();
        // Gram block in LDS: stored entries, with the new row / column
        if (threadIdx.x < KC * KC) {
            const int t = threadIdx.x, i = t / KC, k = t - i * KC;
            double g = pre.g;
            if (i < mcap && k < mcap) {
                if (i == loc && k < NC) g = sA[k];
                else if (k == loc && i < NC) g = sB[i];
            }
            sG[i][k] = g;
        }
        if (threadIdx.x < KC) {
            const int k = threadIdx.x;
            sys[k] = k == loc ? red[0] : pre.y;
        }
        __syncthreads();
        if (threadIdx.x < NC) {
            const int k = threadIdx.x;
            G[loc * KG + k] = sA[k];
            G[k * KG + loc] = sB[k];
            rs[k] = sC[k];
        }
        if (threadIdx.x == 0) {
            ys[loc] = red[0];
            *yy = red[2 * NC];
            *theta = red[2 * NC] / red[0];
            // by age (compile-time indices; the slots are independent)
            int slot[NC];
#pragma unroll
            for (int t = 0; t < NC; ++t) slot[t] = (loc - t + mcap) % mcap;
            double a[NC];
#pragma unroll
            for (int t = 0; t < NC; ++t) {
                double acc = -red[2 * NC + 1 + t];
#pragma unroll
                for (int u = 0; u < t; ++u) acc -= a[u] * sG[slot[t]][slot[u]];
                a[t] = acc / sys[slot[t]];
                alpha[slot[t]] = a[t];
            }
        }
    }
};

// element i of an H0 output (see HSrc)
__device__ __forceinline__ double hAt(const HSrc& q, int i)
{
    if (!q.pinv) return q.h[i];
    const int node = i / 3;
    const size_t j = 3 * static_cast<size_t>(q.pinv[node]) + (i - 3 * node);
    return q.zF32 ? static_cast<double>(static_cast<const float*>(q.z)[j]) : static_cast<const double*>(q.z)[j];
}

// Second loop of the two-loop recursion: dots of Y_k with q (the H0 output,
// stored pairs in slots 0..NC-1) and of dw with q in one sweep.
// P2: element pairs with 16-byte loads of R and dw (as kGramFirst)
template <int NC, bool P2>
__global__ __launch_bounds__(kRB) void kTwoLoopDots(HSrc q, const double* __restrict__ R, int n,
                                                        const double* __restrict__ dw, double* work)
{
    static_assert(NC >= 1 && NC <= kMaxCompact, "pair count");
    double v[NC + 1];
#pragma unroll
    for (int k = 0; k <= NC; ++k) v[k] = 0.0;
    if constexpr (P2) {
        const int np = n >> 1;
        for (int p = blockIdx.x * kRB + threadIdx.x; p < np; p += gridDim.x * kRB) {
            const int i = 2 * p;
            const double q0 = hAt(q, i), q1 = hAt(q, i + 1);
            const double2 dw2 = ld2(dw, i);
            double2 rk[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k) rk[k] = ld2(R + static_cast<size_t>(k) * n, i);
#pragma unroll
            for (int k = 0; k < NC; ++k) v[k] += rk[k].x * q0;
            v[NC] += dw2.x * q0;
#pragma unroll
            for (int k = 0; k < NC; ++k) v[k] += rk[k].y * q1;
            v[NC] += dw2.y * q1;
        }
    } else {
        for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += gridDim.x * kRB) {
            const double qi = hAt(q, i), dwi = dw[i];
            double rk[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k) rk[k] = R[static_cast<size_t>(k) * n + i];
#pragma unroll
            for (int k = 0; k < NC; ++k) v[k] += rk[k] * qi;
            v[NC] += dwi * qi;
        }
    }
    writePartials<NC + 1, 0u>(v, work);
}

// Epilogue of kTwoLoopDots: the beta recurrence oldest -> newest, coef =
// alpha - beta, and the slope of d = q + sum coef_k s_k along dw: dw . d =
// dw . q + sum coef_k rs_k (rs_k = s_k . dw from kGramFirst).
template <int KC, int NC>
struct EpiTwoLoopDots {
    static_assert(KC * KC <= kRF, "one Gram entry per lane");
    const double *G, *ys, *alphaIn, *rs;
    double *outCoef, *slope;
    int newest, mcap;
    // lane t: Gram entry t, ys_t and alpha_t
    struct Pre {
        double g, y, a;
    };
    __device__ Pre prefetch() const
    {
        const int t = threadIdx.x, i = t / KC, k = t - i * KC;
        Pre p{0.0, 1.0, 0.0};
        if (t < KC * KC && i < mcap && k < mcap) p.g = G[i * KG + k];
        if (t < mcap) {
            p.y = ys[t];
            p.a = alphaIn[t];
        }
        return p;
    }
    __device__ void operator()(const double (&v)[NC + 1], Pre pre) const
    {
        __shared__ double red[NC + 1];
        __shared__ double sG[KC][KC], sys[KC], sa[KC], cs[KC];
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k <= NC; ++k) red[k] = v[k];
        }
        // the recurrence reads the Gram block, ys and alpha from LDS
        if (threadIdx.x < KC * KC) {
            const int t = threadIdx.x, i = t / KC, k = t - i * KC;
            sG[i][k] = pre.g;
        }
        if (threadIdx.x < KC) {
            sys[threadIdx.x] = pre.y;
            sa[threadIdx.x] = pre.a;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // beta_i = (y_i.h + sum_{k older} coef_k y_i.s_k) / ys_i ; coef =
            // alpha - beta, oldest -> newest (position t from the oldest;
            // compile-time indices, the slots are independent)
            const int oldest = (newest + mcap - NC + 1) % mcap;
            int slot[NC];
#pragma unroll
            for (int t = 0; t < NC; ++t) slot[t] = (oldest + t) % mcap;
            double b[NC];
#pragma unroll
            for (int t = 0; t < NC; ++t) {
                double acc = red[slot[t]];
#pragma unroll
                for (int u = 0; u < t; ++u) acc += b[u] * sG[slot[u]][slot[t]];
                b[t] = sa[slot[t]] - acc / sys[slot[t]];
                outCoef[slot[t]] = b[t];
                cs[slot[t]] = b[t];
            }
            // the stored pairs are slots 0..NC-1: summed in slot order
            double sl = red[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k) sl += cs[k] * rs[k];
            *slope = sl;
        }
    }
};

// out = base * sgn_b + sum_{k < NC} coef_k R_k * sgn_c; with xout, also
// xout = xbase + out (a line-search step of length 1 from xbase: the same
// value kStep forms from the stored out)
// With pinv, out is written in the factor's order instead: out[3 pinv[i / 3]
// + i % 3] (the permuted right-hand side the triangular solves read).  The
// base is read as an H0 output (HSrc).
template <int NC>
__global__ __launch_bounds__(kThreads) void kMultiAxpy(HSrc base, double sgnB, const double* __restrict__ R,
                                                       const double* coef, double sgnC, int n,
                                                       double* __restrict__ out,
                                                       const double* __restrict__ xbase = nullptr,
                                                       double* __restrict__ xout = nullptr,
                                                       const int* __restrict__ pinv = nullptr)
{
    __shared__ double c[NC];
    if (threadIdx.x < NC) c[threadIdx.x] = sgnC * coef[threadIdx.x];
    __syncthreads();
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const double b = hAt(base, i);
    double rk[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) rk[k] = R[static_cast<size_t>(k) * n + i];
    double v = sgnB * b;
#pragma unroll
    for (int k = 0; k < NC; ++k) v += c[k] * rk[k];
    if (pinv) {
        const int node = i / 3;
        out[3 * static_cast<size_t>(pinv[node]) + (i - 3 * node)] = v;
    } else {
        out[i] = v;
    }
    if (xout) xout[i] = xbase[i] + 1.0 * v;
}

// ---- compensated sum --------------------------------------------------------
// Neumaier accumulation per lane, then (sum, compensation) pairs combined in a
// fixed-shape tree with the exact TwoSum error of every addition; the last
// block to arrive combines the block pairs in block order.
__device__ __forceinline__ void neumaier(double& s, double& c, double x)
{
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

__device__ __forceinline__ void combinePair(double& s, double& c, double s2, double c2)
{
    const double t = s + s2;
    const double bp = t - s;
    const double err = (s - (t - bp)) + (s2 - bp);  // TwoSum
    s = t;
    c = c + c2 + err;
}

__global__ __launch_bounds__(kRB) void kSumComp(const double* __restrict__ a, int n, double* work)
{
    __shared__ double ss[kRB], sc[kRB];
    const int tid = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (int i = blockIdx.x * kRB + tid; i < n; i += gridDim.x * kRB) neumaier(s, c, a[i]);
    ss[tid] = s;
    sc[tid] = c;
    __syncthreads();
    for (int w = kRB / 2; w > 0; w >>= 1) {
        if (tid < w) {
            double s1 = ss[tid], c1 = sc[tid];
            combinePair(s1, c1, ss[tid + w], sc[tid + w]);
            ss[tid] = s1;
            sc[tid] = c1;
        }
        __syncthreads();
    }
    if (tid == 0) {
        work[blockIdx.x] = ss[0];
        work[kReduceBlocks + blockIdx.x] = sc[0];
    }
}

// the block pairs combined in block order
__global__ void kSumCompFinish(const double* __restrict__ work, int nb, double* out)
{
    if (threadIdx.x != 0) return;
    double S = 0.0, C = 0.0;
    for (int b = 0; b < nb; ++b) combinePair(S, C, work[b], work[kReduceBlocks + b]);
    out[0] = S + C;
}

// Potential difference of the energy line search, block partials:
// compensated sum of (Wn[e] - Wo[e]) over elements and of -(x - xp) . fext
// over the DOFs (the constant loads' work), plain sums of |Wo[e]| and of
// xp . fext (the rounding scale).  work: [s | c | |W| | ext] x kReduceBlocks.
__global__ __launch_bounds__(kRB) void kEnergyDiff(const double* __restrict__ Wn, const double* __restrict__ Wo,
                                                   int E, const double* __restrict__ fext,
                                                   const double* __restrict__ x, const double* __restrict__ xp,
                                                   int n, double* work)
{
    __shared__ double ss[kRB], sc[kRB], sa[kRB], se[kRB];
    const int tid = threadIdx.x;
    double s = 0.0, c = 0.0, aw = 0.0, ex = 0.0;
    for (int e = blockIdx.x * kRB + tid; e < E; e += gridDim.x * kRB) {
        neumaier(s, c, Wn[e] - Wo[e]);
        aw += fabs(Wo[e]);
    }
    if (fext)
        for (int i = blockIdx.x * kRB + tid; i < n; i += gridDim.x * kRB) {
            neumaier(s, c, -(x[i] - xp[i]) * fext[i]);
            ex += xp[i] * fext[i];
        }
    ss[tid] = s;
    sc[tid] = c;
    sa[tid] = aw;
    se[tid] = ex;
    __syncthreads();
    for (int w = kRB / 2; w > 0; w >>= 1) {
        if (tid < w) {
            double s1 = ss[tid], c1 = sc[tid];
            combinePair(s1, c1, ss[tid + w], sc[tid + w]);
            ss[tid] = s1;
            sc[tid] = c1;
            sa[tid] += sa[tid + w];
            se[tid] += se[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        work[blockIdx.x] = ss[0];
        work[kReduceBlocks + blockIdx.x] = sc[0];
        work[2 * kReduceBlocks + blockIdx.x] = sa[0];
        work[3 * kReduceBlocks + blockIdx.x] = se[0];
    }
}

// out[0] = the potential difference, out[1] = its rounding scale
// sum |Wo| + |xp . fext| (block partials combined in block order)
__global__ void kEnergyDiffFinish(const double* __restrict__ work, int nb, double* out)
{
    if (threadIdx.x != 0) return;
    double S = 0.0, C = 0.0, A = 0.0, X = 0.0;
    for (int b = 0; b < nb; ++b) {
        combinePair(S, C, work[b], work[kReduceBlocks + b]);
        A += work[2 * kReduceBlocks + b];
        X += work[3 * kReduceBlocks + b];
    }
    out[0] = S + C;
    out[1] = A + fabs(X);
}

// out = a x + b y (y may be null)
__global__ void kAxpby(double* out, double a, const double* x, double b, const double* y, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = y ? a * x[i] + b * y[i] : a * x[i];
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
template <typename T, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void kTrsvFwd(TrsvArgs a, const double* __restrict__ bp,
                                                       const int4* __restrict__ items)
{
    __shared__ TrsvFwdLds<T, WAVES> L;
    trsvFwdItem<T, WAVES>(a, bp, items[2 * blockIdx.x], items[2 * blockIdx.x + 1], L);
}

// the same with the level's extend-add table (KT entries per front row)
template <typename T, int WAVES, int KT>
__global__ __launch_bounds__(64 * WAVES) void kTrsvFwdTab(TrsvArgs a, const double* __restrict__ bp,
                                                          const int4* __restrict__ items,
                                                          const int* __restrict__ tab, int W)
{
    __shared__ TrsvFwdLds<T, WAVES> L;
    const FwdTable ft{tab + static_cast<size_t>(blockIdx.x) * KT * (64 + W), W};
    trsvFwdItem<T, WAVES, KT>(a, bp, items[2 * blockIdx.x], items[2 * blockIdx.x + 1], L, ft);
}

// The fused top: z[tcol[k]] = sum_i Z[i, k] f[i], one wave per column k,
// W columns per block.  Each pass covers RC rows: the block first loads the
// indices of its rows of f (each lane RPT rows: the table's update
// positions and the factor column), then each wave issues its column
// segment (NQ 16-byte loads per lane, VE consecutive rows each), then the
// values of f (the permuted right-hand side plus the children's update
// rows, summed in FP64 in table order) go to LDS, and the column is
// multiplied and reduced in a fixed butterfly.  Rows past nT read Z's zero
// padding against zeros in LDS.
template <typename T>
struct TopVec;
template <>
struct TopVec<float> {
    using V = float4;
};
template <>
struct TopVec<double> {
    using V = double2;
};

template <typename T, int W>
__global__ __launch_bounds__(64 * W) void kTrsvTop(TrsvArgs a, const double* __restrict__ bp, TopArgs t)
{
    using V = typename TopVec<T>::V;
    constexpr int NT = 64 * W;
    constexpr int VE = 16 / sizeof(T);      // values per 16-byte load
    constexpr int NQ = 14;                  // loads per lane per pass
    constexpr int RC = NQ * 64 * VE;        // rows per pass: 3584 (FP32), 1792 (FP64); LDS 43 KB
    constexpr int RPT = (RC + NT - 1) / NT; // rows of f each lane stages per pass
    __shared__ T sf[3][RC];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * W + wave;
    const T* col = static_cast<const T*>(t.Z) + static_cast<size_t>(min(k, t.nT - 1)) * t.ldz;
    const T* u = static_cast<const T*>(a.t);
    T acc0 = T(0), acc1 = T(0), acc2 = T(0);
    for (int p0 = 0; p0 < t.nT; p0 += RC) {
        const int pn = min(RC, t.nT - p0);
        int tc[RPT], q[RPT][kTopMaxK];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int i = threadIdx.x + r * NT;
            const bool on = i < pn;
            tc[r] = on ? t.tcol[p0 + i] : -1;
#pragma unroll
            for (int j = 0; j < kTopMaxK; ++j) q[r][j] = on && j < t.K ? t.tab[static_cast<size_t>(j) * t.nT + p0 + i] : -1;
        }
        V zq[NQ];
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            const int r = p0 + (c * 64 + lane) * VE;
            // ldz is a multiple of 32 rows: a 16-byte vector lies wholly
            // inside the column or wholly past it
            zq[c] = r < t.ldz ? *reinterpret_cast<const V*>(col + r) : V{};
        }
        if (p0 > 0) __syncthreads();  // the previous pass is done with sf
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int i = threadIdx.x + r * NT;
            if (i >= RC) break;
            double f0 = 0.0, f1 = 0.0, f2 = 0.0;
            if (tc[r] >= 0) {
                const double* bk = bp + 3 * static_cast<size_t>(tc[r]);
                f0 = bk[0];
                f1 = bk[1];
                f2 = bk[2];
#pragma unroll
                for (int j = 0; j < kTopMaxK; ++j) {
                    if (q[r][j] >= 0) {
                        const T* up = u + 3 * static_cast<size_t>(q[r][j]);
                        f0 += static_cast<double>(up[0]);
                        f1 += static_cast<double>(up[1]);
                        f2 += static_cast<double>(up[2]);
                    }
                }
            }
            sf[0][i] = static_cast<T>(f0);
            sf[1][i] = static_cast<T>(f1);
            sf[2][i] = static_cast<T>(f2);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            const int r = (c * 64 + lane) * VE;
            const T* zv = reinterpret_cast<const T*>(&zq[c]);
#pragma unroll
            for (int v = 0; v < VE; ++v) {
                acc0 += zv[v] * sf[0][r + v];
                acc1 += zv[v] * sf[1][r + v];
                acc2 += zv[v] * sf[2][r + v];
            }
        }
    }
    acc0 = waveSumT(acc0);
    acc1 = waveSumT(acc1);
    acc2 = waveSumT(acc2);
    if (lane == 0 && k < t.nT) {
        T* zo = static_cast<T*>(a.z) + 3 * static_cast<size_t>(t.tcol[k]);
        zo[0] = acc0;
        zo[1] = acc1;
        zo[2] = acc2;
    }
}

template <typename T, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void kTrsvBwd(TrsvArgs a, const int4* __restrict__ items)
{
    __shared__ T red[WAVES][3];
    trsvBwdItem<T, WAVES>(a, items[2 * blockIdx.x], items[2 * blockIdx.x + 1], red);
}

template <typename T, int W>
__global__ __launch_bounds__(64 * W) void kTrsvBwdStage(TrsvArgs a, const int4* __restrict__ items)
{
    __shared__ TrsvBwdStageLds<T> L;
    trsvBwdStageItem<T, W>(a, items[2 * blockIdx.x], items[2 * blockIdx.x + 1], L);
}

template <typename T>
__global__ __launch_bounds__(256) void kTrsvBwdNarrow(TrsvArgs a, const int4* __restrict__ items)
{
    __shared__ TrsvBwdNarrowLds<T> L;
    trsvBwdNarrowItem<T>(a, items[2 * blockIdx.x], items[2 * blockIdx.x + 1], L);
}

// bp[3 j + c] = b[3 perm[j] + c]
__global__ void kPermute(const int* __restrict__ perm, const double* __restrict__ b, double* __restrict__ bp, int n)
{
    const int j = blockIdx.x * kThreads + threadIdx.x;
    if (j < n) {
        const int o = perm[j];
        bp[3 * j] = b[3 * o];
        bp[3 * j + 1] = b[3 * o + 1];
        bp[3 * j + 2] = b[3 * o + 2];
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

thread_local LaunchTuning g_tuning;

}  // namespace

void setTuning(const LaunchTuning& t) { g_tuning = t; }
const LaunchTuning& tuning() { return g_tuning; }

void launchElementForce(const ElemArgs& a, hipStream_t s)
{
    if (a.mixed)
        launchElem<float, MODE_FORCE>(a, nullptr, s);
    else
        launchElem<double, MODE_FORCE>(a, nullptr, s);
}

void launchElementPatches(const ElemArgs& a, const PatchArgs& pa, const FaceArgs& f, hipStream_t s)
{
    // faces whose Gauss points fit the tile go into the same launch
    const bool fused = f.nFaces > 0 && f.nen == 4 && f.fq <= PressureTile<4>::kMaxQ;
    FaceArgs none = f;
    none.nFaces = 0;
    if (a.mixed)
        launchPatchesT<float>(a, pa, fused ? f : none, s);
    else
        launchPatchesT<double>(a, pa, fused ? f : none, s);
    if (f.nFaces > 0 && !fused) launchPressure(f, s);
}

void launchElementForceFaces(const ElemArgs& a, const FaceArgs& f, hipStream_t s)
{
    // one launch for the C3D4 kernel and faces whose Gauss points fit the tile
    if (a.nen == 4 && f.nen == 4 && f.nFaces > 0 && f.fq <= PressureTile<4>::kMaxQ) {
        if (a.mixed)
            launchTet4Faces<float>(a, f, s);
        else
            launchTet4Faces<double>(a, f, s);
        return;
    }
    launchElementForce(a, s);
    launchPressure(f, s);
}

void launchElementTangent(const ElemArgs& a, const double* v, hipStream_t s)
{
    if (a.mixed)
        launchElem<float, MODE_TANGENT>(a, v, s);
    else
        launchElem<double, MODE_TANGENT>(a, v, s);
}

template <int MODE>
void launchScalar(const ElemArgs& a, const double* lam, double* out, hipStream_t s)
{
    if (a.nElems <= 0) return;
    const int g = blocks(a.nElems);
    switch (a.nen) {
        case 4: kElemScalar<4, MODE><<<g, kThreads, 0, s>>>(a, lam, out); break;
        case 8: kElemScalar<8, MODE><<<g, kThreads, 0, s>>>(a, lam, out); break;
        case 10: kElemScalar<10, MODE><<<g, kThreads, 0, s>>>(a, lam, out); break;
        case 20: kElemScalar<20, MODE><<<g, kThreads, 0, s>>>(a, lam, out); break;
    }
}

void launchElementEnergy(const ElemArgs& a, double* out, hipStream_t s) { launchScalar<0>(a, nullptr, out, s); }

void launchElementParamGrad(const ElemArgs& a, const double* lam, double* out, hipStream_t s)
{
    launchScalar<1>(a, lam, out, s);
}

template <int MODE>
void launchPressureMode(const FaceArgs& a, hipStream_t s)
{
    if (a.nFaces <= 0) return;
    const int g = blocks(a.nFaces);
    switch (a.nen) {
        case 4: kPressure<4, MODE><<<g, kThreads, 0, s>>>(a); break;
        case 8: kPressure<8, MODE><<<g, kThreads, 0, s>>>(a); break;
        case 10: kPressure<10, MODE><<<g, kThreads, 0, s>>>(a); break;
        case 20: kPressure<20, MODE><<<g, kThreads, 0, s>>>(a); break;
    }
}

void launchPressure(const FaceArgs& a, hipStream_t s)
{
    if (a.nFaces <= 0) return;
    if (a.fq > PressureTile<4>::kMaxQ) return launchPressureMode<PMODE_FORCE>(a, s);
    const int g = (a.nFaces + PressureTile<4>::fpb - 1) / PressureTile<4>::fpb;
    switch (a.nen) {
        case 4: kPressureTile<4><<<g, kThreads, 0, s>>>(a); break;
        case 8: kPressureTile<8><<<g, kThreads, 0, s>>>(a); break;
        case 10: kPressureTile<10><<<g, kThreads, 0, s>>>(a); break;
        case 20: kPressureTile<20><<<g, kThreads, 0, s>>>(a); break;
    }
}

void launchPressureTangent(const FaceArgs& a, int transpose, hipStream_t s)
{
    if (transpose)
        launchPressureMode<PMODE_TRANSPOSE>(a, s);
    else
        launchPressureMode<PMODE_TANGENT>(a, s);
}

void launchSumComp(const double* a, int n, double* work, double* out, hipStream_t s)
{
    const int g = reduceGrid(n);
    kSumComp<<<g, kRB, 0, s>>>(a, n, work);
    kSumCompFinish<<<1, 64, 0, s>>>(work, g, out);
}

// ---- allocation-free entry points -----------------------------------------
namespace {

// The scratch of the allocation-free entry points.  In host code `scratch`
// only names the symbol: its device address comes from hipGetSymbolAddress
// (scratchInit) and only that address is passed to kernels.
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
double* g_scratchPtr[kMaxDevices] = {};

// dst[i] = src[i] for i < n, 0 for n <= i < n + 3 (the fixed-node slot)
__global__ void kCopyPad(double* __restrict__ dst, const double* __restrict__ src, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n + 3) dst[i] = i < n ? src[i] : 0.0;
}

}  // namespace

int scratchInit(int device)
{
    if (device < 0 || device >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchPtr[device]) return 0;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    g_scratchPtr[device] = static_cast<double*>(p);
    return 0;
}

double* scratchOf(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    return g_scratchPtr[device];
}

void launchEnergyScratch(ElemArgs a, const double* u, int n, double* sp, double* out, hipStream_t s)
{
    // layout: [x copy, n + 3 | per-element energies, E | sum partials, 2 x kReduceBlocks]
    auto up = [](size_t v) { return (v + 31) / 32 * 32; };
    double* xs = sp;
    double* per = xs + up(static_cast<size_t>(n) + 3);
    double* work = per + up(static_cast<size_t>(a.nElems));
    kCopyPad<<<blocks(n + 3), kThreads, 0, s>>>(xs, u, n);
    a.x = xs;
    launchScalar<0>(a, nullptr, per, s);
    launchSumComp(per, a.nElems, work, out, s);
}

void launchEnergyDiff(const double* Wn, const double* Wo, int E, const double* fext, const double* x,
                      const double* xp, int n, double* work, double* out, hipStream_t s)
{
    const int g = reduceGrid(max(E, n));
    kEnergyDiff<<<g, kRB, 0, s>>>(Wn, Wo, E, fext, x, xp, n, work);
    kEnergyDiffFinish<<<1, 64, 0, s>>>(work, g, out);
}

void launchLaplacian(const LapArgs& a, hipStream_t s)
{
    const long long lanes = static_cast<long long>(a.E) * a.nen;
    const int g = static_cast<int>((lanes + kThreads - 1) / kThreads);
    if (lanes > 0) {
        switch (a.nen) {
            case 4: kLapElem<4><<<g, kThreads, 0, s>>>(a); break;
            case 8: kLapElem<8><<<g, kThreads, 0, s>>>(a); break;
            case 10: kLapElem<10><<<g, kThreads, 0, s>>>(a); break;
            default: kLapElem<20><<<g, kThreads, 0, s>>>(a); break;
        }
    }
    if (a.nnz > 0) kLapSum<<<(a.nnz + kThreads - 1) / kThreads, kThreads, 0, s>>>(a);
}

void launchAxpby(double* out, double a, const double* x, double b, const double* y, int n, hipStream_t s)
{
    if (n > 0) kAxpby<<<blocks(n), kThreads, 0, s>>>(out, a, x, b, y, n);
}

void launchBatchSolve(const BatchArgs& a, int count, hipStream_t s)
{
    switch (tuning().batchThreads) {
        case 1024: launchBatchSolveW1024(a, count, s); break;
        case 512: launchBatchSolveW512(a, count, s); break;
        default: launchBatchSolveW256(a, count, s);
    }
}

size_t gatherDotWork(int nFree) { return 4 * (static_cast<size_t>(nFree) * 8 / kThreads + 2); }

template <int DOT>
void gatherLaunch(const GatherArgs& a, int lpn, int g, hipStream_t s)
{
    switch (lpn) {
        case 1: a.mixed ? kGather<float, 1, DOT><<<g, kThreads, 0, s>>>(a) : kGather<double, 1, DOT><<<g, kThreads, 0, s>>>(a); break;
        case 2: a.mixed ? kGather<float, 2, DOT><<<g, kThreads, 0, s>>>(a) : kGather<double, 2, DOT><<<g, kThreads, 0, s>>>(a); break;
        case 8: a.mixed ? kGather<float, 8, DOT><<<g, kThreads, 0, s>>>(a) : kGather<double, 8, DOT><<<g, kThreads, 0, s>>>(a); break;
        default: a.mixed ? kGather<float, 4, DOT><<<g, kThreads, 0, s>>>(a) : kGather<double, 4, DOT><<<g, kThreads, 0, s>>>(a);
    }
}

void launchGather(const GatherArgs& a, hipStream_t s)
{
    if (a.nFree <= 0) return;
    const int lpn = a.lanes ? a.lanes : tuning().gatherLanes;
    const int g = static_cast<int>((static_cast<long long>(lpn) * a.nFree + kThreads - 1) / kThreads);
    if (!a.dotWork) return gatherLaunch<0>(a, lpn, g, s);
    if (a.defer && a.dotCarry) {
        // the finishing reduction runs later (see TrialFinish)
        if (a.dotXt) gatherLaunch<4>(a, lpn, g, s);
        else gatherLaunch<3>(a, lpn, g, s);
        *a.defer = TrialFinish{a.dotWork, g, a.dotXt ? 4 : 3, a.dotOut, a.dotCarry, nullptr};
        return;
    }
    if (a.dotXt) {
        gatherLaunch<4>(a, lpn, g, s);
        if (a.dotCarry)
            kReduceFinish<4, 0x6u><<<1, kRF, 0, s>>>(a.dotWork, g, EpiStoreCarry<4>{{}, a.dotOut, a.dotCarry}, g);
        else
            kReduceFinish<4, 0x6u><<<1, kRF, 0, s>>>(a.dotWork, g, EpiStore<4>{{}, a.dotOut}, g);
    } else {
        gatherLaunch<3>(a, lpn, g, s);
        if (a.dotCarry)
            kReduceFinish<3, 0x6u><<<1, kRF, 0, s>>>(a.dotWork, g, EpiStoreCarry<3>{{}, a.dotOut, a.dotCarry}, g);
        else
            kReduceFinish<3, 0x6u><<<1, kRF, 0, s>>>(a.dotWork, g, EpiStore<3>{{}, a.dotOut}, g);
    }
}

void launchTrialFinish(const TrialFinish& tf, hipStream_t s)
{
    if (tf.k == 4)
        kReduceFinish<4, 0x6u><<<1, kRF, 0, s>>>(tf.work, tf.nb, EpiStoreCarry<4>{{}, tf.out, tf.carry}, tf.nb);
    else if (tf.k == 3)
        kReduceFinish<3, 0x6u><<<1, kRF, 0, s>>>(tf.work, tf.nb, EpiStoreCarry<3>{{}, tf.out, tf.carry}, tf.nb);
    if (tf.k && tf.ev) (void)hipEventRecord(tf.ev, s);
}

void launchDotMax(const double* a, const double* b, const double* c, int n, double* work, double* out,
                  hipStream_t s, const double* xt, const double* cb)
{
    const int g = reduceGrid(n);
    if (!c) {
        kDotMax<false, false><<<g, kRB, 0, s>>>(a, b, c, nullptr, nullptr, n, work);
        kReduceFinish<3, 0x6u><<<1, kRF, 0, s>>>(work, g, EpiStore<3>{{}, out});
    } else if (!xt) {
        kDotMax<true, false><<<g, kRB, 0, s>>>(a, b, c, cb, nullptr, n, work);
        kReduceFinish<3, 0x6u><<<1, kRF, 0, s>>>(work, g, EpiStore<3>{{}, out});
    } else {
        kDotMax<true, true><<<g, kRB, 0, s>>>(a, b, c, cb, xt, n, work);
        kReduceFinish<4, 0x6u><<<1, kRF, 0, s>>>(work, g, EpiStore<4>{{}, out});
    }
}

void launchDot2(const double* a, const double* b, int n, double* work, double* ab, double* bb, hipStream_t s)
{
    const int g = reduceGrid(n);
    kDot2<<<g, kRB, 0, s>>>(a, b, n, work);
    kReduceFinish<2, 0u><<<1, kRF, 0, s>>>(work, g, EpiStore2{{}, ab, bb});
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
                       const double* aj, double* cout, int mode, int n, double* work, hipStream_t s)
{
    const int g = reduceGrid(n);
    kTwoLoop<<<g, kRB, 0, s>>>(q, vprev, cprev, mode ? 1.0 : -1.0, w, n, work);
    kReduceFinish<1, 0u><<<1, kRF, 0, s>>>(work, g, EpiTwoLoopStep{{}, den, aj, cout, mode});
}

// calls f(std::integral_constant<int, nc>) for 1 <= nc <= 16
template <typename F>
void withPairCount(int nc, F&& f)
{
    switch (nc) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 9: f(std::integral_constant<int, 9>{}); break;
    case 10: f(std::integral_constant<int, 10>{}); break;
    case 11: f(std::integral_constant<int, 11>{}); break;
    case 12: f(std::integral_constant<int, 12>{}); break;
    case 13: f(std::integral_constant<int, 13>{}); break;
    case 14: f(std::integral_constant<int, 14>{}); break;
    case 15: f(std::integral_constant<int, 15>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    default: break;
    }
}
static_assert(kMaxCompact == 16, "withPairCount covers 1..kMaxCompact");

// the history sweeps' paired form: n even and every stream 16-byte aligned
// (then every history slot, n doubles apart, is too)
bool pairedStreams(int n, std::initializer_list<const void*> ptrs)
{
    if (n & 1) return false;
    for (const void* p : ptrs)
        if (reinterpret_cast<uintptr_t>(p) & 15u) return false;
    return true;
}

void launchGramFirst(double* S, double* Y, const double* x, const double* xp, const double* r, const double* rp,
                     int n, int nc, int loc, int mcap, double* G, double* ys, double* yy, double* theta,
                     double* alpha, double* rs, double* out, double* work, hipStream_t s, const int* pinv,
                     const TrialFinish* tf)
{
    withPairCount(nc, [&](auto ncc) {
        constexpr int NC = decltype(ncc)::value;
        const int g = historyGrid(n);
        const bool ride = tf && tf->k;
        const int gl = g + (ride ? 1 : 0);
        const double* tw = ride ? tf->work : nullptr;
        const int tnb = ride ? tf->nb : 0, tk = ride ? tf->k : 0;
        double* to = ride ? tf->out : nullptr;
        const double* tc = ride ? tf->carry : nullptr;
        if (pairedStreams(n, {S, Y, x, xp, r, rp}))
            kGramFirst<NC, true><<<gl, kRB, 0, s>>>(S, Y, x, xp, r, rp, n, loc, mcap, work, g, tw, tnb, tk, to, tc);
        else
            kGramFirst<NC, false><<<gl, kRB, 0, s>>>(S, Y, x, xp, r, rp, n, loc, mcap, work, g, tw, tnb, tk, to, tc);
        if (ride && tf->ev) (void)hipEventRecord(tf->ev, s);
        if (mcap <= 8) {
            if constexpr (NC <= 8)
                kReduceFinish<3 * NC + 1, 0u><<<1, kRF, 0, s>>>(
                    work, g, EpiGramFirst<8, NC>{G, ys, yy, theta, alpha, rs, loc, mcap});
        } else {
            kReduceFinish<3 * NC + 1, 0u><<<1, kRF, 0, s>>>(
                work, g, EpiGramFirst<16, NC>{G, ys, yy, theta, alpha, rs, loc, mcap});
        }
        if (n > 0)
            kMultiAxpy<NC><<<blocks(n), kThreads, 0, s>>>(HSrc{r}, -1.0, Y, alpha, -1.0, n, out, nullptr, nullptr,
                                                           pinv);
    });
}

void launchTwoLoopSecond(HSrc h, const double* S, const double* Y, int n, int nc, int newest, int mcap,
                         const double* G, const double* ys, const double* alpha, double* coef, double* out,
                         double* work, hipStream_t s, const double* dw, const double* rs, double* slope,
                         const double* xbase, double* xout)
{
    withPairCount(nc, [&](auto ncc) {
        constexpr int NC = decltype(ncc)::value;
        const int g = historyGrid(n);
        if (pairedStreams(n, {Y, dw}))
            kTwoLoopDots<NC, true><<<g, kRB, 0, s>>>(h, Y, n, dw, work);
        else
            kTwoLoopDots<NC, false><<<g, kRB, 0, s>>>(h, Y, n, dw, work);
        if (mcap <= 8) {
            if constexpr (NC <= 8)
                kReduceFinish<NC + 1, 0u><<<1, kRF, 0, s>>>(
                    work, g, EpiTwoLoopDots<8, NC>{G, ys, alpha, rs, coef, slope, newest, mcap});
        } else {
            kReduceFinish<NC + 1, 0u><<<1, kRF, 0, s>>>(
                work, g, EpiTwoLoopDots<16, NC>{G, ys, alpha, rs, coef, slope, newest, mcap});
        }
        if (n > 0) kMultiAxpy<NC><<<blocks(n), kThreads, 0, s>>>(h, 1.0, S, coef, 1.0, n, out, xbase, xout);
    });
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

void launchPermute(const TrsvArgs& a, const double* b, double* bp, hipStream_t s)
{
    if (a.n > 0) kPermute<<<blocks(a.n), kThreads, 0, s>>>(a.perm, b, bp, a.n);
}

template <typename T, int WAVES>
void fwdTabLaunch(const TrsvArgs& a, const double* bp, const int4* items, int nitems, const int* tab, int K, int W,
                  hipStream_t s)
{
    constexpr int NT = 64 * WAVES;
    switch (K) {
        case 1: kTrsvFwdTab<T, WAVES, 1><<<nitems, NT, 0, s>>>(a, bp, items, tab, W); break;
        case 2: kTrsvFwdTab<T, WAVES, 2><<<nitems, NT, 0, s>>>(a, bp, items, tab, W); break;
        case 3: kTrsvFwdTab<T, WAVES, 3><<<nitems, NT, 0, s>>>(a, bp, items, tab, W); break;
        default: kTrsvFwdTab<T, WAVES, 4><<<nitems, NT, 0, s>>>(a, bp, items, tab, W); break;
    }
}

void launchTrsvForward(const TrsvArgs& a, const double* bp, const int4* items, int nitems, int wide, hipStream_t s,
                       const int* tab, int K, int W)
{
    if (nitems <= 0) return;
    if (tab && K >= 1 && K <= kFwdTableMaxK) {
        if (a.mixed) {
            if (wide) fwdTabLaunch<float, 16>(a, bp, items, nitems, tab, K, W, s);
            else fwdTabLaunch<float, 4>(a, bp, items, nitems, tab, K, W, s);
        } else {
            if (wide) fwdTabLaunch<double, 16>(a, bp, items, nitems, tab, K, W, s);
            else fwdTabLaunch<double, 4>(a, bp, items, nitems, tab, K, W, s);
        }
        return;
    }
    if (a.mixed) {
        if (wide) kTrsvFwd<float, 16><<<nitems, 1024, 0, s>>>(a, bp, items);
        else kTrsvFwd<float, 4><<<nitems, 256, 0, s>>>(a, bp, items);
    } else {
        if (wide) kTrsvFwd<double, 16><<<nitems, 1024, 0, s>>>(a, bp, items);
        else kTrsvFwd<double, 4><<<nitems, 256, 0, s>>>(a, bp, items);
    }
}

void launchTrsvTop(const TrsvArgs& a, const double* bp, const TopArgs& t, hipStream_t s)
{
    if (t.nT <= 0) return;
    if (t.waves == 8) {
        const int grid = (t.nT + 7) / 8;
        if (a.mixed)
            kTrsvTop<float, 8><<<grid, 512, 0, s>>>(a, bp, t);
        else
            kTrsvTop<double, 8><<<grid, 512, 0, s>>>(a, bp, t);
        return;
    }
    const int grid = (t.nT + 15) / 16;
    if (a.mixed)
        kTrsvTop<float, 16><<<grid, 1024, 0, s>>>(a, bp, t);
    else
        kTrsvTop<double, 16><<<grid, 1024, 0, s>>>(a, bp, t);
}

template <typename T>
void bwdStageLaunch(const TrsvArgs& a, const int4* items, int nitems, int W, hipStream_t s)
{
    switch (W) {
        case 4: kTrsvBwdStage<T, 4><<<nitems, 256, 0, s>>>(a, items); break;
        case 8: kTrsvBwdStage<T, 8><<<nitems, 512, 0, s>>>(a, items); break;
        default: kTrsvBwdStage<T, 16><<<nitems, 1024, 0, s>>>(a, items); break;
    }
}

void launchTrsvBackward(const TrsvArgs& a, const int4* items, int nitems, int wide, hipStream_t s)
{
    if (nitems <= 0) return;
    if (wide >= kBwdStage4) {
        const int W = bwdStageCols(wide);
        if (a.mixed) bwdStageLaunch<float>(a, items, nitems, W, s);
        else bwdStageLaunch<double>(a, items, nitems, W, s);
        return;
    }
    if (wide == kBwdNarrow) {
        if (a.mixed)
            kTrsvBwdNarrow<float><<<nitems, 256, 0, s>>>(a, items);
        else
            kTrsvBwdNarrow<double><<<nitems, 256, 0, s>>>(a, items);
        return;
    }
    if (a.mixed) {
        if (wide) kTrsvBwd<float, 16><<<nitems, 1024, 0, s>>>(a, items);
        else kTrsvBwd<float, 4><<<nitems, 256, 0, s>>>(a, items);
    } else {
        if (wide) kTrsvBwd<double, 16><<<nitems, 1024, 0, s>>>(a, items);
        else kTrsvBwd<double, 4><<<nitems, 256, 0, s>>>(a, items);
    }
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
