// Device kernels of the solve loop (gfx950, wave64).  Launch wrappers only;
// the kernels live in kernels.hip.
#pragma once
#include "../limits.hpp"
#include <hip/hip_runtime.h>

namespace efem {
namespace dev {

// Launch shapes chosen per context by the autotuner (efem_autotune); the API
// installs the context's choice for the calling thread before launching.
struct LaunchTuning {
    int gatherLanes = 4;     // lanes per free node in kGather: 1, 2, 4, 8
    int reduceBlocks = 256;  // maximum blocks of a reduction: 32 .. kReduceBlocks
    int tet4Threads = 256;   // block size of kElemTet4: 64, 128, 256
    int historyBlocks = 256; // maximum blocks of the reductions over the
                             // L-BFGS history (Gram row + first loop, second
                             // loop): 32 .. kReduceBlocks
    int batchThreads = 256;  // lanes per instance of the persistent batched
                             // solver: 256, 512, 1024
};
void setTuning(const LaunchTuning& t);
const LaunchTuning& tuning();

// Device layout of the factor's block inverse: supernode s holds its m x w
// block column-major with leading dimension ldOf(m) = m rounded up to
// kLdAlign rows (zero rows in the padding) at a multiple-of-kLdAlign offset:
// every column starts on a 128-byte line in FP32 (256 bytes in FP64), so a
// forward item's 64-row segment of a column is two whole lines and a
// backward column sweep starts on a line
constexpr int kLdAlign = kBlockAlign;
__host__ __device__ constexpr int ldOf(int m) { return blockLd(m); }

// element internal force into per-(element, slot) buffers
struct ElemArgs {
    int nElems, nen, ng, law, paramsPerElem, mixed;
    const int* connDof;     // [nen][E]
    const void* ginv;       // [ng][9][E] (double or float)
    const double* vol;      // [E]
    const void* gdN;        // [ng][nen][3] (double or float)
    const double* gw;       // [ng]
    const double* params;   // [4] or [E][4]
    const double* x;        // displacement, free DOFs + 3 trailing zeros
    void* slots;            // [E][nen][3] (double or float)
    // C3D4 only (may be null): the reference coordinates per free DOF
    // (node of free id i at Xd[3 i]); an element whose four nodes are free
    // forms its J0^-1 from them (bit for bit the stored one), the others
    // read ginv
    const double* Xd = nullptr;
    int nFree = 0;
};
void launchElementForce(const ElemArgs& a, hipStream_t s);
// tangent product of the element forces: slots <- (d f_e / d u)[v] (the
// material tangent through forward-mode dual numbers)
void launchElementTangent(const ElemArgs& a, const double* v, hipStream_t s);
// FP64 per-element scalars (ginv/gdN must be the FP64 copies):
//   energy     out[e] = sum_q w_q V_e W(F_q)
//   paramGrad  out[4e + k] = lam . d f_e / d p_k
void launchElementEnergy(const ElemArgs& a, double* out, hipStream_t s);
void launchElementParamGrad(const ElemArgs& a, const double* lam, double* out, hipStream_t s);

struct FaceArgs {
    int nFaces, nen, fq, directed;
    double pressure, pdir[3];
    const int* faceDof;     // [nen][F] free id or nFree
    const double* A0;       // [F][3]
    const double* WN;       // [F][fq][nen]
    const double* G;        // [F][fq][nen][3]
    const double* x;
    const double* v;        // direction (tangent) or multiplier (transpose)
    const double* fpress;   // per-face pressure (batches) or null: `pressure`
    double* slots;          // [F][nen][3]
};
void launchPressure(const FaceArgs& a, hipStream_t s);
// the element forces and the pressure faces (f.nFaces > 0): one launch for
// C3D4 elements (face tiles after the element blocks), else the two
// launches; the slots are those of the two kernels
void launchElementForceFaces(const ElemArgs& a, const FaceArgs& f, hipStream_t s);
// load stiffness of the follower pressure applied to a.v, or its transpose
void launchPressureTangent(const FaceArgs& a, int transpose, hipStream_t s);

// C3D4 element forces assembled per patch in LDS (host: Tet4Patches): one
// block of kPatch lanes per patch, a lane per element; after a barrier
// each record (a free node the patch touches) sums its entries in FP64, in
// entry order, into its node-major position: part[3 pos + c].  The
// pressure faces (fa.nFaces > 0) follow in the same launch.  A gather with
// eslots = part (double), rptr and face-only lists then forms r.
constexpr int kPatch = 256;
struct PatchArgs {
    int E = 0, nPatch = 0;
    const int* conn = nullptr;        // [4][E] free ids, patch order
    const int* order = nullptr;       // [E] element of each patch position
    const double* vol = nullptr;      // [E] patch order
    const int* roff = nullptr;        // [nPatch + 1]
    const int* rec = nullptr;         // [nRec][4] first entry, count, position, 0
    const uint16_t* ent = nullptr;    // record entries 4 lane + a
    double* part = nullptr;           // [nRec][3], node-major
};
void launchElementPatches(const ElemArgs& ea, const PatchArgs& pa, const FaceArgs& fa, hipStream_t s);

struct GatherArgs {
    int nFree, nElemSlots, mixed;
    int lanes = 0;          // lanes per node (0: the tuned value)
    // optional: node i's element terms are the contiguous FP64 records
    // [rptr[i], rptr[i + 1]) of eslots; gptr / gslot then list face slots
    // only (nElemSlots = 0)
    const int* rptr = nullptr;
    const int* gptr;
    const int* gslot;
    const void* eslots;     // element slots (double or float)
    const double* fslots;   // face slots
    const double* fext;     // constant nodal forces (may be null)
    double ce, cf;          // r = ce sum(element slots) + cf sum(face slots) - fext
    double* r;
    // optional (dotWork non-null): the line search's scalars of the new r in
    // the same launch -- dotOut[0] = r . dotD, [1] = max|r|, [2] = max|dotX -
    // dotCb| (dotCb may be null), [3] = |dotX - dotXt|^2 (dotXt non-null);
    // dotWork holds 4 partials per gather block (see gatherDotWork)
    const double* dotD = nullptr;
    const double* dotX = nullptr;
    const double* dotCb = nullptr;
    const double* dotXt = nullptr;
    double* dotWork = nullptr;
    double* dotOut = nullptr;
    // optional: dotOut[4] = *dotCarry as well (dotOut then pinned host
    // memory: the host reads the trial's scalars without a copy)
    const double* dotCarry = nullptr;
    // optional (with dotCarry): the finishing reduction is not launched but
    // described here, to run as one extra block of the next history sweep
    // (launchGramFirst) or on its own (launchTrialFinish)
    struct TrialFinish* defer = nullptr;
};

// A trial's finishing reduction of the gather's line-search partials (k = 3
// or 4 components at work[c nb + b]) into pinned memory out[0..k), with
// out[4] = *carry; ev (may be null) is recorded behind the launch that runs
// it.  k = 0: nothing pending.
struct TrialFinish {
    const double* work = nullptr;
    int nb = 0, k = 0;
    double* out = nullptr;
    const double* carry = nullptr;
    hipEvent_t ev = nullptr;
};
void launchTrialFinish(const TrialFinish& tf, hipStream_t s);
void launchGather(const GatherArgs& a, hipStream_t s);
// doubles of dotWork a gather over nFree nodes needs (any lanes per node)
size_t gatherDotWork(int nFree);

// ---- approximate Jacobian ----------------------------------------------------
// The scalar Laplacian L_ij = sum_e sum_q V_e c_e w_q grad N_i . grad N_j
// (ElementConstraint::conEvaluteJacobi, ElementConstraint.cpp:164-205) on the
// device: element matrices kel[e][a][b], then every CSR entry k sums its
// contributions kel[src[ptr[k] .. ptr[k+1])] in the listed order.  The
// arithmetic is that of the host assembly operation for operation (no
// contracted multiply-adds), so the values are the host's bit for bit.
struct LapArgs {
    int E, nen, ng, law, ppe, nnz;
    const double* ginv;    // [ng][9][E] FP64
    const double* vol;     // [E]
    const double* gdN;     // [ng][nen][3] FP64
    const double* gw;      // [ng]
    const double* params;  // [4] or [E][4]
    double* kel;           // [E][nen][nen]
    const int* ptr;        // [nnz + 1]
    const int* src;        // e nen^2 + a nen + b
    double* vals;          // [nnz]
};
void launchLaplacian(const LapArgs& a, hipStream_t s);

// ---- vectors ----------------------------------------------------------------
// deterministic reductions: results written to device scalars; the block
// partials live in `work` (see kReduceWorkDoubles); two launches each
constexpr int kReduceBlocks = 256;
constexpr int kMaxCompact = 16;  // history size handled by the compact two-loop
constexpr int kReduceWorkDoubles = (3 * kMaxCompact + 2) * kReduceBlocks + 64;
// out[0] = a.b, out[1] = max|a|, out[2] = max|c - cb| (c may be null, cb
// null means 0); with xt (and c), out[3] = |c - xt|^2
void launchDotMax(const double* a, const double* b, const double* c, int n, double* work,
                  double* out, hipStream_t s, const double* xt = nullptr, const double* cb = nullptr);
// *ab = a.b, *bb = b.b
void launchDot2(const double* a, const double* b, int n, double* work, double* ab, double* bb,
                hipStream_t s);
// *out = sum a (Neumaier lanes + TwoSum tree, reproducible)
void launchSumComp(const double* a, int n, double* work, double* out, hipStream_t s);
// energy line search: out[0] = sum_e (Wn[e] - Wo[e]) - (x - xp) . fext
// (compensated), out[1] = sum_e |Wo[e]| + |xp . fext| (rounding scale);
// fext may be null; work holds 4 kReduceBlocks doubles
void launchEnergyDiff(const double* Wn, const double* Wo, int E, const double* fext, const double* x,
                      const double* xp, int n, double* work, double* out, hipStream_t s);

// ---- allocation-free entry points -----------------------------------------
// One file-scope device array per GPU (kScratchDoubles doubles, 16 MiB) that
// the allocation-free variants use instead of context buffers.
constexpr size_t kScratchDoubles = size_t(2) << 20;
// the array's device address on `device`: looked up once per device (with
// that device current) by hipGetSymbolAddress, then cached
int scratchInit(int device);
double* scratchOf(int device);
// total strain energy into out (device pointer): u (n values) copied into
// the scratch with the fixed-node slot's three trailing zeros, per-element
// energies there, compensated sum; everything on stream s
void launchEnergyScratch(ElemArgs a, const double* u, int n, double* scratch, double* out, hipStream_t s);
// out = a x + b y (y may be null)
void launchAxpby(double* out, double a, const double* x, double b, const double* y, int n, hipStream_t s);
// x = xp + alpha d
void launchStep(double* x, const double* xp, const double* d, double alpha, int n, hipStream_t s);
// s = x - xp, y = r - rp
void launchDiff2(double* sv, double* yv, const double* x, const double* xp, const double* r,
                 const double* rp, int n, hipStream_t s);
// y = a x  (scalar from host)
void launchScale(double* y, const double* x, double a, int n, hipStream_t s);
// two-loop step: q -= (*cprev) * vprev (if vprev), then out = (w . q) ;
//   *cout = mode ? (*aj - out / *den) : out / *den
void launchTwoLoopStep(double* q, const double* vprev, const double* cprev, const double* w,
                       const double* den, const double* aj, double* cout, int mode, int n,
                       double* work, hipStream_t s);
// compact two-loop (same recursion as LBFGSMat::apply_Hv, evaluated from
// Gram products).  History slots 0..nc-1 of the ring S, Y ([m][n]), newest
// slot `newest`; G[i*16+k] = s_i . y_k; rho[i] = 1 / (s_i . y_i).  The new
// pair and loop 1 run in launchGramFirst (declared below).
// loop 2: beta_i = (y_i . r_i) / ys_i oldest -> newest with r_0 = h (the H0
// output), then out = h + sum (alpha_i - beta_i) s_i and, with rs_k =
// s_k . dw from launchGramFirst, *slope = dw . out, formed as
// dw . h + sum (alpha_k - beta_k) rs_k (dw required; 1 <= nc <= 16)
// xbase, xout (optional): also xout = xbase + out, the first line-search
// trial x + 1 d formed with the direction
// The H0 output h of the second loop: a vector in DOF order, or (pinv set)
// the backward sweep's solution z in the factor's order, read as
// h[i] = z[3 pinv[i / 3] + i % 3] (FP32 or FP64 z) -- the same values the
// unpermute launch would store, without that launch.
struct HSrc {
    const double* h = nullptr;
    const int* pinv = nullptr;
    const void* z = nullptr;
    int zF32 = 0;
};
void launchTwoLoopSecond(HSrc h, const double* S, const double* Y, int n, int nc, int newest,
                         int mcap, const double* G, const double* ys, const double* alpha, double* coef,
                         double* out, double* work, hipStream_t s, const double* dw, const double* rs,
                         double* slope, const double* xbase = nullptr, double* xout = nullptr);
// q += (*c) * v   (or q -= when neg)
void launchAxpyDev(double* q, const double* v, const double* c, int neg, int n, hipStream_t s);
// q /= (*theta)
void launchDivDev(double* q, const double* theta, int n, hipStream_t s);
// *theta = yy / ys
void launchTheta(const double* ys, const double* yy, double* theta, hipStream_t s);
// the new pair and the first loop in one reduction pass (compact two-loop):
// S[loc] = x - xp, Y[loc] = r - rp; Gram row / column of loc, ys[loc],
// *yy = y.y, *theta = yy / ys[loc], rs[k] = s_k . r; alpha newest ->
// oldest; then out = -r - sum alpha_i y_i
// (pinv: out written in the factor's node order, out[3 pinv[node] + c])
void launchGramFirst(double* S, double* Y, const double* x, const double* xp, const double* r, const double* rp,
                     int n, int nc, int loc, int mcap, double* G, double* ys, double* yy, double* theta,
                     double* alpha, double* rs, double* out, double* work, hipStream_t s,
                     const int* pinv = nullptr, const TrialFinish* tf = nullptr);

// ---- supernodal triangular solves (3 right-hand sides) ---------------------
struct TrsvArgs {
    int n, mixed;
    const int* perm;        // [n] new -> old
    const int* rows;        // row indices of every supernode (per-item offset)
    const void* B;          // double or float, per-supernode [L11^-1 ; L21 L11^-1]
    const int* fptr;        // [front rows + 1]: children's update rows that
    const int* fidx;        //   extend-add into each front row
    void* y;                // [n][3] work (T)
    void* t;                // [nt][3] update vectors u_s (T)
    void* z;                // [n][3] work (T)
};
// bp[3 j + c] = b[3 perm[j] + c] (once per apply, before the forward sweep)
void launchPermute(const TrsvArgs& a, const double* b, double* bp, hipStream_t s);
// one level of the forward solve; items are descriptor pairs (int4 x 2):
// {r0, w, m, j0}, {front row offset, update offset, B offset lo, hi}.
// Multifrontal form f_s = [bp_s; 0] + sum_children u_c (extend-add),
// y_s = L11^-1 f_s[:w], u_s = f_s[w:] - (L21 L11^-1) f_s[:w]
// wide = 1: 1024-lane blocks whose 16 waves split the supernode's columns
// (for levels of wide supernodes), else 256 lanes / 4 waves
// tab (may be null): the level's extend-add table, K (1..kFwdTableMaxK)
// update-vector positions per front row: for item i at tab + i K (64 + W),
// K x 64 for its 64 rows, then K x W for the supernode's own columns
// (k-major; absent entries name the zero slot after the last update row);
// the sums are those of fptr / fidx, in the same order
constexpr int kFwdTableMaxK = 4;
void launchTrsvForward(const TrsvArgs& a, const double* bp, const int4* items, int nitems, int wide,
                       hipStream_t s, const int* tab = nullptr, int K = 0, int W = 0);
// one level of the backward solve; descriptors {k0, w, m, j0}, {row offset,
// 0, B offset lo, hi}: four columns per item (wide = 1: four waves per
// column split its rows, for levels of tall supernodes; wide = kBwdNarrow:
// sixteen columns per item, k0 = 0 mod 16, for levels of many short ones)
constexpr int kBwdNarrow = 2;
// wide = kBwdStage4 / 8 / 16: staged items of 4, 8 or 16 columns (k0 a
// multiple of that), one wave per column, the right-hand side [y_s;
// -z[rows]] gathered into LDS once per item
constexpr int kBwdStage4 = 3, kBwdStage8 = 4, kBwdStage16 = 5;
__host__ __device__ constexpr int bwdStageCols(int wide) { return wide == kBwdStage4 ? 4 : wide == kBwdStage8 ? 8 : 16; }
// The fused top of the elimination tree.  T = the supernodes within depth D
// of the roots (closed under ancestors), nT columns numbered T-locally by
// depth, then factor column.  With L_TT the factor's TT block, the TT block
// of (P A P^T)^-1 is Z = (L_TT L_TT^T)^-1 (dense, symmetric, column-major,
// ld ldz, zero padding rows), and T's forward and backward solves are one
// product z_T = Z f_T, f_T = the permuted right-hand side on T plus the
// update rows of T's children outside T (their extend-add):
//   z[tcol[k]] = sum_i Z[i, k] f[i],
//   f[i] = bp[tcol[i]] + sum_{j < K} u[tab[j nT + i]]   (j in order; absent
//   entries name the zero slot after the last update row)
// One wave per column k, `waves` (8 or 16) columns per block.
constexpr int kTopMaxK = 4;   // contributions per T row the table holds
struct TopArgs {
    int nT, ldz, K, waves;
    const void* Z;            // FP32 or FP64, as the factor
    const int* tcol;          // [nT]
    const int* tab;           // [K][nT]
};
void launchTrsvTop(const TrsvArgs& a, const double* bp, const TopArgs& t, hipStream_t s);
// Formation of Z at each factorisation (factor.hip, FP64), from the FP64
// block inverses B_s = [L11^-1; L21 L11^-1] of T's supernodes, depth by
// depth from the roots (selected inversion: with R_s the rows of s below
// its columns, Z[:, s] = -Z[:, R_s] B21_s on the columns already formed
// and Z_ss = B11^T B11 - B21^T Z[R_s, s]).
struct TopSuper {
    int cs, w, m, ld;         // first T-local column, width, rows, block leading dimension
    long long boff;           // block offset in the source
    int rOff, ldz;            // rows below the columns as T-local indices: rtl[rOff .. rOff + m - w);
                              // Z's leading dimension
    long long zoff;           // Z of this supernode's tree at TopBuild::Z + zoff (0: one Z)
};
// one tile of a mirror step of several trees: Z[j, i] = Z[i, j] for rows i
// < p0 and columns p0 <= j < p1 of the tree's Z at zoff (ld ldz)
struct TopMirrorTask {
    long long zoff;
    int ldz, p0, p1, i0, j0, pad;
};
void launchTopMirrorTasks(double* Z, const TopMirrorTask* tasks, int count, hipStream_t s);
constexpr int kTopTile = 64;  // the top kernels' tile
struct TopBuild {
    double* Z;
    int nT, ldz;
    const double* B;          // source blocks
    const TopSuper* sn;
    const int* rtl;
    const int4* tasks;        // tile tasks, see the launchers
};
// tasks {supernode, row tile start, column tile start, row end}:
// Z[r, cs + j] = -sum_k Z[r, rtl[rOff + k]] B21[k, j] for the tile's rows < row end
void launchTopGemm(const TopBuild& b, int first, int count, hipStream_t s);
// Z[j, i] = Z[i, j] for i < p0 <= j < p1
void launchTopMirror(const TopBuild& b, int p0, int p1, hipStream_t s);
// tasks {supernode, i tile start, j tile start, 0}: the diagonal block
// Z[cs + i, cs + j] = sum_{r < m} B[r, i] V[r, j], V = [B11; -Z[R_s, s]]
void launchTopDiag(const TopBuild& b, int first, int count, hipStream_t s);
// device numeric factorisation (factor.hip): runs the plan's launch list
// (kind, first task, count) on stream s; the L, X and U arenas must be zero
// on entry; a pivot that is not positive sets *fail
struct DevFactorArgs {
    double *L, *X, *U;
    const double* A;
    const int *w, *m;
    const long long *dval, *uOff;
    const int *rel, *relOff, *pairs;
    const long long* asmDst;
    const int* asmSrc;
    const int4* tasks;
    int* fail;
    const int* launches;  // host array, 3 ints per launch
    int nlaunch;
};
void launchDeviceFactor(const DevFactorArgs& d, hipStream_t s);
// B (float) = X (double), n values
void launchFactorToF32(const double* X, void* B, long long n, hipStream_t s);
void launchTrsvBackward(const TrsvArgs& a, const int4* items, int nitems, int wide, hipStream_t s);
// out[3 perm[j] + c] = z[j][c]
void launchUnpermute(const TrsvArgs& a, double* out, hipStream_t s);

// ---- persistent batched L-BFGS (batch.hip) ---------------------------------
// One workgroup per instance runs LBFGSSolver::solve (LBFGS.h:96-174) with
// the projection line search to the end on the instance's ranges of the
// merged model; every loop is bounded by maxIt / maxLs.
struct BatchRange {
    int e0, e1, f0, f1, n0, n1;  // elements, faces, free nodes of the instance
};
struct BatchArgs {
    ElemArgs elem;               // merged model (ginv/gdN FP64; slots sized FP64)
    FaceArgs face;               // x = the trial-point buffer `x` below
    GatherArgs gather;           // r = the residual buffer `r` below
    TrsvArgs trsv;               // merged factor forest (valid when usePrecond)
    const int4* items;           // descriptor pairs
    const int* instItems;        // [instance][level] {fwd off, cnt, bwd off, cnt}
    int nlev;
    const BatchRange* ranges;
    double *x, *r, *xp, *rp, *d, *q, *bp;  // merged vectors (+3 dummy zeros)
    double *S, *Y;               // history [m][ntot]
    int ntot, m, maxIt, maxLs, usePrecond, elemF32;
    double eps, epsRel, lnsFea, fp64Switch;
    // FP32 element slots of the batch: instances switch from FP32 to FP64
    // elements one by one, so the two precisions keep separate slot
    // buffers (elem.slots / gather.eslots are the FP64 ones)
    void* slots32;
    const double* xt;            // merged displacement targets, or null
    double targetTol;            // stop an instance at ||x - xt|| <= tol ||xt||
    double* stats;               // [instance][kBatchStats]
    // dense preconditioner (zArena non-null): per instance Z_i = (L_i
    // L_i^T)^-1 over its whole tree (TF, column-major, ld zld[i], at
    // zoff[i]); ztcol[n0 + t] = the merged factor column of the instance's
    // local index t (n_i = n1 - n0 <= kBatchDenseMaxN)
    const void* zArena = nullptr;
    const long long* zoff = nullptr;
    const int* zld = nullptr;
    const int* ztcol = nullptr;
};
// instances up to this many free nodes take the dense preconditioner
constexpr int kBatchDenseMaxN = 1024;
// stats per instance: iterations, r_evals, ls_trials, ||r||_inf, ||x||_inf,
// non-finite flag, FP64-switch iteration, preconditioned flag, target
// reached, last ||x - xt|| / ||xt||
constexpr int kBatchStats = 12;
// one persistent launch, one workgroup of `width` lanes per instance
// (tuning().batchThreads: 256, 512 or 1024)
void launchBatchSolve(const BatchArgs& a, int count, hipStream_t s);
void launchBatchSolveW256(const BatchArgs& a, int count, hipStream_t s);
void launchBatchSolveW512(const BatchArgs& a, int count, hipStream_t s);
void launchBatchSolveW1024(const BatchArgs& a, int count, hipStream_t s);

}  // namespace dev
}  // namespace efem
