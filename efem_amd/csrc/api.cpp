// C ABI of the solve loop (include/efem_gpu.h): context, device buffers,
// residual evaluation, host factorisation + device preconditioner, and the
// L-BFGS iteration with the reference's projection line search.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/efem_gpu.h"
#include "device/kernels.hpp"
#include "limits.hpp"
#include "material_math.hpp"
#include "host/model.hpp"
#include "host/sparse.hpp"
#include "host/dense.hpp"
#include "host/devplan.hpp"
#include "host/vtk.hpp"

namespace {

thread_local std::string g_error;

struct DeviceFactor {
    bool ready = false;
    int n = 0, nsuper = 0, nlevels = 0;
    long long nnz = 0;
    int* perm = nullptr;
    int* pinv = nullptr;    // [n] old -> new
    int* rows = nullptr;
    void* B = nullptr;
    int *fptr = nullptr, *fidx = nullptr;
    void *y = nullptr, *t = nullptr, *z = nullptr;
    int4* items = nullptr;  // descriptor pairs
    int* instItems = nullptr;  // batches: [instance][level] {fwd off, cnt, bwd off, cnt}
    int instLevels = 0;
    // batches: the dense Z_i of every instance (see BatchArgs::zArena), or null
    void* batchZ = nullptr;
    long long* batchZoff = nullptr;
    int* batchZld = nullptr;
    int* batchZtcol = nullptr;
    std::vector<int> fwdOff, fwdCnt, bwdOff, bwdCnt;  // per level, in items
    std::vector<int> fwdWide, bwdWide;               // per level: 16-wave items
    // per level: forward extend-add table (K = 0: none) at fwdTab + fwdTabOff
    int* fwdTab = nullptr;
    std::vector<long long> fwdTabOff;
    std::vector<int> fwdK, fwdW;
    // fused top (single-problem factors): T = the supernodes within depth
    // topDepth of the roots, solved as z_T = Z f_T in one launch in place of
    // their forward and backward items; Z = (L_TT L_TT^T)^-1 (topN x topN,
    // ld topLd, FP32 or FP64 as the factor), topCol the factor column of
    // each T-local index, topTab the K x topN update positions of f_T
    bool fusedTop = false;
    int topDepth = -1, topN = 0, topLd = 0, topK = 0;
    long long sweepEntries = 0;  // stored block entries of the supernodes outside T
    void* Z = nullptr;
    int* topCol = nullptr;
    int* topTab = nullptr;
    // the whole apply captured as one HIP graph over fixed staging buffers
    double *in = nullptr, *out = nullptr, *bp = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // device buffers owned by this factor: freed when the factor is
    // replaced or released
    std::vector<void*> allocs;
};

}  // namespace

// columns the fused top of the elimination tree may hold (default of the
// context option "fused_top_max"): one FP32 pass of the top kernel, a Z of
// at most 51 MB in FP32
constexpr int kFusedTopMaxCols = 3584;
// levels with at least this many 4-column backward items use 16-column
// items (default of the context option "bwd_narrow_items")
constexpr int kNarrowMinBwdItems = 4096;
// levels of wide (tall) supernodes with at most this many forward (backward)
// items use 16-wave items (defaults of "wide_fwd_items" / "wide_bwd_items")
constexpr int kWideMaxFwdItems = 500;
constexpr int kWideMaxBwdItems = 1000;

struct efem_ctx {
    std::string err;
    efem::Model model;
    int precision = EFEM_FP64;
    int device = -1;  // the HIP device current at efem_create
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    std::vector<void*>* allocList = &allocs;  // where devAlloc records
    // pinned staging of the factor's block-inverse values for the upload
    void* pinned = nullptr;
    size_t pinnedBytes = 0;
    // element / face / gather data
    int* connDof = nullptr;
    void* ginv = nullptr;
    double* vol = nullptr;
    void* gdN = nullptr;
    void* ginv64 = nullptr;  // FP64 copies used by EFEM_MIXED after its switch
    void* gdN64 = nullptr;
    bool elemFp64 = false;   // current element-kernel precision in EFEM_MIXED
    double* gw = nullptr;
    double* params = nullptr;
    void* eslots = nullptr;
    double* fslots = nullptr;
    int* faceDof = nullptr;
    double *fA0 = nullptr, *fWN = nullptr, *fG = nullptr;
    double* fpress = nullptr;  // per-face pressure of a batch (null: model.pressure)
    // batched instances: per-instance models (factored one by one) and ranges
    std::vector<efem::Model> parts;
    std::vector<efem::PartRange> ranges;
    efem::dev::BatchRange* dRanges = nullptr;
    double* batchStats = nullptr;
    float* batchSlots32 = nullptr;  // FP32 element slots of a batch (see BatchArgs)
    int *gptr = nullptr, *gslot = nullptr;
    double* fext = nullptr;
    // solver state
    int mcap = 0;
    double *x = nullptr, *r = nullptr, *xp = nullptr, *rp = nullptr, *d = nullptr, *tmp = nullptr;
    double *x3 = nullptr, *r3 = nullptr, *d2 = nullptr;  // spares of the solve loop
    double* gdotWork = nullptr;  // partials of the line-search scalars fused into the gather
    hipEvent_t trialEv[2] = {nullptr, nullptr};          // trial scalars copied
    double *S = nullptr, *Y = nullptr;
    // per-element strain energies of the energy line search: the accepted
    // iterate's and the trial's (allocated by the first such solve)
    double *wOld = nullptr, *wNew = nullptr;
    double* scal = nullptr;  // device scalars
    double* work = nullptr;
    double* hostScal = nullptr;  // pinned
    double* hostScalDev = nullptr;  // its device address (kernels write the trials' scalars there)
    DeviceFactor fac;
    bool factorFailed = false;
    double factorSeconds = 0.0;
    // last factorisation, seconds: {host model (assembly, ordering, symbolic,
    // numeric), of which symbolic, of which numeric, device descriptors and
    // maps, block repack, block upload, other device buffers, total}
    double factorPhases[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<double> trace;
    // profiling of the element kernel and of H0 applies
    bool profiling = false;
    std::vector<hipEvent_t> evs, evsH0;
    int evUsed = 0, evH0Used = 0;
    double* xtarget = nullptr;
    // committed configuration (efem_commit): device copy, or null
    double* base = nullptr;
    bool hasBase = false;
    // tangent / Krylov / adjoint work vectors (n + 3 each, allocated on first use)
    double *ku = nullptr, *kb = nullptr, *kx = nullptr, *kr = nullptr, *krh = nullptr, *kp = nullptr,
           *kv = nullptr, *ks = nullptr, *kt = nullptr, *kph = nullptr, *ksh = nullptr, *kxo = nullptr,
           *kbo = nullptr;
    efem::dev::LaunchTuning tuning;  // efem_autotune / options
    // option "h0_graph": replay the apply from a captured HIP graph.  Off by
    // default: on this part hipGraphLaunch holds the host about as long as
    // the 26 plain launches take, and the loop runs 2 % faster without it
    bool useGraph = false;
    bool compactTwoLoop = true;  // option "two_loop_compact" (history <= 16)
    bool lookahead = true;       // option "lookahead"
    bool fwdTables = true;       // option "fwd_tables"
    bool fusedRootOpt = true;    // option "fused_root": the fused top (below)
    // option "fused_top_max": the fused top takes the deepest depth whose
    // supernodes (with all shallower ones) hold at most this many columns
    int fusedTopMax = kFusedTopMaxCols;
    int topWaves = 16;           // option "top_waves": columns per block of the fused top, 8 or 16
    // option "tet4_geometry": C3D4 elements form J0^-1 from the nodes'
    // coordinates (Xd) instead of reading the stored inverse
    bool tet4Geometry = true;
    double* Xd = nullptr;
    // option "tet4_patches": the C3D4 residual sums each patch's element
    // forces in LDS and the gather reads one FP64 partial per (patch, node)
    // instead of one slot per (element, node); option "patch_gather_lanes"
    // sets the gather's lanes per node on that path
    bool tet4Patches = true;
    int patchLanes = 2;
    // option "batch_dense_h0": batches apply each instance's preconditioner
    // as one product with its dense Z_i = (L_i L_i^T)^-1 (instances of at
    // most kBatchDenseMaxN free nodes), formed on the device at each
    // factorisation; 0: the level-scheduled sweeps per instance
    bool batchDense = true;
    efem::dev::PatchArgs patch;  // device arrays (patch.part null: no plan)
    int *pRptr = nullptr, *pFptr = nullptr, *pFslot = nullptr;
    int pRec = 0;
    // option "bwd_stage": backward sweep items that stage [y_s; -z[rows]]
    // in LDS, 4, 8 or 16 columns each (0: the per-wave gathers), on every
    // level but those of many short supernodes (16-column narrow items)
    int bwdStage = 16;
    int residualTerms = 7;       // option "residual_terms" (efem_residual only)
    int narrowMinItems = kNarrowMinBwdItems;  // option "bwd_narrow_items"
    int denseSubtree = efem::kDefaultDenseSubtree;  // option "dense_subtree"
    // ordering + symbolic analysis of the approximate Jacobian: its pattern
    // is fixed by the mesh and the fixed nodes, so it is computed at the
    // first factorisation and reused by every refactorisation (option
    // "reuse_symbolic" = 0: recomputed every time)
    bool reuseSymbolic = true;
    std::unique_ptr<efem::SupernodalSymbolic> symbolic;
    // options "wide_fwd_items" / "wide_bwd_items": levels of wide (tall)
    // supernodes with at most this many items use 16-wave items
    int wideFwdItems = kWideMaxFwdItems;
    int wideBwdItems = kWideMaxBwdItems;
    // host factor memory reused by the next factorisation of this problem
    std::unique_ptr<efem::HostWorkspace> hostWs{new efem::HostWorkspace};
    // approximate Jacobian assembled on the device (option
    // "device_laplacian"): the host keeps its pattern, the device the
    // contribution map (built at the first factorisation)
    bool devLaplacian = true;
    // numeric factorisation on the device (option "device_factor"; needs the
    // device-assembled Laplacian): the plan (built once per pattern), its
    // device copy and the arenas, kept between refactorisations
    bool devFactorOpt = true;
    std::unique_ptr<efem::DevFactorPlan> dplan;
    int *dfW = nullptr, *dfM = nullptr, *dfRel = nullptr, *dfRelOff = nullptr, *dfPairs = nullptr,
        *dfAsmSrc = nullptr, *dfFail = nullptr;
    long long *dfDval = nullptr, *dfUoff = nullptr, *dfAsmDst = nullptr;
    int4* dfTasks = nullptr;
    double *dfL = nullptr, *dfX = nullptr, *dfU = nullptr;
    std::vector<int> dfLaunches;
    bool lapReady = false;
    efem::Csc lapA;
    int *lapPtr = nullptr, *lapSrc = nullptr;
    double *lapKel = nullptr, *lapVals = nullptr;
};

namespace {

// supernode shapes that switch a level to 16-wave triangular-solve items
constexpr int kWideColumns = 128;
constexpr int kTallRows = 256;

// energy line search: Armijo constant, and the relative size below which a
// change of the potential is rounding (then the projection rule decides)
constexpr double kArmijoC1 = 1e-4;
constexpr double kEnergyRtol = 1e-12;

// scalar slots in ctx->scal
constexpr int kOut = 0;       // [3] reduction outputs
constexpr int kYs = 8;        // [64]
constexpr int kYY = 72;       // [64]
constexpr int kAlpha = 136;   // [64]
constexpr int kCoef = 200;    // [64]
constexpr int kTheta = 264;   // [1]
constexpr int kGram = 272;    // [16][16] s_i . y_k of the compact two-loop
constexpr int kRS = 528;      // [16] s_k . r of the compact two-loop
constexpr int kNumScal = 544;
// pinned host scalars: [0, 16) readScalars, then two slots of 8 for the
// solve loop's trial scalars
constexpr int kTrialSlot = 16;
constexpr int kHostScal = kTrialSlot + 16;

#define HIP_CHECK(ctx, expr)                                                          \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);           \
            return EFEM_ERR_HIP;                                                      \
        }                                                                             \
    } while (0)

template <typename T>
int devAlloc(efem_ctx* c, T** p, size_t count)
{
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    HIP_CHECK(c, hipMalloc(&q, bytes));
    c->allocList->push_back(q);
    *p = static_cast<T*>(q);
    return EFEM_OK;
}

// frees a buffer devAlloc recorded in the context's own list (after the
// context stream has drained: queued kernels may still read it)
template <typename T>
int devFree(efem_ctx* c, T** p)
{
    if (!*p) return EFEM_OK;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    void* q = *p;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), q);
    if (it != c->allocs.end()) c->allocs.erase(it);
    HIP_CHECK(c, hipFree(q));
    *p = nullptr;
    return EFEM_OK;
}

// the L-BFGS history ring S / Y grown to hold m pairs (the old buffers are
// freed, not kept until efem_destroy)
int growHistory(efem_ctx* c, int m)
{
    if (m <= c->mcap) return EFEM_OK;
    const size_t n = static_cast<size_t>(std::max(c->model.ndof(), 1));
    int rc;
    if ((rc = devFree(c, &c->S)) || (rc = devFree(c, &c->Y))) return rc;
    c->mcap = 0;
    if ((rc = devAlloc(c, &c->S, static_cast<size_t>(m) * n)) || (rc = devAlloc(c, &c->Y, static_cast<size_t>(m) * n)))
        return rc;
    c->mcap = m;
    return EFEM_OK;
}

template <typename T>
int upload(efem_ctx* c, T** p, const std::vector<T>& h)
{
    int rc = devAlloc(c, p, h.size());
    if (rc) return rc;
    if (!h.empty()) HIP_CHECK(c, hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return EFEM_OK;
}

bool elem32(const efem_ctx* c) { return (c->precision & EFEM_F32_ELEMENTS) != 0; }
bool fac32(const efem_ctx* c) { return (c->precision & EFEM_F32_PRECOND) != 0; }

// double or float copy
int uploadReal(efem_ctx* c, void** p, const std::vector<double>& h, bool f32)
{
    if (f32) {
        std::vector<float> f(h.begin(), h.end());
        float* q = nullptr;
        int rc = upload(c, &q, f);
        *p = q;
        return rc;
    }
    double* q = nullptr;
    int rc = upload(c, &q, h);
    *p = q;
    return rc;
}

int uploadModel(efem_ctx* c)
{
    const efem::Model& m = c->model;
    int rc;
    if ((rc = upload(c, &c->connDof, m.connDof))) return rc;
    if ((rc = uploadReal(c, &c->ginv, m.ginv, false))) return rc;  // J0^-1 stays FP64
    if ((rc = upload(c, &c->vol, m.vol))) return rc;
    if ((rc = uploadReal(c, &c->gdN, m.gdN, false))) return rc;  // local gradients stay FP64
    if ((rc = upload(c, &c->gw, m.gw))) return rc;
    if ((rc = upload(c, &c->params, m.params))) return rc;
    const size_t nslot = static_cast<size_t>(m.nElems) * m.nen * 3;
    {
        // sized for FP64 slots; the FP32 kernels use the first half
        double* q = nullptr;
        if ((rc = devAlloc(c, &q, nslot))) return rc;
        c->eslots = q;
    }
    c->ginv64 = c->ginv;
    c->gdN64 = c->gdN;
    if (m.nen == 4) {
        // reference coordinates per free DOF: C3D4 elements form J0^-1 from
        // them instead of streaming 72 bytes each
        std::vector<double> xd(3 * static_cast<size_t>(m.nFree) + 3, 0.0);
        for (int n = 0; n < m.nNodes; ++n)
            if (m.fid[n] >= 0)
                for (int d = 0; d < 3; ++d) xd[3 * static_cast<size_t>(m.fid[n]) + d] = m.X[3 * static_cast<size_t>(n) + d];
        if ((rc = upload(c, &c->Xd, xd))) return rc;
    }
    if (m.nFaces > 0) {
        std::vector<int> fdof(static_cast<size_t>(m.nen) * m.nFaces);
        for (int f = 0; f < m.nFaces; ++f)
            for (int a = 0; a < m.nen; ++a)
                fdof[static_cast<size_t>(a) * m.nFaces + f] = m.connDof[static_cast<size_t>(a) * m.nElems + m.faceElem[f]];
        if ((rc = upload(c, &c->faceDof, fdof))) return rc;
        if ((rc = upload(c, &c->fA0, m.fA0))) return rc;
        if ((rc = upload(c, &c->fWN, m.fWN))) return rc;
        if ((rc = upload(c, &c->fG, m.fG))) return rc;
        if ((rc = devAlloc(c, &c->fslots, static_cast<size_t>(m.nFaces) * m.nen * 3))) return rc;
    }
    if ((rc = upload(c, &c->gptr, m.gptr))) return rc;
    if ((rc = upload(c, &c->gslot, m.gslot))) return rc;
    if (m.nen == 4 && c->parts.empty() && m.nElems > 0) {
        efem::Tet4Patches tp;
        efem::buildTet4Patches(m, efem::dev::kPatch, tp);
        efem::dev::PatchArgs& pa = c->patch;
        pa.E = m.nElems;
        pa.nPatch = tp.nPatch;
        int *conn = nullptr, *order = nullptr, *roff = nullptr, *rec = nullptr;
        double *vol = nullptr, *part = nullptr;
        uint16_t* ent = nullptr;
        if ((rc = upload(c, &conn, tp.conn)) || (rc = upload(c, &order, tp.order)) || (rc = upload(c, &vol, tp.vol)) ||
            (rc = upload(c, &roff, tp.roff)) || (rc = upload(c, &rec, tp.rec)) || (rc = upload(c, &ent, tp.ent)) ||
            (rc = devAlloc(c, &part, 3 * static_cast<size_t>(std::max(tp.nRec, 1)))) ||
            (rc = upload(c, &c->pRptr, tp.rptr)) || (rc = upload(c, &c->pFptr, tp.fptr)) ||
            (rc = upload(c, &c->pFslot, tp.fslot)))
            return rc;
        pa.conn = conn;
        pa.order = order;
        pa.vol = vol;
        pa.roff = roff;
        pa.rec = rec;
        pa.ent = ent;
        pa.part = part;
        c->pRec = tp.nRec;
    }
    // constant loads: without any (no nodal force, no gravity) the vector
    // stays off the device and the gather and energy kernels skip its reads
    // (subtracting a zero changes no value)
    if (std::any_of(m.fext.begin(), m.fext.end(), [](double v) { return v != 0.0; }))
        if ((rc = upload(c, &c->fext, m.fext))) return rc;
    const size_t np = static_cast<size_t>(m.ndof()) + 3;  // + dummy zero node
    // zero fills on the context stream: it is a non-blocking stream, so a
    // null-stream hipMemset would not be ordered before its work
    for (double** v : {&c->x, &c->r, &c->xp, &c->rp, &c->d, &c->tmp, &c->x3, &c->r3, &c->d2}) {
        if ((rc = devAlloc(c, v, np))) return rc;
        HIP_CHECK(c, hipMemsetAsync(*v, 0, np * sizeof(double), c->stream));
    }
    if ((rc = devAlloc(c, &c->gdotWork, efem::dev::gatherDotWork(m.nFree)))) return rc;
    if ((rc = devAlloc(c, &c->scal, kNumScal))) return rc;
    HIP_CHECK(c, hipMemsetAsync(c->scal, 0, kNumScal * sizeof(double), c->stream));
    if ((rc = devAlloc(c, &c->work, efem::dev::kReduceWorkDoubles))) return rc;
    HIP_CHECK(c, hipHostMalloc(reinterpret_cast<void**>(&c->hostScal), kHostScal * sizeof(double),
                               hipHostMallocDefault));
    HIP_CHECK(c, hipHostGetDevicePointer(reinterpret_cast<void**>(&c->hostScalDev), c->hostScal, 0));
    for (hipEvent_t& e : c->trialEv) HIP_CHECK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

efem::dev::ElemArgs elemArgs(efem_ctx* c, const double* x, bool f32)
{
    const efem::Model& m = c->model;
    efem::dev::ElemArgs ea;
    ea.nElems = m.nElems;
    ea.nen = m.nen;
    ea.ng = m.ng;
    ea.law = m.law;
    ea.paramsPerElem = m.paramsPerElem;
    ea.mixed = f32;
    ea.connDof = c->connDof;
    ea.ginv = f32 ? c->ginv : c->ginv64;
    ea.vol = c->vol;
    ea.gdN = f32 ? c->gdN : c->gdN64;
    ea.gw = c->gw;
    ea.params = c->params;
    ea.x = x;
    ea.slots = c->eslots;
    ea.Xd = c->tet4Geometry ? c->Xd : nullptr;
    ea.nFree = m.nFree;
    return ea;
}

efem::dev::FaceArgs faceArgs(efem_ctx* c, const double* x, const double* v)
{
    const efem::Model& m = c->model;
    efem::dev::FaceArgs fa;
    fa.nFaces = m.nFaces;
    fa.nen = m.nen;
    fa.fq = m.fq;
    fa.directed = m.directed;
    fa.pressure = m.pressure;
    for (int k = 0; k < 3; ++k) fa.pdir[k] = m.pdir[k];
    fa.faceDof = c->faceDof;
    fa.A0 = c->fA0;
    fa.WN = c->fWN;
    fa.G = c->fG;
    fa.x = x;
    fa.v = v;
    fa.fpress = c->fpress;
    fa.slots = c->fslots;
    return fa;
}

efem::dev::GatherArgs gatherArgs(efem_ctx* c, double* r, bool f32, double ce, double cf, const double* fext)
{
    const efem::Model& m = c->model;
    efem::dev::GatherArgs ga;
    ga.nFree = m.nFree;
    ga.nElemSlots = m.nElems * m.nen;
    ga.mixed = f32;
    ga.gptr = c->gptr;
    ga.gslot = c->gslot;
    ga.eslots = c->eslots;
    ga.fslots = c->fslots;
    ga.fext = fext;
    ga.ce = ce;
    ga.cf = cf;
    ga.r = r;
    return ga;
}

// residual terms (option "residual_terms")
constexpr int kTermElements = 1;  // element internal forces
constexpr int kTermPressure = 2;  // follower pressure / directed traction
constexpr int kTermConstant = 4;  // constant nodal forces and gravity
constexpr int kTermsAll = 7;

// r = f_int(x) - f_ext(x); x has 3 trailing zeros (the fixed-node slot)
// r(x); with dotOut, also the line search's scalars of the new r in the
// gather launch: dotOut = {r . dotD, max|r|, max|x - dotCb|, |x - dotXt|^2}
// terms: kTermsAll, or a subset for one constraint's part (option
// "residual_terms" of efem_residual)
int residualInternal(efem_ctx* c, const double* x, double* r, const double* dotD = nullptr,
                     const double* dotCb = nullptr, const double* dotXt = nullptr, double* dotOut = nullptr,
                     int terms = kTermsAll, const double* dotCarry = nullptr,
                     efem::dev::TrialFinish* defer = nullptr)
{
    const efem::Model& m = c->model;
    const bool f32 = elem32(c) && !c->elemFp64;
    // the element forces and the pressure faces go out together (one
    // launch for C3D4); the profiled element time then holds the faces
    const bool faces = m.nFaces > 0 && (terms & kTermPressure);
    const bool patches = (terms & kTermElements) && c->tet4Patches && c->patch.part;
    if (terms & kTermElements) {
        const efem::dev::ElemArgs ea = elemArgs(c, x, f32);
        const bool prof = c->profiling && c->evUsed + 2 <= static_cast<int>(c->evs.size());
        if (prof) HIP_CHECK(c, hipEventRecord(c->evs[c->evUsed], c->stream));
        if (patches) {
            efem::dev::FaceArgs fa = faceArgs(c, x, nullptr);
            if (!faces) fa.nFaces = 0;
            efem::dev::launchElementPatches(ea, c->patch, fa, c->stream);
        } else if (faces) {
            efem::dev::launchElementForceFaces(ea, faceArgs(c, x, nullptr), c->stream);
        } else {
            efem::dev::launchElementForce(ea, c->stream);
        }
        if (prof) {
            HIP_CHECK(c, hipEventRecord(c->evs[c->evUsed + 1], c->stream));
            c->evUsed += 2;
        }
        HIP_CHECK(c, hipGetLastError());
    }
    if (faces && !(terms & kTermElements)) {
        efem::dev::launchPressure(faceArgs(c, x, nullptr), c->stream);
        HIP_CHECK(c, hipGetLastError());
    }
    // a zero coefficient makes the gather skip those slots
    efem::dev::GatherArgs ga = gatherArgs(c, r, f32, (terms & kTermElements) ? 1.0 : 0.0,
                                          (terms & kTermPressure) ? -1.0 : 0.0,
                                          (terms & kTermConstant) ? c->fext : nullptr);
    if (patches) {
        // the records' FP64 partials (node-major) stand for the element
        // slots; the lists hold the face slots only
        ga.nElemSlots = 0;
        ga.mixed = 0;
        ga.rptr = c->pRptr;
        ga.gptr = c->pFptr;
        ga.gslot = c->pFslot;
        ga.eslots = c->patch.part;
        ga.lanes = c->patchLanes;
    }
    if (dotOut) {
        ga.dotD = dotD;
        ga.dotX = x;
        ga.dotCb = dotCb;
        ga.dotXt = dotXt;
        ga.dotWork = c->gdotWork;
        ga.dotOut = dotOut;
        ga.dotCarry = dotCarry;
        ga.defer = defer;
    }
    efem::dev::launchGather(ga, c->stream);
    HIP_CHECK(c, hipGetLastError());
    return EFEM_OK;
}

// out = K_t(x) v = d r / d x [v] (transpose: K_t^T v).  The element part is
// symmetric (hyperelastic); only the follower-pressure part has a transpose.
int tangentInternal(efem_ctx* c, const double* x, const double* v, double* out, int transpose, bool fp64 = false)
{
    const efem::Model& m = c->model;
    const bool f32 = elem32(c) && !fp64;
    efem::dev::launchElementTangent(elemArgs(c, x, f32), v, c->stream);
    HIP_CHECK(c, hipGetLastError());
    if (m.nFaces > 0) {
        efem::dev::launchPressureTangent(faceArgs(c, x, v), transpose, c->stream);
        HIP_CHECK(c, hipGetLastError());
    }
    efem::dev::launchGather(gatherArgs(c, out, f32, 1.0, -1.0, nullptr), c->stream);
    HIP_CHECK(c, hipGetLastError());
    return EFEM_OK;
}

efem::dev::TrsvArgs trsvArgs(efem_ctx* c)
{
    DeviceFactor& f = c->fac;
    efem::dev::TrsvArgs a;
    a.n = f.n;
    a.mixed = fac32(c);
    a.perm = f.perm;
    a.rows = f.rows;
    a.B = f.B;
    a.fptr = f.fptr;
    a.fidx = f.fidx;
    a.y = f.y;
    a.t = f.t;
    a.z = f.z;
    return a;
}

// permuted: the caller already wrote the permuted right-hand side into f.bp;
// unpermute = false: the solution stays in f.z in the factor's order (the
// caller reads it through pinv) and out is not written
void launchApplySweeps(efem_ctx* c, const double* v, double* out, bool permuted = false, bool unpermute = true)
{
    DeviceFactor& f = c->fac;
    const efem::dev::TrsvArgs a = trsvArgs(c);
    if (!permuted) efem::dev::launchPermute(a, v, f.bp, c->stream);
    // with the fused top, the levels hold only the supernodes outside T
    // (levels wholly inside T launch nothing) and T is one launch of
    // z_T = Z f_T between the two sweeps
    for (int l = 0; l < f.nlevels; ++l)
        efem::dev::launchTrsvForward(a, f.bp, f.items + 2 * f.fwdOff[l], f.fwdCnt[l], f.fwdWide[l], c->stream,
                                     f.fwdK[l] ? f.fwdTab + f.fwdTabOff[l] : nullptr, f.fwdK[l], f.fwdW[l]);
    if (f.fusedTop) {
        efem::dev::TopArgs t;
        t.nT = f.topN;
        t.ldz = f.topLd;
        t.K = f.topK;
        t.waves = c->topWaves;
        t.Z = f.Z;
        t.tcol = f.topCol;
        t.tab = f.topTab;
        efem::dev::launchTrsvTop(a, f.bp, t, c->stream);
    }
    for (int l = f.nlevels - 1; l >= 0; --l)
        efem::dev::launchTrsvBackward(a, f.items + 2 * f.bwdOff[l], f.bwdCnt[l], f.bwdWide[l], c->stream);
    if (unpermute) efem::dev::launchUnpermute(a, out, c->stream);
}

// Replay the captured apply: reads the staging buffer f.in, writes f.out
// (captured on first use).
int applyH0Graph(efem_ctx* c)
{
    DeviceFactor& f = c->fac;
    const bool prof = c->profiling && c->evH0Used + 2 <= static_cast<int>(c->evsH0.size());
    if (prof) HIP_CHECK(c, hipEventRecord(c->evsH0[c->evH0Used], c->stream));
    if (!f.exec) {
        HIP_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        launchApplySweeps(c, f.in, f.out);
        hipError_t e = hipStreamEndCapture(c->stream, &f.graph);
        if (e != hipSuccess) {
            c->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
            return EFEM_ERR_HIP;
        }
        HIP_CHECK(c, hipGraphInstantiate(&f.exec, f.graph, nullptr, nullptr, 0));
    }
    HIP_CHECK(c, hipGraphLaunch(f.exec, c->stream));
    if (prof) {
        HIP_CHECK(c, hipEventRecord(c->evsH0[c->evH0Used + 1], c->stream));
        c->evH0Used += 2;
    }
    HIP_CHECK(c, hipGetLastError());
    return EFEM_OK;
}

// out = K^-1 v on the device (v, out: interleaved free-DOF vectors).  With
// the graph option the forward/backward level launches are replayed from one
// captured HIP graph reading `in` and writing `out` staging buffers.
int applyH0Internal(efem_ctx* c, const double* v, double* out, bool permuted = false, bool unpermute = true)
{
    DeviceFactor& f = c->fac;
    if (c->useGraph && !permuted) {
        const size_t nb = 3 * static_cast<size_t>(f.n) * sizeof(double);
        if (v != f.in) HIP_CHECK(c, hipMemcpyAsync(f.in, v, nb, hipMemcpyDeviceToDevice, c->stream));
        int rc = applyH0Graph(c);
        if (rc) return rc;
        if (out != f.out) HIP_CHECK(c, hipMemcpyAsync(out, f.out, nb, hipMemcpyDeviceToDevice, c->stream));
        return EFEM_OK;
    }
    const bool prof = c->profiling && c->evH0Used + 2 <= static_cast<int>(c->evsH0.size());
    if (prof) HIP_CHECK(c, hipEventRecord(c->evsH0[c->evH0Used], c->stream));
    launchApplySweeps(c, v, out, permuted, unpermute);
    if (prof) {
        HIP_CHECK(c, hipEventRecord(c->evsH0[c->evH0Used + 1], c->stream));
        c->evH0Used += 2;
    }
    HIP_CHECK(c, hipGetLastError());
    return EFEM_OK;
}

// upload a (possibly merged) supernodal factor: blocks, index arrays, the
// forward extend-add map and the level-ordered work-item descriptors.  With
// snOffset (batched instances), also the per-instance item ranges of every
// level for the persistent solver.
// frees the device factor: its buffers and its captured apply graph
void releaseFactor(efem_ctx* c)
{
    DeviceFactor& d = c->fac;
    if (!d.allocs.empty() || d.exec || d.graph) (void)hipStreamSynchronize(c->stream);
    if (d.exec) (void)hipGraphExecDestroy(d.exec);
    if (d.graph) (void)hipGraphDestroy(d.graph);
    for (void* p : d.allocs) (void)hipFree(p);
    d = DeviceFactor();
}

// devAlloc records into `list` while this is alive
struct AllocInto {
    efem_ctx* c;
    AllocInto(efem_ctx* ctx, std::vector<void*>* list) : c(ctx) { c->allocList = list; }
    ~AllocInto() { c->allocList = &c->allocs; }
};

// the context's pinned staging buffer, grown to at least `bytes` (kept until
// efem_destroy, so refactorisations reuse it)
int pinnedStage(efem_ctx* c, size_t bytes, void** p)
{
    if (c->pinnedBytes < bytes) {
        if (c->pinned) (void)hipHostFree(c->pinned);
        c->pinned = nullptr;
        c->pinnedBytes = 0;
        HIP_CHECK(c, hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault));
        c->pinnedBytes = bytes;
    }
    *p = c->pinned;
    return EFEM_OK;
}

// Z of the fused top on the device (FP64 from the FP64 blocks of T's
// supernodes, `super` giving each one's offset and leading dimension, then
// FP32 for an FP32 factor) and the top kernel's tables; every buffer is the
// factor's (the caller's AllocInto)
int formTop(efem_ctx* c, const efem::TopPlan& P, const std::vector<efem::dev::TopSuper>& super, const double* src,
            bool f32)
{
    DeviceFactor& d = c->fac;
    int rc;
    const size_t nz = static_cast<size_t>(P.ld) * P.nT;
    std::vector<int4> tasks4(P.tasks.size() / 4);
    for (size_t i = 0; i < tasks4.size(); ++i)
        tasks4[i] = make_int4(P.tasks[4 * i], P.tasks[4 * i + 1], P.tasks[4 * i + 2], P.tasks[4 * i + 3]);
    double* z64 = nullptr;
    efem::dev::TopSuper* sn = nullptr;
    int* rtl = nullptr;
    int4* tasks = nullptr;
    if ((rc = devAlloc(c, &z64, nz)) || (rc = upload(c, &sn, super)) || (rc = upload(c, &rtl, P.rtl)) ||
        (rc = upload(c, &tasks, tasks4)))
        return rc;
    // the padding rows and the blocks between different roots stay zero
    HIP_CHECK(c, hipMemsetAsync(z64, 0, nz * sizeof(double), c->stream));
    efem::dev::TopBuild b;
    b.Z = z64;
    b.nT = P.nT;
    b.ldz = P.ld;
    b.B = src;
    b.sn = sn;
    b.rtl = rtl;
    b.tasks = tasks;
    for (const efem::TopPlan::Step& st : P.steps) {
        if (st.kind == 0) efem::dev::launchTopGemm(b, st.first, st.count, c->stream);
        else if (st.kind == 1) efem::dev::launchTopMirror(b, st.p0, st.p1, c->stream);
        else efem::dev::launchTopDiag(b, st.first, st.count, c->stream);
    }
    HIP_CHECK(c, hipGetLastError());
    if (f32) {
        float* z32 = nullptr;
        if ((rc = devAlloc(c, &z32, nz))) return rc;
        efem::dev::launchFactorToF32(z64, z32, static_cast<long long>(nz), c->stream);
        HIP_CHECK(c, hipGetLastError());
        d.Z = z32;
    } else {
        d.Z = z64;
    }
    if ((rc = upload(c, &d.topCol, P.tcol)) || (rc = upload(c, &d.topTab, P.tab))) return rc;
    d.fusedTop = true;
    d.topDepth = P.depth;
    d.topN = P.nT;
    d.topLd = P.ld;
    d.topK = P.K;
    return EFEM_OK;
}

// devX: the block inverses already on the device (FP64, device layout, the
// context's X arena, not owned by the factor): F holds the structure only
int uploadFactor(efem_ctx* c, efem::SupernodalFactor& F, const std::vector<int>* snOffset,
                 const double* devX = nullptr)
{
    using clock = std::chrono::steady_clock;
    const auto u0 = clock::now();
    auto since = [](clock::time_point a) { return std::chrono::duration<double>(clock::now() - a).count(); };
    double packS = 0.0, copyS = 0.0;
    releaseFactor(c);
    DeviceFactor& d = c->fac;
    AllocInto into(c, &d.allocs);
    d.n = F.n;
    d.nsuper = F.nsuper;
    d.nlevels = F.nlevels;
    d.nnz = F.nnzL;
    std::vector<int> w(F.nsuper), mm(F.nsuper), tOff(F.nsuper + 1);
    // device block offsets: each m x w block with leading dimension ldOf(m)
    std::vector<long long> dval(F.nsuper + 1, 0);
    int nt = 0;
    for (int s = 0; s < F.nsuper; ++s) {
        w[s] = F.first[s + 1] - F.first[s];
        if (w[s] > efem::kMaxSupernodeWidth) {
            c->err = "internal: supernode wider than the device staging buffer";
            return EFEM_ERR_STATE;
        }
        mm[s] = F.rowPtr[s + 1] - F.rowPtr[s];
        tOff[s] = nt;
        nt += mm[s] - w[s];
        dval[s + 1] = dval[s] + static_cast<long long>(w[s]) * efem::dev::ldOf(mm[s]);
    }
    tOff[F.nsuper] = nt;
    // forward-sweep extend-add map (children's update rows -> front rows)
    std::vector<int> fptr, fidx;
    efem::frontGatherMap(F, tOff, fptr, fidx);
    // the fused top (single problems): its supernodes leave the levels
    efem::TopPlan top;
    const bool fused = !snOffset && c->fusedRootOpt &&
                       efem::planTop(F, nt, c->fusedTopMax, efem::dev::kTopMaxK, efem::dev::kLdAlign, top);
    auto inTop = [&](int s) { return fused && top.inT[s]; };
    // entries of L the sweeps read (the diagonal block's lower triangle and
    // the rows below it), as nnzL counts them
    for (int s = 0; s < F.nsuper; ++s)
        if (!inTop(s))
            d.sweepEntries += static_cast<long long>(w[s]) * (w[s] + 1) / 2 + static_cast<long long>(mm[s] - w[s]) * w[s];
    // work-item descriptors per level
    std::vector<std::vector<int>> byLevel(F.nlevels);
    for (int s = 0; s < F.nsuper; ++s)
        if (!inTop(s)) byLevel[F.level[s]].push_back(s);
    std::vector<int4> items;
    auto desc1 = [&](int a0, int a1, long long val) {
        return make_int4(a0, a1, static_cast<int>(static_cast<unsigned>(val & 0xffffffffLL)),
                         static_cast<int>(val >> 32));
    };
    d.fwdOff.assign(F.nlevels, 0);
    d.fwdCnt.assign(F.nlevels, 0);
    d.bwdOff.assign(F.nlevels, 0);
    d.bwdCnt.assign(F.nlevels, 0);
    // levels of wide (tall) supernodes use 16-wave forward (backward) items:
    // the per-lane column (row) chains are 4x shorter
    d.fwdWide.assign(F.nlevels, 0);
    d.bwdWide.assign(F.nlevels, 0);
    {
        // a level goes wide when it has a wide (tall) supernode and too few
        // items to fill the chip with 4-wave blocks (measured on the beam:
        // 16-wave items lose on levels of >500 forward / >1000 backward items)
        std::vector<int> nf(F.nlevels, 0), nb(F.nlevels, 0), maxw(F.nlevels, 0), maxm(F.nlevels, 0);
        for (int s = 0; s < F.nsuper; ++s) {
            if (inTop(s)) continue;
            const int l = F.level[s];
            nf[l] += (mm[s] + 63) / 64;
            nb[l] += (w[s] + 3) / 4;
            maxw[l] = std::max(maxw[l], w[s]);
            maxm[l] = std::max(maxm[l], mm[s]);
        }
        for (int l = 0; l < F.nlevels; ++l) {
            d.fwdWide[l] = maxw[l] >= kWideColumns && nf[l] <= c->wideFwdItems;
            d.bwdWide[l] = maxm[l] >= kTallRows && nb[l] <= c->wideBwdItems;
            // many 4-column items: 16-column items keep the level's waves
            // within what the chip holds at once (not in batches, whose
            // persistent kernel walks 4-column items)
            if (!snOffset && nb[l] >= c->narrowMinItems) d.bwdWide[l] = efem::dev::kBwdNarrow;
            // staged backward items on the other levels (batches: 4
            // columns, the items their persistent kernel walks)
            if (c->bwdStage > 0 && d.bwdWide[l] != efem::dev::kBwdNarrow)
                d.bwdWide[l] = snOffset || c->bwdStage == 4 ? efem::dev::kBwdStage4
                               : c->bwdStage == 8           ? efem::dev::kBwdStage8
                                                            : efem::dev::kBwdStage16;
        }
    }
    for (int l = 0; l < F.nlevels; ++l) {
        d.fwdOff[l] = static_cast<int>(items.size() / 2);
        for (int s : byLevel[l])
            for (int r0 = 0; r0 < mm[s]; r0 += 64) {
                items.push_back(make_int4(r0, w[s], mm[s], F.first[s]));
                items.push_back(desc1(F.rowPtr[s], tOff[s], dval[s]));
            }
        d.fwdCnt[l] = static_cast<int>(items.size() / 2) - d.fwdOff[l];
    }
    // forward extend-add tables: per level with at most kFwdTableMaxK
    // contributions per front row (not for batches, whose persistent kernel
    // walks fptr / fidx); absent entries name the zero slot nt
    d.fwdTabOff.assign(F.nlevels, 0);
    d.fwdK.assign(F.nlevels, 0);
    d.fwdW.assign(F.nlevels, 0);
    std::vector<int> tab;
    if (!snOffset && c->fwdTables) {
        std::vector<efem::FwdTableItem> titems;
        long long size = 0;
        for (int l = 0; l < F.nlevels; ++l) {
            int K = 0, W = 0;
            for (int s : byLevel[l]) {
                W = std::max(W, w[s]);
                for (int r = 0; r < mm[s]; ++r)
                    K = std::max(K, fptr[F.rowPtr[s] + r + 1] - fptr[F.rowPtr[s] + r]);
            }
            if (K < 1 || K > efem::dev::kFwdTableMaxK) continue;
            d.fwdK[l] = K;
            d.fwdW[l] = W;
            d.fwdTabOff[l] = size;
            for (int s : byLevel[l])
                for (int r0 = 0; r0 < mm[s]; r0 += 64) {
                    titems.push_back({s, r0, K, W, size});
                    size += static_cast<long long>(K) * (64 + W);
                }
        }
        efem::fillForwardTables(F, fptr, fidx, titems, size, nt, tab);
    }
    for (int l = 0; l < F.nlevels; ++l) {
        d.bwdOff[l] = static_cast<int>(items.size() / 2);
        const int cols = d.bwdWide[l] == efem::dev::kBwdNarrow ? 16
                         : d.bwdWide[l] >= efem::dev::kBwdStage4 ? efem::dev::bwdStageCols(d.bwdWide[l])
                                                                 : 4;
        for (int s : byLevel[l])
            for (int k0 = 0; k0 < w[s]; k0 += cols) {
                items.push_back(make_int4(k0, w[s], mm[s], F.first[s]));
                items.push_back(desc1(F.rowPtr[s], 0, dval[s]));
            }
        d.bwdCnt[l] = static_cast<int>(items.size() / 2) - d.bwdOff[l];
    }
    if (snOffset) {
        // items of one instance are contiguous within each level (supernodes
        // are numbered instance by instance and listed in order)
        const int np = static_cast<int>(snOffset->size()) - 1;
        std::vector<int> inst(4 * static_cast<size_t>(np) * F.nlevels, 0);  // fwdOff, fwdCnt, bwdOff, bwdCnt
        std::vector<int> partOf(F.nsuper);
        for (int i = 0; i < np; ++i)
            for (int s = (*snOffset)[i]; s < (*snOffset)[i + 1]; ++s) partOf[s] = i;
        for (int l = 0; l < F.nlevels; ++l) {
            int fo = d.fwdOff[l], bo = d.bwdOff[l];
            for (int s : byLevel[l]) {
                const int i = partOf[s];
                int* q = &inst[4 * (static_cast<size_t>(i) * F.nlevels + l)];
                const int nf = (mm[s] + 63) / 64, nb = (w[s] + 3) / 4;
                if (q[1] == 0) q[0] = fo;
                if (q[3] == 0) q[2] = bo;
                q[1] += nf;
                q[3] += nb;
                fo += nf;
                bo += nb;
            }
        }
        d.instLevels = F.nlevels;
        int rc0 = upload(c, &d.instItems, inst);
        if (rc0) return rc0;
    }
    int rc;
    if ((rc = upload(c, &d.perm, F.perm))) return rc;
    {
        std::vector<int> pinv(F.perm.size());
        for (size_t k = 0; k < F.perm.size(); ++k) pinv[F.perm[k]] = static_cast<int>(k);
        if ((rc = upload(c, &d.pinv, pinv))) return rc;
    }
    if ((rc = upload(c, &d.rows, F.rows))) return rc;
    if (devX) {
        // computed on the device in FP64 in the device layout: used in place
        // (FP64), or converted (FP32)
        const long long nb = dval[F.nsuper];
        c->factorPhases[3] = since(u0);
        const auto p0 = clock::now();
        if (fac32(c)) {
            float* q = nullptr;
            if ((rc = devAlloc(c, &q, static_cast<size_t>(std::max(nb, 1LL))))) return rc;
            efem::dev::launchFactorToF32(devX, q, nb, c->stream);
            HIP_CHECK(c, hipGetLastError());
            d.B = q;
        } else {
            d.B = const_cast<double*>(devX);
        }
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        packS = since(p0);
    } else {
        // block-inverse values, repacked to the device layout (leading
        // dimension ldOf(m), zero padding rows) and converted to FP32 if the
        // preconditioner runs in FP32, by all host threads (packBlocks) into
        // pinned memory, then one DMA
        const long long nb = dval[F.nsuper];
        const size_t es = fac32(c) ? sizeof(float) : sizeof(double);
        void* h = nullptr;
        c->factorPhases[3] = since(u0);
        const auto p0 = clock::now();
        if ((rc = pinnedStage(c, static_cast<size_t>(std::max(nb, 1LL)) * es, &h))) return rc;
        static_assert(efem::dev::ldOf(1) == efem::dev::kLdAlign && efem::dev::ldOf(efem::dev::kLdAlign + 1) ==
                          2 * efem::dev::kLdAlign,
                      "packBlocks pads to multiples of kLdAlign");
        efem::packBlocks(F, dval, efem::dev::kLdAlign, fac32(c), h);
        packS = since(p0);
        const auto q0 = clock::now();
        double* q = nullptr;
        if ((rc = devAlloc(c, &q, (static_cast<size_t>(std::max(nb, 1LL)) * es + 7) / 8))) return rc;
        HIP_CHECK(c, hipMemcpyAsync(q, h, static_cast<size_t>(nb) * es, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        copyS = since(q0);
        d.B = q;
    }
    if (fused) {
        // Z from the FP64 block inverses of T: the X arena in place (device
        // factorisation), else T's blocks of the host factor uploaded
        const double* src = devX;
        std::vector<efem::dev::TopSuper> super(top.super.size());
        std::vector<double> hb;
        for (size_t q = 0; q < top.super.size(); ++q) {
            const efem::TopSuperInfo& t = top.super[q];
            efem::dev::TopSuper& o = super[q];
            o.cs = t.cs;
            o.w = t.w;
            o.m = t.m;
            o.rOff = t.rOff;
            o.ldz = top.ld;
            o.zoff = 0;
            if (devX) {
                o.ld = efem::dev::ldOf(t.m);
                o.boff = dval[t.s];
            } else {
                o.ld = t.m;
                o.boff = static_cast<long long>(hb.size());
                const double* b0 = F.B.get() + F.valPtr[t.s];
                hb.insert(hb.end(), b0, b0 + static_cast<size_t>(t.w) * t.m);
            }
        }
        if (!devX) {
            double* q64 = nullptr;
            if ((rc = upload(c, &q64, hb))) return rc;
            src = q64;
        }
        if ((rc = formTop(c, top, super, src, fac32(c)))) return rc;
    }
    if ((rc = upload(c, &d.fptr, fptr))) return rc;
    if ((rc = upload(c, &d.fidx, fidx))) return rc;
    if (!tab.empty() && (rc = upload(c, &d.fwdTab, tab))) return rc;
    if ((rc = devAlloc(c, &d.in, 3 * static_cast<size_t>(F.n) + 3))) return rc;
    if ((rc = devAlloc(c, &d.out, 3 * static_cast<size_t>(F.n) + 3))) return rc;
    if ((rc = devAlloc(c, &d.bp, 3 * static_cast<size_t>(F.n) + 3))) return rc;
    if ((rc = upload(c, &d.items, items))) return rc;
    const size_t es = fac32(c) ? sizeof(float) : sizeof(double);
    double* q = nullptr;
    if ((rc = devAlloc(c, &q, (3 * static_cast<size_t>(F.n) * es + 7) / 8))) return rc;
    d.y = q;
    // update vectors + the zero slot nt that absent table entries name
    if ((rc = devAlloc(c, &q, (3 * static_cast<size_t>(nt + 1) * es + 7) / 8))) return rc;
    HIP_CHECK(c, hipMemsetAsync(q, 0, (3 * static_cast<size_t>(nt + 1) * es + 7) / 8 * 8, c->stream));
    d.t = q;
    if ((rc = devAlloc(c, &q, (3 * static_cast<size_t>(F.n) * es + 7) / 8))) return rc;
    d.z = q;
    d.ready = true;
    c->factorPhases[4] = packS;
    c->factorPhases[5] = copyS;
    c->factorPhases[6] = since(u0) - c->factorPhases[3] - packS - copyS;
    return EFEM_OK;
}

// the approximate Jacobian's pattern and contribution map (once per
// context), then its values by the device kernels into lapA.x
int deviceLaplacian(efem_ctx* c, bool download = true)
{
    const efem::Model& m = c->model;
    if (!c->lapReady) {
        std::vector<int> ptr, src;
        efem::laplacianMap(m, c->lapA, ptr, src);
        int rc;
        if ((rc = upload(c, &c->lapPtr, ptr))) return rc;
        if ((rc = upload(c, &c->lapSrc, src))) return rc;
        if ((rc = devAlloc(c, &c->lapKel, static_cast<size_t>(m.nElems) * m.nen * m.nen))) return rc;
        if ((rc = devAlloc(c, &c->lapVals, c->lapA.x.size()))) return rc;
        c->lapReady = true;
    }
    efem::dev::LapArgs a{};
    a.E = m.nElems;
    a.nen = m.nen;
    a.ng = m.ng;
    a.law = m.law;
    a.ppe = m.paramsPerElem;
    a.nnz = static_cast<int>(c->lapA.x.size());
    a.ginv = static_cast<const double*>(c->ginv64);
    a.vol = c->vol;
    a.gdN = static_cast<const double*>(c->gdN64);
    a.gw = c->gw;
    a.params = c->params;
    a.kel = c->lapKel;
    a.ptr = c->lapPtr;
    a.src = c->lapSrc;
    a.vals = c->lapVals;
    efem::dev::launchLaplacian(a, c->stream);
    HIP_CHECK(c, hipGetLastError());
    if (download) {
        HIP_CHECK(c, hipMemcpyAsync(c->lapA.x.data(), c->lapVals, c->lapA.x.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
    }
    return EFEM_OK;
}

// Numeric factorisation on the device from the symbolic analysis S: the
// plan and its device copy are built once per pattern; the arenas are
// zeroed and the plan's launches run on the context stream.  *notSpd = a
// pivot was not positive (the reference's fallback case).
int deviceNumeric(efem_ctx* c, const efem::SupernodalSymbolic& S, bool rebuild, bool* notSpd)
{
    if (!c->dplan || rebuild) {
        auto P = std::make_unique<efem::DevFactorPlan>();
        efem::buildDevFactorPlan(S, *P);
        // plan arrays and arenas (replacing an earlier plan's)
        for (int** q : {&c->dfW, &c->dfM, &c->dfRel, &c->dfRelOff, &c->dfPairs, &c->dfAsmSrc, &c->dfFail}) {
            int rc0 = devFree(c, q);
            if (rc0) return rc0;
        }
        for (long long** q : {&c->dfDval, &c->dfUoff, &c->dfAsmDst}) {
            int rc0 = devFree(c, q);
            if (rc0) return rc0;
        }
        for (double** q : {&c->dfL, &c->dfX, &c->dfU}) {
            int rc0 = devFree(c, q);
            if (rc0) return rc0;
        }
        int rc = devFree(c, &c->dfTasks);
        if (rc) return rc;
        std::vector<int4> t4(P->tasks.size() / 4);
        for (size_t i = 0; i < t4.size(); ++i)
            t4[i] = make_int4(P->tasks[4 * i], P->tasks[4 * i + 1], P->tasks[4 * i + 2], P->tasks[4 * i + 3]);
        if ((rc = upload(c, &c->dfW, P->w)) || (rc = upload(c, &c->dfM, P->m)) ||
            (rc = upload(c, &c->dfDval, P->dval)) || (rc = upload(c, &c->dfUoff, P->uOff)) ||
            (rc = upload(c, &c->dfRel, P->rel)) || (rc = upload(c, &c->dfRelOff, P->relOff)) ||
            (rc = upload(c, &c->dfPairs, P->pairs)) || (rc = upload(c, &c->dfAsmDst, P->asmDst)) ||
            (rc = upload(c, &c->dfAsmSrc, P->asmSrc)) || (rc = upload(c, &c->dfTasks, t4)) ||
            (rc = devAlloc(c, &c->dfFail, 1)) || (rc = devAlloc(c, &c->dfL, static_cast<size_t>(P->panelSize))) ||
            (rc = devAlloc(c, &c->dfX, static_cast<size_t>(P->panelSize))) ||
            (rc = devAlloc(c, &c->dfU, static_cast<size_t>(std::max(P->updSize, 1LL)))))
            return rc;
        c->dfLaunches.clear();
        for (const efem::DevLaunch& l : P->launches) {
            c->dfLaunches.push_back(l.kind);
            c->dfLaunches.push_back(l.first);
            c->dfLaunches.push_back(l.count);
        }
        c->dplan = std::move(P);
    }
    const efem::DevFactorPlan& P = *c->dplan;
    HIP_CHECK(c, hipMemsetAsync(c->dfL, 0, static_cast<size_t>(P.panelSize) * sizeof(double), c->stream));
    HIP_CHECK(c, hipMemsetAsync(c->dfX, 0, static_cast<size_t>(P.panelSize) * sizeof(double), c->stream));
    if (P.updSize > 0)
        HIP_CHECK(c, hipMemsetAsync(c->dfU, 0, static_cast<size_t>(P.updSize) * sizeof(double), c->stream));
    HIP_CHECK(c, hipMemsetAsync(c->dfFail, 0, sizeof(int), c->stream));
    efem::dev::DevFactorArgs a{};
    a.L = c->dfL;
    a.X = c->dfX;
    a.U = c->dfU;
    a.A = c->lapVals;
    a.w = c->dfW;
    a.m = c->dfM;
    a.dval = c->dfDval;
    a.uOff = c->dfUoff;
    a.rel = c->dfRel;
    a.relOff = c->dfRelOff;
    a.pairs = c->dfPairs;
    a.asmDst = c->dfAsmDst;
    a.asmSrc = c->dfAsmSrc;
    a.tasks = c->dfTasks;
    a.fail = c->dfFail;
    a.launches = c->dfLaunches.data();
    a.nlaunch = static_cast<int>(c->dfLaunches.size() / 3);
    efem::dev::launchDeviceFactor(a, c->stream);
    HIP_CHECK(c, hipGetLastError());
    int fail = 0;
    HIP_CHECK(c, hipMemcpyAsync(&fail, c->dfFail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    *notSpd = fail != 0;
    return EFEM_OK;
}

// the factor's structure (no values) from the symbolic analysis, for the
// device upload of a factor computed on the device
void factorStructure(const efem::SupernodalSymbolic& S, efem::SupernodalFactor& F)
{
    F = efem::SupernodalFactor();
    F.n = S.n;
    F.perm = S.perm;
    F.nsuper = S.nsuper;
    F.first = S.first;
    F.rowPtr = S.rowPtr;
    F.rows = S.rows;
    F.valPtr = S.valPtr;
    F.parent = S.parent;
    F.level = S.level;
    F.nlevels = S.nlevels;
    F.nnzL = S.nnzL;
    F.flops = S.flops;
}

// Batches: Z_i of every instance over its whole elimination tree, by the
// fused top's selected inversion (T = all supernodes of the instance) from
// the FP64 block inverses of the merged factor F (host layout, uploaded for
// the formation), into one arena; FP32 copy for the FP32 preconditioner.
// Leaves the level path in place (no dense Z) when an instance is larger
// than kBatchDenseMaxN free nodes.
int formBatchZ(efem_ctx* c, const std::vector<efem::SupernodalFactor>& P, const efem::SupernodalFactor& F,
               const std::vector<int>& snOffset)
{
    DeviceFactor& d = c->fac;
    const int np = static_cast<int>(P.size());
    std::vector<efem::TopPlan> plans(np);
    for (int i = 0; i < np; ++i) {
        if (P[i].n > efem::dev::kBatchDenseMaxN) return EFEM_OK;
        if (!efem::planTop(P[i], 0, P[i].n, efem::dev::kTopMaxK, efem::dev::kLdAlign, plans[i]) ||
            plans[i].nT != P[i].n)
            return EFEM_OK;
    }
    AllocInto ai(c, &d.allocs);
    std::vector<long long> zoff(np);
    std::vector<int> zld(np), ztcol(static_cast<size_t>(F.n));
    std::vector<efem::dev::TopSuper> sn;
    std::vector<int> rtl, snBase(np), rtlBase(np);
    std::vector<int4> tasks;
    long long tot = 0;
    for (int i = 0; i < np; ++i) {
        const efem::TopPlan& T = plans[i];
        const int col = F.first[snOffset[i]];  // the instance's first merged column
        zoff[i] = tot;
        zld[i] = T.ld;
        tot += (static_cast<long long>(T.ld) * T.nT + 63) / 64 * 64;
        for (int t = 0; t < T.nT; ++t) ztcol[static_cast<size_t>(col) + t] = col + T.tcol[t];
        snBase[i] = static_cast<int>(sn.size());
        rtlBase[i] = static_cast<int>(rtl.size());
        for (const efem::TopSuperInfo& t : T.super) {
            efem::dev::TopSuper o;
            o.cs = t.cs;
            o.w = t.w;
            o.m = t.m;
            o.ld = t.m;
            o.boff = F.valPtr[snOffset[i] + t.s];
            o.rOff = rtlBase[i] + t.rOff;
            o.ldz = T.ld;
            o.zoff = zoff[i];
            sn.push_back(o);
        }
        rtl.insert(rtl.end(), T.rtl.begin(), T.rtl.end());
    }
    // step j of every tree in one launch: the trees' step sequences follow
    // one pattern by depth (their kinds agree step by step); tile tasks with
    // global supernode indices, mirror steps as tile tasks
    size_t nsteps = 0;
    for (const efem::TopPlan& T : plans) nsteps = std::max(nsteps, T.steps.size());
    struct Merged {
        int kind, first, count;
    };
    std::vector<Merged> merged;
    std::vector<efem::dev::TopMirrorTask> mtasks;
    for (size_t j = 0; j < nsteps; ++j) {
        int kind = -1;
        for (const efem::TopPlan& T : plans)
            if (j < T.steps.size()) {
                if (kind >= 0 && T.steps[j].kind != kind) return EFEM_OK;  // the level path stays
                kind = T.steps[j].kind;
            }
        Merged mg{kind, 0, 0};
        if (kind == 1) {
            mg.first = static_cast<int>(mtasks.size());
            for (int i = 0; i < np; ++i) {
                if (j >= plans[i].steps.size()) continue;
                const int p0 = plans[i].steps[j].p0, p1 = plans[i].steps[j].p1;
                if (p0 <= 0 || p1 <= p0) continue;
                constexpr int TB = efem::dev::kTopTile;
                for (int i0 = 0; i0 < p0; i0 += TB)
                    for (int j0 = p0; j0 < p1; j0 += TB)
                        mtasks.push_back(efem::dev::TopMirrorTask{zoff[i], plans[i].ld, p0, p1, i0, j0, 0});
            }
            mg.count = static_cast<int>(mtasks.size()) - mg.first;
        } else {
            mg.first = static_cast<int>(tasks.size());
            for (int i = 0; i < np; ++i) {
                if (j >= plans[i].steps.size()) continue;
                const efem::TopPlan::Step& st = plans[i].steps[j];
                for (int q = st.first; q < st.first + st.count; ++q)
                    tasks.push_back(make_int4(plans[i].tasks[4 * q] + snBase[i], plans[i].tasks[4 * q + 1],
                                              plans[i].tasks[4 * q + 2], plans[i].tasks[4 * q + 3]));
            }
            mg.count = static_cast<int>(tasks.size()) - mg.first;
        }
        merged.push_back(mg);
    }
    int rc;
    double *b64 = nullptr, *z64 = nullptr;
    efem::dev::TopSuper* snd = nullptr;
    int *rtld = nullptr, *zldd = nullptr, *ztc = nullptr;
    int4* td = nullptr;
    efem::dev::TopMirrorTask* md = nullptr;
    long long* zo = nullptr;
    if ((rc = devAlloc(c, &b64, static_cast<size_t>(std::max(F.Bsize, 1LL)))) || (rc = upload(c, &snd, sn)) ||
        (rc = upload(c, &rtld, rtl)) || (rc = upload(c, &td, tasks)) || (rc = upload(c, &md, mtasks)) ||
        (rc = devAlloc(c, &z64, static_cast<size_t>(tot))) ||
        (rc = upload(c, &zo, zoff)) || (rc = upload(c, &zldd, zld)) || (rc = upload(c, &ztc, ztcol)))
        return rc;
    HIP_CHECK(c, hipMemcpyAsync(b64, F.B.get(), static_cast<size_t>(F.Bsize) * sizeof(double), hipMemcpyHostToDevice,
                                c->stream));
    // padding rows stay zero (the dense apply reads whole 16-byte groups)
    HIP_CHECK(c, hipMemsetAsync(z64, 0, static_cast<size_t>(tot) * sizeof(double), c->stream));
    efem::dev::TopBuild b;
    b.Z = z64;
    b.nT = 0;
    b.ldz = 0;
    b.B = b64;
    b.sn = snd;
    b.rtl = rtld;
    b.tasks = td;
    for (const Merged& mg : merged) {
        if (mg.kind == 0) efem::dev::launchTopGemm(b, mg.first, mg.count, c->stream);
        else if (mg.kind == 1) efem::dev::launchTopMirrorTasks(z64, md + mg.first, mg.count, c->stream);
        else efem::dev::launchTopDiag(b, mg.first, mg.count, c->stream);
    }
    HIP_CHECK(c, hipGetLastError());
    if (fac32(c)) {
        float* z32 = nullptr;
        if ((rc = devAlloc(c, &z32, static_cast<size_t>(tot)))) return rc;
        efem::dev::launchFactorToF32(z64, z32, tot, c->stream);
        HIP_CHECK(c, hipGetLastError());
        d.batchZ = z32;
    } else {
        d.batchZ = z64;
    }
    d.batchZoff = zo;
    d.batchZld = zldd;
    d.batchZtcol = ztc;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

int factorInternal(efem_ctx* c, int* failed)
{
    const auto t0 = std::chrono::steady_clock::now();
    const efem::Model& m = c->model;
    releaseFactor(c);
    c->factorFailed = true;
    if (failed) *failed = 1;
    if (m.nFree == 0) return EFEM_OK;  // jacobi.nonZeros() == 0 -> fallback
    efem::SupernodalFactor F;
    std::string msg;
    std::vector<int> snOffset;
    std::vector<efem::SupernodalFactor> Pparts;  // batches: per-instance structure (blocks merged into F)
    bool ok;
    if (!c->parts.empty()) {
        ok = efem::factorModels(c->parts, Pparts, msg, c->denseSubtree);
        if (ok) efem::mergeFactors(Pparts, F, snOffset);
    } else {
        // element codes e nen^2 + a nen + b must fit the map's ints
        const bool dev = c->devLaplacian && static_cast<long long>(m.nElems) * m.nen * m.nen < (1LL << 31);
        if (c->reuseSymbolic && !c->symbolic) c->symbolic.reset(new efem::SupernodalSymbolic());
        efem::SupernodalSymbolic* sym = c->reuseSymbolic ? c->symbolic.get() : nullptr;
        if (dev && c->devFactorOpt) {
            // assembly, numeric factorisation and block inverses on the
            // device; ordering and symbolic analysis on the host (once per
            // pattern unless reuse_symbolic = 0)
            int rc0 = deviceLaplacian(c, false);
            if (rc0) return rc0;
            efem::SupernodalSymbolic fresh;
            efem::SupernodalSymbolic& S = sym ? *sym : fresh;
            const bool rebuild = !sym || S.n == 0;
            double symS = 0.0;
            if (S.n == 0) {
                const auto ts = std::chrono::steady_clock::now();
                const std::vector<int> perm = efem::nestedDissection(c->lapA, efem::freeCoords(m));
                if (!efem::supernodalSymbolic(c->lapA, perm, S, msg, c->denseSubtree)) {
                    S = efem::SupernodalSymbolic();
                    c->err = msg;
                    return EFEM_OK;
                }
                symS = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
            }
            const auto tn = std::chrono::steady_clock::now();
            bool notSpd = false;
            if ((rc0 = deviceNumeric(c, S, rebuild, &notSpd))) return rc0;
            const double numS = std::chrono::duration<double>(std::chrono::steady_clock::now() - tn).count();
            if (notSpd) {
                c->err = "approximate Jacobian is not positive definite";
                c->factorSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                return EFEM_OK;  // the reference falls back to I/theta
            }
            factorStructure(S, F);
            c->factorPhases[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            c->factorPhases[1] = symS;
            c->factorPhases[2] = numS;
            int rc = uploadFactor(c, F, nullptr, c->dfX);
            if (rc) return rc;
            c->factorFailed = false;
            if (failed) *failed = 0;
            c->factorSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            c->factorPhases[7] = c->factorSeconds;
            return EFEM_OK;
        }
        if (dev) {
            int rc0 = deviceLaplacian(c);
            if (rc0) return rc0;
            ok = efem::factorLaplacian(m, c->lapA, F, msg, c->denseSubtree, c->hostWs.get(), sym);
        } else {
            ok = efem::factorModel(m, F, msg, c->denseSubtree, c->hostWs.get(), sym);
        }
    }
    if (!ok) {
        c->err = msg;
        c->factorSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return EFEM_OK;  // not SPD: the reference falls back to I/theta
    }
    c->factorPhases[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    c->factorPhases[1] = F.seconds - F.numericSeconds;
    c->factorPhases[2] = F.numericSeconds;
    int rc = uploadFactor(c, F, c->parts.empty() ? nullptr : &snOffset);
    if (rc == EFEM_OK && !c->parts.empty() && c->batchDense) rc = formBatchZ(c, Pparts, F, snOffset);
    if (c->parts.empty()) c->hostWs->giveBlocks(std::move(F.B), static_cast<size_t>(F.Bsize));
    if (rc) return rc;
    c->factorFailed = false;
    if (failed) *failed = 0;
    c->factorSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    c->factorPhases[7] = c->factorSeconds;
    return EFEM_OK;
}

int readScalars(efem_ctx* c, int count)
{
    HIP_CHECK(c, hipMemcpyAsync(c->hostScal, c->scal + kOut, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

int ensureKrylov(efem_ctx* c)
{
    if (c->ku) return EFEM_OK;
    const size_t np = static_cast<size_t>(c->model.ndof()) + 3;
    int rc;
    for (double** v : {&c->ku, &c->kb, &c->kx, &c->kr, &c->krh, &c->kp, &c->kv, &c->ks, &c->kt, &c->kph, &c->ksh,
                       &c->kxo, &c->kbo}) {
        if ((rc = devAlloc(c, v, np))) return rc;
        // on the context stream, ordered before the copies that follow
        HIP_CHECK(c, hipMemsetAsync(*v, 0, np * sizeof(double), c->stream));
    }
    return EFEM_OK;
}

// (a . b, b . b) on the host
int dotHost(efem_ctx* c, const double* a, const double* b, double* ab, double* bb)
{
    efem::dev::launchDot2(a, b, c->model.ndof(), c->work, c->scal + kOut, c->scal + kOut + 1, c->stream);
    int rc = readScalars(c, 2);
    if (rc) return rc;
    *ab = c->hostScal[0];
    if (bb) *bb = c->hostScal[1];
    return EFEM_OK;
}

int precondApply(efem_ctx* c, bool h0, const double* in, double* out, efem_krylov_stats& st)
{
    if (!h0) {
        HIP_CHECK(c, hipMemcpyAsync(out, in, c->model.ndof() * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        return EFEM_OK;
    }
    ++st.precond_applies;
    return applyH0Internal(c, in, out);
}

// Solve K_t(u) x = b (transpose: K_t^T x = b) on the device vectors ku, kb ->
// kx with preconditioned CG (symmetric tangent) or BiCGStab.  M = the
// approximate-Jacobian Cholesky factor (or I if it failed / is disabled).
// The operator uses the context's element precision unless fp64.
int krylovCore(efem_ctx* c, int transpose, const efem_krylov_params& prm, efem_krylov_stats& st, bool fp64)
{
    const int n = c->model.ndof();
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    if (prm.use_precond && !c->fac.ready) {
        int failed = 0;
        if ((rc = factorInternal(c, &failed))) return rc;
        st.factor_seconds = c->factorSeconds;
    }
    const bool h0 = prm.use_precond && c->fac.ready;
    int method = prm.method;
    if (method == EFEM_KRYLOV_AUTO) method = c->model.nFaces > 0 ? EFEM_KRYLOV_BICGSTAB : EFEM_KRYLOV_CG;
    HIP_CHECK(c, hipMemsetAsync(c->kx, 0, n * sizeof(double), c->stream));
    double bb = 0.0;
    if ((rc = dotHost(c, c->kb, c->kb, &bb, nullptr))) return rc;
    const double bnorm = std::sqrt(bb);
    st.rel_residual = 0.0;
    if (bnorm == 0.0 || n == 0) {
        st.converged = 1;
        return EFEM_OK;
    }
    const double stop = prm.tol * bnorm;
    const size_t nb = static_cast<size_t>(n) * sizeof(double);
    auto A = [&](const double* v, double* out) {
        ++st.matvecs;
        return tangentInternal(c, c->ku, v, out, transpose, fp64);
    };
    HIP_CHECK(c, hipMemcpyAsync(c->kr, c->kb, nb, hipMemcpyDeviceToDevice, c->stream));
    if (method == EFEM_KRYLOV_CG) {
        // z = M r, p = z
        if ((rc = precondApply(c, h0, c->kr, c->ks, st))) return rc;
        HIP_CHECK(c, hipMemcpyAsync(c->kp, c->ks, nb, hipMemcpyDeviceToDevice, c->stream));
        double rz = 0.0;
        if ((rc = dotHost(c, c->kr, c->ks, &rz, nullptr))) return rc;
        for (int it = 1; it <= prm.max_iterations; ++it) {
            if ((rc = A(c->kp, c->kv))) return rc;
            double pq = 0.0;
            if ((rc = dotHost(c, c->kp, c->kv, &pq, nullptr))) return rc;
            if (!(pq != 0.0) || !std::isfinite(pq)) break;
            const double alpha = rz / pq;
            efem::dev::launchAxpby(c->kx, 1.0, c->kx, alpha, c->kp, n, c->stream);
            efem::dev::launchAxpby(c->kr, 1.0, c->kr, -alpha, c->kv, n, c->stream);
            double rr = 0.0;
            if ((rc = dotHost(c, c->kr, c->kr, &rr, nullptr))) return rc;
            st.iterations = it;
            st.rel_residual = std::sqrt(rr) / bnorm;
            if (std::sqrt(rr) <= stop) {
                st.converged = 1;
                break;
            }
            if ((rc = precondApply(c, h0, c->kr, c->ks, st))) return rc;
            double rz1 = 0.0;
            if ((rc = dotHost(c, c->kr, c->ks, &rz1, nullptr))) return rc;
            const double beta = rz1 / rz;
            rz = rz1;
            efem::dev::launchAxpby(c->kp, 1.0, c->ks, beta, c->kp, n, c->stream);
        }
    } else {
        // BiCGStab (van der Vorst) with M applied to p and s
        HIP_CHECK(c, hipMemcpyAsync(c->krh, c->kb, nb, hipMemcpyDeviceToDevice, c->stream));
        double rho = 1.0, alpha = 1.0, omega = 1.0;
        for (int it = 1; it <= prm.max_iterations; ++it) {
            double rho1 = 0.0, rr = 0.0;
            if ((rc = dotHost(c, c->krh, c->kr, &rho1, &rr))) return rc;
            if (!(rho1 != 0.0) || !std::isfinite(rho1)) break;
            if (it == 1) {
                HIP_CHECK(c, hipMemcpyAsync(c->kp, c->kr, nb, hipMemcpyDeviceToDevice, c->stream));
            } else {
                const double beta = (rho1 / rho) * (alpha / omega);
                efem::dev::launchAxpby(c->kp, 1.0, c->kp, -omega, c->kv, n, c->stream);
                efem::dev::launchAxpby(c->kp, 1.0, c->kr, beta, c->kp, n, c->stream);
            }
            if ((rc = precondApply(c, h0, c->kp, c->kph, st))) return rc;
            if ((rc = A(c->kph, c->kv))) return rc;
            double rv = 0.0;
            if ((rc = dotHost(c, c->krh, c->kv, &rv, nullptr))) return rc;
            if (!(rv != 0.0) || !std::isfinite(rv)) break;
            alpha = rho1 / rv;
            efem::dev::launchAxpby(c->ks, 1.0, c->kr, -alpha, c->kv, n, c->stream);
            double ss = 0.0;
            if ((rc = dotHost(c, c->ks, c->ks, &ss, nullptr))) return rc;
            st.iterations = it;
            if (std::sqrt(ss) <= stop) {
                efem::dev::launchAxpby(c->kx, 1.0, c->kx, alpha, c->kph, n, c->stream);
                st.rel_residual = std::sqrt(ss) / bnorm;
                st.converged = 1;
                break;
            }
            if ((rc = precondApply(c, h0, c->ks, c->ksh, st))) return rc;
            if ((rc = A(c->ksh, c->kt))) return rc;
            double ts = 0.0, tt = 0.0;
            if ((rc = dotHost(c, c->ks, c->kt, &ts, &tt))) return rc;
            if (!(tt > 0.0) || !std::isfinite(tt)) break;
            omega = ts / tt;
            efem::dev::launchAxpby(c->kx, 1.0, c->kx, alpha, c->kph, n, c->stream);
            efem::dev::launchAxpby(c->kx, 1.0, c->kx, omega, c->ksh, n, c->stream);
            efem::dev::launchAxpby(c->kr, 1.0, c->ks, -omega, c->kt, n, c->stream);
            rho = rho1;
            double r2 = 0.0;
            if ((rc = dotHost(c, c->kr, c->kr, &r2, nullptr))) return rc;
            st.rel_residual = std::sqrt(r2) / bnorm;
            if (std::sqrt(r2) <= stop) {
                st.converged = 1;
                break;
            }
            if (omega == 0.0) break;
        }
    }
    HIP_CHECK(c, hipGetLastError());
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return EFEM_OK;
}

// FP64 element tangent: plain Krylov.  FP32 element tangent (EFEM_F32_ELEMENTS):
// iterative refinement -- the FP32-tangent Krylov solve is the inner solver,
// the residual b - K x is recomputed with the FP64 tangent, so the result
// meets the FP64 tolerance.  The inner tolerance is 1e-2 with at most 200
// iterations: on smooth vectors the element sums cancel by orders of
// magnitude, which lifts the FP32 operator's relative noise to ~1e-4, where
// a tighter inner solve stagnates.
int krylovInternal(efem_ctx* c, int transpose, const efem_krylov_params& prm, efem_krylov_stats& st)
{
    if (!elem32(c)) return krylovCore(c, transpose, prm, st, true);
    const auto t0 = std::chrono::steady_clock::now();
    const int n = c->model.ndof();
    const size_t nb = static_cast<size_t>(n) * sizeof(double);
    int rc;
    double bb = 0.0;
    if ((rc = dotHost(c, c->kb, c->kb, &bb, nullptr))) return rc;
    const double bnorm = std::sqrt(bb);
    HIP_CHECK(c, hipMemcpyAsync(c->kbo, c->kb, nb, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(c, hipMemsetAsync(c->kxo, 0, nb, c->stream));
    efem_krylov_params inner = prm;
    inner.tol = std::max(prm.tol, 1e-2);
    st.rel_residual = 1.0;
    for (int outer = 0; outer < 30 && bnorm > 0.0; ++outer) {
        efem_krylov_stats is;
        std::memset(&is, 0, sizeof(is));
        inner.max_iterations = std::min(200, prm.max_iterations - st.iterations);
        if (inner.max_iterations <= 0) break;
        if ((rc = krylovCore(c, transpose, inner, is, false))) return rc;
        st.iterations += is.iterations;
        st.matvecs += is.matvecs;
        st.precond_applies += is.precond_applies;
        st.factor_seconds += is.factor_seconds;
        efem::dev::launchAxpby(c->kxo, 1.0, c->kxo, 1.0, c->kx, n, c->stream);
        // b - K x with the FP64 tangent
        if ((rc = tangentInternal(c, c->ku, c->kxo, c->kt, transpose, true))) return rc;
        ++st.matvecs;
        efem::dev::launchAxpby(c->kb, 1.0, c->kbo, -1.0, c->kt, n, c->stream);
        double rr = 0.0;
        if ((rc = dotHost(c, c->kb, c->kb, &rr, nullptr))) return rc;
        st.rel_residual = std::sqrt(rr) / bnorm;
        if (!std::isfinite(rr)) break;
        if (st.rel_residual <= prm.tol) {
            st.converged = 1;
            break;
        }
    }
    HIP_CHECK(c, hipMemcpyAsync(c->kx, c->kxo, nb, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->kb, c->kbo, nb, hipMemcpyDeviceToDevice, c->stream));
    if (bnorm == 0.0) st.converged = 1;
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return EFEM_OK;
}

int checkKrylovParams(efem_ctx* c, const efem_krylov_params* prm)
{
    if (!prm || !(prm->tol > 0.0) || prm->max_iterations <= 0 || prm->method < 0 ||
        prm->method > EFEM_KRYLOV_BICGSTAB) {
        c->err = "invalid Krylov parameters";
        return EFEM_ERR_ARG;
    }
    return EFEM_OK;
}

// device, stream and model upload of a context whose model is built
int initContext(std::unique_ptr<efem_ctx>& c)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_error = "efem_create: no HIP device";
        return EFEM_ERR_HIP;
    }
    if (hipGetDevice(&c->device) != hipSuccess || efem::dev::scratchInit(c->device) != 0) {
        g_error = "efem_create: device scratch of the allocation-free entry points unavailable";
        return EFEM_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g_error = "efem_create: hipStreamCreate failed";
        return EFEM_ERR_HIP;
    }
    int rc = uploadModel(c.get());
    if (rc) {
        g_error = c->err;
        efem_destroy(c.release());
        return rc;
    }
    return EFEM_OK;
}

}  // namespace

extern "C" {

int efem_create(efem_ctx** out, const efem_problem* pb)
{
    if (!out || !pb) {
        g_error = "efem_create: null argument";
        return EFEM_ERR_ARG;
    }
    *out = nullptr;
    auto c = std::make_unique<efem_ctx>();
    if (pb->precision < 0 || pb->precision > EFEM_MIXED) {
        g_error = "efem_create: unknown precision mode";
        return EFEM_ERR_ARG;
    }
    c->precision = pb->precision;
    std::string msg;
    try {
        if (!efem::buildModel(*pb, c->model, msg)) {
            g_error = msg;
            return EFEM_ERR_ARG;
        }
    } catch (const std::exception& e) {
        g_error = e.what();
        return EFEM_ERR_ARG;
    }
    int rc = initContext(c);
    if (rc) return rc;
    *out = c.release();
    return EFEM_OK;
}

int efem_destroy(efem_ctx* c)
{
    if (!c) return EFEM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    releaseFactor(c);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->pinned) (void)hipHostFree(c->pinned);
    for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->evsH0) (void)hipEventDestroy(e);
    if (c->hostScal) (void)hipHostFree(c->hostScal);
    for (hipEvent_t e : c->trialEv)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return EFEM_OK;
}

const char* efem_last_error(const efem_ctx* c) { return c ? c->err.c_str() : g_error.c_str(); }
const char* efem_global_error(void) { return g_error.c_str(); }
int efem_ndof(const efem_ctx* c) { return c ? c->model.ndof() : 0; }
int efem_precision(const efem_ctx* c) { return c ? c->precision : -1; }

int efem_residual(efem_ctx* c, const double* u, double* r)
{
    if (!c || !u || !r) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    const size_t n = static_cast<size_t>(c->model.ndof());
    HIP_CHECK(c, hipMemcpyAsync(c->tmp, u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = residualInternal(c, c->tmp, c->r, nullptr, nullptr, nullptr, nullptr, c->residualTerms);
    if (rc) return rc;
    HIP_CHECK(c, hipMemcpyAsync(r, c->r, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

int efem_residual_d(efem_ctx* c, const double* u_d, double* r_d)
{
    if (!c || !u_d || !r_d) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    const size_t n = static_cast<size_t>(c->model.ndof());
    HIP_CHECK(c, hipMemcpyAsync(c->tmp, u_d, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    int rc = residualInternal(c, c->tmp, r_d);
    if (rc) return rc;
    // the context stream is not ordered against the caller's streams:
    // r_d is complete (and u_d free for reuse) when this returns
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

int efem_factor(efem_ctx* c, int* failed)
{
    if (!c) return EFEM_ERR_ARG;
    return factorInternal(c, failed);
}

long long efem_factor_nnz(const efem_ctx* c) { return c && c->fac.ready ? c->fac.nnz : 0; }

int efem_factor_layout(const efem_ctx* c, long long* out8)
{
    if (!c || !out8) return EFEM_ERR_ARG;
    const DeviceFactor& f = c->fac;
    std::fill(out8, out8 + 8, 0LL);
    if (!f.ready) return EFEM_ERR_STATE;
    int launches = f.fusedTop ? 1 : 0;
    for (int l = 0; l < f.nlevels; ++l) launches += (f.fwdCnt[l] > 0) + (f.bwdCnt[l] > 0);
    out8[0] = f.nnz;
    out8[1] = f.sweepEntries;
    out8[2] = f.fusedTop ? f.topN : 0;
    out8[3] = f.fusedTop ? f.topLd : 0;
    out8[4] = f.fusedTop ? f.topDepth : -1;
    out8[5] = launches;
    out8[6] = f.nlevels;
    out8[7] = f.nsuper;
    return EFEM_OK;
}

int efem_factor_phases(const efem_ctx* c, double* out8)
{
    if (!c || !out8) return EFEM_ERR_ARG;
    std::copy(c->factorPhases, c->factorPhases + 8, out8);
    return EFEM_OK;
}

int efem_factor_release(efem_ctx* c)
{
    if (!c) return EFEM_ERR_ARG;
    // frees the factor's device buffers; the next solve factors again
    releaseFactor(c);
    return EFEM_OK;
}

// Constraint::updateDx (src/fe/Constraint.cpp:103-113): the solved
// displacement becomes the configuration the next solve's increment is
// measured from
int efem_commit(efem_ctx* c, const double* u)
{
    if (!c) return EFEM_ERR_ARG;
    if (!c->parts.empty()) {
        c->err = "efem_commit: batch context";
        return EFEM_ERR_STATE;
    }
    if (!u) {
        c->hasBase = false;
        return EFEM_OK;
    }
    const size_t n = static_cast<size_t>(c->model.ndof());
    int rc;
    if (!c->base && (rc = devAlloc(c, &c->base, n + 3))) return rc;
    HIP_CHECK(c, hipMemcpyAsync(c->base, u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->hasBase = true;
    return EFEM_OK;
}

int efem_committed(efem_ctx* c, double* u)
{
    if (!c || !u) return EFEM_ERR_ARG;
    const size_t n = static_cast<size_t>(c->model.ndof());
    if (!c->hasBase) {
        std::fill(u, u + n, 0.0);
        return EFEM_OK;
    }
    HIP_CHECK(c, hipMemcpyAsync(u, c->base, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

int efem_write_vtk(const efem_problem* pb, const double* u, const char* path)
{
    if (!pb || !path) return EFEM_ERR_ARG;
    std::string msg;
    if (!efem::writeVtk(*pb, u, path, msg)) {
        g_error = msg;
        return EFEM_ERR_ARG;
    }
    return EFEM_OK;
}

int efem_results_path(const char* working_dir, const char* name, char* out, int cap)
{
    if (!working_dir || !name || !out || cap <= 0) return EFEM_ERR_ARG;
    std::string path, msg;
    if (!efem::resultsPath(working_dir, name, path, msg)) {
        g_error = msg;
        return EFEM_ERR_ARG;
    }
    if (static_cast<int>(path.size()) >= cap) {
        g_error = "efem_results_path: buffer too small (" + std::to_string(path.size() + 1) + " bytes needed)";
        return EFEM_ERR_ARG;
    }
    std::memcpy(out, path.c_str(), path.size() + 1);
    return EFEM_OK;
}

int efem_apply_h0(efem_ctx* c, const double* v, double* out)
{
    if (!c || !v || !out) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    if (!c->fac.ready) {
        c->err = "efem_apply_h0: no factor (call efem_factor first)";
        return EFEM_ERR_STATE;
    }
    const size_t n = static_cast<size_t>(c->model.ndof());
    HIP_CHECK(c, hipMemcpyAsync(c->rp, v, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = applyH0Internal(c, c->rp, c->tmp);
    if (rc) return rc;
    HIP_CHECK(c, hipMemcpyAsync(out, c->tmp, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

// LBFGSSolver::solve (LBFGS.h:96-174), two-loop LBFGSMat::apply_Hv
// (LBFGSMat.h:103-131), projection line search (LineSearchFEA.h:37-68)
int efem_solve(efem_ctx* c, double* u, const efem_lbfgs_params* prm, efem_solve_stats* st)
{
    if (!c || !u || !prm) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    if (!c->parts.empty()) {
        c->err = "efem_solve: batch context, use efem_batch_solve";
        return EFEM_ERR_STATE;
    }
    if (prm->m <= 0 || prm->m > 64 || prm->max_linesearch <= 0 || prm->epsilon < 0 || prm->epsilon_rel < 0 ||
        prm->max_iterations < 0) {
        c->err = "efem_solve: invalid parameters (Param.h:110-127)";
        return EFEM_ERR_ARG;
    }
    if (prm->line_search != EFEM_LS_PROJECTION && prm->line_search != EFEM_LS_ENERGY) {
        c->err = "efem_solve: unknown line-search rule";
        return EFEM_ERR_ARG;
    }
    const bool energyLs = prm->line_search == EFEM_LS_ENERGY;
    if (energyLs && c->model.nFaces > 0) {
        c->err = "efem_solve: the energy line search needs conservative loads (no pressure faces)";
        return EFEM_ERR_ARG;
    }
    efem_solve_stats local;
    efem_solve_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof(s));
    const auto t0 = std::chrono::steady_clock::now();
    const int n = c->model.ndof();
    const int mc = prm->m;
    int rc;
    if ((rc = growHistory(c, mc))) return rc;
    c->trace.clear();
    double* sc = c->scal;
    c->elemFp64 = false;
    if (prm->x_target) {
        if (!c->xtarget && (rc = devAlloc(c, &c->xtarget, static_cast<size_t>(n) + 3))) return rc;
        HIP_CHECK(c, hipMemcpyAsync(c->xtarget, prm->x_target, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        efem::dev::launchDot2(c->xtarget, c->xtarget, n, c->work, sc + kOut + 4, sc + kOut + 5, c->stream);
    }
    HIP_CHECK(c, hipMemcpyAsync(c->x, u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = residualInternal(c, c->x, c->r))) return rc;
    s.r_evals = 1;
    if (energyLs) {
        // strain energies of the start, per element (FP64 in every mode)
        const size_t ne = static_cast<size_t>(std::max(c->model.nElems, 1));
        if (!c->wOld && ((rc = devAlloc(c, &c->wOld, ne)) || (rc = devAlloc(c, &c->wNew, ne)))) return rc;
        efem::dev::launchElementEnergy(elemArgs(c, c->x, false), c->wOld, c->stream);
    }
    // the relative criterion measures the solution from the committed
    // configuration: the reference's solver variable is the increment dx
    // on top of node->dx (Constraint.cpp:93-113)
    const double* base = c->hasBase ? c->base : nullptr;
    if (base) {
        efem::dev::launchAxpby(c->tmp, 1.0, c->x, -1.0, base, n, c->stream);
        efem::dev::launchDotMax(c->tmp, c->tmp, c->r, n, c->work, sc + kOut, c->stream);
    } else {
        efem::dev::launchDotMax(c->x, c->x, c->r, n, c->work, sc + kOut, c->stream);
    }
    if ((rc = readScalars(c, 6))) return rc;
    double rinf = c->hostScal[2], xinf = c->hostScal[1];
    const double xnorm2 = std::sqrt(c->hostScal[0]);
    const double tnorm = std::sqrt(c->hostScal[5]);
    s.rinf = rinf;
    s.xinf = xinf;
    const double rinf0 = rinf;
    if (rinf <= prm->epsilon || rinf <= prm->epsilon_rel * xnorm2) {
        s.iterations = 1;
        s.solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return EFEM_OK;
    }
    if (prm->use_precond && !c->fac.ready) {
        int failed = 0;
        if ((rc = factorInternal(c, &failed))) return rc;
        s.factor_seconds = c->factorSeconds;
    }
    const bool h0 = prm->use_precond && c->fac.ready;
    s.precond_failed = prm->use_precond && !h0;
    s.factor_nnz = h0 ? c->fac.nnz : 0;
    // d = -H0 r
    if (h0) {
        if ((rc = applyH0Internal(c, c->r, c->d))) return rc;
        efem::dev::launchScale(c->d, c->d, -1.0, n, c->stream);
    } else {
        efem::dev::launchScale(c->d, c->r, -1.0, n, c->stream);
    }
    int ncorr = 0, ptr = mc, k = 1;
    const size_t nb = static_cast<size_t>(n) * sizeof(double);
    // the first line search's slope r . d (later ones come with d)
    efem::dev::launchDotMax(c->r, c->d, nullptr, n, c->work, sc + kOut + 4, c->stream);
    const bool compact = c->compactTwoLoop && mc <= efem::dev::kMaxCompact;
    const bool staged = h0 && c->useGraph;

    // add_correction (LBFGSMat.h:81-98) of the pair (x - xp, r - rp), then
    // the next direction d2 = -H r (LBFGSMat::apply_Hv) and its slope r . d2
    // (kOut + 4).  Writes S/Y slot loc, the Gram row / column of loc, the
    // scalar slots and d2; reads x, xp, r, rp and the rest of the history.
    bool preformed = false;  // the next first trial's x already formed
    // A first trial's finishing reduction (its scalars into the pinned
    // slot) is deferred when the look-ahead may form the next direction
    // right behind it: it then runs as one extra block of that direction's
    // Gram / first-loop sweep, which records the trial's event; otherwise it
    // is launched on its own before the host waits for the trial.  A
    // speculative trial that is rolled back drops it.
    efem::dev::TrialFinish pendingTF{};
    int pendingSlot = 0;
    auto flushPending = [&]() -> int {
        if (pendingTF.k) {
            pendingTF.ev = c->trialEv[pendingSlot];
            efem::dev::launchTrialFinish(pendingTF, c->stream);
            pendingTF = efem::dev::TrialFinish{};
            HIP_CHECK(c, hipGetLastError());
        }
        return EFEM_OK;
    };
    auto nextDirection = [&]() -> int {
        preformed = false;
        if (!compact) {
            int rcf = flushPending();
            if (rcf) return rcf;
        }
        const int loc = ptr % mc;
        double* sl = c->S + static_cast<size_t>(loc) * n;
        double* yl = c->Y + static_cast<size_t>(loc) * n;
        if (ncorr < mc) ++ncorr;
        ptr = loc + 1;
        int rc2;
        if (compact) {
            // the new pair, its Gram row / column, theta and the first loop
            // in one pass; with the captured preconditioner the first pass
            // writes its staging input and the second reads its output
            // with launched (not captured) preconditioner sweeps the pass
            // writes the permuted right-hand side the sweeps read
            const bool permuted = h0 && !staged;
            double* q0 = staged ? c->fac.in : permuted ? c->fac.bp : c->d2;
            efem::dev::TrialFinish ride = pendingTF;
            ride.ev = ride.k ? c->trialEv[pendingSlot] : nullptr;
            pendingTF = efem::dev::TrialFinish{};
            efem::dev::launchGramFirst(c->S, c->Y, c->x, c->xp, c->r, c->rp, n, ncorr, loc, mc, sc + kGram,
                                       sc + kYs, sc + kYY + loc, sc + kTheta, sc + kAlpha, sc + kRS, q0, c->work,
                                       c->stream, permuted ? c->fac.pinv : nullptr, ride.k ? &ride : nullptr);
            efem::dev::HSrc h{c->d2};
            if (staged) {
                if ((rc2 = applyH0Graph(c))) return rc2;
                h = efem::dev::HSrc{c->fac.out};
            } else if (h0) {
                // the second loop reads the backward sweep's solution through
                // the inverse permutation: no unpermute launch
                if ((rc2 = applyH0Internal(c, nullptr, nullptr, true, false))) return rc2;
                h = efem::dev::HSrc{nullptr, c->fac.pinv, c->fac.z, fac32(c) ? 1 : 0};
            } else {
                efem::dev::launchDivDev(c->d2, sc + kTheta, n, c->stream);
            }
            // the next first trial x + 1 d goes straight into the spare that
            // becomes x after rotate()
            efem::dev::launchTwoLoopSecond(h, c->S, c->Y, n, ncorr, loc, mc, sc + kGram, sc + kYs, sc + kAlpha,
                                           sc + kCoef, c->d2, c->work, c->stream, c->r, sc + kRS, sc + kOut + 4,
                                           c->x, c->x3);
            preformed = true;
        } else {
            efem::dev::launchDiff2(sl, yl, c->x, c->xp, c->r, c->rp, n, c->stream);
            efem::dev::launchDot2(sl, yl, n, c->work, sc + kYs + loc, sc + kYY + loc, c->stream);
            efem::dev::launchTheta(sc + kYs + loc, sc + kYY + loc, sc + kTheta, c->stream);
            // two-loop recursion in the reference's order
            efem::dev::launchScale(c->d2, c->r, -1.0, n, c->stream);
            int j = ptr % mc, prev = -1;
            for (int i = 0; i < ncorr; ++i) {
                j = (j + mc - 1) % mc;
                efem::dev::launchTwoLoopStep(c->d2, prev >= 0 ? c->Y + static_cast<size_t>(prev) * n : nullptr,
                                             prev >= 0 ? sc + kAlpha + prev : nullptr,
                                             c->S + static_cast<size_t>(j) * n, sc + kYs + j, nullptr,
                                             sc + kAlpha + j, 0, n, c->work, c->stream);
                prev = j;
            }
            if (prev >= 0)
                efem::dev::launchAxpyDev(c->d2, c->Y + static_cast<size_t>(prev) * n, sc + kAlpha + prev, 1, n,
                                         c->stream);
            if (h0) {
                if ((rc2 = applyH0Internal(c, c->d2, c->tmp))) return rc2;
                std::swap(c->d2, c->tmp);
            } else {
                efem::dev::launchDivDev(c->d2, sc + kTheta, n, c->stream);
            }
            prev = -1;
            for (int i = 0; i < ncorr; ++i) {
                efem::dev::launchTwoLoopStep(c->d2, prev >= 0 ? c->S + static_cast<size_t>(prev) * n : nullptr,
                                             prev >= 0 ? sc + kCoef + prev : nullptr,
                                             c->Y + static_cast<size_t>(j) * n, sc + kYs + j, sc + kAlpha + j,
                                             sc + kCoef + j, 1, n, c->work, c->stream);
                prev = j;
                j = (j + 1) % mc;
            }
            if (prev >= 0)
                efem::dev::launchAxpyDev(c->d2, c->S + static_cast<size_t>(prev) * n, sc + kCoef + prev, 0, n,
                                         c->stream);
            efem::dev::launchDotMax(c->r, c->d2, nullptr, n, c->work, sc + kOut + 4, c->stream);
        }
        HIP_CHECK(c, hipGetLastError());
        return EFEM_OK;
    };
    // the current iterate and residual become xp / rp; the trial writes the
    // spare buffers and the previous xp / rp become the spares (every trial
    // step writes all of x, every residual all of r; all keep the zero dummy
    // node)
    auto rotate = [&]() {
        double* t = c->xp;
        c->xp = c->x;
        c->x = c->x3;
        c->x3 = t;
        t = c->rp;
        c->rp = c->r;
        c->r = c->r3;
        c->r3 = t;
    };
    // one trial x = xp + step d, r(x), and its scalars {r . d, ||r||_inf,
    // ||x - u_c||_inf, ||x - x_target||^2, slope r . d of the iteration}
    // copied to pinned slot `slot` behind an event
    auto enqueueTrial = [&](double step, int slot) -> int {
        int rc2 = flushPending();
        if (rc2) return rc2;
        if (!(preformed && step == 1.0)) efem::dev::launchStep(c->x, c->xp, c->d, step, n, c->stream);
        preformed = false;
        // the trial's scalars come with the residual's gather; without the
        // energy rule its finishing block writes them, and the slope, into
        // the pinned slot itself
        double* pinnedSlot = c->hostScalDev + kTrialSlot + 8 * slot;
        const bool defer = !energyLs && c->lookahead && compact;
        efem::dev::TrialFinish tf{};
        rc2 = residualInternal(c, c->x, c->r, c->d, base, prm->x_target ? c->xtarget : nullptr,
                               energyLs ? sc + kOut : pinnedSlot, kTermsAll, energyLs ? nullptr : sc + kOut + 4,
                               defer ? &tf : nullptr);
        if (rc2) return rc2;
        if (tf.k) {
            // the event is recorded behind the launch that finishes it
            pendingTF = tf;
            pendingSlot = slot;
            return EFEM_OK;
        }
        if (energyLs) {
            // the potential's change from xp and its rounding scale into
            // slots 5, 6: per-element energy differences (compensated)
            efem::dev::launchElementEnergy(elemArgs(c, c->x, false), c->wNew, c->stream);
            efem::dev::launchEnergyDiff(c->wNew, c->wOld, c->model.nElems, c->fext, c->x, c->xp, n, c->work,
                                        sc + kOut + 5, c->stream);
        }
        if (energyLs)
            HIP_CHECK(c, hipMemcpyAsync(c->hostScal + kTrialSlot + 8 * slot, sc + kOut, 7 * sizeof(double),
                                        hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(c, hipEventRecord(c->trialEv[slot], c->stream));
        return EFEM_OK;
    };
    // Look-ahead: while the host waits for a first trial's scalars, the
    // device already forms the next direction and runs the next first trial
    // (step 1).  Nearly every first trial is accepted and the solve goes on;
    // otherwise (a rejected trial, a stop, the FP64 switch) the host state is
    // restored and the speculative results, which only touched the spares,
    // the S/Y slot and Gram row / column about to be replaced and scalar
    // slots rewritten before their next use, are ignored.  Either way the
    // arithmetic is that of the plain loop.
    struct LoopState {
        double *x, *xp, *x3, *r, *rp, *r3, *d, *d2, *tmp;
        int ptr, ncorr;
    };
    auto save = [&]() { return LoopState{c->x, c->xp, c->x3, c->r, c->rp, c->r3, c->d, c->d2, c->tmp, ptr, ncorr}; };
    auto restore = [&](const LoopState& q) {
        c->x = q.x;
        c->xp = q.xp;
        c->x3 = q.x3;
        c->r = q.r;
        c->rp = q.rp;
        c->r3 = q.r3;
        c->d = q.d;
        c->d2 = q.d2;
        c->tmp = q.tmp;
        ptr = q.ptr;
        ncorr = q.ncorr;
    };
    rotate();
    if ((rc = enqueueTrial(1.0, 0))) return rc;
    int slot = 0, it = 0, trials = 0;
    double cur = 1.0, dgInit = 0.0;
    for (;;) {
        bool spec = false;
        LoopState before{};
        if (c->lookahead && !energyLs && it == 0 && (prm->max_iterations == 0 || k < prm->max_iterations)) {
            before = save();
            if ((rc = nextDirection())) return rc;
            std::swap(c->d, c->d2);
            rotate();
            if ((rc = enqueueTrial(1.0, slot ^ 1))) return rc;
            spec = true;
        }
        if (pendingTF.k && pendingSlot == slot && (rc = flushPending())) return rc;
        HIP_CHECK(c, hipEventSynchronize(c->trialEv[slot]));
        const double* hsc = c->hostScal + kTrialSlot + 8 * slot;
        const double dg = hsc[0];
        rinf = hsc[1];
        xinf = hsc[2];
        if (it == 0) dgInit = hsc[4];
        ++s.r_evals;
        ++trials;
        const double used = cur;
        const bool proj = std::fabs(dg) <= prm->lns_fea * std::fabs(dgInit);
        // energy rule (no counterpart in the reference): sufficient decrease
        // of the potential, or a change within its rounding, as well
        const bool armijo = !energyLs || hsc[5] <= kArmijoC1 * used * dgInit ||
                            std::fabs(hsc[5]) <= kEnergyRtol * hsc[6];
        if (!(proj && armijo) && it + 1 < prm->max_linesearch) {
            // rejected: the next trial of this iteration
            if (spec) {
                restore(before);
                pendingTF = efem::dev::TrialFinish{};  // the speculative trial's
            }
            cur *= !armijo ? 0.5 : (dg * dgInit > 0 ? 1.5 : 0.5);
            ++it;
            if ((rc = enqueueTrial(cur, slot))) return rc;
            continue;
        }
        // accepted: its energies are the next iterations' reference
        if (energyLs) std::swap(c->wOld, c->wNew);
        // accepted (or the last trial allowed: LineSearchFEA.h:37-68)
        s.ls_trials += trials;
        c->trace.push_back(rinf);
        c->trace.push_back(used);
        c->trace.push_back(trials);
        s.iterations = k;
        s.rinf = rinf;
        s.xinf = xinf;
        bool stop = !std::isfinite(rinf) || !std::isfinite(xinf);
        if (!stop && prm->x_target) {
            // ||x - x_target|| of the accepted step (read with its scalars)
            s.target_err = std::sqrt(hsc[3]) / tnorm;
            if (s.target_err <= prm->target_tol) {
                s.target_reached = 1;
                stop = true;
            }
        }
        if (!stop && (rinf <= prm->epsilon || rinf <= prm->epsilon_rel * xinf)) stop = true;
        if (!stop && prm->max_iterations != 0 && k >= prm->max_iterations) stop = true;
        const bool toFp64 = !stop && elem32(c) && !c->elemFp64 && prm->fp64_switch > 0.0 &&
                            rinf <= prm->fp64_switch * rinf0;
        if (spec && (stop || toFp64)) {
            restore(before);
            pendingTF = efem::dev::TrialFinish{};  // the speculative trial's
            spec = false;
        }
        if (stop) break;
        if (toFp64) {
            // FP64 finish: re-evaluate r(x) with the FP64 element kernels
            c->elemFp64 = true;
            s.fp64_switch_iteration = k;
            if ((rc = residualInternal(c, c->x, c->r))) return rc;
            ++s.r_evals;
        }
        ++k;
        it = 0;
        trials = 0;
        cur = 1.0;
        slot ^= 1;
        if (spec) continue;  // the next first trial is already in flight
        if ((rc = nextDirection())) return rc;
        std::swap(c->d, c->d2);
        rotate();
        if ((rc = enqueueTrial(1.0, slot))) return rc;
    }
    pendingTF = efem::dev::TrialFinish{};
    HIP_CHECK(c, hipMemcpyAsync(u, c->x, nb, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->elemFp64 = false;
    s.solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!std::isfinite(s.rinf) || !std::isfinite(s.xinf)) {
        // Integrator.cpp:28-34: non-finite result -> dx zeroed, failure reported
        std::memset(u, 0, nb);
        c->err = "L-BFGS produced a non-finite residual or displacement";
        return EFEM_ERR_NONFINITE;
    }
    return EFEM_OK;
}

const char* efem_host_dense_kernel(void) { return efem::dense::kernelName(); }

int efem_host_constant_force(const efem_problem* pb, double* out, int n)
{
    if (!pb || !out) return EFEM_ERR_ARG;
    efem::Model m;
    std::string msg;
    try {
        if (!efem::buildModel(*pb, m, msg)) {
            g_error = msg;
            return EFEM_ERR_ARG;
        }
    } catch (const std::exception& e) {
        g_error = e.what();
        return EFEM_ERR_ARG;
    }
    if (n != m.ndof()) {
        g_error = "efem_host_constant_force: n must be 3 x the free nodes (" + std::to_string(m.ndof()) + ")";
        return EFEM_ERR_ARG;
    }
    std::copy(m.fext.begin(), m.fext.end(), out);
    return EFEM_OK;
}

int efem_host_jacobi(const efem_problem* pb, long long* nnz, int* rows, int* cols, double* vals, long long cap)
{
    if (!pb || !nnz) return EFEM_ERR_ARG;
    efem::Model m;
    std::string msg;
    efem::Csc L;
    try {
        if (!efem::buildModel(*pb, m, msg)) {
            g_error = msg;
            return EFEM_ERR_ARG;
        }
        efem::scalarLaplacian(m, L);
    } catch (const std::exception& e) {
        g_error = e.what();
        return EFEM_ERR_ARG;
    }
    // every 3 x 3 block of a nonzero L_ij: L_ij on its diagonal, explicit
    // zeros off it (the pattern ElementConstraint.cpp:193-204 pushes)
    const long long need = 9LL * static_cast<long long>(L.i.size());
    *nnz = need;
    if (!rows || !cols || !vals) return EFEM_OK;
    if (cap < need) {
        g_error = "efem_host_jacobi: cap < " + std::to_string(need);
        return EFEM_ERR_ARG;
    }
    long long q = 0;
    for (int j = 0; j < L.n; ++j)
        for (int p = L.p[j]; p < L.p[j + 1]; ++p)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    rows[q] = 3 * L.i[p] + a;
                    cols[q] = 3 * j + b;
                    vals[q] = a == b ? L.x[p] : 0.0;
                    ++q;
                }
    return EFEM_OK;
}

int efem_device_memory(long long* free_bytes, long long* total_bytes)
{
    if (!free_bytes || !total_bytes) return EFEM_ERR_ARG;
    size_t f = 0, t = 0;
    const hipError_t e = hipMemGetInfo(&f, &t);
    if (e != hipSuccess) {
        g_error = std::string("hipMemGetInfo: ") + hipGetErrorString(e);
        return EFEM_ERR_HIP;
    }
    *free_bytes = static_cast<long long>(f);
    *total_bytes = static_cast<long long>(t);
    return EFEM_OK;
}

int efem_host_factor_solve(const efem_problem* pb, const double* b, double* x, long long* info)
{
    if (!pb || !b || !x) return EFEM_ERR_ARG;
    efem::Model m;
    std::string msg;
    try {
        if (!efem::buildModel(*pb, m, msg)) {
            g_error = msg;
            return EFEM_ERR_ARG;
        }
        if (m.nFree == 0) {
            g_error = "no free DOF";
            return EFEM_ERR_STATE;
        }
        const auto t0 = std::chrono::steady_clock::now();
        efem::Csc A;
        efem::scalarLaplacian(m, A);
        const auto t1 = std::chrono::steady_clock::now();
        const std::vector<int> perm = efem::nestedDissection(A, efem::freeCoords(m));
        const auto t2 = std::chrono::steady_clock::now();
        efem::SupernodalFactor F;
        if (!efem::supernodalCholesky(A, perm, F, msg)) {
            g_error = msg;
            return EFEM_ERR_STATE;
        }
        efem::blockInverseSolve3(F, b, x);
        if (info) {
            auto us = [](auto a, auto b) { return static_cast<long long>(std::chrono::duration<double>(b - a).count() * 1e6); };
            info[0] = F.nnzL;
            info[1] = F.nsuper;
            info[2] = F.nlevels;
            info[3] = static_cast<long long>(F.flops);
            info[4] = static_cast<long long>(F.seconds * 1e6);
            info[5] = static_cast<long long>(F.numericSeconds * 1e6);
            info[6] = us(t0, t1);
            info[7] = us(t1, t2);
        }
    } catch (const std::exception& e) {
        g_error = e.what();
        return EFEM_ERR_ARG;
    }
    return EFEM_OK;
}

int efem_set_option(efem_ctx* c, const char* name, double value)
{
    if (!c || !name) return EFEM_ERR_ARG;
    const std::string k(name);
    if (k == "h0_graph") {
        c->useGraph = value != 0.0;
        return EFEM_OK;
    }
    if (k == "two_loop_compact") {
        c->compactTwoLoop = value != 0.0;
        return EFEM_OK;
    }
    if (k == "lookahead") {
        c->lookahead = value != 0.0;
        return EFEM_OK;
    }
    if (k == "device_laplacian") {
        // assemble the approximate Jacobian on the device (1) or the host (0);
        // the same values either way; takes effect at the next factorisation
        c->devLaplacian = value != 0.0;
        c->symbolic.reset();
        releaseFactor(c);
        return EFEM_OK;
    }
    if (k == "reuse_symbolic") {
        // keep the ordering and symbolic analysis between factorisations (1)
        // or recompute them every time (0)
        c->reuseSymbolic = value != 0.0;
        c->symbolic.reset();
        releaseFactor(c);
        return EFEM_OK;
    }
    if (k == "device_factor") {
        // numeric factorisation on the device (1) or the host (0); takes
        // effect at the next factorisation
        c->devFactorOpt = value != 0.0;
        releaseFactor(c);
        return EFEM_OK;
    }
    if (k == "fused_root") {
        // the top of the elimination tree (at least the roots) solved as one
        // launch z_T = Z f_T; takes effect at the next factorisation
        c->fusedRootOpt = value != 0.0;
        releaseFactor(c);
        return EFEM_OK;
    }
    if (k == "fwd_tables") {
        // per-item extend-add tables for the forward sweep; takes effect at
        // the next factorisation
        c->fwdTables = value != 0.0;
        releaseFactor(c);
        return EFEM_OK;
    }
    const int iv = static_cast<int>(value);
    if (k == "residual_terms" && iv >= 1 && iv <= 7 && static_cast<double>(iv) == value) {
        // efem_residual evaluates only these terms (1 element internal
        // forces, 2 pressure faces, 4 constant loads): one constraint's part
        // of the residual; solves always use all terms
        c->residualTerms = iv;
        return EFEM_OK;
    }
    if (k == "gather_lanes" && (iv == 1 || iv == 2 || iv == 4 || iv == 8)) {
        c->tuning.gatherLanes = iv;
        return EFEM_OK;
    }
    if (k == "reduce_blocks" && iv >= 1 && iv <= efem::dev::kReduceBlocks) {
        c->tuning.reduceBlocks = iv;
        return EFEM_OK;
    }
    if (k == "tet4_threads" && (iv == 64 || iv == 128 || iv == 256)) {
        c->tuning.tet4Threads = iv;
        return EFEM_OK;
    }
    if (k == "history_blocks" && iv >= 1 && iv <= efem::dev::kReduceBlocks) {
        c->tuning.historyBlocks = iv;
        return EFEM_OK;
    }
    if (k == "batch_threads" && (iv == 256 || iv == 512 || iv == 1024)) {
        c->tuning.batchThreads = iv;
        return EFEM_OK;
    }
    if (k == "bwd_narrow_items" && value >= 1.0) {
        // levels with at least this many 4-column backward items use
        // 16-column items; takes effect at the next factorisation
        c->narrowMinItems = value >= 2147483647.0 ? 2147483647 : iv;
        releaseFactor(c);
        return EFEM_OK;
    }
    if ((k == "wide_fwd_items" || k == "wide_bwd_items") && value >= 0.0) {
        // takes effect at the next factorisation
        (k == "wide_fwd_items" ? c->wideFwdItems : c->wideBwdItems) =
            value >= 2147483647.0 ? 2147483647 : iv;
        releaseFactor(c);
        return EFEM_OK;
    }
    if (k == "bwd_stage" && (iv == 0 || iv == 4 || iv == 8 || iv == 16) && static_cast<double>(iv) == value) {
        // takes effect at the next factorisation
        c->bwdStage = iv;
        releaseFactor(c);
        return EFEM_OK;
    }
    if (k == "tet4_patches") {
        c->tet4Patches = value != 0.0;
        return EFEM_OK;
    }
    if (k == "batch_dense_h0") {
        // takes effect at the next factorisation
        c->batchDense = value != 0.0;
        releaseFactor(c);
        return EFEM_OK;
    }
    if (k == "patch_gather_lanes" && (iv == 1 || iv == 2 || iv == 4 || iv == 8)) {
        c->patchLanes = iv;
        return EFEM_OK;
    }
    if (k == "tet4_geometry") {
        c->tet4Geometry = value != 0.0;
        return EFEM_OK;
    }
    if (k == "top_waves" && (iv == 8 || iv == 16) && static_cast<double>(iv) == value) {
        c->topWaves = iv;
        return EFEM_OK;
    }
    if (k == "fused_top_max" && iv >= 1 && iv <= 65536 && static_cast<double>(iv) == value) {
        // columns the fused top may hold: it takes the deepest depth whose
        // supernodes, with all shallower ones, fit (none if the roots alone
        // do not); takes effect at the next factorisation
        c->fusedTopMax = iv;
        releaseFactor(c);
        return EFEM_OK;
    }
    if (k == "dense_subtree" && iv >= 1 && iv <= efem::kMaxSupernodeWidth) {
        // elimination subtrees of at most this many columns become one dense
        // supernode; takes effect at the next factorisation
        c->denseSubtree = iv;
        c->symbolic.reset();
        releaseFactor(c);
        return EFEM_OK;
    }
    c->err = "efem_set_option: unknown option or value " + k;
    return EFEM_ERR_ARG;
}

int efem_trace(const efem_ctx* c, double* out, int cap)
{
    if (!c) return 0;
    const int n = static_cast<int>(c->trace.size() / 3);
    if (out)
        for (int i = 0; i < std::min(n, cap); ++i)
            for (int k = 0; k < 3; ++k) out[3 * i + k] = c->trace[3 * static_cast<size_t>(i) + k];
    return n;
}

int efem_profile_reset(efem_ctx* c)
{
    if (!c) return EFEM_ERR_ARG;
    if (c->evs.empty()) {
        c->evs.resize(8192);
        for (auto& e : c->evs) HIP_CHECK(c, hipEventCreate(&e));
        c->evsH0.resize(8192);
        for (auto& e : c->evsH0) HIP_CHECK(c, hipEventCreate(&e));
    }
    c->evUsed = 0;
    c->evH0Used = 0;
    c->profiling = true;
    return EFEM_OK;
}

int efem_profile_read_h0(efem_ctx* c, double* ms, int* applies)
{
    if (!c) return EFEM_ERR_ARG;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (int i = 0; i + 1 < c->evH0Used; i += 2) {
        float t = 0.f;
        HIP_CHECK(c, hipEventElapsedTime(&t, c->evsH0[i], c->evsH0[i + 1]));
        tot += t;
    }
    if (ms) *ms = tot;
    if (applies) *applies = c->evH0Used / 2;
    return EFEM_OK;
}

int efem_profile_read(efem_ctx* c, double* ms, int* launches)
{
    if (!c) return EFEM_ERR_ARG;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (int i = 0; i + 1 < c->evUsed; i += 2) {
        float t = 0.f;
        HIP_CHECK(c, hipEventElapsedTime(&t, c->evs[i], c->evs[i + 1]));
        tot += t;
    }
    if (ms) *ms = tot;
    if (launches) *launches = c->evUsed / 2;
    return EFEM_OK;
}

int efem_tangent(efem_ctx* c, const double* u, const double* v, double* out, int transpose)
{
    if (!c || !u || !v || !out) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    int rc = ensureKrylov(c);
    if (rc) return rc;
    const size_t nb = static_cast<size_t>(c->model.ndof()) * sizeof(double);
    HIP_CHECK(c, hipMemcpyAsync(c->ku, u, nb, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->kp, v, nb, hipMemcpyHostToDevice, c->stream));
    if ((rc = tangentInternal(c, c->ku, c->kp, c->kv, transpose))) return rc;
    HIP_CHECK(c, hipMemcpyAsync(out, c->kv, nb, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

int efem_tangent_solve(efem_ctx* c, const double* u, const double* b, double* x, int transpose,
                       const efem_krylov_params* prm, efem_krylov_stats* st)
{
    if (!c || !u || !b || !x) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    int rc = checkKrylovParams(c, prm);
    if (rc) return rc;
    efem_krylov_stats local;
    efem_krylov_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof(s));
    if ((rc = ensureKrylov(c))) return rc;
    const size_t nb = static_cast<size_t>(c->model.ndof()) * sizeof(double);
    HIP_CHECK(c, hipMemcpyAsync(c->ku, u, nb, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->kb, b, nb, hipMemcpyHostToDevice, c->stream));
    if ((rc = krylovInternal(c, transpose, *prm, s))) return rc;
    HIP_CHECK(c, hipMemcpyAsync(x, c->kx, nb, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

int efem_newton(efem_ctx* c, double* u, double epsilon, int max_iterations, const efem_krylov_params* kp,
                efem_solve_stats* st)
{
    if (!c || !u || max_iterations <= 0 || !(epsilon >= 0.0)) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    int rc = checkKrylovParams(c, kp);
    if (rc) return rc;
    efem_solve_stats local;
    efem_solve_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof(s));
    if ((rc = ensureKrylov(c))) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    const int n = c->model.ndof();
    const size_t nb = static_cast<size_t>(n) * sizeof(double);
    double* sc = c->scal;
    c->elemFp64 = false;
    c->trace.clear();
    HIP_CHECK(c, hipMemcpyAsync(c->x, u, nb, hipMemcpyHostToDevice, c->stream));
    auto resNorms = [&](const double* x, double* r, double* r2, double* rinf, double* xinf) {
        int e = residualInternal(c, x, r);
        if (e) return e;
        ++s.r_evals;
        efem::dev::launchDotMax(r, r, x, n, c->work, sc + kOut, c->stream);
        if ((e = readScalars(c, 3))) return e;
        *r2 = c->hostScal[0];
        *rinf = c->hostScal[1];
        *xinf = c->hostScal[2];
        return EFEM_OK;
    };
    double r2 = 0.0, rinf = 0.0, xinf = 0.0;
    if ((rc = resNorms(c->x, c->r, &r2, &rinf, &xinf))) return rc;
    double r2prev = r2;
    for (int k = 1; k <= max_iterations && rinf > epsilon && std::isfinite(rinf); ++k) {
        // K_t(x) dx = -r, solved inexactly: forcing term eta_k =
        // 0.9 (||r_k|| / ||r_k-1||)^2 in [tol, 1e-2] (Eisenstat-Walker choice 2)
        efem_krylov_params kpk = *kp;
        const double eta = k == 1 ? 1e-2 : 0.9 * (r2 / r2prev);
        kpk.tol = std::max(kp->tol, std::min(1e-2, eta));
        r2prev = r2;
        HIP_CHECK(c, hipMemcpyAsync(c->ku, c->x, nb, hipMemcpyDeviceToDevice, c->stream));
        efem::dev::launchScale(c->kb, c->r, -1.0, n, c->stream);
        efem_krylov_stats ks;
        std::memset(&ks, 0, sizeof(ks));
        if ((rc = krylovInternal(c, 0, kpk, ks))) return rc;
        s.factor_seconds += ks.factor_seconds;
        s.ls_trials += ks.iterations;  // Krylov iterations, summed
        // backtracking on ||r||_2
        double step = 1.0;
        bool accepted = false;
        for (int t = 0; t < 12; ++t) {
            efem::dev::launchStep(c->xp, c->x, c->kx, step, n, c->stream);
            double r2n = 0.0, rinfn = 0.0, xinfn = 0.0;
            if ((rc = resNorms(c->xp, c->rp, &r2n, &rinfn, &xinfn))) return rc;
            if (std::isfinite(r2n) && r2n < r2) {
                std::swap(c->x, c->xp);
                std::swap(c->r, c->rp);
                r2 = r2n;
                rinf = rinfn;
                xinf = xinfn;
                accepted = true;
                break;
            }
            step *= 0.5;
        }
        s.iterations = k;
        c->trace.push_back(rinf);
        c->trace.push_back(step);
        c->trace.push_back(ks.iterations);
        if (!accepted) break;
    }
    s.rinf = rinf;
    s.xinf = xinf;
    HIP_CHECK(c, hipMemcpyAsync(u, c->x, nb, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    s.solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!std::isfinite(rinf)) {
        std::memset(u, 0, nb);
        c->err = "Newton produced a non-finite residual";
        return EFEM_ERR_NONFINITE;
    }
    return EFEM_OK;
}

int efem_adjoint(efem_ctx* c, const double* u, const double* g, double* lambda, double* grad_params,
                 double* grad_load, const efem_krylov_params* prm, efem_krylov_stats* st)
{
    if (!c || !u || !g) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    int rc = checkKrylovParams(c, prm);
    if (rc) return rc;
    efem_krylov_stats local;
    efem_krylov_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof(s));
    if ((rc = ensureKrylov(c))) return rc;
    const efem::Model& m = c->model;
    const int n = m.ndof();
    const size_t nb = static_cast<size_t>(n) * sizeof(double);
    HIP_CHECK(c, hipMemcpyAsync(c->ku, u, nb, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->kb, g, nb, hipMemcpyHostToDevice, c->stream));
    // K_t^T lambda = dJ/du
    if ((rc = krylovInternal(c, 1, *prm, s))) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    if (lambda) HIP_CHECK(c, hipMemcpyAsync(lambda, c->kx, nb, hipMemcpyDeviceToHost, c->stream));
    // dJ/dtheta = -lambda . dr/dtheta; r = f_int - f_pressure - f_const
    if (grad_params) {
        double* raw = static_cast<double*>(c->eslots);  // [E][4] FP64 (slots hold >= 12 E doubles)
        efem::dev::launchElementParamGrad(elemArgs(c, c->ku, false), c->kx, raw, c->stream);
        HIP_CHECK(c, hipGetLastError());
        std::vector<double> h(4 * static_cast<size_t>(m.nElems));
        HIP_CHECK(c, hipMemcpyAsync(h.data(), raw, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        if (m.paramsPerElem) {
            for (size_t i = 0; i < h.size(); ++i) grad_params[i] = -h[i];
        } else {
            for (int k = 0; k < 4; ++k) {
                double t = 0.0;
                for (int e = 0; e < m.nElems; ++e) t += h[4 * static_cast<size_t>(e) + k];
                grad_params[k] = -t;
            }
        }
    }
    if (grad_load) {
        // dr/dp = -f_pressure at unit load: dJ/dp = lambda . f_pressure(p = 1)
        grad_load[0] = grad_load[1] = grad_load[2] = 0.0;
        for (int comp = 0; comp < (m.directed ? 3 : 1) && m.nFaces > 0; ++comp) {
            efem::dev::FaceArgs fa = faceArgs(c, c->ku, nullptr);
            fa.pressure = 1.0;
            if (m.directed)
                for (int k = 0; k < 3; ++k) fa.pdir[k] = k == comp ? 1.0 : 0.0;
            efem::dev::launchPressure(fa, c->stream);
            efem::dev::launchGather(gatherArgs(c, c->kt, false, 0.0, 1.0, nullptr), c->stream);
            HIP_CHECK(c, hipGetLastError());
            double v = 0.0;
            if ((rc = dotHost(c, c->kx, c->kt, &v, nullptr))) return rc;
            grad_load[comp] = v;
        }
    }
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    s.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return EFEM_OK;
}

int efem_energy(efem_ctx* c, const double* u, double* W)
{
    if (!c || !u || !W) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    int rc = ensureKrylov(c);
    if (rc) return rc;
    const efem::Model& m = c->model;
    const size_t nb = static_cast<size_t>(m.ndof()) * sizeof(double);
    HIP_CHECK(c, hipMemcpyAsync(c->ku, u, nb, hipMemcpyHostToDevice, c->stream));
    double* per = static_cast<double*>(c->eslots);
    efem::dev::launchElementEnergy(elemArgs(c, c->ku, false), per, c->stream);
    efem::dev::launchSumComp(per, m.nElems, c->work, c->scal + kOut, c->stream);
    HIP_CHECK(c, hipGetLastError());
    if ((rc = readScalars(c, 1))) return rc;
    *W = c->hostScal[0];
    return EFEM_OK;
}

// efem_energy without allocation: the context's model, the device's scratch
// array (no hipMalloc / hipFree, no lazily allocated work vectors)
int efem_energy_noalloc(efem_ctx* c, const double* u_d, double* W_d)
{
    if (!c || !u_d || !W_d) return EFEM_ERR_ARG;
    const efem::Model& m = c->model;
    const size_t n = static_cast<size_t>(m.ndof());
    auto up = [](size_t v) { return (v + 31) / 32 * 32; };
    const size_t need = up(n + 3) + up(static_cast<size_t>(m.nElems)) + 2 * efem::dev::kReduceBlocks;
    if (need > efem::dev::kScratchDoubles) {
        c->err = "efem_energy_noalloc: problem too large for the device scratch (" + std::to_string(need) +
                 " doubles needed, " + std::to_string(efem::dev::kScratchDoubles) + " available)";
        return EFEM_ERR_ARG;
    }
    double* sp = efem::dev::scratchOf(c->device);
    if (!sp) {
        c->err = "efem_energy_noalloc: device scratch not initialised";
        return EFEM_ERR_STATE;
    }
    efem::dev::setTuning(c->tuning);
    efem::dev::launchEnergyScratch(elemArgs(c, nullptr, false), u_d, static_cast<int>(n), sp, W_d, c->stream);
    HIP_CHECK(c, hipGetLastError());
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

int efem_set_params(efem_ctx* c, const double* params)
{
    if (!c || !params) return EFEM_ERR_ARG;
    if (!c->parts.empty()) {
        // a batch's preconditioner is built from its instances' own models
        c->err = "efem_set_params: batch context (parameters are fixed at efem_batch_create)";
        return EFEM_ERR_STATE;
    }
    std::string msg;
    if (!efem::setParams(c->model, params, msg)) {
        c->err = msg;
        return EFEM_ERR_ARG;
    }
    // kernels still queued on the (non-blocking) context stream may read the
    // old parameters: drain it before the null-stream copy
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    HIP_CHECK(c, hipMemcpy(c->params, c->model.params.data(), c->model.params.size() * sizeof(double),
                           hipMemcpyHostToDevice));
    c->fac.ready = false;  // the approximate Jacobian depends on the parameters
    return EFEM_OK;
}

int efem_set_load(efem_ctx* c, double pressure, const double* pdir)
{
    if (!c) return EFEM_ERR_ARG;
    if (!c->parts.empty()) {
        // a batch's faces carry their instances' own pressures
        c->err = "efem_set_load: batch context (loads are fixed at efem_batch_create)";
        return EFEM_ERR_STATE;
    }
    efem::Model& m = c->model;
    if (m.nFaces == 0) {
        c->err = "efem_set_load: the problem has no pressure faces";
        return EFEM_ERR_STATE;
    }
    if (m.directed != (pdir ? 1 : 0)) {
        c->err = "efem_set_load: the load kind (follower / directed) is fixed at efem_create";
        return EFEM_ERR_ARG;
    }
    if (pdir)
        for (int k = 0; k < 3; ++k) m.pdir[k] = pdir[k];
    else
        m.pressure = pressure;
    return EFEM_OK;
}

int efem_material_point(int law, const double* params, const double* F, double* P, double* W, const double* dF,
                        double* dP)
{
    if (!params || !F || law < EFEM_NEOHOOKEAN || law > EFEM_YEOH || (dP && !dF)) return EFEM_ERR_ARG;
    double H[9];
    for (int k = 0; k < 9; ++k) H[k] = F[k] - ((k % 4) == 0 ? 1.0 : 0.0);
    if (P) efem::piolaFromHd<double>(law, params, H, P);
    if (W) *W = efem::energyFromH(law, params, H);
    if (dP) efem::piolaTangent<double>(law, params, H, dF, dP);
    return EFEM_OK;
}

int efem_batch_create(efem_ctx** out, const efem_problem* problems, int count, int precision)
{
    if (!out || !problems || count <= 0) {
        g_error = "efem_batch_create: null argument or empty batch";
        return EFEM_ERR_ARG;
    }
    *out = nullptr;
    if (precision < 0 || precision > EFEM_MIXED) {
        g_error = "efem_batch_create: unknown precision mode";
        return EFEM_ERR_ARG;
    }
    auto c = std::make_unique<efem_ctx>();
    c->precision = precision;
    std::string msg;
    std::vector<double> fpress;
    try {
        c->parts.resize(count);
        for (int i = 0; i < count; ++i)
            if (!efem::buildModel(problems[i], c->parts[i], msg)) {
                g_error = "instance " + std::to_string(i) + ": " + msg;
                return EFEM_ERR_ARG;
            }
        if (!efem::concatModels(c->parts, c->model, fpress, c->ranges, msg)) {
            g_error = msg;
            return EFEM_ERR_ARG;
        }
    } catch (const std::exception& e) {
        g_error = e.what();
        return EFEM_ERR_ARG;
    }
    int rc = initContext(c);
    if (rc) return rc;
    efem_ctx* cp = c.get();
    std::vector<efem::dev::BatchRange> br(count);
    for (int i = 0; i < count; ++i) {
        const efem::PartRange& r = c->ranges[i];
        br[i] = {r.e0, r.e1, r.f0, r.f1, r.n0, r.n1};
    }
    if ((rc = upload(cp, &cp->dRanges, br)) || (!fpress.empty() && (rc = upload(cp, &cp->fpress, fpress))) ||
        (rc = devAlloc(cp, &cp->batchStats, efem::dev::kBatchStats * static_cast<size_t>(count))) ||
        (rc = devAlloc(cp, &cp->batchSlots32, static_cast<size_t>(cp->model.nElems) * cp->model.nen * 3))) {
        g_error = cp->err;
        efem_destroy(c.release());
        return rc;
    }
    *out = c.release();
    return EFEM_OK;
}

int efem_batch_count(const efem_ctx* c) { return c ? static_cast<int>(c->parts.size()) : 0; }

int efem_batch_offsets(const efem_ctx* c, int* dof_offsets)
{
    if (!c || !dof_offsets || c->parts.empty()) return EFEM_ERR_ARG;
    for (size_t i = 0; i < c->ranges.size(); ++i) dof_offsets[i] = 3 * c->ranges[i].n0;
    dof_offsets[c->ranges.size()] = 3 * c->ranges.back().n1;
    return EFEM_OK;
}

int efem_batch_solve(efem_ctx* c, double* u, const efem_lbfgs_params* prm, efem_solve_stats* stats)
{
    if (!c || !u || !prm) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    if (c->parts.empty()) {
        c->err = "efem_batch_solve: not a batch context (efem_batch_create)";
        return EFEM_ERR_STATE;
    }
    if (prm->m <= 0 || prm->m > efem::dev::kMaxCompact || prm->max_linesearch <= 0 || prm->epsilon < 0 ||
        prm->epsilon_rel < 0 || prm->max_iterations <= 0) {
        c->err = "efem_batch_solve: invalid parameters (history 1..16, max_iterations > 0)";
        return EFEM_ERR_ARG;
    }
    if (prm->line_search != EFEM_LS_PROJECTION) {
        c->err = "efem_batch_solve: the batched solver runs the projection line search only";
        return EFEM_ERR_ARG;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const int count = static_cast<int>(c->parts.size());
    const int n = c->model.ndof();
    const int mc = prm->m;
    int rc;
    if ((rc = growHistory(c, mc))) return rc;
    double factorSeconds = 0.0;
    if (prm->use_precond && !c->fac.ready) {
        int failed = 0;
        if ((rc = factorInternal(c, &failed))) return rc;
        factorSeconds = c->factorSeconds;
    }
    const bool h0 = prm->use_precond && c->fac.ready;
    if (h0 && !c->fac.bp) {
        c->err = "internal: factor without staging buffer";
        return EFEM_ERR_STATE;
    }
    const size_t nb = static_cast<size_t>(n) * sizeof(double);
    HIP_CHECK(c, hipMemcpyAsync(c->x, u, nb, hipMemcpyHostToDevice, c->stream));
    if (prm->x_target) {
        // per-instance displacement targets (concatenated like u)
        if (!c->xtarget && (rc = devAlloc(c, &c->xtarget, static_cast<size_t>(n) + 3))) return rc;
        HIP_CHECK(c, hipMemcpyAsync(c->xtarget, prm->x_target, nb, hipMemcpyHostToDevice, c->stream));
    }
    efem::dev::BatchArgs a;
    a.elem = elemArgs(c, c->x, false);
    a.face = faceArgs(c, c->x, nullptr);
    a.gather = gatherArgs(c, c->r, false, 1.0, -1.0, c->fext);
    a.trsv = h0 ? trsvArgs(c) : efem::dev::TrsvArgs{};
    a.trsv.mixed = fac32(c);
    a.items = h0 ? c->fac.items : nullptr;
    a.instItems = h0 ? c->fac.instItems : nullptr;
    a.nlev = h0 ? c->fac.instLevels : 0;
    if (h0 && c->batchDense && c->fac.batchZ) {
        a.zArena = c->fac.batchZ;
        a.zoff = c->fac.batchZoff;
        a.zld = c->fac.batchZld;
        a.ztcol = c->fac.batchZtcol;
    }
    a.ranges = c->dRanges;
    a.x = c->x;
    a.r = c->r;
    a.xp = c->xp;
    a.rp = c->rp;
    a.d = c->d;
    a.q = c->tmp;
    a.bp = h0 ? c->fac.bp : nullptr;
    a.S = c->S;
    a.Y = c->Y;
    a.ntot = n;
    a.m = mc;
    a.maxIt = prm->max_iterations;
    a.maxLs = prm->max_linesearch;
    a.usePrecond = h0 ? 1 : 0;
    a.elemF32 = elem32(c) ? 1 : 0;
    a.eps = prm->epsilon;
    a.epsRel = prm->epsilon_rel;
    a.lnsFea = prm->lns_fea;
    a.fp64Switch = elem32(c) ? prm->fp64_switch : 0.0;
    a.slots32 = c->batchSlots32;
    a.xt = prm->x_target ? c->xtarget : nullptr;
    a.targetTol = prm->target_tol;
    a.stats = c->batchStats;
    efem::dev::launchBatchSolve(a, count, c->stream);
    HIP_CHECK(c, hipGetLastError());
    const int ks = efem::dev::kBatchStats;
    std::vector<double> st(ks * static_cast<size_t>(count));
    HIP_CHECK(c, hipMemcpyAsync(u, c->x, nb, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(st.data(), c->batchStats, st.size() * sizeof(double), hipMemcpyDeviceToHost,
                                c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    bool anyNonfinite = false;
    for (int i = 0; i < count; ++i) {
        const double* q = &st[ks * static_cast<size_t>(i)];
        const bool bad = q[5] != 0.0 || !std::isfinite(q[3]) || !std::isfinite(q[4]);
        if (bad) {
            anyNonfinite = true;
            const efem::PartRange& r = c->ranges[i];
            std::memset(u + 3 * static_cast<size_t>(r.n0), 0, 3 * static_cast<size_t>(r.n1 - r.n0) * sizeof(double));
        }
        if (stats) {
            efem_solve_stats& s = stats[i];
            std::memset(&s, 0, sizeof(s));
            s.iterations = static_cast<int>(q[0]);
            s.r_evals = static_cast<int>(q[1]);
            s.ls_trials = static_cast<int>(q[2]);
            s.rinf = q[3];
            s.xinf = q[4];
            s.precond_failed = prm->use_precond && !h0;
            s.fp64_switch_iteration = static_cast<int>(q[6]);
            s.target_reached = static_cast<int>(q[8]);
            s.target_err = q[9];
            s.factor_seconds = factorSeconds;
            s.solve_seconds = wall;
            s.factor_nnz = h0 ? c->fac.nnz : 0;
        }
    }
    if (anyNonfinite) {
        c->err = "batched L-BFGS: non-finite residual or displacement in at least one instance";
        return EFEM_ERR_NONFINITE;
    }
    return EFEM_OK;
}

int efem_time_residual(efem_ctx* c, const double* u, int count, double* ms_per_eval)
{
    if (!c || !u || count <= 0 || !ms_per_eval) return EFEM_ERR_ARG;
    efem::dev::setTuning(c->tuning);
    const size_t nb = static_cast<size_t>(c->model.ndof()) * sizeof(double);
    HIP_CHECK(c, hipMemcpyAsync(c->tmp, u, nb, hipMemcpyHostToDevice, c->stream));
    int rc = residualInternal(c, c->tmp, c->r);  // warm-up
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIP_CHECK(c, hipEventCreate(&e0));
    HIP_CHECK(c, hipEventCreate(&e1));
    HIP_CHECK(c, hipEventRecord(e0, c->stream));
    for (int i = 0; i < count && rc == EFEM_OK; ++i) rc = residualInternal(c, c->tmp, c->r);
    HIP_CHECK(c, hipEventRecord(e1, c->stream));
    HIP_CHECK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_eval = ms / count;
    return rc;
}

int efem_autotune_factor(efem_ctx* c, double* report)
{
    if (!c) return EFEM_ERR_ARG;
    if (!c->parts.empty()) {
        c->err = "efem_autotune_factor: single-problem contexts only";
        return EFEM_ERR_STATE;
    }
    const int n = c->model.ndof();
    if (n == 0) {
        c->err = "efem_autotune_factor: no free DOF";
        return EFEM_ERR_STATE;
    }
    hipEvent_t e0, e1;
    HIP_CHECK(c, hipEventCreate(&e0));
    HIP_CHECK(c, hipEventCreate(&e1));
    // the apply's time does not depend on the values: a zero right-hand side
    HIP_CHECK(c, hipMemsetAsync(c->tmp, 0, static_cast<size_t>(n + 3) * sizeof(double), c->stream));
    const int cands[4] = {64, 128, 192, 256};
    const int keep = c->denseSubtree;
    int bestDs = keep;
    double bestMs = 1e300, bestFactor = 0.0;
    long long bestNnz = 0;
    int rc = EFEM_OK;
    for (int i = 0; i < 4 && rc == EFEM_OK; ++i) {
        c->denseSubtree = cands[i];
        c->symbolic.reset();
        int failed = 0;
        const auto t0 = std::chrono::steady_clock::now();
        rc = factorInternal(c, &failed);
        const double fs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc != EFEM_OK) break;
        if (failed) {
            c->err = "efem_autotune_factor: the approximate Jacobian is not SPD";
            rc = EFEM_ERR_STATE;
            break;
        }
        // one warm-up apply, then 20 timed
        if ((rc = applyH0Internal(c, c->tmp, c->d))) break;
        HIP_CHECK(c, hipEventRecord(e0, c->stream));
        for (int k = 0; k < 20 && rc == EFEM_OK; ++k) rc = applyH0Internal(c, c->tmp, c->d);
        HIP_CHECK(c, hipEventRecord(e1, c->stream));
        HIP_CHECK(c, hipEventSynchronize(e1));
        float t = 0.f;
        HIP_CHECK(c, hipEventElapsedTime(&t, e0, e1));
        const double ms = t / 20.0;
        // a larger subtree must win by 1 %: it stores more entries
        if (ms < 0.99 * bestMs) {
            bestMs = ms;
            bestDs = cands[i];
            bestFactor = fs;
            bestNnz = c->fac.nnz;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->denseSubtree = rc == EFEM_OK ? bestDs : keep;
    c->symbolic.reset();
    releaseFactor(c);
    if (rc != EFEM_OK) return rc;
    int failed = 0;
    if ((rc = factorInternal(c, &failed))) return rc;
    if (report) {
        report[0] = bestDs;
        report[1] = bestMs;
        report[2] = bestFactor;
        report[3] = static_cast<double>(bestNnz);
    }
    return EFEM_OK;
}

int efem_autotune(efem_ctx* c, double* report)
{
    if (!c) return EFEM_ERR_ARG;
    const int n = c->model.ndof();
    hipEvent_t e0, e1;
    HIP_CHECK(c, hipEventCreate(&e0));
    HIP_CHECK(c, hipEventCreate(&e1));
    int rc = EFEM_OK;
    // time `reps` repetitions of f() on the context stream (after one warm-up)
    auto timeIt = [&](auto&& f, int reps, double* ms) -> int {
        int r = f();
        if (r) return r;
        HIP_CHECK(c, hipEventRecord(e0, c->stream));
        for (int i = 0; i < reps && r == EFEM_OK; ++i) r = f();
        HIP_CHECK(c, hipEventRecord(e1, c->stream));
        HIP_CHECK(c, hipEventSynchronize(e1));
        float t = 0.f;
        HIP_CHECK(c, hipEventElapsedTime(&t, e0, e1));
        *ms = t / reps;
        return r;
    };
    efem::dev::LaunchTuning best = c->tuning;
    // residual: gather lanes per node x C3D4 block size (at u = 0; the cost
    // does not depend on the values)
    HIP_CHECK(c, hipMemsetAsync(c->tmp, 0, static_cast<size_t>(n) * sizeof(double), c->stream));
    double bestRes = 1e300;
    const int lanes[4] = {1, 2, 4, 8};
    const int tpbs[3] = {64, 128, 256};
    // with C3D4 patches the residual's shapes are the record gather's lanes
    const bool patchPath = c->tet4Patches && c->patch.part;
    int bestPatchLanes = c->patchLanes;
    for (int li = 0; li < 4 && rc == EFEM_OK; ++li)
        for (int ti = 0; ti < (c->model.nen == 4 && !patchPath ? 3 : 1) && rc == EFEM_OK; ++ti) {
            efem::dev::LaunchTuning t = best;
            if (patchPath) {
                c->patchLanes = lanes[li];
            } else {
                t.gatherLanes = lanes[li];
                if (c->model.nen == 4) t.tet4Threads = tpbs[ti];
            }
            efem::dev::setTuning(t);
            double ms = 0.0;
            rc = timeIt([&]() { return residualInternal(c, c->tmp, c->r); }, 10, &ms);
            if (rc == EFEM_OK && ms < bestRes) {
                bestRes = ms;
                bestPatchLanes = lanes[li];
                best.gatherLanes = t.gatherLanes;
                best.tet4Threads = t.tet4Threads;
            }
        }
    c->patchLanes = bestPatchLanes;
    // reductions: the two-vector dot + norm of the line search
    double bestRed = 1e300;
    const int grids[4] = {32, 64, 128, 256};
    for (int gi = 0; gi < 4 && rc == EFEM_OK; ++gi) {
        efem::dev::LaunchTuning t = best;
        t.reduceBlocks = grids[gi];
        efem::dev::setTuning(t);
        double ms = 0.0;
        rc = timeIt([&]() {
            efem::dev::launchDotMax(c->r, c->tmp, c->x, n, c->work, c->scal + kOut, c->stream);
            return EFEM_OK;
        }, 20, &ms);
        if (rc == EFEM_OK && ms < bestRed) {
            bestRed = ms;
            best.reduceBlocks = grids[gi];
        }
    }
    // history reductions: the new pair + first loop and the second loop of
    // the compact two-loop with m = 8 (history buffers grown if needed;
    // their contents and the scalar slots are scratch outside a solve)
    double bestHist = 1e300;
    const int hm = 8;
    if (rc == EFEM_OK) rc = growHistory(c, hm);
    if (rc == EFEM_OK) {
        HIP_CHECK(c, hipMemsetAsync(c->S, 0, static_cast<size_t>(hm) * n * sizeof(double), c->stream));
        HIP_CHECK(c, hipMemsetAsync(c->Y, 0, static_cast<size_t>(hm) * n * sizeof(double), c->stream));
    }
    double* sc = c->scal;
    for (int gi = 0; gi < 4 && rc == EFEM_OK; ++gi) {
        efem::dev::LaunchTuning t = best;
        t.historyBlocks = grids[gi];
        efem::dev::setTuning(t);
        double ms = 0.0;
        rc = timeIt([&]() {
            efem::dev::launchGramFirst(c->S, c->Y, c->x, c->xp, c->r, c->rp, n, hm, hm - 1, hm, sc + kGram, sc + kYs,
                                       sc + kYY, sc + kTheta, sc + kAlpha, sc + kRS, c->tmp, c->work,
                                       c->stream);
            efem::dev::launchTwoLoopSecond(efem::dev::HSrc{c->tmp}, c->S, c->Y, n, hm, hm - 1, hm, sc + kGram, sc + kYs, sc + kAlpha,
                                           sc + kCoef, c->d, c->work, c->stream, c->r, sc + kRS,
                                           sc + kOut + 4);
            return EFEM_OK;
        }, 10, &ms);
        if (rc == EFEM_OK && ms < bestHist) {
            bestHist = ms;
            best.historyBlocks = grids[gi];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    // batches: the persistent solver's workgroup width, timed on 40
    // iterations of every instance from u = 0 (the history and iterate
    // buffers are scratch outside a solve)
    double bestBatch = 0.0;
    if (!c->parts.empty()) {
        bestBatch = 1e300;
        std::vector<double> u0(static_cast<size_t>(std::max(n, 1)));
        efem_lbfgs_params bp{};
        bp.epsilon = 0.0;
        bp.epsilon_rel = 0.0;
        bp.max_iterations = 40;
        bp.max_linesearch = 20;
        bp.m = 6;
        bp.lns_fea = 1.0;
        bp.use_precond = 1;
        bp.line_search = EFEM_LS_PROJECTION;
        const int widths[3] = {256, 512, 1024};
        efem::dev::LaunchTuning keep = best;
        for (int wi = 0; wi < 3 && rc == EFEM_OK; ++wi) {
            c->tuning = keep;
            c->tuning.batchThreads = widths[wi];
            double ms = 1e300;
            for (int rep = 0; rep < 3 && rc == EFEM_OK; ++rep) {
                std::fill(u0.begin(), u0.end(), 0.0);
                const auto t0 = std::chrono::steady_clock::now();
                rc = efem_batch_solve(c, u0.data(), &bp, nullptr);
                if (rc == EFEM_ERR_NONFINITE) rc = EFEM_OK;  // timing only
                const double t = 1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (rep > 0) ms = std::min(ms, t);  // the first run also factors
            }
            if (rc == EFEM_OK && ms < bestBatch) {
                bestBatch = ms;
                best.batchThreads = widths[wi];
            }
        }
        if (rc) return rc;
    }
    c->tuning = best;
    efem::dev::setTuning(best);
    if (report) {
        report[0] = patchPath ? c->patchLanes : best.gatherLanes;
        report[1] = best.tet4Threads;
        report[2] = best.reduceBlocks;
        report[3] = bestRes;
        report[4] = bestRed;
        report[5] = best.historyBlocks;
        report[6] = bestHist;
        report[7] = best.batchThreads;
        report[8] = bestBatch;
    }
    return EFEM_OK;
}

}  // extern "C"
