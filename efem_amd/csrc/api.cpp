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
#include "host/model.hpp"
#include "host/sparse.hpp"

namespace {

thread_local std::string g_error;

struct DeviceFactor {
    bool ready = false;
    int n = 0, nsuper = 0, nlevels = 0;
    long long nnz = 0;
    int* perm = nullptr;
    int *snFirst = nullptr, *snW = nullptr, *snM = nullptr, *snRow = nullptr, *snT = nullptr;
    long long* snVal = nullptr;
    int* rows = nullptr;
    void* B = nullptr;
    int *fptr = nullptr, *fidx = nullptr;
    void *y = nullptr, *t = nullptr, *z = nullptr;
    int2* items = nullptr;
    std::vector<int> fwdOff, fwdCnt, bwdOff, bwdCnt;  // per level
    // the whole apply captured as one HIP graph over fixed staging buffers
    double *in = nullptr, *out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

}  // namespace

struct efem_ctx {
    std::string err;
    efem::Model model;
    int precision = EFEM_FP64;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
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
    int *gptr = nullptr, *gslot = nullptr;
    double* fext = nullptr;
    // solver state
    int mcap = 0;
    double *x = nullptr, *r = nullptr, *xp = nullptr, *rp = nullptr, *d = nullptr, *tmp = nullptr;
    double *S = nullptr, *Y = nullptr;
    double* scal = nullptr;  // device scalars
    double* work = nullptr;
    unsigned* counter = nullptr;
    double* hostScal = nullptr;  // pinned
    DeviceFactor fac;
    bool factorFailed = false;
    double factorSeconds = 0.0;
    std::vector<double> trace;
    // profiling of the element kernel and of H0 applies
    bool profiling = false;
    std::vector<hipEvent_t> evs, evsH0;
    int evUsed = 0, evH0Used = 0;
    double* xtarget = nullptr;
    bool useGraph = true;    // option "h0_graph"
    bool compactTwoLoop = true;  // option "two_loop_compact" (history <= 16)
};

namespace {

// scalar slots in ctx->scal
constexpr int kOut = 0;       // [3] reduction outputs
constexpr int kYs = 8;        // [64]
constexpr int kYY = 72;       // [64]
constexpr int kAlpha = 136;   // [64]
constexpr int kCoef = 200;    // [64]
constexpr int kTheta = 264;   // [1]
constexpr int kGram = 272;    // [16][16] s_i . y_k of the compact two-loop
constexpr int kNumScal = 528;

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
    c->allocs.push_back(q);
    *p = static_cast<T*>(q);
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
    if ((rc = upload(c, &c->fext, m.fext))) return rc;
    const size_t np = static_cast<size_t>(m.ndof()) + 3;  // + dummy zero node
    for (double** v : {&c->x, &c->r, &c->xp, &c->rp, &c->d, &c->tmp}) {
        if ((rc = devAlloc(c, v, np))) return rc;
        HIP_CHECK(c, hipMemset(*v, 0, np * sizeof(double)));
    }
    if ((rc = devAlloc(c, &c->scal, kNumScal))) return rc;
    HIP_CHECK(c, hipMemset(c->scal, 0, kNumScal * sizeof(double)));
    if ((rc = devAlloc(c, &c->work, efem::dev::kReduceWorkDoubles))) return rc;
    if ((rc = devAlloc(c, &c->counter, 4))) return rc;
    HIP_CHECK(c, hipMemset(c->counter, 0, 4 * sizeof(unsigned)));
    HIP_CHECK(c, hipHostMalloc(reinterpret_cast<void**>(&c->hostScal), 16 * sizeof(double), hipHostMallocDefault));
    return EFEM_OK;
}

// r = f_int(x) - f_ext(x); x has 3 trailing zeros (the fixed-node slot)
int residualInternal(efem_ctx* c, const double* x, double* r)
{
    const efem::Model& m = c->model;
    efem::dev::ElemArgs ea;
    ea.nElems = m.nElems;
    ea.nen = m.nen;
    ea.ng = m.ng;
    ea.law = m.law;
    ea.paramsPerElem = m.paramsPerElem;
    const bool f32 = elem32(c) && !c->elemFp64;
    ea.mixed = f32;
    ea.connDof = c->connDof;
    ea.ginv = f32 ? c->ginv : c->ginv64;
    ea.vol = c->vol;
    ea.gdN = f32 ? c->gdN : c->gdN64;
    ea.gw = c->gw;
    ea.params = c->params;
    ea.x = x;
    ea.slots = c->eslots;
    const bool prof = c->profiling && c->evUsed + 2 <= static_cast<int>(c->evs.size());
    if (prof) HIP_CHECK(c, hipEventRecord(c->evs[c->evUsed], c->stream));
    efem::dev::launchElementForce(ea, c->stream);
    if (prof) {
        HIP_CHECK(c, hipEventRecord(c->evs[c->evUsed + 1], c->stream));
        c->evUsed += 2;
    }
    HIP_CHECK(c, hipGetLastError());
    if (m.nFaces > 0) {
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
        fa.slots = c->fslots;
        efem::dev::launchPressure(fa, c->stream);
        HIP_CHECK(c, hipGetLastError());
    }
    efem::dev::GatherArgs ga;
    ga.nFree = m.nFree;
    ga.nElemSlots = m.nElems * m.nen;
    ga.mixed = f32;
    ga.gptr = c->gptr;
    ga.gslot = c->gslot;
    ga.eslots = c->eslots;
    ga.fslots = c->fslots;
    ga.fext = c->fext;
    ga.r = r;
    efem::dev::launchGather(ga, c->stream);
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
    a.snFirst = f.snFirst;
    a.snW = f.snW;
    a.snM = f.snM;
    a.snRow = f.snRow;
    a.snVal = f.snVal;
    a.snT = f.snT;
    a.rows = f.rows;
    a.B = f.B;
    a.fptr = f.fptr;
    a.fidx = f.fidx;
    a.y = f.y;
    a.t = f.t;
    a.z = f.z;
    return a;
}

void launchApplySweeps(efem_ctx* c, const double* v, double* out)
{
    DeviceFactor& f = c->fac;
    const efem::dev::TrsvArgs a = trsvArgs(c);
    for (int l = 0; l < f.nlevels; ++l)
        efem::dev::launchTrsvForward(a, v, f.items + f.fwdOff[l], f.fwdCnt[l], c->stream);
    for (int l = f.nlevels - 1; l >= 0; --l)
        efem::dev::launchTrsvBackward(a, f.items + f.bwdOff[l], f.bwdCnt[l], c->stream);
    efem::dev::launchUnpermute(a, out, c->stream);
}

// out = K^-1 v on the device (v, out: interleaved free-DOF vectors).  With
// the graph option the forward/backward level launches are replayed from one
// captured HIP graph reading `in` and writing `out` staging buffers.
int applyH0Internal(efem_ctx* c, const double* v, double* out)
{
    DeviceFactor& f = c->fac;
    const bool prof = c->profiling && c->evH0Used + 2 <= static_cast<int>(c->evsH0.size());
    if (prof) HIP_CHECK(c, hipEventRecord(c->evsH0[c->evH0Used], c->stream));
    if (c->useGraph) {
        const size_t nb = 3 * static_cast<size_t>(f.n) * sizeof(double);
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
        HIP_CHECK(c, hipMemcpyAsync(f.in, v, nb, hipMemcpyDeviceToDevice, c->stream));
        HIP_CHECK(c, hipGraphLaunch(f.exec, c->stream));
        HIP_CHECK(c, hipMemcpyAsync(out, f.out, nb, hipMemcpyDeviceToDevice, c->stream));
    } else {
        launchApplySweeps(c, v, out);
    }
    if (prof) {
        HIP_CHECK(c, hipEventRecord(c->evsH0[c->evH0Used + 1], c->stream));
        c->evH0Used += 2;
    }
    HIP_CHECK(c, hipGetLastError());
    return EFEM_OK;
}

int factorInternal(efem_ctx* c, int* failed)
{
    const auto t0 = std::chrono::steady_clock::now();
    const efem::Model& m = c->model;
    c->fac.ready = false;
    c->factorFailed = true;
    if (failed) *failed = 1;
    if (m.nFree == 0) return EFEM_OK;  // jacobi.nonZeros() == 0 -> fallback
    efem::Csc A;
    efem::scalarLaplacian(m, A);
    const std::vector<int> perm = efem::nestedDissection(A, efem::freeCoords(m));
    efem::SupernodalFactor F;
    std::string msg;
    if (!efem::supernodalCholesky(A, perm, F, msg)) {
        c->err = msg;
        c->factorSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return EFEM_OK;  // not SPD: the reference falls back to I/theta
    }
    DeviceFactor& d = c->fac;
    d.n = F.n;
    d.nsuper = F.nsuper;
    d.nlevels = F.nlevels;
    d.nnz = F.nnzL;
    std::vector<int> w(F.nsuper), mm(F.nsuper), rowOff(F.nsuper), tOff(F.nsuper + 1);
    std::vector<long long> valOff(F.nsuper);
    int nt = 0;
    for (int s = 0; s < F.nsuper; ++s) {
        w[s] = F.first[s + 1] - F.first[s];
        if (w[s] > 256) {
            c->err = "internal: supernode wider than the device staging buffer";
            return EFEM_ERR_STATE;
        }
        mm[s] = F.rowPtr[s + 1] - F.rowPtr[s];
        rowOff[s] = F.rowPtr[s];
        valOff[s] = F.valPtr[s];
        tOff[s] = nt;
        nt += mm[s] - w[s];
    }
    tOff[F.nsuper] = nt;
    // forward-sweep extend-add map (children's update rows -> front rows)
    std::vector<int> fptr, fidx;
    efem::frontGatherMap(F, tOff, fptr, fidx);
    // work items per level
    std::vector<std::vector<int>> byLevel(F.nlevels);
    for (int s = 0; s < F.nsuper; ++s) byLevel[F.level[s]].push_back(s);
    std::vector<int2> items;
    d.fwdOff.assign(F.nlevels, 0);
    d.fwdCnt.assign(F.nlevels, 0);
    d.bwdOff.assign(F.nlevels, 0);
    d.bwdCnt.assign(F.nlevels, 0);
    for (int l = 0; l < F.nlevels; ++l) {
        d.fwdOff[l] = static_cast<int>(items.size());
        for (int s : byLevel[l])
            for (int r0 = 0; r0 < mm[s]; r0 += 64) items.push_back(make_int2(s, r0));
        d.fwdCnt[l] = static_cast<int>(items.size()) - d.fwdOff[l];
    }
    for (int l = 0; l < F.nlevels; ++l) {
        d.bwdOff[l] = static_cast<int>(items.size());
        for (int s : byLevel[l])
            for (int k0 = 0; k0 < w[s]; k0 += 4) items.push_back(make_int2(s, k0));
        d.bwdCnt[l] = static_cast<int>(items.size()) - d.bwdOff[l];
    }
    int rc;
    if ((rc = upload(c, &d.perm, F.perm))) return rc;
    if ((rc = upload(c, &d.snFirst, std::vector<int>(F.first.begin(), F.first.end() - 1)))) return rc;
    if ((rc = upload(c, &d.snW, w))) return rc;
    if ((rc = upload(c, &d.snM, mm))) return rc;
    if ((rc = upload(c, &d.snRow, rowOff))) return rc;
    if ((rc = upload(c, &d.snVal, valOff))) return rc;
    if ((rc = upload(c, &d.snT, std::vector<int>(tOff.begin(), tOff.end() - 1)))) return rc;
    if ((rc = upload(c, &d.rows, F.rows))) return rc;
    if (fac32(c)) {
        std::vector<float> Bf(static_cast<size_t>(F.Bsize));
        for (long long i = 0; i < F.Bsize; ++i) Bf[static_cast<size_t>(i)] = static_cast<float>(F.B[i]);
        float* q = nullptr;
        if ((rc = upload(c, &q, Bf))) return rc;
        d.B = q;
    } else {
        double* q = nullptr;
        if ((rc = devAlloc(c, &q, static_cast<size_t>(F.Bsize)))) return rc;
        HIP_CHECK(c, hipMemcpy(q, F.B.get(), static_cast<size_t>(F.Bsize) * sizeof(double), hipMemcpyHostToDevice));
        d.B = q;
    }
    if ((rc = upload(c, &d.fptr, fptr))) return rc;
    if ((rc = upload(c, &d.fidx, fidx))) return rc;
    if ((rc = devAlloc(c, &d.in, 3 * static_cast<size_t>(F.n) + 3))) return rc;
    if ((rc = devAlloc(c, &d.out, 3 * static_cast<size_t>(F.n) + 3))) return rc;
    if (d.exec) {
        (void)hipGraphExecDestroy(d.exec);
        d.exec = nullptr;
    }
    if (d.graph) {
        (void)hipGraphDestroy(d.graph);
        d.graph = nullptr;
    }
    if ((rc = upload(c, &d.items, items))) return rc;
    const size_t es = fac32(c) ? sizeof(float) : sizeof(double);
    double* q = nullptr;
    if ((rc = devAlloc(c, &q, (3 * static_cast<size_t>(F.n) * es + 7) / 8))) return rc;
    d.y = q;
    if ((rc = devAlloc(c, &q, (3 * static_cast<size_t>(std::max(nt, 1)) * es + 7) / 8))) return rc;
    d.t = q;
    if ((rc = devAlloc(c, &q, (3 * static_cast<size_t>(F.n) * es + 7) / 8))) return rc;
    d.z = q;
    d.ready = true;
    c->factorFailed = false;
    if (failed) *failed = 0;
    c->factorSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return EFEM_OK;
}

int readScalars(efem_ctx* c, int count)
{
    HIP_CHECK(c, hipMemcpyAsync(c->hostScal, c->scal + kOut, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
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
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_error = "efem_create: no HIP device";
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
    *out = c.release();
    return EFEM_OK;
}

int efem_destroy(efem_ctx* c)
{
    if (!c) return EFEM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->fac.exec) (void)hipGraphExecDestroy(c->fac.exec);
    if (c->fac.graph) (void)hipGraphDestroy(c->fac.graph);
    for (void* p : c->allocs) (void)hipFree(p);
    for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->evsH0) (void)hipEventDestroy(e);
    if (c->hostScal) (void)hipHostFree(c->hostScal);
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
    const size_t n = static_cast<size_t>(c->model.ndof());
    HIP_CHECK(c, hipMemcpyAsync(c->tmp, u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = residualInternal(c, c->tmp, c->r);
    if (rc) return rc;
    HIP_CHECK(c, hipMemcpyAsync(r, c->r, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return EFEM_OK;
}

int efem_residual_d(efem_ctx* c, const double* u_d, double* r_d)
{
    if (!c || !u_d || !r_d) return EFEM_ERR_ARG;
    const size_t n = static_cast<size_t>(c->model.ndof());
    HIP_CHECK(c, hipMemcpyAsync(c->tmp, u_d, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return residualInternal(c, c->tmp, r_d);
}

int efem_factor(efem_ctx* c, int* failed)
{
    if (!c) return EFEM_ERR_ARG;
    return factorInternal(c, failed);
}

long long efem_factor_nnz(const efem_ctx* c) { return c && c->fac.ready ? c->fac.nnz : 0; }

int efem_factor_release(efem_ctx* c)
{
    if (!c) return EFEM_ERR_ARG;
    // device buffers of the factor stay in c->allocs until efem_destroy;
    // dropping `ready` makes the next solve factor again
    c->fac.ready = false;
    return EFEM_OK;
}

int efem_apply_h0(efem_ctx* c, const double* v, double* out)
{
    if (!c || !v || !out) return EFEM_ERR_ARG;
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
    if (prm->m <= 0 || prm->m > 64 || prm->max_linesearch <= 0 || prm->epsilon < 0 || prm->epsilon_rel < 0 ||
        prm->max_iterations < 0) {
        c->err = "efem_solve: invalid parameters (Param.h:110-127)";
        return EFEM_ERR_ARG;
    }
    efem_solve_stats local;
    efem_solve_stats& s = st ? *st : local;
    std::memset(&s, 0, sizeof(s));
    const auto t0 = std::chrono::steady_clock::now();
    const int n = c->model.ndof();
    const int mc = prm->m;
    int rc;
    if (mc > c->mcap) {
        if ((rc = devAlloc(c, &c->S, static_cast<size_t>(mc) * std::max(n, 1)))) return rc;
        if ((rc = devAlloc(c, &c->Y, static_cast<size_t>(mc) * std::max(n, 1)))) return rc;
        c->mcap = mc;
    }
    c->trace.clear();
    double* sc = c->scal;
    c->elemFp64 = false;
    if (prm->x_target) {
        if (!c->xtarget && (rc = devAlloc(c, &c->xtarget, static_cast<size_t>(n) + 3))) return rc;
        HIP_CHECK(c, hipMemcpyAsync(c->xtarget, prm->x_target, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        efem::dev::launchDot2(c->xtarget, c->xtarget, n, c->work, c->counter, sc + kOut + 4, sc + kOut + 5, c->stream);
    }
    HIP_CHECK(c, hipMemcpyAsync(c->x, u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = residualInternal(c, c->x, c->r))) return rc;
    s.r_evals = 1;
    efem::dev::launchDotMax(c->x, c->x, c->r, n, c->work, c->counter, sc + kOut, c->stream);
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
    for (;;) {
        HIP_CHECK(c, hipMemcpyAsync(c->xp, c->x, nb, hipMemcpyDeviceToDevice, c->stream));
        HIP_CHECK(c, hipMemcpyAsync(c->rp, c->r, nb, hipMemcpyDeviceToDevice, c->stream));
        efem::dev::launchDotMax(c->r, c->d, nullptr, n, c->work, c->counter, sc + kOut, c->stream);
        if ((rc = readScalars(c, 1))) return rc;
        const double dgInit = c->hostScal[0];
        double cur = 1.0, used = 1.0;
        int trials = 0;
        for (int it = 0; it < prm->max_linesearch; ++it) {
            efem::dev::launchStep(c->x, c->xp, c->d, cur, n, c->stream);
            if ((rc = residualInternal(c, c->x, c->r))) return rc;
            ++s.r_evals;
            ++trials;
            used = cur;
            efem::dev::launchDotMax(c->r, c->d, c->x, n, c->work, c->counter, sc + kOut, c->stream);
            if ((rc = readScalars(c, 3))) return rc;
            const double dg = c->hostScal[0];
            rinf = c->hostScal[1];
            xinf = c->hostScal[2];
            if (std::fabs(dg) <= prm->lns_fea * std::fabs(dgInit)) break;
            cur *= dg * dgInit > 0 ? 1.5 : 0.5;
        }
        s.ls_trials += trials;
        c->trace.push_back(rinf);
        c->trace.push_back(used);
        c->trace.push_back(trials);
        s.iterations = k;
        s.rinf = rinf;
        s.xinf = xinf;
        if (!std::isfinite(rinf) || !std::isfinite(xinf)) break;
        if (prm->x_target) {
            // ||x - x_target||: diff into tmp, then its norm
            efem::dev::launchDiff2(c->tmp, c->tmp, c->x, c->xtarget, c->x, c->xtarget, n, c->stream);
            efem::dev::launchDot2(c->tmp, c->tmp, n, c->work, c->counter, sc + kOut + 6, sc + kOut + 7, c->stream);
            HIP_CHECK(c, hipMemcpyAsync(c->hostScal + 6, sc + kOut + 6, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIP_CHECK(c, hipStreamSynchronize(c->stream));
            s.target_err = std::sqrt(c->hostScal[6]) / tnorm;
            if (s.target_err <= prm->target_tol) {
                s.target_reached = 1;
                break;
            }
        }
        if (rinf <= prm->epsilon || rinf <= prm->epsilon_rel * xinf) break;
        if (prm->max_iterations != 0 && k >= prm->max_iterations) break;
        if (elem32(c) && !c->elemFp64 && prm->fp64_switch > 0.0 &&
            rinf <= prm->fp64_switch * rinf0) {
            // FP64 finish: re-evaluate r(x) with the FP64 element kernels
            c->elemFp64 = true;
            s.fp64_switch_iteration = k;
            if ((rc = residualInternal(c, c->x, c->r))) return rc;
            ++s.r_evals;
        }
        // add_correction (LBFGSMat.h:81-98)
        const int loc = ptr % mc;
        double* sl = c->S + static_cast<size_t>(loc) * n;
        double* yl = c->Y + static_cast<size_t>(loc) * n;
        efem::dev::launchDiff2(sl, yl, c->x, c->xp, c->r, c->rp, n, c->stream);
        if (ncorr < mc) ++ncorr;
        ptr = loc + 1;
        const bool compact = c->compactTwoLoop && mc <= efem::dev::kMaxCompact;
        if (compact) {
            efem::dev::launchGramUpdate(c->S, c->Y, n, ncorr, loc, sc + kGram, sc + kYs, sc + kYY + loc, c->work,
                                        c->counter, c->stream);
        } else {
            efem::dev::launchDot2(sl, yl, n, c->work, c->counter, sc + kYs + loc, sc + kYY + loc, c->stream);
        }
        efem::dev::launchTheta(sc + kYs + loc, sc + kYY + loc, sc + kTheta, c->stream);
        if (compact) {
            // two-loop from Gram products: 2 multi-dot passes + 2 multi-axpys
            efem::dev::launchTwoLoopFirst(c->r, c->S, c->Y, n, ncorr, loc, mc, sc + kGram, sc + kYs, sc + kAlpha,
                                          c->d, c->work, c->counter, c->stream);
            if (h0) {
                if ((rc = applyH0Internal(c, c->d, c->tmp))) return rc;
                std::swap(c->d, c->tmp);
            } else {
                efem::dev::launchDivDev(c->d, sc + kTheta, n, c->stream);
            }
            efem::dev::launchTwoLoopSecond(c->d, c->S, c->Y, n, ncorr, loc, mc, sc + kGram, sc + kYs, sc + kAlpha,
                                           sc + kCoef, c->d, c->work, c->counter, c->stream);
        } else {
            // two-loop recursion: d = -H r (LBFGSMat::apply_Hv)
            efem::dev::launchScale(c->d, c->r, -1.0, n, c->stream);
            int j = ptr % mc, prev = -1;
            for (int i = 0; i < ncorr; ++i) {
                j = (j + mc - 1) % mc;
                efem::dev::launchTwoLoopStep(c->d, prev >= 0 ? c->Y + static_cast<size_t>(prev) * n : nullptr,
                                             prev >= 0 ? sc + kAlpha + prev : nullptr, c->S + static_cast<size_t>(j) * n,
                                             sc + kYs + j, nullptr, sc + kAlpha + j, 0, n, c->work, c->counter,
                                             c->stream);
                prev = j;
            }
            if (prev >= 0)
                efem::dev::launchAxpyDev(c->d, c->Y + static_cast<size_t>(prev) * n, sc + kAlpha + prev, 1, n, c->stream);
            if (h0) {
                if ((rc = applyH0Internal(c, c->d, c->tmp))) return rc;
                std::swap(c->d, c->tmp);
            } else {
                efem::dev::launchDivDev(c->d, sc + kTheta, n, c->stream);
            }
            prev = -1;
            for (int i = 0; i < ncorr; ++i) {
                efem::dev::launchTwoLoopStep(c->d, prev >= 0 ? c->S + static_cast<size_t>(prev) * n : nullptr,
                                             prev >= 0 ? sc + kCoef + prev : nullptr, c->Y + static_cast<size_t>(j) * n,
                                             sc + kYs + j, sc + kAlpha + j, sc + kCoef + j, 1, n, c->work, c->counter,
                                             c->stream);
                prev = j;
                j = (j + 1) % mc;
            }
            if (prev >= 0)
                efem::dev::launchAxpyDev(c->d, c->S + static_cast<size_t>(prev) * n, sc + kCoef + prev, 0, n, c->stream);
        }
        HIP_CHECK(c, hipGetLastError());
        ++k;
    }
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
    c->err = "efem_set_option: unknown option " + k;
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

}  // extern "C"
