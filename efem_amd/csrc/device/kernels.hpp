// Device kernels of the solve loop (gfx950, wave64).  Launch wrappers only;
// the kernels live in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace efem {
namespace dev {

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
};
void launchElementForce(const ElemArgs& a, hipStream_t s);

struct FaceArgs {
    int nFaces, nen, fq, directed;
    double pressure, pdir[3];
    const int* faceDof;     // [nen][F] free id or nFree
    const double* A0;       // [F][3]
    const double* WN;       // [F][fq][nen]
    const double* G;        // [F][fq][nen][3]
    const double* x;
    double* slots;          // [F][nen][3]
};
void launchPressure(const FaceArgs& a, hipStream_t s);

struct GatherArgs {
    int nFree, nElemSlots, mixed;
    const int* gptr;
    const int* gslot;
    const void* eslots;     // element slots (double or float)
    const double* fslots;   // face slots
    const double* fext;     // constant nodal forces
    double* r;
};
void launchGather(const GatherArgs& a, hipStream_t s);

// ---- vectors ----------------------------------------------------------------
// deterministic reductions: results written to device scalars; the block
// partials and an arrival counter live in `work` (see kReduceWorkDoubles)
constexpr int kReduceBlocks = 512;
constexpr int kReduceWorkDoubles = 4 * kReduceBlocks + 16;

// out[0] = a.b, out[1] = max|a|, out[2] = max|c|  (c may be null)
void launchDotMax(const double* a, const double* b, const double* c, int n, double* work,
                  unsigned* counter, double* out, hipStream_t s);
// *ab = a.b, *bb = b.b
void launchDot2(const double* a, const double* b, int n, double* work, unsigned* counter,
                double* ab, double* bb, hipStream_t s);
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
                       double* work, unsigned* counter, hipStream_t s);
// q += (*c) * v   (or q -= when neg)
void launchAxpyDev(double* q, const double* v, const double* c, int neg, int n, hipStream_t s);
// q /= (*theta)
void launchDivDev(double* q, const double* theta, int n, hipStream_t s);
// *theta = yy / ys
void launchTheta(const double* ys, const double* yy, double* theta, hipStream_t s);

// ---- supernodal triangular solves (3 right-hand sides) ---------------------
struct TrsvArgs {
    int n, mixed;
    const int* perm;        // [n] new -> old
    const int* snFirst;     // [S]
    const int* snW;         // [S]
    const int* snM;         // [S]
    const int* snRow;       // [S] offset into rows
    const long long* snVal; // [S] offset into B
    const int* snT;         // [S] offset into t (rows)
    const int* rows;
    const void* B;          // double or float
    const int* cptr;        // [n+1] contributions to new row j
    const int* cidx;        // t row positions
    void* y;                // [n][3] work (T)
    void* t;                // [nt][3] work (T)
    void* z;                // [n][3] work (T)
};
// one level of the forward solve over work items (supernode, first row)
void launchTrsvForward(const TrsvArgs& a, const double* b, const int2* items, int nitems,
                       hipStream_t s);
void launchTrsvBackward(const TrsvArgs& a, const int2* items, int nitems, hipStream_t s);
// out[3 perm[j] + c] = z[j][c]
void launchUnpermute(const TrsvArgs& a, double* out, hipStream_t s);

}  // namespace dev
}  // namespace efem
