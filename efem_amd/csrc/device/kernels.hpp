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
constexpr int kMaxCompact = 16;  // history size handled by the compact two-loop
constexpr int kReduceWorkDoubles = (2 * kMaxCompact + 2) * kReduceBlocks + 64;

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
// compact two-loop (same recursion as LBFGSMat::apply_Hv, evaluated from
// Gram products).  History slots 0..nc-1 of the ring S, Y ([m][n]), newest
// slot `newest`; G[i*16+k] = s_i . y_k; rho[i] = 1 / (s_i . y_i).
// Gram update after a new pair at slot loc: G[loc][k], G[k][loc], ys, yy.
void launchGramUpdate(const double* S, const double* Y, int n, int nc, int loc, double* G, double* ys,
                      double* yy, double* work, unsigned* counter, hipStream_t s);
// loop 1: alpha_i = (s_i . q_i) / ys_i newest -> oldest with q_0 = -r, then
// out = -r - sum alpha_i y_i
void launchTwoLoopFirst(const double* r, const double* S, const double* Y, int n, int nc, int newest,
                        int mcap, const double* G, const double* ys, double* alpha, double* out, double* work,
                        unsigned* counter, hipStream_t s);
// loop 2: beta_i = (y_i . r_i) / ys_i oldest -> newest with r_0 = h (the H0
// output), then out = h + sum (alpha_i - beta_i) s_i
void launchTwoLoopSecond(const double* h, const double* S, const double* Y, int n, int nc, int newest,
                         int mcap, const double* G, const double* ys, const double* alpha, double* coef,
                         double* out, double* work, unsigned* counter, hipStream_t s);
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
    const int* snT;         // [S] offset into the update vectors (rows)
    const int* rows;
    const void* B;          // double or float
    const int* fptr;        // [sum m_s + 1] per front row (snRow[s] + r): children's
    const int* fidx;        //   update rows that extend-add into it
    void* y;                // [n][3] work (T)
    void* t;                // [nt][3] update vectors u_s (T)
    void* z;                // [n][3] work (T)
};
// one level of the forward solve over work items (supernode, first row):
// multifrontal form f_s = [b_s; 0] + sum_children u_c (extend-add),
// y_s = L11^-1 f_s[:w], u_s = f_s[w:] - (L21 L11^-1) f_s[:w]
void launchTrsvForward(const TrsvArgs& a, const double* b, const int2* items, int nitems,
                       hipStream_t s);
void launchTrsvBackward(const TrsvArgs& a, const int2* items, int nitems, hipStream_t s);
// out[3 perm[j] + c] = z[j][c]
void launchUnpermute(const TrsvArgs& a, double* out, hipStream_t s);

}  // namespace dev
}  // namespace efem
