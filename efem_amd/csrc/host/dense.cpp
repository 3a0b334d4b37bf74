// Dense kernels (see dense.hpp).  Blocked formulations of the LAPACK-style
// operations written from the textbook algorithms (Golub & Van Loan,
// "Matrix Computations", ch. 3-4); the GEMM is a packed-panel kernel in the
// style of Goto & van de Geijn ("Anatomy of High-Performance Matrix
// Multiplication", ACM TOMS 34(3), 2008): A and B blocks are copied into
// contiguous register-tile strips, and a micro-kernel picked at run time
// (AVX-512 16 x 8, AVX2 8 x 6, portable 8 x 6) accumulates one tile.
#include "dense.hpp"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

namespace efem {
namespace dense {
namespace {

constexpr int NB = 32;   // block size of the triangular algorithms
constexpr int KC = 256;  // k block of the packed panels
constexpr int MC = 128;  // rows of a packed A block (multiple of every MR)
constexpr int NC = 96;   // columns of a packed B block (multiple of every NR)

// micro-kernel: C(MR x NR, ldc) += Ap Bp over kc, Ap packed [k][MR], Bp
// packed [k][NR] (alpha folded into Bp)
using Micro = void (*)(int kc, const double* Ap, const double* Bp, double* C, int ldc);

template <int MR, int NR>
void microPortable(int kc, const double* Ap, const double* Bp, double* C, int ldc)
{
    double acc[NR][MR] = {};
    for (int k = 0; k < kc; ++k)
        for (int j = 0; j < NR; ++j) {
            const double b = Bp[k * NR + j];
            for (int i = 0; i < MR; ++i) acc[j][i] += Ap[k * MR + i] * b;
        }
    for (int j = 0; j < NR; ++j)
        for (int i = 0; i < MR; ++i) C[i + static_cast<size_t>(j) * ldc] += acc[j][i];
}

#if defined(__x86_64__)
// 8 x 6: 12 ymm accumulators, two A vectors and one broadcast per k
__attribute__((target("avx2,fma"))) void microAvx2(int kc, const double* Ap, const double* Bp, double* C, int ldc)
{
    __m256d acc[6][2];
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[j][0] = acc[j][1] = _mm256_setzero_pd();
    for (int k = 0; k < kc; ++k) {
        const __m256d a0 = _mm256_load_pd(Ap + 8 * k), a1 = _mm256_load_pd(Ap + 8 * k + 4);
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const __m256d b = _mm256_broadcast_sd(Bp + 6 * k + j);
            acc[j][0] = _mm256_fmadd_pd(a0, b, acc[j][0]);
            acc[j][1] = _mm256_fmadd_pd(a1, b, acc[j][1]);
        }
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double* cj = C + static_cast<size_t>(j) * ldc;
        _mm256_storeu_pd(cj, _mm256_add_pd(_mm256_loadu_pd(cj), acc[j][0]));
        _mm256_storeu_pd(cj + 4, _mm256_add_pd(_mm256_loadu_pd(cj + 4), acc[j][1]));
    }
}

// 16 x 8: 16 zmm accumulators, two A vectors and eight broadcasts per k
__attribute__((target("avx512f"))) void microAvx512(int kc, const double* Ap, const double* Bp, double* C, int ldc)
{
    __m512d acc[8][2];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j][0] = acc[j][1] = _mm512_setzero_pd();
    for (int k = 0; k < kc; ++k) {
        const __m512d a0 = _mm512_load_pd(Ap + 16 * k), a1 = _mm512_load_pd(Ap + 16 * k + 8);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const __m512d b = _mm512_set1_pd(Bp[8 * k + j]);
            acc[j][0] = _mm512_fmadd_pd(a0, b, acc[j][0]);
            acc[j][1] = _mm512_fmadd_pd(a1, b, acc[j][1]);
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        double* cj = C + static_cast<size_t>(j) * ldc;
        _mm512_storeu_pd(cj, _mm512_add_pd(_mm512_loadu_pd(cj), acc[j][0]));
        _mm512_storeu_pd(cj + 8, _mm512_add_pd(_mm512_loadu_pd(cj + 8), acc[j][1]));
    }
}
#endif

struct Kernel {
    int mr, nr;
    Micro micro;
};

// the widest kernel the CPU runs; EFEM_DENSE_KERNEL=avx2|portable caps it
// (read once per process)
const Kernel& kernel()
{
    static const Kernel k = [] {
        const char* cap = std::getenv("EFEM_DENSE_KERNEL");
        const std::string want = cap ? cap : "";
#if defined(__x86_64__)
        if (want != "avx2" && want != "portable" && __builtin_cpu_supports("avx512f"))
            return Kernel{16, 8, microAvx512};
        if (want != "portable" && __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma"))
            return Kernel{8, 6, microAvx2};
#endif
        return Kernel{8, 6, microPortable<8, 6>};
    }();
    return k;
}

}  // namespace

const char* kernelName()
{
    const Kernel& k = kernel();
#if defined(__x86_64__)
    if (k.micro == microAvx512) return "avx512 16x8";
    if (k.micro == microAvx2) return "avx2 8x6";
#endif
    return "portable 8x6";
}

namespace {

// per-thread packing buffers (gemm has no OpenMP scheduling point inside,
// so a task never shares them)
struct PackBuffers {
    double* a = nullptr;
    double* b = nullptr;
    PackBuffers()
    {
        a = static_cast<double*>(::operator new(sizeof(double) * MC * KC, std::align_val_t(64)));
        b = static_cast<double*>(::operator new(sizeof(double) * KC * NC, std::align_val_t(64)));
    }
    ~PackBuffers()
    {
        ::operator delete(a, std::align_val_t(64));
        ::operator delete(b, std::align_val_t(64));
    }
};

// rows [i0, i0 + mc) x k [k0, k0 + kc) of A into MR-row strips [k][MR], zero
// rows past mc
void packA(double* Ap, const double* A, int lda, int i0, int mc, int k0, int kc, int mr)
{
    for (int s = 0; s < mc; s += mr) {
        const int rows = std::min(mr, mc - s);
        double* dst = Ap + static_cast<size_t>(s) * kc;
        for (int k = 0; k < kc; ++k) {
            const double* src = A + (i0 + s) + static_cast<size_t>(k0 + k) * lda;
            double* d = dst + k * mr;
            int i = 0;
            for (; i < rows; ++i) d[i] = src[i];
            for (; i < mr; ++i) d[i] = 0.0;
        }
    }
}

// op(B) rows k [k0, k0 + kc) x columns [j0, j0 + nc) into NR-column strips
// [k][NR] times alpha, zero columns past nc
void packB(double* Bp, const double* B, int ldb, bool bT, int j0, int nc, int k0, int kc, int nr, double alpha)
{
    for (int s = 0; s < nc; s += nr) {
        const int cols = std::min(nr, nc - s);
        double* dst = Bp + static_cast<size_t>(s) * kc;
        if (bT) {
            for (int k = 0; k < kc; ++k) {
                const double* src = B + (j0 + s) + static_cast<size_t>(k0 + k) * ldb;
                double* d = dst + k * nr;
                int j = 0;
                for (; j < cols; ++j) d[j] = alpha * src[j];
                for (; j < nr; ++j) d[j] = 0.0;
            }
        } else {
            for (int j = 0; j < nr; ++j) {
                if (j < cols) {
                    const double* src = B + k0 + static_cast<size_t>(j0 + s + j) * ldb;
                    for (int k = 0; k < kc; ++k) dst[k * nr + j] = alpha * src[k];
                } else {
                    for (int k = 0; k < kc; ++k) dst[k * nr + j] = 0.0;
                }
            }
        }
    }
}

}  // namespace

void gemm(double* C, int ldc, const double* A, int lda, const double* B, int ldb, bool bT, int mi, int nj, int kk,
          double alpha)
{
    if (mi <= 0 || nj <= 0 || kk <= 0) return;
    const Kernel& K = kernel();
    thread_local PackBuffers buf;
    alignas(64) double edge[16 * 8];
    for (int jc = 0; jc < nj; jc += NC) {
        const int nc = std::min(NC, nj - jc);
        for (int pc = 0; pc < kk; pc += KC) {
            const int kc = std::min(KC, kk - pc);
            packB(buf.b, B, ldb, bT, jc, nc, pc, kc, K.nr, alpha);
            for (int ic = 0; ic < mi; ic += MC) {
                const int mc = std::min(MC, mi - ic);
                packA(buf.a, A, lda, ic, mc, pc, kc, K.mr);
                for (int jr = 0; jr < nc; jr += K.nr) {
                    const int nr = std::min(K.nr, nc - jr);
                    const double* Bp = buf.b + static_cast<size_t>(jr) * kc;
                    for (int ir = 0; ir < mc; ir += K.mr) {
                        const int mr = std::min(K.mr, mc - ir);
                        const double* Ap = buf.a + static_cast<size_t>(ir) * kc;
                        double* Cij = C + (ic + ir) + static_cast<size_t>(jc + jr) * ldc;
                        if (mr == K.mr && nr == K.nr) {
                            K.micro(kc, Ap, Bp, Cij, ldc);
                            continue;
                        }
                        std::fill(edge, edge + K.mr * K.nr, 0.0);
                        K.micro(kc, Ap, Bp, edge, K.mr);
                        for (int j = 0; j < nr; ++j)
                            for (int i = 0; i < mr; ++i) Cij[i + static_cast<size_t>(j) * ldc] += edge[i + j * K.mr];
                    }
                }
            }
        }
    }
}

// unblocked Cholesky of an n x n block (n <= NB typically)
static bool potrfUnblocked(double* A, int n, int ld)
{
    for (int j = 0; j < n; ++j) {
        double* cj = A + static_cast<size_t>(j) * ld;
        for (int k = 0; k < j; ++k) {
            const double* ck = A + static_cast<size_t>(k) * ld;
            const double s = ck[j];
            if (s == 0.0) continue;
            for (int i = j; i < n; ++i) cj[i] -= ck[i] * s;
        }
        const double d = cj[j];
        if (!(d > 0.0)) return false;
        const double r = std::sqrt(d);
        cj[j] = r;
        const double ir = 1.0 / r;
        for (int i = j + 1; i < n; ++i) cj[i] *= ir;
    }
    return true;
}

bool potrf(double* A, int n, int ld, bool par)
{
    // left-looking by column blocks: A[j:, J] -= A[j:, :j] A[J, :j]^T, then
    // factor the diagonal block and solve the rows below it
    for (int j = 0; j < n; j += NB) {
        const int jb = std::min(NB, n - j);
        double* Ajj = A + j + static_cast<size_t>(j) * ld;
        const int rows = n - j;
        if (par && rows > MC && j > 0) {
            // row ranges of MC (the packing block): the same tiles per entry
#pragma omp taskloop grainsize(1)
            for (int r0 = 0; r0 < rows; r0 += MC)
                gemm(Ajj + r0, ld, A + j + r0, ld, A + j, ld, true, std::min(MC, rows - r0), jb, j, -1.0);
        } else {
            gemm(Ajj, ld, A + j, ld, A + j, ld, true, rows, jb, j, -1.0);
        }
        if (!potrfUnblocked(Ajj, jb, ld)) return false;
        if (j + jb < n) {
            if (par && rows - jb > MC) {
                // independent rows
#pragma omp taskloop grainsize(1)
                for (int r0 = jb; r0 < rows; r0 += MC)
                    trsmRightLT(Ajj, ld, Ajj, ld, jb, r0, std::min(rows, r0 + MC));
            } else {
                trsmRightLT(Ajj, ld, Ajj, ld, jb, jb, rows);
            }
        }
    }
    return true;
}

void trsmRightLT(double* X, int ld, const double* L, int ldl, int n, int i0, int i1)
{
    if (i1 <= i0) return;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = std::min(NB, n - k0);
        // X[:, K] -= X[:, :k0] L[K, :k0]^T
        gemm(X + i0 + static_cast<size_t>(k0) * ld, ld, X + i0, ld, L + k0, ldl, true, i1 - i0, kb, k0, -1.0);
        // X[:, K] <- X[:, K] L[K, K]^-T (column by column)
        for (int k = k0; k < k0 + kb; ++k) {
            double* xk = X + static_cast<size_t>(k) * ld;
            for (int p = k0; p < k; ++p) {
                const double s = L[k + static_cast<size_t>(p) * ldl];
                if (s == 0.0) continue;
                const double* xp = X + static_cast<size_t>(p) * ld;
                for (int i = i0; i < i1; ++i) xk[i] -= xp[i] * s;
            }
            const double ir = 1.0 / L[k + static_cast<size_t>(k) * ldl];
            for (int i = i0; i < i1; ++i) xk[i] *= ir;
        }
    }
}

void trtri(const double* L, int ldl, double* Linv, int ldi, int n, bool par)
{
    // X = L^-1 by row blocks (forward substitution on the identity):
    // X[I, :i0] = -L[I, I]^-1 (L[I, :i0] X[:i0, :i0]), X[I, I] = L[I, I]^-1
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) Linv[i + static_cast<size_t>(j) * ldi] = 0.0;
    for (int i0 = 0; i0 < n; i0 += NB) {
        const int ib = std::min(NB, n - i0);
        double* XI = Linv + i0;  // rows I of X
        // X[I, :i0] = -(L[I, :i0] X[:i0, :i0]), in column ranges of NC (the
        // packing block: the same tiles per entry) when split into tasks
        if (par && i0 > NC) {
#pragma omp taskloop grainsize(1)
            for (int c0 = 0; c0 < i0; c0 += NC)
                gemm(XI + static_cast<size_t>(c0) * ldi, ldi, L + i0, ldl, Linv + static_cast<size_t>(c0) * ldi, ldi,
                     false, ib, std::min(NC, i0 - c0), i0, -1.0);
        } else {
            gemm(XI, ldi, L + i0, ldl, Linv, ldi, false, ib, i0, i0, -1.0);
        }
        // diagonal block: X[I, I] = L[I, I]^-1 (columns of the identity)
        for (int k = 0; k < ib; ++k) XI[k + static_cast<size_t>(i0 + k) * ldi] = 1.0;
        // X[I, :i0 + ib] <- L[I, I]^-1 X[I, :i0 + ib]  (lower solve per column)
        auto solveCols = [&](int c0, int c1) {
            for (int c = c0; c < c1; ++c) {
                double* x = XI + static_cast<size_t>(c) * ldi;
                for (int r = 0; r < ib; ++r) {
                    double v = x[r];
                    for (int p = 0; p < r; ++p) v -= L[(i0 + r) + static_cast<size_t>(i0 + p) * ldl] * x[p];
                    x[r] = v / L[(i0 + r) + static_cast<size_t>(i0 + r) * ldl];
                }
            }
        };
        if (par && i0 + ib > 2 * NC) {
#pragma omp taskloop grainsize(1)
            for (int c0 = 0; c0 < i0 + ib; c0 += NC) solveCols(c0, std::min(i0 + ib, c0 + NC));
        } else {
            solveCols(0, i0 + ib);
        }
    }
}

void trmmRightLower(double* M, int ldm, const double* L21, int ldl, const double* Linv, int ldi, int n, int rows)
{
    if (rows <= 0) return;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = std::min(NB, n - k0), k1 = k0 + kb;
        // full part: M[:, K] += L21[:, k1:] Linv[k1:, K]
        gemm(M + static_cast<size_t>(k0) * ldm, ldm, L21 + static_cast<size_t>(k1) * ldl, ldl,
             Linv + k1 + static_cast<size_t>(k0) * ldi, ldi, false, rows, kb, n - k1, 1.0);
        // triangular part: M[:, k] += sum_{p in [k, k1)} L21[:, p] Linv[p, k]
        for (int k = k0; k < k1; ++k) {
            double* mk = M + static_cast<size_t>(k) * ldm;
            for (int p = k; p < k1; ++p) {
                const double s = Linv[p + static_cast<size_t>(k) * ldi];
                if (s == 0.0) continue;
                const double* lp = L21 + static_cast<size_t>(p) * ldl;
                for (int i = 0; i < rows; ++i) mk[i] += lp[i] * s;
            }
        }
    }
}

}  // namespace dense
}  // namespace efem
