// Dense kernels of the host multifrontal factorisation, on column-major
// blocks (element (i, j) at A[i + j * ld]).  All heavy work goes through
// one packed GEMM whose register tile fits the CPU's vector unit (AVX-512,
// AVX2 or portable C++, picked at run time); the triangular operations are
// blocked so that their O(n^3) part is GEMM.
#pragma once

namespace efem {
namespace dense {

// name of the micro-kernel in use ("avx512 16x8", "avx2 8x6", "portable 8x6")
const char* kernelName();

// C(mi x nj) += alpha * A(mi x kk) * op(B), op(B)(k, j) = B[k + j ldb]
// (bT = false) or B[j + k ldb] (bT = true)
void gemm(double* C, int ldc, const double* A, int lda, const double* B, int ldb, bool bT, int mi, int nj, int kk,
          double alpha);

// in-place lower Cholesky of the leading n x n block; false if a pivot is
// not positive
// par: the panel updates and solves split into OpenMP tasks over 128-row
// (96-column) ranges aligned to the packing blocks, so every entry sees the
// same arithmetic as without them
bool potrf(double* A, int n, int ld, bool par = false);

// rows [i0, i1) of X (m x n block at X, ld) <- X L^-T, L the lower n x n
// block at L (ldl): X = A21 L11^-T of the supernode
void trsmRightLT(double* X, int ld, const double* L, int ldl, int n, int i0, int i1);

// Linv (n x n, ldi, lower) <- L^-1, L lower n x n (ldl); the strict upper
// part of Linv is set to 0
void trtri(const double* L, int ldl, double* Linv, int ldi, int n, bool par = false);

// M (rows x n, ldm) += L21 Linv, L21 rows x n (ldl), Linv lower n x n (ldi):
// M(i, k) = sum_{p >= k} L21(i, p) Linv(p, k)
void trmmRightLower(double* M, int ldm, const double* L21, int ldl, const double* Linv, int ldi, int n, int rows);

}  // namespace dense
}  // namespace efem
