// Device numeric factorisation of the approximate Jacobian (the scalar
// Laplacian, K = L (x) I3): supernodal multifrontal LL^T in FP64, level by
// level from the leaves, on the structure of the host symbolic analysis,
// producing the block inverses B_s = [L11^-1; L21 L11^-1] in the layout the
// triangular solves read (replaces the host numeric phase and the upload of
// updateApproxJacobi, src/simulation/nlsolver/LBFGS.h:65-81).
//
// Every dense step works on 64 x 64 tiles of a front, batched over the
// fronts of a level (one workgroup per tile task, FP64 FMA on the vector
// ALUs; the tiles of one task are staged in LDS).  No atomics on values:
// every front entry is written by one task per launch, extend-add walks a
// front column's contributions in children order.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace efem {
namespace dev {
namespace {

constexpr int TB = 64;  // tile edge

struct FacArgs {
    double* L;               // panels [L11; L21]
    double* X;               // block inverses (FP64 B)
    double* U;               // update matrices
    const double* A;         // values of A (the device-assembled Laplacian)
    const int* w;
    const int* m;
    const long long* dval;
    const long long* uOff;
    const int* rel;
    const int* relOff;
    const int* pairs;
    const long long* asmDst;
    const int* asmSrc;
    const int4* tasks;
    int* fail;
};

__global__ __launch_bounds__(256) void kFacAsm(FacArgs a, int first, int count)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) a.L[a.asmDst[first + i]] = a.A[a.asmSrc[first + i]];
}

// front column t of s += its children's contributions, children in order
__global__ __launch_bounds__(256) void kFacExtAdd(FacArgs a, int first)
{
    const int4 t4 = a.tasks[first + blockIdx.x];
    const int s = t4.x, t = t4.y;
    const int ws = a.w[s], ms = a.m[s], nr = ms - ws, ld = blockLd(ms);
    for (int p = t4.z; p < t4.w; ++p) {
        const int c = a.pairs[2 * p], b = a.pairs[2 * p + 1];
        const int cm = a.m[c] - a.w[c];
        const double* src = a.U + a.uOff[c] + static_cast<size_t>(b) * cm;
        const int* rel = a.rel + a.relOff[c];
        for (int q = b + threadIdx.x; q < cm; q += 256) {
            const int row = rel[q];
            if (t < ws)
                a.L[a.dval[s] + row + static_cast<size_t>(t) * ld] += src[q];
            else
                a.U[a.uOff[s] + (row - ws) + static_cast<size_t>(t - ws) * nr] += src[q];
        }
        __syncthreads();  // the next child may add to the same rows
    }
}

// Cholesky of the diagonal tile k of the panel (kw x kw), in LDS
__global__ __launch_bounds__(256) void kFacPotrf(FacArgs a, int first)
{
    __shared__ double T[TB][TB + 1];
    const int4 t4 = a.tasks[first + blockIdx.x];
    const int s = t4.x, k = t4.y;
    const int ws = a.w[s], ld = blockLd(a.m[s]);
    const int c0 = TB * k, kw = min(TB, ws - c0);
    double* P = a.L + a.dval[s] + c0 + static_cast<size_t>(c0) * ld;
    const int tid = threadIdx.x;
    for (int e = tid; e < TB * TB; e += 256) {
        const int i = e % TB, j = e / TB;
        T[i][j] = (i < kw && j < kw && i >= j) ? P[i + static_cast<size_t>(j) * ld] : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < kw; ++c) {
        if (tid == 0) {
            const double d = T[c][c];
            if (!(d > 0.0)) {
                atomicOr(a.fail, 1);
                T[c][c] = 1.0;
            } else {
                T[c][c] = sqrt(d);
            }
        }
        __syncthreads();
        const double dc = T[c][c];
        for (int r = c + 1 + tid; r < kw; r += 256) T[r][c] /= dc;
        __syncthreads();
        // trailing update of the tile's lower part
        const int nt = kw - c - 1;
        for (int e = tid; e < nt * nt; e += 256) {
            const int r = c + 1 + e % nt, q = c + 1 + e / nt;
            if (r >= q) T[r][q] -= T[r][c] * T[q][c];
        }
        __syncthreads();
    }
    for (int e = tid; e < TB * TB; e += 256) {
        const int i = e % TB, j = e / TB;
        if (i < kw && j < kw && i >= j) P[i + static_cast<size_t>(j) * ld] = T[i][j];
    }
}

// rows [r0, r0 + 64) of the panel's column tile k <- (rows) L_kk^-T
__global__ __launch_bounds__(64) void kFacTrsm(FacArgs a, int first)
{
    __shared__ double Lk[TB][TB + 1];
    const int4 t4 = a.tasks[first + blockIdx.x];
    const int s = t4.x, k = t4.y, r0 = t4.z;
    const int ws = a.w[s], ms = a.m[s], ld = blockLd(ms);
    const int c0 = TB * k, kw = min(TB, ws - c0);
    double* P = a.L + a.dval[s];
    const int lane = threadIdx.x;
    for (int j = 0; j < TB; ++j) Lk[lane][j] = (lane < kw && j < kw && lane >= j) ?
                                                   P[c0 + lane + static_cast<size_t>(c0 + j) * ld] : 0.0;
    __syncthreads();
    const int r = r0 + lane;
    if (r >= ms) return;
    double x[TB];
#pragma unroll
    for (int j = 0; j < TB; ++j) x[j] = j < kw ? P[r + static_cast<size_t>(c0 + j) * ld] : 0.0;
    // forward substitution: x_j = (x_j - sum_{c<j} x_c L_jc) / L_jj
#pragma unroll
    for (int j = 0; j < TB; ++j) {
        if (j < kw) {
            double v = x[j];
#pragma unroll
            for (int c = 0; c < j; ++c) v -= x[c] * Lk[j][c];
            x[j] = v / Lk[j][j];
        }
    }
#pragma unroll
    for (int j = 0; j < TB; ++j)
        if (j < kw) P[r + static_cast<size_t>(c0 + j) * ld] = x[j];
}

// 64 x 64 tile product C -= A B^T over kk columns; A rows ra.., B rows rb..
// (both ld, columns from ca), C at (rc, cc) of the same ld-matrix or given
// pointer; threads in a 16 x 16 grid, 4 x 4 outputs each
struct TileAcc {
    double v[4][4];
};

__device__ __forceinline__ void tileMulAbt(const double* __restrict__ Ap, size_t lda, int arows,
                                           const double* __restrict__ Bp, size_t ldb, int brows, int kk,
                                           double (*As)[TB + 1], double (*Bs)[TB + 1], TileAcc& acc)
{
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
    for (int k0 = 0; k0 < kk; k0 += TB) {
        const int kc = min(TB, kk - k0);
        __syncthreads();
        for (int e = tid; e < TB * TB; e += 256) {
            const int i = e % TB, j = e / TB;
            As[i][j] = (i < arows && j < kc) ? Ap[i + static_cast<size_t>(k0 + j) * lda] : 0.0;
            Bs[i][j] = (i < brows && j < kc) ? Bp[i + static_cast<size_t>(k0 + j) * ldb] : 0.0;
        }
        __syncthreads();
        for (int j = 0; j < kc; ++j) {
            double av[4], bv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = As[4 * tr + q][j];
                bv[q] = Bs[4 * tc + q][j];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc.v[p][q] += av[p] * bv[q];
        }
    }
}

// panel update: tile (it, jt) of the panel (jt > k, inside the w columns)
// -= (rows of it, cols of k) (rows of jt, cols of k)^T
__global__ __launch_bounds__(256) void kFacUpdate(FacArgs a, int first)
{
    __shared__ double As[TB][TB + 1], Bs[TB][TB + 1];
    const int4 t4 = a.tasks[first + blockIdx.x];
    const int s = t4.x, k = t4.y, it = t4.z, jt = t4.w;
    const int ws = a.w[s], ms = a.m[s], ld = blockLd(ms);
    const int c0 = TB * k, kw = min(TB, ws - c0);
    const int ri = TB * it, rj = TB * jt;
    double* P = a.L + a.dval[s];
    TileAcc acc{};
    tileMulAbt(P + ri + static_cast<size_t>(c0) * ld, ld, min(TB, ms - ri), P + rj + static_cast<size_t>(c0) * ld,
               ld, min(TB, ws - rj), kw, As, Bs, acc);
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = ri + 4 * tr + p, j = rj + 4 * tc + q;
            if (i < ms && j < ws && i >= j) P[i + static_cast<size_t>(j) * ld] -= acc.v[p][q];
        }
}

// update matrix tile (it, jt) of U_s -= L21(it) L21(jt)^T over the w columns
__global__ __launch_bounds__(256) void kFacSyrk(FacArgs a, int first)
{
    __shared__ double As[TB][TB + 1], Bs[TB][TB + 1];
    const int4 t4 = a.tasks[first + blockIdx.x];
    const int s = t4.x, it = t4.y, jt = t4.z;
    const int ws = a.w[s], ms = a.m[s], nr = ms - ws, ld = blockLd(ms);
    const int ri = TB * it, rj = TB * jt;
    const double* P = a.L + a.dval[s] + ws;  // L21
    TileAcc acc{};
    tileMulAbt(P + ri, ld, min(TB, nr - ri), P + rj, ld, min(TB, nr - rj), ws, As, Bs, acc);
    double* Us = a.U + a.uOff[s];
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = ri + 4 * tr + p, j = rj + 4 * tc + q;
            if (i < nr && j < nr && i >= j) Us[i + static_cast<size_t>(j) * nr] -= acc.v[p][q];
        }
}

// block inverse, column tile J, rows of tile it: X = R L11^-1 with
// R = [I; L21]: y = R(rows, J) - sum_{K > J} X(rows, K) L11(K, J), then
// x L_JJ = y by backward substitution over the tile's columns
__global__ __launch_bounds__(256) void kFacBsolve(FacArgs a, int first)
{
    __shared__ double As[TB][TB + 1], Bs[TB][TB + 1];
    const int4 t4 = a.tasks[first + blockIdx.x];
    const int s = t4.x, J = t4.y, it = t4.z;
    const int ws = a.w[s], ms = a.m[s], ld = blockLd(ms);
    const int c0 = TB * J, kw = min(TB, ws - c0);
    const int ri = TB * it, rows = min(TB, ms - ri);
    const double* P = a.L + a.dval[s];
    double* X = a.X + a.dval[s];
    // sum_{K > J} X(rows, K) L11(K, J) = X(rows, c0 + kw .. ws) * (L11(c0 + kw .., J))
    // as C = A B^T with A = X rows, B = L11(J cols) transposed: B(j, k) =
    // L11(c0 + kw + k, c0 + j); done here with an explicit loop over the
    // trailing columns through the same tile machinery on transposed staging
    TileAcc acc{};
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
    for (int k0 = c0 + kw; k0 < ws; k0 += TB) {
        const int kc = min(TB, ws - k0);
        __syncthreads();
        for (int e = tid; e < TB * TB; e += 256) {
            const int i = e % TB, j = e / TB;
            // As[i][j] = X(ri + i, k0 + j); Bs[q][j] = L11(k0 + j, c0 + q)
            As[i][j] = (i < rows && j < kc) ? X[ri + i + static_cast<size_t>(k0 + j) * ld] : 0.0;
            Bs[i][j] = (i < kw && j < kc) ? P[k0 + j + static_cast<size_t>(c0 + i) * ld] : 0.0;
        }
        __syncthreads();
        for (int j = 0; j < kc; ++j) {
            double av[4], bv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = As[4 * tr + q][j];
                bv[q] = Bs[4 * tc + q][j];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc.v[p][q] += av[p] * bv[q];
        }
    }
    __syncthreads();
    // y into As (rows x kw), L_JJ into Bs
    for (int e = tid; e < TB * TB; e += 256) {
        const int i = e % TB, j = e / TB;
        Bs[i][j] = (i < kw && j < kw && i >= j) ? P[c0 + i + static_cast<size_t>(c0 + j) * ld] : 0.0;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = 4 * tr + p, j = 4 * tc + q, gi = ri + i, gj = c0 + j;
            double r = 0.0;
            if (i < rows && j < kw) r = gi < ws ? (gi == gj ? 1.0 : 0.0) : P[gi + static_cast<size_t>(gj) * ld];
            As[i][j] = r - acc.v[p][q];
        }
    __syncthreads();
    if (tid >= TB) return;
    const int i = tid;
    if (i >= rows) return;
    double x[TB];
#pragma unroll
    for (int j = 0; j < TB; ++j) x[j] = As[i][j];
    // x_j = (y_j - sum_{c > j} x_c L_cj) / L_jj, j from the last column down
#pragma unroll
    for (int jj = 0; jj < TB; ++jj) {
        const int j = TB - 1 - jj;
        if (j < kw) {
            double v = x[j];
#pragma unroll
            for (int c = j + 1; c < TB; ++c)
                if (c < kw) v -= x[c] * Bs[c][j];
            x[j] = v / Bs[j][j];
        }
    }
#pragma unroll
    for (int j = 0; j < TB; ++j)
        if (j < kw) X[ri + i + static_cast<size_t>(c0 + j) * ld] = x[j];
}

__global__ __launch_bounds__(256) void kFacToF32(const double* __restrict__ X, float* __restrict__ B, long long n)
{
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) B[i] = static_cast<float>(X[i]);
}

// M = X^T X of one w x w block (ld), tile (it, jt); M column-major, ld
__global__ __launch_bounds__(256) void kFacGram(const double* __restrict__ Xr, int wr, int ld, double* __restrict__ M)
{
    __shared__ double As[TB][TB + 1], Bs[TB][TB + 1];
    const int it = blockIdx.x, jt = blockIdx.y;
    const int ri = TB * it, rj = TB * jt;
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
    TileAcc acc{};
    // M(i, j) = sum_k X(k, i) X(k, j), k >= max(i, j) (X lower triangular)
    for (int k0 = min(ri, rj) / TB * TB; k0 < wr; k0 += TB) {
        const int kc = min(TB, wr - k0);
        __syncthreads();
        for (int e = tid; e < TB * TB; e += 256) {
            const int i = e % TB, j = e / TB;
            As[i][j] = (ri + i < wr && j < kc) ? Xr[k0 + j + static_cast<size_t>(ri + i) * ld] : 0.0;
            Bs[i][j] = (rj + i < wr && j < kc) ? Xr[k0 + j + static_cast<size_t>(rj + i) * ld] : 0.0;
        }
        __syncthreads();
        for (int j = 0; j < kc; ++j) {
            double av[4], bv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = As[4 * tr + q][j];
                bv[q] = Bs[4 * tc + q][j];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc.v[p][q] += av[p] * bv[q];
        }
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = ri + 4 * tr + p, j = rj + 4 * tc + q;
            if (i < wr && j < wr) M[i + static_cast<size_t>(j) * ld] = acc.v[p][q];
        }
}

}  // namespace

void launchDeviceFactor(const DevFactorArgs& d, hipStream_t s)
{
    FacArgs a;
    a.L = d.L;
    a.X = d.X;
    a.U = d.U;
    a.A = d.A;
    a.w = d.w;
    a.m = d.m;
    a.dval = d.dval;
    a.uOff = d.uOff;
    a.rel = d.rel;
    a.relOff = d.relOff;
    a.pairs = d.pairs;
    a.asmDst = d.asmDst;
    a.asmSrc = d.asmSrc;
    a.tasks = d.tasks;
    a.fail = d.fail;
    for (int q = 0; q < d.nlaunch; ++q) {
        const int kind = d.launches[3 * q], first = d.launches[3 * q + 1], count = d.launches[3 * q + 2];
        if (count <= 0) continue;
        switch (kind) {
            case 0: kFacAsm<<<(count + 255) / 256, 256, 0, s>>>(a, first, count); break;
            case 1: kFacExtAdd<<<count, 256, 0, s>>>(a, first); break;
            case 2: kFacPotrf<<<count, 256, 0, s>>>(a, first); break;
            case 3: kFacTrsm<<<count, 64, 0, s>>>(a, first); break;
            case 4: kFacUpdate<<<count, 256, 0, s>>>(a, first); break;
            case 5: kFacSyrk<<<count, 256, 0, s>>>(a, first); break;
            default: kFacBsolve<<<count, 256, 0, s>>>(a, first); break;
        }
    }
}

void launchFactorToF32(const double* X, void* B, long long n, hipStream_t s)
{
    if (n > 0) kFacToF32<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(X, static_cast<float*>(B), n);
}

void launchFactorGram(const double* Xr, int wr, int ld, double* M, hipStream_t s)
{
    if (wr <= 0) return;
    const int nt = (wr + TB - 1) / TB;
    kFacGram<<<dim3(nt, nt), 256, 0, s>>>(Xr, wr, ld, M);
}

}  // namespace dev
}  // namespace efem
