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

constexpr int TB = 64;           // tile edge
constexpr int kTrsmRows = 256;   // rows of one triangular-solve task (the plan's)

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

// one lane's copy of a value held by another lane of the wave
__device__ __forceinline__ double readLane(double v, int lane)
{
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u), lane));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), lane));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// PIN(v) marks a register value as updated where it stands: placed after
// each column of a right-looking elimination it keeps the compiler from
// sinking the column's updates to their first use (it would then hold every
// loaded factor entry live until there, and spill); emits no instruction
#define PIN(v) asm volatile("" : "+v"(v))

// Cholesky of the diagonal tile k of the panel (kw x kw): one wave, lane =
// row, right-looking in registers; each finished column is broadcast
// through LDS for the trailing update
__global__ __launch_bounds__(TB) void kFacPotrf(FacArgs a, int first)
{
    __shared__ double Lc[TB][TB];
    const int4 t4 = a.tasks[first + blockIdx.x];
    const int s = t4.x, k = t4.y;
    const int ws = a.w[s], ld = blockLd(a.m[s]);
    const int c0 = TB * k, kw = min(TB, ws - c0);
    double* P = a.L + a.dval[s] + c0 + static_cast<size_t>(c0) * ld;
    const int r = threadIdx.x;
    // the tile's lower part (zeros elsewhere) staged through LDS: coalesced
    // column loads, then each lane takes its row
#pragma unroll 16
    for (int j = 0; j < TB; ++j) {
        const double e = P[min(r, kw - 1) + static_cast<size_t>(min(j, kw - 1)) * ld];
        Lc[r][j] = (r < kw && j <= r) ? e : 0.0;
    }
    __syncthreads();
    double v[TB];
#pragma unroll
    for (int j = 0; j < TB; ++j) v[j] = Lc[r][j];
    __syncthreads();
    bool bad = false;
#pragma unroll
    for (int c = 0; c < TB; ++c) {
        const bool pad = c >= kw;
        const double d = readLane(v[c], c);  // the updated diagonal entry
        const bool notPos = !pad && !(d > 0.0);
        bad = bad || notPos;
        // sqrt(d) and its reciprocal from the hardware 1/sqrt estimate by
        // Newton steps (a shorter dependent chain than sqrt and a division:
        // the column loop is latency bound)
        const double dd = pad || notPos ? 1.0 : d;
        double y = __builtin_amdgcn_rsq(dd);
        y = fma(0.5 * y, fma(-dd * y, y, 1.0), y);
        double l = dd * y;
        l = fma(0.5 * y, fma(-l, l, dd), l);
        y = fma(y, fma(-l, y, 1.0), y);
        const double inv = pad ? 0.0 : y;
        // lanes below the diagonal get L(r, c); lanes above only touch their
        // own upper part, which nothing reads
        const double lc = r == c ? (pad ? 0.0 : l) : v[c] * inv;
        v[c] = lc;
        Lc[c][r] = lc;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the trailing update, scheduled in chunks of 16 (bounds the LDS
        // reads in flight)
#pragma unroll
        for (int q = c + 1; q < TB; ++q) {
            v[q] -= lc * Lc[c][q];
            if ((q & 15) == 15) __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int q = c + 1; q < TB; ++q) PIN(v[q]);
    }
    if (r == 0 && bad) atomicOr(a.fail, 1);
    if (r < kw)
#pragma unroll
        for (int j = 0; j < TB; ++j)
            if (j <= r) P[r + static_cast<size_t>(j) * ld] = v[j];
}

// rows [r0, r0 + kTrsmRows) of the panel's column tile k (all below the
// diagonal tile) <- (rows) L_kk^-T: thread = row, right-looking forward
// substitution against L_kk staged in LDS column by column
__global__ __launch_bounds__(kTrsmRows) void kFacTrsm(FacArgs a, int first)
{
    __shared__ double Lc[TB][TB];  // Lc[c][j] = L_kk(j, c)
    __shared__ double invd[TB];    // 1 / L_kk(c, c) (0 for padding columns)
    const int4 t4 = a.tasks[first + blockIdx.x];
    const int s = t4.x, k = t4.y, r0 = t4.z;
    const int ws = a.w[s], ms = a.m[s], ld = blockLd(ms);
    const int c0 = TB * k, kw = min(TB, ws - c0);
    double* P = a.L + a.dval[s];
    const int tid = threadIdx.x;
    for (int e = tid; e < TB * TB; e += kTrsmRows) {
        const int j = e % TB, c = e / TB;
        Lc[c][j] = (j < kw && c < kw && j > c) ? P[c0 + j + static_cast<size_t>(c0 + c) * ld] : 0.0;
        if (j == c) invd[c] = c < kw ? 1.0 / P[c0 + c + static_cast<size_t>(c0 + c) * ld] : 0.0;
    }
    __syncthreads();
    const int gr = r0 + tid;
    if (gr >= ms) return;
    double b[TB];
#pragma unroll
    for (int j = 0; j < TB; ++j) {
        const double e = P[gr + static_cast<size_t>(c0 + min(j, kw - 1)) * ld];
        b[j] = j < kw ? e : 0.0;
    }
#pragma unroll
    for (int c = 0; c < TB; ++c) {
        const double x = b[c] * invd[c];
        b[c] = x;
#pragma unroll
        for (int j = c + 1; j < TB; ++j) b[j] -= x * Lc[c][j];
#pragma unroll
        for (int j = c + 1; j < TB; ++j) PIN(b[j]);
    }
    if (kw == TB) {
#pragma unroll
        for (int j = 0; j < TB; ++j) P[gr + static_cast<size_t>(c0 + j) * ld] = b[j];
    } else {
#pragma unroll
        for (int j = 0; j < TB; ++j)
            if (j < kw) P[gr + static_cast<size_t>(c0 + j) * ld] = b[j];
    }
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

// L11^-1, column tile J, rows of tile it >= J (inside the w columns):
// y = I(rows, J) - sum_{J < K <= it} X(rows, K) L11(K, J), then x L_JJ = y by
// right-looking backward substitution, lane = row
__global__ __launch_bounds__(256) void kFacBsolve(FacArgs a, int first)
{
    __shared__ double As[TB][TB + 1], Bs[TB][TB + 1];
    const int4 t4 = a.tasks[first + blockIdx.x];
    const int s = t4.x, J = t4.y, it = t4.z;
    const int ws = a.w[s], ld = blockLd(a.m[s]);
    const int c0 = TB * J, kw = min(TB, ws - c0);
    const int ri = TB * it, rows = min(TB, ws - ri);
    const double* P = a.L + a.dval[s];
    double* X = a.X + a.dval[s];
    TileAcc acc{};
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
    // X(rows, k) vanishes for k past the tile's last row (L11^-1 is lower)
    for (int k0 = c0 + kw; k0 < ri + rows; k0 += TB) {
        const int kc = min(TB, ri + rows - k0);
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
    // y into As, L_JJ row-major into Bs with its reciprocal diagonal on the
    // diagonal (unit on padding columns)
    for (int e = tid; e < TB * TB; e += 256) {
        const int i = e % TB, j = e / TB;
        const double l = (i < kw && j < kw && j >= i) ? P[c0 + j + static_cast<size_t>(c0 + i) * ld]
                                                      : (i == j ? 1.0 : 0.0);
        Bs[j][i] = i == j ? 1.0 / l : l;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = 4 * tr + p, j = 4 * tc + q;
            const double d = (i < rows && j < kw && ri + i == c0 + j) ? 1.0 : 0.0;
            As[i][j] = (i < rows && j < kw) ? d - acc.v[p][q] : 0.0;
        }
    __syncthreads();
    if (tid >= TB) return;
    const int i = tid;
    double y[TB];
#pragma unroll
    for (int j = 0; j < TB; ++j) y[j] = As[i][j];
#pragma unroll
    for (int jj = 0; jj < TB; ++jj) {
        const int j = TB - 1 - jj;
        const double x = y[j] * Bs[j][j];
        y[j] = x;
#pragma unroll
        for (int c = 0; c < j; ++c) y[c] -= x * Bs[j][c];
#pragma unroll
        for (int c = 0; c < j; ++c) PIN(y[c]);
    }
    if (i >= rows) return;
#pragma unroll
    for (int j = 0; j < TB; ++j)
        if (j < kw) X[ri + i + static_cast<size_t>(c0 + j) * ld] = y[j];
}

// L21 L11^-1, tile (it, jt): rows w + 64 it.., columns 64 jt.. of X, the
// sum over k >= 64 jt (L11^-1 is lower)
__global__ __launch_bounds__(256) void kFacGemm21(FacArgs a, int first)
{
    __shared__ double As[TB][TB + 1], Bs[TB][TB + 1];
    const int4 t4 = a.tasks[first + blockIdx.x];
    const int s = t4.x, it = t4.y, jt = t4.z;
    const int ws = a.w[s], ms = a.m[s], nr = ms - ws, ld = blockLd(ms);
    const int ri = TB * it, rows = min(TB, nr - ri), cj = TB * jt, cols = min(TB, ws - cj);
    const double* P = a.L + a.dval[s] + ws + ri;  // L21 rows of the tile
    double* X = a.X + a.dval[s];
    TileAcc acc{};
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
    for (int k0 = cj; k0 < ws; k0 += TB) {
        const int kc = min(TB, ws - k0);
        __syncthreads();
        for (int e = tid; e < TB * TB; e += 256) {
            const int i = e % TB, j = e / TB;
            // As[i][j] = L21(ri + i, k0 + j); Bs[j][i] = X11(k0 + i, cj + j)
            As[i][j] = (i < rows && j < kc) ? P[i + static_cast<size_t>(k0 + j) * ld] : 0.0;
            Bs[j][i] = (j < cols && i < kc) ? X[k0 + i + static_cast<size_t>(cj + j) * ld] : 0.0;
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
            const int i = 4 * tr + p, j = 4 * tc + q;
            if (i < rows && j < cols) X[ws + ri + i + static_cast<size_t>(cj + j) * ld] = acc.v[p][q];
        }
}

__global__ __launch_bounds__(256) void kFacToF32(const double* __restrict__ X, float* __restrict__ B, long long n)
{
    const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) B[i] = static_cast<float>(X[i]);
}

// ---- the fused top: Z = (L_TT L_TT^T)^-1 by selected inversion -----------
// Tile (r0, j0) of the column block of T supernode s:
// Z[r, cs + j] = -sum_k Z[r, rtl[rOff + k]] B21_s[k, j] for r0 <= r < row end,
// k over the rows of s below its columns (their columns of Z are complete)
__global__ __launch_bounds__(256) void kTopGemm(TopBuild b, int first)
{
    __shared__ double As[TB][TB + 1], Bs[TB][TB + 1];
    const int4 t4 = b.tasks[first + blockIdx.x];
    const TopSuper sn = b.sn[t4.x];
    double* const Zs = b.Z + sn.zoff;
    const int r0 = t4.y, j0 = t4.z;
    const int nr = sn.m - sn.w;
    const int arows = min(TB, t4.w - r0), bcols = min(TB, sn.w - j0);
    const int* rt = b.rtl + sn.rOff;
    const double* B21 = b.B + sn.boff + sn.w;
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
    TileAcc acc{};
    for (int k0 = 0; k0 < nr; k0 += TB) {
        const int kc = min(TB, nr - k0);
        __syncthreads();
        for (int e = tid; e < TB * TB; e += 256) {
            const int i = e % TB, j = e / TB;
            As[i][j] = (i < arows && j < kc) ? Zs[(r0 + i) + static_cast<size_t>(rt[k0 + j]) * sn.ldz] : 0.0;
            Bs[j][i] = (i < kc && j < bcols) ? B21[(k0 + i) + static_cast<size_t>(j0 + j) * sn.ld] : 0.0;
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
            const int i = 4 * tr + p, j = 4 * tc + q;
            if (i < arows && j < bcols)
                Zs[(r0 + i) + static_cast<size_t>(sn.cs + j0 + j) * sn.ldz] = -acc.v[p][q];
        }
}

// Z[j, i] = Z[i, j] for i < p0 <= j < p1 (64 x 64 tiles through LDS)
// one 64 x 64 tile of a mirror step per block, the trees' steps together
__global__ __launch_bounds__(256) void kTopMirrorTasks(double* Z, const TopMirrorTask* __restrict__ tasks)
{
    __shared__ double t[TB][TB + 1];
    const TopMirrorTask m = tasks[blockIdx.x];
    double* const Zt = Z + m.zoff;
    for (int e = threadIdx.x; e < TB * TB; e += 256) {
        const int i = e % TB, j = e / TB;
        if (m.i0 + i < m.p0 && m.j0 + j < m.p1) t[i][j] = Zt[(m.i0 + i) + static_cast<size_t>(m.j0 + j) * m.ldz];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < TB * TB; e += 256) {
        const int j = e % TB, i = e / TB;
        if (m.i0 + i < m.p0 && m.j0 + j < m.p1) Zt[(m.j0 + j) + static_cast<size_t>(m.i0 + i) * m.ldz] = t[i][j];
    }
}

__global__ __launch_bounds__(256) void kTopMirror(TopBuild b, int p0, int p1)
{
    __shared__ double t[TB][TB + 1];
    const int i0 = blockIdx.x * TB, j0 = p0 + blockIdx.y * TB;
    for (int e = threadIdx.x; e < TB * TB; e += 256) {
        const int i = e % TB, j = e / TB;
        if (i0 + i < p0 && j0 + j < p1) t[i][j] = b.Z[(i0 + i) + static_cast<size_t>(j0 + j) * b.ldz];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < TB * TB; e += 256) {
        const int j = e % TB, i = e / TB;
        if (i0 + i < p0 && j0 + j < p1) b.Z[(j0 + j) + static_cast<size_t>(i0 + i) * b.ldz] = t[i][j];
    }
}

// Tile (i0, j0) of the diagonal block of s: Z[cs + i, cs + j] =
// sum_{r < m} B_s[r, i] V[r, j], V[r, j] = B_s[r, j] (r < w: B11^T B11, the
// zero upper triangle included) or -Z[rtl[rOff + r - w], cs + j] (r >= w:
// -B21^T Z[R_s, s], the column block's rows on R_s formed before)
__global__ __launch_bounds__(256) void kTopDiag(TopBuild b, int first)
{
    __shared__ double As[TB][TB + 1], Vs[TB][TB + 1];
    const int4 t4 = b.tasks[first + blockIdx.x];
    const TopSuper sn = b.sn[t4.x];
    double* const Zs = b.Z + sn.zoff;
    const int i0 = t4.y, j0 = t4.z;
    const int ni = min(TB, sn.w - i0), nj = min(TB, sn.w - j0);
    const int* rt = b.rtl + sn.rOff;
    const double* Bp = b.B + sn.boff;
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
    TileAcc acc{};
    for (int k0 = 0; k0 < sn.m; k0 += TB) {
        const int kc = min(TB, sn.m - k0);
        __syncthreads();
        for (int e = tid; e < TB * TB; e += 256) {
            const int kk = e % TB, c = e / TB, r = k0 + kk;
            As[c][kk] = (kk < kc && c < ni) ? Bp[r + static_cast<size_t>(i0 + c) * sn.ld] : 0.0;
            double v = 0.0;
            if (kk < kc && c < nj)
                v = r < sn.w ? Bp[r + static_cast<size_t>(j0 + c) * sn.ld]
                             : -Zs[rt[r - sn.w] + static_cast<size_t>(sn.cs + j0 + c) * sn.ldz];
            Vs[c][kk] = v;
        }
        __syncthreads();
        for (int k = 0; k < kc; ++k) {
            double av[4], bv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = As[4 * tr + q][k];
                bv[q] = Vs[4 * tc + q][k];
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
            const int i = 4 * tr + p, j = 4 * tc + q;
            if (i < ni && j < nj)
                Zs[(sn.cs + i0 + i) + static_cast<size_t>(sn.cs + j0 + j) * sn.ldz] = acc.v[p][q];
        }
}

}  // namespace

void launchTopGemm(const TopBuild& b, int first, int count, hipStream_t s)
{
    if (count > 0) kTopGemm<<<count, 256, 0, s>>>(b, first);
}

void launchTopMirror(const TopBuild& b, int p0, int p1, hipStream_t s)
{
    if (p0 > 0 && p1 > p0) kTopMirror<<<dim3((p0 + TB - 1) / TB, (p1 - p0 + TB - 1) / TB), 256, 0, s>>>(b, p0, p1);
}

void launchTopMirrorTasks(double* Z, const TopMirrorTask* tasks, int count, hipStream_t s)
{
    if (count > 0) kTopMirrorTasks<<<count, 256, 0, s>>>(Z, tasks);
}

void launchTopDiag(const TopBuild& b, int first, int count, hipStream_t s)
{
    if (count > 0) kTopDiag<<<count, 256, 0, s>>>(b, first);
}

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
            case 2: kFacPotrf<<<count, TB, 0, s>>>(a, first); break;
            case 3: kFacTrsm<<<count, kTrsmRows, 0, s>>>(a, first); break;
            case 4: kFacUpdate<<<count, 256, 0, s>>>(a, first); break;
            case 5: kFacSyrk<<<count, 256, 0, s>>>(a, first); break;
            case 6: kFacBsolve<<<count, 256, 0, s>>>(a, first); break;
            case 7: kFacGemm21<<<count, 256, 0, s>>>(a, first); break;
            default: break;
        }
    }
}

void launchFactorToF32(const double* X, void* B, long long n, hipStream_t s)
{
    if (n > 0) kFacToF32<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(X, static_cast<float*>(B), n);
}

}  // namespace dev
}  // namespace efem
