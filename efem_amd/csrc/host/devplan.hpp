// Plan of the device numeric factorisation: everything the device
// multifrontal LL^T needs that depends on the approximate Jacobian's pattern
// only (so a context builds it once, beside the reused symbolic analysis).
//
// The device factors the same supernodal structure as the host
// (SupernodalSymbolic): per supernode s (w columns, m rows, nr = m - w) the
// panel [L11; L21] lives in the L arena at dval[s] (column-major, leading
// dimension ldOf(m), the device B layout), its update matrix U_s (nr x nr,
// lower, ld nr) in the U arena at uOff[s], and the block inverse
// B_s = [L11^-1; L21 L11^-1] is written to the X arena at the same dval[s]
// (X is the FP64 B the triangular solves read).  Levels run bottom-up; each
// is a fixed sequence of batched launches over 64 x 64 tiles:
//   assemble A's lower entries -> extend-add children's update matrices ->
//   per 64-column step k: Cholesky of the diagonal tile, triangular solve
//   of the rows below it, update of the panel's remaining columns -> U_s -= L21 L21^T -> per 64-column step J (last to first):
//   L11^-1 by a right-side triangular solve of L11 X = I -> L21 L11^-1 as
//   one tile product.
#pragma once
#include <cstdint>
#include <vector>

#include "sparse.hpp"

namespace efem {

// one launch of the device factorisation: kind, first task, task count
struct DevLaunch {
    int kind;
    int first;
    int count;
};

enum DevLaunchKind {
    DF_ASM = 0,     // tasks = A entries (asmDst, asmSrc)
    DF_EXTADD = 1,  // tasks {s, target front column, pair begin, pair end}
    DF_POTRF = 2,   // tasks {s, k, 0, 0}: Cholesky of the diagonal tile k
    DF_TRSM = 3,    // tasks {s, k, first row, 0}: 256 rows below the tile
                    // solved against its factor
    DF_UPDATE = 4,  // tasks {s, k, row tile, column tile}
    DF_SYRK = 5,    // tasks {s, row tile, column tile, 0} of U_s
    DF_BSOLVE = 6,  // tasks {s, J, row tile < w / 64, 0}: L11^-1 by column tiles
    DF_GEMM21 = 7,  // tasks {s, row tile of L21, column tile, 0}: L21 L11^-1
};

struct DevFactorPlan {
    int n = 0, nsuper = 0;
    long long panelSize = 0;  // doubles of the L and X arenas (the device B layout)
    long long updSize = 0;    // doubles of the U arena
    std::vector<int> w, m;               // per supernode
    std::vector<long long> dval, uOff;   // per supernode (uOff -1: no update matrix)
    std::vector<long long> asmDst;       // L arena position of each A lower entry
    std::vector<int> asmSrc;             // its index in A's values
    std::vector<int> pairs;              // extend-add pairs {child, child column} (2 ints)
    std::vector<int> rel;                // child update rows -> parent front rows
    std::vector<int> relOff;             // per supernode offset into rel (-1: none)
    std::vector<int> tasks;              // 4 ints per task (asm tasks are not listed)
    std::vector<DevLaunch> launches;     // in execution order
};

// builds the plan from the symbolic analysis (host, once per pattern)
void buildDevFactorPlan(const SupernodalSymbolic& S, DevFactorPlan& P);

}  // namespace efem

namespace efem {

// The fused top of the elimination tree: T = the supernodes within depth D
// of the roots, D the deepest depth whose supernodes (with all shallower
// ones) hold at most maxCols columns; T is closed under ancestors.  With
// L_TT the factor's TT block, Z = (L_TT L_TT^T)^-1 is the TT block of
// (P A P^T)^-1, so T's forward and backward solves are the one product
// z_T = Z f_T, f_T the right-hand side on T plus the update rows of T's
// children outside T.  T-local numbering: by depth, then factor column (a
// supernode's columns stay contiguous).
struct TopSuperInfo {
    int s;      // supernode
    int cs;     // first T-local column
    int w, m;   // width, rows
    int rOff;   // rows below its columns, T-local, at rtl[rOff .. rOff + m - w)
};
struct TopPlan {
    int depth = -1, nT = 0, ld = 0, K = 0;
    std::vector<char> inT;              // per supernode
    std::vector<int> tcol;              // [nT] factor column of each T-local index
    std::vector<int> tab;               // [K][nT] update-vector positions of f_T (zeroSlot if absent)
    std::vector<TopSuperInfo> super;    // T supernodes in T-local order
    std::vector<int> rtl;
    // Z is formed depth by depth from the roots (selected inversion, with
    // R_s the rows of s below its columns and B21_s = L21 L11^-1):
    //   Z[r, s] = -Z[r, R_s] B21_s for the rows r of shallower depths (kind 0),
    //   Z[s, r] = Z[r, s] (kind 1, mirror of the rows < p0 into [p0, p1)),
    //   the same product for the rows of this depth (kind 0), then the
    //   diagonal block Z_ss = B11^T B11 - B21^T Z[R_s, s] (kind 2).
    // Tile tasks, 4 ints each: kind 0 {T supernode, row tile start, column
    // tile start, row end}; kind 2 {T supernode, i tile start, j tile
    // start, 0}; tiles of 64.
    std::vector<int> tasks;
    struct Step {
        int kind, first, count, p0, p1;
    };
    std::vector<Step> steps;
};
// false (and no fused top) if not even the roots fit, or a T row has more
// than maxK children's update rows; zeroSlot = the index of the all-zero
// update row; Z's leading dimension = nT rounded up to `align` rows
bool planTop(const SupernodalFactor& F, int zeroSlot, int maxCols, int maxK, int align, TopPlan& P);

}  // namespace efem
