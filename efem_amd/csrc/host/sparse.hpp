// Sparse Cholesky of the approximate Jacobian for the device preconditioner.
//
// The reference factors the 3N_f x 3N_f interleaved matrix
// K = sum_e sum_q V w c (T^T T) (x) I3 with Eigen's SimplicialLLT + AMD
// (src/simulation/nlsolver/LBFGS.h:22,65-81, src/fe/ElementConstraint.cpp:
// 164-205).  Because K = Pi (L (x) I3) Pi^T, this module factors the scalar
// N_f x N_f Laplacian L once and applies it to the three interleaved
// components: identical K^-1 up to rounding, 1/9 of the factor storage and
// 1/27 of the factor flops.
//
// Pipeline (host): nested-dissection ordering -> elimination tree and
// postorder -> column counts -> relaxed supernodes -> supernodal multifrontal
// LL^T -> per-supernode block B_s = [L11^-1 ; L21 L11^-1] used by the device
// level-scheduled triangular solves (forward y_s = L11^-1 b'_s and
// t_s = L21 L11^-1 b'_s; backward z_s = L11^-T y_s - (L21 L11^-1)^T z_R).
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace efem {

struct Csc {            // full symmetric, sorted rows, no duplicates
    int n = 0;
    std::vector<int> p, i;
    std::vector<double> x;
};

struct SupernodalFactor {
    int n = 0;
    std::vector<int> perm;       // perm[new] = old
    int nsuper = 0;
    std::vector<int> first;      // [nsuper+1] first column of each supernode
    std::vector<int> rowPtr;     // [nsuper+1] offsets into rows
    std::vector<int> rows;       // row set of each supernode (its columns first)
    std::vector<long long> valPtr;  // [nsuper+1] offsets into blocks
    std::unique_ptr<double[]> B; // column-major m_s x ns blocks [L11^-1 ; L21 L11^-1]
    long long Bsize = 0;
    std::vector<int> parent;     // supernodal elimination tree (-1 root)
    std::vector<int> level;      // height from the leaves
    int nlevels = 0;
    long long nnzL = 0;          // scalar nonzeros of L incl. relaxed zeros
    double flops = 0;
    double seconds = 0;
    double numericSeconds = 0;
};

// Symbolic analysis of P A P^T: what the numeric factorisation needs that
// depends on A's pattern and the ordering only, not on A's values, so a
// refactorisation of a matrix with the same pattern reuses it.
struct SupernodalSymbolic {
    int n = 0;
    long long anz = 0;                   // entries of the A it was built from
    std::vector<int> perm;               // final (postordered) ordering, perm[new] = old
    std::vector<int> Lp, Li;             // lower pattern of P A P^T by columns
    std::vector<int> Lsrc;               // entry of A.x behind each lower entry
    int nsuper = 0;
    std::vector<int> first, rowPtr, rows, parent, level;  // as in SupernodalFactor
    std::vector<long long> valPtr;
    std::vector<std::vector<int>> children;
    std::vector<long long> subtreeCols;  // columns in each supernode's subtree
    int nlevels = 0;
    long long nnzL = 0;
    double flops = 0;
    double seconds = 0;
};

// scalar Laplacian from element gradients (see model.cpp)
// order: nested dissection on the node graph using coordinates xyz[n][3]
std::vector<int> nestedDissection(const Csc& A, const std::vector<double>& xyz);

// Host memory kept between factorisations of one problem: the block storage
// B and the update matrices come back to it when a factorisation is done,
// so a refactorisation writes into pages that are already mapped instead of
// faulting in fresh ones (hundreds of MB on the larger meshes).
class HostWorkspace {
public:
    HostWorkspace() = default;
    HostWorkspace(const HostWorkspace&) = delete;
    HostWorkspace& operator=(const HostWorkspace&) = delete;
    ~HostWorkspace();
    // block storage of at least `count` doubles (contents undefined)
    std::unique_ptr<double[]> takeBlocks(size_t count);
    void giveBlocks(std::unique_ptr<double[]> b, size_t count);
    // an update buffer of at least `count` doubles (thread safe)
    double* takeUpdate(size_t count);
    void giveUpdate(double* p);

private:
    std::unique_ptr<double[]> blocks_;
    size_t blockCount_ = 0;
    std::mutex mu_;
    std::multimap<size_t, double*> free_;    // capacity -> buffer
    std::map<double*, size_t> capacity_;     // every buffer it owns
};

// Every maximal elimination subtree of at most `denseSubtree` columns
// becomes one dense supernode (fewer levels for the device triangular
// solves, a few percent more stored entries); 1..kMaxSupernodeWidth.
constexpr int kDefaultDenseSubtree = 64;

// ordering-dependent analysis: elimination tree, postorder, column counts,
// (relaxed) supernodes, row sets, block offsets, levels
bool supernodalSymbolic(const Csc& A, const std::vector<int>& perm, SupernodalSymbolic& S, std::string& msg,
                        int denseSubtree = kDefaultDenseSubtree);
// the numeric multifrontal factorisation of A (whose pattern S was built
// from); false if a pivot is not positive
bool supernodalNumeric(const Csc& A, const SupernodalSymbolic& S, SupernodalFactor& F, std::string& msg,
                       HostWorkspace* ws = nullptr);
// both; returns false if a pivot is not positive (the reference's fallback case)
bool supernodalCholesky(const Csc& A, const std::vector<int>& perm, SupernodalFactor& F,
                        std::string& msg, int denseSubtree = kDefaultDenseSubtree,
                        HostWorkspace* ws = nullptr);

// Forest of independent factors (batched instances): part i's columns,
// supernodes, rows and blocks follow those of part i-1; levels stay per part.
// snOffset[i] = first supernode of part i (size parts + 1).
void mergeFactors(std::vector<SupernodalFactor>& parts, SupernodalFactor& out, std::vector<int>& snOffset);

// the device algorithm run on the host with 3 interleaved right-hand sides:
// forward (multifrontal) f_s = [b_s; 0] + sum_children u_c,
// y_s = L11^-1 f_s[:w], u_s = f_s[w:] - (L21 L11^-1) f_s[:w];
// backward z_s = L11^-T y_s - (L21 L11^-1)^T z_R
void blockInverseSolve3(const SupernodalFactor& F, const double* b, double* x);

// One forward work item's extend-add table (64 rows from r0 of supernode s,
// K entries per front row, own-column stride W) at tab[base, base + K (64 + W))
struct FwdTableItem {
    int s, r0, K, W;
    long long base;
};
// Fills tab (resized to size) for the items, by all host threads: entry
// [k][lane] is the k-th update position (fidx order) of front row r0 + lane
// when that row is below the supernode's columns, entry K 64 + [k][c] that of
// the supernode's own column c; every other entry is `zeroSlot`.
void fillForwardTables(const SupernodalFactor& F, const std::vector<int>& fptr, const std::vector<int>& fidx,
                       const std::vector<FwdTableItem>& items, long long size, int zeroSlot,
                       std::vector<int>& tab);

// Copy of B in another layout, by all host threads: supernode s's m x w
// block to dst + off[s] with leading dimension ld = m rounded up to a
// multiple of `align` (padding rows zero), as float (f32) or double.
void packBlocks(const SupernodalFactor& F, const std::vector<long long>& off, int align, bool f32, void* dst);

// multifrontal extend-add map of the forward sweep: front row
// (rowPtr[s] + r) of supernode s receives the update rows uOff[c] + i of
// its children c; CSR over all front rows
void frontGatherMap(const SupernodalFactor& F, const std::vector<int>& uOff, std::vector<int>& fptr,
                    std::vector<int>& fidx);

}  // namespace efem
