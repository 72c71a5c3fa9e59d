// deform.h — the host Laplacian deformation of the FSI coupling
// (LaplacianDeformation, src/deformation/LaplacianDeformation.cpp, and its
// driver deformStep, src/main.cpp:862-944), with the sparse LDL^T it needs
// (the reference uses Eigen's SimplicialLDLT / CHOLMOD and its own
// IncrementalCholesky, src/deformation/IncrementalCholesky.cpp).
//
// Host code (g++), FP64 throughout.  The C-ABI (fsi_deform_*) is in
// deform.cpp.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace fsi {
namespace deform {

// Compressed sparse column matrix; row indices ascending within a column.
struct Csc {
    int rows = 0, cols = 0;
    std::vector<int> p, i;
    std::vector<double> x;
};

struct Triplet {
    int r, c;
    double v;
};

// Sums duplicate entries.
Csc from_triplets(int rows, int cols, std::vector<Triplet> t);
Csc transpose(const Csc& a);
Csc multiply(const Csc& a, const Csc& b);               // a * b
void gemv(const Csc& a, const double* x, double* y);    // y = a * x
void gemv_t(const Csc& a, const double* x, double* y);  // y = a^T * x

// Reverse Cuthill-McKee ordering of the pattern of a symmetric matrix
// (Cuthill & McKee 1969; reversed, George 1971), started in each connected
// component from a pseudo-peripheral vertex of minimum degree.  perm[new] =
// old.
std::vector<int> rcm_order(const Csc& a);

// Sparse LDL^T of a symmetric matrix, up-looking, one row of L per step
// (T. A. Davis, "Algorithm 849: a concise sparse Cholesky factorization
// package", ACM TOMS 31(4), 2005): the elimination tree and the column counts
// from the pattern, then row k of L by a sparse triangular solve over the
// tree reach of column k.  A fill-reducing order is applied first.
//
// Constraints (Herholz et al., "Sparse Cholesky updates for interactive mesh
// parameterization", ACM TOG 2020, Eq. 14): a constrained index k has row
// and column k of the working matrix replaced by the identity, which only
// zeroes entries of A, so the pattern of L never changes.  After a change of
// the constrained set only the steps on tree paths from the changed columns
// to the root are recomputed (update).
class Ldlt {
public:
    // Pattern analysis of the symmetric matrix a (both triangles stored).
    // order: apply reverse Cuthill-McKee (else the natural order).
    void analyze(const Csc& a, bool order);
    // Numeric factorization of a under the current constraint mask.  Returns
    // false on a zero pivot.
    bool factorize(const Csc& a);
    // Recompute only the steps reached from the columns `changed` (indices of
    // a, original order) after set_constrained.  Returns false on a zero pivot.
    bool update(const Csc& a, const std::vector<int>& changed);
    // constrained[k] != 0 detaches index k (original order).
    void set_constrained(const std::vector<char>& constrained);
    // Solves in place (original order).
    void solve(double* b) const;
    int size() const { return n_; }
    bool ok() const { return ok_; }
    // Steps recomputed by the last factorize / update.
    int last_steps() const { return last_steps_; }
    // Entries of L below the diagonal.
    size_t nnz() const { return li_.size(); }

private:
    int n_ = 0;
    bool ok_ = false;
    int last_steps_ = 0;
    std::vector<int> perm_, inv_;  // perm_[new] = old
    Csc ap_;                       // permuted pattern (upper triangle, column k holds rows <= k)
    std::vector<int> parent_, lp_, li_;
    std::vector<double> lx_, d_;
    std::vector<char> con_;        // permuted order
    mutable std::vector<double> y_;
    mutable std::vector<int> pattern_, flag_;
    std::vector<int> ap_src_;      // for each entry of ap_, its position in a.x
    std::vector<int> next_;        // next slot of each column during a full pass
    template <bool SEQ>
    bool step(const Csc& a, int k);  // row k of L and D[k]
};

// Cotangent Laplacian mesh deformation (LaplacianDeformation).
class Laplacian {
public:
    Laplacian(std::vector<double> verts, const std::vector<int32_t>& faces);

    // initialize(): cotangent weights, L, delta = L * V (rest pose = the
    // vertices given at construction).
    void initialize();
    void add_anchor(int v);
    void set_control_points(const std::vector<int>& idx, const std::vector<double>& targets);
    void clear_control_points();
    void set_use_local(bool on) { use_local_ = on; }
    void set_local_kring(int k) { kring_ = k; }
    void set_use_incremental(bool on) { use_inc_ = on; }
    bool precompute();
    bool solve();

    int n() const { return (int)(v_.size() / 3); }
    const std::vector<double>& vertices() const { return v_; }
    const std::vector<double>& rest() const { return rest_; }
    const std::vector<int>& anchors() const { return anchors_; }
    bool is_anchor(int v) const;
    int local_patch_size() const { return local_.ready ? (int)local_.interior.size() : 0; }
    int last_factor_steps() const { return last_steps_; }
    double edge_weight(int a, int b) const;
    const Csc& laplacian() const { return L_; }
    // Max |x - x_fresh| of the incremental factor against a fresh full
    // factorization of the same A^B (x axis); -1 when not in that mode.
    double verify_incremental();

private:
    std::vector<double> v_, rest_;
    std::vector<int32_t> faces_;
    std::vector<std::vector<int>> nbr_;       // one ring, ascending
    std::vector<std::vector<double>> wnbr_;   // cotangent weight of each ring edge
    Csc L_;
    std::vector<double> delta_[3];
    std::vector<int> anchors_;
    std::vector<double> anchor_pos_;
    std::vector<int> ctrl_;
    std::vector<double> ctrl_t_;
    bool use_local_ = true, use_inc_ = false;
    int kring_ = 3;
    int last_steps_ = 0;

    struct Local {
        std::vector<int> ring;               // the k-ring the structures below belong to
        std::vector<int> interior, boundary, g2l;
        Csc Lii, Lib, LiiT, LtL;             // LtL = Lii^T Lii
        std::vector<int> diag;               // position of each diagonal entry in LtL.x
        std::vector<int> handle_local, handle_global;
        std::vector<int> handle_ctrl;        // index in ctrl_ of each handle's first occurrence
        Csc AtA;                             // LtL + I_handles
        Ldlt ldlt;
        bool analyzed = false, ready = false;
    } local_;

    // global soft path
    Csc At_;
    Ldlt glob_;
    bool glob_ready_ = false;

    // incremental path
    Csc A_;                 // L^T L + 1e-8 I
    std::vector<double> b_[3];
    Ldlt inc_;
    bool inc_ready_ = false;
    std::vector<char> mask_;
    std::vector<int> prev_;

    void compute_weights();
    std::vector<int> kring(const std::vector<int>& seeds, int k) const;
    bool precompute_local();
    bool solve_local();
    std::vector<int> constrained_sorted() const;
};

}  // namespace deform
}  // namespace fsi
