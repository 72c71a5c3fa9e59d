// deform.cpp — sparse LDL^T, the cotangent Laplacian deformation and its
// per-frame driver on the host, behind fsi_deform_* (include/fsi_pbf.h).
//
// Follows LaplacianDeformation (src/deformation/LaplacianDeformation.cpp):
//   - cotangent weights per edge, 0.5 * sum of cot of the opposite angles of
//     its faces (cos / |sin|, faces with |sin| <= 1e-6 skipped), at least 1e-4
//     (:77-124); L_ii = 1, L_ij = -w_ij / sum_j w_ij (:126-168); delta = L V at
//     rest (:170-195);
//   - the local patch solve (default): the k-ring of the handles (anchors
//     excluded) is solved with the boundary ring held at its current
//     positions, normal equations Lii^T Lii + I_handles (:197-391);
//   - the global soft solve: [L; constraint rows] in the least-squares sense
//     (:487-548, 571-623);
//   - the hard-constraint path (Eq. 14 of Herholz and Sorkine-Hornung,
//     "Sparse Cholesky updates for interactive mesh parameterization", 2020):
//     A = L^T L + 1e-8 I with constrained rows / columns detached, refactored
//     only on elimination-tree paths when the constrained set changes
//     (:25-75, 438-485, 553-569, 625-669).
// and deformStep (src/main.cpp:862-944).  The reference's mesh is an OpenMesh
// PolyMesh; here a triangle list with one-ring tables built once.  Orders
// that the reference leaves to OpenMesh or std::unordered_set (ring order,
// patch numbering) are ascending vertex order here; they change only the
// rounding of the FP64 solves.
#include "deform.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <utility>

#include "../../include/fsi_pbf.h"

namespace fsi {
int fail(int code, const std::string& msg);  // pbf_core.hip: records fsi_last_error()

namespace deform {

// ---------------------------------------------------------------- sparse

Csc from_triplets(int rows, int cols, std::vector<Triplet> t) {
    std::sort(t.begin(), t.end(), [](const Triplet& a, const Triplet& b) {
        return a.c != b.c ? a.c < b.c : a.r < b.r;
    });
    Csc m;
    m.rows = rows;
    m.cols = cols;
    m.p.assign((size_t)cols + 1, 0);
    for (size_t k = 0; k < t.size();) {
        size_t e = k;
        double v = 0.0;
        while (e < t.size() && t[e].r == t[k].r && t[e].c == t[k].c) v += t[e++].v;
        m.i.push_back(t[k].r);
        m.x.push_back(v);
        m.p[(size_t)t[k].c + 1]++;
        k = e;
    }
    for (int c = 0; c < cols; ++c) m.p[(size_t)c + 1] += m.p[(size_t)c];
    return m;
}

Csc transpose(const Csc& a) {
    std::vector<Triplet> t;
    t.reserve(a.x.size());
    for (int c = 0; c < a.cols; ++c)
        for (int q = a.p[(size_t)c]; q < a.p[(size_t)c + 1]; ++q) t.push_back({c, a.i[(size_t)q], a.x[(size_t)q]});
    return from_triplets(a.cols, a.rows, std::move(t));
}

Csc multiply(const Csc& a, const Csc& b) {
    // column j of a*b = sum_k a(:, k) b(k, j), accumulated densely per column
    Csc m;
    m.rows = a.rows;
    m.cols = b.cols;
    m.p.assign((size_t)b.cols + 1, 0);
    std::vector<double> acc((size_t)a.rows, 0.0);
    std::vector<int> mark((size_t)a.rows, -1), rows;
    for (int j = 0; j < b.cols; ++j) {
        rows.clear();
        for (int q = b.p[(size_t)j]; q < b.p[(size_t)j + 1]; ++q) {
            const int k = b.i[(size_t)q];
            const double bkj = b.x[(size_t)q];
            for (int s = a.p[(size_t)k]; s < a.p[(size_t)k + 1]; ++s) {
                const int r = a.i[(size_t)s];
                if (mark[(size_t)r] != j) {
                    mark[(size_t)r] = j;
                    acc[(size_t)r] = 0.0;
                    rows.push_back(r);
                }
                acc[(size_t)r] += a.x[(size_t)s] * bkj;
            }
        }
        std::sort(rows.begin(), rows.end());
        for (int r : rows) {
            m.i.push_back(r);
            m.x.push_back(acc[(size_t)r]);
        }
        m.p[(size_t)j + 1] = (int)m.i.size();
    }
    return m;
}

void gemv(const Csc& a, const double* x, double* y) {
    std::fill(y, y + a.rows, 0.0);
    for (int c = 0; c < a.cols; ++c)
        for (int q = a.p[(size_t)c]; q < a.p[(size_t)c + 1]; ++q) y[a.i[(size_t)q]] += a.x[(size_t)q] * x[c];
}

void gemv_t(const Csc& a, const double* x, double* y) {
    for (int c = 0; c < a.cols; ++c) {
        double s = 0.0;
        for (int q = a.p[(size_t)c]; q < a.p[(size_t)c + 1]; ++q) s += a.x[(size_t)q] * x[a.i[(size_t)q]];
        y[c] = s;
    }
}

std::vector<int> rcm_order(const Csc& a) {
    const int n = a.cols;
    std::vector<int> deg((size_t)n, 0);
    for (int c = 0; c < n; ++c)
        for (int q = a.p[(size_t)c]; q < a.p[(size_t)c + 1]; ++q)
            if (a.i[(size_t)q] != c) deg[(size_t)c]++;
    auto less = [&](int x, int y) { return deg[(size_t)x] != deg[(size_t)y] ? deg[(size_t)x] < deg[(size_t)y] : x < y; };
    std::vector<char> done((size_t)n, 0);
    std::vector<int> level((size_t)n, -1), order, q, nb;
    order.reserve((size_t)n);
    // breadth-first search from s over vertices not yet ordered, neighbours
    // taken by increasing degree; q holds the visit order, level[] the depth.
    // Returns the depth of the last vertex (the eccentricity of s).
    auto bfs = [&](int s) {
        q.clear();
        q.push_back(s);
        level[(size_t)s] = 0;
        for (size_t h = 0; h < q.size(); ++h) {
            const int v = q[h];
            nb.clear();
            for (int k = a.p[(size_t)v]; k < a.p[(size_t)v + 1]; ++k) {
                const int u = a.i[(size_t)k];
                if (u != v && !done[(size_t)u] && level[(size_t)u] < 0) {
                    level[(size_t)u] = level[(size_t)v] + 1;
                    nb.push_back(u);
                }
            }
            std::sort(nb.begin(), nb.end(), less);
            q.insert(q.end(), nb.begin(), nb.end());
        }
        return level[(size_t)q.back()];
    };
    auto reset = [&]() {
        for (int v : q) level[(size_t)v] = -1;
    };
    for (int seed = 0; seed < n; ++seed) {
        if (done[(size_t)seed]) continue;
        // pseudo-peripheral start (George & Liu 1979): from a minimum-degree
        // vertex, move to a minimum-degree vertex of the last level while the
        // eccentricity grows
        bfs(seed);
        int start = *std::min_element(q.begin(), q.end(), less);
        reset();
        for (int it = 0; it < 8; ++it) {
            const int ecc = bfs(start);
            int cand = -1;
            for (int v : q)
                if (level[(size_t)v] == ecc && (cand < 0 || less(v, cand))) cand = v;
            reset();
            const int e2 = bfs(cand);
            reset();
            if (e2 <= ecc) break;
            start = cand;
        }
        bfs(start);
        for (int v : q) {
            done[(size_t)v] = 1;
            order.push_back(v);
        }
        reset();
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// ---------------------------------------------------------------- LDL^T

void Ldlt::analyze(const Csc& a, bool order) {
    n_ = a.cols;
    ok_ = false;
    perm_ = order ? rcm_order(a) : std::vector<int>();
    if (!order) {
        perm_.resize((size_t)n_);
        for (int k = 0; k < n_; ++k) perm_[(size_t)k] = k;
    }
    inv_.assign((size_t)n_, 0);
    for (int k = 0; k < n_; ++k) inv_[(size_t)perm_[(size_t)k]] = k;
    // upper triangle of P A P^T, with each entry's source position in a
    std::vector<std::vector<std::pair<int, int>>> cols((size_t)n_);
    for (int c = 0; c < n_; ++c)
        for (int q = a.p[(size_t)c]; q < a.p[(size_t)c + 1]; ++q) {
            const int r = inv_[(size_t)a.i[(size_t)q]], cc = inv_[(size_t)c];
            if (r <= cc) cols[(size_t)cc].push_back({r, q});
        }
    ap_ = Csc();
    ap_.rows = ap_.cols = n_;
    ap_.p.assign((size_t)n_ + 1, 0);
    ap_src_.clear();
    for (int c = 0; c < n_; ++c) {
        auto& v = cols[(size_t)c];
        std::sort(v.begin(), v.end());
        for (auto& e : v) {
            ap_.i.push_back(e.first);
            ap_src_.push_back(e.second);
        }
        ap_.p[(size_t)c + 1] = (int)ap_.i.size();
    }
    // elimination tree and the row patterns of L, appended per column in
    // increasing row order (so every column's rows are ascending)
    parent_.assign((size_t)n_, -1);
    std::vector<int> flag((size_t)n_, -1), cnt((size_t)n_, 0);
    for (int k = 0; k < n_; ++k) {
        flag[(size_t)k] = k;
        for (int q = ap_.p[(size_t)k]; q < ap_.p[(size_t)k + 1]; ++q) {
            for (int i = ap_.i[(size_t)q]; flag[(size_t)i] != k; i = parent_[(size_t)i]) {
                if (parent_[(size_t)i] == -1) parent_[(size_t)i] = k;
                cnt[(size_t)i]++;
                flag[(size_t)i] = k;
            }
        }
    }
    lp_.assign((size_t)n_ + 1, 0);
    for (int k = 0; k < n_; ++k) lp_[(size_t)k + 1] = lp_[(size_t)k] + cnt[(size_t)k];
    li_.assign((size_t)lp_[(size_t)n_], 0);
    std::vector<int> fill(lp_.begin(), lp_.end() - 1);
    std::fill(flag.begin(), flag.end(), -1);
    for (int k = 0; k < n_; ++k) {
        flag[(size_t)k] = k;
        for (int q = ap_.p[(size_t)k]; q < ap_.p[(size_t)k + 1]; ++q)
            for (int i = ap_.i[(size_t)q]; flag[(size_t)i] != k; i = parent_[(size_t)i]) {
                li_[(size_t)fill[(size_t)i]++] = k;
                flag[(size_t)i] = k;
            }
    }
    lx_.assign(li_.size(), 0.0);
    d_.assign((size_t)n_, 0.0);
    con_.assign((size_t)n_, 0);
    y_.assign((size_t)n_, 0.0);
    pattern_.assign((size_t)n_, 0);
    flag_.assign((size_t)n_, -1);
}

void Ldlt::set_constrained(const std::vector<char>& constrained) {
    for (int k = 0; k < n_; ++k) con_[(size_t)inv_[(size_t)k]] = (k < (int)constrained.size()) ? constrained[(size_t)k] : 0;
}

template <bool SEQ>
bool Ldlt::step(const Csc& a, int k) {
    // y = column k of A^B (rows <= k); pattern = reach of its rows in the tree
    int top = n_;
    flag_[(size_t)k] = k;
    y_[(size_t)k] = 0.0;
    for (int q = ap_.p[(size_t)k]; q < ap_.p[(size_t)k + 1]; ++q) {
        int i = ap_.i[(size_t)q];
        double v = a.x[(size_t)ap_src_[(size_t)q]];
        if (con_[(size_t)i] || con_[(size_t)k]) v = (i == k) ? 1.0 : 0.0;
        y_[(size_t)i] += v;
        int len = 0;
        for (; flag_[(size_t)i] != k; i = parent_[(size_t)i]) {
            pattern_[(size_t)len++] = i;
            flag_[(size_t)i] = k;
        }
        while (len > 0) pattern_[(size_t)--top] = pattern_[(size_t)--len];
    }
    double dk = y_[(size_t)k];
    y_[(size_t)k] = 0.0;
    for (; top < n_; ++top) {
        const int i = pattern_[(size_t)top];
        const double yi = y_[(size_t)i];
        y_[(size_t)i] = 0.0;
        // slot of row k in column i; the rows above it are L(:, i) of earlier
        // steps.  In a full pass (SEQ) the rows of a column are reached in
        // increasing k, so the slot is the column's next one.
        int s;
        if (SEQ) {
            s = next_[(size_t)i]++;
        } else {
            const int* b = li_.data() + lp_[(size_t)i];
            const int* e = li_.data() + lp_[(size_t)i + 1];
            s = (int)(std::lower_bound(b, e, k) - li_.data());
        }
        for (int p = lp_[(size_t)i]; p < s; ++p) y_[(size_t)li_[(size_t)p]] -= lx_[(size_t)p] * yi;
        const double lki = yi / d_[(size_t)i];
        dk -= lki * yi;
        lx_[(size_t)s] = lki;
    }
    d_[(size_t)k] = dk;
    return dk != 0.0 && std::isfinite(dk);
}

bool Ldlt::factorize(const Csc& a) {
    ok_ = false;
    std::fill(flag_.begin(), flag_.end(), -1);
    next_.assign(lp_.begin(), lp_.end() - 1);
    for (int k = 0; k < n_; ++k)
        if (!step<true>(a, k)) return false;
    last_steps_ = n_;
    ok_ = true;
    return true;
}

bool Ldlt::update(const Csc& a, const std::vector<int>& changed) {
    if (!ok_) return factorize(a);
    ok_ = false;
    std::vector<char> mark((size_t)n_, 0);
    std::vector<int> steps;
    for (int c : changed) {
        if (c < 0 || c >= n_) continue;
        for (int k = inv_[(size_t)c]; k != -1 && !mark[(size_t)k]; k = parent_[(size_t)k]) {
            mark[(size_t)k] = 1;
            steps.push_back(k);
        }
    }
    std::sort(steps.begin(), steps.end());
    // flags from earlier passes hold step numbers; clear them so a step k
    // recomputed now does not stop at a mark left by the previous step k
    std::fill(flag_.begin(), flag_.end(), -1);
    for (int k : steps)
        if (!step<false>(a, k)) return false;
    last_steps_ = (int)steps.size();
    ok_ = true;
    return true;
}

void Ldlt::solve(double* b) const {
    std::vector<double> x((size_t)n_);
    for (int k = 0; k < n_; ++k) x[(size_t)k] = b[perm_[(size_t)k]];
    for (int j = 0; j < n_; ++j)
        for (int p = lp_[(size_t)j]; p < lp_[(size_t)j + 1]; ++p) x[(size_t)li_[(size_t)p]] -= lx_[(size_t)p] * x[(size_t)j];
    for (int j = 0; j < n_; ++j) x[(size_t)j] /= d_[(size_t)j];
    for (int j = n_ - 1; j >= 0; --j)
        for (int p = lp_[(size_t)j]; p < lp_[(size_t)j + 1]; ++p) x[(size_t)j] -= lx_[(size_t)p] * x[(size_t)li_[(size_t)p]];
    for (int k = 0; k < n_; ++k) b[perm_[(size_t)k]] = x[(size_t)k];
}

// ---------------------------------------------------------------- Laplacian

// The reference's mesh stores its points in single precision (OpenMesh's
// default Vec3f points, read at LaplacianDeformation.cpp:179,350 and written
// at :388-389,568,617): every vertex position the deformation reads or
// writes is a float, promoted to double for the arithmetic.
static inline double mesh_point(double x) { return (double)(float)x; }

static std::vector<double> mesh_points(std::vector<double> v) {
    for (double& x : v) x = mesh_point(x);
    return v;
}

Laplacian::Laplacian(std::vector<double> verts, const std::vector<int32_t>& faces)
    : v_(mesh_points(std::move(verts))), rest_(v_) {
    const int nv = n();
    nbr_.assign((size_t)nv, {});
    for (size_t f = 0; f + 2 < faces.size(); f += 3)
        for (int e = 0; e < 3; ++e) {
            const int a = faces[f + (size_t)e], b = faces[f + (size_t)((e + 1) % 3)];
            nbr_[(size_t)a].push_back(b);
            nbr_[(size_t)b].push_back(a);
        }
    for (auto& r : nbr_) {
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
    }
    faces_ = faces;
}

double Laplacian::edge_weight(int a, int b) const {
    const auto& r = nbr_[(size_t)a];
    const auto it = std::lower_bound(r.begin(), r.end(), b);
    if (it == r.end() || *it != b) return 0.0;
    return wnbr_[(size_t)a][(size_t)(it - r.begin())];
}

void Laplacian::compute_weights() {
    // sum of cot of the angle opposite each edge over the faces holding it
    std::map<std::pair<int, int>, double> cot;
    auto P = [&](int v, int k) { return rest_[3 * (size_t)v + (size_t)k]; };
    for (size_t f = 0; f + 2 < faces_.size(); f += 3)
        for (int e = 0; e < 3; ++e) {
            const int s = faces_[f + (size_t)e], t = faces_[f + (size_t)((e + 1) % 3)];
            const int o = faces_[f + (size_t)((e + 2) % 3)];
            const double a0 = P(s, 0) - P(o, 0), a1 = P(s, 1) - P(o, 1), a2 = P(s, 2) - P(o, 2);
            const double b0 = P(t, 0) - P(o, 0), b1 = P(t, 1) - P(o, 1), b2 = P(t, 2) - P(o, 2);
            const double cosv = a0 * b0 + a1 * b1 + a2 * b2;
            const double c0 = a1 * b2 - a2 * b1, c1 = a2 * b0 - a0 * b2, c2 = a0 * b1 - a1 * b0;
            const double sinv = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
            double& acc = cot[{std::min(s, t), std::max(s, t)}];
            if (sinv > 1e-6) acc += cosv / sinv;
        }
    wnbr_.assign(nbr_.size(), {});
    for (size_t v = 0; v < nbr_.size(); ++v)
        for (int u : nbr_[v]) {
            const auto it = cot.find({std::min((int)v, u), std::max((int)v, u)});
            const double sum = it == cot.end() ? 0.0 : it->second;
            wnbr_[v].push_back(std::max(0.5 * sum, 1e-4));
        }
}

void Laplacian::initialize() {
    compute_weights();
    const int nv = n();
    std::vector<Triplet> t;
    for (int i = 0; i < nv; ++i) {
        double wsum = 0.0;
        for (double w : wnbr_[(size_t)i]) wsum += w;
        t.push_back({i, i, 1.0});
        for (size_t k = 0; k < nbr_[(size_t)i].size(); ++k) t.push_back({i, nbr_[(size_t)i][k], -wnbr_[(size_t)i][k] / wsum});
    }
    L_ = from_triplets(nv, nv, std::move(t));
    std::vector<double> x((size_t)nv);
    for (int a = 0; a < 3; ++a) {
        for (int i = 0; i < nv; ++i) x[(size_t)i] = rest_[3 * (size_t)i + (size_t)a];
        delta_[a].assign((size_t)nv, 0.0);
        gemv(L_, x.data(), delta_[a].data());
    }
}

bool Laplacian::is_anchor(int v) const { return std::find(anchors_.begin(), anchors_.end(), v) != anchors_.end(); }

void Laplacian::add_anchor(int v) {
    if (is_anchor(v)) return;
    anchors_.push_back(v);
    for (int a = 0; a < 3; ++a) anchor_pos_.push_back(v_[3 * (size_t)v + (size_t)a]);
}

void Laplacian::set_control_points(const std::vector<int>& idx, const std::vector<double>& targets) {
    ctrl_ = idx;
    ctrl_t_ = targets;
}

void Laplacian::clear_control_points() {
    ctrl_.clear();
    ctrl_t_.clear();
}

std::vector<int> Laplacian::kring(const std::vector<int>& seeds, int k) const {
    std::vector<char> seen((size_t)n(), 0);
    std::vector<int> all, frontier;
    for (int s : seeds)
        if (!seen[(size_t)s]) {
            seen[(size_t)s] = 1;
            frontier.push_back(s);
            all.push_back(s);
        }
    for (int hop = 0; hop < k; ++hop) {
        std::vector<int> next;
        for (int v : frontier)
            for (int u : nbr_[(size_t)v])
                if (!seen[(size_t)u]) {
                    seen[(size_t)u] = 1;
                    next.push_back(u);
                    all.push_back(u);
                }
        frontier.swap(next);
    }
    std::sort(all.begin(), all.end());
    return all;
}

bool Laplacian::precompute_local() {
    Local& L = local_;
    L.ready = false;
    if (ctrl_.empty()) return true;  // nothing to solve: the mesh stays as it is
    const std::vector<int> ring = kring(ctrl_, kring_);
    if (!L.analyzed || ring != L.ring) {
        // a new patch: its sub-Laplacians, Lii^T Lii and the factor's pattern
        // (kept while the handles move inside the same k-ring)
        L.analyzed = false;
        L.interior.clear();
        L.boundary.clear();
        L.g2l.assign((size_t)n(), -1);
        std::vector<int> g2b((size_t)n(), -1);
        for (int v : ring) {
            if (is_anchor(v)) {
                g2b[(size_t)v] = (int)L.boundary.size();
                L.boundary.push_back(v);
            } else {
                L.g2l[(size_t)v] = (int)L.interior.size();
                L.interior.push_back(v);
            }
        }
        const int ni = (int)L.interior.size();
        if (ni == 0) return true;
        std::vector<Triplet> tii, tib;
        for (int li = 0; li < ni; ++li) {
            const int gi = L.interior[(size_t)li];
            double wsum = 0.0;
            for (double w : wnbr_[(size_t)gi]) wsum += w;
            if (wsum < 1e-12) wsum = 1e-12;
            tii.push_back({li, li, 1.0});
            for (size_t k = 0; k < nbr_[(size_t)gi].size(); ++k) {
                const int gj = nbr_[(size_t)gi][k];
                const double w = wnbr_[(size_t)gi][k] / wsum;
                if (L.g2l[(size_t)gj] >= 0) {
                    tii.push_back({li, L.g2l[(size_t)gj], -w});
                } else {
                    if (g2b[(size_t)gj] < 0) {
                        g2b[(size_t)gj] = (int)L.boundary.size();
                        L.boundary.push_back(gj);
                    }
                    tib.push_back({li, g2b[(size_t)gj], -w});
                }
            }
        }
        L.Lii = from_triplets(ni, ni, std::move(tii));
        L.Lib = from_triplets(ni, (int)L.boundary.size(), std::move(tib));
        L.LiiT = transpose(L.Lii);
        L.LtL = multiply(L.LiiT, L.Lii);  // unit diagonal of Lii: every diagonal entry is present
        L.diag.assign((size_t)ni, -1);
        for (int c = 0; c < ni; ++c)
            for (int q = L.LtL.p[(size_t)c]; q < L.LtL.p[(size_t)c + 1]; ++q)
                if (L.LtL.i[(size_t)q] == c) L.diag[(size_t)c] = q;
        L.ldlt.analyze(L.LtL, true);
        L.ring = ring;
        L.analyzed = true;
    }
    L.handle_local.clear();
    L.handle_global.clear();
    L.handle_ctrl.clear();
    std::vector<int> first((size_t)n(), -1);
    for (size_t t = 0; t < ctrl_.size(); ++t)
        if (first[(size_t)ctrl_[t]] < 0) first[(size_t)ctrl_[t]] = (int)t;
    for (int h : ctrl_)
        if (L.g2l[(size_t)h] >= 0) {
            L.handle_local.push_back(L.g2l[(size_t)h]);
            L.handle_global.push_back(h);
            L.handle_ctrl.push_back(first[(size_t)h]);
        }
    if (L.handle_local.empty()) return true;
    // Lii^T Lii + I_handles (a handle listed twice counts twice)
    L.AtA = L.LtL;
    for (int c : L.handle_local) L.AtA.x[(size_t)L.diag[(size_t)c]] += 1.0;
    if (!L.ldlt.factorize(L.AtA)) return fail(FSI_ERR_STATE, "local LDL^T factorization failed") == FSI_OK;
    last_steps_ = L.ldlt.last_steps();
    L.ready = true;
    return true;
}

bool Laplacian::solve_local() {
    const Local& L = local_;
    if (!L.ready) return true;
    const int ni = (int)L.interior.size(), nb = (int)L.boundary.size();
    std::vector<double> xb((size_t)nb), b((size_t)ni), rhs[3];
    for (int a = 0; a < 3; ++a) {
        for (int k = 0; k < nb; ++k) xb[(size_t)k] = v_[3 * (size_t)L.boundary[(size_t)k] + (size_t)a];
        // b = rest delta of the interior - Lib * x_boundary (current)
        gemv(L.Lib, xb.data(), b.data());
        for (int li = 0; li < ni; ++li) b[(size_t)li] = delta_[a][(size_t)L.interior[(size_t)li]] - b[(size_t)li];
        rhs[a].assign((size_t)ni, 0.0);
        gemv_t(L.Lii, b.data(), rhs[a].data());
    }
    for (size_t c = 0; c < L.handle_local.size(); ++c) {
        // the target of the handle's first occurrence in the control list
        // (found again if the list was reordered since the precompute)
        const int gv = L.handle_global[c];
        size_t t = (size_t)L.handle_ctrl[c];
        if (t >= ctrl_.size() || ctrl_[t] != gv) {
            t = 0;
            while (t < ctrl_.size() && ctrl_[t] != gv) ++t;
        }
        if (t >= ctrl_.size()) continue;  // no longer a control point
        for (int a = 0; a < 3; ++a) rhs[a][(size_t)L.handle_local[c]] += ctrl_t_[3 * t + (size_t)a];
    }
    for (int a = 0; a < 3; ++a) {
        L.ldlt.solve(rhs[a].data());
        for (int li = 0; li < ni; ++li)
            v_[3 * (size_t)L.interior[(size_t)li] + (size_t)a] = mesh_point(rhs[a][(size_t)li]);
    }
    return true;
}

std::vector<int> Laplacian::constrained_sorted() const {
    std::vector<int> c = anchors_;
    c.insert(c.end(), ctrl_.begin(), ctrl_.end());
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    return c;
}

bool Laplacian::precompute() {
    const int nv = n();
    if (use_local_) return precompute_local();
    if (use_inc_) {
        const std::vector<int> cur = constrained_sorted();
        if (cur.empty()) return fail(FSI_ERR_STATE, "no constraints to precompute with") == FSI_OK;
        if (!inc_ready_) {
            Csc LtL = multiply(transpose(L_), L_);
            std::vector<Triplet> t;
            for (int c = 0; c < nv; ++c)
                for (int q = LtL.p[(size_t)c]; q < LtL.p[(size_t)c + 1]; ++q)
                    if (LtL.x[(size_t)q] != 0.0) t.push_back({LtL.i[(size_t)q], c, LtL.x[(size_t)q]});
            for (int k = 0; k < nv; ++k) t.push_back({k, k, 1e-8});
            A_ = from_triplets(nv, nv, std::move(t));
            for (int a = 0; a < 3; ++a) {
                b_[a].assign((size_t)nv, 0.0);
                gemv_t(L_, delta_[a].data(), b_[a].data());
            }
            inc_.analyze(A_, true);
            inc_ready_ = true;
            prev_.clear();
        }
        mask_.assign((size_t)nv, 0);
        for (int v : cur) mask_[(size_t)v] = 1;
        inc_.set_constrained(mask_);
        bool okf;
        if (prev_.empty()) {
            okf = inc_.factorize(A_);
        } else {
            std::vector<int> changed, cols;
            std::set_symmetric_difference(cur.begin(), cur.end(), prev_.begin(), prev_.end(), std::back_inserter(changed));
            std::vector<char> seen((size_t)nv, 0);
            for (int v : changed) {
                if (!seen[(size_t)v]) {
                    seen[(size_t)v] = 1;
                    cols.push_back(v);
                }
                for (int q = A_.p[(size_t)v]; q < A_.p[(size_t)v + 1]; ++q) {
                    const int r = A_.i[(size_t)q];
                    if (!seen[(size_t)r]) {
                        seen[(size_t)r] = 1;
                        cols.push_back(r);
                    }
                }
            }
            okf = inc_.update(A_, cols);
        }
        if (!okf) return fail(FSI_ERR_STATE, "incremental LDL^T: zero pivot") == FSI_OK;
        last_steps_ = inc_.last_steps();
        prev_ = cur;
        return true;
    }
    // global soft path: A = [L; one row per anchor, then per handle]
    const int nc = (int)(anchors_.size() + ctrl_.size());
    if (nc == 0) return fail(FSI_ERR_STATE, "no constraints to precompute with") == FSI_OK;
    std::vector<Triplet> t;
    for (int c = 0; c < nv; ++c)
        for (int q = L_.p[(size_t)c]; q < L_.p[(size_t)c + 1]; ++q) t.push_back({L_.i[(size_t)q], c, L_.x[(size_t)q]});
    int row = nv;
    for (int a : anchors_) t.push_back({row++, a, 1.0});
    for (int h : ctrl_) t.push_back({row++, h, 1.0});
    const Csc A = from_triplets(nv + nc, nv, std::move(t));
    At_ = transpose(A);
    const Csc AtA = multiply(At_, A);
    glob_.analyze(AtA, true);
    glob_ready_ = false;
    if (!glob_.factorize(AtA)) return fail(FSI_ERR_STATE, "LDL^T factorization failed") == FSI_OK;
    last_steps_ = glob_.last_steps();
    glob_ready_ = true;
    return true;
}

bool Laplacian::solve() {
    const int nv = n();
    if (use_local_) return solve_local();
    if (use_inc_) {
        if (!inc_ready_ || !inc_.ok()) return fail(FSI_ERR_STATE, "solve before precompute") == FSI_OK;
        std::vector<double> tg((size_t)nv * 3, 0.0);
        for (size_t i = 0; i < anchors_.size(); ++i)
            for (int a = 0; a < 3; ++a) tg[3 * (size_t)anchors_[i] + (size_t)a] = anchor_pos_[3 * i + (size_t)a];
        for (size_t i = 0; i < ctrl_.size(); ++i)
            for (int a = 0; a < 3; ++a) tg[3 * (size_t)ctrl_[i] + (size_t)a] = ctrl_t_[3 * i + (size_t)a];
        for (int a = 0; a < 3; ++a) {
            std::vector<double> rhs = b_[a];
            for (int k = 0; k < nv; ++k) {
                if (!mask_[(size_t)k]) continue;
                const double ck = tg[3 * (size_t)k + (size_t)a];
                for (int q = A_.p[(size_t)k]; q < A_.p[(size_t)k + 1]; ++q) rhs[(size_t)A_.i[(size_t)q]] -= A_.x[(size_t)q] * ck;
            }
            for (int k = 0; k < nv; ++k)
                if (mask_[(size_t)k]) rhs[(size_t)k] = tg[3 * (size_t)k + (size_t)a];
            inc_.solve(rhs.data());
            for (int k = 0; k < nv; ++k) v_[3 * (size_t)k + (size_t)a] = mesh_point(rhs[(size_t)k]);
        }
        return true;
    }
    if (!glob_ready_) return fail(FSI_ERR_STATE, "solve before precompute") == FSI_OK;
    const int nc = (int)(anchors_.size() + ctrl_.size());
    std::vector<double> b((size_t)(nv + nc)), rhs((size_t)nv);
    for (int a = 0; a < 3; ++a) {
        for (int k = 0; k < nv; ++k) b[(size_t)k] = delta_[a][(size_t)k];
        int row = nv;
        for (size_t i = 0; i < anchors_.size(); ++i) b[(size_t)row++] = anchor_pos_[3 * i + (size_t)a];
        for (size_t i = 0; i < ctrl_.size(); ++i) b[(size_t)row++] = ctrl_t_[3 * i + (size_t)a];
        gemv(At_, b.data(), rhs.data());
        glob_.solve(rhs.data());
        for (int k = 0; k < nv; ++k) v_[3 * (size_t)k + (size_t)a] = mesh_point(rhs[(size_t)k]);
    }
    return true;
}

double Laplacian::verify_incremental() {
    if (use_local_ || !use_inc_ || !inc_ready_ || !inc_.ok()) return -1.0;
    const int nv = n();
    Ldlt fresh;
    fresh.analyze(A_, true);
    fresh.set_constrained(mask_);
    if (!fresh.factorize(A_)) return -2.0;
    std::vector<double> r1((size_t)nv), r2;
    for (int k = 0; k < nv; ++k) r1[(size_t)k] = b_[0][(size_t)k] + (mask_[(size_t)k] ? 1.0 : 0.0);
    r2 = r1;
    inc_.solve(r1.data());
    fresh.solve(r2.data());
    double m = 0.0;
    for (int k = 0; k < nv; ++k) m = std::max(m, std::fabs(r1[(size_t)k] - r2[(size_t)k]));
    return m;
}

}  // namespace deform
}  // namespace fsi

// ---------------------------------------------------------------- C-ABI

struct fsi_deform {
    fsi::deform::Laplacian lap;
    std::vector<char> anchor;     // per vertex
    std::vector<int> last;        // handle set of the last step (ascending)
    uint32_t refactors = 0, handles = 0;
    fsi_deform(std::vector<double> v, const std::vector<int32_t>& f) : lap(std::move(v), f) {}
};

using fsi::fail;

namespace {
int check(const fsi_deform* d) { return d ? FSI_OK : fail(FSI_ERR_INVALID_ARG, "deform handle is NULL"); }
}  // namespace

extern "C" {

int fsi_deform_params_default(fsi_deform_params* p) {
    if (!p) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    // src/main.cpp:251, 300-304, 343
    p->stiffness = 1.0f;
    p->max_displacement = 0.02f;
    p->max_total_disp = 0.40f;
    p->restore_enabled = 1u;
    p->restore_strength = 0.05f;
    p->control_threshold = 0.50f;
    p->control_threshold_low = 0.30f;
    return FSI_OK;
}

int fsi_deform_create(const double* verts, uint32_t nverts, const int32_t* faces, uint32_t nfaces, fsi_deform** out) {
    if (!out || (nverts && !verts) || (nfaces && !faces)) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    if (nverts == 0 || nverts > (1u << 26) || nfaces > (1u << 27)) return fail(FSI_ERR_INVALID_ARG, "mesh size");
    for (size_t k = 0; k < 3 * (size_t)nfaces; ++k)
        if (faces[k] < 0 || (uint32_t)faces[k] >= nverts) return fail(FSI_ERR_INVALID_ARG, "face index out of range");
    fsi_deform* d = new (std::nothrow) fsi_deform(std::vector<double>(verts, verts + 3 * (size_t)nverts),
                                                  std::vector<int32_t>(faces, faces + 3 * (size_t)nfaces));
    if (!d) return fail(FSI_ERR_INVALID_ARG, "out of host memory");
    d->lap.initialize();
    d->anchor.assign(nverts, 0);
    *out = d;
    return FSI_OK;
}

int fsi_deform_destroy(fsi_deform* d) {
    delete d;
    return FSI_OK;
}

int fsi_deform_set_mode(fsi_deform* d, int use_local, int kring, int use_incremental) {
    if (int rc = check(d)) return rc;
    if (kring < 0) return fail(FSI_ERR_INVALID_ARG, "kring < 0");
    d->lap.set_use_local(use_local != 0);
    d->lap.set_local_kring(kring);
    d->lap.set_use_incremental(use_incremental != 0);
    return FSI_OK;
}

int fsi_deform_add_anchor(fsi_deform* d, int32_t vertex) {
    if (int rc = check(d)) return rc;
    if (vertex < 0 || vertex >= d->lap.n()) return fail(FSI_ERR_INVALID_ARG, "vertex out of range");
    d->lap.add_anchor(vertex);
    d->anchor[(size_t)vertex] = 1;
    return FSI_OK;
}

int fsi_deform_set_control_points(fsi_deform* d, const int32_t* idx, const double* targets, uint32_t count) {
    if (int rc = check(d)) return rc;
    if (count && (!idx || !targets)) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    for (uint32_t k = 0; k < count; ++k)
        if (idx[k] < 0 || idx[k] >= d->lap.n()) return fail(FSI_ERR_INVALID_ARG, "vertex out of range");
    d->lap.set_control_points(std::vector<int>(idx, idx + count), std::vector<double>(targets, targets + 3 * (size_t)count));
    return FSI_OK;
}

int fsi_deform_precompute(fsi_deform* d) {
    if (int rc = check(d)) return rc;
    if (!d->lap.precompute()) return FSI_ERR_STATE;
    d->refactors++;
    return FSI_OK;
}

int fsi_deform_solve(fsi_deform* d) {
    if (int rc = check(d)) return rc;
    return d->lap.solve() ? FSI_OK : FSI_ERR_STATE;
}

int fsi_deform_read_vertices(fsi_deform* d, double* verts) {
    if (int rc = check(d)) return rc;
    if (!verts) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    std::memcpy(verts, d->lap.vertices().data(), sizeof(double) * d->lap.vertices().size());
    return FSI_OK;
}

int fsi_deform_step(fsi_deform* d, const float* forces, float dt, const fsi_deform_params* p, uint32_t* handle_count,
                    int* solved) {
    if (int rc = check(d)) return rc;
    if (!forces || !p) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    const int nv = d->lap.n();
    const std::vector<double>& cur = d->lap.vertices();
    const std::vector<double>& rest = d->lap.rest();
    std::vector<int> hs;
    std::vector<double> ts;
    for (int i = 0; i < nv; ++i) {
        if (d->anchor[(size_t)i]) continue;  // anchors have their own rows
        const float fx = forces[3 * (size_t)i], fy = forces[3 * (size_t)i + 1], fz = forces[3 * (size_t)i + 2];
        const float fmag = std::sqrt(fx * fx + fy * fy + fz * fz);
        const bool was = std::binary_search(d->last.begin(), d->last.end(), i);
        // hysteresis: enter at the high threshold, leave below the low one
        const bool pushed = fmag >= p->control_threshold || (was && fmag >= p->control_threshold_low);
        double off0 = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double e = cur[3 * (size_t)i + (size_t)a] - rest[3 * (size_t)i + (size_t)a];
            off0 += e * e;
        }
        const bool relaxing = p->restore_enabled && was && std::sqrt(off0) > 1e-3;
        if (!pushed && !relaxing) continue;
        double t[3] = {cur[3 * (size_t)i], cur[3 * (size_t)i + 1], cur[3 * (size_t)i + 2]};
        if (pushed) {
            // the fluid's push, clamped per frame (FP32 as PVec3)
            const float g = p->stiffness * dt;
            float dx = fx * g, dy = fy * g, dz = fz * g;
            const float dm = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (dm > p->max_displacement && dm > 1e-12f) {
                const float s = p->max_displacement / dm;
                dx = dx * s;
                dy = dy * s;
                dz = dz * s;
            }
            t[0] += (double)dx;
            t[1] += (double)dy;
            t[2] += (double)dz;
        }
        if (p->restore_enabled)
            for (int a = 0; a < 3; ++a) t[a] += (double)p->restore_strength * (rest[3 * (size_t)i + (size_t)a] - t[a]);
        double fr[3], o2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            fr[a] = t[a] - rest[3 * (size_t)i + (size_t)a];
            o2 += fr[a] * fr[a];
        }
        const double off = std::sqrt(o2);
        if (off > p->max_total_disp)
            for (int a = 0; a < 3; ++a) t[a] = rest[3 * (size_t)i + (size_t)a] + fr[a] * (p->max_total_disp / off);
        hs.push_back(i);
        ts.insert(ts.end(), t, t + 3);
    }
    const bool changed = hs != d->last;
    d->lap.set_control_points(hs, ts);
    if (changed) {
        if (!d->lap.precompute()) return FSI_ERR_STATE;
        d->refactors++;
    }
    d->last = hs;
    d->handles = (uint32_t)hs.size();
    int did = 0;
    if (!hs.empty() || changed) {
        if (!d->lap.solve()) return FSI_ERR_STATE;
        did = 1;
    }
    if (handle_count) *handle_count = (uint32_t)hs.size();
    if (solved) *solved = did;
    return FSI_OK;
}

int fsi_deform_info(fsi_deform* d, uint32_t out[4]) {
    if (int rc = check(d)) return rc;
    if (!out) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    out[0] = (uint32_t)d->lap.local_patch_size();
    out[1] = (uint32_t)d->lap.last_factor_steps();
    out[2] = d->handles;
    out[3] = d->refactors;
    return FSI_OK;
}

int fsi_deform_verify_incremental(fsi_deform* d, double* max_diff) {
    if (int rc = check(d)) return rc;
    if (!max_diff) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    *max_diff = d->lap.verify_incremental();
    return FSI_OK;
}

}  // extern "C"
