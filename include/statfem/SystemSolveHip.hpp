// SystemSolveHip — a drop-in of statFEM's SystemSolve concept
// (corlib/SystemSolveEigenSparse.hpp:54-122) whose operator, right-hand side
// and solves live on an MI355X, over the C ABI of statfem_hip.h.
//
// Header-only C++17. The element types are template parameters so that the
// reference's Eigen::MatrixXd / Eigen::VectorXd / std::vector<unsigned> work
// unchanged: a matrix needs rows(), cols() and operator()(i, j); a vector
// needs size() and operator()(i) or operator[](i).
//
// Semantics kept from the reference:
//  - insertToMatrix appends (row, col, value) triplets; the CSR operator is the
//    Triplet's (corlib/Triplet.hpp:126-170,264-339), built on the GPU at the
//    first solve after a change (sfh_system_create_triplets);
//  - applyConstraints / fixDOF act in call order like Triplet::zeroRow
//    (corlib/Triplet.hpp:219-260): entries of the row inserted before the
//    constraint become zero and the diagonal becomes `scalar`, entries
//    inserted afterwards add to it; F[dof] = factor * value * scalar
//    (corlib/SystemSolveEigenSparse.hpp:197-223);
//  - clearMatrix keeps the pattern (a re-assembly with the same insertion
//    sequence only re-merges values on the GPU), clearRhs zeroes F;
//  - errors end the program like FTL_VERIFY (corlib/verify.hpp:38-41) unless
//    STATFEM_HIP_VERIFY is defined by the includer.
//
// Solves are the library's mixed-precision multi-RHS PCG to a relative FP64
// residual (default 1e-10, see options()); reuseFactoredMatrix is accepted
// for interface compatibility (there is no factorisation to reuse).
//
// A solver handed to AssemblerHip (AssemblerHip.hpp) holds a whole-mesh
// system instead: the operator is assembled on the device for all elements at
// once, constraints act on its rows there, and per-element insertToMatrix is
// refused (insertToRhs still accumulates into the host right-hand side).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>

#include "statfem_hip.h"

#ifndef STATFEM_HIP_VERIFY
#define STATFEM_HIP_VERIFY(cond, what)                                                               \
    do {                                                                                             \
        if (!(cond)) {                                                                               \
            std::fprintf(stderr, "statfem_hip: %s (%s)\n", (what), sfh_last_error());                \
            std::abort();                                                                            \
        }                                                                                            \
    } while (0)
#endif

namespace statfem_hip {

class AssemblerHip;

namespace detail {
template <class V>
auto at(const V& v, std::size_t i, int) -> decltype(v(i)) { return v(i); }
template <class V>
auto at(const V& v, std::size_t i, long) -> decltype(v[i]) { return v[i]; }
template <class V>
double get(const V& v, std::size_t i) { return static_cast<double>(at(v, i, 0)); }

inline void check(int status, const char* what) { STATFEM_HIP_VERIFY(status == SFH_OK, what); }
}  // namespace detail

// One device context per process (device 0 unless set before first use).
inline sfh_ctx* default_context(int device = 0) {
    struct Holder {
        sfh_ctx* c = nullptr;
        explicit Holder(int d) { detail::check(sfh_ctx_create(d, &c), "sfh_ctx_create"); }
        ~Holder() { sfh_ctx_destroy(c); }
    };
    static Holder h(device);
    return h.c;
}

// Owning device buffer of doubles.
class DeviceVector {
public:
    DeviceVector() = default;
    DeviceVector(sfh_ctx* c, std::size_t n) : ctx_(c), n_(n) {
        detail::check(sfh_malloc(c, n * sizeof(double), &p_), "sfh_malloc");
    }
    DeviceVector(const DeviceVector&) = delete;
    DeviceVector& operator=(const DeviceVector&) = delete;
    DeviceVector(DeviceVector&& o) noexcept { swap(o); }
    DeviceVector& operator=(DeviceVector&& o) noexcept { swap(o); return *this; }
    ~DeviceVector() { if (p_) sfh_free(ctx_, p_); }
    double* data() const { return static_cast<double*>(p_); }
    std::size_t size() const { return n_; }
    void upload(const double* h) { detail::check(sfh_memcpy_h2d(ctx_, p_, h, n_ * sizeof(double)), "upload"); }
    void download(double* h) const { detail::check(sfh_memcpy_d2h(ctx_, h, p_, n_ * sizeof(double)), "download"); }
private:
    void swap(DeviceVector& o) { std::swap(ctx_, o.ctx_); std::swap(n_, o.n_); std::swap(p_, o.p_); }
    sfh_ctx* ctx_ = nullptr;
    std::size_t n_ = 0;
    void* p_ = nullptr;
};

class SystemSolveHip {
public:
    using Constraints = std::vector<std::pair<unsigned, double>>;

    explicit SystemSolveHip(const unsigned& numDofs, sfh_ctx* ctx = nullptr)
        : ctx_(ctx ? ctx : default_context()), n_(numDofs), F_(numDofs, 0.0), X_(numDofs, 0.0),
          rowEpoch_(numDofs, -1) {
        sfh_solve_opts_default(&opts_);
    }
    ~SystemSolveHip() { if (sys_) sfh_system_destroy(sys_); }
    SystemSolveHip(const SystemSolveHip&) = delete;
    SystemSolveHip& operator=(const SystemSolveHip&) = delete;

    // ---------------------------------------------------------- insertion
    template <class MAT, class IDX>
    void insertToMatrix(const MAT& eStiff, const IDX& rowIndices, const IDX& colIndices) {
        STATFEM_HIP_VERIFY(!meshMode_, "insertToMatrix on a system assembled by AssemblerHip");
        for (std::size_t i = 0; i < rowIndices.size(); ++i)
            for (std::size_t j = 0; j < colIndices.size(); ++j)
                push(static_cast<int>(rowIndices[i]), static_cast<int>(colIndices[j]), eStiff(i, j));
    }
    template <class VEC, class IDX>
    void insertToRhs(const VEC& eForce, const IDX& dofIndices) {
        for (std::size_t i = 0; i < dofIndices.size(); ++i) F_[dofIndices[i]] += detail::get(eForce, i);
    }

    // --------------------------------------------------------- constraints
    void applyConstraints(const Constraints& constraints, const double factor, const double scalar) {
        if (meshMode_) {   // rows of the device operator, in one call
            std::vector<int> dofs;
            std::vector<double> vals;
            for (const auto& c : constraints) {
                dofs.push_back(static_cast<int>(c.first));
                vals.push_back(c.second);
                F_[c.first] = factor * c.second * scalar;
            }
            if (!dofs.empty())
                detail::check(sfh_apply_constraints(sys_, static_cast<int>(dofs.size()), dofs.data(), vals.data(),
                                                    factor, scalar), "sfh_apply_constraints");
            return;
        }
        for (const auto& c : constraints) {
            zeroRow(c.first, scalar);
            F_[c.first] = factor * c.second * scalar;
        }
    }
    void fixDOF(const unsigned index, const double scalar = 1.) {
        if (meshMode_) {
            applyConstraints(Constraints{{index, 0.0}}, 1.0, scalar);
            return;
        }
        zeroRow(index, scalar);
        F_[index] = 0.;
    }

    // The reference's PETSc-compatibility hooks; nothing to do here.
    void finishRhsAssembly() {}
    void finishMatAssembly() {}
    void finishAssembly() {}

    // -------------------------------------------------------------- solves
    void solveSystem() { solveSystem(false); }
    void solveSystem(const bool /*reuseFactoredMatrix*/) {
        prepare();
        DeviceVector Xd(ctx_, n_);
        sfh_solve_stats st;
        detail::check(sfh_solve_rhs(sys_, Xd.data(), &opts_, &st), "solveSystem");
        Xd.download(X_.data());
        stats_ = st;
    }
    template <class VEC>
    void solveSystemTranspose(const VEC& rhsVec, VEC& uVec, const bool /*reuse*/ = false) {
        prepare();
        std::vector<double> b(n_);
        for (unsigned i = 0; i < n_; ++i) b[i] = detail::get(rhsVec, i);
        std::vector<double> x(n_);
        solveMany(1, true, b.data(), x.data());
        resizeLike(uVec, n_);
        for (unsigned i = 0; i < n_; ++i) set(uVec, i, x[i]);
    }
    // Batched extension: Rt (N x k) = A^-T Pt (N x k) in one multi-RHS solve,
    // the loop of ForwardGaussRHS::updateSolution (statX/infer/ForwardGaussRHS.hpp:105-112).
    template <class MAT>
    void solveSystemTransposeMulti(const MAT& Pt, MAT& Rt, const bool /*reuse*/ = false) {
        prepare();
        const std::size_t k = static_cast<std::size_t>(Pt.cols());
        std::vector<double> b(n_ * k), x(n_ * k);
        for (unsigned i = 0; i < n_; ++i)
            for (std::size_t c = 0; c < k; ++c) b[i * k + c] = Pt(i, c);
        solveMany(static_cast<int>(k), true, b.data(), x.data());
        for (unsigned i = 0; i < n_; ++i)
            for (std::size_t c = 0; c < k; ++c) Rt(i, c) = x[i * k + c];
    }

    // --------------------------------------------------------- accessors
    template <class IDX, class VEC>
    void giveSolution(const IDX& dofIndices, VEC& result) const { gather(X_, dofIndices, result); }
    template <class IDX, class VEC>
    void giveForce(const IDX& dofIndices, VEC& result) const { gather(F_, dofIndices, result); }

    void clearMatrix() {
        if (meshMode_) {
            detail::check(sfh_clear_matrix(sys_), "sfh_clear_matrix");
            return;
        }
        for (auto& v : vals_) v = 0.0;
        cursor_ = 0;        // a re-assembly that repeats the insertion sequence keeps the pattern
        replay_ = !rows_.empty();
        constrained_.clear();
        rowEpoch_.assign(n_, -1);
        matrixChanged_ = true;
    }
    void clearRhs() { for (auto& f : F_) f = 0.0; }

    double normMatrix() {
        prepare();
        std::vector<double> v(static_cast<std::size_t>(nnz_));
        detail::check(sfh_matrix_values(sys_, v.data()), "sfh_matrix_values");
        double s = 0.0;
        for (double a : v) s += a * a;
        return std::sqrt(s);
    }
    double normRhs() const { double s = 0; for (double f : F_) s += f * f; return std::sqrt(s); }
    double norm1Rhs() const { double s = 0; for (double f : F_) s += std::fabs(f); return s; }
    double normSol() const { double s = 0; for (double x : X_) s += x * x; return std::sqrt(s); }
    unsigned numNonZeros() { prepare(); return static_cast<unsigned>(nnz_); }

    // ------------------------------------------------------ library extras
    sfh_solve_opts& options() { return opts_; }
    const sfh_solve_stats& lastStats() const { return stats_; }
    sfh_system* handle() { prepare(); return sys_; }
    sfh_ctx* context() const { return ctx_; }
    unsigned numDofs() const { return n_; }
    bool wholeMesh() const { return meshMode_; }

private:
    friend class AssemblerHip;
    // AssemblerHip hands over its whole-mesh system (the solver owns it from here)
    void bindMesh(sfh_system* sys) {
        if (sys_) sfh_system_destroy(sys_);
        sys_ = sys;
        meshMode_ = true;
        int nd = 0;
        detail::check(sfh_system_info(sys_, &nd, &nnz_), "sfh_system_info");
        STATFEM_HIP_VERIFY(static_cast<unsigned>(nd) == n_, "mesh DOFs differ from the solver's numDofs");
        rows_.clear(); cols_.clear(); vals_.clear(); epoch_.clear();
        patternChanged_ = matrixChanged_ = false;
    }
    void addForce(const std::vector<double>& f) {
        for (unsigned i = 0; i < n_; ++i) F_[i] += f[i];
    }

    void push(int r, int c, double v) {
        const int epoch = static_cast<int>(rowsSeen());
        if (replay_) {   // same insertion sequence after clearMatrix: only values change
            if (cursor_ < rows_.size() && rows_[cursor_] == r && cols_[cursor_] == c) {
                vals_[cursor_] = v;
                epoch_[cursor_] = epoch;
                ++cursor_;
                matrixChanged_ = true;
                return;
            }
            replay_ = false;   // the sequence changed: rebuild the pattern
            rows_.resize(cursor_);
            cols_.resize(cursor_);
            vals_.resize(cursor_);
            epoch_.resize(cursor_);
            patternChanged_ = true;
        }
        rows_.push_back(r);
        cols_.push_back(c);
        vals_.push_back(v);
        epoch_.push_back(epoch);
        ++cursor_;
        patternChanged_ = true;
        matrixChanged_ = true;
    }
    std::size_t rowsSeen() const { return events_; }

    // Triplet::zeroRow: later-inserted entries survive; the diagonal holds scalar
    void zeroRow(unsigned row, double scalar) {
        ++events_;
        rowEpoch_[row] = static_cast<int>(events_);
        push(static_cast<int>(row), static_cast<int>(row), scalar);
        epoch_[cursor_ - 1] = static_cast<int>(events_);   // the diagonal belongs to this constraint
        constrained_.push_back(row);
        ++events_;
    }

    void prepare() {
        if (meshMode_) {
            detail::check(sfh_system_set_rhs(sys_, F_.data()), "sfh_system_set_rhs");
            return;
        }
        if (replay_ && cursor_ != rows_.size()) {   // fewer insertions than before: new pattern
            rows_.resize(cursor_); cols_.resize(cursor_); vals_.resize(cursor_); epoch_.resize(cursor_);
            patternChanged_ = true;
        }
        replay_ = false;
        if (!matrixChanged_ && sys_) {
            detail::check(sfh_system_set_rhs(sys_, F_.data()), "sfh_system_set_rhs");
            return;
        }
        // values after the row zeroings, in call order
        std::vector<double> v(vals_);
        for (std::size_t i = 0; i < v.size(); ++i) {
            const int last = rowEpoch_[rows_[i]];
            if (last >= 0 && epoch_[i] < last) v[i] = 0.0;
        }
        const long long n = static_cast<long long>(rows_.size());
        STATFEM_HIP_VERIFY(n > 0, "no matrix entries were inserted");
        if (patternChanged_ || !sys_) {
            if (sys_) sfh_system_destroy(sys_);
            sys_ = nullptr;
            detail::check(sfh_system_create_triplets(ctx_, static_cast<int>(n_), n, rows_.data(), cols_.data(),
                                                     v.data(), &sys_), "sfh_system_create_triplets");
            int nd = 0;
            detail::check(sfh_system_info(sys_, &nd, &nnz_), "sfh_system_info");
        } else {
            detail::check(sfh_system_set_triplet_values(sys_, n, v.data()), "sfh_system_set_triplet_values");
        }
        // rows whose last event is their constraint are identity rows: the solver eliminates them
        std::vector<char> hasLater(n_, 0);
        for (std::size_t i = 0; i < rows_.size(); ++i) {
            const int last = rowEpoch_[rows_[i]];
            if (last >= 0 && epoch_[i] > last) hasLater[rows_[i]] = 1;
        }
        std::vector<double> dsum(n_, 0.0);
        for (std::size_t i = 0; i < rows_.size(); ++i)
            if (rows_[i] == cols_[i]) dsum[rows_[i]] += v[i];
        std::vector<char> seen(n_, 0);
        std::vector<std::pair<double, int>> clean;   // (diagonal value, row) of identity rows
        for (unsigned r : constrained_)
            if (!seen[r] && !hasLater[r]) {
                seen[r] = 1;
                clean.emplace_back(dsum[r], static_cast<int>(r));
            }
        std::sort(clean.begin(), clean.end());
        for (std::size_t a = 0; a < clean.size();) {   // one call per distinct diagonal value
            std::size_t b = a;
            std::vector<int> rows;
            while (b < clean.size() && clean[b].first == clean[a].first) rows.push_back(clean[b++].second);
            detail::check(sfh_apply_constraints(sys_, static_cast<int>(rows.size()), rows.data(), nullptr, 1.0,
                                                clean[a].first), "sfh_apply_constraints");
            a = b;
        }
        detail::check(sfh_system_set_rhs(sys_, F_.data()), "sfh_system_set_rhs");
        patternChanged_ = false;
        matrixChanged_ = false;
    }

    void solveMany(int k, bool trans, const double* B, double* X) {
        DeviceVector Bd(ctx_, n_ * k), Xd(ctx_, n_ * k);
        Bd.upload(B);
        sfh_solve_stats st;
        detail::check(sfh_solve(sys_, trans ? 1 : 0, k, Bd.data(), Xd.data(), &opts_, &st), "sfh_solve");
        Xd.download(X);
        stats_ = st;
    }

    template <class IDX, class VEC>
    static void gather(const std::vector<double>& src, const IDX& dofs, VEC& out) {
        resizeLike(out, dofs.size());
        for (std::size_t i = 0; i < dofs.size(); ++i) set(out, i, src[dofs[i]]);
    }
    template <class VEC>
    static auto resizeLike(VEC& v, std::size_t n) -> decltype(v.resize(n), void()) { v.resize(n); }
    template <class VEC>
    static auto set(VEC& v, std::size_t i, double x) -> decltype(v[i] = x, void()) { v[i] = x; }

    sfh_ctx* ctx_;
    unsigned n_;
    sfh_system* sys_ = nullptr;
    long long nnz_ = 0;
    std::vector<int> rows_, cols_, epoch_;
    std::vector<double> vals_;
    std::vector<double> F_, X_;
    std::vector<int> rowEpoch_;
    std::vector<unsigned> constrained_;
    std::size_t cursor_ = 0, events_ = 0;
    bool replay_ = false, patternChanged_ = true, matrixChanged_ = true, meshMode_ = false;
    sfh_solve_opts opts_;
    sfh_solve_stats stats_{};
};

}  // namespace statfem_hip
