// Header-only shim that plugs libcfdxhip into the reference's solver
// interface. Include it inside the reference tree (it needs
// cfdx/core/linalg/preconditioner.h and cg_solver.h) and link libcfdxhip.so.
//
//   cfdx::core::Preconditioner            core/linalg/preconditioner.h:18-28
//   NativeAMGPreconditioner lifecycle     core/linalg/hypre_amg.h:26-101
//   solve_cg_prepared                     core/linalg/cg_solver.h:653-667
//
// setup/update_values/apply return false on failure and keep the reason in
// last_error(), as the reference does; nothing throws for numerical failure.
// One instance per thread (the reference's instances are not reentrant
// either).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

#include "cfdx/core/linalg/cg_solver.h"
#include "cfdx/core/linalg/preconditioner.h"
#include "cfdx/core/linalg/solver_status.h"
#include "cfdxhip.h"

namespace cfdxhip_shim {

class HipAmgPreconditioner final : public cfdx::core::Preconditioner {
public:
    explicit HipAmgPreconditioner(const cfdxhip_amg_options* options = nullptr) {
        cfdxhip_amg_options o;
        cfdxhip_amg_options_default(&o);
        if (options) o = *options;
        if (cfdxhip_amg_create(&o, &h_) != CFDXHIP_OK) {
            h_ = nullptr;
            error_ = "cfdxhip_amg_create failed";
        }
    }
    ~HipAmgPreconditioner() override { cfdxhip_amg_destroy(h_); }
    HipAmgPreconditioner(const HipAmgPreconditioner&) = delete;
    HipAmgPreconditioner& operator=(const HipAmgPreconditioner&) = delete;

    bool setup(const cfdx::core::SparseMatrix& A) override {
        if (!h_ || A.n_rows() != A.n_cols() || A.n_rows() == 0) return fail("invalid matrix or handle");
        n_ = A.n_rows();
        return check(cfdxhip_amg_setup(h_, n_, A.row_offsets_data(), A.columns_data(), A.values_data()));
    }
    bool update_values(const cfdx::core::SparseMatrix& A) override {
        if (!h_ || A.n_rows() != n_) return fail("numeric update requires an existing compatible hierarchy");
        return check(cfdxhip_amg_update_values(h_, A.values_data()));
    }
    bool apply(const cfdx::core::Vector& r, cfdx::core::Vector& z) const override {
        if (!h_ || r.size() != n_) return fail("native AMG preconditioner has not been set up");
        if (z.size() != n_) z.resize(n_);
        return check(cfdxhip_amg_apply(h_, r.data(), z.data()));
    }
    const char* name() const override { return "CfdxHipSmoothedAggregationAMG"; }
    const std::string& last_error() const noexcept { return error_; }

    // solve_cg_prepared(A, b, x, *this, max_iter, tol) on the device: the
    // whole PCG runs on the GPU with the FP64 true residual every iteration.
    cfdx::core::SolverResult solve_cg(const cfdx::core::Vector& b, cfdx::core::Vector& x, std::size_t max_iter,
                                      double tolerance) const {
        cfdxhip_result r{};
        cfdx::core::SolverResult out;
        if (!h_ || b.size() != n_ || x.size() != n_ ||
            !check(cfdxhip_amg_solve_cg(h_, b.data(), x.data(), max_iter, tolerance, &r)))
            return out;  // NOT_APPLICABLE
        out.status = static_cast<cfdx::core::SolverStatus>(r.status);
        out.iterations = r.iterations;
        out.residual = r.residual;
        out.residual_relative = r.residual_relative;
        out.residual_replacements = r.residual_replacements;
        return out;
    }

private:
    bool fail(const char* m) const {
        error_ = m;
        return false;
    }
    bool check(cfdxhip_status s) const {
        if (s == CFDXHIP_OK) {
            error_.clear();
            return true;
        }
        error_ = cfdxhip_amg_last_error(h_);
        return false;
    }
    cfdxhip_amg* h_ = nullptr;
    std::size_t n_ = 0;
    mutable std::string error_;
};

}  // namespace cfdxhip_shim
