// solver_group: several factors' PCG solves run together, one launch per
// kernel kind for all of them (no counterpart in the reference, whose
// cpu_solver / cuda_pcg answer one factor per call). Meant for many small
// factors, none of which fills the GPU alone.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "gpu_solver.hpp"

namespace apxchol_amd {

class solver_group {
public:
    // Members: solvers of the same vector precision and factor storage, each
    // on the hybrid sweep schedule; at most kMaxMembers of them. The group
    // uses the members' device arrays while it solves: a member solves
    // nothing else meanwhile (as a solver never runs two solves at once).
    static constexpr int kMaxMembers = 64;
    explicit solver_group(std::vector<std::shared_ptr<const gpu_solver>> members);
    ~solver_group();
    solver_group(const solver_group&) = delete;
    solver_group& operator=(const solver_group&) = delete;

    // One right-hand side per member (d_b[i], d_x[i]: member i's n doubles,
    // device, its caller order); every member stops as its own
    // solve_device(b, x, tol, max_iter) would, with the same bits.
    std::vector<solve_result> solve(const std::vector<const double*>& d_b, const std::vector<double*>& d_x,
                                    double tol, int max_iter, hipStream_t s) const;
    int size() const { return static_cast<int>(members_.size()); }
    // Launches of one grouped iteration (all members together).
    int launches_per_iteration() const { return launches_per_iteration_; }

private:
    struct impl;
    std::vector<std::shared_ptr<const gpu_solver>> members_;
    std::unique_ptr<impl> p_;
    int launches_per_iteration_ = 0;
};

}  // namespace apxchol_amd
