// solver_group: the PCG iterations of several member factors as grouped
// launches (kernels/group.hip). Each kernel kind of the iteration is one
// launch over every member; a member's part of it is the member's own launch
// (same kernel body, same grid, same partials), so every member gets the bits
// of its own solve. Sweeps are grouped step by step: the members' hybrid
// schedules (launched levels, tail runs) are aligned at their last step, and
// each group step is at most one level launch, one coupling launch and one
// tail launch for all members that have a step there.
#include "group.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

namespace apxchol_amd {

namespace {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("apxchol_amd group: ") + what + ": " + hipGetErrorString(e));
}

class devmem {
public:
    devmem() = default;
    ~devmem() { if (p_) (void)hipFree(p_); }
    devmem(const devmem&) = delete;
    devmem& operator=(const devmem&) = delete;
    template <class T>
    void upload(const std::vector<T>& v) {
        if (p_) { (void)hipFree(p_); p_ = nullptr; }
        if (v.empty()) return;
        check(hipMalloc(&p_, v.size() * sizeof(T)), "hipMalloc");
        check(hipMemcpy(p_, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    }
    void alloc(std::size_t bytes) {
        if (p_) { (void)hipFree(p_); p_ = nullptr; }
        if (bytes) check(hipMalloc(&p_, bytes), "hipMalloc");
    }
    template <class T> T* as() const { return static_cast<T*>(p_); }

private:
    void* p_ = nullptr;
};

// Starts of a grouped launch over entries with these own grids.
std::vector<int> prefix(const std::vector<int>& grids) {
    std::vector<int> s(grids.size() + 1, 0);
    for (std::size_t i = 0; i < grids.size(); ++i) s[i + 1] = s[i] + grids[i];
    return s;
}

}  // namespace

struct solver_group::impl {
    vec_prec prec = vec_prec::fp64;
    bool fp16 = false;
    int count = 0;
    std::vector<group_member_view> views;
    devmem mem;                  // group_member[count]
    devmem states;               // pcg_state[count]
    pcg_state* host_st = nullptr;
    devmem starts;               // every starts array, back to back
    int spmv_s = 0, spmv_blocks = 0, vec_s = 0, vec_blocks = 0;
    // One sweep step: its level units, then its tail units, in units[dir].
    struct step {
        int lv_off = 0, lv_count = 0, lv_s = 0, lv_blocks = 0;
        int tl_off = 0, tl_count = 0, cp_s = 0, cp_blocks = 0;
        std::size_t lds = 0;
    };
    std::vector<step> steps[2];
    devmem units[2];
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~impl() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (host_st) (void)hipHostFree(host_st);
    }

    const int* S(int off) const { return starts.as<int>() + off; }
    const group_member* M() const { return mem.as<group_member>(); }

    void sweep(int dir, hipStream_t s) const {
        const group_unit* u = units[dir].as<group_unit>();
        for (const step& t : steps[dir]) {
            launch_group_sweep_level(prec, fp16, dir == 1, u + t.lv_off, S(t.lv_s), t.lv_count, t.lv_blocks, s);
            launch_group_sweep_couple(prec, fp16, dir == 1, u + t.tl_off, S(t.cp_s), t.tl_count, t.cp_blocks, s);
            launch_group_sweep_tail(prec, u + t.tl_off, t.tl_count, t.lds, s);
        }
    }
    // The preconditioner application of every member, then the p update
    // (gpu_solver::enqueue_precondition).
    void precondition(bool first, hipStream_t s) const {
        if (first) launch_group_fill_outputs(prec, M(), S(vec_s), count, vec_blocks, s);
        sweep(0, s);
        sweep(1, s);
        launch_group_dot_center(prec, M(), S(vec_s), count, vec_blocks, s);
        launch_group_finish(group_fin_op(first ? "dot_first" : "dot"), M(), S(vec_s), count, s);
        launch_group_update_p(prec, M(), S(vec_s), count, vec_blocks, first, s);
    }
    // One PCG iteration of every member (gpu_solver::enqueue_iteration).
    void iteration(hipStream_t s) const {
        launch_group_spmv(M(), S(spmv_s), count, spmv_blocks, s);
        launch_group_finish(group_fin_op("pAp"), M(), S(spmv_s), count, s);
        launch_group_update_xr(prec, M(), S(vec_s), count, vec_blocks, s);
        launch_group_finish(group_fin_op("xr"), M(), S(vec_s), count, s);
        precondition(false, s);
    }
};

solver_group::solver_group(std::vector<std::shared_ptr<const gpu_solver>> members)
    : members_(std::move(members)), p_(std::make_unique<impl>()) {
    auto& P = *p_;
    if (members_.empty()) throw std::invalid_argument("solver_group: no members");
    if (static_cast<int>(members_.size()) > kMaxMembers)
        throw std::invalid_argument("solver_group: at most " + std::to_string(kMaxMembers) + " members");
    P.count = static_cast<int>(members_.size());
    for (const auto& m : members_) {
        if (!m) throw std::invalid_argument("solver_group: null member");
        if (m->rows() == 0) throw std::invalid_argument("solver_group: a member has no rows");
        P.views.push_back(m->group_view());
    }
    P.prec = P.views[0].prec;
    P.fp16 = P.views[0].fp16;
    for (const auto& v : P.views)
        if (v.prec != P.prec || v.fp16 != P.fp16)
            throw std::invalid_argument("solver_group: members must share the vector precision and factor storage");
    check(hipHostMalloc(reinterpret_cast<void**>(&P.host_st), static_cast<std::size_t>(P.count) * sizeof(pcg_state),
                        hipHostMallocDefault), "hipHostMalloc");
    P.states.alloc(static_cast<std::size_t>(P.count) * sizeof(pcg_state));
    pcg_state* st = P.states.as<pcg_state>();

    // Members with their states; the grids of their SpMV and vector passes.
    std::vector<group_member> mem;
    std::vector<int> g_spmv, g_vec;
    for (int i = 0; i < P.count; ++i) {
        group_member g = P.views[i].mem;
        g.st = st + i;
        mem.push_back(g);
        g_spmv.push_back(group_spmv_blocks(g));
        g_vec.push_back(group_vec_blocks(g.n));
    }
    P.mem.upload(mem);
    std::vector<int> all;
    auto add = [&](const std::vector<int>& grids, int& blocks) {
        const int off = static_cast<int>(all.size());
        const auto sv = prefix(grids);
        all.insert(all.end(), sv.begin(), sv.end());
        blocks = sv.back();
        return off;
    };
    P.spmv_s = add(g_spmv, P.spmv_blocks);
    P.vec_s = add(g_vec, P.vec_blocks);
    int launches = 4 + 3;   // spmv, finish, update_xr, finish; dot_center, finish, update_p
    // Sweep steps, aligned at each member's last step.
    for (int dir = 0; dir < 2; ++dir) {
        std::size_t S = 0;
        for (const auto& v : P.views) S = std::max(S, v.steps[dir].size());
        std::vector<group_unit> units;
        for (std::size_t j = 0; j < S; ++j) {
            impl::step t;
            std::vector<group_unit> lv, tl;
            for (int i = 0; i < P.count; ++i) {
                const auto& ms = P.views[i].steps[dir];
                if (j + ms.size() < S) continue;
                group_unit u = ms[j + ms.size() - S];
                u.st = st + i;
                (u.kind == kGroupTail ? tl : lv).push_back(u);
            }
            std::vector<int> g_lv, g_cp;
            for (const auto& u : lv) g_lv.push_back(u.grid);
            for (const auto& u : tl) {
                g_cp.push_back(u.grid);
                t.lds = std::max(t.lds, u.lds);
            }
            t.lv_off = static_cast<int>(units.size());
            t.lv_count = static_cast<int>(lv.size());
            units.insert(units.end(), lv.begin(), lv.end());
            t.tl_off = static_cast<int>(units.size());
            t.tl_count = static_cast<int>(tl.size());
            units.insert(units.end(), tl.begin(), tl.end());
            t.lv_s = add(g_lv, t.lv_blocks);
            t.cp_s = add(g_cp, t.cp_blocks);
            launches += (t.lv_count > 0) + 2 * (t.tl_count > 0);
            P.steps[dir].push_back(t);
        }
        P.units[dir].upload(units);
    }
    P.starts.upload(all);
    launches_per_iteration_ = launches;
}

solver_group::~solver_group() = default;

std::vector<solve_result> solver_group::solve(const std::vector<const double*>& d_b, const std::vector<double*>& d_x,
                                              double tol, int max_iter, hipStream_t s) const {
    auto& P = *p_;
    const int N = P.count;
    if (static_cast<int>(d_b.size()) != N || static_cast<int>(d_x.size()) != N)
        throw std::invalid_argument("solver_group::solve: one right-hand side and one solution per member");
    for (int i = 0; i < N; ++i)
        if (!d_b[i] || !d_x[i]) throw std::invalid_argument("solver_group::solve: null vector");
    if (tol < 0.0) tol = 1e-8;
    if (max_iter < 0) max_iter = 200;
    for (const auto& m : members_) m->group_view();   // refuses serving members
    pcg_state* st = P.states.as<pcg_state>();
    for (int i = 0; i < N; ++i) {
        pcg_state& h = P.host_st[i];
        std::memset(&h, 0, sizeof h);
        h.tol = tol;
        h.max_iter = max_iter;
        h.stag_window = P.views[i].stagnation_window;
        h.center_k = P.views[i].center_k;
        h.laplacian = P.views[i].sddm ? 0 : 1;
        h.n = P.views[i].mem.n;
        h.prev_check = 1.0;
    }
    check(hipMemcpyAsync(st, P.host_st, static_cast<std::size_t>(N) * sizeof(pcg_state), hipMemcpyHostToDevice, s),
          "state upload");
    // Start of every member's solve (gpu_solver::solve_device, x0 = 0).
    for (int i = 0; i < N; ++i) {
        const auto& g = P.views[i].mem;
        launch_scatter_init(g.n, P.views[i].perm, d_b[i], nullptr, g.r, g.x, s);
        launch_init_residual(g.n, g.r, nullptr, st + i, g.part, s);
    }
    P.precondition(true, s);
    if (!P.exec) {
        static std::mutex capture_mu;
        std::lock_guard<std::mutex> lock(capture_mu);
        if (!P.cap) check(hipStreamCreateWithFlags(&P.cap, hipStreamNonBlocking), "hipStreamCreate");
        check(hipStreamBeginCapture(P.cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            P.iteration(P.cap);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(P.cap, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(P.cap, &P.graph), "hipStreamEndCapture");
        check(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    const int J = 2;
    int enqueued = 0;
    for (;;) {
        for (int j = 0; j < J; ++j) check(hipGraphLaunch(P.exec, s), "hipGraphLaunch");
        enqueued += J;
        check(hipMemcpyAsync(P.host_st, st, static_cast<std::size_t>(N) * sizeof(pcg_state), hipMemcpyDeviceToHost, s),
              "state read");
        check(hipStreamSynchronize(s), "solve");
        bool all = true;
        for (int i = 0; i < N; ++i) all = all && P.host_st[i].done;
        if (all || enqueued > max_iter + J) break;
    }
    for (int i = 0; i < N; ++i) {
        const auto& g = P.views[i].mem;
        if (!P.views[i].sddm) launch_sum_vec(g.n, g.x, st + i, g.part, s);
        launch_gather_shift(g.n, P.views[i].perm, g.x, d_x[i], st + i, !P.views[i].sddm, s);
    }
    check(hipMemcpyAsync(P.host_st, st, static_cast<std::size_t>(N) * sizeof(pcg_state), hipMemcpyDeviceToHost, s),
          "state read");
    check(hipStreamSynchronize(s), "solve");
    std::vector<solve_result> res(static_cast<std::size_t>(N));
    for (int i = 0; i < N; ++i) {
        const pcg_state& h = P.host_st[i];
        if (h.status == 5)
            check(hipMemsetAsync(d_x[i], 0, sizeof(double) * static_cast<std::size_t>(P.views[i].mem.n), s), "zero x");
        res[i].iterations = h.iter;
        res[i].residual = h.residual;
        res[i].status = h.status;
        res[i].lumped_offdiag = P.views[i].lumped_offdiag;
    }
    check(hipStreamSynchronize(s), "solve");
    return res;
}

}  // namespace apxchol_amd
