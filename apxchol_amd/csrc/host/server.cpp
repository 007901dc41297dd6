// solve_server: resident factors, one blocked session per factor, requests
// admitted into free columns between iteration groups.
#include "server.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace apxchol_amd {

namespace {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("apxchol_amd server: ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

solve_server::solve_server(const server_options& opts) : opts_(opts) {
    check(hipGetDevice(&device_), "hipGetDevice");
    if (opts_.memory_budget == 0) {
        std::size_t fr = 0, tot = 0;
        check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
        opts_.memory_budget = static_cast<std::uint64_t>(static_cast<double>(fr) * 0.9);
    }
    if (opts_.slots < 0 || opts_.slots > 64) throw std::invalid_argument("server: slots must be in [0, 64]");
    opts_.iters_per_check = std::max(1, opts_.iters_per_check);
}

namespace {

// Frees a lane's streams and buffers (after its worker has stopped).
template <class Lane>
void free_lane(Lane& l) {
    if (l.block_stream) (void)hipStreamSynchronize(l.block_stream);
    for (auto& b : l.blocks) {
        if (b.d_b) (void)hipFree(b.d_b);
        if (b.d_x) (void)hipFree(b.d_x);
        b.stream = nullptr;
        b.d_b = b.d_x = nullptr;
    }
    if (l.block_stream) (void)hipStreamDestroy(l.block_stream);
    l.block_stream = nullptr;
    if (l.single_stream) (void)hipStreamSynchronize(l.single_stream);
    if (l.d_sb) (void)hipFree(l.d_sb);
    if (l.d_sx) (void)hipFree(l.d_sx);
    if (l.single_stream) (void)hipStreamDestroy(l.single_stream);
    if (l.h_stage) (void)hipHostFree(l.h_stage);
    l.single_stream = nullptr;
    l.d_sb = l.d_sx = l.h_stage = nullptr;
}

}  // namespace

solve_server::~solve_server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    work_cv_.notify_all();
    for (auto& l : lanes_)
        if (l->worker.joinable()) l->worker.join();
    for (auto& l : lanes_) {
        free_lane(*l);
        if (l->solver) {
            l->solver->set_split_couplings(true);
            l->solver->release_serving();
        }
    }
}

int solve_server::add(std::shared_ptr<gpu_solver> solver) {
    if (!solver) throw std::invalid_argument("server: null solver");
    // One server, one session per solver: a second add (to this or another
    // server) or any other device call while it serves would share the
    // session's block buffers and graphs.
    solver->claim_serving();
    struct release_on_throw {
        const gpu_solver* s;
        ~release_on_throw() {
            if (!s) return;
            s->set_split_couplings(true);
            s->release_serving();
        }
    } guard{solver.get()};
    // No side-stream branches in the lanes' graphs (see set_split_couplings).
    solver->set_split_couplings(false);
    auto pow2 = [](int want) {
        int kp = 1;
        while (kp < want) kp <<= 1;
        return kp;
    };
    // slots = 0: the factor's autotuned block width (32 if not tuned).
    const int kp = pow2(opts_.slots > 0 ? opts_.slots : (solver->preferred_block_width() > 0 ? solver->preferred_block_width() : 32));
    int small_slots;
    {
        std::lock_guard<std::mutex> lk(mu_);
        small_slots = opts_.small_slots;
    }
    // The narrow block only when it is narrower than the main one.
    const int kp_small = small_slots > 0 && pow2(small_slots) < kp ? pow2(small_slots) : 0;
    const std::int64_t n = solver->rows();
    const std::uint64_t nn = static_cast<std::uint64_t>(std::max<std::int64_t>(n, 1));
    // Solver arrays + each block (r, x, p, Ap, y, z and its b / x staging) +
    // the single lane's b / x.
    const std::uint64_t need = solver->stats().device_bytes +
                               static_cast<std::uint64_t>(8) * nn * static_cast<std::uint64_t>(kp + kp_small) * sizeof(double) +
                               static_cast<std::uint64_t>(2) * nn * sizeof(double);
    auto ln = std::make_unique<lane>();
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (used_ + need > opts_.memory_budget)
            throw std::runtime_error("server: factor does not fit the memory budget (" + std::to_string(used_ + need) +
                                     " > " + std::to_string(opts_.memory_budget) + " bytes)");
        used_ += need;
    }
    try {
        ln->solver = solver;
        ln->n = n;
        check(hipSetDevice(device_), "hipSetDevice");
        // Sessions in admission order: the narrow block (solver session 1),
        // then the main one (session 0).
        if (kp_small > 0) ln->blocks.push_back(block_lane{1, kp_small});
        ln->blocks.push_back(block_lane{0, kp});
        // The blocks share one stream (one factor's block work runs in order;
        // every stream beyond the process's few hardware queues shares one
        // with others).
        check(hipStreamCreateWithFlags(&ln->block_stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& b : ln->blocks) {
            b.slot.assign(static_cast<std::size_t>(b.kp), nullptr);
            b.stream = ln->block_stream;
            check(hipMalloc(&b.d_b, nn * b.kp * sizeof(double)), "hipMalloc");
            check(hipMalloc(&b.d_x, nn * b.kp * sizeof(double)), "hipMalloc");
        }
        check(hipStreamCreateWithFlags(&ln->single_stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipMalloc(&ln->d_sb, nn * sizeof(double)), "hipMalloc");
        check(hipMalloc(&ln->d_sx, nn * sizeof(double)), "hipMalloc");
        // Pinned staging of one vector: every host <-> device copy of the lane
        // is a DMA from / to it, never a copy from pageable memory (which the
        // runtime stages internally and may serialize behind other streams).
        check(hipHostMalloc(&ln->h_stage, nn * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        for (auto& b : ln->blocks) ln->solver->session_begin(b.kp, b.stream, b.which);
        ln->solver->single_prepare();
    } catch (...) {
        free_lane(*ln);
        std::lock_guard<std::mutex> lk(mu_);
        used_ -= need;
        throw;
    }
    lane* raw = ln.get();
    int id;
    {
        std::lock_guard<std::mutex> lk(mu_);
        lanes_.push_back(std::move(ln));
        id = static_cast<int>(lanes_.size()) - 1;
    }
    raw->worker = std::thread([this, raw] { run(*raw); });
    guard.s = nullptr;   // the lane owns the claim now; ~solve_server releases it
    return id;
}

std::int64_t solve_server::submit(int factor, const double* b, double tol, int max_iter) {
    auto rq = std::make_shared<request>();
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (factor < 0 || factor >= static_cast<int>(lanes_.size())) throw std::invalid_argument("server: bad factor id");
        // A lane ended by a device failure takes no more requests.
        if (!lanes_[factor]->error.empty())
            throw std::runtime_error("server: factor " + std::to_string(factor) + " failed: " + lanes_[factor]->error);
        const std::int64_t n = lanes_[factor]->n;
        rq->b.assign(b, b + n);
        rq->id = next_id_++;
        rq->factor = factor;
        rq->tol = tol;
        rq->max_iter = max_iter;
        rq->t_submit = std::chrono::steady_clock::now();
        if (!any_submit_) { first_submit_ = rq->t_submit; any_submit_ = true; }
        ++submitted_;
        requests_[rq->id] = rq;
        lanes_[factor]->queue.push_back(rq);
    }
    work_cv_.notify_all();
    return rq->id;
}

bool solve_server::wait(std::int64_t id, double* x, solve_result* res, double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) throw std::invalid_argument("server: unknown request id");
    auto rq = it->second;
    auto ready = [&] { return rq->done; };
    if (timeout_s < 0) done_cv_.wait(lk, ready);
    else if (!done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return false;
    if (rq->res.status == -1) {   // its lane ended on a device failure
        const std::string why = lanes_[rq->factor]->error;
        requests_.erase(id);
        throw std::runtime_error("server: request " + std::to_string(id) + " failed: " + why);
    }
    if (x) std::copy(rq->x.begin(), rq->x.end(), x);
    if (res) *res = rq->res;
    requests_.erase(id);
    return true;
}

void solve_server::set_single_lane(bool on) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        opts_.single_lane = on;
    }
    work_cv_.notify_all();
}

void solve_server::set_small_slots(int slots) {
    if (slots < 0 || slots > 64) throw std::invalid_argument("server: narrow block slots must be in [0, 64]");
    std::lock_guard<std::mutex> lk(mu_);
    opts_.small_slots = slots;
}

server_stats solve_server::stats() const {
    std::lock_guard<std::mutex> lk(mu_);
    server_stats s;
    s.submitted = submitted_;
    s.completed = static_cast<std::int64_t>(latencies_ms_.size());
    s.factors = static_cast<std::int64_t>(lanes_.size());
    s.memory_used = used_;
    s.memory_budget = opts_.memory_budget;
    s.single_lane_completed = single_completed_;
    if (!latencies_ms_.empty()) {
        std::vector<double> v = latencies_ms_;
        std::sort(v.begin(), v.end());
        auto q = [&](double p) {
            const std::size_t k = static_cast<std::size_t>(p * static_cast<double>(v.size() - 1) + 0.5);
            return v[std::min(k, v.size() - 1)];
        };
        s.p50_ms = q(0.50);
        s.p99_ms = q(0.99);
        double sum = 0.0;
        for (double t : v) sum += t;
        s.mean_ms = sum / static_cast<double>(v.size());
        const double wall = std::chrono::duration<double>(last_done_ - first_submit_).count();
        s.solves_per_s = wall > 0.0 ? static_cast<double>(v.size()) / wall : 0.0;
    }
    return s;
}

// Marks requests done (their x and result filled) and wakes the waiters.
void solve_server::finish(lane& ln, const std::vector<std::shared_ptr<request>>& done) {
    std::lock_guard<std::mutex> lk(mu_);
    const auto now = std::chrono::steady_clock::now();
    for (const auto& rq : done) {
        rq->t_done = now;
        rq->done = true;
        const double ms = std::chrono::duration<double, std::milli>(now - rq->t_submit).count();
        rq->res.timings.pcg_ms = ms;
        latencies_ms_.push_back(ms);
        last_done_ = now;
        for (auto& b : ln.blocks)
            for (auto& c : b.slot)
                if (c == rq) c = nullptr;
        if (ln.single == rq) {
            ln.single = nullptr;
            ++single_completed_;
        }
    }
    done_cv_.notify_all();
}

void solve_server::run(lane& ln) {
    (void)hipSetDevice(device_);
    const std::size_t n = static_cast<std::size_t>(ln.n);
    double* stage = static_cast<double*>(ln.h_stage);
    const int J = opts_.iters_per_check;
    // Host <-> device copies of one vector through the pinned staging vector
    // (reused, so each copy completes before the next).
    auto upload = [&](const std::vector<double>& b, double* dst, hipStream_t s) {
        std::copy(b.begin(), b.end(), stage);
        check(hipMemcpyAsync(dst, stage, n * sizeof(double), hipMemcpyHostToDevice, s), "upload b");
        check(hipStreamSynchronize(s), "upload b");
    };
    auto download = [&](const double* src, std::vector<double>& x, hipStream_t s) {
        check(hipMemcpyAsync(stage, src, n * sizeof(double), hipMemcpyDeviceToHost, s), "download x");
        check(hipStreamSynchronize(s), "download x");
        x.assign(stage, stage + n);
    };
    const std::size_t nb = ln.blocks.size();
    for (;;) {
        // Per block: the columns admitted this turn, and whether any is busy.
        std::vector<std::vector<std::pair<int, std::shared_ptr<request>>>> admit(nb);
        std::vector<char> busy(nb, 0);
        std::shared_ptr<request> single_new, single;
        {
            std::unique_lock<std::mutex> lk(mu_);
            work_cv_.wait(lk, [&] {
                if (stop_ || !ln.queue.empty() || ln.single) return true;
                for (auto& b : ln.blocks)
                    for (auto& r : b.slot)
                        if (r) return true;
                return false;
            });
            if (stop_) return;
            // The oldest waiting request takes the single lane when it is
            // free; the rest fill free columns, the narrow block first.
            if (opts_.single_lane && !ln.single && !ln.queue.empty() && ln.n > 0) {
                ln.single = ln.queue.front();
                ln.queue.pop_front();
                single_new = ln.single;
            }
            single = ln.single;
            // The narrow block takes the waiting requests only while the main
            // block is idle and they all fit: a running main block costs the
            // same with more columns, so under load it takes them.
            auto active = [](const block_lane& b) {
                for (auto& r : b.slot)
                    if (r) return true;
                return false;
            };
            auto free_slots = [](const block_lane& b) {
                std::size_t f = 0;
                for (auto& r : b.slot) f += r ? 0 : 1;
                return f;
            };
            const bool use_narrow = nb > 1 && !active(ln.blocks[nb - 1]) && ln.queue.size() <= free_slots(ln.blocks[0]);
            for (std::size_t k = 0; k < nb; ++k) {
                auto& b = ln.blocks[k];
                if (nb > 1 && k == 0 && !use_narrow) {
                    if (active(b)) busy[k] = 1;
                    continue;
                }
                for (int c = 0; c < b.kp && !ln.queue.empty(); ++c)
                    if (!b.slot[c]) {
                        b.slot[c] = ln.queue.front();
                        ln.queue.pop_front();
                        admit[k].emplace_back(c, b.slot[c]);
                    }
                for (auto& r : b.slot)
                    if (r) busy[k] = 1;
            }
        }
        try {
            if (single_new) {
                upload(single_new->b, static_cast<double*>(ln.d_sb), ln.single_stream);
                ln.solver->single_begin(static_cast<const double*>(ln.d_sb), nullptr, single_new->tol,
                                        single_new->max_iter, ln.single_stream);
            }
            for (std::size_t k = 0; k < nb; ++k) {
                if (admit[k].empty()) continue;
                auto& b = ln.blocks[k];
                std::vector<gpu_solver::session_column> cols;
                for (auto& [c, rq] : admit[k]) {
                    double* dst = static_cast<double*>(b.d_b) + static_cast<std::size_t>(c) * n;
                    upload(rq->b, dst, b.stream);
                    cols.push_back({c, dst, rq->tol, rq->max_iter});
                }
                ln.solver->session_admit(cols, b.stream, b.which);
            }
            // One iteration group of each, on their own streams, so the device
            // runs them side by side; then each is read back.
            if (single) ln.solver->single_step(J, ln.single_stream);
            for (std::size_t k = 0; k < nb; ++k)
                if (busy[k]) ln.solver->session_step(J, ln.blocks[k].stream, ln.blocks[k].which);
            std::vector<std::shared_ptr<request>> finished;
            if (single) {
                gpu_solver::session_slot st1;
                if (ln.solver->single_poll(st1, ln.single_stream)) {
                    solve_result res;
                    ln.solver->single_collect(static_cast<double*>(ln.d_sx), res, ln.single_stream);
                    download(static_cast<const double*>(ln.d_sx), single->x, ln.single_stream);
                    single->res.iterations = res.iterations;
                    single->res.residual = res.residual;
                    single->res.status = res.status;
                    single->res.lumped_offdiag = res.lumped_offdiag;
                    finished.push_back(single);
                }
            }
            for (std::size_t k = 0; k < nb; ++k) {
                if (!busy[k]) continue;
                auto& b = ln.blocks[k];
                ln.solver->session_read(b.states, b.stream, b.which);
                std::vector<int> fin;
                std::vector<double*> dst;
                for (int c = 0; c < b.kp; ++c)
                    if (b.slot[c] && b.states[c].done) {
                        fin.push_back(c);
                        dst.push_back(static_cast<double*>(b.d_x) + static_cast<std::size_t>(c) * n);
                    }
                if (fin.empty()) continue;
                ln.solver->session_collect(fin, dst, b.stream, b.which);
                for (std::size_t j = 0; j < fin.size(); ++j) {
                    auto rq = b.slot[fin[j]];
                    download(dst[j], rq->x, b.stream);
                    const auto& st = b.states[fin[j]];
                    rq->res.iterations = st.iterations;
                    rq->res.residual = st.residual;
                    rq->res.status = st.status;
                    rq->res.lumped_offdiag = ln.solver->factor().lumped_offdiag;
                    if (st.status == 5) std::fill(rq->x.begin(), rq->x.end(), 0.0);
                    finished.push_back(rq);
                }
            }
            if (!finished.empty()) finish(ln, finished);
        } catch (const std::exception& e) {
            // A device failure ends this lane: its requests fail (status -1)
            // and later submissions to it are refused with the reason.
            std::lock_guard<std::mutex> lk(mu_);
            ln.error = e.what();
            if (ln.error.empty()) ln.error = "device failure";
            auto fail = [&](std::shared_ptr<request>& rq) {
                if (!rq) return;
                rq->res.status = -1;
                rq->done = true;
                rq = nullptr;
            };
            for (auto& b : ln.blocks)
                for (auto& r : b.slot) fail(r);
            fail(ln.single);
            for (auto& r : ln.queue) fail(r);
            ln.queue.clear();
            done_cv_.notify_all();
            return;
        }
    }
}

}  // namespace apxchol_amd
