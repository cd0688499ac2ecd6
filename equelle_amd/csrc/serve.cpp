#include "serve.hpp"

#include "core.hpp"

namespace eqhip {

using Clock = std::chrono::steady_clock;

static double secs(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

HeatServer::HeatServer(const ServerConfig& cfg) : cfg_(cfg), batch_(cfg.batch) {
    EQ_REQUIRE(cfg.hbm_budget > 0, "hbm_budget must be positive");
    EQ_REQUIRE(cfg.max_in_flight > 0, "max_in_flight must be positive");
}

int HeatServer::submit(int nx, int ny, double k, double cv, const double* u0, int steps, double dt) {
    EQ_REQUIRE(nx > 0 && ny > 0 && (long)nx * ny < (1L << 31), "bad request grid " << nx << "x" << ny);
    EQ_REQUIRE(steps > 0 && dt > 0, "a request needs steps > 0 and dt > 0");
    const double bytes = (double)HeatBatch::estimate_bytes(nx * ny, cfg_.batch.precision);
    EQ_REQUIRE(bytes <= cfg_.hbm_budget, "request of " << bytes << " bytes exceeds the HBM budget of "
                                                       << cfg_.hbm_budget);
    Req r;
    r.id = next_id_++;
    r.nx = nx;
    r.ny = ny;
    r.k = k;
    r.cv = cv;
    r.dt = dt;
    r.steps = steps;
    r.u.assign(u0, u0 + (size_t)nx * ny);
    r.bytes = bytes;
    r.t_submit = Clock::now();
    const int id = r.id;
    reqs_.emplace(id, std::move(r));
    queue_.push_back(id);
    return id;
}

std::vector<int> HeatServer::progress() {
    // admission: FIFO while the head fits (no overtaking, so large requests cannot starve)
    while (!queue_.empty() && (int)running_.size() < cfg_.max_in_flight) {
        Req& r = reqs_.at(queue_.front());
        if (used_ + r.bytes > cfg_.hbm_budget) break;
        r.slot = batch_.add(r.nx, r.ny, r.k, r.cv, r.u.data());
        r.u.clear();
        r.u.shrink_to_fit();
        used_ += r.bytes;
        r.t_admit = Clock::now();
        r.info.queued_s = secs(r.t_submit, r.t_admit);
        running_[r.slot] = r.id;
        batch_.begin_step(r.slot, r.dt);
        queue_.pop_front();
    }
    std::vector<int> finished;
    if (running_.empty()) return finished;
    for (int slot : batch_.iterate()) {
        Req& r = reqs_.at(running_.at(slot));
        const StepStats& st = batch_.stats(slot);
        r.info.steps += 1;
        r.info.newton += st.newton;
        r.info.krylov += st.krylov;
        const bool bad = st.failed || st.stop == 0;
        r.info.failed |= bad;
        if (!bad && r.info.steps < r.steps) {
            batch_.begin_step(slot, r.dt);   // next time step joins the running batch
            continue;
        }
        r.u.resize((size_t)r.nx * r.ny);
        batch_.get_state(slot, r.u.data());
        batch_.remove(slot);
        running_.erase(slot);
        used_ -= r.bytes;
        r.slot = -1;
        r.t_done = Clock::now();
        r.info.latency_s = secs(r.t_submit, r.t_done);
        finished.push_back(r.id);
    }
    return finished;
}

void HeatServer::take_result(int id, double* u, RequestInfo* info) {
    auto it = reqs_.find(id);
    EQ_REQUIRE(it != reqs_.end(), "unknown request " << id);
    Req& r = it->second;
    EQ_REQUIRE(r.slot < 0 && !r.u.empty() && r.t_done != Clock::time_point(), "request " << id << " has not finished");
    if (u) std::copy(r.u.begin(), r.u.end(), u);
    if (info) *info = r.info;
    reqs_.erase(it);
}

}  // namespace eqhip
