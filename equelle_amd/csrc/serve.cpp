#include "serve.hpp"

#include <utility>

#include "core.hpp"

namespace eqhip {

using Clock = std::chrono::steady_clock;

static double secs(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

static FusedConfig twophase_config(const ServerConfig& cfg) {
    FusedConfig c = cfg.batch;
    c.precision = cfg.twophase_precision;
    return c;
}

HeatServer::HeatServer(const ServerConfig& cfg)
    : cfg_(cfg), batch_{HeatBatch(cfg.batch), HeatBatch(twophase_config(cfg))} {
    EQ_REQUIRE(cfg.hbm_budget > 0, "hbm_budget must be positive");
    EQ_REQUIRE(cfg.max_in_flight > 0, "max_in_flight must be positive");
    EQ_REQUIRE(cfg.max_in_flight_twophase > 0, "max_in_flight_twophase must be positive");
}

int HeatServer::submit(int nx, int ny, double k, double cv, const double* u0, int steps, double dt) {
    EQ_REQUIRE(nx > 0 && ny > 0 && (long)nx * ny < (1L << 31), "bad request grid " << nx << "x" << ny);
    EQ_REQUIRE(steps > 0 && dt > 0, "a request needs steps > 0 and dt > 0");
    const double bytes = (double)HeatBatch::estimate_bytes(HeatBatch::heat_cells(nx, ny, cfg_.batch), cfg_.batch.precision,
                                                             false, cfg_.batch.reduced_form(false));
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
    queue_[0].push_back(id);
    return id;
}

int HeatServer::submit_twophase(const TwoPhaseSpec& spec, const double* perm, const double* poro, const double* sw,
                                int nsrc, const int* src_cells, const double* src_vals, int steps, double dt) {
    const long n = (long)spec.nx * spec.ny * spec.nz;
    EQ_REQUIRE(spec.nx > 0 && spec.ny > 0 && spec.nz > 0 && n < (1L << 31), "bad request grid");
    EQ_REQUIRE(steps > 0 && dt > 0, "a request needs steps > 0 and dt > 0");
    const double bytes = (double)HeatBatch::estimate_bytes((int)n, cfg_.twophase_precision, true,
                                                             twophase_config(cfg_).reduced_form(true));
    EQ_REQUIRE(bytes <= cfg_.hbm_budget, "request of " << bytes << " bytes exceeds the HBM budget of "
                                                       << cfg_.hbm_budget);
    Req r;
    r.id = next_id_++;
    r.twophase = true;
    r.spec = spec;
    r.nx = spec.nx;
    r.ny = spec.ny * spec.nz;   // n = nx * ny for the result copies
    r.dt = dt;
    r.steps = steps;
    r.u.assign(sw, sw + n);
    r.perm.assign(perm, perm + n);
    r.poro.assign(poro, poro + n);
    r.src_cells.assign(src_cells, src_cells + nsrc);
    r.src_vals.assign(src_vals, src_vals + nsrc);
    r.bytes = bytes;
    r.t_submit = Clock::now();
    const int id = r.id;
    reqs_.emplace(id, std::move(r));
    queue_[1].push_back(id);
    return id;
}

std::vector<int> HeatServer::progress() {
    // admission: FIFO per request kind while the head fits (no overtaking within a kind, so
    // large requests cannot starve), the older of the two heads first; two-phase requests
    // also up to their in-flight cap
    for (bool admitted = true; admitted;) {
        admitted = false;
        int order[2] = {0, 1};
        if (!queue_[0].empty() && !queue_[1].empty() && queue_[1].front() < queue_[0].front()) std::swap(order[0], order[1]);
        for (int cls : order) {
            if (queue_[cls].empty() || in_flight() >= cfg_.max_in_flight) continue;
            if (cls == 1 && (int)running_[1].size() >= cfg_.max_in_flight_twophase) continue;
            Req& r = reqs_.at(queue_[cls].front());
            if (used_ + r.bytes > cfg_.hbm_budget) continue;
            if (r.twophase) {
                r.slot = batch_[1].add_twophase(r.spec, r.perm.data(), r.poro.data(), r.u.data(),
                                                (int)r.src_cells.size(), r.src_cells.data(), r.src_vals.data());
                r.perm = std::vector<double>();
                r.poro = std::vector<double>();
            } else {
                r.slot = batch_[0].add(r.nx, r.ny, r.k, r.cv, r.u.data());
            }
            r.u.clear();
            r.u.shrink_to_fit();
            used_ += r.bytes;
            r.t_admit = Clock::now();
            r.info.queued_s = secs(r.t_submit, r.t_admit);
            running_[cls][r.slot] = r.id;
            if (r.max_newton) batch_[cls].set_max_newton(r.slot, r.max_newton);
            batch_[cls].begin_step(r.slot, r.dt);
            queue_[cls].pop_front();
            admitted = true;
            break;   // re-order the heads
        }
    }
    std::vector<int> finished;
    int cls;
    if (running_[0].empty() && running_[1].empty()) return finished;
    if (running_[0].empty()) cls = 1;
    else if (running_[1].empty()) cls = 0;
    else cls = busy_s_[0] <= busy_s_[1] ? 0 : 1;   // the class with less device time so far
    HeatBatch& batch = batch_[cls];
    const auto t0 = Clock::now();
    const std::vector<int> done = batch.iterate();   // ends with a device-to-host copy of the norms
    busy_s_[cls] += secs(t0, Clock::now());
    for (int slot : done) {
        Req& r = reqs_.at(running_[cls].at(slot));
        const StepStats& st = batch.stats(slot);
        r.info.steps += 1;
        r.info.newton += st.newton;
        r.info.krylov += st.krylov;
        const bool bad = st.failed || st.stop == 0;
        r.info.failed |= bad;
        if (!bad && r.info.steps < r.steps) {
            batch.begin_step(slot, r.dt);   // next time step joins the running batch
            continue;
        }
        r.u.resize((size_t)r.nx * r.ny);
        if (r.twophase) {
            r.p.resize(r.u.size());
            batch.get_twophase_state(slot, r.p.data(), r.u.data());
        } else {
            batch.get_state(slot, r.u.data());
        }
        batch.remove(slot);
        running_[cls].erase(slot);
        used_ -= r.bytes;
        r.slot = -1;
        r.t_done = Clock::now();
        r.info.latency_s = secs(r.t_submit, r.t_done);
        finished.push_back(r.id);
    }
    return finished;
}

void HeatServer::set_max_newton(int id, int max_newton) {
    auto it = reqs_.find(id);
    EQ_REQUIRE(it != reqs_.end(), "unknown request " << id);
    EQ_REQUIRE(max_newton >= 0, "max_newton must be >= 0, got " << max_newton);
    Req& r = it->second;
    r.max_newton = max_newton;
    if (r.slot >= 0) batch_[r.twophase ? 1 : 0].set_max_newton(r.slot, max_newton);
}

void HeatServer::take_result(int id, double* u, double* p, RequestInfo* info) {
    auto it = reqs_.find(id);
    EQ_REQUIRE(it != reqs_.end(), "unknown request " << id);
    Req& r = it->second;
    EQ_REQUIRE(r.slot < 0 && !r.u.empty() && r.t_done != Clock::time_point(), "request " << id << " has not finished");
    if (u) std::copy(r.u.begin(), r.u.end(), u);
    if (p && r.twophase) std::copy(r.p.begin(), r.p.end(), p);
    if (info) *info = r.info;
    reqs_.erase(it);
}

}  // namespace eqhip
