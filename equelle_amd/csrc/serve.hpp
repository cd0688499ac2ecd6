// Request loop over a continuously re-formed batch of PDE solves.
//
// A request is one PDE instance to be advanced by a number of implicit time steps: a 2D
// nonlinear-heat instance (grid, conductivity, initial field) or a two-phase instance
// (grid, permeability, porosity, initial saturation, sources).  Requests wait in a FIFO
// queue until their HBM footprint fits the budget; admitted requests become slots of a
// fused batch, their state stays device-resident between steps, and a progress() call
// advances all in-flight slots of one batch by one Newton iteration in shared launches.
// A slot whose time step converged starts its next step in the same call; a finished
// request's field is copied to the host and its slot (and memory) released, so the
// next queued request can join mid-flight.
//
// Heat and two-phase requests live in two batches: a Newton iteration's Krylov solve runs
// until every member converged, and a 3D pressure solve takes hundreds of Krylov
// iterations where a heat solve takes a few, so one shared batch would hold every heat
// request behind the slowest pressure solve.  progress() ticks the batch that has used
// less device time so far (time-sharing between the two classes).  Per-instance
// reductions are deterministic, so a request's result does not depend on which others
// share the launches.
#pragma once
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "heat_fused.hpp"

namespace eqhip {

struct ServerConfig {
    FusedConfig batch;
    // Krylov precision of the two-phase batch.  FP64 by default: with the multigrid CG
    // pressure solve both precisions converge, and FP64 needs fewer Krylov iterations.
    std::string twophase_precision = "fp64";
    double hbm_budget = 64e9;   // bytes of instance state admitted at once
    int max_in_flight = 4096;   // requests in flight, both kinds
    // two-phase requests in flight at once (the rest wait in their FIFO queue): the batch
    // advances its requests in lock step, so admitting fewer at a time lets the early ones
    // finish early instead of every request sharing the device until the backlog ends
    int max_in_flight_twophase = 4096;
};

struct RequestInfo {
    int steps = 0;               // time steps completed
    int newton = 0, krylov = 0;  // totals over the steps
    bool failed = false;         // a step stopped without converging
    double queued_s = 0;         // submit -> admission
    double latency_s = 0;        // submit -> result available
};

class HeatServer {
public:
    explicit HeatServer(const ServerConfig& cfg);
    // Queues a request; returns its id.  Throws if it can never fit the budget.
    int submit(int nx, int ny, double k, double cv, const double* u0, int steps, double dt);
    int submit_twophase(const TwoPhaseSpec& spec, const double* perm, const double* poro, const double* sw,
                        int nsrc, const int* src_cells, const double* src_vals, int steps, double dt);
    // One scheduling tick: admits what fits, advances every in-flight request by one
    // Newton iteration, retires finished requests.  Returns the ids finished now.
    std::vector<int> progress();
    // Copies a finished request's field (heat: u; two-phase: sw, and p when given; n
    // values each) and releases it.
    void take_result(int id, double* u, double* p, RequestInfo* info);
    // Newton iterations per solve allowed for one request (at most the server's max_iter;
    // 0: the server's limit), queued or in flight.  A step that reaches it without
    // converging fails the request; the other requests are unaffected.
    void set_max_newton(int id, int max_newton);
    int queued() const { return (int)(queue_[0].size() + queue_[1].size()); }
    int in_flight() const { return (int)(running_[0].size() + running_[1].size()); }
    double used_bytes() const { return used_; }
    const ServerConfig& config() const { return cfg_; }

private:
    struct Req {
        int id = 0, nx = 0, ny = 0, slot = -1;
        bool twophase = false;
        TwoPhaseSpec spec;
        double k = 0, cv = 1, dt = 0;
        int steps = 0;
        int max_newton = 0;
        std::vector<double> u;    // initial field until admitted, result after completion
        std::vector<double> perm, poro, src_vals, p;
        std::vector<int> src_cells;
        double bytes = 0;
        std::chrono::steady_clock::time_point t_submit, t_admit, t_done;
        RequestInfo info;
    };
    ServerConfig cfg_;
    HeatBatch batch_[2];                // [0] heat, [1] two-phase
    double busy_s_[2] = {0.0, 0.0};     // device time spent per class
    std::map<int, Req> reqs_;
    std::deque<int> queue_[2];   // FIFO per request kind: heat, two-phase
    std::map<int, int> running_[2];     // slot -> request id, per class
    double used_ = 0;
    int next_id_ = 1;
};

}  // namespace eqhip
