// Fused nonlinear-heat engine on structured 2D grids: many independent
// instances of different sizes advanced together (batching, serving) with
// per-instance Newton / Krylov convergence.
//
// The program is the one of heat_program.cpp / tests/generated/nonlinear_heat.equelle
// (zero-flux boundary): F(u) = u - u0 + dt/(cv vol) div(-k_f geo grad u),
// k(u) = k (1 + u^2), k_f = mean of the two cells.  Instead of one AD op per
// built-in, one assembly kernel computes F and the five Jacobian stencil
// coefficients (diagonal, west, east, south, north) per cell; the Jacobian is
// the same 5-point matrix the AD engine produces (boundary coefficients are
// structural zeros that are not stored in its CSR form).
//
// HBM layout per instance (n = nx*ny cells, x fastest): u, u0, F, du and the
// coefficients C, W, E, S, N as separate FP64 arrays (FP32 copies in mixed
// precision), Krylov vectors in the Krylov precision.  Kernels walk a device
// table of instance descriptors plus a block -> (instance, local block) map,
// so a batch is a virtual concatenation whose membership can change every
// Newton iteration without moving any state.  All reductions are per instance
// and deterministic (fixed block partition and order), so a result does not
// depend on which other instances share the launch.
#pragma once
#include <string>
#include <vector>

#include "core.hpp"

namespace eqhip {

struct FusedConfig {
    std::string precision = "mixed";   // fp64 | mixed
    std::string precond = "ilu0";      // jacobi | ilu0 (red-black) | none
    double rel_res_tol = 1e-8;         // Newton: ||F_k|| <= rel ||F_0||
    double abs_res_tol = 0.0;          //         or ||F_k|| <= abs
    int max_iter = 20;                 // Newton iterations per step
    double solver_tol = 1e-6;          // linear: ||b - J du|| <= tol ||b|| (FP64 defect)
    int solver_max_iter = 1000;
    double inner_tol = 1e-6;           // mixed: FP32 inner solve tolerance of heat instances
    double inner_tol_twophase = 1e-4;  //        and of two-phase instances (singular, badly scaled pressure)
    // Inexact Newton: the linear solve of a Newton iteration stops at
    //   eta = min(eta_max, max(solver_tol, forcing * target / ||F_k||)),
    // target = max(abs_res_tol, rel_res_tol ||F_0||): only as far as the next residual needs
    // to go to meet the Newton test (with a margin 1 / forcing), never tighter than
    // solver_tol.  The first iteration of a step takes solver_tol.  forcing = 0: solver_tol
    // in every iteration.  Linear-system instances always take solver_tol.
    double forcing = 0.1;
    double eta_max = 1e-2;
    // the first iteration of a step stops at max(solver_tol, forcing0 * rho), rho =
    // ||F_1|| / ||F_0|| of the latest step (the nonlinear error a first iteration leaves);
    // 0: solver_tol
    double forcing0 = 0.1;
    // Krylov form with preconditioner ilu0, per instance kind: "reduced" (BiCGStab on the
    // red-black Schur complement, preconditioned by its diagonal) or "full" (BiCGStab on
    // the whole system with the two-pass red-black ILU(0)); equal in exact arithmetic.
    // "auto" (two-phase): the full form when the pressure stage runs CG with the multigrid
    // V-cycle; otherwise reduced in FP64, full in mixed precision, where FP32 rounding in
    // the reduced operator costs convergence on the singular pressure system.
    std::string krylov_form_heat = "reduced";
    std::string krylov_form_twophase = "auto";
    // Whether instances of the kind run the reduced form (preconditioner ilu0 only).
    bool reduced_form(bool twophase) const {
        if (precond != "ilu0") return false;
        const std::string& f = twophase ? krylov_form_twophase : krylov_form_heat;
        if (f == "auto" && twophase && twophase_pressure_cg && twophase_pressure_mg) return false;
        return f == "reduced" || (f == "auto" && precision != "mixed");
    }
    // 2D heat instances with nx % 8 == 0 in the reduced form: Krylov kernels give each lane
    // four consecutive pair ranks of one row (16-byte loads) instead of ranks strided by the
    // block size.  Other instances always take the strided form.
    bool pair_vectors = true;
    // Of those, instances with nx % 512 == 0 apply the reduced operator in tiles of four
    // rows x 256 pair ranks per workgroup (red unknowns shared through LDS).
    bool operator_tiles = true;
    // Two-phase instances solve the pressure stage (a symmetric M-matrix) by preconditioned
    // CG instead of BiCGStab: in the reduced form with the Schur complement's diagonal, in
    // the full form with an aggregation multigrid V-cycle (twophase_pressure_mg).
    bool twophase_pressure_cg = true;
    bool twophase_pressure_mg = true;
    // factor of the V-cycle's coarse-grid correction (z += mg_scale P z_c on every level):
    // piecewise-constant prolongation under-corrects smooth errors, and over-correction
    // by 1.8 cuts the CG iterations of the 128 x 128 x 64 pressure stage 2.7-fold (820 ->
    // 307 over ten steps); 0 < mg_scale < 2 keeps the preconditioner symmetric positive
    // definite
    double mg_scale = 1.8;
    // two-phase instances whose pressure stage runs multigrid CG on the full system solve
    // their transport stage by BiCGStab on the red-black reduced system (equal to the full
    // system with red-black ILU(0) in exact arithmetic, fewer bytes per iteration)
    bool twophase_split_form = true;
    int max_refine = 20;
    bool graph = true;                 // capture BiCGStab iterations in a hipGraph
    // with graphs: queue the work after a Newton iteration's Krylov solve behind the predicted
    // iterations without waiting for the convergence count (gated on device; resolved at
    // the iteration's norm check)
    bool speculate = true;
    // A whole speculative Newton iteration as one captured graph (1) or as direct launches
    // (0: on an idle stream the first kernel starts at once, where a graph launch waits).
    bool newton_graph = false;
    int cpt = 0;                       // cells per thread (1, 2, 4, 8); 0 = chosen by the tuner
    // a fixed shape per kernel class (assembly, defect, SpMV, preconditioner, vector
    // updates; 0 = none): set with the XCD order by parameter "shape=a,d,s,p,v,xcd"
    int cpt_fixed[5] = {0, 0, 0, 0, 0};
    int xcd = 1;                       // XCD-contiguous block order when cpt is fixed
    bool tune = true;                  // time the launch shapes per membership signature
};

struct LaunchShape {
    int cpt = 0, xcd = 0;              // cells per thread of the SpMV; XCD block order
    int signatures = 0;                // memberships tuned so far
    int cpt_class[5] = {};             // cells per thread: assembly, defect, SpMV, preconditioner, vector updates
};

struct StepStats {
    int newton = 0;                    // Newton iterations of the step (two-phase: both solves)
    int krylov = 0;
    int refinements = 0;
    int stop = 0;                      // 1 abs, 2 rel, 0 max_iter / failure
    double r0 = 0, rn = 0;
    bool failed = false;               // non-finite residual or Krylov breakdown
    int newton_p = 0, krylov_p = 0;    // two-phase: the pressure solve's share
    int stop_p = 0;                    // two-phase: the pressure solve's stop reason
    int newton_stage = 0;              // iterations of the solve in progress
};

// A two-phase instance (examples/dsl/twophase.equelle with Corey exponent 1 or 2) on a
// Cartesian grid; nz = 1 gives a 2D grid.
struct TwoPhaseSpec {
    int nx = 1, ny = 1, nz = 1;
    double dx = 1, dy = 1, dz = 1;
    double watervisc = 0.0005, oilvisc = 0.005, min_poro = 0.0001;
    int corey = 1;
};

// Result of HeatBatch::solve_linear.
struct LinearSolveStats {
    int iterations = 0;    // Krylov iterations (all refinement rounds)
    int refinements = 0;   // FP64 refinement rounds (mixed precision)
    double relres = 0.0;   // FP64 defect / ||b|| (mixed), Krylov residual / its right-hand side (fp64)
};

class HeatBatch {
public:
    explicit HeatBatch(const FusedConfig& cfg);
    ~HeatBatch();
    HeatBatch(const HeatBatch&) = delete;
    HeatBatch& operator=(const HeatBatch&) = delete;
    // Adds an instance (unit cells unless dx, dy given); returns its slot id.
    int add(int nx, int ny, double k, double cv, const double* u_host, double dx = 1.0, double dy = 1.0);
    // Adds a two-phase instance: per-cell permeability, porosity, initial saturation and
    // rate sources (ascending cells); pressure starts at 0.  Each step is a pressure
    // Newton solve followed by a saturation Newton solve, both on the shared launches.
    int add_twophase(const TwoPhaseSpec& spec, const double* perm, const double* poro, const double* sw,
                     int nsrc, const int* src_cells, const double* src_vals);
    void remove(int slot);
    // A linear-system instance on an nx x ny x nz Cartesian grid (nz = 1: 2D): no physics;
    // solve_linear gathers a CSR Jacobian with the grid's 5- or 7-point stencil pattern
    // (device pointers rp, ci, val) and right-hand side b, solves J x = b with the batch's
    // Krylov method, precision and tolerances (solver_tol, inner_tol, max_refine), and
    // writes x (device memory, cell order).  Every CSR entry must couple a cell with itself
    // or a grid neighbour; the caller checks the pattern.
    int add_linear(int nx, int ny, int nz);
    // A linear-system instance for a symmetric stencil M-matrix (e.g. the pressure block of
    // a two-phase Jacobian, zero row sums allowed) that carries the aggregation multigrid
    // hierarchy: set_linear_matrix gathers its coefficients from CSR and builds the coarse
    // levels; vcycle applies one V-cycle z = M r (cell order, FP64 in and out; the cycle
    // runs in the batch's Krylov precision), a fixed linear operator usable as a
    // preconditioner.
    int add_linear_multigrid(int nx, int ny, int nz);
    void set_linear_matrix(int slot, const int* rp, const int* ci, const double* val);
    void vcycle(int slot, const double* r, double* z);
    // tol > 0: this solve's tolerance in place of solver_tol
    LinearSolveStats solve_linear(int slot, const int* rp, const int* ci, const double* val, const double* b,
                                  double* x, double tol = 0.0);
    // Where a caller writes a linear-system instance's right-hand side and stencil
    // coefficients in place, instead of solve_linear's gather from CSR: F (cell order), the
    // Krylov-precision coefficients K[q] (q = C, W, E, S, N, B, T; red-black rank order: cell
    // c = (i, j, k) at ((i + j + k) & 1) h + c / 2; float when fp32, else double) and, in
    // mixed precision, the FP64 coefficients in cell order (C64[q], else null).  Absent
    // neighbours take coefficient 0.  Valid while the slot lives.
    struct LinearTargets {
        double* F = nullptr;
        void* K[7] = {};
        double* C64[7] = {};
        int nx = 0, ny = 0, nz = 1, h = 0, ncoef = 5;
        bool fp32 = false;
    };
    LinearTargets linear_targets(int slot);
    // solve_linear on coefficients and right-hand side already written at linear_targets
    LinearSolveStats solve_linear_prepared(int slot, double* x, double tol = 0.0);
    // Newton iterations allowed per solve for this slot (at most the batch's max_iter;
    // 0 restores the batch's limit).  A solve that reaches it stops with reason 0.
    void set_max_newton(int slot, int max_newton);
    int size_of(int slot) const;
    // Starts one implicit time step of length dt for the slot.
    void begin_step(int slot, double dt);
    // Advances every in-flight slot by one Newton iteration (continuous batching); in mixed
    // precision a slot whose FP64 refinement needs another round runs that round instead
    // and takes its Newton update in a later call.  Returns the slots whose step finished.
    std::vector<int> iterate();
    bool in_flight(int slot) const;
    const StepStats& stats(int slot) const;
    // Runs whole steps for the given slots (loops iterate()).
    void step(const std::vector<int>& slots, double dt);
    void get_state(int slot, double* u) const;   // two-phase: the saturation
    void get_twophase_state(int slot, double* p, double* sw) const;
    void set_state(int slot, const double* u);
    // Two-phase state (pressure and/or saturation; null leaves a field as it is).
    void set_twophase_state(int slot, const double* p, const double* sw);
    // Test hook: F and the Jacobian in CSR form at the slot's current state (u0 = prev).
    void export_residual(int slot, const double* u, const double* u0, double dt, std::vector<double>& F,
                         std::vector<int>& rp, std::vector<int>& ci, std::vector<double>& v);
    // Test hook: the pressure (stage 0) or transport (stage 1) residual and Jacobian of a
    // two-phase slot at (p, sw0, sw, dt); the slot's state is overwritten.
    void export_twophase_residual(int slot, int stage, const double* p, const double* sw0, const double* sw,
                                  double dt, std::vector<double>& F, std::vector<int>& rp, std::vector<int>& ci,
                                  std::vector<double>& v);
    size_t device_bytes(int slot) const;
    // Average execution time (ms) of one launch of a kernel over every instance, in the
    // membership's tuned shape, with every instance active: each of `reps` launches on the
    // runtime stream is bracketed by its own HIP start and stop events (hipExtLaunchKernel)
    // and waited for.  event_ms gets the average of `reps` further launches queued back to
    // back between one event pair (dispatch gaps included).  Kernels: operator_red,
    // operator_black (reduced-system instances), spmv, ilu_black, ilu_red (full-system
    // instances), assemble, defect (mixed), b1, b2, b3, load_rhs.
    double time_kernel(const std::string& name, int reps, double* event_ms = nullptr);
    // Device bytes of one instance; `reduced`: its Krylov space is the reduced system.
    static size_t estimate_bytes(int n, const std::string& precision, bool twophase = false, bool reduced = true);
    // Cells a heat instance of nx x ny stores (rows padded for the four-ranks-per-lane kernels).
    static int heat_cells(int nx, int ny, const FusedConfig& cfg);
    // Launch shape used by the most recent launch and the number of tuned signatures.
    LaunchShape launch_shape() const;

    struct Impl;

private:
    LinearSolveStats solve_linear_impl(int slot, const int* rp, const int* ci, const double* val, const double* rhs,
                                       double* x, double tol);
    Impl* p_;
};

}  // namespace eqhip
