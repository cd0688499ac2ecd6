// Newton-Krylov linear solver on the device: J du = r for the Jacobian of an
// AD residual.  Replaces the CUSP dispatch of the reference
// (backends/cuda/src/linearSolver.cu:107-201, options :24-65).
//
//  * Matrix storage for the iterations: SELL-64 (slices of 64 rows = one wave64,
//    entries column-major inside a slice) built once per sparsity pattern.
//  * Krylov methods: BiCGStab (default), CG, GMRES(m) with CGS2 orthogonalisation.
//  * Preconditioners: Jacobi ("diagonal"), ILU(0) in a multicolour ordering
//    (colour classes are independent sets, so each colour is one parallel sweep),
//    none.
//  * precision=fp64: everything FP64.  precision=mixed: the Krylov iteration and
//    the preconditioner run in FP32 on an FP32 copy of J; an FP64 defect
//    r = b - J x drives iterative refinement until the FP64 tolerance is met.
//  * All Krylov scalars stay on the device; the host reads one norm per iteration.
#pragma once
#include <functional>
#include <string>

#include "core.hpp"

namespace eqhip {

struct SolverConfig {
    std::string method = "bicgstab";   // bicgstab | cg | gmres
    std::string precond = "jacobi";    // jacobi (= "diagonal") | ilu0 | none
    std::string precision = "fp64";    // fp64 | mixed
    int max_iter = 1000;               // solver_max_iter
    double tol = 1e-8;                 // solver_tol (relative to ||b||)
    int gmres_restart = 30;
    double inner_tol = 1e-4;           // mixed: relative tolerance of each FP32 solve
    int max_refine = 30;               // mixed: refinement steps
    int verbose = 0;
};

struct LinReport {
    bool converged = false;
    int iterations = 0;     // total Krylov iterations
    int refinements = 0;    // FP64 defect corrections (mixed)
    double residual = 0.0;  // final ||b - A x|| / ||b|| (FP64)
};

// Constrained-pressure-residual (CPR) preconditioning of a stacked two-unknown system
// [p; s] (newtonSolveSystem's Jacobian, the pressure unknowns and equations first): stage
// 1 applies an approximate inverse of the pressure block (the rows and columns of the
// first np unknowns) to the pressure rows of the residual, stage 2 applies ILU(0) of the
// whole system to the residual that leaves; M^-1 r = x1 + ILU(0)^-1 (r - J x1), x1 =
// [M_p^-1 r_p; 0].  A fixed linear operator when stage 1 is (one multigrid V-cycle).
struct CprStage {
    int np = 0;
    std::function<void(const PatternP& P, const double* vals)> setup;   // the whole Jacobian (CSR values)
    std::function<void(const double* r, double* z)> apply;              // z = M_p^-1 r (np each, device)
};

// Solves residual.jacobian * du = residual.value; returns du (no Jacobian).  With cpr
// (FP64 only): the CPR preconditioner with the configured Krylov method.
SP linear_solve(const SolverConfig& cfg, const SP& residual, LinReport* rep, const CprStage* cpr = nullptr);

// SpMV through the solver's SELL-64 storage (test hook / benchmark):
// y = A x for the Jacobian of `a`, in FP64 or FP32 (x, y converted on the fly).
void sell_spmv_test(const SP& a, const double* x_host, double* y_host, bool fp32);

}  // namespace eqhip
