// Workload programs for the oracle, written in the style of the code the
// Equelle compiler emits (examples/app/out_heateq.cpp:28-73): the same
// built-in calls in the same order, over the oracle runtime.
// TEST INFRASTRUCTURE ONLY (see eqo.hpp).
#pragma once
#include <vector>

#include "eqo.hpp"

namespace eqo {

// Heat conduction, out_heateq.cpp:28-58, extended with the nonlinear
// conductivity used by the nonlinear-heat workload:
//   k_model=const      flux = -(k |f|/|c1-c2|) grad u          (out_heateq.cpp:33-36)
//   k_model=quadratic  k(u) = k (1 + u^2) on cells, face value = mean of the two
//                      cells (On FirstCell / On SecondCell), flux = -(|f|/|c1-c2|) k_f grad u
// Dirichlet faces (dir_boundary/dir_val) follow out_heateq.cpp:37-50; with no
// Dirichlet faces the boundary is zero-flux.
struct HeatProgram {
    HeatProgram(const Runtime& er, const Params& p, const std::vector<int>& dir_faces,
                const std::vector<double>& dir_vals);
    ADB residual(const ADB& u, const ADB& u0, double dt) const;

    const Runtime& er;
    double k = 2.85, cv = 2.17e6;
    bool quadratic = false;
    IndexSet ifaces, first, second, all_cells, all_faces, bfaces, dir_boundary, dir_cells;
    ADB itrans, vol, dir_val, dir_sign, dir_trans;
    bool has_dir = false;
};

struct HeatResult {
    std::vector<double> u;
    std::vector<NewtonReport> reports;
};
HeatResult run_heat(const Params& p, const std::vector<double>& u_initial, const std::vector<int>& dir_faces,
                    const std::vector<double>& dir_vals, const std::vector<double>& dts);

// Two-phase incompressible flow with operator splitting, examples/dsl/twophase.equelle
// (generated form examples/app/out_twophase.cpp): per time step one newtonSolve for
// the pressure (total-mobility upwinded on the sign of -grad p), then one for the
// water saturation (fractional flow upwinded on the total flux).  Parameters:
//   watervisc (5e-4), oilvisc (5e-3), min_poro (1e-4) as in the program;
//   corey = 1 (krw = sw, kro = 1 - sw, the program) or 2 (krw = sw^2, kro = (1-sw)^2).
// Sources: a subset of cells (ascending) with rates; positive = injection of water.
struct TwoPhaseProgram {
    TwoPhaseProgram(const Runtime& er, const Params& p, const std::vector<double>& perm,
                    const std::vector<double>& poro_in, const std::vector<int>& source_cells,
                    const std::vector<double>& source_values);
    ADB upwind(const ADB& flux, const ADB& x) const;
    ADB total_flux(const ADB& pressure, const ADB& total_mobility) const;
    ADB pressure_residual(const ADB& pressure, const ADB& total_mobility) const;
    ADB water_mob(const ADB& sw) const;
    ADB oil_mob(const ADB& sw) const;
    ADB transport_residual(const ADB& sw, const ADB& sw0, const ADB& flux, double dt) const;

    const Runtime& er;
    double watervisc = 5e-4, oilvisc = 5e-3;
    int corey = 1;
    IndexSet all_cells, ifaces, first, second;
    ADB pv, trans, source, insource_sw;
};

struct TwoPhaseResult {
    std::vector<double> p, sw;
    std::vector<NewtonReport> pressure_reports, transport_reports;
};
TwoPhaseResult run_twophase(const Params& p, const std::vector<double>& perm, const std::vector<double>& poro,
                            const std::vector<double>& sw_initial, const std::vector<int>& source_cells,
                            const std::vector<double>& source_values, const std::vector<double>& dts);

// Fully implicit two-phase flow, examples/dsl/twophase_fully_implicit.equelle (generated
// form examples/app/out_twophase_fully_implicit.cpp:125-141): per time step one
// newtonSolveSystem on (p, sw) with initial guess (p0, 0.5) and the residuals
//   pressureResLocal(p, sw)  = computePressureResidual(p, mw(sw) + mo(sw))       (:122-128)
//   transportResLocal(p, sw) = computeTransportResidual(sw, sw0,
//                                computeTotalFlux(p, mw(sw) + mo(sw)), dt)     (:130-135)
// (the sequential program's building blocks, with the mobility now a function of the
// unknown sw).  Reports: one per step in pressure_reports (the system solve).
ADB twophase_fi_pressure(const TwoPhaseProgram& prog, const ADB& p, const ADB& sw);
ADB twophase_fi_transport(const TwoPhaseProgram& prog, const ADB& p, const ADB& sw, const ADB& sw0, double dt);
TwoPhaseResult run_twophase_fi(const Params& p, const std::vector<double>& perm, const std::vector<double>& poro,
                               const std::vector<double>& sw_initial, const std::vector<int>& source_cells,
                               const std::vector<double>& source_values, const std::vector<double>& dts);

// Wall time of `iters` Newton iterations (assembly + linear solve + update).
double time_newton_iterations(const Runtime& rt, const HeatProgram& prog, const ADB& u0, double dt, int iters,
                              int* lin_iters);

}  // namespace eqo
