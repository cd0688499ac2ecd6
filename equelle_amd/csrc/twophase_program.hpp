// The two-phase program of examples/dsl/twophase.equelle (generated form
// examples/app/out_twophase.cpp:26-113) written the way the Equelle CUDA printer
// emits it: er.<builtin>(...) calls on EquelleRuntimeCUDA.  Operator splitting per
// time step: newtonSolve for the pressure (total mobility upwinded on the sign of
// -grad p), then newtonSolve for the water saturation (fractional flow upwinded on the
// total flux).  Relative permeabilities: corey=1 krw = sw, kro = 1 - sw (the program),
// corey=2 krw = sw^2, kro = (1 - sw)^2.
#pragma once
#include <EquelleRuntimeCUDA.hpp>

namespace eqhip {

struct TwoPhaseProgramCUDA {
    TwoPhaseProgramCUDA(equelleCUDA::EquelleRuntimeCUDA& er, const std::vector<double>& perm,
                        const std::vector<double>& poro_in, const std::vector<int>& source_cells,
                        const std::vector<double>& source_values);
    equelleCUDA::CollOfScalar upwind(const equelleCUDA::CollOfScalar& flux, const equelleCUDA::CollOfScalar& x) const;
    equelleCUDA::CollOfScalar totalFlux(const equelleCUDA::CollOfScalar& pressure,
                                        const equelleCUDA::CollOfScalar& total_mobility) const;
    equelleCUDA::CollOfScalar pressureResidual(const equelleCUDA::CollOfScalar& pressure,
                                               const equelleCUDA::CollOfScalar& total_mobility) const;
    equelleCUDA::CollOfScalar waterMob(const equelleCUDA::CollOfScalar& sw) const;
    equelleCUDA::CollOfScalar oilMob(const equelleCUDA::CollOfScalar& sw) const;
    equelleCUDA::CollOfScalar transportResidual(const equelleCUDA::CollOfScalar& sw,
                                                const equelleCUDA::CollOfScalar& sw0,
                                                const equelleCUDA::CollOfScalar& flux, double dt) const;
    // The fully implicit program (examples/dsl/twophase_fully_implicit.equelle:117-149):
    // both residuals as functions of (p, sw), for newtonSolveSystem.
    equelleCUDA::CollOfScalar pressureResLocal(const equelleCUDA::CollOfScalar& p,
                                               const equelleCUDA::CollOfScalar& sw) const;
    equelleCUDA::CollOfScalar transportResLocal(const equelleCUDA::CollOfScalar& p, const equelleCUDA::CollOfScalar& sw,
                                                const equelleCUDA::CollOfScalar& sw0, double dt) const;

    equelleCUDA::EquelleRuntimeCUDA& er;
    double watervisc, oilvisc;
    int corey;
    equelleCUDA::CollOfScalar pv, trans, source, insource_sw;
};

}  // namespace eqhip
