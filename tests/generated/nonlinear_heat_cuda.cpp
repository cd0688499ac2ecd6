
// This program was created by the Equelle compiler from SINTEF.
// (Written by hand in the CUDA printer's conventions for nonlinear_heat.equelle in this
// folder; the compiler is not built here.)

#include <opm/common/utility/parameters/ParameterGroup.hpp>
#include <opm/common/ErrorMacros.hpp>
#include <opm/grid/UnstructuredGrid.h>
#include <opm/grid/GridManager.hpp>
#include <algorithm>
#include <iterator>
#include <iostream>
#include <cmath>
#include <array>

#include "EquelleRuntimeCUDA.hpp"

void ensureRequirements(const EquelleRuntimeCUDA& er);
void equelleGeneratedCode(equelleCUDA::EquelleRuntimeCUDA& er);

#ifndef EQUELLE_NO_MAIN
int main(int argc, char** argv)
{
    // Get user parameters.
    Opm::ParameterGroup param(argc, argv, false);

    // Create the Equelle runtime.
    equelleCUDA::EquelleRuntimeCUDA er(param);
    equelleGeneratedCode(er);
    return 0;
}
#endif // EQUELLE_NO_MAIN

void equelleGeneratedCode(equelleCUDA::EquelleRuntimeCUDA& er) {
    using namespace equelleCUDA;
    ensureRequirements(er);

    // ============= Generated code starts here ================

    const Scalar k = er.inputScalarWithDefault("k", double(0.1));
    const Scalar cv = er.inputScalarWithDefault("cv", double(1));
    const CollOfFace ifaces = er.interiorFaces();
    const CollOfCell first = er.firstCell(ifaces);
    const CollOfCell second = er.secondCell(ifaces);
    const CollOfScalar geo = (er.norm(ifaces) / er.norm((er.centroid(first) - er.centroid(second))));
    const CollOfFace dir_boundary = er.inputDomainSubsetOf("dir_boundary", er.boundaryFaces());
    const CollOfScalar dir_val = er.inputCollectionOfScalar("dir_val", dir_boundary);
    const CollOfFace bf = er.boundaryFaces();
    const CollOfCell bf_cells = er.trinaryIf(er.isEmpty(er.firstCell(bf)), er.secondCell(bf), er.firstCell(bf));
    const CollOfScalar bf_sign = er.trinaryIf(er.isEmpty(er.firstCell(bf)), er.operatorExtend(-double(1), bf), er.operatorExtend(double(1), bf));
    const CollOfScalar btrans = (k * (er.norm(bf) / er.norm((er.centroid(bf) - er.centroid(bf_cells)))));
    const CollOfCell dir_cells = er.operatorOn(bf_cells, er.boundaryFaces(), dir_boundary);
    const CollOfScalar dir_sign = er.operatorOn(bf_sign, er.boundaryFaces(), dir_boundary);
    const CollOfScalar dir_trans = er.operatorOn(btrans, er.boundaryFaces(), dir_boundary);
    auto computeResidual = [&](const CollOfScalar& u, const CollOfScalar& u0, const Scalar& dt) -> CollOfScalar {
        const CollOfScalar kc = (k * (er.operatorExtend(double(1), er.allCells()) + (u * u)));
        const CollOfScalar kf = (double(0.5) * (er.operatorOn(kc, er.allCells(), first) + er.operatorOn(kc, er.allCells(), second)));
        const CollOfScalar ifluxes = ((-geo * kf) * er.gradient(u));
        const CollOfScalar u_dir = er.operatorOn(u, er.allCells(), dir_cells);
        const CollOfScalar dir_fluxes = ((dir_trans * dir_sign) * (u_dir - dir_val));
        const CollOfScalar bfluxes = er.operatorExtend(dir_fluxes, dir_boundary, er.boundaryFaces());
        const CollOfScalar fluxes = (er.operatorExtend(ifluxes, er.interiorFaces(), er.allFaces()) + er.operatorExtend(bfluxes, er.boundaryFaces(), er.allFaces()));
        return ((u - u0) + ((dt / (cv * er.norm(er.allCells()))) * er.divergence(fluxes)));
    };
    const CollOfScalar u_initial = er.inputCollectionOfScalar("u_initial", er.allCells());
    const SeqOfScalar timesteps = er.inputSequenceOfScalar("timesteps");
    auto u0 = u_initial;
    for (const Scalar& dt : timesteps) {
        auto residualLocal = [&](const CollOfScalar& u) -> CollOfScalar {
            return computeResidual(u, u0, dt);
        };
        const CollOfScalar u = er.newtonSolve(residualLocal, u0);
        er.output("u", u);
        er.output("maximum of u", er.maxReduce(u));
        u0 = u;
    }

    // ============= Generated code ends here ================

}

void ensureRequirements(const equelleCUDA::EquelleRuntimeCUDA& er)
{
    (void)er;
}
