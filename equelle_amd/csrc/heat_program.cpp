// Heat program in printer style (see heat_program.hpp).
#include "heat_program.hpp"

#include "runtime_core.hpp"

namespace eqhip {

using namespace equelleCUDA;

HeatProgramCUDA::HeatProgramCUDA(EquelleRuntimeCUDA& er_, const std::vector<int>& dir_faces,
                                 const std::vector<double>& dir_vals)
    : er(er_) {
    k = er.inputScalarWithDefault("k", double(2.85));
    cv = er.inputScalarWithDefault("cv", double(2170000));
    quadratic = er.core().params().get_string("k_model", "const") == "quadratic";
    ifaces = er.interiorFaces();
    first = er.firstCell(ifaces);
    second = er.secondCell(ifaces);
    const CollOfScalar geo = (er.norm(ifaces) / er.norm((er.centroid(first) - er.centroid(second))));
    itrans = quadratic ? geo : (k * geo);
    vol = er.norm(er.allCells());
    has_dir = !dir_faces.empty();
    if (has_dir) {
        dir_boundary = CollOfFace(dir_faces);
        er.boundaryFaces().contains(dir_boundary, "dir_boundary");
        dir_val = CollOfScalar(dir_vals);
        const CollOfFace bf = er.boundaryFaces();
        const CollOfCell bf_cells = er.trinaryIf(er.isEmpty(er.firstCell(bf)), er.secondCell(bf), er.firstCell(bf));
        const CollOfScalar bf_sign =
            er.trinaryIf(er.isEmpty(er.firstCell(bf)), er.operatorExtend(-double(1), bf), er.operatorExtend(double(1), bf));
        const CollOfScalar btrans = (k * (er.norm(bf) / er.norm((er.centroid(bf) - er.centroid(bf_cells)))));
        dir_cells = er.operatorOn(bf_cells, er.boundaryFaces(), dir_boundary);
        dir_sign = er.operatorOn(bf_sign, er.boundaryFaces(), dir_boundary);
        dir_trans = er.operatorOn(btrans, er.boundaryFaces(), dir_boundary);
    }
}

CollOfScalar HeatProgramCUDA::residual(const CollOfScalar& u, const CollOfScalar& u0, double dt) const {
    CollOfScalar ifluxes;
    if (quadratic) {
        const CollOfScalar kc = (k * (er.operatorExtend(double(1), er.allCells()) + (u * u)));
        const CollOfScalar kf =
            (double(0.5) * (er.operatorOn(kc, er.allCells(), first) + er.operatorOn(kc, er.allCells(), second)));
        ifluxes = ((-itrans * kf) * er.gradient(u));
    } else {
        ifluxes = (-itrans * er.gradient(u));
    }
    CollOfScalar fluxes = er.operatorExtend(ifluxes, er.interiorFaces(), er.allFaces());
    if (has_dir) {
        const CollOfScalar u_dirbdycells = er.operatorOn(u, er.allCells(), dir_cells);
        const CollOfScalar dir_fluxes = ((dir_trans * dir_sign) * (u_dirbdycells - dir_val));
        const CollOfScalar bfluxes = er.operatorExtend(dir_fluxes, dir_boundary, er.boundaryFaces());
        fluxes = (fluxes + er.operatorExtend(bfluxes, er.boundaryFaces(), er.allFaces()));
    }
    return ((u - u0) + ((dt / (cv * vol)) * er.divergence(fluxes)));
}

}  // namespace eqhip
