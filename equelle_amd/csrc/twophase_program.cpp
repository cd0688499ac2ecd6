// Two-phase program in printer style (see twophase_program.hpp).
#include "twophase_program.hpp"

#include "runtime_core.hpp"

namespace eqhip {

using namespace equelleCUDA;

TwoPhaseProgramCUDA::TwoPhaseProgramCUDA(EquelleRuntimeCUDA& er_, const std::vector<double>& perm_v,
                                         const std::vector<double>& poro_v, const std::vector<int>& source_cells_v,
                                         const std::vector<double>& source_values_v)
    : er(er_) {
    const CollOfScalar perm = (CollOfScalar(perm_v) * double(1));
    const CollOfScalar poro_in = CollOfScalar(poro_v);
    watervisc = (er.inputScalarWithDefault("watervisc", double(0.0005)) * double(1));
    oilvisc = (er.inputScalarWithDefault("oilvisc", double(0.005)) * double(1));
    corey = er.core().params().get_int("corey", 1);
    if (corey != 1 && corey != 2) OPM_THROW(std::runtime_error, "corey must be 1 or 2");
    if (perm.size() != er.allCells().size() || poro_in.size() != er.allCells().size())
        OPM_THROW(std::runtime_error, "perm and poro need one value per cell");
    const CollOfScalar min_poro = er.operatorExtend(er.inputScalarWithDefault("min_poro", double(0.0001)), er.allCells());
    const CollOfScalar poro = er.trinaryIf((poro_in < min_poro), min_poro, poro_in);
    pv = (poro * er.norm(er.allCells()));
    // computeTransmissibilities(perm)
    const CollOfFace interior_faces = er.interiorFaces();
    const CollOfCell first = er.firstCell(interior_faces);
    const CollOfCell second = er.secondCell(interior_faces);
    const CollOfVector cdiff1 = (er.centroid(first) - er.centroid(interior_faces));
    const CollOfVector cdiff2 = (er.centroid(second) - er.centroid(interior_faces));
    const CollOfScalar p1 = er.operatorOn(perm, er.allCells(), first);
    const CollOfScalar p2 = er.operatorOn(perm, er.allCells(), second);
    const CollOfScalar a = er.norm(interior_faces);
    const CollOfScalar halftrans1 = ((-a * p1) * (er.dot(er.normal(interior_faces), cdiff1) / er.dot(cdiff1, cdiff1)));
    const CollOfScalar halftrans2 = ((a * p2) * (er.dot(er.normal(interior_faces), cdiff2) / er.dot(cdiff2, cdiff2)));
    trans = (double(1) / ((double(1) / halftrans1) + (double(1) / halftrans2)));
    // source = source_values Extend AllCells()
    const CollOfCell source_cells = CollOfCell(source_cells_v);
    er.allCells().contains(source_cells, "source_cells");
    const CollOfScalar source_values = (CollOfScalar(source_values_v) * double(1));
    source = er.operatorExtend(source_values, source_cells, er.allCells());
    insource_sw = er.operatorExtend(double(1), er.allCells());
}

CollOfScalar TwoPhaseProgramCUDA::upwind(const CollOfScalar& flux, const CollOfScalar& x) const {
    const CollOfScalar x1 = er.operatorOn(x, er.allCells(), er.firstCell(er.interiorFaces()));
    const CollOfScalar x2 = er.operatorOn(x, er.allCells(), er.secondCell(er.interiorFaces()));
    const CollOfScalar zero = (double(0) * flux);
    return er.trinaryIf((flux >= zero), x1, x2);
}

CollOfScalar TwoPhaseProgramCUDA::totalFlux(const CollOfScalar& pressure, const CollOfScalar& total_mobility) const {
    const CollOfScalar ngradp = -er.gradient(pressure);
    const CollOfScalar face_total_mobility = upwind(ngradp, total_mobility);
    return ((trans * face_total_mobility) * ngradp);
}

CollOfScalar TwoPhaseProgramCUDA::pressureResidual(const CollOfScalar& pressure,
                                                   const CollOfScalar& total_mobility) const {
    const CollOfScalar flux = totalFlux(pressure, total_mobility);
    return (er.divergence(flux) - source);
}

CollOfScalar TwoPhaseProgramCUDA::waterMob(const CollOfScalar& sw) const {
    const CollOfScalar krw = corey == 2 ? (sw * sw) : sw;
    return (krw * (double(1) / watervisc));
}

CollOfScalar TwoPhaseProgramCUDA::oilMob(const CollOfScalar& sw) const {
    const CollOfScalar so = (er.operatorExtend(double(1), er.allCells()) - sw);
    const CollOfScalar kro = corey == 2 ? (so * so) : so;
    return (kro * (double(1) / oilvisc));
}

CollOfScalar TwoPhaseProgramCUDA::transportResidual(const CollOfScalar& sw, const CollOfScalar& sw0,
                                                    const CollOfScalar& flux, double dt) const {
    const Scalar zero = double(0);
    const CollOfScalar insource = er.trinaryIf((source > zero), source, er.operatorExtend(zero, er.allCells()));
    const CollOfScalar outsource = er.trinaryIf((source < zero), source, er.operatorExtend(zero, er.allCells()));
    const CollOfScalar mw = waterMob(sw);
    const CollOfScalar mo = oilMob(sw);
    const CollOfScalar fracflow = (mw / (mw + mo));
    const CollOfScalar face_fracflow = upwind(flux, fracflow);
    const CollOfScalar water_flux = (face_fracflow * flux);
    const CollOfScalar q = ((insource * insource_sw) + (outsource * fracflow));
    return ((sw - sw0) + ((dt / pv) * (er.divergence(water_flux) - q)));
}

CollOfScalar TwoPhaseProgramCUDA::pressureResLocal(const CollOfScalar& p, const CollOfScalar& sw) const {
    const CollOfScalar total_mobility = (waterMob(sw) + oilMob(sw));
    return pressureResidual(p, total_mobility);
}

CollOfScalar TwoPhaseProgramCUDA::transportResLocal(const CollOfScalar& p, const CollOfScalar& sw,
                                                    const CollOfScalar& sw0, double dt) const {
    const CollOfScalar total_mobility = (waterMob(sw) + oilMob(sw));
    const CollOfScalar flux = totalFlux(p, total_mobility);
    return transportResidual(sw, sw0, flux, dt);
}

}  // namespace eqhip
