// Workload programs for the oracle (test infrastructure only; see eqo_programs.hpp).
#include "eqo_programs.hpp"

#include <chrono>
#include <cmath>

namespace eqo {

static ADB vec_norm(const std::vector<double>& a, const std::vector<double>& b, int n, int dim) {
    // er.norm(er.centroid(x) - er.centroid(y)): 2-norm of the difference vectors
    // (EquelleRuntimeCPU.cpp:230-238).
    ADB o;
    o.v.resize(n);
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int d = 0; d < dim; ++d) {
            const double x = a[(size_t)i * dim + d] - b[(size_t)i * dim + d];
            s += x * x;
        }
        o.v[i] = std::sqrt(s);
    }
    return o;
}

HeatProgram::HeatProgram(const Runtime& er_, const Params& p, const std::vector<int>& dir_faces,
                         const std::vector<double>& dir_vals)
    : er(er_) {
    k = p.get_double("k", 2.85);
    cv = p.get_double("cv", 2.17e6);
    quadratic = p.get_string("k_model", "const") == "quadratic";
    const int dim = er.grid().dim;
    ifaces = er.interiorFaces();
    first = er.firstCell(ifaces);
    second = er.secondCell(ifaces);
    all_cells = er.allCells();
    all_faces = er.allFaces();
    // itrans = k * (|ifaces| / |Centroid(first) - Centroid(second)|)
    ADB geo = adb_div(er.norm(ifaces), vec_norm(er.centroid(first), er.centroid(second), ifaces.size(), dim));
    itrans = quadratic ? geo : adb_scal_mul(k, geo);
    vol = er.norm(all_cells);
    bfaces = er.boundaryFaces();
    has_dir = !dir_faces.empty();
    if (has_dir) {
        dir_boundary.codim = 1;
        dir_boundary.idx = dir_faces;
        subset_indices(bfaces, dir_boundary);   // inputDomainSubsetOf's subset check
        dir_val.v = dir_vals;
        IndexSet fb = er.firstCell(bfaces), sb = er.secondCell(bfaces);
        std::vector<uint8_t> emp = er.isEmpty(fb);
        IndexSet bf_cells = er.trinaryIf(emp, sb, fb);
        ADB bf_sign = er.trinaryIf(emp, er.operatorExtend(-1.0, bfaces), er.operatorExtend(1.0, bfaces));
        ADB btrans = adb_scal_mul(k, adb_div(er.norm(bfaces),
                                             vec_norm(er.centroid(bfaces), er.centroid(bf_cells), bfaces.size(), dim)));
        dir_cells = er.operatorOn(bf_cells, bfaces, dir_boundary);
        dir_sign = er.operatorOn(bf_sign, bfaces, dir_boundary);
        dir_trans = er.operatorOn(btrans, bfaces, dir_boundary);
    }
}

ADB HeatProgram::residual(const ADB& u, const ADB& u0, double dt) const {
    // computeInteriorFlux
    ADB ifluxes;
    if (quadratic) {
        ADB kc = adb_scal_mul(k, adb_add(er.operatorExtend(1.0, all_cells), adb_mul(u, u)));
        ADB kf = adb_scal_mul(0.5, adb_add(er.operatorOn(kc, all_cells, first), er.operatorOn(kc, all_cells, second)));
        ifluxes = adb_mul(adb_mul(adb_neg(itrans), kf), er.gradient(u));
    } else {
        ifluxes = adb_mul(adb_neg(itrans), er.gradient(u));
    }
    ADB fluxes = er.operatorExtend(ifluxes, ifaces, all_faces);
    if (has_dir) {
        // computeBoundaryFlux
        ADB u_dirbdycells = er.operatorOn(u, all_cells, dir_cells);
        ADB dir_fluxes = adb_mul(adb_mul(dir_trans, dir_sign), adb_sub(u_dirbdycells, dir_val));
        ADB bfluxes = er.operatorExtend(dir_fluxes, dir_boundary, bfaces);
        fluxes = adb_add(fluxes, er.operatorExtend(bfluxes, bfaces, all_faces));
    }
    // residual = (u - u0) + (dt / (cv * vol)) * Divergence(fluxes)
    ADB coef = adb_scal_div_coll(dt, adb_scal_mul(cv, vol));
    return adb_add(adb_sub(u, u0), adb_mul(coef, er.divergence(fluxes)));
}

HeatResult run_heat(const Params& p, const std::vector<double>& u_initial, const std::vector<int>& dir_faces,
                    const std::vector<double>& dir_vals, const std::vector<double>& dts) {
    Runtime er(p);
    HeatProgram prog(er, p, dir_faces, dir_vals);
    HeatResult res;
    ADB u0;
    u0.v = u_initial;
    for (double dt : dts) {
        NewtonReport rep;
        auto rescomp = [&](const ADB& u) { return prog.residual(u, u0, dt); };
        ADB u = er.newtonSolve(rescomp, u0, &rep);
        res.reports.push_back(rep);
        u0 = u;
    }
    res.u = u0.v;
    return res;
}

// ---------------------------------------------------------------------------
// Two-phase program (twophase.equelle:6-113; generated form out_twophase.cpp).
static ADB vec_dot(const std::vector<double>& a, const std::vector<double>& b, int n, int dim) {
    ADB o;
    o.v.resize(n);
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int d = 0; d < dim; ++d) s += a[(size_t)i * dim + d] * b[(size_t)i * dim + d];
        o.v[i] = s;
    }
    return o;
}
static std::vector<double> vec_sub(const std::vector<double>& a, const std::vector<double>& b) {
    std::vector<double> o(a.size());
    for (size_t i = 0; i < a.size(); ++i) o[i] = a[i] - b[i];
    return o;
}
static std::vector<uint8_t> cmp(const ADB& a, const ADB& b, int op) {   // 0: >=, 1: >, 2: <
    std::vector<uint8_t> o(a.v.size());
    for (size_t i = 0; i < a.v.size(); ++i)
        o[i] = op == 0 ? a.v[i] >= b.v[i] : op == 1 ? a.v[i] > b.v[i] : a.v[i] < b.v[i];
    return o;
}

TwoPhaseProgram::TwoPhaseProgram(const Runtime& er_, const Params& p, const std::vector<double>& perm_in,
                                 const std::vector<double>& poro_v, const std::vector<int>& source_cells,
                                 const std::vector<double>& source_values)
    : er(er_) {
    const int dim = er.grid().dim;
    all_cells = er.allCells();
    if ((int)perm_in.size() != all_cells.size() || (int)poro_v.size() != all_cells.size())
        throw std::runtime_error("perm/poro must have one value per cell");
    watervisc = p.get_double("watervisc", 0.0005);
    oilvisc = p.get_double("oilvisc", 0.005);
    corey = p.get_int("corey", 1);
    if (corey != 1 && corey != 2) throw std::runtime_error("corey must be 1 or 2");
    ADB perm;
    perm.v = perm_in;
    ADB poro_in;
    poro_in.v = poro_v;
    // poro = (poro_in < min_poro) ? min_poro : poro_in ; pv = poro * |AllCells()|
    ADB min_poro = er.operatorExtend(p.get_double("min_poro", 0.0001), all_cells);
    ADB poro = er.trinaryIf(cmp(poro_in, min_poro, 2), min_poro, poro_in);
    pv = adb_mul(poro, er.norm(all_cells));
    // computeTransmissibilities
    ifaces = er.interiorFaces();
    first = er.firstCell(ifaces);
    second = er.secondCell(ifaces);
    const int nif = ifaces.size();
    const std::vector<double> cf = er.centroid(ifaces), nrm = er.normal(ifaces);
    const std::vector<double> cdiff1 = vec_sub(er.centroid(first), cf), cdiff2 = vec_sub(er.centroid(second), cf);
    ADB p1 = er.operatorOn(perm, all_cells, first), p2 = er.operatorOn(perm, all_cells, second);
    ADB a = er.norm(ifaces);
    ADB ht1 = adb_mul(adb_mul(adb_neg(a), p1), adb_div(vec_dot(nrm, cdiff1, nif, dim), vec_dot(cdiff1, cdiff1, nif, dim)));
    ADB ht2 = adb_mul(adb_mul(a, p2), adb_div(vec_dot(nrm, cdiff2, nif, dim), vec_dot(cdiff2, cdiff2, nif, dim)));
    trans = adb_scal_div_coll(1.0, adb_add(adb_scal_div_coll(1.0, ht1), adb_scal_div_coll(1.0, ht2)));
    // source = source_values Extend AllCells() (source_cells: InputDomainSubsetOf, ascending)
    IndexSet sc;
    sc.codim = 0;
    sc.idx = source_cells;
    for (size_t i = 1; i < source_cells.size(); ++i)
        if (source_cells[i] <= source_cells[i - 1]) throw std::runtime_error("source_cells must be ascending");
    subset_indices(all_cells, sc);
    ADB sv;
    sv.v = source_values;
    source = er.operatorExtend(sv, sc, all_cells);
    insource_sw = er.operatorExtend(1.0, all_cells);
}

ADB TwoPhaseProgram::upwind(const ADB& flux, const ADB& x) const {
    ADB x1 = er.operatorOn(x, all_cells, er.firstCell(er.interiorFaces()));
    ADB x2 = er.operatorOn(x, all_cells, er.secondCell(er.interiorFaces()));
    ADB zero = adb_scal_mul(0.0, flux);
    return er.trinaryIf(cmp(flux, zero, 0), x1, x2);
}

ADB TwoPhaseProgram::total_flux(const ADB& pressure, const ADB& total_mobility) const {
    ADB ngradp = adb_neg(er.gradient(pressure));
    ADB face_total_mobility = upwind(ngradp, total_mobility);
    return adb_mul(adb_mul(trans, face_total_mobility), ngradp);
}

ADB TwoPhaseProgram::pressure_residual(const ADB& pressure, const ADB& total_mobility) const {
    return adb_sub(er.divergence(total_flux(pressure, total_mobility)), source);
}

ADB TwoPhaseProgram::water_mob(const ADB& sw) const {
    ADB krw = corey == 2 ? adb_mul(sw, sw) : sw;
    return adb_scal_mul(1.0 / watervisc, krw);   // krw / watervisc
}

ADB TwoPhaseProgram::oil_mob(const ADB& sw) const {
    ADB s = adb_sub(er.operatorExtend(1.0, all_cells), sw);
    ADB kro = corey == 2 ? adb_mul(s, s) : s;
    return adb_scal_mul(1.0 / oilvisc, kro);     // kro / oilvisc
}

ADB TwoPhaseProgram::transport_residual(const ADB& sw, const ADB& sw0, const ADB& flux, double dt) const {
    ADB zero = er.operatorExtend(0.0, all_cells);
    ADB insource = er.trinaryIf(cmp(source, zero, 1), source, zero);
    ADB outsource = er.trinaryIf(cmp(source, zero, 2), source, zero);
    ADB mw = water_mob(sw), mo = oil_mob(sw);
    ADB fracflow = adb_div(mw, adb_add(mw, mo));
    ADB face_fracflow = upwind(flux, fracflow);
    ADB water_flux = adb_mul(face_fracflow, flux);
    ADB q = adb_add(adb_mul(insource, insource_sw), adb_mul(outsource, fracflow));
    ADB coef = adb_scal_div_coll(dt, pv);
    return adb_add(adb_sub(sw, sw0), adb_mul(coef, adb_sub(er.divergence(water_flux), q)));
}

TwoPhaseResult run_twophase(const Params& p, const std::vector<double>& perm, const std::vector<double>& poro,
                            const std::vector<double>& sw_initial, const std::vector<int>& source_cells,
                            const std::vector<double>& source_values, const std::vector<double>& dts) {
    Runtime er(p);
    TwoPhaseProgram prog(er, p, perm, poro, source_cells, source_values);
    TwoPhaseResult res;
    ADB sw0;
    sw0.v = sw_initial;
    ADB p0 = er.operatorExtend(0.0, prog.all_cells);
    for (double dt : dts) {
        // total_mobility = computeWaterMob(sw0) + computeOilMob(sw0)
        ADB total_mobility = adb_add(prog.water_mob(sw0), prog.oil_mob(sw0));
        NewtonReport rp, rs;
        ADB pr = er.newtonSolve([&](const ADB& x) { return prog.pressure_residual(x, total_mobility); }, p0, &rp);
        ADB flux = prog.total_flux(pr, total_mobility);
        ADB sw = er.newtonSolve([&](const ADB& x) { return prog.transport_residual(x, sw0, flux, dt); },
                                er.operatorExtend(0.5, prog.all_cells), &rs);
        res.pressure_reports.push_back(rp);
        res.transport_reports.push_back(rs);
        p0 = pr;
        sw0 = sw;
    }
    res.p = p0.v;
    res.sw = sw0.v;
    return res;
}

ADB twophase_fi_pressure(const TwoPhaseProgram& prog, const ADB& p, const ADB& sw) {
    const ADB total_mobility = adb_add(prog.water_mob(sw), prog.oil_mob(sw));
    return prog.pressure_residual(p, total_mobility);
}

ADB twophase_fi_transport(const TwoPhaseProgram& prog, const ADB& p, const ADB& sw, const ADB& sw0, double dt) {
    const ADB total_mobility = adb_add(prog.water_mob(sw), prog.oil_mob(sw));
    const ADB flux = prog.total_flux(p, total_mobility);
    return prog.transport_residual(sw, sw0, flux, dt);
}

TwoPhaseResult run_twophase_fi(const Params& p, const std::vector<double>& perm, const std::vector<double>& poro,
                               const std::vector<double>& sw_initial, const std::vector<int>& source_cells,
                               const std::vector<double>& source_values, const std::vector<double>& dts) {
    Runtime er(p);
    TwoPhaseProgram prog(er, p, perm, poro, source_cells, source_values);
    TwoPhaseResult res;
    ADB sw0;
    sw0.v = sw_initial;
    ADB p0 = er.operatorExtend(0.0, prog.all_cells);
    for (double dt : dts) {
        NewtonReport rep;
        auto pres = [&](const ADB& pp, const ADB& ss) { return twophase_fi_pressure(prog, pp, ss); };
        auto tres = [&](const ADB& pp, const ADB& ss) { return twophase_fi_transport(prog, pp, ss, sw0, dt); };
        std::pair<ADB, ADB> nv = er.newtonSolveSystem(pres, tres, p0, er.operatorExtend(0.5, prog.all_cells), &rep);
        res.pressure_reports.push_back(rep);
        p0 = nv.first;
        sw0 = nv.second;
    }
    res.p = p0.v;
    res.sw = sw0.v;
    return res;
}

double time_newton_iterations(const Runtime& rt, const HeatProgram& prog, const ADB& u0, double dt, int iters,
                              int* lin_iters) {
    const auto t0 = std::chrono::steady_clock::now();
    ADB u = primary_variable(u0);
    int li = 0;
    for (int it = 0; it < iters; ++it) {
        ADB r = prog.residual(u, u0, dt);
        LinSolveReport rep;
        std::vector<double> du = rt.solveForUpdate(r, &rep);
        li += rep.iterations;
        for (int i = 0; i < u.size(); ++i) u.v[i] -= du[i];
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (lin_iters) *lin_iters = li;
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // namespace eqo
