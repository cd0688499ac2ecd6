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
