// Test program for recorded residual programs (not an Equelle example): newtonSolve on
// residual functors of several forms, written against the EquelleRuntimeCUDA facade like
// generated code.  Each form's solution goes to form<k>-NNNNN.output; the process prints
// the recorded-program counters (eqhip_residual_stats) after each form.  Run with
// fuse_residual=1 and =0 and compare the outputs.
//   form 0: the nonlinear heat residual with sqrt, norm(CollOfScalar), a trinaryIf on a
//           comparison of the unknown, division, On/Extend of an index subset;
//   form 1: a coefficient from the unknown's values alone (u.value(): no derivative);
//   form 2: a reduction of the unknown inside the residual (MaxReduce: a value the
//           recording cannot hold, so the solve runs op by op).
#include <opm/common/utility/parameters/ParameterGroup.hpp>
#include <opm/common/ErrorMacros.hpp>
#include <opm/grid/UnstructuredGrid.h>
#include <opm/grid/GridManager.hpp>
#include <cmath>
#include <cstdio>
#include <iostream>

#include "EquelleRuntimeCUDA.hpp"
#include "eqhip.h"

static void print_stats(int form) {
    long s[6] = {0, 0, 0, 0, 0, 0};
    double secs = 0.0;
    eqhip_residual_stats(s, &secs);
    std::printf("form %d stats recorded=%ld symbolic=%ld compiled=%ld evaluations=%ld aborted=%ld rejected=%ld\n", form,
                s[0], s[1], s[2], s[3], s[4], s[5]);
    std::fflush(stdout);
}

int main(int argc, char** argv) {
    using namespace equelleCUDA;
    Opm::ParameterGroup param(argc, argv, false);
    EquelleRuntimeCUDA er(param);
    const Scalar k = er.inputScalarWithDefault("k", double(0.1));
    const CollOfFace ifaces = er.interiorFaces();
    const CollOfCell first = er.firstCell(ifaces);
    const CollOfCell second = er.secondCell(ifaces);
    const CollOfScalar geo = er.norm(ifaces) / er.norm(er.centroid(first) - er.centroid(second));
    const CollOfCell bcells = er.boundaryCells();
    const int n = er.allCells().size();
    std::vector<double> init(n);
    for (int c = 0; c < n; ++c) init[c] = 0.5 + 0.4 * std::sin(0.37 * c) + 0.3 * std::cos(0.11 * c);
    const CollOfScalar u_initial(init);
    const double dts[3] = {0.5, 1.0, 1.0};

    // form 0
    {
        CollOfScalar u0 = u_initial;
        for (double dt : dts) {
            auto res = [&](const CollOfScalar& u) -> CollOfScalar {
                const CollOfScalar kc =
                    k * (er.operatorExtend(double(1), er.allCells()) +
                         er.sqrt(u * u + er.operatorExtend(double(0.25), er.allCells())));
                const CollOfScalar kf = 0.5 * (er.operatorOn(kc, er.allCells(), first) +
                                               er.operatorOn(kc, er.allCells(), second));
                const CollOfScalar up = er.trinaryIf(er.operatorOn(u, er.allCells(), first) >
                                                         er.operatorOn(u, er.allCells(), second),
                                                     er.operatorOn(u, er.allCells(), first),
                                                     er.operatorOn(u, er.allCells(), second));
                const CollOfScalar one = er.operatorExtend(double(1), ifaces);
                const CollOfScalar flux =
                    (-geo * kf) * er.gradient(u) * (one + 0.1 * er.norm(up)) / (one + 0.05 * up * up);
                // a sink on the boundary cells, extended back to all cells
                const CollOfScalar sink = er.operatorExtend(0.01 * er.operatorOn(u, er.allCells(), bcells),
                                                            bcells, er.allCells());
                return (u - u0) + dt * (er.divergence(flux) + sink);
            };
            const CollOfScalar u = er.newtonSolve(res, u0);
            er.output("form0", u);
            u0 = u;
        }
        print_stats(0);
    }
    // form 1
    {
        CollOfScalar u0 = u_initial;
        for (double dt : dts) {
            auto res = [&](const CollOfScalar& u) -> CollOfScalar {
                const CollOfScalar frozen = CollOfScalar(u.value());   // no derivative through the coefficient
                const CollOfScalar kc = k * (er.operatorExtend(double(1), er.allCells()) + frozen * frozen);
                const CollOfScalar kf = 0.5 * (er.operatorOn(kc, er.allCells(), first) +
                                               er.operatorOn(kc, er.allCells(), second));
                return (u - u0) + dt * er.divergence((-geo * kf) * er.gradient(u));
            };
            const CollOfScalar u = er.newtonSolve(res, u0);
            er.output("form1", u);
            u0 = u;
        }
        print_stats(1);
    }
    // form 2
    {
        CollOfScalar u0 = u_initial;
        for (double dt : dts) {
            auto res = [&](const CollOfScalar& u) -> CollOfScalar {
                const Scalar umax = er.maxReduce(u);   // a value of the unknown on the host
                const CollOfScalar kc = (k / (1.0 + umax)) * (er.operatorExtend(double(1), er.allCells()) + u * u);
                const CollOfScalar kf = 0.5 * (er.operatorOn(kc, er.allCells(), first) +
                                               er.operatorOn(kc, er.allCells(), second));
                return (u - u0) + dt * er.divergence((-geo * kf) * er.gradient(u));
            };
            const CollOfScalar u = er.newtonSolve(res, u0);
            er.output("form2", u);
            u0 = u;
        }
        print_stats(2);
    }
    return 0;
}
