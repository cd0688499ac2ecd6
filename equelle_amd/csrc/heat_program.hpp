// The heat program of examples/app/out_heateq.cpp:28-58 written the way the
// Equelle CUDA printer emits it (er.<builtin>(...) calls on EquelleRuntimeCUDA),
// with the nonlinear-conductivity variant used by the nonlinear-heat workload:
//   k_model=const      flux = -(k |f| / |c1 - c2|) grad u
//   k_model=quadratic  k(u) = k (1 + u^2) per cell, face value = mean of the two
//                      cells (On FirstCell / On SecondCell), flux = -(|f|/|c1-c2|) k_f grad u
// Dirichlet faces follow out_heateq.cpp:37-50; without them the boundary is zero-flux.
#pragma once
#include <EquelleRuntimeCUDA.hpp>

namespace eqhip {

struct HeatProgramCUDA {
    HeatProgramCUDA(equelleCUDA::EquelleRuntimeCUDA& er, const std::vector<int>& dir_faces,
                    const std::vector<double>& dir_vals);
    equelleCUDA::CollOfScalar residual(const equelleCUDA::CollOfScalar& u, const equelleCUDA::CollOfScalar& u0,
                                       double dt) const;

    equelleCUDA::EquelleRuntimeCUDA& er;
    double k, cv;
    bool quadratic;
    bool has_dir;
    equelleCUDA::CollOfFace ifaces, dir_boundary;
    equelleCUDA::CollOfCell first, second, dir_cells;
    equelleCUDA::CollOfScalar itrans, vol, dir_val, dir_sign, dir_trans;
};

}  // namespace eqhip
