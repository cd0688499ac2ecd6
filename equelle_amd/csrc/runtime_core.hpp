// Runtime core behind EquelleRuntimeCUDA: the device grid, HelperOps matrices,
// grid built-ins and the Newton / Krylov settings read from the parameters.
// Follows backends/cuda/src/EquelleRuntimeCUDA.cpp:29-346 and
// backends/cuda/src/deviceGrid.cu (topology, geometry) for behaviour.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "core.hpp"
#include "engine.hpp"
#include "heat_fused.hpp"
#include "krylov.hpp"

namespace eqhip {

// Opm::ParameterGroup key=value subset.
struct ParamMap {
    std::map<std::string, std::string> kv;
    void parse(const std::string& text);
    bool has(const std::string& k) const { return kv.count(k) != 0; }
    double get_double(const std::string& k, double def) const;
    int get_int(const std::string& k, int def) const;
    bool get_bool(const std::string& k, bool def) const;
    std::string get_string(const std::string& k, const std::string& def) const;
    std::string require(const std::string& k) const;
};

// Cartesian grid with OPM GridManager numbering (host build, device copy).
struct GridHost {
    int dim = 2, n[3] = {0, 1, 1};
    double h[3] = {1, 1, 1};
    int nc = 0, nf = 0;
    std::vector<int> face_cells, cell_facepos, cell_faces;
    std::vector<double> cell_centroids, cell_volumes, face_centroids, face_areas, face_normals;
};
GridHost build_cartesian(int dim, int nx, int ny, int nz, double dx, double dy, double dz);

struct DeviceGrid {
    int dim = 2, nc = 0, nf = 0, nif = 0;
    GridHost host;
    DevPtr<int> face_cells, cell_facepos, cell_faces, signed_faces, if_c1, if_c2;
    DevPtr<double> cell_centroids, cell_volumes, face_centroids, face_areas, face_normals;
    IP all_cells, all_faces, interior_faces, boundary_faces, interior_cells, boundary_cells;
    DevCsr grad, ngrad, div, fulldiv;   // HelperOps (deviceHelperOps.cpp:84-148)
    explicit DeviceGrid(GridHost g);
};

class RuntimeCore {
public:
    explicit RuntimeCore(const ParamMap& p);
    const ParamMap& params() const { return params_; }
    DeviceGrid& grid() { return *grid_; }

    IP firstCell(const IP& faces) const;
    IP secondCell(const IP& faces) const;
    SP norm(const IP& s) const;
    VP centroid(const IP& s) const;
    VP normal(const IP& faces) const;
    SP gradient(const SP& x) const;
    SP gradient_matrix(const SP& x) const;
    SP divergence(const SP& x) const;
    // EquelleRuntimeCPU::negGradient / interiorDivergence (EquelleRuntimeCPU.cpp:345-365):
    // ngrad * x on AllCells() and div * x on InteriorFaces()
    SP neg_gradient(const SP& x) const;
    SP interior_divergence(const SP& x) const;
    SP divergence_matrix(const SP& x) const;
    // tol > 0: this solve's relative tolerance in place of solver_tol (inexact Newton)
    SP solve_for_update(const SP& residual, LinReport* rep, double tol = 0.0) const;
    // The stencil solver's in-place targets for a residual whose Jacobian has pattern P,
    // when its Newton step is solved there (a fused evaluation then writes its Jacobian
    // straight into them; StencilSink in tape.hpp).
    bool stencil_sink(const PatternP& P, struct StencilSink* out) const;
    // Index set [start, start + len) of a combined unknown (newtonSolveSystem ranges);
    // cached, so the position maps of subset / superset are built once.
    IP range_set(int start, int len) const;

    SolverConfig solver;
    bool stencil_solver = true;   // stencil-pattern Jacobians: solve on the stencil kernels
    bool fuse_residual = true;    // newtonSolve residuals: recorded once, then one fused kernel
    // newtonSolve's inexact Newton: the linear solve of an iteration stops at
    // min(eta_max, max(solver_tol, forcing * target / ||F_k||)), target = max(abs_res_tol,
    // rel_res_tol ||F_0||); forcing = 0: solver_tol in every iteration
    double forcing = 0.1;
    double eta_max = 1e-2;
    // the first iteration: max(solver_tol, forcing0 * rho), rho = ||F_1|| / ||F_0|| of the
    // call site's previous solve (0: none; forcing0 = 0: solver_tol)
    double forcing0 = 0.1;
    double linear_tolerance(double rn, double r0) const;
    double first_tolerance(double rho) const;
    int max_iter = 10;
    double abs_res_tol = 1e-6;
    double rel_res_tol = 0.0;
    int verbose = 0;
    bool output_to_file = false;

private:
    bool stencil_pattern(const PatternP& P) const;
    ParamMap params_;
    mutable std::map<uint64_t, bool> stencil_patterns_;   // pattern id -> grid stencil
    mutable std::unique_ptr<HeatBatch> stencil_;         // linear-system instance of the grid
    mutable int stencil_slot_ = 0;
    mutable uint64_t sink_token_ = 0;                    // the last fused evaluation written in place
    // CPR: the multigrid instance of the pressure block and the blocks per Jacobian pattern
    mutable std::unique_ptr<HeatBatch> mg_;
    mutable int mg_slot_ = 0;
    struct PressureBlock {
        PatternP pat, owner;
        DevPtr<int> map;
    };
    mutable std::map<uint64_t, PressureBlock> pblocks_;
    const PressureBlock& pressure_block(const PatternP& P, int np) const;
    bool stencil_solves(const PatternP& P) const;        // the Newton step of pattern P is solved on the stencil kernels
    void ensure_stencil() const;
    std::unique_ptr<DeviceGrid> grid_;
    mutable std::map<std::pair<int, int>, IP> ranges_;
    // firstCell / secondCell results per (face set, side).  Generated code calls them inside
    // the residual function (e.g. FirstCell(InteriorFaces()) in an upwind helper), so a
    // cached result keeps one index set id per Newton iteration and the position maps and
    // row-gather patterns keyed by it are built once.
    mutable std::map<std::pair<uint64_t, int>, std::pair<std::weak_ptr<const IndexData>, IP>> face_cells_;
    IP face_cell(const IP& faces, int which) const;
};

}  // namespace eqhip
