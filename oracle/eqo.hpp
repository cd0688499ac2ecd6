// eqo — CPU oracle for the Equelle runtime hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a serial C++ restatement of the
// reference serial backend `EquelleRuntimeCPU` (OPM AutoDiffBlock<double> +
// Eigen + OPM GridManager/HelperOps), written without OPM or Eigen.  It is the
// checker the HIP runtime is compared against; only tests/, the smoke check and
// the CPU-baseline leg of bench.py may load it.  Nothing in the product links it.
//
// What it follows (paths relative to the reference tree):
//   grid + numbering     OPM GridManager Cartesian grids, as pinned by
//                        backends/cuda/test/gridTest.cpp:107-171 and
//                        backends/cuda/test/valsOnGrid.cpp:225-350
//   HelperOps            backends/cuda/src/deviceHelperOps.cpp:84-148
//   built-ins            backends/serial/src/EquelleRuntimeCPU.cpp:80-446
//   On/Extend/trinaryIf  backends/serial/include/equelle/EquelleRuntimeCPU_impl.hpp:15-207
//   newtonSolve          backends/serial/include/equelle/EquelleRuntimeCPU_impl.hpp:210-279
//   AD product rules     backends/cuda/src/collOfScalar.cu:273-403 (same rules
//                        as AutoDiffBlock; CSR with explicit zeros kept, columns
//                        ascending, empty Jacobian = constant)
//   CSR algebra          backends/cuda/src/cudaMatrix.cu:640-800 (empty-matrix rules)
//   sqrt                 backends/serial/include/equelle/equelleTypes.hpp:195-211
//
// Parity status of each part is listed in tests/test_oracle_golden.py.
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include <map>
#include <stdexcept>
#include <utility>

namespace eqo {

// ----------------------------------------------------------------------------
// Grid (UnstructuredGrid subset used on the path)
// ----------------------------------------------------------------------------
struct Grid {
    int dim = 2;
    int n[3] = {0, 1, 1};
    double h[3] = {1.0, 1.0, 1.0};
    int nc = 0, nf = 0;
    std::vector<int> face_cells;       // 2*nf, -1 = outside
    std::vector<int> cell_facepos;     // nc+1
    std::vector<int> cell_faces;       // 2*dim*nc, order x-,x+,y-,y+(,z-,z+)
    std::vector<double> cell_centroids;  // dim*nc
    std::vector<double> cell_volumes;    // nc
    std::vector<double> face_centroids;  // dim*nf
    std::vector<double> face_areas;      // nf
    std::vector<double> face_normals;    // dim*nf, area-scaled (UnstructuredGrid convention)
};

Grid make_cartesian(int dim, int nx, int ny, int nz, double dx, double dy, double dz);

// ----------------------------------------------------------------------------
// CSR matrix.  rp.empty() <=> "empty" matrix = zero matrix of whatever size
// (the convention of the reference CudaMatrix and of the zero AD Jacobian).
// ----------------------------------------------------------------------------
struct Csr {
    int rows = 0, cols = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
    bool empty() const { return rp.empty(); }
    int nnz() const { return (int)ci.size(); }
};

struct Triplet { int r, c; double v; };
Csr csr_from_triplets(int rows, int cols, std::vector<Triplet> t);  // sorted, duplicates summed
Csr csr_identity(int n);
Csr csr_diag(const std::vector<double>& d);
Csr csr_transpose(const Csr& a);
// alpha*A + beta*B, union pattern, explicit zeros kept (cusparse csrgeam semantics).
Csr csr_add(const Csr& a, double alpha, const Csr& b, double beta);
// A*B, structural product pattern, explicit zeros kept.  Empty operand -> empty.
Csr csr_mul(const Csr& a, const Csr& b);
// diag(d)*A keeping A's pattern.  Empty -> empty.
Csr csr_diag_scale(const std::vector<double>& d, const Csr& a);
Csr csr_scale(const Csr& a, double s);
std::vector<double> csr_spmv(const Csr& a, const std::vector<double>& x);

// ----------------------------------------------------------------------------
// HelperOps (deviceHelperOps.cpp:84-148)
// ----------------------------------------------------------------------------
struct HelperOps {
    std::vector<int> internal_faces;
    Csr ngrad, grad, div, fullngrad, fulldiv;
};
HelperOps make_helper_ops(const Grid& g);

// ----------------------------------------------------------------------------
// Forward-AD collection of scalars (value + CSR Jacobian wrt the primary var).
// ----------------------------------------------------------------------------
struct ADB {
    std::vector<double> v;
    Csr J;            // empty => constant
    bool ad = false;  // "useAutoDiff" flag of the reference CollOfScalar
    int size() const { return (int)v.size(); }
};

ADB adb_add(const ADB& a, const ADB& b);
ADB adb_sub(const ADB& a, const ADB& b);
ADB adb_mul(const ADB& a, const ADB& b);
ADB adb_div(const ADB& a, const ADB& b);
ADB adb_scal_mul(double s, const ADB& a);
ADB adb_scal_div_coll(double s, const ADB& a);   // s / a
ADB adb_neg(const ADB& a);
ADB adb_norm(const ADB& a);   // norm(CollOfScalar): elementwise absolute value
ADB adb_sqrt(const ADB& a);
ADB adb_matmul(const Csr& m, const ADB& a);      // Matrix * CollOfScalar
ADB adb_gather(const ADB& a, const std::vector<int>& rows);           // On
ADB adb_scatter(const ADB& a, const std::vector<int>& rows, int n);   // Extend
ADB adb_select(const std::vector<uint8_t>& p, const ADB& t, const ADB& f);  // trinaryIf
// Rows [start, start + len) of u, and a placed at rows [start, start + a.size()) of a
// collection of `total` rows (zero rows elsewhere): the subset / superset of
// EquelleRuntimeCPU_impl.hpp:28-110 on a contiguous range (ESpan), as used by
// newtonSolveSystem.
ADB adb_subset(const ADB& u, int start, int len);
ADB adb_superset(const ADB& a, int start, int total);

// ----------------------------------------------------------------------------
// Index sets (CollOfIndices<codim>): explicit ascending-or-not index lists;
// `full` marks AllCells()/AllFaces() like the reference.
// ----------------------------------------------------------------------------
struct IndexSet {
    int codim = 0;
    bool full = false;
    std::vector<int> idx;
    int size() const { return (int)idx.size(); }
};

// Positions of `sub`'s elements inside `super` (subsetIndices, _impl.hpp:28-83).
std::vector<int> subset_indices(const IndexSet& super, const IndexSet& sub);

// ----------------------------------------------------------------------------
// Linear solvers on the host (the role of Opm::LinearSolverFactory).
// ----------------------------------------------------------------------------
struct LinSolveReport { bool converged = false; int iterations = 0; double residual = 0.0; };
// Dense-banded LU with partial pivoting inside the band (exact direct solve).
LinSolveReport solve_banded_direct(const Csr& a, const std::vector<double>& b, std::vector<double>& x);
// BiCGStab with ILU(0), FP64, relative tolerance on ||r||/||b||.
LinSolveReport solve_bicgstab_ilu0(const Csr& a, const std::vector<double>& b, std::vector<double>& x,
                                   double rtol, int maxit);

// ----------------------------------------------------------------------------
// Runtime: the EquelleRuntimeCPU built-ins over the structures above.
// ----------------------------------------------------------------------------
struct Params {
    std::map<std::string, std::string> kv;
    void parse(const std::string& text);  // "k=v k=v" or newline separated
    bool has(const std::string& k) const { return kv.count(k) != 0; }
    double get_double(const std::string& k, double def) const;
    int get_int(const std::string& k, int def) const;
    std::string get_string(const std::string& k, const std::string& def) const;
};

struct NewtonReport {
    int iterations = 0;
    double initial_norm = 0.0, final_norm = 0.0;
    int linear_iterations = 0;
    bool converged = false;
    int stop_reason = 0;  // 1 = abs tol, 2 = rel tol, 0 = max_iter
};

class Runtime {
public:
    explicit Runtime(const Params& p);
    const Grid& grid() const { return grid_; }
    const HelperOps& ops() const { return ops_; }

    IndexSet allCells() const;
    IndexSet boundaryCells() const;
    IndexSet interiorCells() const;
    IndexSet allFaces() const;
    IndexSet boundaryFaces() const;
    IndexSet interiorFaces() const;
    IndexSet firstCell(const IndexSet& faces) const;
    IndexSet secondCell(const IndexSet& faces) const;
    std::vector<uint8_t> isEmpty(const IndexSet& s) const;

    ADB norm(const IndexSet& s) const;                 // volumes / areas
    std::vector<double> centroid(const IndexSet& s) const;  // n*dim, row-major per entity
    std::vector<double> normal(const IndexSet& s) const;    // n*dim unit normals

    ADB gradient(const ADB& x) const;
    ADB negGradient(const ADB& x) const;
    ADB divergence(const ADB& x) const;
    ADB interiorDivergence(const ADB& x) const;

    ADB operatorExtend(const ADB& x, const IndexSet& from, const IndexSet& to) const;
    ADB operatorExtend(double s, const IndexSet& to) const;
    ADB operatorOn(const ADB& x, const IndexSet& from, const IndexSet& to) const;
    IndexSet operatorOn(const IndexSet& x, const IndexSet& from, const IndexSet& to) const;
    ADB trinaryIf(const std::vector<uint8_t>& p, const ADB& t, const ADB& f) const;
    IndexSet trinaryIf(const std::vector<uint8_t>& p, const IndexSet& t, const IndexSet& f) const;

    double minReduce(const ADB& x) const;
    double maxReduce(const ADB& x) const;
    double sumReduce(const ADB& x) const;
    double prodReduce(const ADB& x) const;
    double twoNorm(const ADB& x) const;

    // Solve J du = r (EquelleRuntimeCPU.cpp:413-440).
    // blocks > 1: the unknowns form `blocks` equal blocks (a newtonSolveSystem system)
    std::vector<double> solveForUpdate(const ADB& residual, LinSolveReport* rep, int blocks = 1) const;

    template <class F>
    ADB newtonSolve(const F& rescomp, const ADB& u0, NewtonReport* rep = nullptr, int blocks = 1) const;
    // newtonSolveSystem for two equations (EquelleRuntimeCPU_impl.hpp:282-339): the
    // unknowns are stacked [u0; u1] (ranges [0, n0) and [n0, n0 + n1)), each residual
    // function is called with the two subsets of the combined primary variable, the
    // residuals are placed back by superset and added, and newtonSolve runs on the
    // combined residual.
    template <class F0, class F1>
    std::pair<ADB, ADB> newtonSolveSystem(const F0& f0, const F1& f1, const ADB& u0, const ADB& u1,
                                          NewtonReport* rep = nullptr) const;
    // The combined residual at (u0, u1) (the functor newtonSolveSystem iterates on).
    template <class F0, class F1>
    ADB systemResidual(const F0& f0, const F1& f1, const ADB& u) const;

    int max_iter = 10;
    double abs_res_tol = 1e-6;
    double rel_res_tol = 0.0;   // extension: 0 disables the relative test
    std::string linear_solver = "auto";   // auto | direct | bicgstab
    double solver_tol = 1e-12;
    int solver_max_iter = 5000;

private:
    Grid grid_;
    HelperOps ops_;
};

ADB primary_variable(const ADB& u0);  // identity Jacobian (EquelleRuntimeCPU.cpp:565-571)

template <class F>
ADB Runtime::newtonSolve(const F& rescomp, const ADB& u0, NewtonReport* rep, int blocks) const {
    ADB u = primary_variable(u0);
    ADB residual = rescomp(u);
    double r0 = twoNorm(residual);
    double rn = r0;
    int iter = 0, lin_iters = 0;
    auto done = [&](double r) {
        if (r <= abs_res_tol) return 1;
        if (rel_res_tol > 0.0 && r <= rel_res_tol * r0) return 2;
        return 0;
    };
    while (!done(rn) && iter < max_iter) {
        LinSolveReport lr;
        std::vector<double> du = solveForUpdate(residual, &lr, blocks);
        lin_iters += lr.iterations;
        for (int i = 0; i < u.size(); ++i) u.v[i] -= du[i];   // u = u - du (J unchanged)
        residual = rescomp(u);
        rn = twoNorm(residual);
        ++iter;
    }
    if (rep) {
        rep->iterations = iter;
        rep->initial_norm = r0;
        rep->final_norm = rn;
        rep->linear_iterations = lin_iters;
        rep->stop_reason = done(rn);
        rep->converged = rep->stop_reason != 0;
    }
    ADB out;
    out.v = u.v;
    return out;
}

template <class F0, class F1>
ADB Runtime::systemResidual(const F0& f0, const F1& f1, const ADB& u) const {
    const int total = u.size(), n0 = total / 2;
    const ADB a = adb_subset(u, 0, n0);
    const ADB b = adb_subset(u, n0, total - n0);
    return adb_add(adb_superset(f0(a, b), 0, total), adb_superset(f1(a, b), n0, total));
}

template <class F0, class F1>
std::pair<ADB, ADB> Runtime::newtonSolveSystem(const F0& f0, const F1& f1, const ADB& u0, const ADB& u1,
                                               NewtonReport* rep) const {
    if (u0.size() != u1.size()) throw std::runtime_error("newtonSolveSystem: blocks of different sizes");
    const int n0 = u0.size(), total = n0 + u1.size();
    // combined initial guess: superset(u0) + superset(u1)
    ADB guess = adb_add(adb_superset(u0, 0, total), adb_superset(u1, n0, total));
    ADB u = newtonSolve([&](const ADB& x) { return systemResidual(f0, f1, x); }, guess, rep, 2);
    ADB a, b;
    a.v.assign(u.v.begin(), u.v.begin() + n0);
    b.v.assign(u.v.begin() + n0, u.v.end());
    return {a, b};
}

}  // namespace eqo
