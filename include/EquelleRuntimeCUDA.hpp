// equelleCUDA::EquelleRuntimeCUDA — the runtime class that code emitted by the
// Equelle compiler's CUDA printer (compiler/PrintCUDABackendASTVisitor.cpp:18-59)
// is written against.  Member names and signatures follow
// backends/cuda/include/EquelleRuntimeCUDA.hpp:89-331; the implementation is the
// MI355X runtime in libeqhip.so (HIP kernels on gfx950).
//
// Extensions over the reference (all optional, via parameters):
//   rel_res_tol      relative Newton stop ||F(u_k)|| <= rel_res_tol ||F(u_0)||
//                    (beside the reference's absolute abs_res_tol)
//   precision        fp64 | mixed (FP32 Krylov + FP64 iterative refinement)
//   preconditioner   diagonal | ilu0 | none | cpr (newtonSolveSystem: one multigrid V-cycle
//                    on the pressure block, then ILU(0) of the coupled system)
//   forcing, forcing0, eta_max   inexact Newton: each iteration's linear tolerance
//   fuse_residual    0: evaluate newtonSolve residuals op by op (1: recorded programs)
//   gmres_restart, inner_tol, max_refine
#ifndef EQUELLERUNTIMECUDA_HEADER_INCLUDED
#define EQUELLERUNTIMECUDA_HEADER_INCLUDED

#include <opm/common/ErrorMacros.hpp>
#include <opm/common/utility/parameters/ParameterGroup.hpp>
#include <opm/grid/GridManager.hpp>
#include <opm/grid/UnstructuredGrid.h>

#include <chrono>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "equelle_cuda_types.hpp"

namespace eqhip {
class RuntimeCore;
class ResidualSession;
}

namespace equelleCUDA {

// Statistics of the last newtonSolve (not in the reference; the reference only prints them).
struct NewtonReport {
    int iterations = 0;
    int linear_iterations = 0;
    double initial_norm = 0.0;
    double final_norm = 0.0;
    int stop_reason = 0;   // 1 = abs_res_tol met, 2 = rel_res_tol met, 0 = max_iter reached
    double seconds = 0.0;
};

class EquelleRuntimeCUDA;

// The residual program of one newtonSolve (not in the reference): the first evaluation
// records the operations on the unknown, later evaluations run one fused kernel
// (libeqhip: recorded residual programs).  `site_state` is one int per call site.
class FusedResidual {
public:
    FusedResidual(const EquelleRuntimeCUDA& er, int* site_state, const CollOfScalar& u);
    ~FusedResidual();
    void begin(int pass);
    // true: `out` is the residual of this pass; false: run the next pass
    bool finish(const CollOfScalar& residual, CollOfScalar* out);
    // a recording pass stopped (rethrows exceptions that are not the recording's own)
    void aborted(const std::exception& e);
    bool fused() const;
    CollOfScalar eval(const CollOfScalar& u, bool jacobian);

private:
    std::unique_ptr<eqhip::ResidualSession> s_;
};

class EquelleRuntimeCUDA {
public:
    EquelleRuntimeCUDA(const Opm::ParameterGroup& param);
    ~EquelleRuntimeCUDA();

    // ----------- Topology and geometry
    CollOfCell allCells() const;
    CollOfCell boundaryCells() const;
    CollOfCell interiorCells() const;
    CollOfFace allFaces() const;
    CollOfFace boundaryFaces() const;
    CollOfFace interiorFaces() const;
    CollOfCell firstCell(CollOfFace faces) const;
    CollOfCell secondCell(CollOfFace faces) const;
    template <int codim>
    CollOfScalar norm(const CollOfIndices<codim>& set) const {
        return normOfSet(set);
    }
    CollOfScalar norm(const CollOfVector& vectors) const;
    /// Elementwise absolute value of a collection, derivatives included
    /// (EquelleRuntimeCPU::norm(const CollOfScalar&), EquelleRuntimeCPU.cpp:241-254).
    CollOfScalar norm(const CollOfScalar& scalars) const;
    template <int codim>
    CollOfVector centroid(const CollOfIndices<codim>& set) const {
        return centroidOfSet(set);
    }
    CollOfVector normal(const CollOfFace& faces) const;

    // ----------- Operators and math functions
    CollOfScalar dot(const CollOfVector& v1, const CollOfVector& v2) const;
    CollOfScalar negGradient(const CollOfScalar& cell_scalarfield) const;
    CollOfScalar interiorDivergence(const CollOfScalar& face_fluxes) const;
    CollOfScalar gradient(const CollOfScalar& cell_scalarfield) const;
    CollOfScalar gradient_matrix(const CollOfScalar& cell_scalarfield) const;
    CollOfScalar divergence(const CollOfScalar& fluxes) const;
    CollOfScalar divergence_matrix(const CollOfScalar& fluxes) const;
    template <int codim>
    CollOfBool isEmpty(const CollOfIndices<codim>& set) const {
        return set.isEmpty();
    }

    // ----------- Extend and On
    template <int codim>
    CollOfScalar operatorExtend(const CollOfScalar& data_in, const CollOfIndices<codim>& from_set,
                                const CollOfIndices<codim>& to_set) const {
        return extendImpl(data_in, from_set, to_set);
    }
    template <int codim>
    CollOfScalar operatorExtend(const Scalar& data, const CollOfIndices<codim>& set) {
        return CollOfScalar(set.size(), data);
    }
    template <int codim>
    CollOfScalar operatorOn(const CollOfScalar& data_in, const CollOfIndices<codim>& from_set,
                            const CollOfIndices<codim>& to_set) {
        return onImpl(data_in, from_set, to_set);
    }
    template <int codim_data, int codim_set>
    CollOfIndices<codim_data> operatorOn(const CollOfIndices<codim_data>& in_data,
                                         const CollOfIndices<codim_set>& from_set,
                                         const CollOfIndices<codim_set>& to_set) {
        return CollOfIndices<codim_data>(onIndicesImpl(in_data, from_set, to_set));
    }
    CollOfScalar trinaryIf(const CollOfBool& predicate, const CollOfScalar& iftrue,
                           const CollOfScalar& iffalse) const;
    template <int codim>
    CollOfIndices<codim> trinaryIf(const CollOfBool& predicate, const CollOfIndices<codim>& iftrue,
                                   const CollOfIndices<codim>& iffalse) const {
        return CollOfIndices<codim>(trinaryIndicesImpl(predicate, iftrue, iffalse));
    }

    // ----------- Reductions and special functions
    Scalar minReduce(const CollOfScalar& x) const;
    Scalar maxReduce(const CollOfScalar& x) const;
    Scalar sumReduce(const CollOfScalar& x) const;
    Scalar prodReduce(const CollOfScalar& x) const;
    CollOfScalar sqrt(const CollOfScalar& x) const;

    // ----------- Implicit methods (EquelleRuntimeCUDA_impl.hpp:216-300)
    template <class ResidualFunctor>
    CollOfScalar newtonSolve(const ResidualFunctor& rescomp, const CollOfScalar& u_initialguess);
    // newtonSolveSystem for two equations, as the serial backend has it
    // (backends/serial/include/equelle/EquelleRuntimeCPU_impl.hpp:282-339; commented out in
    // the CUDA backend, EquelleRuntimeCUDA.hpp:252-254): the unknowns are stacked [u0; u1],
    // each residual function is called with the two subsets of the combined primary
    // variable, the residuals are placed back by superset and added, and newtonSolve runs
    // on the combined residual (one Jacobian with the coupling blocks).
    template <class F0, class F1>
    std::tuple<CollOfScalar, CollOfScalar> newtonSolveSystem(const std::tuple<F0, F1>& rescomp,
                                                             const std::tuple<CollOfScalar, CollOfScalar>& u_initialguess);
    // The combined residual of a two-equation system at the stacked unknown u.
    template <class F0, class F1>
    CollOfScalar systemResidual(const std::tuple<F0, F1>& rescomp, const CollOfScalar& u) const;
    // Rows [start, start + len) of u (values and Jacobian rows), and `a` placed at rows
    // [start, start + a.size()) of a collection of `total` rows: the subset / superset of
    // EquelleRuntimeCPU_impl.hpp:85-110 on a contiguous range.
    CollOfScalar subsetRange(const CollOfScalar& u, int start, int len) const;
    CollOfScalar supersetRange(const CollOfScalar& a, int start, int total) const;

    // ----------- Output
    void output(const String& tag, Scalar val) const;
    void output(const String& tag, const CollOfScalar& coll);

    // ----------- Input
    Scalar inputScalarWithDefault(const String& name, const Scalar default_value);
    template <int codim>
    CollOfIndices<codim> inputDomainSubsetOf(const String& name, CollOfIndices<codim> superset) {
        CollOfIndices<codim> subset(readSubsetFile(name));
        superset.contains(subset, name);
        return subset;
    }
    template <int codim>
    CollOfScalar inputCollectionOfScalar(const String& name, const CollOfIndices<codim>& coll) {
        return inputCollectionImpl(name, coll.size());
    }
    SeqOfScalar inputSequenceOfScalar(const String& name);
    void ensureGridDimensionMin(const int minimum_grid_dimension) const;

    // ----------- Test hooks
    UnstructuredGrid getGrid() const;
    CudaMatrix getGradMatrix() const;
    CudaMatrix getDivMatrix() const;
    CudaMatrix getFulldivMatrix() const;
    Scalar twoNormTester(const CollOfScalar& val) const { return twoNorm(val); }

    // ----------- Extensions
    const NewtonReport& lastNewtonReport() const { return last_newton_; }
    // tol > 0: the solve's relative tolerance in place of solver_tol
    CollOfScalar solveForUpdate(const CollOfScalar& residual, int* krylov_iterations = nullptr, double tol = 0.0) const;
    // newtonSolve's inexact-Newton tolerance for iteration k at ||F_k|| = rn (||F_0|| = r0);
    // first_ratio: ||F_1|| / ||F_0|| of the call site's previous solve (0: none)
    double linearTolerance(int k, double rn, double r0, double first_ratio) const;
    eqhip::RuntimeCore& core() const { return *core_; }

private:
    Scalar twoNorm(const CollOfScalar& vals) const;
    CollOfScalar normOfSet(const IndexSetBase& set) const;
    CollOfVector centroidOfSet(const IndexSetBase& set) const;
    CollOfScalar extendImpl(const CollOfScalar& data, const IndexSetBase& from, const IndexSetBase& to) const;
    CollOfScalar onImpl(const CollOfScalar& data, const IndexSetBase& from, const IndexSetBase& to) const;
    std::shared_ptr<const eqhip::IndexData> onIndicesImpl(const IndexSetBase& data, const IndexSetBase& from,
                                                          const IndexSetBase& to) const;
    std::shared_ptr<const eqhip::IndexData> trinaryIndicesImpl(const CollOfBool& p, const IndexSetBase& t,
                                                               const IndexSetBase& f) const;
    std::vector<int> readSubsetFile(const String& name) const;
    CollOfScalar inputCollectionImpl(const String& name, int size) const;
    int newtonStop(double rn, double r0) const;
    void newtonIterationMessage(int iter, double rn) const;
    void newtonDoneMessage(int iter, int stop, double seconds) const;

    std::unique_ptr<eqhip::RuntimeCore> core_;
    const Opm::ParameterGroup& param_;
    std::map<std::string, int> outputcount_;
    NewtonReport last_newton_;
    std::map<const void*, double> first_ratio_;   // per newtonSolve call site
    int max_iter_;
    int verbose_;
};

template <class ResidualFunctor>
CollOfScalar EquelleRuntimeCUDA::newtonSolve(const ResidualFunctor& rescomp, const CollOfScalar& u_initialguess) {
    const auto t0 = std::chrono::steady_clock::now();
    // Primary variable with identity Jacobian; the residual and its Jacobian are
    // assembled by the AD engine on the device.  The first evaluation records the
    // operations on u; later evaluations run them as one fused kernel (one recorded
    // program per call site, i.e. per residual functor type).
    static int site_state = 0;
    // ||F_1|| / ||F_0|| of this call site's previous solve on this runtime (0: none)
    double& site_ratio = first_ratio_[&site_state];
    CollOfScalar u = CollOfScalar(u_initialguess, true);
    FusedResidual fr(*this, &site_state, u);
    CollOfScalar residual;
    for (int pass = 0;; ++pass) {
        fr.begin(pass);
        try {
            if (fr.finish(rescomp(u), &residual)) break;
        } catch (const std::exception& e) {
            fr.aborted(e);
        }
    }
    auto evaluate = [&](const CollOfScalar& x, bool jacobian) {
        if (fr.fused()) return fr.eval(x, jacobian);
        return jacobian ? rescomp(x) : rescomp(CollOfScalar(x.value()));
    };
    const double r0 = twoNorm(residual);
    double rn = r0;
    int iter = 0;
    NewtonReport rep;
    newtonIterationMessage(iter, rn);
    // Where this solve is expected to converge (the previous newtonSolve's iteration count;
    // every iteration of the first solve), the new residual is first evaluated on the
    // values alone (a constant u: no Jacobian is assembled); the value kernels are those of
    // the AD evaluation, so the norm and the stop decision are bit-identical, and the
    // Jacobian is assembled only when another Newton step follows.
    const int expected = last_newton_.iterations;
    while (!newtonStop(rn, r0) && iter < max_iter_) {
        int lin = 0;
        CollOfScalar du = solveForUpdate(residual, &lin, linearTolerance(iter, rn, r0, site_ratio));
        rep.linear_iterations += lin;
        // du is constant: u keeps its identity Jacobian
        u = u - du;
        ++iter;
        if (iter >= expected) {
            const CollOfScalar values = evaluate(u, false);
            rn = twoNorm(values);
            if (iter == 1 && r0 > 0.0) site_ratio = rn / r0;
            newtonIterationMessage(iter, rn);
            if (!newtonStop(rn, r0) && iter < max_iter_) residual = evaluate(u, true);
            continue;
        }
        residual = evaluate(u, true);
        rn = twoNorm(residual);
        if (iter == 1 && r0 > 0.0) site_ratio = rn / r0;
        newtonIterationMessage(iter, rn);
    }
    rep.iterations = iter;
    rep.initial_norm = r0;
    rep.final_norm = rn;
    rep.stop_reason = newtonStop(rn, r0);
    rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    last_newton_ = rep;
    newtonDoneMessage(iter, rep.stop_reason, rep.seconds);
    return CollOfScalar(u.value());
}

template <class F0, class F1>
CollOfScalar EquelleRuntimeCUDA::systemResidual(const std::tuple<F0, F1>& rescomp, const CollOfScalar& u) const {
    const int total = u.size(), n0 = total / 2;
    const CollOfScalar a = subsetRange(u, 0, n0);
    const CollOfScalar b = subsetRange(u, n0, total - n0);
    return supersetRange(std::get<0>(rescomp)(a, b), 0, total) + supersetRange(std::get<1>(rescomp)(a, b), n0, total);
}

template <class F0, class F1>
std::tuple<CollOfScalar, CollOfScalar> EquelleRuntimeCUDA::newtonSolveSystem(
    const std::tuple<F0, F1>& rescomp, const std::tuple<CollOfScalar, CollOfScalar>& u_initialguess) {
    const CollOfScalar& g0 = std::get<0>(u_initialguess);
    const CollOfScalar& g1 = std::get<1>(u_initialguess);
    if (g0.size() != g1.size())
        OPM_THROW(std::runtime_error, "newtonSolveSystem: unknowns of sizes " << g0.size() << " and " << g1.size());
    const int n0 = g0.size(), total = 2 * n0;
    const CollOfScalar guess = supersetRange(g0, 0, total) + supersetRange(g1, n0, total);
    const CollOfScalar u = newtonSolve([&](const CollOfScalar& x) { return systemResidual(rescomp, x); }, guess);
    return std::tuple<CollOfScalar, CollOfScalar>(subsetRange(u, 0, n0), subsetRange(u, n0, n0));
}

}  // namespace equelleCUDA

// Generated code declares `void ensureRequirements(const EquelleRuntimeCUDA& er);`
// before its `using namespace equelleCUDA;` (PrintCUDABackendASTVisitor.cpp:38),
// so the class name is visible at global scope as in the reference.
using equelleCUDA::EquelleRuntimeCUDA;

#endif  // EQUELLERUNTIMECUDA_HEADER_INCLUDED
