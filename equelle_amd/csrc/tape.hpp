// Recorded residual programs: the op sequence of a newtonSolve residual, lowered to one
// fused HIP kernel per call site.
//
// The reference evaluates the residual functor op by op on every Newton iteration
// (backends/cuda/include/EquelleRuntimeCUDA_impl.hpp:216-300, the CollOfScalar operators
// of backends/cuda/src/collOfScalar.cu:273-403): every operation writes its values and
// its CSR Jacobian to device memory.  Here the first evaluation of a newtonSolve records
// the operations that depend on the unknown (a tape); the tape is lowered to a HIP kernel
// (hiprtc, gfx950) in which one thread evaluates one residual row: every intermediate is
// a value plus its derivatives with respect to the unknowns of that row's Jacobian
// pattern (stencil-local forward-mode AD in registers), On / Extend follow the position
// maps, and operator products (gradient, divergence, Matrix * CollOfScalar) loop over the
// operator's CSR row.  The kernel writes the residual and the Jacobian values straight
// into the row's entries of the pattern the recorded evaluation produced (sparsity
// patterns never depend on values, so the pattern is fixed per tape).
//
// * Later newtonSolve calls at the same call site record symbolically: the operations
//   on the unknown create tape nodes without device work, collections that do not
//   depend on the unknown are computed as before and become the kernel's inputs.
// * A kernel is checked once against the recorded evaluation (values and Jacobian) before
//   it replaces it.  Operations outside the recorded set (reductions or host copies of a
//   value that depends on the unknown, user-supplied Jacobians, vectors scaled by it)
//   stop the recording, and that newtonSolve runs op by op.
#pragma once
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "core.hpp"
#include "engine.hpp"

namespace eqhip {

// Thrown by a symbolic recording pass when the residual reads a value it does not hold;
// newtonSolve catches it and evaluates the residual op by op.
struct TapeAbort : std::runtime_error {
    TapeAbort() : std::runtime_error("residual recording: operation outside the recorded set") {}
};

enum TapeOp : uint8_t {
    T_PRIM,    // the unknown (identity Jacobian)
    T_LEAF,    // a collection that does not depend on the unknown
    T_BLEAF,   // a CollOfBool that does not depend on the unknown
    T_ADD,
    T_SUB,
    T_MUL,
    T_DIV,
    T_SMUL,    // constant * a
    T_SDIV,    // constant / a
    T_SQRT,
    T_NORM,    // |a| with the sign-factor Jacobian
    T_VAL,     // the values of a, no Jacobian
    T_CMP,     // a op b -> CollOfBool
    T_CMPS,    // a op constant, or constant op a -> CollOfBool
    T_SEL,     // p ? t : f
    T_ON,      // row gather through a position map
    T_EXT,     // row gather through an inverse position map (-1: zero row)
    T_MAT,     // constant CSR matrix * a
};

struct TapeNode {
    uint8_t op = 0;
    int8_t sub = 0;       // comparison code (T_CMPS: +16 when the constant is on the left);
                          // T_EXT: 1 when the operand is empty (the result is zero);
                          // T_LEAF: 1 for a uniform collection (res: its constant slot)
    int a = -1, b = -1, c = -1;
    int res = -1;         // leaf, map, matrix or constant slot
    bool ad = false;      // carries derivatives
    bool boolean = false;
};

class Tape;

// The active recording of this thread (null when none, or while suspended).
Tape* tape_current();
// Suspends recording for the lifetime of the object (eager evaluations inside a hook).
struct TapeSuspend {
    TapeSuspend();
    ~TapeSuspend();
    TapeSuspend(const TapeSuspend&) = delete;
    TapeSuspend& operator=(const TapeSuspend&) = delete;
};

class Tape {
public:
    Tape(uint64_t session, bool symbolic);
    bool symbolic() const { return symbolic_; }
    bool valid() const { return valid_; }
    uint64_t session() const { return session_; }

    // node of a value in this recording, -1 if it is not one
    int node_of(const ScalarData& s) const { return s.tsess == session_ ? s.tnode : -1; }
    int node_of(const BoolData& b) const { return b.tsess == session_ ? b.tnode : -1; }
    // node of an operand: its own, or a new leaf for a value that does not depend on the unknown
    int operand(const SP& s);
    int operand(const BP& b);
    int add(const TapeNode& nd);
    // an operation node; derivative flow and kind follow from the operation and operands
    int node(uint8_t op, int a, int b = -1, int c = -1, int res = -1, int sub = 0);
    int constant(double v);
    int map(const DevPtr<int>& m, uint64_t id, std::shared_ptr<const void> keep);
    int matrix(const DevCsr& m);
    void tag(const ScalarData& s, int id);
    void tag(const BoolData& b, int id);
    SP symbolic_scalar(int id, int n, bool ad);
    BP symbolic_bool(int id, int n);
    void primary(const SP& u);   // the unknown (node 0)
    // a read the tape cannot record: a symbolic pass throws TapeAbort, a recording
    // evaluation is marked not replayable
    void unsupported(const char* what);

    // recorded data (read by the lowering)
    std::vector<TapeNode> nodes;
    std::vector<SP> leaves;
    std::vector<BP> bleaves;
    std::vector<std::pair<DevPtr<int>, uint64_t>> maps;
    std::vector<DevCsr> mats;
    // per matrix: longest row, and (for a matrix of +-1 values) each entry's column with
    // its sign folded in (~column where the value is -1)
    std::vector<int> mat_maxrow;
    std::vector<char> mat_regular;   // every row has mat_maxrow entries (rows need no row pointers)
    std::vector<DevPtr<int>> mat_sci;
    std::vector<double> consts;
    std::vector<std::shared_ptr<const void>> keep;
    int n_primary = 0;
    std::string why;   // first unsupported operation

private:
    uint64_t session_;
    bool symbolic_;
    bool valid_ = true;
};

// Recording hooks for the engine: `operands` traced in the active tape?
inline Tape* tape_for(const ScalarData& a) {
    Tape* t = tape_current();
    return t && t->node_of(a) >= 0 ? t : nullptr;
}
inline Tape* tape_for(const ScalarData& a, const ScalarData& b) {
    Tape* t = tape_current();
    return t && (t->node_of(a) >= 0 || t->node_of(b) >= 0) ? t : nullptr;
}
// A read of the values (or Jacobian) of `a` outside the recorded operations.
inline void tape_read(const ScalarData& a, const char* what) {
    if (Tape* t = tape_for(a)) t->unsupported(what);
}
inline void tape_read(const BoolData& b, const char* what) {
    Tape* t = tape_current();
    if (t && t->node_of(b) >= 0) t->unsupported(what);
}

// Recording hook of an operation returning a CollOfScalar: record node ND (an int
// expression) and return a symbolic result, or evaluate EVAL with recording suspended and
// tag its result.
#define EQ_TAPE_RETURN(TP, ND, N, AD, EVAL)                                     \
    do {                                                                        \
        const int id_ = (ND);                                                   \
        if ((TP)->symbolic()) return (TP)->symbolic_scalar(id_, (N), (AD));     \
        ::eqhip::SP r_;                                                         \
        {                                                                       \
            ::eqhip::TapeSuspend ts_;                                           \
            r_ = (EVAL);                                                        \
        }                                                                       \
        (TP)->tag(*r_, id_);                                                    \
        return r_;                                                              \
    } while (0)

struct ResidualStats {
    long recorded = 0;       // recording evaluations
    long symbolic = 0;       // symbolic recording passes that found their kernel
    long compiled = 0;       // kernels built (hiprtc)
    long evaluations = 0;    // fused evaluations
    long aborted = 0;        // recordings stopped by an unsupported operation
    long rejected = 0;       // kernels that disagreed with the recorded evaluation
    double compile_seconds = 0.0;
};
ResidualStats residual_stats();

// Where a fused evaluation may write its Jacobian instead of CSR values: the stencil linear
// solver's right-hand side and coefficients in place (HeatBatch::linear_targets), for a
// residual whose Jacobian has the grid's 5- or 7-point stencil pattern.
struct StencilSink {
    double* F = nullptr;
    void* K[7] = {};
    double* C64[7] = {};
    int nx = 0, ny = 0, nz = 1, h = 0, ncoef = 5;
    bool fp32 = false;
    uint64_t token = 0;   // names the evaluation for the solve that reads it
};
// true (with the targets) when the Newton step of a residual with pattern P is solved there
using SinkFn = std::function<bool(const PatternP& P, StencilSink* out)>;

// One newtonSolve's residual program.
class ResidualSession {
public:
    // `site_state`: one int per newtonSolve call site (0: record on the next call; 1: a
    // kernel exists, record symbolically; -1: op by op)
    ResidualSession(int* site_state, const SP& u, bool enabled);
    ~ResidualSession();
    // Starts a recording pass (pass 0: symbolic where the call site has a kernel; pass 1:
    // a recording evaluation).
    void begin(int pass);
    // Ends a pass with the functor's result: the residual to use, or null when a symbolic
    // pass found no kernel (run pass 1).
    SP finish(const SP& residual);
    void abort();   // the pass threw TapeAbort
    static bool is_abort(const std::exception& e) { return dynamic_cast<const TapeAbort*>(&e) != nullptr; }
    bool fused() const;
    // the linear solver's in-place targets for fused Jacobian evaluations
    void set_sink(SinkFn f);
    // Residual (and, with `jacobian`, its Jacobian) at the values of `u`, by the kernel.
    SP eval(const SP& u, bool jacobian);

    struct Impl;

private:
    std::unique_ptr<Impl> impl_;
};

// Lowering check without a device: the kernel source of a small heat-like tape, and its
// hiprtc compilation (0 on success, the compiler log in `log`).
std::string residual_codegen_sample(int K, bool jac);
int residual_codegen_selftest(std::string* log);

}  // namespace eqhip
