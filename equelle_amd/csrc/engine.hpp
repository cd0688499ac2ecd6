// The forward-mode sparse AD engine and the grid-aware built-ins.
//
// Every function returns a fresh immutable value.  Jacobian structure follows
// the reference conventions (backends/cuda/src/collOfScalar.cu:273-403,
// backends/cuda/src/cudaMatrix.cu:640-800): union patterns for sums, structural
// products for Matrix*J, explicit zeros kept, columns ascending, empty Jacobian
// = constant.  The structural part is computed once per distinct input pattern
// (cached) and only numeric kernels run afterwards.
#pragma once
#include <string>
#include <vector>

#include "core.hpp"

namespace eqhip {

// ---------------------------------------------------------------- scalars
SP s_from_host(const double* h, int n);
SP s_uniform(int n, double v);
SP s_from_device(DevPtr<double> v, int n);
SP s_primary(const SP& a);
SP s_with_jacobian(const double* v, int n, const int* rp, const int* ci, const double* jv, int cols);
SP s_value_only(const SP& a);   // drops the Jacobian (newtonSolve result)
SP s_add(const SP& a, const SP& b);
SP s_sub(const SP& a, const SP& b);
SP s_mul(const SP& a, const SP& b);
SP s_div(const SP& a, const SP& b);
SP s_scal_mul(double s, const SP& a);
SP s_div_scal(const SP& a, double s);   // a / s = (1/s)*a  (collOfScalar.cu:385-387)
SP s_scal_div(double s, const SP& a);   // s / a
SP s_neg(const SP& a);
SP s_norm(const SP& a);   // norm(CollOfScalar): elementwise |a| with the sign-factor Jacobian
SP s_sqrt(const SP& a);
BP s_compare(int op, const SP& a, const SP& b);
BP s_compare_scalar(int op, const SP& a, double s, int scalar_left);
SP s_select(const BP& p, const SP& t, const SP& f);
double s_reduce(int op, const SP& a);
double s_two_norm(const SP& a);
std::vector<double> s_to_host(const SP& a);
// The Jacobian's values with its pending factors applied (null without a Jacobian); the
// first request on a value with pending factors runs one kernel and keeps the result.
DevPtr<double> jac_effective(const ScalarData& a);
const double* jac_values(const ScalarData& a);

// Matrix * CollOfScalar with a device CSR matrix (values on device, pattern cached)
struct DevCsr {
    PatternP pat;
    DevPtr<double> val;
    // host copy of the values' signs when every value is +1 or -1 (set by csr_from_host),
    // with an id for the symbolic cache; products with it skip the matrix values
    std::shared_ptr<const std::vector<int8_t>> sign;
    uint64_t sign_id = 0;
};
// `values`: precomputed output values (e.g. from a grid sweep); null => CSR SpMV.
SP s_matmul(const DevCsr& m, const SP& a, DevPtr<double> values = nullptr);
DevCsr csr_from_host(int rows, int cols, const int* rp, const int* ci, const double* v);

// Union / structural-product helpers exposed for tests of the CSR algebra.
DevCsr csr_add(const DevCsr& a, double alpha, const DevCsr& b, double beta);
DevCsr csr_mul(const DevCsr& a, const DevCsr& b);

// ---------------------------------------------------------------- index sets
struct PosMap {
    uint64_t id = 0;
    std::vector<int> pos;   // positions of sub's elements inside super
    std::vector<int> inv;   // |super| entries: index into sub or -1
    DevPtr<int> d_pos;
    DevPtr<int> d_inv;      // |super| entries: index into sub or -1
    int super_size = 0;
};
using PMP = std::shared_ptr<const PosMap>;
PMP position_map(const IP& super, const IP& sub);   // cached, throws if not a subset

SP s_on(const SP& a, const IP& from, const IP& to);
SP s_extend(const SP& a, const IP& from, const IP& to);
IP i_on(const IP& data, const IP& from, const IP& to);
IP i_select(const BP& p, const IP& t, const IP& f);
BP i_is_empty(const IP& s);
IP i_from_device(int codim, DevPtr<int> d, int n);

// ---------------------------------------------------------------- vectors
VP v_from_host(const double* h, int n, int dim);
VP v_binary(int op, const VP& a, const VP& b);        // EW_ADD / EW_SUB
VP v_scale(const VP& a, const SP& s, double sc, int divide);
SP v_norm(const VP& a);
SP v_dot(const VP& a, const VP& b);
SP v_col(const VP& a, int col);
std::vector<double> v_to_host(const VP& a);

BP b_from_host(const uint8_t* h, int n);
std::vector<uint8_t> b_to_host(const BP& b);

}  // namespace eqhip
