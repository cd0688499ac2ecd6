// Implementation of the equelleCUDA C++ surface (include/EquelleRuntimeCUDA.hpp,
// include/equelle_cuda_types.hpp) over the device engine.
#include <EquelleRuntimeCUDA.hpp>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iterator>
#include <sstream>

#include "core.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "tape.hpp"
#include "runtime_core.hpp"

namespace eqhip {
struct MatrixData {
    int rows = 0, cols = 0;
    DevCsr m;   // m.pat null => empty (zero) matrix
};
}  // namespace eqhip

namespace equelleCUDA {

using eqhip::BP;
using eqhip::IP;
using eqhip::SP;
using eqhip::VP;

// ---------------------------------------------------------------- SeqOfScalar
SeqOfScalar operator*(const SeqOfScalar& seq, const Scalar a) {
    SeqOfScalar out(seq);
    for (double& x : out) x *= a;
    return out;
}

// ---------------------------------------------------------------- CollOfBool
CollOfBool::CollOfBool() {}
CollOfBool::CollOfBool(const std::vector<bool>& host) {
    std::vector<uint8_t> h(host.begin(), host.end());
    d_ = eqhip::b_from_host(h.data(), (int)h.size());
}
CollOfBool::CollOfBool(std::shared_ptr<const eqhip::BoolData> d) : d_(std::move(d)) {}
int CollOfBool::size() const { return d_ ? d_->n : 0; }
std::vector<bool> CollOfBool::toHost() const {
    if (!d_) return {};
    std::vector<uint8_t> h = eqhip::b_to_host(d_);
    return std::vector<bool>(h.begin(), h.end());
}

// ---------------------------------------------------------------- index sets
IndexSetBase::IndexSetBase(int codim, const std::vector<int>& host) : d_(eqhip::make_index(codim, false, host)) {}
bool IndexSetBase::isFull() const { return d_ && d_->full; }
int IndexSetBase::size() const { return d_ ? d_->size() : 0; }
std::vector<int> IndexSetBase::stdToHost() const { return d_ ? d_->host : std::vector<int>(); }
CollOfBool IndexSetBase::isEmpty() const { return CollOfBool(eqhip::i_is_empty(d_)); }
int IndexSetBase::codimension() const { return d_ ? d_->codim : 0; }
void index_contains(const IndexSetBase& super, const IndexSetBase& sub, const std::string& name) {
    try {
        eqhip::position_map(super.data(), sub.data());
    } catch (const std::exception&) {
        OPM_THROW(std::runtime_error, "Input set " << name << " is not a subset of the given superset");
    }
}

// ---------------------------------------------------------------- CudaArray
CudaArray::CudaArray() {}
CudaArray::CudaArray(int size) : d_(eqhip::s_uniform(size, 0.0)) {}
CudaArray::CudaArray(int size, double value) : d_(eqhip::s_uniform(size, value)) {}
CudaArray::CudaArray(const std::vector<double>& host) : d_(eqhip::s_from_host(host.data(), (int)host.size())) {}
CudaArray::CudaArray(std::shared_ptr<const eqhip::ScalarData> d) : d_(eqhip::s_value_only(d)) {}
int CudaArray::size() const { return d_ ? d_->n : 0; }
std::vector<double> CudaArray::copyToHost() const { return d_ ? eqhip::s_to_host(d_) : std::vector<double>(); }

// ---------------------------------------------------------------- CudaMatrix
static std::shared_ptr<const eqhip::MatrixData> mat(int rows, int cols, eqhip::DevCsr m) {
    auto d = std::make_shared<eqhip::MatrixData>();
    d->rows = rows;
    d->cols = cols;
    d->m = std::move(m);
    return d;
}
CudaMatrix::CudaMatrix() : d_(mat(0, 0, eqhip::DevCsr())) {}
CudaMatrix::CudaMatrix(const double* val, const int* rowPtr, const int* colInd, int nnz, int rows, int cols) {
    if (rowPtr[rows] != nnz) OPM_THROW(std::runtime_error, "CudaMatrix: rowPtr[rows] != nnz");
    d_ = mat(rows, cols, eqhip::csr_from_host(rows, cols, rowPtr, colInd, val));
}
CudaMatrix::CudaMatrix(int size) {
    eqhip::DevCsr m;
    m.pat = eqhip::identity_pattern(size);
    m.val = eqhip::dev_alloc<double>(size);
    eqhip::launch_fill(m.val->p, 1.0, size);
    d_ = mat(size, size, m);
}
CudaMatrix::CudaMatrix(const CollOfScalar& diag) {
    const SP& s = diag.data();
    eqhip::tape_read(*s, "a matrix made from the unknown");
    eqhip::DevCsr m;
    m.pat = eqhip::identity_pattern(s->n);
    m.val = s->val;
    d_ = mat(s->n, s->n, m);
}
CudaMatrix::CudaMatrix(const CudaArray& diag) : CudaMatrix(CollOfScalar(diag)) {}
CudaMatrix::CudaMatrix(std::shared_ptr<const eqhip::MatrixData> d) : d_(std::move(d)) {}
int CudaMatrix::nnz() const { return d_->m.pat ? d_->m.pat->nnz : 0; }
int CudaMatrix::rows() const { return d_->rows; }
int CudaMatrix::cols() const { return d_->cols; }
bool CudaMatrix::isEmpty() const { return !d_->m.pat; }
hostMat CudaMatrix::toHost() const {
    hostMat h;
    h.rows = d_->rows;
    h.cols = d_->cols;
    h.nnz = nnz();
    if (d_->m.pat) {
        h.rowPtr = d_->m.pat->rp();
        h.colInd = d_->m.pat->ci();
        h.vals = eqhip::dev_to_host(d_->m.val->p, h.nnz);
    }
    return h;
}
CudaMatrix CudaMatrix::transpose() const {
    if (!d_->m.pat) return CudaMatrix();
    // structural transpose on the host (symbolic), values permuted on the device
    const eqhip::Pattern& p = *d_->m.pat;
    std::vector<int> rp(p.cols + 1, 0), ci(p.nnz), src(p.nnz);
    for (int c : p.ci()) rp[c + 1]++;
    for (int i = 0; i < p.cols; ++i) rp[i + 1] += rp[i];
    std::vector<int> pos(rp.begin(), rp.end() - 1);
    for (int r = 0; r < p.rows; ++r)
        for (int k = p.rp()[r]; k < p.rp()[r + 1]; ++k) {
            const int q = pos[p.ci()[k]]++;
            ci[q] = r;
            src[q] = k;
        }
    eqhip::DevCsr t;
    t.pat = eqhip::make_pattern(p.cols, p.rows, std::move(rp), std::move(ci));
    t.val = eqhip::dev_alloc<double>(p.nnz);
    auto dsrc = eqhip::dev_from_host(src.data(), src.size());
    eqhip::launch_gather(d_->m.val->p, dsrc->p, t.val->p, p.nnz);
    return CudaMatrix(mat(d_->cols, d_->rows, t));
}
CudaMatrix operator+(const CudaMatrix& lhs, const CudaMatrix& rhs) {
    if (lhs.isEmpty()) return rhs;
    if (rhs.isEmpty()) return lhs;
    if (lhs.rows() != rhs.rows() || lhs.cols() != rhs.cols())
        OPM_THROW(std::runtime_error, "CudaMatrix + CudaMatrix: matrices of different sizes");
    return CudaMatrix(mat(lhs.rows(), lhs.cols(), eqhip::csr_add(lhs.data()->m, 1.0, rhs.data()->m, 1.0)));
}
CudaMatrix operator-(const CudaMatrix& lhs, const CudaMatrix& rhs) {
    if (lhs.isEmpty()) return -1.0 * rhs;
    if (rhs.isEmpty()) return lhs;
    if (lhs.rows() != rhs.rows() || lhs.cols() != rhs.cols())
        OPM_THROW(std::runtime_error, "CudaMatrix - CudaMatrix: matrices of different sizes");
    return CudaMatrix(mat(lhs.rows(), lhs.cols(), eqhip::csr_add(lhs.data()->m, 1.0, rhs.data()->m, -1.0)));
}
CudaMatrix operator*(const CudaMatrix& lhs, const CudaMatrix& rhs) {
    if (lhs.isEmpty() || rhs.isEmpty()) return CudaMatrix();
    if (lhs.cols() != rhs.rows())
        OPM_THROW(std::runtime_error, "Error in CudaMatrix * CudaMatrix size checking: lhs.cols_ = "
                                          << lhs.cols() << " rhs.rows_ = " << rhs.rows());
    return CudaMatrix(mat(lhs.rows(), rhs.cols(), eqhip::csr_mul(lhs.data()->m, rhs.data()->m)));
}
CudaArray operator*(const CudaMatrix& m, const CudaArray& vec) {
    if (m.isEmpty() || m.cols() != vec.size())
        OPM_THROW(std::runtime_error, "Error in matrix * vector operation as matrix is of size "
                                          << m.rows() << " by " << m.cols() << " and the vector of size "
                                          << vec.size());
    return CudaArray(eqhip::s_matmul(m.data()->m, vec.data()));
}
CudaMatrix operator*(const Scalar lhs, const CudaMatrix& rhs) {
    if (rhs.isEmpty()) OPM_THROW(std::runtime_error, "Calling CudaMatrix * Scalar with empty matrix...");
    return CudaMatrix(mat(rhs.rows(), rhs.cols(), eqhip::csr_add(rhs.data()->m, lhs, eqhip::DevCsr(), 0.0)));
}
CudaMatrix operator*(const CudaMatrix& lhs, const Scalar rhs) { return rhs * lhs; }
CudaMatrix operator-(const CudaMatrix& arg) { return -1.0 * arg; }

// ---------------------------------------------------------------- CollOfScalar
CollOfScalar::CollOfScalar() : d_(eqhip::s_uniform(0, 0.0)) {}
CollOfScalar::CollOfScalar(const int size) : d_(eqhip::s_uniform(size, 0.0)) {}
CollOfScalar::CollOfScalar(const int size, const double value) : d_(eqhip::s_uniform(size, value)) {}
CollOfScalar::CollOfScalar(const std::vector<double>& h) : d_(eqhip::s_from_host(h.data(), (int)h.size())) {}
CollOfScalar::CollOfScalar(const CudaArray& val) : d_(val.data() ? val.data() : eqhip::s_uniform(0, 0.0)) {}
CollOfScalar::CollOfScalar(const CudaArray& val, const CudaMatrix& der) {
    eqhip::tape_read(*val.data(), "a collection made from values of the unknown and a matrix");
    auto s = std::make_shared<eqhip::ScalarData>(*val.data());
    s->tsess = 0;
    s->tnode = -1;
    s->uniform = false;
    s->ad = true;
    s->jscale = 1.0;
    s->jrow = nullptr;
    s->jones = false;
    s->jmap = nullptr;
    s->jmap_id = 0;
    s->jeff = nullptr;
    if (!der.isEmpty()) {
        if (der.rows() != s->n) OPM_THROW(std::runtime_error, "CollOfScalar: Jacobian rows != collection size");
        s->pat = der.data()->m.pat;
        s->jval = der.data()->m.val;
    }
    d_ = s;
}
CollOfScalar::CollOfScalar(const CollOfScalar& val, const bool primaryVariable) {
    if (!primaryVariable)
        OPM_THROW(std::runtime_error, "Trying to create a primary variable with primaryVarible = " << primaryVariable);
    d_ = eqhip::s_primary(val.d_);
}
CollOfScalar::CollOfScalar(std::shared_ptr<const eqhip::ScalarData> d) : d_(std::move(d)) {}
CollOfScalar& CollOfScalar::operator*=(const Scalar lhs) {
    d_ = eqhip::s_scal_mul(lhs, d_);
    return *this;
}
CollOfScalar& CollOfScalar::operator*=(const CollOfScalar& rhs) {
    d_ = eqhip::s_mul(d_, rhs.d_);
    return *this;
}
int CollOfScalar::size() const { return d_->n; }
std::vector<double> CollOfScalar::copyToHost() const { return eqhip::s_to_host(d_); }
hostMat CollOfScalar::matrixToHost() const {
    if (!d_->ad) OPM_THROW(std::runtime_error, "Trying to copy empty matrix to host\n");
    return derivative().toHost();
}
bool CollOfScalar::useAutoDiff() const { return d_->ad; }
CudaMatrix CollOfScalar::derivative() const {
    if (!d_->pat) return CudaMatrix();
    eqhip::DevCsr m;
    m.pat = d_->pat;
    m.val = eqhip::jac_effective(*d_);
    return CudaMatrix(mat(d_->pat->rows, d_->pat->cols, m));
}
CudaArray CollOfScalar::value() const { return CudaArray(d_); }
double CollOfScalar::reduce(const EquelleReduce r) const {
    const int op = r == SUM ? eqhip::RED_SUM : r == PRODUCT ? eqhip::RED_PROD : r == MAX ? eqhip::RED_MAX
                                                                                             : eqhip::RED_MIN;
    return eqhip::s_reduce(op, d_);
}

#define EQ_WRAP(expr)                                                  \
    try {                                                              \
        return expr;                                                   \
    } catch (const std::runtime_error&) {                              \
        throw;                                                         \
    }
CollOfScalar operator+(const CollOfScalar& a, const CollOfScalar& b) { return CollOfScalar(eqhip::s_add(a.data(), b.data())); }
CollOfScalar operator-(const CollOfScalar& a, const CollOfScalar& b) { return CollOfScalar(eqhip::s_sub(a.data(), b.data())); }
CollOfScalar operator*(const CollOfScalar& a, const CollOfScalar& b) { return CollOfScalar(eqhip::s_mul(a.data(), b.data())); }
CollOfScalar operator/(const CollOfScalar& a, const CollOfScalar& b) { return CollOfScalar(eqhip::s_div(a.data(), b.data())); }
CollOfScalar operator*(const Scalar s, const CollOfScalar& a) { return CollOfScalar(eqhip::s_scal_mul(s, a.data())); }
CollOfScalar operator*(const CollOfScalar& a, const Scalar s) { return CollOfScalar(eqhip::s_scal_mul(s, a.data())); }
CollOfScalar operator/(const CollOfScalar& a, const Scalar s) { return CollOfScalar(eqhip::s_div_scal(a.data(), s)); }
CollOfScalar operator/(const Scalar s, const CollOfScalar& a) { return CollOfScalar(eqhip::s_scal_div(s, a.data())); }
CollOfScalar operator-(const CollOfScalar& a) { return CollOfScalar(eqhip::s_neg(a.data())); }
#undef EQ_WRAP

#define EQ_CMP(OP, CODE, SWAPCODE)                                                                  \
    CollOfBool operator OP(const CollOfScalar& a, const CollOfScalar& b) {                          \
        return CollOfBool(eqhip::s_compare(CODE, a.data(), b.data()));                              \
    }                                                                                               \
    CollOfBool operator OP(const CollOfScalar& a, const Scalar s) {                                 \
        return CollOfBool(eqhip::s_compare_scalar(CODE, a.data(), s, 0));                           \
    }                                                                                               \
    CollOfBool operator OP(const Scalar s, const CollOfScalar& a) {                                 \
        return CollOfBool(eqhip::s_compare_scalar(CODE, a.data(), s, 1));                           \
    }
EQ_CMP(>, eqhip::CMP_GT, 0)
EQ_CMP(>=, eqhip::CMP_GE, 0)
EQ_CMP(<, eqhip::CMP_LT, 0)
EQ_CMP(<=, eqhip::CMP_LE, 0)
EQ_CMP(==, eqhip::CMP_EQ, 0)
EQ_CMP(!=, eqhip::CMP_NE, 0)
#undef EQ_CMP

CollOfScalar operator*(const CudaMatrix& m, const CollOfScalar& c) {
    if (m.isEmpty()) OPM_THROW(std::runtime_error, "Matrix * CollOfScalar with an empty matrix");
    return CollOfScalar(eqhip::s_matmul(m.data()->m, c.data()));
}

// ---------------------------------------------------------------- CollOfVector
CollOfVector::CollOfVector() {}
CollOfVector::CollOfVector(const int size, const int dim) {
    std::vector<double> z((size_t)size * dim, 0.0);
    d_ = eqhip::v_from_host(z.data(), size, dim);
}
CollOfVector::CollOfVector(const std::vector<double>& host, const int dim) {
    if (dim <= 0 || host.size() % dim) OPM_THROW(std::runtime_error, "CollOfVector: size not a multiple of dim");
    d_ = eqhip::v_from_host(host.data(), (int)(host.size() / dim), dim);
}
CollOfVector::CollOfVector(std::shared_ptr<const eqhip::VectorData> d) : d_(std::move(d)) {}
CollOfScalar CollOfVector::norm() const { return CollOfScalar(eqhip::v_norm(d_)); }
CollOfScalar CollOfVector::dot(const CollOfVector& rhs) const { return CollOfScalar(eqhip::v_dot(d_, rhs.d_)); }
int CollOfVector::dim() const { return d_ ? d_->dim : 0; }
int CollOfVector::numVectors() const { return d_ ? d_->n : 0; }
int CollOfVector::numElements() const { return numVectors() * dim(); }
CollOfScalar CollOfVector::operator[](const int index) const { return CollOfScalar(eqhip::v_col(d_, index)); }
CollOfScalar CollOfVector::col(const int index) const { return (*this)[index]; }
std::vector<double> CollOfVector::toHost() const { return eqhip::v_to_host(d_); }
CollOfVector operator+(const CollOfVector& a, const CollOfVector& b) {
    return CollOfVector(eqhip::v_binary(eqhip::EW_ADD, a.data(), b.data()));
}
CollOfVector operator-(const CollOfVector& a, const CollOfVector& b) {
    return CollOfVector(eqhip::v_binary(eqhip::EW_SUB, a.data(), b.data()));
}
CollOfVector operator-(const CollOfVector& a) { return CollOfVector(eqhip::v_scale(a.data(), nullptr, -1.0, 0)); }
CollOfVector operator*(const CollOfVector& v, const Scalar s) { return CollOfVector(eqhip::v_scale(v.data(), nullptr, s, 0)); }
CollOfVector operator*(const Scalar s, const CollOfVector& v) { return v * s; }
CollOfVector operator*(const CollOfVector& v, const CollOfScalar& s) {
    return CollOfVector(eqhip::v_scale(v.data(), s.data(), 1.0, 0));
}
CollOfVector operator*(const CollOfScalar& s, const CollOfVector& v) { return v * s; }
CollOfVector operator/(const CollOfVector& v, const Scalar s) { return CollOfVector(eqhip::v_scale(v.data(), nullptr, s, 1)); }
CollOfVector operator/(const CollOfVector& v, const CollOfScalar& s) {
    return CollOfVector(eqhip::v_scale(v.data(), s.data(), 1.0, 1));
}

// ---------------------------------------------------------------- fused residual
FusedResidual::FusedResidual(const EquelleRuntimeCUDA& er, int* site_state, const CollOfScalar& u)
    : s_(new eqhip::ResidualSession(site_state, u.data(), er.core().fuse_residual)) {
    // a Jacobian solved on the stencil kernels is written there by the fused evaluation
    const eqhip::RuntimeCore* core = &er.core();
    s_->set_sink([core](const eqhip::PatternP& P, eqhip::StencilSink* out) { return core->stencil_sink(P, out); });
}
FusedResidual::~FusedResidual() {}
void FusedResidual::begin(int pass) { s_->begin(pass); }
bool FusedResidual::finish(const CollOfScalar& residual, CollOfScalar* out) {
    SP r = s_->finish(residual.data());
    if (!r) return false;
    *out = CollOfScalar(r);
    return true;
}
void FusedResidual::aborted(const std::exception& e) {
    if (!eqhip::ResidualSession::is_abort(e)) {
        s_->abort();
        throw;
    }
    s_->abort();
}
bool FusedResidual::fused() const { return s_->fused(); }
CollOfScalar FusedResidual::eval(const CollOfScalar& u, bool jacobian) { return CollOfScalar(s_->eval(u.data(), jacobian)); }

// ---------------------------------------------------------------- runtime
static eqhip::ParamMap to_params(const Opm::ParameterGroup& p) {
    eqhip::ParamMap m;
    m.parse(p.toString());
    return m;
}

EquelleRuntimeCUDA::EquelleRuntimeCUDA(const Opm::ParameterGroup& param)
    : core_(new eqhip::RuntimeCore(to_params(param))), param_(param) {
    max_iter_ = core_->max_iter;
    verbose_ = core_->verbose;
}
EquelleRuntimeCUDA::~EquelleRuntimeCUDA() {}

CollOfCell EquelleRuntimeCUDA::allCells() const { return CollOfCell(core_->grid().all_cells); }
CollOfCell EquelleRuntimeCUDA::boundaryCells() const { return CollOfCell(core_->grid().boundary_cells); }
CollOfCell EquelleRuntimeCUDA::interiorCells() const { return CollOfCell(core_->grid().interior_cells); }
CollOfFace EquelleRuntimeCUDA::allFaces() const { return CollOfFace(core_->grid().all_faces); }
CollOfFace EquelleRuntimeCUDA::boundaryFaces() const { return CollOfFace(core_->grid().boundary_faces); }
CollOfFace EquelleRuntimeCUDA::interiorFaces() const { return CollOfFace(core_->grid().interior_faces); }
CollOfCell EquelleRuntimeCUDA::firstCell(CollOfFace faces) const { return CollOfCell(core_->firstCell(faces.data())); }
CollOfCell EquelleRuntimeCUDA::secondCell(CollOfFace faces) const {
    return CollOfCell(core_->secondCell(faces.data()));
}
CollOfScalar EquelleRuntimeCUDA::norm(const CollOfVector& v) const { return v.norm(); }
CollOfVector EquelleRuntimeCUDA::normal(const CollOfFace& faces) const {
    return CollOfVector(core_->normal(faces.data()));
}
CollOfScalar EquelleRuntimeCUDA::dot(const CollOfVector& a, const CollOfVector& b) const { return a.dot(b); }
// The reference CUDA backend throws on these two (EquelleRuntimeCUDA.cpp:311-319); they
// follow the serial backend here (EquelleRuntimeCPU.cpp:345-365).
CollOfScalar EquelleRuntimeCUDA::negGradient(const CollOfScalar& x) const {
    return CollOfScalar(core_->neg_gradient(x.data()));
}
CollOfScalar EquelleRuntimeCUDA::interiorDivergence(const CollOfScalar& x) const {
    return CollOfScalar(core_->interior_divergence(x.data()));
}
CollOfScalar EquelleRuntimeCUDA::norm(const CollOfScalar& x) const { return CollOfScalar(eqhip::s_norm(x.data())); }
CollOfScalar EquelleRuntimeCUDA::gradient(const CollOfScalar& x) const { return CollOfScalar(core_->gradient(x.data())); }
CollOfScalar EquelleRuntimeCUDA::gradient_matrix(const CollOfScalar& x) const {
    return CollOfScalar(core_->gradient_matrix(x.data()));
}
CollOfScalar EquelleRuntimeCUDA::divergence(const CollOfScalar& x) const {
    return CollOfScalar(core_->divergence(x.data()));
}
CollOfScalar EquelleRuntimeCUDA::divergence_matrix(const CollOfScalar& x) const {
    return CollOfScalar(core_->divergence_matrix(x.data()));
}
CollOfScalar EquelleRuntimeCUDA::trinaryIf(const CollOfBool& p, const CollOfScalar& t, const CollOfScalar& f) const {
    if (t.size() != f.size() || t.size() != p.size())
        OPM_THROW(std::runtime_error, "Collections are not of the same size");
    return CollOfScalar(eqhip::s_select(p.data(), t.data(), f.data()));
}
Scalar EquelleRuntimeCUDA::minReduce(const CollOfScalar& x) const { return x.reduce(MIN); }
Scalar EquelleRuntimeCUDA::maxReduce(const CollOfScalar& x) const { return x.reduce(MAX); }
Scalar EquelleRuntimeCUDA::sumReduce(const CollOfScalar& x) const { return x.reduce(SUM); }
Scalar EquelleRuntimeCUDA::prodReduce(const CollOfScalar& x) const { return x.reduce(PRODUCT); }
CollOfScalar EquelleRuntimeCUDA::sqrt(const CollOfScalar& x) const { return CollOfScalar(eqhip::s_sqrt(x.data())); }

Scalar EquelleRuntimeCUDA::twoNorm(const CollOfScalar& v) const { return eqhip::s_two_norm(v.data()); }

CollOfScalar EquelleRuntimeCUDA::normOfSet(const IndexSetBase& set) const { return CollOfScalar(core_->norm(set.data())); }
CollOfVector EquelleRuntimeCUDA::centroidOfSet(const IndexSetBase& set) const {
    return CollOfVector(core_->centroid(set.data()));
}
CollOfScalar EquelleRuntimeCUDA::extendImpl(const CollOfScalar& data, const IndexSetBase& from,
                                            const IndexSetBase& to) const {
    return CollOfScalar(eqhip::s_extend(data.data(), from.data(), to.data()));
}
CollOfScalar EquelleRuntimeCUDA::onImpl(const CollOfScalar& data, const IndexSetBase& from,
                                        const IndexSetBase& to) const {
    return CollOfScalar(eqhip::s_on(data.data(), from.data(), to.data()));
}
std::shared_ptr<const eqhip::IndexData> EquelleRuntimeCUDA::onIndicesImpl(const IndexSetBase& data,
                                                                          const IndexSetBase& from,
                                                                          const IndexSetBase& to) const {
    return eqhip::i_on(data.data(), from.data(), to.data());
}
std::shared_ptr<const eqhip::IndexData> EquelleRuntimeCUDA::trinaryIndicesImpl(const CollOfBool& p,
                                                                               const IndexSetBase& t,
                                                                               const IndexSetBase& f) const {
    return eqhip::i_select(p.data(), t.data(), f.data());
}

CollOfScalar EquelleRuntimeCUDA::subsetRange(const CollOfScalar& u, int start, int len) const {
    if (start < 0 || len < 0 || start + len > u.size())
        OPM_THROW(std::runtime_error, "subset [" << start << ", " << start + len << ") of " << u.size() << " rows");
    return CollOfScalar(eqhip::s_on(u.data(), core_->range_set(0, u.size()), core_->range_set(start, len)));
}

CollOfScalar EquelleRuntimeCUDA::supersetRange(const CollOfScalar& a, int start, int total) const {
    if (start < 0 || start + a.size() > total)
        OPM_THROW(std::runtime_error, "superset of " << a.size() << " rows at " << start << " in " << total);
    return CollOfScalar(eqhip::s_extend(a.data(), core_->range_set(start, a.size()), core_->range_set(0, total)));
}

double EquelleRuntimeCUDA::linearTolerance(int k, double rn, double r0, double first_ratio) const {
    return k == 0 ? core_->first_tolerance(first_ratio) : core_->linear_tolerance(rn, r0);
}

CollOfScalar EquelleRuntimeCUDA::solveForUpdate(const CollOfScalar& residual, int* iters, double tol) const {
    eqhip::LinReport rep;
    SP du = core_->solve_for_update(residual.data(), &rep, tol);
    if (iters) *iters = rep.iterations;
    return CollOfScalar(du);
}

int EquelleRuntimeCUDA::newtonStop(double rn, double r0) const {
    if (rn <= core_->abs_res_tol) return 1;
    if (core_->rel_res_tol > 0.0 && rn <= core_->rel_res_tol * r0) return 2;
    return 0;
}
void EquelleRuntimeCUDA::newtonIterationMessage(int iter, double rn) const {
    if (verbose_ > 1)
        std::cout << "    newtonSolve: iter = " << iter << " (max = " << max_iter_ << "), norm(residual) = " << rn
                  << " (tol = " << core_->abs_res_tol << ")" << std::endl;
}
void EquelleRuntimeCUDA::newtonDoneMessage(int iter, int stop, double seconds) const {
    if (verbose_ > 0) {
        if (!stop) std::cout << "Newton solver failed to converge in " << max_iter_ << " iterations" << std::endl;
        else std::cout << "Newton solver converged in " << iter << " iterations" << std::endl;
    }
    if (verbose_ > 1) std::cout << "Newton solver took: " << seconds << " seconds." << std::endl;
}

void EquelleRuntimeCUDA::output(const String& tag, Scalar val) const { std::cout << tag << " = " << val << std::endl; }

void EquelleRuntimeCUDA::output(const String& tag, const CollOfScalar& coll) {
    const std::vector<double> host = coll.copyToHost();
    if (core_->output_to_file) {
        int count = 0;
        auto it = outputcount_.find(tag);
        if (it == outputcount_.end()) outputcount_[tag] = 1;
        else count = it->second++;
        std::ostringstream fname;
        fname << tag << "-" << std::setw(5) << std::setfill('0') << count << ".output";
        std::ofstream file(fname.str().c_str());
        if (!file) OPM_THROW(std::runtime_error, "Failed to open " << fname.str());
        file.precision(16);
        std::copy(host.begin(), host.end(), std::ostream_iterator<double>(file, "\n"));
    } else {
        std::cout << "\n" << "Values in " << tag << std::endl;
        for (double v : host) std::cout << v << "  ";
        std::cout << std::endl;
    }
}

Scalar EquelleRuntimeCUDA::inputScalarWithDefault(const String& name, const Scalar def) {
    return param_.getDefault(name, def);
}

std::vector<int> EquelleRuntimeCUDA::readSubsetFile(const String& name) const {
    const String filename = param_.get<String>(name + "_filename");
    std::ifstream is(filename.c_str());
    if (!is) OPM_THROW(std::runtime_error, "Could not find file " << filename);
    std::vector<int> v((std::istream_iterator<int>(is)), std::istream_iterator<int>());
    if (!std::is_sorted(v.begin(), v.end()))
        OPM_THROW(std::runtime_error, "Input set " << name << " was not sorted in ascending order");
    return v;
}

CollOfScalar EquelleRuntimeCUDA::inputCollectionImpl(const String& name, int size) const {
    if (param_.getDefault(name + "_from_file", false)) {
        const String filename = param_.get<String>(name + "_filename");
        std::ifstream is(filename.c_str());
        if (!is) OPM_THROW(std::runtime_error, "Could not find file " << filename);
        std::vector<double> data((std::istream_iterator<double>(is)), std::istream_iterator<double>());
        if ((int)data.size() != size)
            OPM_THROW(std::runtime_error, "Unexpected size of input data for " << name << " in file " << filename);
        return CollOfScalar(data);
    }
    return CollOfScalar(size, param_.get<double>(name));
}

SeqOfScalar EquelleRuntimeCUDA::inputSequenceOfScalar(const String& name) {
    const String filename = param_.get<String>(name + "_filename");
    std::ifstream is(filename.c_str());
    if (!is) OPM_THROW(std::runtime_error, "Could not find file " << filename);
    return SeqOfScalar((std::istream_iterator<Scalar>(is)), std::istream_iterator<Scalar>());
}

void EquelleRuntimeCUDA::ensureGridDimensionMin(const int d) const {
    if (core_->grid().dim < d)
        OPM_THROW(std::runtime_error, "Equelle simulator requires minimum " << d << " dimensions, but grid only has "
                                                                            << core_->grid().dim << " dimensions.");
}

UnstructuredGrid EquelleRuntimeCUDA::getGrid() const {
    eqhip::GridHost& h = core_->grid().host;
    UnstructuredGrid g{};
    g.dimensions = h.dim;
    g.number_of_cells = h.nc;
    g.number_of_faces = h.nf;
    g.face_cells = h.face_cells.data();
    g.cell_facepos = h.cell_facepos.data();
    g.cell_faces = h.cell_faces.data();
    g.cell_centroids = h.cell_centroids.data();
    g.cell_volumes = h.cell_volumes.data();
    g.face_centroids = h.face_centroids.data();
    g.face_areas = h.face_areas.data();
    g.face_normals = h.face_normals.data();
    g.cartdims[0] = h.n[0];
    g.cartdims[1] = h.n[1];
    g.cartdims[2] = h.n[2];
    return g;
}
CudaMatrix EquelleRuntimeCUDA::getGradMatrix() const {
    const auto& g = core_->grid();
    return CudaMatrix(mat(g.nif, g.nc, g.grad));
}
CudaMatrix EquelleRuntimeCUDA::getDivMatrix() const {
    const auto& g = core_->grid();
    return CudaMatrix(mat(g.nc, g.nif, g.div));
}
CudaMatrix EquelleRuntimeCUDA::getFulldivMatrix() const {
    const auto& g = core_->grid();
    return CudaMatrix(mat(g.nc, g.nf, g.fulldiv));
}

}  // namespace equelleCUDA
