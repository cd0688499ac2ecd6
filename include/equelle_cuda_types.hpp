// Value types of the equelleCUDA API surface (device-resident, MI355X runtime).
//
// Names, constructors and operators follow the reference headers that generated
// code and tests name:
//   CollOfScalar      backends/cuda/cuda_include/CollOfScalar.hpp:21-434
//   CollOfIndices     backends/cuda/cuda_include/CollOfIndices.hpp:39-178
//   CollOfVector      backends/cuda/cuda_include/CollOfVector.hpp:38-168
//   CudaArray         backends/cuda/cuda_include/CudaArray.hpp
//   CudaMatrix/hostMat backends/cuda/cuda_include/CudaMatrix.hpp:28-267
//   SeqOfScalar, Scalar, Bool, String, CollOfBool, EquelleReduce
//                     backends/cuda/cuda_include/equelleTypedefs.hpp, SeqOfScalar.hpp
// The headers are plain C++ (no HIP types): storage is owned by the library
// through shared handles, copies are shallow (values are immutable), and every
// operation runs on the device.
#ifndef EQUELLE_CUDA_TYPES_HPP_INCLUDED
#define EQUELLE_CUDA_TYPES_HPP_INCLUDED

#include <memory>
#include <string>
#include <tuple>
#include <vector>

namespace eqhip {
struct ScalarData;
struct IndexData;
struct BoolData;
struct VectorData;
struct MatrixData;
}  // namespace eqhip

namespace equelleCUDA {

typedef double Scalar;
typedef bool Bool;
typedef std::string String;
enum EquelleReduce { SUM, PRODUCT, MAX, MIN };

typedef std::vector<Scalar> SeqOfScalar;
SeqOfScalar operator*(const SeqOfScalar& seq, const Scalar a);

class CollOfScalar;
class CudaMatrix;

// ---------------------------------------------------------------- CollOfBool
class CollOfBool {
public:
    CollOfBool();
    explicit CollOfBool(const std::vector<bool>& host);
    explicit CollOfBool(std::shared_ptr<const eqhip::BoolData> d);
    int size() const;
    std::vector<bool> toHost() const;
    const std::shared_ptr<const eqhip::BoolData>& data() const { return d_; }

private:
    std::shared_ptr<const eqhip::BoolData> d_;
};

// ---------------------------------------------------------------- CollOfIndices
class IndexSetBase {
public:
    bool isFull() const;
    int size() const;
    std::vector<int> stdToHost() const;
    CollOfBool isEmpty() const;
    int codimension() const;
    const std::shared_ptr<const eqhip::IndexData>& data() const { return d_; }

protected:
    IndexSetBase() = default;
    IndexSetBase(int codim, const std::vector<int>& host);
    explicit IndexSetBase(std::shared_ptr<const eqhip::IndexData> d) : d_(std::move(d)) {}
    std::shared_ptr<const eqhip::IndexData> d_;
};

template <int codim>
class CollOfIndices : public IndexSetBase {
public:
    CollOfIndices() = default;
    explicit CollOfIndices(const std::vector<int>& host) : IndexSetBase(codim, host) {}
    explicit CollOfIndices(std::shared_ptr<const eqhip::IndexData> d) : IndexSetBase(std::move(d)) {}
    // Converting constructor between codims (operatorOn<codim_data, codim_set> result).
    template <int other>
    explicit CollOfIndices(const CollOfIndices<other>& o) : IndexSetBase(o.data()) {}
    // CollOfIndices::contains (CollOfIndices_impl.hpp): throws unless subset is a subset of *this
    void contains(const CollOfIndices<codim>& subset, const std::string& name) const;
};
typedef CollOfIndices<0> CollOfCell;
typedef CollOfIndices<1> CollOfFace;
void index_contains(const IndexSetBase& super, const IndexSetBase& sub, const std::string& name);
template <int codim>
void CollOfIndices<codim>::contains(const CollOfIndices<codim>& subset, const std::string& name) const {
    index_contains(*this, subset, name);
}

// ---------------------------------------------------------------- CudaArray
class CudaArray {
public:
    CudaArray();
    explicit CudaArray(int size);
    CudaArray(int size, double value);
    explicit CudaArray(const std::vector<double>& host);
    explicit CudaArray(std::shared_ptr<const eqhip::ScalarData> d);
    int size() const;
    std::vector<double> copyToHost() const;
    const std::shared_ptr<const eqhip::ScalarData>& data() const { return d_; }

private:
    std::shared_ptr<const eqhip::ScalarData> d_;
};

// ---------------------------------------------------------------- CudaMatrix
struct hostMat {
    std::vector<double> vals;
    std::vector<int> rowPtr;
    std::vector<int> colInd;
    int nnz;
    int rows;
    int cols;
};

class CudaMatrix {
public:
    CudaMatrix();                                             // empty (= zero) matrix
    CudaMatrix(const double* val, const int* rowPtr, const int* colInd, int nnz, int rows, int cols);
    explicit CudaMatrix(int size);                            // identity
    explicit CudaMatrix(const CollOfScalar& diag);            // diagonal of the values
    explicit CudaMatrix(const CudaArray& diag);
    explicit CudaMatrix(std::shared_ptr<const eqhip::MatrixData> d);
    int nnz() const;
    int rows() const;
    int cols() const;
    bool isEmpty() const;
    hostMat toHost() const;
    CudaMatrix transpose() const;
    const std::shared_ptr<const eqhip::MatrixData>& data() const { return d_; }

    friend CudaMatrix operator+(const CudaMatrix& lhs, const CudaMatrix& rhs);
    friend CudaMatrix operator-(const CudaMatrix& lhs, const CudaMatrix& rhs);
    friend CudaMatrix operator*(const CudaMatrix& lhs, const CudaMatrix& rhs);
    friend CudaArray operator*(const CudaMatrix& mat, const CudaArray& vec);
    friend CudaMatrix operator*(const CudaMatrix& lhs, const Scalar rhs);
    friend CudaMatrix operator*(const Scalar lhs, const CudaMatrix& rhs);
    friend CudaMatrix operator-(const CudaMatrix& arg);

private:
    std::shared_ptr<const eqhip::MatrixData> d_;
};

// ---------------------------------------------------------------- CollOfScalar
class CollOfScalar {
public:
    CollOfScalar();
    explicit CollOfScalar(const int size);
    explicit CollOfScalar(const int size, const double value);
    explicit CollOfScalar(const std::vector<double>& host_vec);
    explicit CollOfScalar(const CudaArray& val);
    CollOfScalar(const CudaArray& val, const CudaMatrix& der);
    // Primary variable: copies the values and takes an identity Jacobian.
    explicit CollOfScalar(const CollOfScalar& val, const bool primaryVariable);
    explicit CollOfScalar(std::shared_ptr<const eqhip::ScalarData> d);

    CollOfScalar& operator*=(const Scalar lhs);
    CollOfScalar& operator*=(const CollOfScalar& rhs);

    int size() const;
    std::vector<double> copyToHost() const;
    hostMat matrixToHost() const;
    bool useAutoDiff() const;
    CudaMatrix derivative() const;
    CudaArray value() const;
    double reduce(const EquelleReduce reduce) const;
    const std::shared_ptr<const eqhip::ScalarData>& data() const { return d_; }

private:
    std::shared_ptr<const eqhip::ScalarData> d_;
};

CollOfScalar operator+(const CollOfScalar& lhs, const CollOfScalar& rhs);
CollOfScalar operator-(const CollOfScalar& lhs, const CollOfScalar& rhs);
CollOfScalar operator*(const CollOfScalar& lhs, const CollOfScalar& rhs);
CollOfScalar operator/(const CollOfScalar& lhs, const CollOfScalar& rhs);
CollOfScalar operator*(const Scalar lhs, const CollOfScalar& rhs);
CollOfScalar operator*(const CollOfScalar& lhs, const Scalar rhs);
CollOfScalar operator/(const CollOfScalar& lhs, const Scalar rhs);
CollOfScalar operator/(const Scalar lhs, const CollOfScalar& rhs);
CollOfScalar operator-(const CollOfScalar& arg);
CollOfBool operator>(const CollOfScalar& lhs, const CollOfScalar& rhs);
CollOfBool operator>(const CollOfScalar& lhs, const Scalar rhs);
CollOfBool operator>(const Scalar lhs, const CollOfScalar& rhs);
CollOfBool operator<(const CollOfScalar& lhs, const CollOfScalar& rhs);
CollOfBool operator<(const CollOfScalar& lhs, const Scalar rhs);
CollOfBool operator<(const Scalar lhs, const CollOfScalar& rhs);
CollOfBool operator>=(const CollOfScalar& lhs, const CollOfScalar& rhs);
CollOfBool operator>=(const CollOfScalar& lhs, const Scalar rhs);
CollOfBool operator>=(const Scalar lhs, const CollOfScalar& rhs);
CollOfBool operator<=(const CollOfScalar& lhs, const CollOfScalar& rhs);
CollOfBool operator<=(const CollOfScalar& lhs, const Scalar rhs);
CollOfBool operator<=(const Scalar lhs, const CollOfScalar& rhs);
CollOfBool operator==(const CollOfScalar& lhs, const CollOfScalar& rhs);
CollOfBool operator==(const CollOfScalar& lhs, const Scalar rhs);
CollOfBool operator==(const Scalar lhs, const CollOfScalar& rhs);
CollOfBool operator!=(const CollOfScalar& lhs, const CollOfScalar& rhs);
CollOfBool operator!=(const CollOfScalar& lhs, const Scalar rhs);
CollOfBool operator!=(const Scalar lhs, const CollOfScalar& rhs);
// Matrix * CollOfScalar (CollOfScalar.hpp:434)
CollOfScalar operator*(const CudaMatrix& mat, const CollOfScalar& coll);

// ---------------------------------------------------------------- CollOfVector
class CollOfVector {
public:
    CollOfVector();
    CollOfVector(const int size, const int dim);
    CollOfVector(const std::vector<double>& host, const int dim);
    explicit CollOfVector(std::shared_ptr<const eqhip::VectorData> d);
    CollOfScalar norm() const;
    CollOfScalar dot(const CollOfVector& rhs) const;
    int dim() const;
    int numVectors() const;
    int numElements() const;
    CollOfScalar operator[](const int index) const;
    CollOfScalar col(const int index) const;
    std::vector<double> toHost() const;
    const std::shared_ptr<const eqhip::VectorData>& data() const { return d_; }

private:
    std::shared_ptr<const eqhip::VectorData> d_;
};
CollOfVector operator+(const CollOfVector& lhs, const CollOfVector& rhs);
CollOfVector operator-(const CollOfVector& lhs, const CollOfVector& rhs);
CollOfVector operator-(const CollOfVector& arg);
CollOfVector operator*(const CollOfVector& vec, const Scalar s);
CollOfVector operator*(const Scalar s, const CollOfVector& vec);
CollOfVector operator*(const CollOfVector& vec, const CollOfScalar& s);
CollOfVector operator*(const CollOfScalar& s, const CollOfVector& vec);
CollOfVector operator/(const CollOfVector& vec, const Scalar s);
CollOfVector operator/(const CollOfVector& vec, const CollOfScalar& s);

// ---------------------------------------------------------------- makeArray
// Arrays of collections (EquelleRuntimeCUDA.hpp:41-72).
template <typename T>
std::tuple<T> makeArray(const T& t) {
    return std::tuple<T>{t};
}
// The elements may differ in type: generated code passes residual lambdas to
// newtonSolveSystem as makeArray(pressureResLocal, transportResLocal)
// (examples/app/out_twophase_fully_implicit.cpp:135).
template <typename T1, typename T2>
std::tuple<T1, T2> makeArray(const T1& t1, const T2& t2) {
    return std::tuple<T1, T2>{t1, t2};
}
template <typename T1, typename T2, typename T3>
std::tuple<T1, T2, T3> makeArray(const T1& t1, const T2& t2, const T3& t3) {
    return std::tuple<T1, T2, T3>{t1, t2, t3};
}
template <typename T1, typename T2, typename T3, typename T4>
std::tuple<T1, T2, T3, T4> makeArray(const T1& t1, const T2& t2, const T3& t3, const T4& t4) {
    return std::tuple<T1, T2, T3, T4>{t1, t2, t3, t4};
}

}  // namespace equelleCUDA

#endif  // EQUELLE_CUDA_TYPES_HPP_INCLUDED
