// Internal core of the MI355X Equelle runtime: device context, stream-ordered
// caching pool, immutable CSR sparsity patterns and the value types behind the
// public CollOfScalar / CollOfIndices / CollOfBool / CollOfVector handles.
//
// Design (gfx950):
//  * One HIP stream per process (default priority).  Every kernel and copy of
//    the runtime is ordered on it, so a pool block freed by the host can be
//    reused by the next allocation without any synchronisation.
//  * Jacobian sparsity patterns are immutable, content-hashed and shared.  The
//    AD rules never depend on values (union / structural product, explicit
//    zeros kept), so every symbolic result is cached by the ids of its input
//    patterns and only numeric kernels run after the first Newton iteration.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace eqhip {

[[noreturn]] void throw_error(const std::string& msg);
void check_hip(hipError_t e, const char* what, const char* file, int line);
#define EQ_HIP(x) ::eqhip::check_hip((x), #x, __FILE__, __LINE__)
#define EQ_REQUIRE(cond, msg)                                         \
    do {                                                              \
        if (!(cond)) {                                                \
            std::ostringstream eq_os_;                                \
            eq_os_ << msg;                                            \
            ::eqhip::throw_error(eq_os_.str());                       \
        }                                                             \
    } while (0)

// ----------------------------------------------------------------------------
// Device context + caching pool
// ----------------------------------------------------------------------------
class Device {
public:
    static Device& get();
    hipStream_t stream() const { return stream_; }
    void* alloc(size_t bytes);
    void release(void* p, size_t bytes);
    void sync();
    void trim();                         // return cached blocks to the driver
    size_t bytes_in_use() const { return in_use_; }
    size_t bytes_cached() const { return cached_; }
    int num_cus() const { return num_cus_; }
    // Pinned host scratch for small D2H reads (reductions, norms).
    double* host_scratch() { return host_scratch_; }

private:
    Device();
    ~Device();
    static size_t round_size(size_t bytes);
    hipStream_t stream_ = nullptr;
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    size_t in_use_ = 0, cached_ = 0;
    int num_cus_ = 256;
    double* host_scratch_ = nullptr;
};

// A pooled device block; shared ownership through std::shared_ptr.
template <class T>
struct DevArray {
    T* p = nullptr;
    size_t n = 0;
    explicit DevArray(size_t count) : n(count) {
        if (n) p = static_cast<T*>(Device::get().alloc(n * sizeof(T)));
    }
    ~DevArray() {
        if (p) Device::get().release(p, n * sizeof(T));
    }
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
};
template <class T>
using DevPtr = std::shared_ptr<DevArray<T>>;
template <class T>
DevPtr<T> dev_alloc(size_t n) {
    return std::make_shared<DevArray<T>>(n);
}
template <class T>
DevPtr<T> dev_from_host(const T* h, size_t n) {
    auto d = dev_alloc<T>(n);
    if (n) {
        // pageable source: wait so the caller may free it on return
        EQ_HIP(hipMemcpyAsync(d->p, h, n * sizeof(T), hipMemcpyHostToDevice, Device::get().stream()));
        Device::get().sync();
    }
    return d;
}
template <class T>
std::vector<T> dev_to_host(const T* d, size_t n) {
    std::vector<T> h(n);
    if (n) {
        EQ_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, Device::get().stream()));
        Device::get().sync();
    }
    return h;
}

// ----------------------------------------------------------------------------
// Sparsity pattern (CSR structure only), immutable and content-deduplicated.
// ----------------------------------------------------------------------------
struct Pattern {
    uint64_t id = 0;
    uint64_t hash = 0;
    int rows = 0, cols = 0, nnz = 0;
    DevPtr<int> d_rp, d_ci;           // the structure (symbolic passes run on the device)
    mutable DevPtr<int> d_row;        // row index of every entry (lazy)
    mutable int canon = -1;           // columns strictly ascending in every row (lazy; -1 unknown)
    mutable std::mutex mu;
    const int* row_of() const;        // device row-of-entry map
    bool ascending() const;           // columns strictly ascending in every row
    // host copies, for host-side consumers (copies to the caller, the CSR Krylov set-up):
    // kept for patterns made from host arrays, downloaded on first use otherwise
    const std::vector<int>& rp() const;
    const std::vector<int>& ci() const;
    mutable std::vector<int> h_rp, h_ci;
    mutable std::atomic<bool> on_host{false};
};
using PatternP = std::shared_ptr<const Pattern>;

// Content hash of a pattern, computed the same way on the host and on the device: a sum
// over the entries of pattern_mix((position << 32) | value) (row pointers at positions
// 0..rows, column indices at positions 2^31 + k), plus the dimensions.
inline uint64_t pattern_mix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// Returns the canonical (deduplicated) pattern with this structure.
PatternP make_pattern(int rows, int cols, std::vector<int> rp, std::vector<int> ci);
// The same from device arrays (rows + 1 row pointers, nnz columns); `ascending`: columns
// known to be strictly ascending in every row.
PatternP make_pattern_dev(int rows, int cols, int nnz, DevPtr<int> rp, DevPtr<int> ci, bool ascending);
PatternP identity_pattern(int n);

// Symbolic results cached by (op, input ids): output pattern plus device index
// maps that drive the numeric kernels.
uint64_t next_id();
struct SymEntry {
    uint64_t id = next_id();   // names the entry's index maps (compositions are cached by it)
    PatternP out;
    DevPtr<int> ia, ib;        // per output entry: index into operand a / b values, -1 if absent
    DevPtr<int> seg;           // spgemm: segment offsets (out.nnz+1)
    DevPtr<int> mi, ai;        // spgemm: (matrix entry, operand entry) pairs
    DevPtr<int> sai;           // spgemm with a +-1 matrix: operand entry, ~entry where the sign is -1
    std::vector<std::weak_ptr<const Pattern>> inputs;
};
using SymP = std::shared_ptr<const SymEntry>;
SymP sym_lookup(const std::string& key);
void sym_store(const std::string& key, SymP e);
std::string sym_key(const char* op, uint64_t a, uint64_t b = 0, uint64_t c = 0);

// ----------------------------------------------------------------------------
// Value types
// ----------------------------------------------------------------------------
struct ScalarData {
    int n = 0;
    DevPtr<double> val;
    bool ad = false;       // useAutoDiff flag of the reference CollOfScalar
    PatternP pat;          // null => empty (zero) Jacobian
    DevPtr<double> jval;   // pat->nnz stored values
    // The Jacobian is jscale * diag(jrow) * (pat, jval): scalings by constants and
    // collections without derivatives are kept as factors and applied by the next kernel
    // that reads the values (or once, by jac_values, where a caller needs them).
    double jscale = 1.0;
    DevPtr<double> jrow;   // n row factors, or null
    // stored value of entry k: jval[jmap[k]] (jmap null: jval[k]); set by row gathers (On,
    // Extend), which then share the operand's values instead of copying them
    DevPtr<int> jmap;
    uint64_t jmap_id = 0;
    bool jones = false;    // (pat, jval) is the n x n identity with unit values (a primary variable)
    mutable DevPtr<double> jeff;   // the factors applied (materialized on first request)
    // Residual recording: the recording session and node this value is (0 / -1: none).
    // A value made by a symbolic recording pass has no device values (val is null).
    mutable uint64_t tsess = 0;
    mutable int tnode = -1;
    // nonzero: the Jacobian's values were written into the stencil linear solver's layout by
    // a fused residual evaluation (this token names it), not into jval
    uint64_t jsink = 0;
    // every value equals uvalue (a collection made from one Scalar)
    bool uniform = false;
    double uvalue = 0.0;
    bool jpending() const { return jscale != 1.0 || (bool)jrow; }
    int jcols() const { return pat ? pat->cols : 0; }
};
using SP = std::shared_ptr<const ScalarData>;

struct IndexData {
    uint64_t id = 0;
    int codim = 0;
    bool full = false;
    std::vector<int> host;   // indices (full sets hold 0..n-1 too)
    DevPtr<int> dev;
    int size() const { return (int)host.size(); }
};
using IP = std::shared_ptr<const IndexData>;
IP make_index(int codim, bool full, std::vector<int> host);

struct BoolData {
    int n = 0;
    DevPtr<uint8_t> val;
    mutable uint64_t tsess = 0;   // recording session and node, as for ScalarData
    mutable int tnode = -1;
};
using BP = std::shared_ptr<const BoolData>;

struct VectorData {   // n vectors of dimension dim, AoS on the device
    int n = 0, dim = 0;
    DevPtr<double> val;
};
using VP = std::shared_ptr<const VectorData>;

uint64_t next_id();

}  // namespace eqhip
