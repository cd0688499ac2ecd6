// Device context, caching pool, pattern registry and symbolic cache.
#include "core.hpp"

#include <atomic>
#include <cstring>

#include "kernels.hpp"

namespace eqhip {

void throw_error(const std::string& msg) { throw std::runtime_error(msg); }

void check_hip(hipError_t e, const char* what, const char* file, int line) {
    if (e != hipSuccess) {
        std::ostringstream os;
        os << "HIP error " << hipGetErrorName(e) << " (" << hipGetErrorString(e) << ") in " << what << " at "
           << file << ":" << line;
        throw std::runtime_error(os.str());
    }
}

uint64_t next_id() {
    static std::atomic<uint64_t> ctr{1};
    return ctr++;
}

// ----------------------------------------------------------------------------
Device& Device::get() {
    static Device* d = new Device();   // lives until process exit (no static-destruction races)
    return *d;
}

Device::Device() {
    int ndev = 0;
    EQ_HIP(hipGetDeviceCount(&ndev));
    EQ_REQUIRE(ndev > 0, "no HIP device visible");
    hipDeviceProp_t prop;
    EQ_HIP(hipGetDeviceProperties(&prop, 0));
    num_cus_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    EQ_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    EQ_HIP(hipHostMalloc((void**)&host_scratch_, 4096, hipHostMallocDefault));
    noalloc_scratch_init();   // device address of the allocation-free reduction's scratch
}

Device::~Device() {}

size_t Device::round_size(size_t bytes) {
    if (bytes <= 256) return 256;
    if (bytes <= (1u << 20)) {   // powers of two up to 1 MiB
        size_t s = 256;
        while (s < bytes) s <<= 1;
        return s;
    }
    // larger: round up to 1/8 of the enclosing power of two (<= 12.5% slack)
    size_t p = 1u << 20;
    while (p < bytes) p <<= 1;
    const size_t step = p / 16;
    return (bytes + step - 1) / step * step;
}

void* Device::alloc(size_t bytes) {
    const size_t sz = round_size(bytes);
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.find(sz);
        if (it != free_.end()) {
            void* p = it->second;
            free_.erase(it);
            cached_ -= sz;
            in_use_ += sz;
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        EQ_HIP(hipMalloc(&p, sz));
    }
    std::lock_guard<std::mutex> lk(mu_);
    in_use_ += sz;
    return p;
}

void Device::release(void* p, size_t bytes) {
    const size_t sz = round_size(bytes);
    std::lock_guard<std::mutex> lk(mu_);
    free_.emplace(sz, p);
    in_use_ -= sz;
    cached_ += sz;
}

void Device::sync() { EQ_HIP(hipStreamSynchronize(stream_)); }

void Device::trim() {
    sync();
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached_ = 0;
}

// ----------------------------------------------------------------------------
// Pattern registry: content hash -> weak canonical pattern.
// ----------------------------------------------------------------------------
namespace {
struct Registry {
    std::mutex mu;
    std::unordered_multimap<uint64_t, std::weak_ptr<const Pattern>> by_hash;
    std::unordered_map<int, PatternP> identity;
    std::unordered_map<std::string, SymP> sym;
};
Registry& reg() {
    static Registry* r = new Registry();
    return *r;
}
// The host half of the content hash (pattern_mix over the entries, plus the dimensions);
// launch_pattern_hash computes the entry sum on the device.
uint64_t dims_hash(int rows, int cols) {
    return pattern_mix((uint64_t)(uint32_t)rows) * 3 + pattern_mix(((uint64_t)(uint32_t)cols) | (1ull << 40)) * 5;
}
uint64_t entries_hash(const std::vector<int>& rp, const std::vector<int>& ci) {
    uint64_t h = 0;
    for (size_t k = 0; k < rp.size(); ++k) h += pattern_mix(((uint64_t)k << 32) | (uint32_t)rp[k]);
    for (size_t k = 0; k < ci.size(); ++k) h += pattern_mix(((uint64_t)(k | 0x80000000ull) << 32) | (uint32_t)ci[k]);
    return h;
}
bool dev_equal(const int* a, const int* b, long n) { return n <= 0 || !int_arrays_differ(a, b, n); }
// A registered pattern with this structure, or null (registry lock held by the caller).
PatternP find_equal(Registry& r, uint64_t h, int rows, int cols, int nnz, const std::vector<int>* hrp,
                    const std::vector<int>* hci, const int* drp, const int* dci) {
    auto range = r.by_hash.equal_range(h);
    for (auto it = range.first; it != range.second;) {
        PatternP p = it->second.lock();
        if (!p) {
            it = r.by_hash.erase(it);
            continue;
        }
        if (p->rows == rows && p->cols == cols && p->nnz == nnz) {
            const bool same = hrp ? (p->rp() == *hrp && p->ci() == *hci)
                                  : (dev_equal(p->d_rp->p, drp, rows + 1L) && dev_equal(p->d_ci->p, dci, nnz));
            if (same) return p;
        }
        ++it;
    }
    return nullptr;
}
}  // namespace

const int* Pattern::row_of() const {
    std::lock_guard<std::mutex> lk(mu);
    if (!d_row) {
        d_row = dev_alloc<int>(nnz > 0 ? nnz : 1);
        if (nnz > 0) launch_row_of(d_rp->p, rows, d_row->p);
    }
    return d_row->p;
}

const std::vector<int>& Pattern::rp() const {
    if (on_host.load(std::memory_order_acquire)) return h_rp;
    std::lock_guard<std::mutex> lk(mu);
    if (!on_host.load(std::memory_order_relaxed)) {
        h_rp = dev_to_host(d_rp->p, (size_t)rows + 1);
        h_ci = dev_to_host(d_ci->p, (size_t)nnz);
        on_host.store(true, std::memory_order_release);
    }
    return h_rp;
}
const std::vector<int>& Pattern::ci() const {
    rp();
    return h_ci;
}

bool Pattern::ascending() const {
    {
        std::lock_guard<std::mutex> lk(mu);
        if (canon >= 0) return canon == 1;
    }
    const std::vector<int>& R = rp();
    const std::vector<int>& Cc = ci();
    int ok = 1;
    for (int r = 0; r < rows && ok; ++r)
        for (int k = R[r] + 1; k < R[r + 1]; ++k)
            if (Cc[k] <= Cc[k - 1]) {
                ok = 0;
                break;
            }
    std::lock_guard<std::mutex> lk(mu);
    canon = ok;
    return ok == 1;
}

PatternP make_pattern(int rows, int cols, std::vector<int> rp, std::vector<int> ci) {
    const int nnz = rp.empty() ? 0 : rp.back();
    const uint64_t h = dims_hash(rows, cols) + entries_hash(rp, ci);
    Registry& r = reg();
    {
        std::lock_guard<std::mutex> lk(r.mu);
        if (PatternP p = find_equal(r, h, rows, cols, nnz, &rp, &ci, nullptr, nullptr)) return p;
    }
    auto p = std::make_shared<Pattern>();
    p->id = next_id();
    p->hash = h;
    p->rows = rows;
    p->cols = cols;
    p->nnz = nnz;
    p->d_rp = dev_from_host(rp.data(), rp.size());
    p->d_ci = dev_from_host(ci.data(), ci.size() ? ci.size() : 0);
    p->h_rp = std::move(rp);
    p->h_ci = std::move(ci);
    p->on_host.store(true);
    std::lock_guard<std::mutex> lk(r.mu);
    r.by_hash.emplace(h, std::weak_ptr<const Pattern>(p));
    return p;
}

PatternP make_pattern_dev(int rows, int cols, int nnz, DevPtr<int> rp, DevPtr<int> ci, bool ascending) {
    const uint64_t h = dims_hash(rows, cols) + pattern_entries_hash(rp->p, rows + 1, ci->p, nnz);
    Registry& r = reg();
    {
        std::lock_guard<std::mutex> lk(r.mu);
        if (PatternP p = find_equal(r, h, rows, cols, nnz, nullptr, nullptr, rp->p, ci->p)) return p;
    }
    auto p = std::make_shared<Pattern>();
    p->id = next_id();
    p->hash = h;
    p->rows = rows;
    p->cols = cols;
    p->nnz = nnz;
    p->d_rp = std::move(rp);
    p->d_ci = std::move(ci);
    p->canon = ascending ? 1 : -1;
    std::lock_guard<std::mutex> lk(r.mu);
    r.by_hash.emplace(h, std::weak_ptr<const Pattern>(p));
    return p;
}

PatternP identity_pattern(int n) {
    Registry& r = reg();
    {
        std::lock_guard<std::mutex> lk(r.mu);
        auto it = r.identity.find(n);
        if (it != r.identity.end()) return it->second;
    }
    std::vector<int> rp(n + 1), ci(n);
    for (int i = 0; i <= n; ++i) rp[i] = i;
    for (int i = 0; i < n; ++i) ci[i] = i;
    PatternP p = make_pattern(n, n, std::move(rp), std::move(ci));
    std::lock_guard<std::mutex> lk(r.mu);
    r.identity[n] = p;
    return p;
}

std::string sym_key(const char* op, uint64_t a, uint64_t b, uint64_t c) {
    std::ostringstream os;
    os << op << ':' << a << ':' << b << ':' << c;
    return os.str();
}

SymP sym_lookup(const std::string& key) {
    Registry& r = reg();
    std::lock_guard<std::mutex> lk(r.mu);
    auto it = r.sym.find(key);
    if (it == r.sym.end()) return nullptr;
    for (auto& w : it->second->inputs)
        if (w.expired()) {   // an input pattern died: the entry can never hit again
            r.sym.erase(it);
            return nullptr;
        }
    return it->second;
}

void sym_store(const std::string& key, SymP e) {
    Registry& r = reg();
    std::lock_guard<std::mutex> lk(r.mu);
    if (r.sym.size() > 4096) {   // drop dead entries
        for (auto it = r.sym.begin(); it != r.sym.end();) {
            bool dead = false;
            for (auto& w : it->second->inputs) dead |= w.expired();
            it = dead ? r.sym.erase(it) : std::next(it);
        }
    }
    r.sym[key] = std::move(e);
}

IP make_index(int codim, bool full, std::vector<int> host) {
    auto d = std::make_shared<IndexData>();
    d->id = next_id();
    d->codim = codim;
    d->full = full;
    d->dev = dev_from_host(host.data(), host.size());
    d->host = std::move(host);
    return d;
}

}  // namespace eqhip
