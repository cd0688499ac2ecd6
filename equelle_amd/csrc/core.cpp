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
uint64_t fnv(uint64_t h, const void* data, size_t bytes) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}
uint64_t hash_words(uint64_t h, const std::vector<int>& v) {
    // 64-bit mixing over 32-bit words (fast path for large patterns)
    for (int x : v) {
        h ^= (uint32_t)x + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    }
    return h;
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

bool Pattern::ascending() const {
    std::lock_guard<std::mutex> lk(mu);
    if (canon < 0) {
        canon = 1;
        for (int r = 0; r < rows && canon; ++r)
            for (int k = rp[r] + 1; k < rp[r + 1]; ++k)
                if (ci[k] <= ci[k - 1]) {
                    canon = 0;
                    break;
                }
    }
    return canon == 1;
}

PatternP make_pattern(int rows, int cols, std::vector<int> rp, std::vector<int> ci) {
    uint64_t h = 1469598103934665603ull;
    const int dims[2] = {rows, cols};
    h = fnv(h, dims, sizeof(dims));
    h = hash_words(h, rp);
    h = hash_words(h, ci);
    Registry& r = reg();
    {
        std::lock_guard<std::mutex> lk(r.mu);
        auto range = r.by_hash.equal_range(h);
        for (auto it = range.first; it != range.second;) {
            PatternP p = it->second.lock();
            if (!p) {
                it = r.by_hash.erase(it);
                continue;
            }
            if (p->rows == rows && p->cols == cols && p->rp == rp && p->ci == ci) return p;
            ++it;
        }
    }
    auto p = std::make_shared<Pattern>();
    p->id = next_id();
    p->hash = h;
    p->rows = rows;
    p->cols = cols;
    p->nnz = rp.empty() ? 0 : rp.back();
    p->d_rp = dev_from_host(rp.data(), rp.size());
    p->d_ci = dev_from_host(ci.data(), ci.size() ? ci.size() : 0);
    p->rp = std::move(rp);
    p->ci = std::move(ci);
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
