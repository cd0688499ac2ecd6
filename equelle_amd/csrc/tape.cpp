// Recorded residual programs and their fused kernels (see tape.hpp).
#include "tape.hpp"

#include <hip/hiprtc.h>

#include "kernels.hpp"

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>

namespace eqhip {

namespace {

thread_local Tape* g_tape = nullptr;
thread_local int g_suspend = 0;
std::atomic<uint64_t> g_session{1};
std::mutex g_mu;
ResidualStats g_stats;

// Largest Jacobian row a kernel carries derivatives for (slots per value).
constexpr int KMAX = 32;

}  // namespace

Tape* tape_current() { return g_suspend ? nullptr : g_tape; }
TapeSuspend::TapeSuspend() { ++g_suspend; }
TapeSuspend::~TapeSuspend() { --g_suspend; }

ResidualStats residual_stats() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_stats;
}

// ============================================================================ recording
Tape::Tape(uint64_t session, bool symbolic) : session_(session), symbolic_(symbolic) {}

int Tape::add(const TapeNode& nd) {
    nodes.push_back(nd);
    return (int)nodes.size() - 1;
}

int Tape::node(uint8_t op, int a, int b, int c, int res, int sub) {
    TapeNode nd;
    nd.op = op;
    nd.a = a;
    nd.b = b;
    nd.c = c;
    nd.res = res;
    nd.sub = (int8_t)sub;
    auto dad = [&](int i) { return i >= 0 && nodes[i].ad; };
    switch (op) {
        case T_CMP:
        case T_CMPS: nd.boolean = true; break;
        case T_VAL: break;
        case T_SEL: nd.ad = dad(b) || dad(c); break;
        default: nd.ad = dad(a) || dad(b) || dad(c); break;
    }
    return add(nd);
}

int Tape::operand(const SP& s) {
    const int id = node_of(*s);
    if (id >= 0) return id;
    // a value that does not depend on the unknown: an input of the kernel
    for (size_t i = 0; i < leaves.size(); ++i)
        if (leaves[i].get() == s.get()) {
            for (size_t k = 0; k < nodes.size(); ++k)
                if (nodes[k].op == T_LEAF && nodes[k].res == (int)i) return (int)k;
        }
    if (s->pat) unsupported("a collection with its own Jacobian");
    TapeNode nd;
    nd.op = T_LEAF;
    if (s->uniform) {   // one value: a constant of the kernel, not an array it reads
        nd.sub = 1;
        nd.res = constant(s->uvalue);
        keep.push_back(s);
        return add(nd);
    }
    nd.res = (int)leaves.size();
    leaves.push_back(s);
    return add(nd);
}

int Tape::operand(const BP& b) {
    const int id = node_of(*b);
    if (id >= 0) return id;
    for (size_t i = 0; i < bleaves.size(); ++i)
        if (bleaves[i].get() == b.get()) {
            for (size_t k = 0; k < nodes.size(); ++k)
                if (nodes[k].op == T_BLEAF && nodes[k].res == (int)i) return (int)k;
        }
    TapeNode nd;
    nd.op = T_BLEAF;
    nd.boolean = true;
    nd.res = (int)bleaves.size();
    bleaves.push_back(b);
    return add(nd);
}

int Tape::constant(double v) {
    consts.push_back(v);
    return (int)consts.size() - 1;
}

int Tape::map(const DevPtr<int>& m, uint64_t id, std::shared_ptr<const void> k) {
    for (size_t i = 0; i < maps.size(); ++i)
        if (maps[i].second == id) return (int)i;
    maps.emplace_back(m, id);
    keep.push_back(std::move(k));
    return (int)maps.size() - 1;
}

namespace {
// Longest row and signed-column array of a matrix, cached by pattern (and sign) id.
struct MatInfo {
    int maxrow = 0;
    bool regular = false;
    DevPtr<int> sci;
};
MatInfo matrix_info(const DevCsr& m) {
    static std::mutex mu;
    static std::map<std::pair<uint64_t, uint64_t>, MatInfo> cache;
    const auto key = std::make_pair(m.pat->id, m.sign ? m.sign_id : 0);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    MatInfo mi;
    const std::vector<int>& rp = m.pat->rp();
    for (int r = 0; r < m.pat->rows; ++r) mi.maxrow = std::max(mi.maxrow, rp[r + 1] - rp[r]);
    mi.regular = (long)mi.maxrow * m.pat->rows == m.pat->nnz;
    if (m.sign && m.pat->nnz > 0) {
        const std::vector<int>& ci = m.pat->ci();
        std::vector<int> sci(ci.size());
        for (size_t k = 0; k < ci.size(); ++k) sci[k] = (*m.sign)[k] > 0 ? ci[k] : ~ci[k];
        mi.sci = dev_from_host(sci.data(), sci.size());
    }
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = mi;
    return mi;
}
}  // namespace

int Tape::matrix(const DevCsr& m) {
    for (size_t i = 0; i < mats.size(); ++i)
        if (mats[i].pat.get() == m.pat.get() && mats[i].val.get() == m.val.get()) return (int)i;
    const MatInfo mi = matrix_info(m);
    mats.push_back(m);
    mat_maxrow.push_back(mi.maxrow);
    mat_regular.push_back(mi.regular);
    mat_sci.push_back(mi.sci);
    return (int)mats.size() - 1;
}

void Tape::tag(const ScalarData& s, int id) {
    if (s.tsess == session_) unsupported("a result shared with an operand");
    s.tsess = session_;
    s.tnode = id;
}
void Tape::tag(const BoolData& b, int id) {
    if (b.tsess == session_) unsupported("a result shared with an operand");
    b.tsess = session_;
    b.tnode = id;
}

SP Tape::symbolic_scalar(int id, int n, bool ad) {
    auto s = std::make_shared<ScalarData>();
    s->n = n;
    s->ad = ad;
    s->tsess = session_;
    s->tnode = id;
    return s;
}

BP Tape::symbolic_bool(int id, int n) {
    auto b = std::make_shared<BoolData>();
    b->n = n;
    b->tsess = session_;
    b->tnode = id;
    return b;
}

void Tape::primary(const SP& u) {
    TapeNode nd;
    nd.op = T_PRIM;
    nd.ad = true;
    n_primary = u->n;
    tag(*u, add(nd));
}

void Tape::unsupported(const char* what) {
    if (valid_) why = what;
    valid_ = false;
    if (symbolic_) throw TapeAbort();
}

// ============================================================================ lowering
namespace {

// Word layout of the kernel argument block (one 8-byte word per pointer, integer or
// constant); the generated code addresses the same words.
enum : int {
    W_U = 0, W_RP, W_CI, W_OUTV, W_OUTJ, W_N, W_ERR,
    // stencil output (mode 1): the linear solver's right-hand side and coefficients in place
    W_MODE, W_NX, W_NY, W_NZ, W_H, W_F32, W_SF, W_K0, W_C0 = W_K0 + 7, W_FIXED = W_C0 + 7
};

struct Layout {
    int leaf0, bleaf0, map0, mat0, const0, words;
    explicit Layout(const Tape& t) {
        leaf0 = W_FIXED;
        bleaf0 = leaf0 + (int)t.leaves.size();
        map0 = bleaf0 + (int)t.bleaves.size();
        mat0 = map0 + (int)t.maps.size();
        const0 = mat0 + 3 * (int)t.mats.size();
        words = const0 + (int)t.consts.size();
    }
};

// The tape's structure: node kinds and wiring (no constants, maps or matrix contents).
std::string structure_key(const Tape& t) {
    std::ostringstream o;
    o << "L" << t.leaves.size() << "B" << t.bleaves.size() << "M" << t.maps.size() << "X" << t.mats.size() << "C"
      << t.consts.size() << ":";
    for (size_t i = 0; i < t.mats.size(); ++i)
        o << (t.mats[i].pat->nnz > 0 ? t.mat_maxrow[i] : 0) << (t.mat_sci[i] ? "s" : "v")
          << (t.mat_regular[i] ? "r" : "") << ",";
    o << ":";
    for (const TapeNode& n : t.nodes)
        o << (int)n.op << "," << (int)n.sub << "," << n.a << "," << n.b << "," << n.c << "," << n.res << ","
          << (int)n.ad << (int)n.boolean << ";";
    return o.str();
}

// The structure plus the identities that fix the residual's Jacobian pattern: the maps,
// the matrix patterns and the sizes.
std::string pattern_key(const Tape& t, const std::string& skey, int root) {
    std::ostringstream o;
    o << skey << "|n" << t.n_primary << "|r" << root << "|";
    for (const auto& m : t.maps) o << m.second << ",";
    o << "|";
    for (const DevCsr& m : t.mats) o << m.pat->id << ",";
    o << "|";
    for (const SP& l : t.leaves) o << l->n << ",";
    return o.str();
}

const char* kPrelude = R"(
typedef unsigned long long u64;
struct Args { u64 w[NW_]; };
#define PD(i) ((const double*)A.w[i])
#define PI(i) ((const int*)A.w[i])
#define PB(i) ((const unsigned char*)A.w[i])
#define CD(i) __builtin_bit_cast(double, A.w[i])
struct D { double v; double d[K_]; };
__device__ __forceinline__ D mk(double v) { D r; r.v = v;
#pragma unroll
  for (int s = 0; s < K_; ++s) r.d[s] = 0.0; return r; }
__device__ __forceinline__ double V(double x) { return x; }
__device__ __forceinline__ double V(const D& x) { return x.v; }
template <class T> __device__ __forceinline__ T zero();
template <> __device__ __forceinline__ double zero<double>() { return 0.0; }
template <> __device__ __forceinline__ D zero<D>() { return mk(0.0); }
template <class T> __device__ __forceinline__ T cv(double x);
template <> __device__ __forceinline__ double cv<double>(double x) { return x; }
template <> __device__ __forceinline__ D cv<D>(double x) { return mk(x); }
template <class T> __device__ __forceinline__ T cv(const D& x) { return x; }
__device__ __forceinline__ D prim(double v, int e, const int* col) { D r; r.v = v;
#pragma unroll
  for (int s = 0; s < K_; ++s) r.d[s] = col[s] == e ? 1.0 : 0.0; return r; }
__device__ __forceinline__ bool seen(int e, const int* col) { bool f = false;
#pragma unroll
  for (int s = 0; s < K_; ++s) f |= col[s] == e; return f; }
#define DLOOP _Pragma("unroll") for (int s = 0; s < K_; ++s)
__device__ __forceinline__ D operator+(const D& a, const D& b) { D r; r.v = a.v + b.v; DLOOP r.d[s] = a.d[s] + b.d[s]; return r; }
__device__ __forceinline__ D operator+(const D& a, double b) { D r = a; r.v = a.v + b; return r; }
__device__ __forceinline__ D operator+(double a, const D& b) { D r = b; r.v = a + b.v; return r; }
__device__ __forceinline__ D operator-(const D& a, const D& b) { D r; r.v = a.v - b.v; DLOOP r.d[s] = a.d[s] - b.d[s]; return r; }
__device__ __forceinline__ D operator-(const D& a, double b) { D r = a; r.v = a.v - b; return r; }
__device__ __forceinline__ D operator-(double a, const D& b) { D r; r.v = a - b.v; DLOOP r.d[s] = -b.d[s]; return r; }
__device__ __forceinline__ D operator*(const D& a, const D& b) { D r; r.v = a.v * b.v; DLOOP r.d[s] = b.v * a.d[s] + a.v * b.d[s]; return r; }
__device__ __forceinline__ D operator*(const D& a, double b) { D r; r.v = a.v * b; DLOOP r.d[s] = b * a.d[s]; return r; }
__device__ __forceinline__ D operator*(double a, const D& b) { D r; r.v = a * b.v; DLOOP r.d[s] = a * b.d[s]; return r; }
__device__ __forceinline__ D operator/(const D& a, const D& b) { D r; r.v = a.v / b.v; const double c = 1.0 / (b.v * b.v);
  DLOOP r.d[s] = c * (b.v * a.d[s] - a.v * b.d[s]); return r; }
__device__ __forceinline__ D operator/(const D& a, double b) { D r; r.v = a.v / b; const double c = 1.0 / (b * b);
  DLOOP r.d[s] = (c * b) * a.d[s]; return r; }
__device__ __forceinline__ D operator/(double a, const D& b) { D r; r.v = a / b.v; const double c = 1.0 / (b.v * b.v);
  DLOOP r.d[s] = c * (-a * b.d[s]); return r; }
__device__ __forceinline__ double sdiv(double s, double a) { return s / a; }
__device__ __forceinline__ D sdiv(double s, const D& a) { D r; r.v = s / a.v; const double c = -(s / (a.v * a.v));
  DLOOP r.d[s] = c * a.d[s]; return r; }
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ D dsqrt(const D& a) { D r; r.v = sqrt(a.v); const double c = 0.5 / r.v;
  DLOOP r.d[s] = c * a.d[s]; return r; }
__device__ __forceinline__ double dnorm(double a) { const double f = a >= 0.0 ? 1.0 : -1.0; return f * a; }
__device__ __forceinline__ D dnorm(const D& a) { const double f = a.v >= 0.0 ? 1.0 : -1.0; D r; r.v = f * a.v;
  DLOOP r.d[s] = f * a.d[s]; return r; }
)";

const char* cmp_str(int op) {
    switch (op) {
        case CMP_GT: return ">";
        case CMP_GE: return ">=";
        case CMP_LT: return "<";
        case CMP_LE: return "<=";
        case CMP_EQ: return "==";
        default: return "!=";
    }
}

// Emits the body that evaluates the root node at row `r`: every node at an element index
// (a C++ expression of an int variable), values shared within a scope (a block of
// straight-line code at one set of element indices).
class Emitter {
public:
    Emitter(const Tape& t, const Layout& L, bool jac, int root) : t_(t), L_(L), jac_(jac), need_(t.nodes.size(), 0) {
        scopes_.emplace_back();
        // derivatives are carried only where the root's Jacobian uses them: not through
        // value-only reads, comparisons or predicates
        if (jac && t.nodes[root].ad) need_[root] = 1;
        for (int i = root; i >= 0; --i) {
            if (!need_[i]) continue;
            const TapeNode& n = t.nodes[i];
            if (n.op == T_VAL || n.op == T_CMP || n.op == T_CMPS) continue;
            for (int o : {n.op == T_SEL ? -1 : n.a, n.b, n.c})
                if (o >= 0 && t.nodes[o].ad) need_[o] = 1;
        }
    }

    std::string type(int id) const {
        const TapeNode& n = t_.nodes[id];
        if (n.boolean) return "bool";
        return jac_ && n.ad && need_[id] ? "D" : "double";
    }

    std::string emit(int id, const std::string& e) {
        const auto key = std::make_pair(id, e);
        for (auto it = scopes_.rbegin(); it != scopes_.rend(); ++it) {
            auto f = it->find(key);
            if (f != it->end()) return f->second;
        }
        const TapeNode& n = t_.nodes[id];
        const std::string T = type(id);
        const std::string v = "t" + std::to_string(tmp_++);
        // an operand carrying derivatives read where they are not needed: its value
        auto opnd = [&](int o, const std::string& at) {
            const std::string x = emit(o, at);
            return T != "D" && type(o) == "D" ? "V(" + x + ")" : x;
        };
        auto bin = [&](const char* op) {
            const std::string a = opnd(n.a, e), b = opnd(n.b, e);
            line(T + " " + v + " = " + a + " " + op + " " + b + ";");
        };
        switch (n.op) {
            case T_PRIM:
                if (T == "D") {
                    // an unknown outside the row's pattern on a path that contributes
                    const std::string p = pred(e);
                    line("const D " + v + " = prim(PD(" + w(W_U) + ")[" + e + "], " + e + ", col); miss |= " +
                         (p.empty() ? "" : p + " && ") + "!seen(" + e + ", col);");
                }
                else
                    line("const double " + v + " = PD(" + w(W_U) + ")[" + e + "];");
                break;
            case T_LEAF:
                if (n.sub & 1)
                    line("const double " + v + " = CD(" + w(L_.const0 + n.res) + ");");
                else
                    line("const double " + v + " = PD(" + w(L_.leaf0 + n.res) + ")[" + e + "];");
                break;
            case T_BLEAF: line("const bool " + v + " = PB(" + w(L_.bleaf0 + n.res) + ")[" + e + "] != 0;"); break;
            case T_ADD: bin("+"); break;
            case T_SUB: bin("-"); break;
            case T_MUL: bin("*"); break;
            case T_DIV: bin("/"); break;
            case T_SMUL: {
                const std::string a = opnd(n.a, e);
                line(T + " " + v + " = CD(" + w(L_.const0 + n.res) + ") * " + a + ";");
                break;
            }
            case T_SDIV: {
                const std::string a = opnd(n.a, e);
                line(T + " " + v + " = sdiv(CD(" + w(L_.const0 + n.res) + "), " + a + ");");
                break;
            }
            case T_SQRT: {
                const std::string a = opnd(n.a, e);
                line(T + " " + v + " = dsqrt(" + a + ");");
                break;
            }
            case T_NORM: {
                const std::string a = opnd(n.a, e);
                line(T + " " + v + " = dnorm(" + a + ");");
                break;
            }
            case T_VAL: {
                const std::string a = opnd(n.a, e);
                line("const double " + v + " = V(" + a + ");");
                break;
            }
            case T_CMP: {
                const std::string a = opnd(n.a, e), b = opnd(n.b, e);
                line("const bool " + v + " = V(" + a + ") " + cmp_str(n.sub & 15) + " V(" + b + ");");
                break;
            }
            case T_CMPS: {
                const std::string a = opnd(n.a, e);
                const std::string c = "CD(" + w(L_.const0 + n.res) + ")";
                if (n.sub & 16)
                    line("const bool " + v + " = " + c + " " + cmp_str(n.sub & 15) + " V(" + a + ");");
                else
                    line("const bool " + v + " = V(" + a + ") " + cmp_str(n.sub & 15) + " " + c + ";");
                break;
            }
            case T_SEL: {
                const std::string p = opnd(n.a, e), a = opnd(n.b, e), b = opnd(n.c, e);
                line(T + " " + v + " = " + p + " ? cv<" + T + ">(" + a + ") : cv<" + T + ">(" + b + ");");
                break;
            }
            case T_ON: {
                const std::string ix = "e" + std::to_string(tmp_++);
                line("const int " + ix + " = PI(" + w(L_.map0 + n.res) + ")[" + e + "];");
                guard(ix, e, "");
                const std::string a = opnd(n.a, ix);
                line("const " + T + " " + v + " = " + a + ";");
                break;
            }
            case T_EXT: {
                if (n.sub & 1) {   // empty operand: every row is a zero row
                    line("const " + T + " " + v + " = zero<" + T + ">();");
                    break;
                }
                // straight-line: rows outside the subset evaluate the operand at its element 0
                // and select zero, so the loads of all branches can be in flight together
                const std::string raw = "e" + std::to_string(tmp_++), ix = "e" + std::to_string(tmp_++);
                line("const int " + raw + " = PI(" + w(L_.map0 + n.res) + ")[" + e + "];");
                line("const int " + ix + " = " + raw + " >= 0 ? " + raw + " : 0;");
                guard(ix, e, raw + " >= 0");
                const std::string a = opnd(n.a, ix);
                line("const " + T + " " + v + " = " + raw + " >= 0 ? cv<" + T + ">(" + a + ") : zero<" + T + ">();");
                break;
            }
            case T_MAT: {
                const int m = L_.mat0 + 3 * n.res;
                const int R = t_.mats[n.res].pat->nnz > 0 ? t_.mat_maxrow[n.res] : 0;
                const bool sgn = (bool)t_.mat_sci[n.res];
                const bool reg = t_.mat_regular[n.res] && R <= 8;
                line(T + " " + v + " = zero<" + T + ">();");
                if (R == 0) break;
                const std::string qb = "q" + std::to_string(tmp_++), qe = "q" + std::to_string(tmp_++);
                if (reg)   // every row has R entries: row e starts at R e
                    line("const int " + qb + " = " + std::to_string(R) + " * " + e + ";");
                else
                    line("const int " + qb + " = PI(" + w(m) + ")[" + e + "], " + qe + " = PI(" + w(m) + ")[" + e +
                         " + 1];");
                // one entry: its column (sign folded in for a +-1 matrix) and contribution
                auto entry = [&](const std::string& q, const std::string& on) {
                    const std::string ix = "e" + std::to_string(tmp_++);
                    if (sgn) {
                        const std::string sc = "e" + std::to_string(tmp_++);
                        line("const int " + sc + " = PI(" + w(m + 1) + ")[" + q + "];");
                        line("const int " + ix + " = " + sc + " >= 0 ? " + sc + " : ~" + sc + ";");
                        guard(ix, e, on);
                        const std::string a = opnd(n.a, ix);
                        const std::string add = "(" + sc + " >= 0 ? " + v + " + " + a + " : " + v + " - " + a + ")";
                        line(v + " = " + (on.empty() ? add : on + " ? " + add + " : " + v) + ";");
                    } else {
                        line("const int " + ix + " = PI(" + w(m + 1) + ")[" + q + "];");
                        guard(ix, e, on);
                        const std::string a = opnd(n.a, ix);
                        const std::string add = v + " + PD(" + w(m + 2) + ")[" + q + "] * " + a;
                        line(v + " = " + (on.empty() ? add : on + " ? " + add + " : " + v) + ";");
                    }
                };
                if (R <= 8) {
                    // short rows: unrolled, every entry's loads issued together (entries past
                    // the row's end read entry 0 and are not added)
                    for (int j = 0; j < R; ++j) {
                        const std::string q = "q" + std::to_string(tmp_++);
                        if (reg) {
                            line("const int " + q + " = " + qb + " + " + std::to_string(j) + ";");
                            entry(q, "");
                            continue;
                        }
                        const std::string on = "o" + std::to_string(tmp_++);
                        line("const bool " + on + " = " + qb + " + " + std::to_string(j) + " < " + qe + ";");
                        line("const int " + q + " = " + on + " ? " + qb + " + " + std::to_string(j) + " : 0;");
                        entry(q, on);
                    }
                } else {
                    const std::string q = "q" + std::to_string(tmp_++);
                    line("for (int " + q + " = " + qb + "; " + q + " < " + qe + "; ++" + q + ") {");
                    push();
                    entry(q, "");
                    pop();
                    line("}");
                }
                break;
            }
            default: throw_error("residual lowering: unknown tape operation");
        }
        scopes_.back()[key] = v;
        return v;
    }

    std::string body() const { return out_.str(); }

private:
    static std::string w(int i) { return std::to_string(i); }
    // Whether element variable `e` is on a contributing path: "" (always), or a bool variable.
    std::string pred(const std::string& e) const {
        auto it = pred_.find(e);
        return it == pred_.end() ? std::string() : it->second;
    }
    // `ix` is reached from `from` when `cond` holds ("" : always)
    void guard(const std::string& ix, const std::string& from, const std::string& cond) {
        const std::string p = pred(from);
        if (cond.empty() && p.empty()) return;
        if (cond.empty()) {
            pred_[ix] = p;
            return;
        }
        const std::string b = "p" + std::to_string(tmp_++);
        line("const bool " + b + " = " + (p.empty() ? "" : p + " && ") + "(" + cond + ");");
        pred_[ix] = b;
    }
    void line(const std::string& s) { out_ << std::string(2 * (scopes_.size() + 1), ' ') << s << "\n"; }
    void push() { scopes_.emplace_back(); }
    void pop() { scopes_.pop_back(); }

    const Tape& t_;
    const Layout& L_;
    bool jac_;
    std::vector<char> need_;
    std::map<std::string, std::string> pred_;
    int tmp_ = 0;
    std::vector<std::map<std::pair<int, std::string>, std::string>> scopes_;
    std::ostringstream out_;
};

std::string kernel_source(const Tape& t, int root, int K, bool jac, bool have_pat) {
    const Layout L(t);
    Emitter em(t, L, jac, root);
    const std::string rv = em.emit(root, "r");
    const bool dres = jac && em.type(root) == "D";
    // a 5- or 7-point stencil row can go straight into the stencil solver's layout
    const bool stencil = jac && have_pat && (K == 5 || K == 7);
    std::ostringstream s;
    s << "#define NW_ " << L.words << "\n#define K_ " << K << "\n" << kPrelude;
    s << "extern \"C\" __global__ void __launch_bounds__(256) eqhip_residual(Args A) {\n";
    s << "  const int r = blockIdx.x * 256 + threadIdx.x;\n";
    s << "  if (r >= (int)A.w[" << W_N << "]) return;\n";
    if (jac) {
        s << "  int col[K_];\n  int b0 = 0, len = 0;\n";
        if (stencil) s << "  const int mode = (int)A.w[" << W_MODE << "];\n  int si = 0, sj = 0, sk = 0;\n";
        s << "  " << (stencil ? "if (mode == 0) " : "") << "{\n";
        if (have_pat) {
            s << "    b0 = PI(" << W_RP << ")[r];\n    len = PI(" << W_RP << ")[r + 1] - b0;\n";
            s << "#pragma unroll\n    for (int s = 0; s < K_; ++s) col[s] = s < len ? PI(" << W_CI
              << ")[b0 + s] : -1;\n";
        } else {
            s << "#pragma unroll\n    for (int s = 0; s < K_; ++s) col[s] = -1;\n";
        }
        s << "  }\n";
        if (stencil) {
            // one slot per neighbour in ascending column order (B, S, W, C, E, N, T; 2D:
            // S, W, C, E, N), -1 where the neighbour is outside the grid
            s << "  else {\n";
            s << "    const int nx = (int)A.w[" << W_NX << "], ny = (int)A.w[" << W_NY << "], nz = (int)A.w[" << W_NZ
              << "], nxy = nx * ny;\n";
            s << "    si = r % nx; sj = (r / nx) % ny; sk = r / nxy;\n";
            if (K == 7)
                s << "    col[0] = sk > 0 ? r - nxy : -1; col[1] = sj > 0 ? r - nx : -1; col[2] = si > 0 ? r - 1 : -1;"
                     " col[3] = r; col[4] = si < nx - 1 ? r + 1 : -1; col[5] = sj < ny - 1 ? r + nx : -1;"
                     " col[6] = sk < nz - 1 ? r + nxy : -1;\n";
            else
                s << "    col[0] = sj > 0 ? r - nx : -1; col[1] = si > 0 ? r - 1 : -1; col[2] = r;"
                     " col[3] = si < nx - 1 ? r + 1 : -1; col[4] = sj < ny - 1 ? r + nx : -1;\n";
            s << "  }\n";
        }
        s << "  bool miss = false;\n";
    }
    s << em.body();
    s << "  ((double*)A.w[" << W_OUTV << "])[r] = V(" << rv << ");\n";
    if (jac) {
        if (stencil) s << "  if (mode == 0) {\n";
        if (have_pat && dres) {
            s << "    double* J = (double*)A.w[" << W_OUTJ << "] + b0;\n";
            s << "#pragma unroll\n    for (int s = 0; s < K_; ++s) if (s < len) J[s] = " << rv << ".d[s];\n";
        } else if (have_pat) {
            s << "    double* J = (double*)A.w[" << W_OUTJ << "] + b0;\n";
            s << "    for (int s = 0; s < len; ++s) J[s] = 0.0;\n";
        }
        if (stencil) {
            // the stencil solver's right-hand side, its coefficients (q = C, W, E, S, N, B, T)
            // in red-black rank order and, in mixed precision, FP64 in cell order
            s << "  } else {\n";
            s << "    ((double*)A.w[" << W_SF << "])[r] = V(" << rv << ");\n";
            s << "    const int sl = ((si + sj + sk) & 1) * (int)A.w[" << W_H << "] + (r >> 1);\n";
            s << "    const bool f32 = A.w[" << W_F32 << "] != 0;\n";
            static const int cls7[7] = {5, 3, 1, 0, 2, 4, 6}, cls5[5] = {3, 1, 0, 2, 4};
            for (int q = 0; q < K; ++q) {
                const int c = K == 7 ? cls7[q] : cls5[q];
                const std::string v = dres ? rv + ".d[" + std::to_string(q) + "]" : std::string("0.0");
                s << "    { const double v = " << v << ";\n";
                s << "      if (f32) ((float*)A.w[" << W_K0 + c << "])[sl] = (float)v; else ((double*)A.w["
                  << W_K0 + c << "])[sl] = v;\n";
                s << "      if (A.w[" << W_C0 + c << "]) ((double*)A.w[" << W_C0 + c << "])[r] = v; }\n";
            }
            s << "  }\n";
        }
        // an unknown outside the row's pattern would lose its derivative: flag it
        s << "  if (miss) ((int*)A.w[" << W_ERR << "])[0] = 1;\n";
    }
    s << "}\n";
    return s.str();
}

struct Kernel {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    bool verified = false;
    bool bad = false;
};

struct RootInfo {
    PatternP pat;
    int K = 1;
    int n = 0;
    bool ad = false;
};

std::map<std::string, std::shared_ptr<Kernel>>& kernel_cache() {
    static std::map<std::string, std::shared_ptr<Kernel>> c;
    return c;
}
std::map<std::string, RootInfo>& root_cache() {
    static std::map<std::string, RootInfo> c;
    return c;
}

// Compiles generated source for gfx950; returns the code object (empty on failure, with
// the log in `log`).
std::vector<char> compile_source(const std::string& src, std::string* log) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "eqhip_residual.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        *log = "hiprtcCreateProgram failed";
        return {};
    }
    const char* opts[] = {"--offload-arch=gfx950", "-O3",     "-std=c++17", "-Xclang",
                          "-target-feature",       "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string lg(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &lg[0]);
    *log = lg;
    std::vector<char> code;
    if (rc == HIPRTC_SUCCESS) {
        size_t cs = 0;
        hiprtcGetCodeSize(prog, &cs);
        code.resize(cs);
        hiprtcGetCode(prog, code.data());
    }
    hiprtcDestroyProgram(&prog);
    return code;
}

std::shared_ptr<Kernel> get_kernel(const std::string& key, const Tape& t, int root, int K, bool jac, bool have_pat) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = kernel_cache().find(key);
        if (it != kernel_cache().end()) return it->second;
    }
    auto k = std::make_shared<Kernel>();
    const auto t0 = std::chrono::steady_clock::now();
    std::string log;
    const std::string src = kernel_source(t, root, K, jac, have_pat);
    // EQHIP_DUMP_RESIDUAL=<directory>: keep each generated source (for inspection)
    if (const char* dir = std::getenv("EQHIP_DUMP_RESIDUAL")) {
        const std::string path = std::string(dir) + "/eqhip_residual_" + std::to_string(std::hash<std::string>()(key)) +
                                 (jac ? "_jac" : "_val") + ".hip";
        if (FILE* f = std::fopen(path.c_str(), "w")) {
            std::fputs(src.c_str(), f);
            std::fclose(f);
        }
    }
    const std::vector<char> code = compile_source(src, &log);
    if (code.empty()) {
        fprintf(stderr, "eqhip: residual kernel did not compile; evaluating op by op\n%s\n", log.c_str());
        k->bad = true;
    } else if (hipModuleLoadData(&k->mod, code.data()) != hipSuccess ||
               hipModuleGetFunction(&k->fn, k->mod, "eqhip_residual") != hipSuccess) {
        (void)hipGetLastError();
        fprintf(stderr, "eqhip: residual kernel did not load; evaluating op by op\n");
        k->bad = true;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_stats.compiled++;
    g_stats.compile_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    kernel_cache()[key] = k;
    return k;
}

int max_row(const Pattern& p) {
    const std::vector<int>& rp = p.rp();
    int m = 0;
    for (int r = 0; r < p.rows; ++r) m = std::max(m, rp[r + 1] - rp[r]);
    return m;
}

// |a - b| <= rtol max(|a|, |b|) + atol max|a| elementwise
bool close(const std::vector<double>& a, const std::vector<double>& b, double rtol, double atol, double* worst) {
    double scale = 0.0;
    for (double x : a) scale = std::max(scale, std::fabs(x));
    *worst = 0.0;
    bool ok = a.size() == b.size();
    for (size_t i = 0; ok && i < a.size(); ++i) {
        const double d = std::fabs(a[i] - b[i]);
        const double lim = rtol * std::max(std::fabs(a[i]), std::fabs(b[i])) + atol * scale;
        if (!(d <= lim)) ok = false;
        if (scale > 0.0) *worst = std::max(*worst, d / scale);
    }
    return ok;
}

}  // namespace

// Source of the kernel for a small heat-like tape (lowering check without a device).
std::string residual_codegen_sample(int K, bool jac) {
    Tape t(0, false);
    TapeNode prim;
    prim.op = T_PRIM;
    prim.ad = true;
    t.add(prim);
    auto leaf = [&](int i) {
        TapeNode n;
        n.op = T_LEAF;
        n.res = i;
        t.leaves.push_back(nullptr);
        return t.add(n);
    };
    auto node = [&](uint8_t op, int a, int b, int res = -1) {
        TapeNode n;
        n.op = op;
        n.a = a;
        n.b = b;
        n.res = res;
        n.ad = (a >= 0 && t.nodes[a].ad) || (b >= 0 && t.nodes[b].ad);
        return t.add(n);
    };
    t.maps.resize(3);
    t.mats.resize(2);
    t.mat_maxrow = {2, 4};
    t.mat_regular = {1, 0};
    t.mat_sci.resize(2);
    for (DevCsr& m : t.mats) {
        auto p = std::make_shared<Pattern>();
        p->nnz = 1;
        m.pat = p;
    }
    t.consts = {0.1, 0.5};
    const int one = leaf(0), u0 = leaf(1), coef = leaf(2), trans = leaf(3);
    const int uu = node(T_MUL, 0, 0);
    const int kc = node(T_SMUL, node(T_ADD, one, uu), -1, 0);
    const int kf = node(T_SMUL, node(T_ADD, node(T_ON, kc, -1, 0), node(T_ON, kc, -1, 1)), -1, 1);
    const int flux = node(T_MUL, node(T_MUL, trans, kf), node(T_MAT, 0, -1, 0));
    const int div = node(T_MAT, node(T_EXT, flux, -1, 2), -1, 1);
    const int root = node(T_ADD, node(T_SUB, 0, u0), node(T_MUL, coef, div));
    return kernel_source(t, root, K, jac, true);
}

// Compiles the sample tape's kernel with hiprtc (no device needed); returns 0 on success
// and the compiler log in `log`.
int residual_codegen_selftest(std::string* log) {
    for (int jac = 0; jac < 2; ++jac) {
        const std::vector<char> code = compile_source(residual_codegen_sample(5, jac != 0), log);
        if (code.empty()) return 1;
    }
    return 0;
}

// ============================================================================ session
struct ResidualSession::Impl {
    int* site = nullptr;
    SP u;
    bool enabled = false;
    std::unique_ptr<Tape> tape;   // the current pass's recording
    std::unique_ptr<Tape> bound;  // the recording the kernel's inputs come from
    std::shared_ptr<Kernel> kj, kv;
    RootInfo root;
    std::vector<uint64_t> words;
    DevPtr<int> err;
    SinkFn sink;

    void stop() {
        if (g_tape == tape.get()) g_tape = nullptr;
    }

    void bind(const Tape& t) {
        const Layout L(t);
        words.assign(L.words, 0);
        if (root.pat) {
            words[W_RP] = (uint64_t)(uintptr_t)root.pat->d_rp->p;
            words[W_CI] = (uint64_t)(uintptr_t)root.pat->d_ci->p;
        }
        words[W_N] = (uint64_t)root.n;
        if (!err) err = dev_alloc<int>(1);
        words[W_ERR] = (uint64_t)(uintptr_t)err->p;
        for (size_t i = 0; i < t.leaves.size(); ++i) words[L.leaf0 + i] = (uint64_t)(uintptr_t)t.leaves[i]->val->p;
        for (size_t i = 0; i < t.bleaves.size(); ++i) words[L.bleaf0 + i] = (uint64_t)(uintptr_t)t.bleaves[i]->val->p;
        for (size_t i = 0; i < t.maps.size(); ++i) words[L.map0 + i] = (uint64_t)(uintptr_t)t.maps[i].first->p;
        for (size_t i = 0; i < t.mats.size(); ++i) {
            words[L.mat0 + 3 * i] = (uint64_t)(uintptr_t)t.mats[i].pat->d_rp->p;
            words[L.mat0 + 3 * i + 1] =
                (uint64_t)(uintptr_t)(t.mat_sci[i] ? t.mat_sci[i]->p : t.mats[i].pat->d_ci->p);
            words[L.mat0 + 3 * i + 2] = (uint64_t)(uintptr_t)t.mats[i].val->p;
        }
        for (size_t i = 0; i < t.consts.size(); ++i) std::memcpy(&words[L.const0 + i], &t.consts[i], 8);
    }

    // `sink`: the Jacobian may go straight into the stencil solver's layout
    SP run(const SP& x, bool jac, bool sink_ok = true) {
        EQ_REQUIRE(x->n == bound->n_primary, "fused residual: unknown of size " << x->n << ", recorded for "
                                                                                 << bound->n_primary);
        const Kernel& k = jac ? *kj : *kv;
        auto v = dev_alloc<double>(std::max(1, root.n));
        std::vector<uint64_t> a = words;
        StencilSink st;
        const bool to_sink = jac && sink_ok && sink && root.pat && (root.K == 5 || root.K == 7) && sink(root.pat, &st) &&
                             st.ncoef == root.K && (long)st.nx * st.ny * st.nz == root.n;
        DevPtr<double> j;
        if (jac && root.pat && !to_sink) j = dev_alloc<double>(std::max(1, root.pat->nnz));
        if (to_sink) {
            a[W_MODE] = 1;
            a[W_NX] = (uint64_t)st.nx;
            a[W_NY] = (uint64_t)st.ny;
            a[W_NZ] = (uint64_t)st.nz;
            a[W_H] = (uint64_t)st.h;
            a[W_F32] = st.fp32 ? 1 : 0;
            a[W_SF] = (uint64_t)(uintptr_t)st.F;
            for (int q = 0; q < 7; ++q) {
                a[W_K0 + q] = (uint64_t)(uintptr_t)st.K[q];
                a[W_C0 + q] = (uint64_t)(uintptr_t)st.C64[q];
            }
        }
        a[W_U] = (uint64_t)(uintptr_t)x->val->p;
        a[W_OUTV] = (uint64_t)(uintptr_t)v->p;
        a[W_OUTJ] = j ? (uint64_t)(uintptr_t)j->p : 0;
        void* params[] = {a.data()};   // the one by-value argument: the word block
        if (root.n > 0)
            EQ_HIP(hipModuleLaunchKernel(k.fn, (root.n + 255) / 256, 1, 1, 256, 1, 1, 0, Device::get().stream(),
                                         params, nullptr));
        {
            std::lock_guard<std::mutex> lk(g_mu);
            g_stats.evaluations++;
        }
        auto s = std::make_shared<ScalarData>();
        s->n = root.n;
        s->val = v;
        if (jac) {
            s->ad = root.ad;
            s->pat = root.pat;
            s->jval = j;
            s->jsink = to_sink ? st.token : 0;
        }
        return s;
    }

    // The kernels against the recorded evaluation (values, Jacobian, unknowns in the pattern).
    bool verify(const SP& eager) {
        launch_fill_int(err->p, 0, 1);
        const SP a = run(u, true, false), b = run(u, false);
        const std::vector<double> ve = s_to_host(eager), va = s_to_host(a), vb = s_to_host(b);
        double w1 = 0, w2 = 0, w3 = 0;
        bool ok = close(ve, va, 1e-10, 1e-13, &w1) && close(ve, vb, 1e-10, 1e-13, &w2);
        if (ok && root.pat) {
            const std::vector<double> je = dev_to_host(jac_values(*eager), root.pat->nnz);
            const std::vector<double> ja = dev_to_host(a->jval->p, root.pat->nnz);
            ok = close(je, ja, 1e-10, 1e-13, &w3);
        }
        const int miss = dev_to_host(err->p, 1)[0];
        if (!ok || miss)
            fprintf(stderr,
                    "eqhip: fused residual disagrees with the recorded evaluation (values %.3g, %.3g, Jacobian %.3g,"
                    " unknown outside the pattern %d); evaluating op by op\n",
                    w1, w2, w3, miss);
        return ok && !miss;
    }
};

ResidualSession::ResidualSession(int* site_state, const SP& u, bool enabled) : impl_(new Impl) {
    impl_->site = site_state;
    impl_->u = u;
    impl_->enabled = enabled && *site_state >= 0;
}

ResidualSession::~ResidualSession() { impl_->stop(); }

void ResidualSession::begin(int pass) {
    Impl& I = *impl_;
    I.stop();
    I.tape.reset();
    if (!I.enabled || pass > 1) return;
    if (pass == 0 && *I.site != 1) pass = 1;
    I.tape.reset(new Tape(g_session.fetch_add(1), pass == 0));
    I.tape->primary(I.u);
    g_tape = I.tape.get();
}

void ResidualSession::abort() {
    Impl& I = *impl_;
    I.stop();
    std::lock_guard<std::mutex> lk(g_mu);
    g_stats.aborted++;
    if (I.tape && I.tape->symbolic() && *I.site == 1) *I.site = 0;   // record again next time
    I.tape.reset();
}

bool ResidualSession::fused() const { return impl_->bound && impl_->kj && impl_->kv; }

void ResidualSession::set_sink(SinkFn f) { impl_->sink = std::move(f); }

SP ResidualSession::eval(const SP& u, bool jacobian) { return impl_->run(u, jacobian); }

SP ResidualSession::finish(const SP& res) {
    Impl& I = *impl_;
    I.stop();
    if (!I.tape) return res;
    std::unique_ptr<Tape> t = std::move(I.tape);
    const int root = t->node_of(*res);
    if (!t->valid() || root < 0) {
        // not replayable: this call site runs op by op from now on
        std::lock_guard<std::mutex> lk(g_mu);
        if (!t->valid()) g_stats.aborted++;
        *I.site = -1;
        return t->symbolic() ? nullptr : res;
    }
    const std::string skey = structure_key(*t);
    const std::string pkey = pattern_key(*t, skey, root);
    if (t->symbolic()) {
        RootInfo ri;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            auto it = root_cache().find(pkey);
            if (it == root_cache().end()) return nullptr;
            ri = it->second;
        }
        const std::string kk = skey + "|K" + std::to_string(ri.K) + (ri.pat ? "P" : "");
        std::shared_ptr<Kernel> kj, kv;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            auto a = kernel_cache().find(kk + "J"), b = kernel_cache().find(kk + "V");
            if (a == kernel_cache().end() || b == kernel_cache().end()) return nullptr;
            kj = a->second;
            kv = b->second;
        }
        if (kj->bad || kv->bad || !kj->verified || !kv->verified) return nullptr;
        I.root = ri;
        I.kj = kj;
        I.kv = kv;
        I.bound = std::move(t);
        I.bind(*I.bound);
        {
            std::lock_guard<std::mutex> lk(g_mu);
            g_stats.symbolic++;
        }
        return I.run(I.u, true);
    }
    // a recording evaluation: `res` holds the values and the Jacobian pattern
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_stats.recorded++;
    }
    RootInfo ri;
    ri.pat = res->pat;
    ri.n = res->n;
    ri.ad = res->ad;
    ri.K = ri.pat ? std::max(1, max_row(*ri.pat)) : 1;
    if (ri.K > KMAX) {
        *I.site = -1;
        return res;
    }
    const std::string kk = skey + "|K" + std::to_string(ri.K) + (ri.pat ? "P" : "");
    std::shared_ptr<Kernel> kj = get_kernel(kk + "J", *t, root, ri.K, true, (bool)ri.pat);
    std::shared_ptr<Kernel> kv = get_kernel(kk + "V", *t, root, ri.K, false, (bool)ri.pat);
    if (kj->bad || kv->bad) {
        *I.site = -1;
        return res;
    }
    I.root = ri;
    I.kj = kj;
    I.kv = kv;
    I.bound = std::move(t);
    I.bind(*I.bound);
    bool known;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        known = root_cache().count(pkey) != 0;
    }
    // checked once per kernel and grid (pattern key) against this recorded evaluation
    if (!known || !kj->verified || !kv->verified) {
        if (!I.verify(res)) {
            std::lock_guard<std::mutex> lk(g_mu);
            g_stats.rejected++;
            kj->bad = kv->bad = true;
            I.kj.reset();
            I.kv.reset();
            I.bound.reset();
            *I.site = -1;
            return res;
        }
        kj->verified = kv->verified = true;
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        root_cache()[pkey] = ri;
    }
    *I.site = 1;
    return res;
}

}  // namespace eqhip
