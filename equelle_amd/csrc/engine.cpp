// Forward-mode sparse AD engine: symbolic phase (count / prefix-sum / fill kernels on the
// device, cached per pattern; a host pass only for structural products with rows of more
// than SPGEMM_CMAX contributions) and numeric phase (HIP kernels).  See engine.hpp.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "kernels.hpp"
#include "tape.hpp"

namespace eqhip {

namespace {

SP make_scalar(int n, DevPtr<double> val, bool ad, PatternP pat, DevPtr<double> jval) {
    auto s = std::make_shared<ScalarData>();
    s->n = n;
    s->val = std::move(val);
    s->ad = ad;
    s->pat = std::move(pat);
    s->jval = std::move(jval);
    return s;
}

void check_size(const SP& a, const SP& b, const char* op) {
    EQ_REQUIRE(a->n == b->n, "CollOfScalar sizes differ in operator" << op << " (" << a->n << " vs " << b->n << ")");
}

// The last entry of a device prefix-sum array (one synchronising read).
int read_last(const DevPtr<int>& a, long idx) { return dev_to_host(a->p + idx, 1)[0]; }

// Union of two patterns with maps from output entries to operand entries: two passes on
// the device (entries per row, prefix sum, fill).
SymP sym_union(const PatternP& a, const PatternP& b) {
    const std::string key = sym_key("union", a->id, b->id);
    if (SymP e = sym_lookup(key)) return e;
    EQ_REQUIRE(a->rows == b->rows && a->cols == b->cols,
               "Jacobians of different sizes (" << a->rows << "x" << a->cols << " vs " << b->rows << "x" << b->cols
                                                << ")");
    const int rows = a->rows;
    auto cnt = dev_alloc<int>(rows + 1);
    launch_union_count(rows, a->d_rp->p, a->d_ci->p, b->d_rp->p, b->d_ci->p, cnt->p);
    auto rp = dev_alloc<int>(rows + 1);
    exclusive_scan(cnt->p, rp->p, rows + 1);
    const int nnz = read_last(rp, rows);
    auto ci = dev_alloc<int>(std::max(1, nnz)), ia = dev_alloc<int>(std::max(1, nnz)),
         ib = dev_alloc<int>(std::max(1, nnz));
    launch_union_fill(rows, a->d_rp->p, a->d_ci->p, b->d_rp->p, b->d_ci->p, rp->p, ci->p, ia->p, ib->p);
    auto e = std::make_shared<SymEntry>();
    e->out = make_pattern_dev(rows, a->cols, nnz, rp, ci, a->ascending() && b->ascending());
    // an operand whose pattern is the union itself (as many entries) maps every entry to
    // itself: no map (null), and its numeric kernels read its values in entry order
    if (nnz != a->nnz) e->ia = ia;
    if (nnz != b->nnz) e->ib = ib;
    e->inputs = {a, b};
    sym_store(key, e);
    return e;
}

// Structural product M * A on the host (rows with more contributions than the device
// pass takes): output pattern + (matrix entry, operand entry) pairs.
std::shared_ptr<SymEntry> spgemm_host(const PatternP& m, const PatternP& a, const std::vector<int8_t>* sign) {
    const std::vector<int>&mrp = m->rp(), &mci = m->ci(), &arp = a->rp(), &aci = a->ci();
    std::vector<int> rp(m->rows + 1, 0), ci, seg(1, 0), mi, ai;
    struct Tri {
        int col, t, e;
    };
    std::vector<Tri> row;
    for (int r = 0; r < m->rows; ++r) {
        row.clear();
        for (int t = mrp[r]; t < mrp[r + 1]; ++t) {
            const int k = mci[t];
            for (int e = arp[k]; e < arp[k + 1]; ++e) row.push_back({aci[e], t, e});
        }
        // output columns ascending; contributions to one entry in matrix-entry order
        std::sort(row.begin(), row.end(), [](const Tri& x, const Tri& y) {
            return x.col != y.col ? x.col < y.col : x.t < y.t;
        });
        for (size_t i = 0; i < row.size(); ++i) {
            if (i == 0 || row[i].col != row[i - 1].col) {
                if (i > 0) seg.push_back((int)mi.size());
                ci.push_back(row[i].col);
            }
            mi.push_back(row[i].t);
            ai.push_back(row[i].e);
        }
        if (!row.empty()) seg.push_back((int)mi.size());
        rp[r + 1] = (int)ci.size();
    }
    auto e = std::make_shared<SymEntry>();
    e->out = make_pattern(m->rows, a->cols, std::move(rp), std::move(ci));
    e->seg = dev_from_host(seg.data(), seg.size());
    if (sign) {
        for (size_t t = 0; t < ai.size(); ++t)
            if ((*sign)[mi[t]] < 0) ai[t] = ~ai[t];
        e->sai = dev_from_host(ai.data(), ai.size());
    } else {
        e->mi = dev_from_host(mi.data(), mi.size());
        e->ai = dev_from_host(ai.data(), ai.size());
    }
    return e;
}

// Structural product M * A: output pattern + (matrix entry, operand entry) pairs.  With
// `sign` (M's values, all +1 or -1): the operand entry of each pair with M's sign folded
// in (sai = ai, or ~ai where M's entry is -1) instead of the two index arrays.  On the
// device in two passes (distinct columns and contributions per row, prefix sums, fill);
// rows of more than SPGEMM_CMAX contributions send the product to the host pass.
constexpr int SPGEMM_CMAX = 256;
SymP sym_spgemm(const PatternP& m, const PatternP& a, const std::vector<int8_t>* sign = nullptr,
                uint64_t sign_id = 0) {
    const std::string key = sym_key(sign ? "spgemm_s" : "spgemm", m->id, a->id, sign_id);
    if (SymP e = sym_lookup(key)) return e;
    EQ_REQUIRE(m->cols == a->rows, "Error in CudaMatrix * CudaMatrix size checking: lhs.cols_ = "
                                       << m->cols << " rhs.rows_ = " << a->rows);
    const int rows = m->rows;
    auto ucnt = dev_alloc<int>(rows + 1), ccnt = dev_alloc<int>(rows + 1), over = dev_alloc<int>(1);
    launch_fill_int(over->p, 0, 1);
    launch_spgemm_count(rows, m->d_rp->p, m->d_ci->p, a->d_rp->p, a->d_ci->p, SPGEMM_CMAX, ucnt->p, ccnt->p, over->p);
    std::shared_ptr<SymEntry> out;
    if (dev_to_host(over->p, 1)[0] != 0) {
        out = spgemm_host(m, a, sign);
    } else {
        auto rp = dev_alloc<int>(rows + 1), cp = dev_alloc<int>(rows + 1);
        exclusive_scan(ucnt->p, rp->p, rows + 1);
        exclusive_scan(ccnt->p, cp->p, rows + 1);
        const int nnz = read_last(rp, rows), nc = read_last(cp, rows);
        auto e = std::make_shared<SymEntry>();
        auto ci = dev_alloc<int>(std::max(1, nnz));
        e->seg = dev_alloc<int>(nnz + 1);
        launch_fill_int(e->seg->p + nnz, nc, 1);
        DevPtr<int8_t> dsign;
        if (sign) {
            dsign = dev_from_host(sign->data(), sign->size());
            e->sai = dev_alloc<int>(std::max(1, nc));
        } else {
            e->mi = dev_alloc<int>(std::max(1, nc));
            e->ai = dev_alloc<int>(std::max(1, nc));
        }
        launch_spgemm_fill(rows, m->d_rp->p, m->d_ci->p, a->d_rp->p, a->d_ci->p, dsign ? dsign->p : nullptr, rp->p,
                           cp->p, ci->p, e->seg->p, e->mi ? e->mi->p : nullptr, e->ai ? e->ai->p : nullptr,
                           e->sai ? e->sai->p : nullptr);
        e->out = make_pattern_dev(rows, a->cols, nnz, rp, ci, true);
        out = e;
    }
    out->inputs = {m, a};
    sym_store(key, out);
    return out;
}

// Row gather of a pattern: output row i = input row map[i] (map: device array, -1 =>
// empty row), in two passes on the device.
SymP sym_rows(const PatternP& a, const int* map, uint64_t map_id, int out_rows) {
    const std::string key = sym_key("rows", a->id, map_id);
    if (SymP e = sym_lookup(key)) return e;
    auto cnt = dev_alloc<int>(out_rows + 1);
    launch_rows_count(out_rows, a->d_rp->p, map, cnt->p);
    auto rp = dev_alloc<int>(out_rows + 1);
    exclusive_scan(cnt->p, rp->p, out_rows + 1);
    const int nnz = read_last(rp, out_rows);
    auto ci = dev_alloc<int>(std::max(1, nnz)), src = dev_alloc<int>(std::max(1, nnz));
    launch_rows_fill(out_rows, a->d_rp->p, a->d_ci->p, map, rp->p, ci->p, src->p);
    auto e = std::make_shared<SymEntry>();
    e->out = make_pattern_dev(out_rows, a->cols, nnz, rp, ci, a->ascending());
    e->ia = src;
    e->inputs = {a};
    sym_store(key, e);
    return e;
}

// A Jacobian: pattern, stored values and the factors still to apply (jscale, jrow).
struct JacOut {
    PatternP pat;
    DevPtr<double> val;
    double s = 1.0;
    DevPtr<double> row;
    bool ones = false;
    DevPtr<int> map;
    uint64_t map_id = 0;
};

SP make_ad(int n, DevPtr<double> val, const JacOut& j) {
    auto s = std::make_shared<ScalarData>();
    s->n = n;
    s->val = std::move(val);
    s->ad = true;
    s->pat = j.pat;
    s->jval = j.val;
    s->jscale = j.s;
    s->jrow = j.row;
    s->jones = j.ones;
    s->jmap = j.map;
    s->jmap_id = j.map_id;
    return s;
}

// Id of map `which` (0: ia, 1: ib) of a symbolic entry.
uint64_t map_id(const SymEntry& e, int which) { return (e.id << 1) | (uint64_t)which; }

// The composition inner[outer[k]] (-1 where outer is -1) of two entry maps, cached by
// their ids; `pats` are the patterns the maps belong to (the entry dies with them).
std::pair<DevPtr<int>, uint64_t> compose_maps(const DevPtr<int>& outer, uint64_t outer_id, long n,
                                              const DevPtr<int>& inner, uint64_t inner_id,
                                              std::vector<std::weak_ptr<const Pattern>> pats) {
    const std::string key = sym_key("compose", outer_id, inner_id);
    if (SymP e = sym_lookup(key)) return {e->ia, map_id(*e, 0)};
    auto e = std::make_shared<SymEntry>();
    e->ia = dev_alloc<int>(std::max(1L, n));
    if (n > 0) launch_gather_int(inner->p, outer->p, e->ia->p, (int)n);
    e->inputs = std::move(pats);
    sym_store(key, e);
    return {e->ia, map_id(*e, 0)};
}

// The index into x's stored values of every entry of pattern `out`, whose entries map to
// x's (logical) entries by `ia` (id ia_id; null: the same pattern; -1: absent), with x's
// own entry map composed in; null when neither maps.
const int* value_index(const SP& x, const DevPtr<int>& ia, uint64_t ia_id, const PatternP& out) {
    if (!x->jmap) return ia ? ia->p : nullptr;
    if (!ia) return x->jmap->p;
    return compose_maps(ia, ia_id, (long)out->nnz, x->jmap, x->jmap_id, {out, x->pat}).first->p;
}

// The elementwise product of the given row-factor arrays (null entries skipped; null if
// none): one array is shared, more are multiplied into a fresh one.
DevPtr<double> row_product(std::initializer_list<DevPtr<double>> fs, int n) {
    DevPtr<double> out;
    for (const DevPtr<double>& f : fs) {
        if (!f) continue;
        if (!out) {
            out = f;
            continue;
        }
        auto m = dev_alloc<double>(std::max(1, n));
        launch_ew(EW_MUL, out->p, f->p, m->p, n);
        out = m;
    }
    return out;
}

const double* rowp(const DevPtr<double>& r) { return r ? r->p : nullptr; }

// out = C(r) * (alpha*A(r)*Ja + beta*B(r)*Jb) with the empty-matrix rules of the
// reference (cudaMatrix.cu:646-694): empty operands drop out.  With two operands one
// kernel applies every factor (the operands' pending ones included) over the union
// pattern; with one, its values are shared and the factors stay pending.
JacOut jac_lincomb(const SP& a, const DevPtr<double>& A, double alpha, const SP& b, const DevPtr<double>& B,
                   double beta, const DevPtr<double>& C) {
    const PatternP pa = a ? a->pat : nullptr;
    const PatternP pb = b ? b->pat : nullptr;
    JacOut o;
    if (!pa && !pb) return o;
    if (pa && pb) {
        const int *ia = nullptr, *ib = nullptr;
        if (pa.get() != pb.get()) {
            SymP e = sym_union(pa, pb);
            o.pat = e->out;
            ia = value_index(a, e->ia, map_id(*e, 0), o.pat);   // null map: identity
            ib = value_index(b, e->ib, map_id(*e, 1), o.pat);
        } else {
            o.pat = pa;
            ia = value_index(a, nullptr, 0, o.pat);
            ib = value_index(b, nullptr, 0, o.pat);
        }
        o.val = dev_alloc<double>(o.pat->nnz);
        const bool rows = A || B || C || a->jrow || b->jrow;
        launch_jlin(o.pat->nnz, rows ? o.pat->row_of() : nullptr, a->jval->p, ia, rowp(A), rowp(a->jrow),
                    alpha * a->jscale, b->jval->p, ib, rowp(B), rowp(b->jrow), beta * b->jscale, rowp(C), o.val->p);
        return o;
    }
    const SP& x = pa ? a : b;
    o.pat = x->pat;
    o.val = x->jval;
    o.map = x->jmap;
    o.map_id = x->jmap_id;
    o.s = (pa ? alpha : beta) * x->jscale;
    o.row = row_product({C, pa ? A : B, x->jrow}, x->n);
    o.ones = x->jones;
    return o;
}

SP binary_values(int op, const SP& a, const SP& b) {
    auto v = dev_alloc<double>(a->n);
    launch_ew(op, a->val->p, b->val->p, v->p, a->n);
    return make_scalar(a->n, v, false, nullptr, nullptr);
}

}  // namespace

// ===========================================================================
SP s_from_host(const double* h, int n) { return make_scalar(n, dev_from_host(h, n), false, nullptr, nullptr); }

SP s_uniform(int n, double v) {
    auto d = dev_alloc<double>(n);
    launch_fill(d->p, v, n);
    auto s = std::make_shared<ScalarData>();
    s->n = n;
    s->val = d;
    s->uniform = true;
    s->uvalue = v;
    return s;
}

SP s_from_device(DevPtr<double> v, int n) { return make_scalar(n, std::move(v), false, nullptr, nullptr); }

SP s_primary(const SP& a) {
    tape_read(*a, "a primary variable made inside the residual");
    JacOut j;
    j.pat = identity_pattern(a->n);
    j.val = dev_alloc<double>(a->n);
    launch_fill(j.val->p, 1.0, a->n);
    j.ones = true;
    return make_ad(a->n, a->val, j);
}

SP s_with_jacobian(const double* v, int n, const int* rp, const int* ci, const double* jv, int cols) {
    const int nnz = rp[n];
    PatternP p = make_pattern(n, cols, std::vector<int>(rp, rp + n + 1), std::vector<int>(ci, ci + nnz));
    return make_scalar(n, dev_from_host(v, n), true, p, dev_from_host(jv, nnz));
}

SP s_value_only(const SP& a) {
    if (Tape* tp = tape_for(*a)) EQ_TAPE_RETURN(tp, tp->node(T_VAL, tp->operand(a)), a->n, false, s_value_only(a));
    return make_scalar(a->n, a->val, false, nullptr, nullptr);
}

SP s_add(const SP& a, const SP& b) {
    check_size(a, b, "+");
    if (Tape* tp = tape_for(*a, *b))
        EQ_TAPE_RETURN(tp, tp->node(T_ADD, tp->operand(a), tp->operand(b)), a->n, a->ad || b->ad, s_add(a, b));
    SP v = binary_values(EW_ADD, a, b);
    if (!(a->ad || b->ad)) return v;
    return make_ad(a->n, v->val, jac_lincomb(a, nullptr, 1.0, b, nullptr, 1.0, nullptr));
}

SP s_sub(const SP& a, const SP& b) {
    check_size(a, b, "-");
    if (Tape* tp = tape_for(*a, *b))
        EQ_TAPE_RETURN(tp, tp->node(T_SUB, tp->operand(a), tp->operand(b)), a->n, a->ad || b->ad, s_sub(a, b));
    SP v = binary_values(EW_SUB, a, b);
    if (!(a->ad || b->ad)) return v;
    return make_ad(a->n, v->val, jac_lincomb(a, nullptr, 1.0, b, nullptr, -1.0, nullptr));
}

SP s_mul(const SP& a, const SP& b) {
    check_size(a, b, "*");
    if (Tape* tp = tape_for(*a, *b))
        EQ_TAPE_RETURN(tp, tp->node(T_MUL, tp->operand(a), tp->operand(b)), a->n, a->ad || b->ad, s_mul(a, b));
    SP v = binary_values(EW_MUL, a, b);
    if (!(a->ad || b->ad)) return v;
    // (uv)' = diag(v) u' + diag(u) v'
    return make_ad(a->n, v->val, jac_lincomb(a, b->val, 1.0, b, a->val, 1.0, nullptr));
}

SP s_div(const SP& a, const SP& b) {
    check_size(a, b, "/");
    if (Tape* tp = tape_for(*a, *b))
        EQ_TAPE_RETURN(tp, tp->node(T_DIV, tp->operand(a), tp->operand(b)), a->n, a->ad || b->ad, s_div(a, b));
    SP v = binary_values(EW_DIV, a, b);
    if (!(a->ad || b->ad)) return v;
    // (u/v)' = diag(1/v^2) (diag(v) u' - diag(u) v')
    auto c = dev_alloc<double>(a->n);
    launch_coef(4, nullptr, b->val->p, 0.0, c->p, a->n);
    return make_ad(a->n, v->val, jac_lincomb(a, b->val, 1.0, b, a->val, -1.0, c));
}

SP s_scal_mul(double s, const SP& a) {
    if (Tape* tp = tape_for(*a))
        EQ_TAPE_RETURN(tp, tp->node(T_SMUL, tp->operand(a), -1, -1, tp->constant(s)), a->n, a->ad, s_scal_mul(s, a));
    auto v = dev_alloc<double>(a->n);
    launch_ew_scalar(EW_MUL, s, a->val->p, v->p, a->n, 1);
    if (!a->ad) return make_scalar(a->n, v, false, nullptr, nullptr);
    EQ_REQUIRE(a->pat, "Calling CudaMatrix * Scalar with empty matrix...");
    return make_ad(a->n, v, jac_lincomb(a, nullptr, s, nullptr, nullptr, 0.0, nullptr));
}

SP s_div_scal(const SP& a, double s) { return s_scal_mul(1.0 / s, a); }

SP s_scal_div(double s, const SP& a) {
    if (Tape* tp = tape_for(*a))
        EQ_TAPE_RETURN(tp, tp->node(T_SDIV, tp->operand(a), -1, -1, tp->constant(s)), a->n, a->ad, s_scal_div(s, a));
    auto v = dev_alloc<double>(a->n);
    launch_ew_scalar(EW_DIV, s, a->val->p, v->p, a->n, 1);
    if (!a->ad) return make_scalar(a->n, v, false, nullptr, nullptr);
    // (s/u)' = -diag(s/u^2) u'
    auto c = dev_alloc<double>(a->n);
    launch_coef(2, a->val->p, nullptr, s, c->p, a->n);
    return make_ad(a->n, v, jac_lincomb(a, c, 1.0, nullptr, nullptr, 0.0, nullptr));
}

SP s_neg(const SP& a) { return s_scal_mul(-1.0, a); }

SP s_norm(const SP& a) {
    if (Tape* tp = tape_for(*a)) EQ_TAPE_RETURN(tp, tp->node(T_NORM, tp->operand(a)), a->n, a->ad, s_norm(a));
    // factors (+1 where a >= 0, else -1) times a, as a constant times a collection
    // (EquelleRuntimeCPU.cpp:241-254): values |a|, Jacobian diag(factors) a' (a's pattern)
    auto c = dev_alloc<double>(std::max(1, a->n));
    launch_coef(5, a->val->p, nullptr, 0.0, c->p, a->n);
    return s_mul(make_scalar(a->n, c, false, nullptr, nullptr), a);
}

SP s_sqrt(const SP& a) {
    if (Tape* tp = tape_for(*a)) EQ_TAPE_RETURN(tp, tp->node(T_SQRT, tp->operand(a)), a->n, a->ad, s_sqrt(a));
    auto v = dev_alloc<double>(a->n);
    launch_sqrt(a->val->p, v->p, a->n);
    if (!a->ad) return make_scalar(a->n, v, false, nullptr, nullptr);
    auto c = dev_alloc<double>(a->n);
    launch_coef(3, v->p, nullptr, 0.0, c->p, a->n);   // 1/(2 sqrt(x))
    return make_ad(a->n, v, jac_lincomb(a, c, 1.0, nullptr, nullptr, 0.0, nullptr));
}

BP s_compare(int op, const SP& a, const SP& b) {
    check_size(a, b, " comparison");
    if (Tape* tp = tape_for(*a, *b)) {
        const int id = tp->node(T_CMP, tp->operand(a), tp->operand(b), -1, -1, op);
        if (tp->symbolic()) return tp->symbolic_bool(id, a->n);
        BP r;
        {
            TapeSuspend ts;
            r = s_compare(op, a, b);
        }
        tp->tag(*r, id);
        return r;
    }
    auto o = std::make_shared<BoolData>();
    o->n = a->n;
    o->val = dev_alloc<uint8_t>(a->n);
    launch_compare(op, a->val->p, b->val->p, 0.0, 0, o->val->p, a->n);
    return o;
}

BP s_compare_scalar(int op, const SP& a, double s, int scalar_left) {
    if (Tape* tp = tape_for(*a)) {
        const int id = tp->node(T_CMPS, tp->operand(a), -1, -1, tp->constant(s), op | (scalar_left ? 16 : 0));
        if (tp->symbolic()) return tp->symbolic_bool(id, a->n);
        BP r;
        {
            TapeSuspend ts;
            r = s_compare_scalar(op, a, s, scalar_left);
        }
        tp->tag(*r, id);
        return r;
    }
    auto o = std::make_shared<BoolData>();
    o->n = a->n;
    o->val = dev_alloc<uint8_t>(a->n);
    launch_compare(op, a->val->p, nullptr, s, scalar_left, o->val->p, a->n);
    return o;
}

SP s_select(const BP& p, const SP& t, const SP& f) {
    EQ_REQUIRE(p->n == t->n && p->n == f->n, "Collections are not of the same size");
    if (Tape* tp = tape_current())
        if (tp->node_of(*p) >= 0 || tp->node_of(*t) >= 0 || tp->node_of(*f) >= 0)
            EQ_TAPE_RETURN(tp, tp->node(T_SEL, tp->operand(p), tp->operand(t), tp->operand(f)), t->n, t->ad || f->ad,
                           s_select(p, t, f));
    auto v = dev_alloc<double>(t->n);
    launch_select(p->val->p, t->val->p, f->val->p, v->p, t->n);
    if (!(t->ad || f->ad)) return make_scalar(t->n, v, false, nullptr, nullptr);
    // D(p) A' + (I - D(p)) B': union of both patterns regardless of p, zeros kept.
    const PatternP pa = t->pat, pb = f->pat;
    if (!pa && !pb) return make_scalar(t->n, v, true, nullptr, nullptr);
    PatternP out;
    const int *ia = nullptr, *ib = nullptr;
    SymP e;
    if (pa && pb && pa.get() != pb.get()) {
        e = sym_union(pa, pb);
        out = e->out;
        ia = value_index(t, e->ia, map_id(*e, 0), out);
        ib = value_index(f, e->ib, map_id(*e, 1), out);
    } else {
        out = pa ? pa : pb;
        if (pa) ia = value_index(t, nullptr, 0, out);
        if (pb) ib = value_index(f, nullptr, 0, out);
    }
    auto j = dev_alloc<double>(out->nnz);
    launch_jselect(out->nnz, out->row_of(), p->val->p, pa ? t->jval->p : nullptr, ia, pa ? t->jscale : 1.0,
                   pa ? rowp(t->jrow) : nullptr, pb ? f->jval->p : nullptr, ib, pb ? f->jscale : 1.0,
                   pb ? rowp(f->jrow) : nullptr, j->p);
    return make_scalar(t->n, v, true, out, j);
}

double s_reduce(int op, const SP& a) {
    tape_read(*a, "a reduction");
    EQ_REQUIRE(a->n > 0 || op == RED_SUM || op == RED_PROD, "reduction of an empty collection");
    return reduce(op, a->val->p, a->n);
}

double s_two_norm(const SP& a) {
    tape_read(*a, "a norm");
    return std::sqrt(reduce(RED_SUMSQ, a->val->p, a->n)); }

std::vector<double> s_to_host(const SP& a) {
    tape_read(*a, "a copy to the host");
    return dev_to_host(a->val->p, a->n);
}

// ---------------------------------------------------------------- matrices
DevCsr csr_from_host(int rows, int cols, const int* rp, const int* ci, const double* v) {
    DevCsr m;
    m.pat = make_pattern(rows, cols, std::vector<int>(rp, rp + rows + 1), std::vector<int>(ci, ci + rp[rows]));
    m.val = dev_from_host(v, rp[rows]);
    const int nnz = rp[rows];
    bool unit = nnz > 0;
    for (int k = 0; k < nnz && unit; ++k) unit = v[k] == 1.0 || v[k] == -1.0;
    if (unit) {   // a signed incidence matrix (the grid's divergence and gradient operators)
        auto sg = std::make_shared<std::vector<int8_t>>(nnz);
        for (int k = 0; k < nnz; ++k) (*sg)[k] = v[k] > 0.0 ? 1 : -1;
        m.sign = sg;
        m.sign_id = next_id();
    }
    return m;
}

SP s_matmul(const DevCsr& m, const SP& a, DevPtr<double> values) {
    EQ_REQUIRE(m.pat->cols == a->n, "Error in matrix * vector operation as matrix is of size "
                                        << m.pat->rows << " by " << m.pat->cols << " and the vector of size " << a->n);
    if (Tape* tp = tape_for(*a))
        EQ_TAPE_RETURN(tp, tp->node(T_MAT, tp->operand(a), -1, -1, tp->matrix(m)), m.pat->rows, a->ad,
                       s_matmul(m, a, values));
    DevPtr<double> v = values;
    if (!v) {
        v = dev_alloc<double>(m.pat->rows);
        launch_csr_spmv(m.pat->rows, m.pat->d_rp->p, m.pat->d_ci->p, m.val->p, a->val->p, v->p);
    }
    if (!a->ad || !a->pat) return make_scalar(m.pat->rows, v, a->ad, nullptr, nullptr);
    JacOut o;
    o.s = a->jscale;
    if (a->jones && !a->jrow && m.pat->ascending()) {
        // M * (s I) = s M: the structural product is M's own pattern, one contribution per entry
        o.pat = m.pat;
        o.val = m.val;
        return make_ad(m.pat->rows, v, o);
    }
    SymP e = sym_spgemm(m.pat, a->pat, m.sign.get(), m.sign_id);
    o.pat = e->out;
    o.val = dev_alloc<double>(e->out->nnz);
    o.s = 1.0;
    DevPtr<double> av = a->jval;
    if (a->jmap) {   // the operand's values in entry order
        av = dev_alloc<double>(std::max(1, a->pat->nnz));
        launch_gather(a->jval->p, a->jmap->p, av->p, a->pat->nnz);
    }
    if (m.sign)   // M's entries are +-1: the signs are folded into the operand indices
        launch_spgemm_signed(e->out->nnz, e->seg->p, e->sai->p, av->p, a->jrow ? a->pat->row_of() : nullptr,
                             rowp(a->jrow), a->jscale, o.val->p);
    else
        launch_spgemm_num(e->out->nnz, e->seg->p, e->mi->p, e->ai->p, m.val->p, av->p,
                          a->jrow ? a->pat->row_of() : nullptr, rowp(a->jrow), a->jscale, o.val->p);
    return make_ad(m.pat->rows, v, o);
}

DevCsr csr_add(const DevCsr& a, double alpha, const DevCsr& b, double beta) {
    if (!a.pat && !b.pat) return DevCsr();
    DevCsr o;
    if (!a.pat || !b.pat || a.pat.get() == b.pat.get()) {
        o.pat = a.pat ? a.pat : b.pat;
        o.val = dev_alloc<double>(o.pat->nnz);
        launch_jlin(o.pat->nnz, nullptr, a.pat ? a.val->p : nullptr, nullptr, nullptr, nullptr, alpha,
                    b.pat ? b.val->p : nullptr, nullptr, nullptr, nullptr, beta, nullptr, o.val->p);
        return o;
    }
    SymP e = sym_union(a.pat, b.pat);
    o.pat = e->out;
    o.val = dev_alloc<double>(o.pat->nnz);
    launch_jlin(o.pat->nnz, nullptr, a.val->p, e->ia ? e->ia->p : nullptr, nullptr, nullptr, alpha, b.val->p,
                e->ib ? e->ib->p : nullptr, nullptr, nullptr,
                beta, nullptr, o.val->p);
    return o;
}

DevCsr csr_mul(const DevCsr& a, const DevCsr& b) {
    if (!a.pat || !b.pat) return DevCsr();
    SymP e = sym_spgemm(a.pat, b.pat);
    DevCsr o;
    o.pat = e->out;
    o.val = dev_alloc<double>(o.pat->nnz);
    launch_spgemm_num(o.pat->nnz, e->seg->p, e->mi->p, e->ai->p, a.val->p, b.val->p, nullptr, nullptr, 1.0,
                      o.val->p);
    return o;
}

// Row gather (On / Extend) of a Jacobian: the stored values of the selected rows, the
// pending factors carried along (row factors gathered with the rows; -1 => empty row).
static JacOut gather_rows(const SP& a, const SymP& e, const int* rowmap, int m) {
    JacOut o;
    o.pat = e->out;
    o.val = a->jval;   // shared: the selected entries are addressed through the entry map
    if (!a->jmap) {
        o.map = e->ia;
        o.map_id = map_id(*e, 0);
    } else {
        std::tie(o.map, o.map_id) =
            compose_maps(e->ia, map_id(*e, 0), (long)e->out->nnz, a->jmap, a->jmap_id, {e->out, a->pat});
    }
    o.s = a->jscale;
    if (a->jrow) {
        o.row = dev_alloc<double>(std::max(1, m));
        launch_gather(a->jrow->p, rowmap, o.row->p, m);
    }
    return o;
}

DevPtr<double> jac_effective(const ScalarData& a) {
    tape_read(a, "a Jacobian read");
    if (!a.pat) return nullptr;
    EQ_REQUIRE(!a.jsink, "this residual's Jacobian was written into the stencil linear solver's layout "
                         "(a fused newtonSolve evaluation); it has no CSR values");
    if (!a.jpending() && !a.jmap) return a.jval;
    if (!a.jeff) {
        auto v = dev_alloc<double>(std::max(1, a.pat->nnz));
        launch_jscale(a.pat->nnz, a.jrow ? a.pat->row_of() : nullptr, a.jscale, rowp(a.jrow),
                      a.jmap ? a.jmap->p : nullptr, a.jval->p, v->p);
        a.jeff = v;
    }
    return a.jeff;
}
const double* jac_values(const ScalarData& a) {
    DevPtr<double> v = jac_effective(a);
    return v ? v->p : nullptr;
}

// ---------------------------------------------------------------- index sets
PMP position_map(const IP& super, const IP& sub) {
    static std::mutex mu;
    static std::map<std::pair<uint64_t, uint64_t>, std::weak_ptr<const PosMap>> cache;
    const auto key = std::make_pair(super->id, sub->id);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end())
            if (PMP p = it->second.lock()) return p;
    }
    // subsetIndices (EquelleRuntimeCPU_impl.hpp:28-83): superset sorted and unique.
    const std::vector<int>& sup = super->host;
    const std::vector<int>& s = sub->host;
    std::vector<int> pos(s.size(), -1);
    if (super->full) {   // super holds 0..n-1 in order: the position of an element is its value
        const int ns = (int)sup.size();
        for (size_t i = 0; i < s.size(); ++i) {
            EQ_REQUIRE(s[i] >= 0 && s[i] < ns, "set of size " << s.size() << " is not a subset of the given superset (size "
                                                             << ns << ")");
            pos[i] = s[i];
        }
    } else if (!s.empty()) {
        std::vector<std::pair<int, int>> si(s.size());
        for (size_t i = 0; i < s.size(); ++i) si[i] = {s[i], (int)i};
        std::sort(si.begin(), si.end());
        size_t k = 0;
        for (size_t i = 0; i < sup.size() && k < si.size(); ++i)
            while (k < si.size() && si[k].first == sup[i]) pos[si[k++].second] = (int)i;
        EQ_REQUIRE(k == si.size(), "set of size " << s.size() << " is not a subset of the given superset (size "
                                                  << sup.size() << ")");
    }
    auto p = std::make_shared<PosMap>();
    p->id = next_id();
    p->super_size = (int)sup.size();
    p->d_pos = dev_from_host(pos.data(), pos.size());
    std::vector<int> inv(sup.size(), -1);
    for (size_t i = 0; i < pos.size(); ++i) inv[pos[i]] = (int)i;
    p->d_inv = dev_from_host(inv.data(), inv.size());
    p->pos = std::move(pos);
    p->inv = std::move(inv);
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = p;
    // keep the maps alive as long as both sets live
    static std::vector<std::pair<std::weak_ptr<const IndexData>, PMP>> owners;
    owners.emplace_back(super, p);
    owners.emplace_back(sub, p);
    if (owners.size() > 8192)
        owners.erase(std::remove_if(owners.begin(), owners.end(), [](auto& x) { return x.first.expired(); }),
                     owners.end());
    return p;
}

SP s_on(const SP& a, const IP& from, const IP& to) {
    EQ_REQUIRE(a->n == from->size(), "data_in (size " << a->n << ") and from_set (size " << from->size()
                                                      << ") have to be of the same size in On function.");
    PMP pm = position_map(from, to);
    const int m = to->size();
    if (Tape* tp = tape_for(*a))
        EQ_TAPE_RETURN(tp, tp->node(T_ON, tp->operand(a), -1, -1, tp->map(pm->d_pos, pm->id, pm)), m, a->ad,
                       s_on(a, from, to));
    auto v = dev_alloc<double>(m);
    launch_gather(a->val->p, pm->d_pos->p, v->p, m);
    if (!a->ad || !a->pat) return make_scalar(m, v, a->ad, nullptr, nullptr);
    SymP e = sym_rows(a->pat, pm->d_pos->p, pm->id, m);
    return make_ad(m, v, gather_rows(a, e, pm->d_pos->p, m));
}

SP s_extend(const SP& a, const IP& from, const IP& to) {
    EQ_REQUIRE(a->n == from->size(), "data_in (size " << a->n << ") and from_set (size " << from->size()
                                                      << ") have to be of the same size in Extend function.");
    EQ_REQUIRE(from->size() <= to->size(), "From_set (size " << from->size() << ") has to be a subset of to_set (size "
                                                             << to->size() << ")");
    PMP pm = position_map(to, from);
    const int m = to->size();
    if (Tape* tp = tape_for(*a))
        EQ_TAPE_RETURN(tp,
                       tp->node(T_EXT, tp->operand(a), -1, -1, tp->map(pm->d_inv, pm->id ^ 0x8000000000000000ull, pm),
                                a->n == 0 ? 1 : 0),
                       m, a->ad, s_extend(a, from, to));
    auto v = dev_alloc<double>(m);
    launch_gather(a->val->p, pm->d_inv->p, v->p, m);
    if (!a->ad || !a->pat) return make_scalar(m, v, a->ad, nullptr, nullptr);
    SymP e = sym_rows(a->pat, pm->d_inv->p, pm->id ^ 0x8000000000000000ull, m);
    return make_ad(m, v, gather_rows(a, e, pm->d_inv->p, m));
}

IP i_from_device(int codim, DevPtr<int> d, int n) {
    auto out = std::make_shared<IndexData>();
    out->id = next_id();
    out->codim = codim;
    out->host = dev_to_host(d->p, n);
    out->dev = std::move(d);
    return out;
}

IP i_on(const IP& data, const IP& from, const IP& to) {
    EQ_REQUIRE(data->size() == from->size(), "data_in(size " << data->size() << ") and from_set (size "
                                                             << from->size()
                                                             << ") have to be of the same size in On function for "
                                                                "CollOfIndices.");
    PMP pm = position_map(from, to);
    auto d = dev_alloc<int>(to->size());
    launch_gather_int(data->dev->p, pm->d_pos->p, d->p, to->size());
    return i_from_device(data->codim, d, to->size());
}

IP i_select(const BP& p, const IP& t, const IP& f) {
    tape_read(*p, "an index selection by a predicate on the unknown");
    EQ_REQUIRE(p->n == t->size() && p->n == f->size(), "The sets are not of the same size");
    auto d = dev_alloc<int>(p->n);
    launch_select_int(p->val->p, t->dev->p, f->dev->p, d->p, p->n);
    return i_from_device(t->codim, d, p->n);
}

BP i_is_empty(const IP& s) {
    auto o = std::make_shared<BoolData>();
    o->n = s->size();
    o->val = dev_alloc<uint8_t>(o->n);
    launch_is_empty(s->dev->p, o->val->p, o->n);
    return o;
}

BP b_from_host(const uint8_t* h, int n) {
    auto o = std::make_shared<BoolData>();
    o->n = n;
    o->val = dev_from_host(h, n);
    return o;
}

std::vector<uint8_t> b_to_host(const BP& b) {
    tape_read(*b, "a copy to the host");
    return dev_to_host(b->val->p, b->n); }

// ---------------------------------------------------------------- vectors
VP v_from_host(const double* h, int n, int dim) {
    auto v = std::make_shared<VectorData>();
    v->n = n;
    v->dim = dim;
    v->val = dev_from_host(h, (size_t)n * dim);
    return v;
}

VP v_binary(int op, const VP& a, const VP& b) {
    EQ_REQUIRE(a->n == b->n && a->dim == b->dim, "CollOfVector sizes differ");
    auto v = std::make_shared<VectorData>();
    v->n = a->n;
    v->dim = a->dim;
    v->val = dev_alloc<double>((size_t)a->n * a->dim);
    launch_vec_ew(op, a->val->p, b->val->p, a->n * a->dim, v->val->p);
    return v;
}

VP v_scale(const VP& a, const SP& s, double sc, int divide) {
    if (s) tape_read(*s, "a vector scaled by the unknown");
    if (s) EQ_REQUIRE(s->n == a->n, "CollOfVector and CollOfScalar sizes differ");
    auto v = std::make_shared<VectorData>();
    v->n = a->n;
    v->dim = a->dim;
    v->val = dev_alloc<double>((size_t)a->n * a->dim);
    launch_vec_scale(a->val->p, s ? s->val->p : nullptr, sc, a->n, a->dim, divide, v->val->p);
    return v;
}

SP v_norm(const VP& a) {
    auto d = dev_alloc<double>(a->n);
    launch_vec_norm(a->val->p, a->n, a->dim, d->p);
    return make_scalar(a->n, d, false, nullptr, nullptr);
}

SP v_dot(const VP& a, const VP& b) {
    EQ_REQUIRE(a->n == b->n && a->dim == b->dim, "Non-matching size of Vector collections for dot().");
    auto d = dev_alloc<double>(a->n);
    launch_vec_dot(a->val->p, b->val->p, a->n, a->dim, d->p);
    return make_scalar(a->n, d, false, nullptr, nullptr);
}

SP v_col(const VP& a, int col) {
    EQ_REQUIRE(col >= 0 && col < a->dim, "CollOfVector index " << col << " out of range");
    auto d = dev_alloc<double>(a->n);
    launch_vec_col(a->val->p, a->n, a->dim, col, d->p);
    return make_scalar(a->n, d, false, nullptr, nullptr);
}

std::vector<double> v_to_host(const VP& a) { return dev_to_host(a->val->p, (size_t)a->n * a->dim); }

}  // namespace eqhip
