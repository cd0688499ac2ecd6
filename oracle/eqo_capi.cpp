// eqo C API — handle-based access to the CPU oracle for the test suite.
//
// TEST INFRASTRUCTURE ONLY (see eqo.hpp).  The entry points have the same
// names after the prefix and the same argument meaning as the HIP runtime's
// C-ABI (include/eqhip.h), so one Python driver can replay an Equelle op
// sequence on both and compare results.  Return value: 0 = ok, nonzero =
// error; the message is available from eqo_last_error().
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "eqo.hpp"
#include "eqo_programs.hpp"

using namespace eqo;

namespace {
thread_local std::string g_err;

struct Rt {
    Params params;
    std::unique_ptr<Runtime> rt;
    std::vector<std::unique_ptr<ADB>> scal;
    std::vector<std::unique_ptr<IndexSet>> sets;
    std::vector<std::unique_ptr<std::vector<uint8_t>>> bools;
    template <class T>
    static int put(std::vector<std::unique_ptr<T>>& v, T&& x) {
        for (size_t i = 0; i < v.size(); ++i)
            if (!v[i]) { v[i].reset(new T(std::move(x))); return (int)i + 1; }
        v.emplace_back(new T(std::move(x)));
        return (int)v.size();
    }
    template <class T>
    static T& get(std::vector<std::unique_ptr<T>>& v, int h) {
        if (h < 1 || h > (int)v.size() || !v[h - 1]) throw std::runtime_error("invalid handle " + std::to_string(h));
        return *v[h - 1];
    }
    ADB& S(int h) { return get(scal, h); }
    IndexSet& I(int h) { return get(sets, h); }
    std::vector<uint8_t>& B(int h) { return get(bools, h); }
};

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    } catch (...) {
        g_err = "unknown error";
        return 1;
    }
}
Rt* R(void* p) {
    if (!p) throw std::runtime_error("null runtime");
    return static_cast<Rt*>(p);
}
}  // namespace

extern "C" {

const char* eqo_last_error(void) { return g_err.c_str(); }

int eqo_runtime_create(const char* params, void** out) {
    return guard([&] {
        auto* r = new Rt();
        r->params.parse(params ? params : "");
        r->rt.reset(new Runtime(r->params));
        *out = r;
    });
}

int eqo_runtime_destroy(void* rt) {
    delete static_cast<Rt*>(rt);
    return 0;
}

int eqo_grid_info(void* rt, int* info) {
    return guard([&] {
        const Grid& g = R(rt)->rt->grid();
        info[0] = g.dim;
        info[1] = g.nc;
        info[2] = g.nf;
        info[3] = (int)R(rt)->rt->ops().internal_faces.size();
        info[4] = g.nf - info[3];
    });
}

// Raw grid arrays: which = 0 face_cells (2nf int), 1 cell_facepos (nc+1 int), 2 cell_faces (int)
int eqo_grid_int_array(void* rt, int which, int* out) {
    return guard([&] {
        const Grid& g = R(rt)->rt->grid();
        const std::vector<int>& v = which == 0 ? g.face_cells : which == 1 ? g.cell_facepos : g.cell_faces;
        std::memcpy(out, v.data(), v.size() * sizeof(int));
    });
}
// which = 0 cell_centroids, 1 cell_volumes, 2 face_centroids, 3 face_areas, 4 face_normals (area-scaled)
int eqo_grid_double_array(void* rt, int which, double* out) {
    return guard([&] {
        const Grid& g = R(rt)->rt->grid();
        const std::vector<double>* v[] = {&g.cell_centroids, &g.cell_volumes, &g.face_centroids, &g.face_areas,
                                          &g.face_normals};
        std::memcpy(out, v[which]->data(), v[which]->size() * sizeof(double));
    });
}

// HelperOps matrices: which = 0 grad, 1 div, 2 fulldiv, 3 ngrad.  Query sizes with
// eqo_opmat_info (rows, cols, nnz) then copy.
static const Csr& opmat(Rt* r, int which) {
    const HelperOps& h = r->rt->ops();
    switch (which) {
        case 0: return h.grad;
        case 1: return h.div;
        case 2: return h.fulldiv;
        default: return h.ngrad;
    }
}
int eqo_opmat_info(void* rt, int which, int* info) {
    return guard([&] {
        const Csr& m = opmat(R(rt), which);
        info[0] = m.rows;
        info[1] = m.cols;
        info[2] = m.nnz();
    });
}
int eqo_opmat_copy(void* rt, int which, int* rp, int* ci, double* v) {
    return guard([&] {
        const Csr& m = opmat(R(rt), which);
        std::memcpy(rp, m.rp.data(), m.rp.size() * sizeof(int));
        std::memcpy(ci, m.ci.data(), m.ci.size() * sizeof(int));
        std::memcpy(v, m.v.data(), m.v.size() * sizeof(double));
    });
}

// ---------------- index sets ----------------
// which: 0 allCells 1 boundaryCells 2 interiorCells 3 allFaces 4 boundaryFaces 5 interiorFaces
int eqo_set_builtin(void* rt, int which, int* h) {
    return guard([&] {
        Runtime& x = *R(rt)->rt;
        IndexSet s;
        switch (which) {
            case 0: s = x.allCells(); break;
            case 1: s = x.boundaryCells(); break;
            case 2: s = x.interiorCells(); break;
            case 3: s = x.allFaces(); break;
            case 4: s = x.boundaryFaces(); break;
            case 5: s = x.interiorFaces(); break;
            default: throw std::runtime_error("unknown builtin set");
        }
        *h = Rt::put(R(rt)->sets, std::move(s));
    });
}
int eqo_set_from_host(void* rt, int codim, const int* idx, int n, int* h) {
    return guard([&] {
        IndexSet s;
        s.codim = codim;
        s.idx.assign(idx, idx + n);
        *h = Rt::put(R(rt)->sets, std::move(s));
    });
}
int eqo_set_size(void* rt, int h, int* n) {
    return guard([&] { *n = R(rt)->I(h).size(); });
}
int eqo_set_copy(void* rt, int h, int* out) {
    return guard([&] {
        const IndexSet& s = R(rt)->I(h);
        std::memcpy(out, s.idx.data(), s.idx.size() * sizeof(int));
    });
}
int eqo_first_cell(void* rt, int faces, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->sets, R(rt)->rt->firstCell(R(rt)->I(faces))); });
}
int eqo_second_cell(void* rt, int faces, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->sets, R(rt)->rt->secondCell(R(rt)->I(faces))); });
}
int eqo_set_on(void* rt, int data, int from, int to, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->sets, r->rt->operatorOn(r->I(data), r->I(from), r->I(to)));
    });
}
int eqo_set_trinary(void* rt, int pred, int t, int f, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->sets, r->rt->trinaryIf(r->B(pred), r->I(t), r->I(f)));
    });
}
int eqo_is_empty(void* rt, int set, int* pred) {
    return guard([&] { *pred = Rt::put(R(rt)->bools, R(rt)->rt->isEmpty(R(rt)->I(set))); });
}
int eqo_set_contains(void* rt, int super, int sub) {
    return guard([&] { subset_indices(R(rt)->I(super), R(rt)->I(sub)); });
}

// ---------------- booleans ----------------
int eqo_bool_size(void* rt, int h, int* n) {
    return guard([&] { *n = (int)R(rt)->B(h).size(); });
}
int eqo_bool_copy(void* rt, int h, unsigned char* out) {
    return guard([&] {
        const auto& b = R(rt)->B(h);
        std::memcpy(out, b.data(), b.size());
    });
}
// op: 0 >, 1 >=, 2 <, 3 <=, 4 ==, 5 !=
static bool cmp(int op, double a, double b) {
    switch (op) {
        case 0: return a > b;
        case 1: return a >= b;
        case 2: return a < b;
        case 3: return a <= b;
        case 4: return a == b;
        default: return a != b;
    }
}
int eqo_compare(void* rt, int op, int a, int b, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        const ADB &x = r->S(a), &y = r->S(b);
        if (x.size() != y.size()) throw std::runtime_error("compare: size mismatch");
        std::vector<uint8_t> o(x.size());
        for (int i = 0; i < x.size(); ++i) o[i] = cmp(op, x.v[i], y.v[i]);
        *h = Rt::put(r->bools, std::move(o));
    });
}
// scalar_left: 0 => a OP s, 1 => s OP a
int eqo_compare_scalar(void* rt, int op, int a, double s, int scalar_left, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        const ADB& x = r->S(a);
        std::vector<uint8_t> o(x.size());
        for (int i = 0; i < x.size(); ++i) o[i] = scalar_left ? cmp(op, s, x.v[i]) : cmp(op, x.v[i], s);
        *h = Rt::put(r->bools, std::move(o));
    });
}

// ---------------- collections of scalars ----------------
int eqo_scalar_from_host(void* rt, const double* v, int n, int* h) {
    return guard([&] {
        ADB a;
        a.v.assign(v, v + n);
        *h = Rt::put(R(rt)->scal, std::move(a));
    });
}
int eqo_scalar_uniform(void* rt, int n, double s, int* h) {
    return guard([&] {
        ADB a;
        a.v.assign(n, s);
        *h = Rt::put(R(rt)->scal, std::move(a));
    });
}
int eqo_scalar_with_jacobian(void* rt, const double* v, int n, const int* rp, const int* ci, const double* jv,
                             int cols, int* h) {
    return guard([&] {
        ADB a;
        a.v.assign(v, v + n);
        a.ad = true;
        a.J.rows = n;
        a.J.cols = cols;
        a.J.rp.assign(rp, rp + n + 1);
        a.J.ci.assign(ci, ci + rp[n]);
        a.J.v.assign(jv, jv + rp[n]);
        *h = Rt::put(R(rt)->scal, std::move(a));
    });
}
int eqo_primary(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, primary_variable(R(rt)->S(a))); });
}
// info = {size, autodiff flag, jac rows, jac cols, jac nnz}; empty Jacobian => rows = cols = nnz = 0
int eqo_scalar_info(void* rt, int h, int* info) {
    return guard([&] {
        const ADB& a = R(rt)->S(h);
        info[0] = a.size();
        info[1] = a.ad ? 1 : 0;
        info[2] = a.J.empty() ? 0 : a.J.rows;
        info[3] = a.J.empty() ? 0 : a.J.cols;
        info[4] = a.J.nnz();
    });
}
int eqo_scalar_copy(void* rt, int h, double* out) {
    return guard([&] {
        const ADB& a = R(rt)->S(h);
        std::memcpy(out, a.v.data(), a.v.size() * sizeof(double));
    });
}
int eqo_jacobian_copy(void* rt, int h, int* rp, int* ci, double* v) {
    return guard([&] {
        const Csr& m = R(rt)->S(h).J;
        if (m.empty()) throw std::runtime_error("Trying to copy empty matrix to host");
        std::memcpy(rp, m.rp.data(), m.rp.size() * sizeof(int));
        std::memcpy(ci, m.ci.data(), m.ci.size() * sizeof(int));
        std::memcpy(v, m.v.data(), m.v.size() * sizeof(double));
    });
}
int eqo_release(void* rt, int kind, int h) {
    return guard([&] {
        Rt* r = R(rt);
        if (kind == 0) { r->S(h); r->scal[h - 1].reset(); }
        else if (kind == 1) { r->I(h); r->sets[h - 1].reset(); }
        else { r->B(h); r->bools[h - 1].reset(); }
    });
}
// op: 0 +, 1 -, 2 *, 3 /
int eqo_binop(void* rt, int op, int a, int b, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        const ADB &x = r->S(a), &y = r->S(b);
        ADB o = op == 0 ? adb_add(x, y) : op == 1 ? adb_sub(x, y) : op == 2 ? adb_mul(x, y) : adb_div(x, y);
        *h = Rt::put(r->scal, std::move(o));
    });
}
// op: 2 * (s*a or a*s), 3 / (a/s if scalar_left=0, s/a if 1)
int eqo_scalar_binop(void* rt, int op, double s, int a, int scalar_left, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        const ADB& x = r->S(a);
        ADB o;
        if (op == 2) o = adb_scal_mul(s, x);
        else if (op == 3 && !scalar_left) o = adb_scal_mul(1.0 / s, x);   // (1/rhs)*lhs, collOfScalar.cu:385-387
        else if (op == 3) o = adb_scal_div_coll(s, x);
        else throw std::runtime_error("scalar_binop: unsupported op");
        *h = Rt::put(r->scal, std::move(o));
    });
}
int eqo_negate(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, adb_neg(R(rt)->S(a))); });
}
int eqo_sqrt(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, adb_sqrt(R(rt)->S(a))); });
}
int eqo_trinary(void* rt, int pred, int t, int f, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->scal, r->rt->trinaryIf(r->B(pred), r->S(t), r->S(f)));
    });
}
int eqo_gradient(void* rt, int a, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        if (r->S(a).size() != r->rt->grid().nc) throw std::runtime_error("Gradient need input defined on AllCells()");
        *h = Rt::put(r->scal, r->rt->gradient(r->S(a)));
    });
}
int eqo_divergence(void* rt, int a, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        const int n = r->S(a).size();
        if (n != r->rt->grid().nf && n != (int)r->rt->ops().internal_faces.size())
            throw std::runtime_error("Input for divergence has to be on AllFaces or on InteriorFaces()");
        *h = Rt::put(r->scal, r->rt->divergence(r->S(a)));
    });
}
int eqo_neg_gradient(void* rt, int a, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        if (r->S(a).size() != r->rt->grid().nc) throw std::runtime_error("negGradient needs input defined on AllCells()");
        *h = Rt::put(r->scal, r->rt->negGradient(r->S(a)));
    });
}
int eqo_interior_divergence(void* rt, int a, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        if (r->S(a).size() != (int)r->rt->ops().internal_faces.size())
            throw std::runtime_error("interiorDivergence needs input defined on InteriorFaces()");
        *h = Rt::put(r->scal, r->rt->interiorDivergence(r->S(a)));
    });
}
int eqo_norm_scalar(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, adb_norm(R(rt)->S(a))); });
}
int eqo_extend(void* rt, int a, int from, int to, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->scal, r->rt->operatorExtend(r->S(a), r->I(from), r->I(to)));
    });
}
int eqo_extend_scalar(void* rt, double s, int to, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->scal, r->rt->operatorExtend(s, r->I(to)));
    });
}
int eqo_on(void* rt, int a, int from, int to, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->scal, r->rt->operatorOn(r->S(a), r->I(from), r->I(to)));
    });
}
// kind: 0 sum, 1 product, 2 max, 3 min, 4 twoNorm
int eqo_reduce(void* rt, int kind, int a, double* out) {
    return guard([&] {
        Rt* r = R(rt);
        const ADB& x = r->S(a);
        switch (kind) {
            case 0: *out = r->rt->sumReduce(x); break;
            case 1: *out = r->rt->prodReduce(x); break;
            case 2: *out = r->rt->maxReduce(x); break;
            case 3: *out = r->rt->minReduce(x); break;
            default: *out = r->rt->twoNorm(x); break;
        }
    });
}
int eqo_norm_set(void* rt, int set, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, R(rt)->rt->norm(R(rt)->I(set))); });
}
int eqo_centroid(void* rt, int set, double* out) {
    return guard([&] {
        std::vector<double> c = R(rt)->rt->centroid(R(rt)->I(set));
        std::memcpy(out, c.data(), c.size() * sizeof(double));
    });
}
int eqo_normal(void* rt, int set, double* out) {
    return guard([&] {
        std::vector<double> c = R(rt)->rt->normal(R(rt)->I(set));
        std::memcpy(out, c.data(), c.size() * sizeof(double));
    });
}
// Solve J du = r for a residual with a Jacobian; du is a constant collection.
int eqo_solve_for_update(void* rt, int residual, int* du, int* iters) {
    return guard([&] {
        Rt* r = R(rt);
        LinSolveReport rep;
        ADB d;
        d.v = r->rt->solveForUpdate(r->S(residual), &rep);
        if (iters) *iters = rep.iterations;
        *du = Rt::put(r->scal, std::move(d));
    });
}

// ---------------- CSR algebra (matrix.cpp known answers) ----------------
// op: 0 A+B, 1 A-B, 2 A*B ; output CSR is returned through a scalar handle's Jacobian
// so the caller copies it with eqo_scalar_info / eqo_jacobian_copy.
int eqo_csr_binop(int op, int ar, int ac, const int* arp, const int* aci, const double* av, int br, int bc,
                  const int* brp, const int* bci, const double* bv, int* out_info, int* orp, int* oci, double* ov,
                  int cap) {
    return guard([&] {
        auto mk = [](int r, int c, const int* rp, const int* ci, const double* v) {
            Csr m;
            if (!rp) return m;
            m.rows = r;
            m.cols = c;
            m.rp.assign(rp, rp + r + 1);
            m.ci.assign(ci, ci + rp[r]);
            m.v.assign(v, v + rp[r]);
            return m;
        };
        Csr A = mk(ar, ac, arp, aci, av), B = mk(br, bc, brp, bci, bv);
        Csr C = op == 0 ? csr_add(A, 1.0, B, 1.0) : op == 1 ? csr_add(A, 1.0, B, -1.0) : csr_mul(A, B);
        out_info[0] = C.empty() ? 0 : C.rows;
        out_info[1] = C.empty() ? 0 : C.cols;
        out_info[2] = C.nnz();
        if (C.empty()) return;
        if (C.nnz() > cap) throw std::runtime_error("csr_binop: output capacity too small");
        std::memcpy(orp, C.rp.data(), C.rp.size() * sizeof(int));
        std::memcpy(oci, C.ci.data(), C.ci.size() * sizeof(int));
        std::memcpy(ov, C.v.data(), C.v.size() * sizeof(double));
    });
}

// y = op(A) x for a host CSR matrix (matrix.cpp:360-378 known answers).
int eqo_csr_spmv(int rows, int cols, const int* rp, const int* ci, const double* v, const double* x, double* y,
                 int transpose) {
    return guard([&] {
        Csr a;
        a.rows = rows;
        a.cols = cols;
        a.rp.assign(rp, rp + rows + 1);
        a.ci.assign(ci, ci + rp[rows]);
        a.v.assign(v, v + rp[rows]);
        if (transpose) a = csr_transpose(a);
        std::vector<double> xx(x, x + a.cols);
        std::vector<double> yy = csr_spmv(a, xx);
        std::memcpy(y, yy.data(), yy.size() * sizeof(double));
    });
}

// ---------------- workload programs (heat) ----------------
int eqo_heat_run(const char* params, const double* u_initial, int n, const int* dir_faces, const double* dir_vals,
                 int ndir, const double* dts, int nsteps, double* u_out, int* newton_iters, double* final_norms) {
    return guard([&] {
        Params p;
        p.parse(params ? params : "");
        HeatResult res = run_heat(p, std::vector<double>(u_initial, u_initial + n),
                                  std::vector<int>(dir_faces, dir_faces + ndir),
                                  std::vector<double>(dir_vals, dir_vals + ndir), std::vector<double>(dts, dts + nsteps));
        std::memcpy(u_out, res.u.data(), res.u.size() * sizeof(double));
        for (int s = 0; s < nsteps; ++s) {
            if (newton_iters) newton_iters[s] = res.reports[s].iterations;
            if (final_norms) final_norms[s] = res.reports[s].final_norm;
        }
    });
}

// Two-phase program: nsteps time steps; p_out/sw_out final fields;
// iters[2*s] / iters[2*s+1] = pressure / transport Newton iterations of step s.
int eqo_twophase_run(const char* params, const double* perm, const double* poro, const double* sw_initial, int n,
                     const int* src_cells, const double* src_vals, int nsrc, const double* dts, int nsteps,
                     double* p_out, double* sw_out, int* iters) {
    return guard([&] {
        Params p;
        p.parse(params ? params : "");
        TwoPhaseResult res = run_twophase(p, std::vector<double>(perm, perm + n), std::vector<double>(poro, poro + n),
                                          std::vector<double>(sw_initial, sw_initial + n),
                                          std::vector<int>(src_cells, src_cells + nsrc),
                                          std::vector<double>(src_vals, src_vals + nsrc),
                                          std::vector<double>(dts, dts + nsteps));
        std::memcpy(p_out, res.p.data(), n * sizeof(double));
        std::memcpy(sw_out, res.sw.data(), n * sizeof(double));
        for (int s = 0; s < nsteps && iters; ++s) {
            iters[2 * s] = res.pressure_reports[s].iterations;
            iters[2 * s + 1] = res.transport_reports[s].iterations;
        }
    });
}

// One residual + Jacobian of the two-phase program: which = 0 the pressure residual
// at pressure `p` (total mobility from sw0), which = 1 the transport residual at
// saturation `sw` (total flux from p and sw0, step dt).  info = {n, nnz}.
int eqo_twophase_residual_params(const char* params, const double* perm, const double* poro, int n, const int* src_cells,
                          const double* src_vals, int nsrc, int which, const double* pr, const double* sw0,
                          const double* sw, double dt, int* info, double* F, int* rp, int* ci, double* jv) {
    return guard([&] {
        Params p;
        p.parse(params ? params : "");
        Runtime rt(p);
        TwoPhaseProgram prog(rt, p, std::vector<double>(perm, perm + n), std::vector<double>(poro, poro + n),
                             std::vector<int>(src_cells, src_cells + nsrc),
                             std::vector<double>(src_vals, src_vals + nsrc));
        ADB S0, P;
        S0.v.assign(sw0, sw0 + n);
        P.v.assign(pr, pr + n);
        ADB tm = adb_add(prog.water_mob(S0), prog.oil_mob(S0));
        ADB r;
        if (which == 0) {
            r = prog.pressure_residual(primary_variable(P), tm);
        } else {
            ADB S;
            S.v.assign(sw, sw + n);
            r = prog.transport_residual(primary_variable(S), S0, prog.total_flux(P, tm), dt);
        }
        info[0] = r.size();
        info[1] = r.J.nnz();
        if (!F) return;
        std::memcpy(F, r.v.data(), r.v.size() * sizeof(double));
        std::memcpy(rp, r.J.rp.data(), r.J.rp.size() * sizeof(int));
        std::memcpy(ci, r.J.ci.data(), r.J.ci.size() * sizeof(int));
        std::memcpy(jv, r.J.v.data(), r.J.v.size() * sizeof(double));
    });
}

// Fully implicit two-phase program: nsteps time steps; p_out/sw_out final fields;
// iters[s] = Newton iterations of step s (one newtonSolveSystem per step).
int eqo_twophase_fi_run(const char* params, const double* perm, const double* poro, const double* sw_initial, int n,
                        const int* src_cells, const double* src_vals, int nsrc, const double* dts, int nsteps,
                        double* p_out, double* sw_out, int* iters) {
    return guard([&] {
        Params p;
        p.parse(params ? params : "");
        TwoPhaseResult res = run_twophase_fi(p, std::vector<double>(perm, perm + n),
                                             std::vector<double>(poro, poro + n),
                                             std::vector<double>(sw_initial, sw_initial + n),
                                             std::vector<int>(src_cells, src_cells + nsrc),
                                             std::vector<double>(src_vals, src_vals + nsrc),
                                             std::vector<double>(dts, dts + nsteps));
        std::memcpy(p_out, res.p.data(), n * sizeof(double));
        std::memcpy(sw_out, res.sw.data(), n * sizeof(double));
        for (int s = 0; s < nsteps && iters; ++s) iters[s] = res.pressure_reports[s].iterations;
    });
}

// The combined residual + Jacobian of the fully implicit system at (p, sw) (previous
// saturation sw0, step dt): 2n rows, block order [pressure; transport], columns
// [p; sw] of the combined primary variable.  info = {2n, nnz}.
int eqo_twophase_fi_residual(const char* params, const double* perm, const double* poro, int n, const int* src_cells,
                             const double* src_vals, int nsrc, const double* pr, const double* sw0, const double* sw,
                             double dt, int* info, double* F, int* rp, int* ci, double* jv) {
    return guard([&] {
        Params p;
        p.parse(params ? params : "");
        Runtime rt(p);
        TwoPhaseProgram prog(rt, p, std::vector<double>(perm, perm + n), std::vector<double>(poro, poro + n),
                             std::vector<int>(src_cells, src_cells + nsrc),
                             std::vector<double>(src_vals, src_vals + nsrc));
        ADB S0, U;
        S0.v.assign(sw0, sw0 + n);
        U.v.assign(pr, pr + n);
        U.v.insert(U.v.end(), sw, sw + n);
        auto pres = [&](const ADB& pp, const ADB& ss) { return twophase_fi_pressure(prog, pp, ss); };
        auto tres = [&](const ADB& pp, const ADB& ss) { return twophase_fi_transport(prog, pp, ss, S0, dt); };
        ADB r = rt.systemResidual(pres, tres, primary_variable(U));
        info[0] = r.size();
        info[1] = r.J.nnz();
        if (!F) return;
        std::memcpy(F, r.v.data(), r.v.size() * sizeof(double));
        std::memcpy(rp, r.J.rp.data(), r.J.rp.size() * sizeof(int));
        std::memcpy(ci, r.J.ci.data(), r.J.ci.size() * sizeof(int));
        std::memcpy(jv, r.J.v.data(), r.J.v.size() * sizeof(double));
    });
}

// One residual + Jacobian evaluation of the heat program at state u (prev state u0, step dt).
// Sizes are returned in info = {n, nnz}; pass null arrays to query first.
int eqo_heat_residual(const char* params, const double* u, const double* u0, int n, const int* dir_faces,
                      const double* dir_vals, int ndir, double dt, int* info, double* F, int* rp, int* ci, double* jv) {
    return guard([&] {
        Params p;
        p.parse(params ? params : "");
        Runtime rt(p);
        HeatProgram prog(rt, p, std::vector<int>(dir_faces, dir_faces + ndir),
                         std::vector<double>(dir_vals, dir_vals + ndir));
        ADB uu;
        uu.v.assign(u, u + n);
        ADB U = primary_variable(uu);
        ADB U0;
        U0.v.assign(u0, u0 + n);
        ADB r = prog.residual(U, U0, dt);
        info[0] = r.size();
        info[1] = r.J.nnz();
        if (!F) return;
        std::memcpy(F, r.v.data(), r.v.size() * sizeof(double));
        std::memcpy(rp, r.J.rp.data(), r.J.rp.size() * sizeof(int));
        std::memcpy(ci, r.J.ci.data(), r.J.ci.size() * sizeof(int));
        std::memcpy(jv, r.J.v.data(), r.J.v.size() * sizeof(double));
    });
}

// Residual values only (no Jacobian) of the same program at (u, u0, dt): F has n entries.
int eqo_heat_residual_values(const char* params, const double* u, const double* u0, int n, const int* dir_faces,
                             const double* dir_vals, int ndir, double dt, double* F) {
    return guard([&] {
        Params p;
        p.parse(params ? params : "");
        Runtime rt(p);
        HeatProgram prog(rt, p, std::vector<int>(dir_faces, dir_faces + ndir),
                         std::vector<double>(dir_vals, dir_vals + ndir));
        ADB U, U0;
        U.v.assign(u, u + n);
        U0.v.assign(u0, u0 + n);
        ADB r = prog.residual(U, U0, dt);
        if ((int)r.v.size() != n) throw std::runtime_error("residual size differs from the state size");
        std::memcpy(F, r.v.data(), r.v.size() * sizeof(double));
    });
}

// Timed sample of the serial path: `iters` Newton iterations (residual + Jacobian
// assembly + linear solve) of the heat program at state u.  Returns wall seconds.
int eqo_heat_newton_sample(const char* params, const double* u0, int n, int iters, double dt, double* seconds,
                           int* lin_iters) {
    return guard([&] {
        Params p;
        p.parse(params ? params : "");
        Runtime rt(p);
        HeatProgram prog(rt, p, {}, {});
        ADB U0;
        U0.v.assign(u0, u0 + n);
        *seconds = time_newton_iterations(rt, prog, U0, dt, iters, lin_iters);
    });
}

}  // extern "C"
