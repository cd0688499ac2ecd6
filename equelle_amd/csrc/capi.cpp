// C-ABI of the runtime (include/eqhip.h): handle tables over the device engine.
#include <eqhip.h>

#include <algorithm>
#include <cstdio>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "core.hpp"
#include "engine.hpp"
#include "heat_fused.hpp"
#include "heat_program.hpp"
#include "kernels.hpp"
#include "krylov.hpp"
#include "runtime_core.hpp"
#include "tape.hpp"
#include "serve.hpp"
#include "twophase_program.hpp"
#include "timing.hpp"

using namespace eqhip;

namespace {
thread_local std::string g_err;

struct Rt {
    std::unique_ptr<RuntimeCore> core;
    std::vector<SP> scal;
    std::vector<IP> sets;
    std::vector<BP> bools;
    template <class T>
    static int put(std::vector<T>& v, T x) {
        for (size_t i = 0; i < v.size(); ++i)
            if (!v[i]) {
                v[i] = std::move(x);
                return (int)i + 1;
            }
        v.push_back(std::move(x));
        return (int)v.size();
    }
    template <class T>
    static const T& get(const std::vector<T>& v, int h) {
        if (h < 1 || h > (int)v.size() || !v[h - 1]) throw std::runtime_error("invalid handle " + std::to_string(h));
        return v[h - 1];
    }
    const SP& S(int h) const { return get(scal, h); }
    const IP& I(int h) const { return get(sets, h); }
    const BP& B(int h) const { return get(bools, h); }
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

Opm::ParameterGroup to_group(const char* params) {
    ParamMap m;
    m.parse(params ? params : "");
    Opm::ParameterGroup g;
    for (auto& e : m.kv) g.insertParameter(e.first, e.second);
    return g;
}
}  // namespace

extern "C" {

const char* eqhip_last_error(void) { return g_err.c_str(); }

int eqhip_runtime_create(const char* params, void** out) {
    return guard([&] {
        ParamMap m;
        m.parse(params ? params : "");
        auto* r = new Rt();
        try {
            r->core.reset(new RuntimeCore(m));
        } catch (...) {
            delete r;
            throw;
        }
        *out = r;
    });
}

int eqhip_runtime_destroy(void* rt) {
    return guard([&] { delete static_cast<Rt*>(rt); });
}

int eqhip_grid_info(void* rt, int* info) {
    return guard([&] {
        DeviceGrid& g = R(rt)->core->grid();
        info[0] = g.dim;
        info[1] = g.nc;
        info[2] = g.nf;
        info[3] = g.nif;
        info[4] = g.nf - g.nif;
    });
}

int eqhip_grid_int_array(void* rt, int which, int* out) {
    return guard([&] {
        DeviceGrid& g = R(rt)->core->grid();
        const DevPtr<int>& d = which == 0 ? g.face_cells : which == 1 ? g.cell_facepos : g.cell_faces;
        std::vector<int> h = dev_to_host(d->p, d->n);
        std::memcpy(out, h.data(), h.size() * sizeof(int));
    });
}

int eqhip_grid_double_array(void* rt, int which, double* out) {
    return guard([&] {
        DeviceGrid& g = R(rt)->core->grid();
        const DevPtr<double>* a[] = {&g.cell_centroids, &g.cell_volumes, &g.face_centroids, &g.face_areas,
                                     &g.face_normals};
        EQ_REQUIRE(which >= 0 && which < 5, "grid_double_array: bad selector " << which);
        std::vector<double> h = dev_to_host((*a[which])->p, (*a[which])->n);
        std::memcpy(out, h.data(), h.size() * sizeof(double));
    });
}

static const DevCsr& opmat(Rt* r, int which) {
    DeviceGrid& g = r->core->grid();
    EQ_REQUIRE(which >= 0 && which < 3, "opmat: bad selector " << which);
    return which == 0 ? g.grad : which == 1 ? g.div : g.fulldiv;
}
int eqhip_opmat_info(void* rt, int which, int* info) {
    return guard([&] {
        const DevCsr& m = opmat(R(rt), which);
        info[0] = m.pat->rows;
        info[1] = m.pat->cols;
        info[2] = m.pat->nnz;
    });
}
int eqhip_opmat_copy(void* rt, int which, int* rp, int* ci, double* v) {
    return guard([&] {
        const DevCsr& m = opmat(R(rt), which);
        std::memcpy(rp, m.pat->rp().data(), m.pat->rp().size() * sizeof(int));
        std::memcpy(ci, m.pat->ci().data(), m.pat->ci().size() * sizeof(int));
        std::vector<double> h = dev_to_host(m.val->p, m.pat->nnz);
        std::memcpy(v, h.data(), h.size() * sizeof(double));
    });
}

// ---------------- index sets
int eqhip_set_builtin(void* rt, int which, int* h) {
    return guard([&] {
        DeviceGrid& g = R(rt)->core->grid();
        const IP* s[] = {&g.all_cells, &g.boundary_cells, &g.interior_cells,
                         &g.all_faces, &g.boundary_faces, &g.interior_faces};
        EQ_REQUIRE(which >= 0 && which < 6, "unknown builtin set " << which);
        *h = Rt::put(R(rt)->sets, *s[which]);
    });
}
int eqhip_set_from_host(void* rt, int codim, const int* idx, int n, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->sets, make_index(codim, false, std::vector<int>(idx, idx + n))); });
}
int eqhip_set_size(void* rt, int h, int* n) {
    return guard([&] { *n = R(rt)->I(h)->size(); });
}
int eqhip_set_copy(void* rt, int h, int* out) {
    return guard([&] {
        const IP& s = R(rt)->I(h);
        std::vector<int> d = dev_to_host(s->dev->p, s->size());   // device copy is the source of truth
        std::memcpy(out, d.data(), d.size() * sizeof(int));
    });
}
int eqhip_first_cell(void* rt, int faces, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->sets, R(rt)->core->firstCell(R(rt)->I(faces))); });
}
int eqhip_second_cell(void* rt, int faces, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->sets, R(rt)->core->secondCell(R(rt)->I(faces))); });
}
int eqhip_set_on(void* rt, int data, int from, int to, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->sets, i_on(r->I(data), r->I(from), r->I(to)));
    });
}
int eqhip_set_trinary(void* rt, int pred, int t, int f, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->sets, i_select(r->B(pred), r->I(t), r->I(f)));
    });
}
int eqhip_is_empty(void* rt, int set, int* pred) {
    return guard([&] { *pred = Rt::put(R(rt)->bools, i_is_empty(R(rt)->I(set))); });
}
int eqhip_set_contains(void* rt, int super, int sub) {
    return guard([&] { position_map(R(rt)->I(super), R(rt)->I(sub)); });
}

// ---------------- booleans
int eqhip_bool_size(void* rt, int h, int* n) {
    return guard([&] { *n = R(rt)->B(h)->n; });
}
int eqhip_bool_copy(void* rt, int h, unsigned char* out) {
    return guard([&] {
        std::vector<uint8_t> b = b_to_host(R(rt)->B(h));
        std::memcpy(out, b.data(), b.size());
    });
}
int eqhip_compare(void* rt, int op, int a, int b, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->bools, s_compare(op, R(rt)->S(a), R(rt)->S(b))); });
}
int eqhip_compare_scalar(void* rt, int op, int a, double s, int scalar_left, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->bools, s_compare_scalar(op, R(rt)->S(a), s, scalar_left)); });
}

// ---------------- scalars
int eqhip_scalar_from_host(void* rt, const double* v, int n, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, s_from_host(v, n)); });
}
int eqhip_scalar_uniform(void* rt, int n, double s, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, s_uniform(n, s)); });
}
int eqhip_scalar_with_jacobian(void* rt, const double* v, int n, const int* rp, const int* ci, const double* jv,
                               int cols, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, s_with_jacobian(v, n, rp, ci, jv, cols)); });
}
int eqhip_primary(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, s_primary(R(rt)->S(a))); });
}
int eqhip_scalar_info(void* rt, int h, int* info) {
    return guard([&] {
        const SP& a = R(rt)->S(h);
        info[0] = a->n;
        info[1] = a->ad ? 1 : 0;
        info[2] = a->pat ? a->pat->rows : 0;
        info[3] = a->pat ? a->pat->cols : 0;
        info[4] = a->pat ? a->pat->nnz : 0;
    });
}
int eqhip_scalar_copy(void* rt, int h, double* out) {
    return guard([&] {
        std::vector<double> v = s_to_host(R(rt)->S(h));
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    });
}
int eqhip_jacobian_copy(void* rt, int h, int* rp, int* ci, double* v) {
    return guard([&] {
        const SP& a = R(rt)->S(h);
        EQ_REQUIRE(a->pat, "Trying to copy empty matrix to host");
        std::memcpy(rp, a->pat->rp().data(), a->pat->rp().size() * sizeof(int));
        std::memcpy(ci, a->pat->ci().data(), a->pat->ci().size() * sizeof(int));
        std::vector<double> j = dev_to_host(jac_values(*a), a->pat->nnz);
        std::memcpy(v, j.data(), j.size() * sizeof(double));
    });
}
int eqhip_release(void* rt, int kind, int h) {
    return guard([&] {
        Rt* r = R(rt);
        if (kind == 0) {
            r->S(h);
            r->scal[h - 1].reset();
        } else if (kind == 1) {
            r->I(h);
            r->sets[h - 1].reset();
        } else {
            r->B(h);
            r->bools[h - 1].reset();
        }
    });
}
int eqhip_binop(void* rt, int op, int a, int b, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        const SP &x = r->S(a), &y = r->S(b);
        SP o = op == 0 ? s_add(x, y) : op == 1 ? s_sub(x, y) : op == 2 ? s_mul(x, y) : s_div(x, y);
        *h = Rt::put(r->scal, o);
    });
}
int eqhip_scalar_binop(void* rt, int op, double s, int a, int scalar_left, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        const SP& x = r->S(a);
        SP o;
        if (op == 2) o = s_scal_mul(s, x);
        else if (op == 3 && !scalar_left) o = s_div_scal(x, s);
        else if (op == 3) o = s_scal_div(s, x);
        else throw std::runtime_error("scalar_binop: unsupported op");
        *h = Rt::put(r->scal, o);
    });
}
int eqhip_negate(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, s_neg(R(rt)->S(a))); });
}
int eqhip_sqrt(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, s_sqrt(R(rt)->S(a))); });
}
int eqhip_trinary(void* rt, int pred, int t, int f, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->scal, s_select(r->B(pred), r->S(t), r->S(f)));
    });
}
int eqhip_gradient(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, R(rt)->core->gradient(R(rt)->S(a))); });
}
int eqhip_divergence(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, R(rt)->core->divergence(R(rt)->S(a))); });
}
int eqhip_neg_gradient(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, R(rt)->core->neg_gradient(R(rt)->S(a))); });
}
int eqhip_interior_divergence(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, R(rt)->core->interior_divergence(R(rt)->S(a))); });
}
int eqhip_norm_scalar(void* rt, int a, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, s_norm(R(rt)->S(a))); });
}
int eqhip_extend(void* rt, int a, int from, int to, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->scal, s_extend(r->S(a), r->I(from), r->I(to)));
    });
}
int eqhip_extend_scalar(void* rt, double s, int to, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, s_uniform(R(rt)->I(to)->size(), s)); });
}
int eqhip_on(void* rt, int a, int from, int to, int* h) {
    return guard([&] {
        Rt* r = R(rt);
        *h = Rt::put(r->scal, s_on(r->S(a), r->I(from), r->I(to)));
    });
}
int eqhip_reduce(void* rt, int kind, int a, double* out) {
    return guard([&] {
        const SP& x = R(rt)->S(a);
        if (kind == 4) *out = s_two_norm(x);
        else *out = s_reduce(kind == 0 ? RED_SUM : kind == 1 ? RED_PROD : kind == 2 ? RED_MAX : RED_MIN, x);
    });
}
int eqhip_reduce_noalloc(void* rt, int kind, int a, double* out) {
    return guard([&] {
        const SP& x = R(rt)->S(a);
        EQ_REQUIRE(kind >= 0 && kind <= 4, "reduction kind must be 0..4, got " << kind);
        EQ_REQUIRE(x->n > 0 || kind == 0 || kind == 1 || kind == 4, "reduction of an empty collection");
        const double* v = x->val ? x->val->p : nullptr;
        if (kind == 4) *out = std::sqrt(reduce_noalloc(RED_SUMSQ, v, x->n));
        else *out = reduce_noalloc(kind == 0 ? RED_SUM : kind == 1 ? RED_PROD : kind == 2 ? RED_MAX : RED_MIN, v, x->n);
    });
}
int eqhip_norm_set(void* rt, int set, int* h) {
    return guard([&] { *h = Rt::put(R(rt)->scal, R(rt)->core->norm(R(rt)->I(set))); });
}
int eqhip_centroid(void* rt, int set, double* out) {
    return guard([&] {
        std::vector<double> v = v_to_host(R(rt)->core->centroid(R(rt)->I(set)));
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    });
}
int eqhip_normal(void* rt, int set, double* out) {
    return guard([&] {
        std::vector<double> v = v_to_host(R(rt)->core->normal(R(rt)->I(set)));
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    });
}
int eqhip_solve_for_update(void* rt, int residual, int* du, int* iters) {
    return guard([&] {
        LinReport rep;
        SP d = R(rt)->core->solve_for_update(R(rt)->S(residual), &rep);
        if (iters) *iters = rep.iterations;
        *du = Rt::put(R(rt)->scal, d);
    });
}

// ---------------- CSR algebra
int eqhip_csr_binop(int op, int ar, int ac, const int* arp, const int* aci, const double* av, int br, int bc,
                    const int* brp, const int* bci, const double* bv, int* out_info, int* orp, int* oci, double* ov,
                    int cap) {
    return guard([&] {
        DevCsr A, B;
        if (arp) A = csr_from_host(ar, ac, arp, aci, av);
        if (brp) B = csr_from_host(br, bc, brp, bci, bv);
        DevCsr C;
        if (op == 2) {
            if (A.pat && B.pat) EQ_REQUIRE(ac == br, "Error in CudaMatrix * CudaMatrix size checking");
            C = csr_mul(A, B);
        } else {
            if (A.pat && B.pat) EQ_REQUIRE(ar == br && ac == bc, "CudaMatrix +/-: matrices of different sizes");
            C = csr_add(A, 1.0, B, op == 1 ? -1.0 : 1.0);
        }
        out_info[0] = C.pat ? C.pat->rows : 0;
        out_info[1] = C.pat ? C.pat->cols : 0;
        out_info[2] = C.pat ? C.pat->nnz : 0;
        if (!C.pat) return;
        EQ_REQUIRE(C.pat->nnz <= cap, "csr_binop: output capacity too small");
        std::memcpy(orp, C.pat->rp().data(), C.pat->rp().size() * sizeof(int));
        std::memcpy(oci, C.pat->ci().data(), C.pat->ci().size() * sizeof(int));
        std::vector<double> v = dev_to_host(C.val->p, C.pat->nnz);
        std::memcpy(ov, v.data(), v.size() * sizeof(double));
    });
}
int eqhip_csr_spmv(int rows, int cols, const int* rp, const int* ci, const double* v, const double* x, double* y,
                   int transpose) {
    return guard([&] {
        DevCsr A = csr_from_host(rows, cols, rp, ci, v);
        if (transpose) {
            // structural transpose (symbolic on host), values gathered on the device
            const Pattern& p = *A.pat;
            std::vector<int> trp(p.cols + 1, 0), tci(p.nnz), src(p.nnz);
            for (int c : p.ci()) trp[c + 1]++;
            for (int i = 0; i < p.cols; ++i) trp[i + 1] += trp[i];
            std::vector<int> pos(trp.begin(), trp.end() - 1);
            for (int r = 0; r < p.rows; ++r)
                for (int k = p.rp()[r]; k < p.rp()[r + 1]; ++k) {
                    const int q = pos[p.ci()[k]]++;
                    tci[q] = r;
                    src[q] = k;
                }
            DevCsr T;
            T.pat = make_pattern(p.cols, p.rows, std::move(trp), std::move(tci));
            T.val = dev_alloc<double>(p.nnz);
            auto dsrc = dev_from_host(src.data(), src.size());
            launch_gather(A.val->p, dsrc->p, T.val->p, p.nnz);
            A = T;
        }
        SP xs = s_from_host(x, A.pat->cols);
        SP ys = s_matmul(A, xs);
        std::vector<double> h = s_to_host(ys);
        std::memcpy(y, h.data(), h.size() * sizeof(double));
    });
}
int eqhip_sell_spmv(void* rt, int h, const double* x, double* y, int fp32) {
    return guard([&] { sell_spmv_test(R(rt)->S(h), x, y, fp32 != 0); });
}

// ---------------- heat program
int eqhip_heat_run(const char* params, const double* u_initial, int n, const int* dir_faces, const double* dir_vals,
                   int ndir, const double* dts, int nsteps, double* u_out, int* newton_iters, double* final_norms) {
    return guard([&] {
        Opm::ParameterGroup param = to_group(params);
        equelleCUDA::EquelleRuntimeCUDA er(param);
        HeatProgramCUDA prog(er, std::vector<int>(dir_faces, dir_faces + ndir),
                             std::vector<double>(dir_vals, dir_vals + ndir));
        equelleCUDA::CollOfScalar u0(std::vector<double>(u_initial, u_initial + n));
        for (int s = 0; s < nsteps; ++s) {
            const double dt = dts[s];
            auto rescomp = [&](const equelleCUDA::CollOfScalar& u) { return prog.residual(u, u0, dt); };
            equelleCUDA::CollOfScalar u = er.newtonSolve(rescomp, u0);
            if (newton_iters) newton_iters[s] = er.lastNewtonReport().iterations;
            if (final_norms) final_norms[s] = er.lastNewtonReport().final_norm;
            u0 = u;
        }
        std::vector<double> h = u0.copyToHost();
        std::memcpy(u_out, h.data(), h.size() * sizeof(double));
    });
}

// ---------------- heat simulation sessions (device-resident state between steps)
namespace {
struct HeatSim {
    Opm::ParameterGroup param;
    std::unique_ptr<equelleCUDA::EquelleRuntimeCUDA> er;
    std::unique_ptr<HeatProgramCUDA> prog;
    equelleCUDA::CollOfScalar u;
};
}  // namespace

int eqhip_heat_create(const char* params, const double* u_initial, int n, const int* dir_faces,
                      const double* dir_vals, int ndir, void** sim) {
    return guard([&] {
        auto s = std::make_unique<HeatSim>();
        s->param = to_group(params);
        s->er.reset(new equelleCUDA::EquelleRuntimeCUDA(s->param));
        EQ_REQUIRE(n == s->er->core().grid().nc, "u_initial has " << n << " values for "
                                                               << s->er->core().grid().nc << " cells");
        s->prog.reset(new HeatProgramCUDA(*s->er, std::vector<int>(dir_faces, dir_faces + ndir),
                                          std::vector<double>(dir_vals, dir_vals + ndir)));
        s->u = equelleCUDA::CollOfScalar(std::vector<double>(u_initial, u_initial + n));
        *sim = s.release();
    });
}

int eqhip_heat_step(void* sim, double dt, int* stats, double* norms) {
    return guard([&] {
        EQ_REQUIRE(sim, "null simulation");
        HeatSim& s = *static_cast<HeatSim*>(sim);
        const equelleCUDA::CollOfScalar u0 = s.u;
        auto rescomp = [&](const equelleCUDA::CollOfScalar& u) { return s.prog->residual(u, u0, dt); };
        s.u = s.er->newtonSolve(rescomp, u0);
        const equelleCUDA::NewtonReport& r = s.er->lastNewtonReport();
        if (stats) {
            stats[0] = r.iterations;
            stats[1] = r.linear_iterations;
            stats[2] = r.stop_reason;
        }
        if (norms) {
            norms[0] = r.initial_norm;
            norms[1] = r.final_norm;
        }
    });
}

int eqhip_heat_state(void* sim, double* u) {
    return guard([&] {
        EQ_REQUIRE(sim, "null simulation");
        std::vector<double> h = static_cast<HeatSim*>(sim)->u.copyToHost();
        std::memcpy(u, h.data(), h.size() * sizeof(double));
    });
}

int eqhip_heat_set_state(void* sim, const double* u) {
    return guard([&] {
        EQ_REQUIRE(sim, "null simulation");
        HeatSim* s = static_cast<HeatSim*>(sim);
        s->u = equelleCUDA::CollOfScalar(std::vector<double>(u, u + s->u.size()));
    });
}

int eqhip_heat_eval_time(void* sim, double dt, int reps, double* ms_per_eval, int* info) {
    return guard([&] {
        EQ_REQUIRE(sim, "null simulation");
        EQ_REQUIRE(reps > 0, "reps must be positive");
        HeatSim* s = static_cast<HeatSim*>(sim);
        const equelleCUDA::CollOfScalar u0 = s->u;
        const equelleCUDA::CollOfScalar u(s->u, true);   // the primary variable of the Newton solve
        auto rescomp = [&](const equelleCUDA::CollOfScalar& x) { return s->prog->residual(x, u0, dt); };
        // the recorded program, as newtonSolve builds it (one call site of its own)
        static int site_state = 0;
        equelleCUDA::FusedResidual fr(*s->er, &site_state, u);
        equelleCUDA::CollOfScalar r;
        for (int pass = 0;; ++pass) {
            fr.begin(pass);
            try {
                if (fr.finish(rescomp(u), &r)) break;
            } catch (const std::exception& e) {
                fr.aborted(e);
            }
        }
        const bool fused = fr.fused();
        auto eval = [&] {
            if (fused) return fr.eval(u, true);
            equelleCUDA::CollOfScalar q = rescomp(u);
            (void)jac_values(*q.data());   // the Jacobian's values, pending factors applied
            return q;
        };
        r = eval();   // symbolic passes and caches outside the timing
        Device::get().sync();
        const auto t0 = std::chrono::steady_clock::now();
        for (int q = 0; q < reps; ++q) r = eval();
        Device::get().sync();
        *ms_per_eval = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
        info[0] = r.size();
        info[1] = r.data()->pat ? r.data()->pat->nnz : 0;
        info[2] = fused ? 1 : 0;
    });
}

int eqhip_residual_stats(long* out, double* seconds) {
    return guard([&] {
        const ResidualStats st = residual_stats();
        const long v[6] = {st.recorded, st.symbolic, st.compiled, st.evaluations, st.aborted, st.rejected};
        for (int i = 0; i < 6; ++i) out[i] = v[i];
        if (seconds) *seconds = st.compile_seconds;
    });
}

int eqhip_residual_codegen_selftest(char* log, int len) {
    std::string lg;
    const int rc = residual_codegen_selftest(&lg);
    if (log && len > 0) {
        std::strncpy(log, lg.c_str(), (size_t)len - 1);
        log[len - 1] = '\0';
    }
    return rc;
}

int eqhip_heat_destroy(void* sim) {
    return guard([&] { delete static_cast<HeatSim*>(sim); });
}

// ---------------- two-phase sessions (generic AD engine, printer-style program)
}  // extern "C"
namespace {
struct TwoPhaseSim {
    Opm::ParameterGroup param;
    std::unique_ptr<equelleCUDA::EquelleRuntimeCUDA> er;
    std::unique_ptr<TwoPhaseProgramCUDA> prog;
    equelleCUDA::CollOfScalar p, sw;
    bool fully_implicit = false;   // twophase_fully_implicit.equelle: one newtonSolveSystem per step
};
TwoPhaseSim& TP(void* s) {
    EQ_REQUIRE(s, "null two-phase simulation");
    return *static_cast<TwoPhaseSim*>(s);
}
}  // namespace
extern "C" {

int eqhip_twophase_create(const char* params, const double* perm, const double* poro, const double* sw_initial,
                          int n, const int* src_cells, const double* src_vals, int nsrc, void** sim) {
    return guard([&] {
        auto s = std::make_unique<TwoPhaseSim>();
        s->param = to_group(params);
        s->er.reset(new equelleCUDA::EquelleRuntimeCUDA(s->param));
        EQ_REQUIRE(n == s->er->core().grid().nc, "fields have " << n << " values for " << s->er->core().grid().nc
                                                              << " cells");
        s->prog.reset(new TwoPhaseProgramCUDA(*s->er, std::vector<double>(perm, perm + n),
                                              std::vector<double>(poro, poro + n),
                                              std::vector<int>(src_cells, src_cells + nsrc),
                                              std::vector<double>(src_vals, src_vals + nsrc)));
        s->sw = equelleCUDA::CollOfScalar(std::vector<double>(sw_initial, sw_initial + n));
        s->p = s->er->operatorExtend(double(0), s->er->allCells());
        s->fully_implicit = s->er->core().params().get_bool("fully_implicit", false);
        *sim = s.release();
    });
}

int eqhip_twophase_step(void* sim, double dt, int* stats, double* norms) {
    return guard([&] {
        TwoPhaseSim& s = TP(sim);
        equelleCUDA::EquelleRuntimeCUDA& er = *s.er;
        const TwoPhaseProgramCUDA& prog = *s.prog;
        if (s.fully_implicit) {
            // twophase_fully_implicit.equelle:117-149 (out_twophase_fully_implicit.cpp:125-141)
            const equelleCUDA::CollOfScalar sw0 = s.sw;
            auto pressureResLocal = [&](const equelleCUDA::CollOfScalar& p, const equelleCUDA::CollOfScalar& sw) {
                return prog.pressureResLocal(p, sw);
            };
            auto transportResLocal = [&](const equelleCUDA::CollOfScalar& p, const equelleCUDA::CollOfScalar& sw) {
                return prog.transportResLocal(p, sw, sw0, dt);
            };
            const std::tuple<equelleCUDA::CollOfScalar, equelleCUDA::CollOfScalar> newvals = er.newtonSolveSystem(
                equelleCUDA::makeArray(pressureResLocal, transportResLocal),
                equelleCUDA::makeArray(s.p, er.operatorExtend(double(0.5), er.allCells())));
            const equelleCUDA::NewtonReport& r = er.lastNewtonReport();
            s.p = std::get<0>(newvals);
            s.sw = std::get<1>(newvals);
            if (stats) {   // the system solve in the pressure slots, its stop reason in both
                stats[0] = r.iterations;
                stats[1] = r.linear_iterations;
                stats[2] = r.stop_reason;
                stats[3] = 0;
                stats[4] = 0;
                stats[5] = r.stop_reason;
            }
            if (norms) {
                norms[0] = norms[2] = r.initial_norm;
                norms[1] = norms[3] = r.final_norm;
            }
            return;
        }
        // twophase.equelle:99-113
        const equelleCUDA::CollOfScalar total_mobility = (prog.waterMob(s.sw) + prog.oilMob(s.sw));
        auto pressureResLocal = [&](const equelleCUDA::CollOfScalar& pressure) {
            return prog.pressureResidual(pressure, total_mobility);
        };
        const equelleCUDA::CollOfScalar p = er.newtonSolve(pressureResLocal, s.p);
        const equelleCUDA::NewtonReport rp = er.lastNewtonReport();
        const equelleCUDA::CollOfScalar flux = prog.totalFlux(p, total_mobility);
        const equelleCUDA::CollOfScalar sw0 = s.sw;
        auto transportResLocal = [&](const equelleCUDA::CollOfScalar& sw) {
            return prog.transportResidual(sw, sw0, flux, dt);
        };
        const equelleCUDA::CollOfScalar sw = er.newtonSolve(transportResLocal, er.operatorExtend(double(0.5), er.allCells()));
        const equelleCUDA::NewtonReport& rs = er.lastNewtonReport();
        s.p = p;
        s.sw = sw;
        if (stats) {
            stats[0] = rp.iterations;
            stats[1] = rp.linear_iterations;
            stats[2] = rp.stop_reason;
            stats[3] = rs.iterations;
            stats[4] = rs.linear_iterations;
            stats[5] = rs.stop_reason;
        }
        if (norms) {
            norms[0] = rp.initial_norm;
            norms[1] = rp.final_norm;
            norms[2] = rs.initial_norm;
            norms[3] = rs.final_norm;
        }
    });
}

int eqhip_twophase_set_state(void* sim, const double* p, const double* sw) {
    return guard([&] {
        TwoPhaseSim& s = TP(sim);
        const int n = s.sw.size();
        if (p) s.p = equelleCUDA::CollOfScalar(std::vector<double>(p, p + n));
        if (sw) s.sw = equelleCUDA::CollOfScalar(std::vector<double>(sw, sw + n));
    });
}
int eqhip_twophase_state(void* sim, double* p, double* sw) {
    return guard([&] {
        TwoPhaseSim& s = TP(sim);
        if (p) {
            std::vector<double> h = s.p.copyToHost();
            std::memcpy(p, h.data(), h.size() * sizeof(double));
        }
        if (sw) {
            std::vector<double> h = s.sw.copyToHost();
            std::memcpy(sw, h.data(), h.size() * sizeof(double));
        }
    });
}

int eqhip_twophase_residual(void* sim, int which, const double* pr, const double* sw0, const double* sw, double dt,
                            int* info, double* F, int* rp, int* ci, double* jv) {
    return guard([&] {
        TwoPhaseSim& s = TP(sim);
        const int n = s.er->core().grid().nc;
        const equelleCUDA::CollOfScalar S0(std::vector<double>(sw0, sw0 + n));
        const equelleCUDA::CollOfScalar P(std::vector<double>(pr, pr + n));
        const equelleCUDA::CollOfScalar tm = (s.prog->waterMob(S0) + s.prog->oilMob(S0));
        equelleCUDA::CollOfScalar r;
        if (which == 2) {   // fully implicit system at (p, sw): 2n rows, columns [p; sw]
            std::vector<double> u(pr, pr + n);
            u.insert(u.end(), sw, sw + n);
            auto pres = [&](const equelleCUDA::CollOfScalar& p, const equelleCUDA::CollOfScalar& ss) {
                return s.prog->pressureResLocal(p, ss);
            };
            auto tres = [&](const equelleCUDA::CollOfScalar& p, const equelleCUDA::CollOfScalar& ss) {
                return s.prog->transportResLocal(p, ss, S0, dt);
            };
            r = s.er->systemResidual(equelleCUDA::makeArray(pres, tres),
                                     equelleCUDA::CollOfScalar(equelleCUDA::CollOfScalar(u), true));
        } else if (which == 0) {
            r = s.prog->pressureResidual(equelleCUDA::CollOfScalar(P, true), tm);
        } else {
            const equelleCUDA::CollOfScalar S(std::vector<double>(sw, sw + n));
            r = s.prog->transportResidual(equelleCUDA::CollOfScalar(S, true), S0, s.prog->totalFlux(P, tm), dt);
        }
        info[0] = r.size();
        info[1] = r.data()->pat ? r.data()->pat->nnz : 0;
        if (!F) return;
        std::vector<double> fv = r.copyToHost();
        std::memcpy(F, fv.data(), fv.size() * sizeof(double));
        equelleCUDA::hostMat m = r.matrixToHost();
        std::memcpy(rp, m.rowPtr.data(), m.rowPtr.size() * sizeof(int));
        std::memcpy(ci, m.colInd.data(), m.colInd.size() * sizeof(int));
        std::memcpy(jv, m.vals.data(), m.vals.size() * sizeof(double));
    });
}

int eqhip_twophase_destroy(void* sim) {
    return guard([&] { delete static_cast<TwoPhaseSim*>(sim); });
}

// ---------------- fused heat batches
static FusedConfig fused_config(const char* params) {
    ParamMap m;
    m.parse(params ? params : "");
    FusedConfig c;
    c.precision = m.get_string("precision", c.precision);
    std::string pc = m.get_string("preconditioner", m.get_string("precond", c.precond));
    std::transform(pc.begin(), pc.end(), pc.begin(), ::tolower);
    c.precond = pc == "diagonal" ? "jacobi" : pc;
    c.rel_res_tol = m.get_double("rel_res_tol", c.rel_res_tol);
    c.abs_res_tol = m.get_double("abs_res_tol", c.abs_res_tol);
    c.max_iter = m.get_int("max_iter", c.max_iter);
    c.solver_tol = m.get_double("solver_tol", c.solver_tol);
    c.solver_max_iter = m.get_int("solver_max_iter", c.solver_max_iter);
    c.inner_tol = m.get_double("inner_tol", c.inner_tol);
    c.inner_tol_twophase = m.get_double("inner_tol_twophase", c.inner_tol_twophase);
    c.forcing = m.get_double("forcing", c.forcing);
    c.eta_max = m.get_double("eta_max", c.eta_max);
    c.forcing0 = m.get_double("forcing0", c.forcing0);
    EQ_REQUIRE(c.forcing >= 0.0 && c.eta_max > 0.0 && c.eta_max < 1.0, "forcing must be >= 0 and eta_max in (0, 1)");
    c.krylov_form_heat = m.get_string("krylov_form_heat", c.krylov_form_heat);
    c.krylov_form_twophase = m.get_string("krylov_form_twophase", c.krylov_form_twophase);
    c.pair_vectors = m.get_int("pair_vectors", c.pair_vectors ? 1 : 0) != 0;
    c.operator_tiles = m.get_int("operator_tiles", c.operator_tiles ? 1 : 0) != 0;
    c.twophase_pressure_cg = m.get_int("twophase_pressure_cg", c.twophase_pressure_cg ? 1 : 0) != 0;
    c.twophase_pressure_mg = m.get_int("twophase_pressure_mg", c.twophase_pressure_mg ? 1 : 0) != 0;
    c.mg_scale = m.get_double("mg_scale", c.mg_scale);
    c.twophase_split_form = m.get_int("twophase_split_form", c.twophase_split_form ? 1 : 0) != 0;
    EQ_REQUIRE(c.mg_scale > 0.0 && c.mg_scale < 2.0, "mg_scale must lie in (0, 2), got " << c.mg_scale);
    c.max_refine = m.get_int("max_refine", c.max_refine);
    c.graph = m.get_int("graph", 1) != 0;
    c.speculate = m.get_int("speculate", 1) != 0;
    c.newton_graph = m.get_int("newton_graph", c.newton_graph ? 1 : 0) != 0;
    c.cpt = m.get_int("cpt", c.cpt);
    EQ_REQUIRE(c.cpt == 0 || c.cpt == 1 || c.cpt == 2 || c.cpt == 4 || c.cpt == 8, "cpt must be 0, 1, 2, 4 or 8");
    c.xcd = m.get_int("xcd", c.xcd) != 0 ? 1 : 0;
    const std::string shape = m.get_string("shape", "");
    if (!shape.empty()) {   // a,d,s,p,v,xcd: cells per thread of every kernel class, XCD order
        int v[6];
        EQ_REQUIRE(std::sscanf(shape.c_str(), "%d,%d,%d,%d,%d,%d", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]) == 6,
                   "shape must be six comma-separated integers, got " << shape);
        for (int q = 0; q < 5; ++q) {
            EQ_REQUIRE(v[q] == 1 || v[q] == 2 || v[q] == 4 || v[q] == 8, "shape: cells per thread must be 1, 2, 4 or 8");
            c.cpt_fixed[q] = v[q];
        }
        c.xcd = v[5] != 0 ? 1 : 0;
    }
    c.tune = m.get_int("tune", 1) != 0;
    return c;
}
static HeatBatch& B(void* b) {
    EQ_REQUIRE(b, "null batch");
    return *static_cast<HeatBatch*>(b);
}

int eqhip_batch_create(const char* params, void** batch) {
    return guard([&] { *batch = new HeatBatch(fused_config(params)); });
}
int eqhip_batch_destroy(void* batch) {
    return guard([&] { delete static_cast<HeatBatch*>(batch); });
}
int eqhip_batch_add(void* batch, int nx, int ny, double k, double cv, const double* u, int* slot) {
    return guard([&] { *slot = B(batch).add(nx, ny, k, cv, u); });
}
int eqhip_batch_add_twophase(void* batch, const char* params, int nx, int ny, int nz, const double* perm,
                             const double* poro, const double* sw, int nsrc, const int* src_cells,
                             const double* src_vals, int* slot) {
    return guard([&] {
        ParamMap m;
        m.parse(params ? params : "");
        TwoPhaseSpec sp;
        sp.nx = nx;
        sp.ny = ny;
        sp.nz = nz;
        sp.dx = m.get_double("dx", 1.0);
        sp.dy = m.get_double("dy", 1.0);
        sp.dz = m.get_double("dz", 1.0);
        sp.watervisc = m.get_double("watervisc", sp.watervisc);
        sp.oilvisc = m.get_double("oilvisc", sp.oilvisc);
        sp.min_poro = m.get_double("min_poro", sp.min_poro);
        sp.corey = m.get_int("corey", sp.corey);
        *slot = B(batch).add_twophase(sp, perm, poro, sw, nsrc, src_cells, src_vals);
    });
}
int eqhip_batch_twophase_state(void* batch, int slot, double* p, double* sw) {
    return guard([&] { B(batch).get_twophase_state(slot, p, sw); });
}
int eqhip_batch_twophase_stats(void* batch, int slot, int* stats) {
    return guard([&] {
        const StepStats& st = B(batch).stats(slot);
        stats[0] = st.newton_p;
        stats[1] = st.krylov_p;
        stats[2] = st.stop_p;
    });
}
int eqhip_batch_twophase_residual(void* batch, int slot, int stage, const double* p, const double* sw0,
                                  const double* sw, double dt, int* info, double* F, int* rp, int* ci, double* v) {
    return guard([&] {
        std::vector<double> f, vv;
        std::vector<int> r, c;
        B(batch).export_twophase_residual(slot, stage, p, sw0, sw, dt, f, r, c, vv);
        info[0] = (int)f.size();
        info[1] = (int)c.size();
        std::copy(f.begin(), f.end(), F);
        std::copy(r.begin(), r.end(), rp);
        std::copy(c.begin(), c.end(), ci);
        std::copy(vv.begin(), vv.end(), v);
    });
}
int eqhip_batch_remove(void* batch, int slot) {
    return guard([&] { B(batch).remove(slot); });
}
int eqhip_batch_begin_step(void* batch, int slot, double dt) {
    return guard([&] { B(batch).begin_step(slot, dt); });
}
int eqhip_batch_iterate(void* batch, int* finished, int cap, int* nfinished) {
    return guard([&] {
        std::vector<int> done = B(batch).iterate();
        EQ_REQUIRE((int)done.size() <= cap || !finished, "finished-slot buffer too small");
        if (finished) std::copy(done.begin(), done.end(), finished);
        *nfinished = (int)done.size();
    });
}
int eqhip_batch_in_flight(void* batch, int slot, int* flag) {
    return guard([&] { *flag = B(batch).in_flight(slot) ? 1 : 0; });
}
int eqhip_batch_step(void* batch, const int* slots, int nslots, double dt) {
    return guard([&] { B(batch).step(std::vector<int>(slots, slots + nslots), dt); });
}
int eqhip_batch_stats(void* batch, int slot, int* stats, double* norms) {
    return guard([&] {
        const StepStats& s = B(batch).stats(slot);
        if (stats) {
            stats[0] = s.newton;
            stats[1] = s.krylov;
            stats[2] = s.refinements;
            stats[3] = s.stop;
            stats[4] = s.failed ? 1 : 0;
        }
        if (norms) {
            norms[0] = s.r0;
            norms[1] = s.rn;
        }
    });
}
int eqhip_batch_stats_many(void* batch, const int* slots, int nslots, int* stats) {
    return guard([&] {
        EQ_REQUIRE(nslots >= 0 && (nslots == 0 || (slots && stats)), "batch_stats_many: bad arguments");
        for (int q = 0; q < nslots; ++q) {
            const StepStats& s = B(batch).stats(slots[q]);
            int* o = stats + 5 * (size_t)q;
            o[0] = s.newton;
            o[1] = s.krylov;
            o[2] = s.refinements;
            o[3] = s.stop;
            o[4] = s.failed ? 1 : 0;
        }
    });
}
int eqhip_batch_state(void* batch, int slot, double* u) {
    return guard([&] { B(batch).get_state(slot, u); });
}
int eqhip_batch_set_state(void* batch, int slot, const double* u) {
    return guard([&] { B(batch).set_state(slot, u); });
}
int eqhip_batch_set_twophase_state(void* batch, int slot, const double* p, const double* sw) {
    return guard([&] { B(batch).set_twophase_state(slot, p, sw); });
}
int eqhip_batch_bytes(void* batch, int slot, double* bytes) {
    return guard([&] { *bytes = (double)B(batch).device_bytes(slot); });
}
int eqhip_batch_shape(void* batch, int* info) {
    return guard([&] {
        LaunchShape sh = B(batch).launch_shape();
        info[0] = sh.cpt;
        info[1] = sh.xcd;
        info[2] = sh.signatures;
    });
}
int eqhip_batch_shape_classes(void* batch, int* cpt) {
    return guard([&] {
        LaunchShape sh = B(batch).launch_shape();
        for (int c = 0; c < 5; ++c) cpt[c] = sh.cpt_class[c];
    });
}
int eqhip_batch_kernel_time(void* batch, const char* kernel, int reps, double* avg_ms, double* event_ms) {
    return guard([&] { *avg_ms = B(batch).time_kernel(kernel ? kernel : "", reps, event_ms); });
}
int eqhip_batch_residual(void* batch, int slot, const double* u, const double* u0, double dt, int* info, double* F,
                         int* rp, int* ci, double* v) {
    return guard([&] {
        std::vector<double> f, vv;
        std::vector<int> r, c;
        B(batch).export_residual(slot, u, u0, dt, f, r, c, vv);
        info[0] = (int)f.size();
        info[1] = (int)c.size();
        std::copy(f.begin(), f.end(), F);
        std::copy(r.begin(), r.end(), rp);
        std::copy(c.begin(), c.end(), ci);
        std::copy(vv.begin(), vv.end(), v);
    });
}

// ---------------- request loop
static HeatServer& SV(void* s) {
    EQ_REQUIRE(s, "null server");
    return *static_cast<HeatServer*>(s);
}
int eqhip_server_create(const char* params, void** server) {
    return guard([&] {
        ParamMap m;
        m.parse(params ? params : "");
        ServerConfig c;
        c.batch = fused_config(params);
        c.hbm_budget = m.get_double("hbm_budget", c.hbm_budget);
        c.max_in_flight = m.get_int("max_in_flight", c.max_in_flight);
        c.max_in_flight_twophase = m.get_int("max_in_flight_twophase", c.max_in_flight_twophase);
        c.twophase_precision = m.get_string("twophase_precision", c.twophase_precision);
        EQ_REQUIRE(c.twophase_precision == "fp64" || c.twophase_precision == "mixed",
                   "twophase_precision must be fp64 or mixed");
        *server = new HeatServer(c);
    });
}
int eqhip_server_destroy(void* server) {
    return guard([&] { delete static_cast<HeatServer*>(server); });
}
int eqhip_server_submit(void* server, int nx, int ny, double k, double cv, const double* u0, int steps, double dt,
                        int* request) {
    return guard([&] { *request = SV(server).submit(nx, ny, k, cv, u0, steps, dt); });
}
int eqhip_server_submit_twophase(void* server, const char* params, int nx, int ny, int nz, const double* perm,
                                 const double* poro, const double* sw, int nsrc, const int* src_cells,
                                 const double* src_vals, int steps, double dt, int* request) {
    return guard([&] {
        ParamMap m;
        m.parse(params ? params : "");
        TwoPhaseSpec sp;
        sp.nx = nx;
        sp.ny = ny;
        sp.nz = nz;
        sp.dx = m.get_double("dx", 1.0);
        sp.dy = m.get_double("dy", 1.0);
        sp.dz = m.get_double("dz", 1.0);
        sp.watervisc = m.get_double("watervisc", sp.watervisc);
        sp.oilvisc = m.get_double("oilvisc", sp.oilvisc);
        sp.min_poro = m.get_double("min_poro", sp.min_poro);
        sp.corey = m.get_int("corey", sp.corey);
        *request = SV(server).submit_twophase(sp, perm, poro, sw, nsrc, src_cells, src_vals, steps, dt);
    });
}
int eqhip_server_progress(void* server, int* finished, int cap, int* nfinished) {
    return guard([&] {
        std::vector<int> done = SV(server).progress();
        EQ_REQUIRE((int)done.size() <= cap || !finished, "finished-request buffer too small");
        if (finished) std::copy(done.begin(), done.end(), finished);
        *nfinished = (int)done.size();
    });
}
int eqhip_server_result(void* server, int request, double* u, double* p, int* stats, double* times) {
    return guard([&] {
        RequestInfo info;
        SV(server).take_result(request, u, p, &info);
        if (stats) {
            stats[0] = info.steps;
            stats[1] = info.newton;
            stats[2] = info.krylov;
            stats[3] = info.failed ? 1 : 0;
        }
        if (times) {
            times[0] = info.queued_s;
            times[1] = info.latency_s;
        }
    });
}
int eqhip_server_set_max_newton(void* server, int request, int max_newton) {
    return guard([&] { SV(server).set_max_newton(request, max_newton); });
}
int eqhip_server_status(void* server, int* counts, double* bytes) {
    return guard([&] {
        HeatServer& s = SV(server);
        if (counts) {
            counts[0] = s.queued();
            counts[1] = s.in_flight();
        }
        if (bytes) {
            bytes[0] = s.used_bytes();
            bytes[1] = s.config().hbm_budget;
        }
    });
}

// ---------------- timing / device
int eqhip_timing(int mode) {
    return guard([&] {
        if (mode == 2) timing_reset();
        else timing_enable(mode != 0);
    });
}
int eqhip_timing_query(const char* name, int* count, double* total_ms) {
    return guard([&] { timing_query(name ? name : "", count, total_ms); });
}
int eqhip_device_sync(void) {
    return guard([&] { Device::get().sync(); });
}

int eqhip_heat_residual(const char* params, const double* u, const double* u0, int n, const int* dir_faces,
                        const double* dir_vals, int ndir, double dt, int* info, double* F, int* rp, int* ci,
                        double* jv) {
    return guard([&] {
        Opm::ParameterGroup param = to_group(params);
        equelleCUDA::EquelleRuntimeCUDA er(param);
        HeatProgramCUDA prog(er, std::vector<int>(dir_faces, dir_faces + ndir),
                             std::vector<double>(dir_vals, dir_vals + ndir));
        equelleCUDA::CollOfScalar U(equelleCUDA::CollOfScalar(std::vector<double>(u, u + n)), true);
        equelleCUDA::CollOfScalar U0(std::vector<double>(u0, u0 + n));
        equelleCUDA::CollOfScalar r = prog.residual(U, U0, dt);
        info[0] = r.size();
        info[1] = r.data()->pat ? r.data()->pat->nnz : 0;
        if (!F) return;
        std::vector<double> fv = r.copyToHost();
        std::memcpy(F, fv.data(), fv.size() * sizeof(double));
        equelleCUDA::hostMat m = r.matrixToHost();
        std::memcpy(rp, m.rowPtr.data(), m.rowPtr.size() * sizeof(int));
        std::memcpy(ci, m.colInd.data(), m.colInd.size() * sizeof(int));
        std::memcpy(jv, m.vals.data(), m.vals.size() * sizeof(double));
    });
}

}  // extern "C"
