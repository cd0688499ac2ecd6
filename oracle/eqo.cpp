// eqo — CPU oracle (test infrastructure only; see eqo.hpp for the citations).
#include "eqo.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <sstream>

namespace eqo {

static void fail(const std::string& m) { throw std::runtime_error(m); }

// ============================================================================
// Grid: OPM GridManager Cartesian numbering.
//   2D: cells c = i + nx*j; x-faces f = i + (nx+1)*j (i = 0..nx) first, then
//       y-faces f = (nx+1)*ny + i + nx*j (j = 0..ny).  Pinned by
//       gridTest.cpp:107-171 (4x3 face ids and face_cells) and valsOnGrid.cpp.
//   3D: x-faces i + (nx+1)*(j + ny*k); y-faces nxf + i + nx*(j + (ny+1)*k);
//       z-faces nxf + nyf + i + nx*(j + ny*k).  face_cells = (-axis cell,
//       +axis cell), -1 outside; normals area-scaled and pointing +axis.
// ============================================================================
Grid make_cartesian(int dim, int nx, int ny, int nz, double dx, double dy, double dz) {
    if (dim != 2 && dim != 3) fail("Cannot handle " + std::to_string(dim) + " dimensions.");
    if (nx <= 0 || ny <= 0 || (dim == 3 && nz <= 0)) fail("grid sizes must be positive");
    Grid g;
    g.dim = dim;
    g.n[0] = nx; g.n[1] = ny; g.n[2] = dim == 3 ? nz : 1;
    g.h[0] = dx; g.h[1] = dy; g.h[2] = dim == 3 ? dz : 1.0;
    nz = g.n[2];
    const long nc = (long)nx * ny * nz;
    const long nxf = (long)(nx + 1) * ny * nz;
    const long nyf = (long)nx * (ny + 1) * nz;
    const long nzf = dim == 3 ? (long)nx * ny * (nz + 1) : 0;
    const long nf = nxf + nyf + nzf;
    g.nc = (int)nc;
    g.nf = (int)nf;
    g.face_cells.assign(2 * nf, -1);
    g.face_centroids.assign(dim * nf, 0.0);
    g.face_areas.assign(nf, 0.0);
    g.face_normals.assign(dim * nf, 0.0);
    g.cell_centroids.assign(dim * nc, 0.0);
    g.cell_volumes.assign(nc, 0.0);
    g.cell_facepos.resize(nc + 1);
    g.cell_faces.resize(2 * dim * nc);
    auto cell = [&](int i, int j, int k) { return i + nx * (j + ny * k); };
    const double ax = dim == 3 ? dy * dz : dy;
    const double ay = dim == 3 ? dx * dz : dx;
    const double az = dx * dy;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                long f = i + (long)(nx + 1) * (j + (long)ny * k);
                g.face_cells[2 * f] = i > 0 ? cell(i - 1, j, k) : -1;
                g.face_cells[2 * f + 1] = i < nx ? cell(i, j, k) : -1;
                g.face_areas[f] = ax;
                g.face_normals[dim * f + 0] = ax;
                g.face_centroids[dim * f + 0] = i * dx;
                g.face_centroids[dim * f + 1] = (j + 0.5) * dy;
                if (dim == 3) g.face_centroids[dim * f + 2] = (k + 0.5) * dz;
            }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long f = nxf + i + (long)nx * (j + (long)(ny + 1) * k);
                g.face_cells[2 * f] = j > 0 ? cell(i, j - 1, k) : -1;
                g.face_cells[2 * f + 1] = j < ny ? cell(i, j, k) : -1;
                g.face_areas[f] = ay;
                g.face_normals[dim * f + 1] = ay;
                g.face_centroids[dim * f + 0] = (i + 0.5) * dx;
                g.face_centroids[dim * f + 1] = j * dy;
                if (dim == 3) g.face_centroids[dim * f + 2] = (k + 0.5) * dz;
            }
    if (dim == 3)
        for (int k = 0; k <= nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    long f = nxf + nyf + i + (long)nx * (j + (long)ny * k);
                    g.face_cells[2 * f] = k > 0 ? cell(i, j, k - 1) : -1;
                    g.face_cells[2 * f + 1] = k < nz ? cell(i, j, k) : -1;
                    g.face_areas[f] = az;
                    g.face_normals[dim * f + 2] = az;
                    g.face_centroids[dim * f + 0] = (i + 0.5) * dx;
                    g.face_centroids[dim * f + 1] = (j + 0.5) * dy;
                    g.face_centroids[dim * f + 2] = k * dz;
                }
    const double vol = dim == 3 ? dx * dy * dz : dx * dy;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long c = cell(i, j, k);
                g.cell_volumes[c] = vol;
                g.cell_centroids[dim * c + 0] = (i + 0.5) * dx;
                g.cell_centroids[dim * c + 1] = (j + 0.5) * dy;
                if (dim == 3) g.cell_centroids[dim * c + 2] = (k + 0.5) * dz;
                long p = 2 * dim * c;
                g.cell_facepos[c] = (int)p;
                g.cell_faces[p + 0] = (int)(i + (long)(nx + 1) * (j + (long)ny * k));
                g.cell_faces[p + 1] = (int)(i + 1 + (long)(nx + 1) * (j + (long)ny * k));
                g.cell_faces[p + 2] = (int)(nxf + i + (long)nx * (j + (long)(ny + 1) * k));
                g.cell_faces[p + 3] = (int)(nxf + i + (long)nx * (j + 1 + (long)(ny + 1) * k));
                if (dim == 3) {
                    g.cell_faces[p + 4] = (int)(nxf + nyf + i + (long)nx * (j + (long)ny * k));
                    g.cell_faces[p + 5] = (int)(nxf + nyf + i + (long)nx * (j + (long)ny * (k + 1)));
                }
            }
    g.cell_facepos[nc] = (int)(2 * dim * nc);
    return g;
}

// ============================================================================
// CSR algebra
// ============================================================================
Csr csr_from_triplets(int rows, int cols, std::vector<Triplet> t) {
    std::sort(t.begin(), t.end(), [](const Triplet& a, const Triplet& b) {
        return a.r != b.r ? a.r < b.r : a.c < b.c;
    });
    Csr m;
    m.rows = rows;
    m.cols = cols;
    m.rp.assign(rows + 1, 0);
    for (size_t k = 0; k < t.size(); ++k) {
        if (!m.ci.empty() && k > 0 && t[k].r == t[k - 1].r && t[k].c == t[k - 1].c) {
            m.v.back() += t[k].v;   // setFromTriplets sums duplicates
            continue;
        }
        m.ci.push_back(t[k].c);
        m.v.push_back(t[k].v);
        m.rp[t[k].r + 1]++;
    }
    for (int r = 0; r < rows; ++r) m.rp[r + 1] += m.rp[r];
    return m;
}

Csr csr_identity(int n) {
    Csr m;
    m.rows = m.cols = n;
    m.rp.resize(n + 1);
    m.ci.resize(n);
    m.v.assign(n, 1.0);
    for (int i = 0; i <= n; ++i) m.rp[i] = i;
    for (int i = 0; i < n; ++i) m.ci[i] = i;
    return m;
}

Csr csr_diag(const std::vector<double>& d) {
    Csr m = csr_identity((int)d.size());
    m.v = d;
    return m;
}

Csr csr_transpose(const Csr& a) {
    if (a.empty()) return Csr();
    Csr t;
    t.rows = a.cols;
    t.cols = a.rows;
    t.rp.assign(a.cols + 1, 0);
    for (int c : a.ci) t.rp[c + 1]++;
    for (int r = 0; r < a.cols; ++r) t.rp[r + 1] += t.rp[r];
    t.ci.resize(a.nnz());
    t.v.resize(a.nnz());
    std::vector<int> pos(t.rp.begin(), t.rp.end() - 1);
    for (int r = 0; r < a.rows; ++r)
        for (int k = a.rp[r]; k < a.rp[r + 1]; ++k) {
            int p = pos[a.ci[k]]++;
            t.ci[p] = r;
            t.v[p] = a.v[k];
        }
    return t;
}

Csr csr_scale(const Csr& a, double s) {
    Csr m = a;
    for (double& x : m.v) x *= s;
    return m;
}

Csr csr_add(const Csr& a, double alpha, const Csr& b, double beta) {
    if (a.empty() && b.empty()) return Csr();
    if (a.empty()) return csr_scale(b, beta);
    if (b.empty()) return csr_scale(a, alpha);
    if (a.rows != b.rows || a.cols != b.cols) fail("csr_add: matrices of different sizes");
    Csr m;
    m.rows = a.rows;
    m.cols = a.cols;
    m.rp.assign(a.rows + 1, 0);
    m.ci.reserve(a.nnz() + b.nnz());
    m.v.reserve(a.nnz() + b.nnz());
    for (int r = 0; r < a.rows; ++r) {
        int ka = a.rp[r], ea = a.rp[r + 1], kb = b.rp[r], eb = b.rp[r + 1];
        while (ka < ea || kb < eb) {
            if (kb >= eb || (ka < ea && a.ci[ka] < b.ci[kb])) {
                m.ci.push_back(a.ci[ka]);
                m.v.push_back(alpha * a.v[ka]);
                ++ka;
            } else if (ka >= ea || b.ci[kb] < a.ci[ka]) {
                m.ci.push_back(b.ci[kb]);
                m.v.push_back(beta * b.v[kb]);
                ++kb;
            } else {
                m.ci.push_back(a.ci[ka]);
                m.v.push_back(alpha * a.v[ka] + beta * b.v[kb]);
                ++ka;
                ++kb;
            }
        }
        m.rp[r + 1] = m.nnz();
    }
    return m;
}

Csr csr_mul(const Csr& a, const Csr& b) {
    if (a.empty() || b.empty()) return Csr();
    if (a.cols != b.rows) fail("csr_mul: matrices of illegal sizes");
    Csr m;
    m.rows = a.rows;
    m.cols = b.cols;
    m.rp.assign(a.rows + 1, 0);
    std::vector<int> mark(b.cols, -1);
    std::vector<double> acc(b.cols, 0.0);
    std::vector<int> cols;
    for (int r = 0; r < a.rows; ++r) {
        cols.clear();
        for (int ka = a.rp[r]; ka < a.rp[r + 1]; ++ka) {
            const int k = a.ci[ka];
            const double av = a.v[ka];
            for (int kb = b.rp[k]; kb < b.rp[k + 1]; ++kb) {
                const int c = b.ci[kb];
                if (mark[c] != r) {
                    mark[c] = r;
                    acc[c] = 0.0;
                    cols.push_back(c);
                }
                acc[c] += av * b.v[kb];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int c : cols) {
            m.ci.push_back(c);
            m.v.push_back(acc[c]);
        }
        m.rp[r + 1] = m.nnz();
    }
    return m;
}

Csr csr_diag_scale(const std::vector<double>& d, const Csr& a) {
    if (a.empty()) return Csr();
    if ((int)d.size() != a.rows) fail("csr_diag_scale: size mismatch");
    Csr m = a;
    for (int r = 0; r < a.rows; ++r)
        for (int k = a.rp[r]; k < a.rp[r + 1]; ++k) m.v[k] *= d[r];
    return m;
}

std::vector<double> csr_spmv(const Csr& a, const std::vector<double>& x) {
    if ((int)x.size() != a.cols) fail("csr_spmv: size mismatch");
    std::vector<double> y(a.rows, 0.0);
    for (int r = 0; r < a.rows; ++r) {
        double s = 0.0;
        for (int k = a.rp[r]; k < a.rp[r + 1]; ++k) s += a.v[k] * x[a.ci[k]];
        y[r] = s;
    }
    return y;
}

// ============================================================================
// HelperOps: deviceHelperOps.cpp:84-148 (the construction of Opm::HelperOps).
// ============================================================================
HelperOps make_helper_ops(const Grid& g) {
    HelperOps h;
    std::vector<Triplet> ng, fng;
    for (int f = 0; f < g.nf; ++f) {
        const int c0 = g.face_cells[2 * f], c1 = g.face_cells[2 * f + 1];
        if (c0 >= 0 && c1 >= 0) {
            const int i = (int)h.internal_faces.size();
            h.internal_faces.push_back(f);
            ng.push_back({i, c0, 1.0});
            ng.push_back({i, c1, -1.0});
        }
        if (c0 >= 0) fng.push_back({f, c0, 1.0});
        if (c1 >= 0) fng.push_back({f, c1, -1.0});
    }
    const int nif = (int)h.internal_faces.size();
    h.ngrad = csr_from_triplets(nif, g.nc, ng);
    h.grad = csr_scale(h.ngrad, -1.0);
    h.div = csr_transpose(h.ngrad);
    h.fullngrad = csr_from_triplets(g.nf, g.nc, fng);
    h.fulldiv = csr_transpose(h.fullngrad);
    return h;
}

// ============================================================================
// AD rules (collOfScalar.cu:273-403, equelleTypes.hpp:86-211)
// ============================================================================
static void check_same(const ADB& a, const ADB& b, const char* op) {
    if (a.size() != b.size()) fail(std::string("size mismatch in ") + op);
}

ADB adb_add(const ADB& a, const ADB& b) {
    check_same(a, b, "+");
    ADB o;
    o.v.resize(a.size());
    for (int i = 0; i < a.size(); ++i) o.v[i] = a.v[i] + b.v[i];
    if (a.ad || b.ad) { o.J = csr_add(a.J, 1.0, b.J, 1.0); o.ad = true; }
    return o;
}

ADB adb_sub(const ADB& a, const ADB& b) {
    check_same(a, b, "-");
    ADB o;
    o.v.resize(a.size());
    for (int i = 0; i < a.size(); ++i) o.v[i] = a.v[i] - b.v[i];
    if (a.ad || b.ad) { o.J = csr_add(a.J, 1.0, b.J, -1.0); o.ad = true; }
    return o;
}

ADB adb_mul(const ADB& a, const ADB& b) {
    check_same(a, b, "*");
    ADB o;
    o.v.resize(a.size());
    for (int i = 0; i < a.size(); ++i) o.v[i] = a.v[i] * b.v[i];
    if (a.ad || b.ad) {
        // (uv)' = diag(v) u' + diag(u) v'
        o.J = csr_add(csr_diag_scale(b.v, a.J), 1.0, csr_diag_scale(a.v, b.J), 1.0);
        o.ad = true;
    }
    return o;
}

ADB adb_div(const ADB& a, const ADB& b) {
    check_same(a, b, "/");
    ADB o;
    o.v.resize(a.size());
    for (int i = 0; i < a.size(); ++i) o.v[i] = a.v[i] / b.v[i];
    if (a.ad || b.ad) {
        // (u/v)' = diag(1/v^2) (diag(v) u' - diag(u) v')
        std::vector<double> inv2(a.size());
        for (int i = 0; i < a.size(); ++i) inv2[i] = 1.0 / (b.v[i] * b.v[i]);
        o.J = csr_diag_scale(inv2, csr_add(csr_diag_scale(b.v, a.J), 1.0, csr_diag_scale(a.v, b.J), -1.0));
        o.ad = true;
    }
    return o;
}

ADB adb_scal_mul(double s, const ADB& a) {
    ADB o;
    o.v.resize(a.size());
    for (int i = 0; i < a.size(); ++i) o.v[i] = s * a.v[i];
    if (a.ad) { o.J = csr_scale(a.J, s); o.ad = true; }
    return o;
}

ADB adb_scal_div_coll(double s, const ADB& a) {
    ADB o;
    o.v.resize(a.size());
    for (int i = 0; i < a.size(); ++i) o.v[i] = s / a.v[i];
    if (a.ad) {
        // (s/u)' = -diag(s/u^2) u'
        std::vector<double> d(a.size());
        for (int i = 0; i < a.size(); ++i) d[i] = -(s / (a.v[i] * a.v[i]));
        o.J = csr_diag_scale(d, a.J);
        o.ad = true;
    }
    return o;
}

ADB adb_neg(const ADB& a) { return adb_scal_mul(-1.0, a); }

// norm(CollOfScalar) (EquelleRuntimeCPU.cpp:241-254): a constant collection of factors,
// +1 where the value is >= 0 and -1 elsewhere, times the collection
ADB adb_norm(const ADB& a) {
    ADB f;
    f.v.resize(a.size());
    for (int i = 0; i < a.size(); ++i) f.v[i] = a.v[i] >= 0.0 ? 1.0 : -1.0;
    return adb_mul(f, a);
}

ADB adb_sqrt(const ADB& a) {
    ADB o;
    o.v.resize(a.size());
    for (int i = 0; i < a.size(); ++i) o.v[i] = std::sqrt(a.v[i]);
    if (a.ad) {
        std::vector<double> d(a.size());
        for (int i = 0; i < a.size(); ++i) d[i] = 0.5 / o.v[i];
        o.J = csr_diag_scale(d, a.J);
        o.ad = true;
    }
    return o;
}

ADB adb_matmul(const Csr& m, const ADB& a) {
    ADB o;
    o.v = csr_spmv(m, a.v);
    if (a.ad) { o.J = csr_mul(m, a.J); o.ad = true; }
    return o;
}

ADB adb_gather(const ADB& a, const std::vector<int>& rows) {
    ADB o;
    o.v.resize(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) o.v[i] = a.v[rows[i]];
    o.ad = a.ad;
    if (a.ad && !a.J.empty()) {
        Csr& m = o.J;
        m.rows = (int)rows.size();
        m.cols = a.J.cols;
        m.rp.assign(rows.size() + 1, 0);
        for (size_t i = 0; i < rows.size(); ++i) {
            const int r = rows[i];
            for (int k = a.J.rp[r]; k < a.J.rp[r + 1]; ++k) {
                m.ci.push_back(a.J.ci[k]);
                m.v.push_back(a.J.v[k]);
            }
            m.rp[i + 1] = m.nnz();
        }
    }
    return o;
}

ADB adb_scatter(const ADB& a, const std::vector<int>& rows, int n) {
    ADB o;
    o.v.assign(n, 0.0);
    for (size_t i = 0; i < rows.size(); ++i) o.v[rows[i]] = a.v[i];
    o.ad = a.ad;
    if (a.ad && !a.J.empty()) {
        std::vector<int> src(n, -1);
        for (size_t i = 0; i < rows.size(); ++i) src[rows[i]] = (int)i;
        Csr& m = o.J;
        m.rows = n;
        m.cols = a.J.cols;
        m.rp.assign(n + 1, 0);
        for (int r = 0; r < n; ++r) {
            const int s = src[r];
            if (s >= 0)
                for (int k = a.J.rp[s]; k < a.J.rp[s + 1]; ++k) {
                    m.ci.push_back(a.J.ci[k]);
                    m.v.push_back(a.J.v[k]);
                }
            m.rp[r + 1] = m.nnz();
        }
    }
    return o;
}

ADB adb_select(const std::vector<uint8_t>& p, const ADB& t, const ADB& f) {
    // EquelleRuntimeCPU_impl.hpp:163-180: iftrue*trueones + iffalse*falseones.
    const int n = (int)p.size();
    if (t.size() != n || f.size() != n) fail("Collections are not of the same size");
    ADB o;
    o.v.resize(n);
    std::vector<double> one(n), zero(n);
    for (int i = 0; i < n; ++i) {
        o.v[i] = p[i] ? t.v[i] : f.v[i];
        one[i] = p[i] ? 1.0 : 0.0;
        zero[i] = p[i] ? 0.0 : 1.0;
    }
    if (t.ad || f.ad) {
        o.J = csr_add(csr_diag_scale(one, t.J), 1.0, csr_diag_scale(zero, f.J), 1.0);
        o.ad = true;
    }
    return o;
}

ADB primary_variable(const ADB& u0) {
    ADB u;
    u.v = u0.v;
    u.J = csr_identity(u0.size());
    u.ad = true;
    return u;
}

// ============================================================================
// Index sets
// ============================================================================
std::vector<int> subset_indices(const IndexSet& super, const IndexSet& sub) {
    // EquelleRuntimeCPU_impl.hpp:28-83: for each element of sub, its position in super.
    std::vector<int> out(sub.size(), -1);
    if (sub.size() == 0) return out;
    std::vector<std::pair<int, int>> s(sub.size());
    for (int i = 0; i < sub.size(); ++i) s[i] = {sub.idx[i], i};
    std::sort(s.begin(), s.end());
    size_t k = 0;
    for (int i = 0; i < super.size() && k < s.size(); ++i)
        while (k < s.size() && s[k].first == super.idx[i]) out[s[k++].second] = i;
    if (k != s.size()) fail("subset_indices: set is not a subset of the superset");
    return out;
}

// ============================================================================
// Linear solvers
// ============================================================================
LinSolveReport solve_banded_direct(const Csr& a, const std::vector<double>& b, std::vector<double>& x) {
    const int n = a.rows;
    int kl = 0, ku = 0;
    for (int r = 0; r < n; ++r)
        for (int k = a.rp[r]; k < a.rp[r + 1]; ++k) {
            kl = std::max(kl, r - a.ci[k]);
            ku = std::max(ku, a.ci[k] - r);
        }
    // LAPACK-style band storage with room for fill from partial pivoting.
    const int ldab = 2 * kl + ku + 1;
    std::vector<double> ab((size_t)ldab * n, 0.0);
    auto AB = [&](int i, int j) -> double& { return ab[(size_t)j * ldab + (kl + ku + i - j)]; };
    for (int r = 0; r < n; ++r)
        for (int k = a.rp[r]; k < a.rp[r + 1]; ++k) AB(r, a.ci[k]) += a.v[k];
    std::vector<double> rhs = b;
    std::vector<int> piv(n);
    const int kv = ku + kl;
    double amax = 0.0;
    for (double v : a.v) amax = std::max(amax, std::fabs(v));
    // A connected pure-Neumann operator (the two-phase pressure equation with rate
    // sources, alone or as the pressure block of the fully implicit system) is singular
    // with the constants (of the pressure unknowns) as null space: elimination meets a
    // round-off pivot in the column where the dependency closes (the last one for the
    // pressure alone).  The right-hand side is consistent up to rounding and any
    // solution is valid, so such a column's unknown is set to 0 and its row, a
    // combination of the others, is left out (the residual returned below shows any
    // inconsistency).
    std::vector<char> free_col(n, 0);
    for (int j = 0; j < n; ++j) {
        const int km = std::min(kl, n - 1 - j);
        int p = j;
        double best = std::fabs(AB(j, j));
        for (int i = j + 1; i <= j + km; ++i)
            if (std::fabs(AB(i, j)) > best) { best = std::fabs(AB(i, j)); p = i; }
        piv[j] = p;
        if (n > 1 && best <= 1e-11 * amax) {
            free_col[j] = 1;
            continue;
        }
        if (best == 0.0) { LinSolveReport r; return r; }
        const int ju = std::min(n - 1, j + kv);
        if (p != j) {
            for (int c = j; c <= ju; ++c) std::swap(AB(j, c), AB(p, c));
            std::swap(rhs[j], rhs[p]);
        }
        const double d = AB(j, j);
        for (int i = j + 1; i <= j + km; ++i) {
            const double l = AB(i, j) / d;
            AB(i, j) = l;
            if (l == 0.0) continue;
            for (int c = j + 1; c <= ju; ++c) AB(i, c) -= l * AB(j, c);
            rhs[i] -= l * rhs[j];
        }
    }
    x.assign(n, 0.0);
    for (int j = n - 1; j >= 0; --j) {
        double s = rhs[j];
        const int ju = std::min(n - 1, j + kv);
        for (int c = j + 1; c <= ju; ++c) s -= AB(j, c) * x[c];
        x[j] = free_col[j] ? 0.0 : s / AB(j, j);
    }
    LinSolveReport rep;
    rep.converged = true;
    rep.iterations = 1;
    std::vector<double> ax = csr_spmv(a, x);
    double rr = 0, bb = 0;
    for (int i = 0; i < n; ++i) { rr += (b[i] - ax[i]) * (b[i] - ax[i]); bb += b[i] * b[i]; }
    rep.residual = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    return rep;
}

namespace {
struct Ilu0 {
    const Csr* a;
    std::vector<double> lu;
    std::vector<int> diag;
    explicit Ilu0(const Csr& m) : a(&m), lu(m.v), diag(m.rows, -1) {
        const int n = m.rows;
        for (int r = 0; r < n; ++r)
            for (int k = m.rp[r]; k < m.rp[r + 1]; ++k)
                if (m.ci[k] == r) diag[r] = k;
        std::vector<int> pos(m.cols, -1);
        for (int i = 0; i < n; ++i) {
            for (int k = m.rp[i]; k < m.rp[i + 1]; ++k) pos[m.ci[k]] = k;
            for (int k = m.rp[i]; k < m.rp[i + 1] && m.ci[k] < i; ++k) {
                const int j = m.ci[k];
                if (diag[j] < 0) continue;
                const double l = lu[k] / lu[diag[j]];
                lu[k] = l;
                for (int kk = diag[j] + 1; kk < m.rp[j + 1]; ++kk) {
                    const int p = pos[m.ci[kk]];
                    if (p >= 0) lu[p] -= l * lu[kk];
                }
            }
            for (int k = m.rp[i]; k < m.rp[i + 1]; ++k) pos[m.ci[k]] = -1;
        }
    }
    void apply(const std::vector<double>& r, std::vector<double>& z) const {
        const Csr& m = *a;
        const int n = m.rows;
        z.resize(n);
        for (int i = 0; i < n; ++i) {
            double s = r[i];
            for (int k = m.rp[i]; k < m.rp[i + 1] && m.ci[k] < i; ++k) s -= lu[k] * z[m.ci[k]];
            z[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = z[i];
            for (int k = diag[i] + 1; k < m.rp[i + 1]; ++k) s -= lu[k] * z[m.ci[k]];
            z[i] = s / lu[diag[i]];
        }
    }
};
double dotv(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}
}  // namespace

LinSolveReport solve_bicgstab_ilu0(const Csr& a, const std::vector<double>& b, std::vector<double>& x,
                                   double rtol, int maxit) {
    const int n = a.rows;
    Ilu0 M(a);
    x.assign(n, 0.0);
    std::vector<double> r = b, rh = b, p(n, 0.0), v(n, 0.0), s(n), t(n), ph, sh;
    double rho = 1, alpha = 1, omega = 1;
    const double bn = std::sqrt(dotv(b, b));
    LinSolveReport rep;
    if (bn == 0.0) { rep.converged = true; return rep; }
    double rn = bn;
    int it = 0;
    for (; it < maxit && rn > rtol * bn; ++it) {
        const double rho1 = dotv(rh, r);
        if (rho1 == 0.0) break;
        const double beta = (rho1 / rho) * (alpha / omega);
        for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        M.apply(p, ph);
        v = csr_spmv(a, ph);
        alpha = rho1 / dotv(rh, v);
        for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
        M.apply(s, sh);
        t = csr_spmv(a, sh);
        const double tt = dotv(t, t);
        omega = tt > 0 ? dotv(t, s) / tt : 0.0;
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * ph[i] + omega * sh[i];
            r[i] = s[i] - omega * t[i];
        }
        rho = rho1;
        rn = std::sqrt(dotv(r, r));
        if (omega == 0.0) { ++it; break; }
    }
    rep.iterations = it;
    rep.residual = rn / bn;
    rep.converged = rn <= rtol * bn;
    return rep;
}

// ============================================================================
// Params (Opm::ParameterGroup key=value subset)
// ============================================================================
void Params::parse(const std::string& text) {
    std::string tok;
    std::istringstream is(text);
    while (is >> tok) {
        const size_t e = tok.find('=');
        if (e == std::string::npos) continue;
        kv[tok.substr(0, e)] = tok.substr(e + 1);
    }
}
double Params::get_double(const std::string& k, double def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::strtod(it->second.c_str(), nullptr);
}
int Params::get_int(const std::string& k, int def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : (int)std::strtol(it->second.c_str(), nullptr, 10);
}
std::string Params::get_string(const std::string& k, const std::string& def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : it->second;
}

// ============================================================================
// Runtime (EquelleRuntimeCPU.cpp)
// ============================================================================
Runtime::Runtime(const Params& p) {
    // createGridManager (EquelleRuntimeCPU.cpp:20-49): defaults 6x1(x1), unit cells.
    const int dim = p.get_int("grid_dim", 2);
    int nz = 1;
    double dz = 1.0;
    if (dim == 3) { nz = p.get_int("nz", 1); dz = p.get_double("dz", 1.0); }
    grid_ = make_cartesian(dim, p.get_int("nx", 6), p.get_int("ny", 1), nz,
                           p.get_double("dx", 1.0), p.get_double("dy", 1.0), dz);
    ops_ = make_helper_ops(grid_);
    max_iter = p.get_int("max_iter", 10);
    abs_res_tol = p.get_double("abs_res_tol", 1e-6);
    rel_res_tol = p.get_double("rel_res_tol", 0.0);
    linear_solver = p.get_string("oracle_linear_solver", "auto");
    solver_tol = p.get_double("oracle_solver_tol", 1e-12);
    solver_max_iter = p.get_int("oracle_solver_max_iter", 5000);
}

static IndexSet make_set(int codim, bool full, std::vector<int> idx) {
    IndexSet s;
    s.codim = codim;
    s.full = full;
    s.idx = std::move(idx);
    return s;
}

IndexSet Runtime::allCells() const {
    std::vector<int> v(grid_.nc);
    std::iota(v.begin(), v.end(), 0);
    return make_set(0, true, v);
}
IndexSet Runtime::allFaces() const {
    std::vector<int> v(grid_.nf);
    std::iota(v.begin(), v.end(), 0);
    return make_set(1, true, v);
}
IndexSet Runtime::interiorFaces() const { return make_set(1, false, ops_.internal_faces); }
IndexSet Runtime::boundaryFaces() const {
    std::vector<int> v;
    for (int f = 0; f < grid_.nf; ++f)
        if (grid_.face_cells[2 * f] < 0 || grid_.face_cells[2 * f + 1] < 0) v.push_back(f);
    return make_set(1, false, v);
}
static bool is_boundary_cell(const Grid& g, int c) {
    for (int k = g.cell_facepos[c]; k < g.cell_facepos[c + 1]; ++k) {
        const int f = g.cell_faces[k];
        if (g.face_cells[2 * f] < 0 || g.face_cells[2 * f + 1] < 0) return true;
    }
    return false;
}
IndexSet Runtime::boundaryCells() const {
    std::vector<int> v;
    for (int c = 0; c < grid_.nc; ++c)
        if (is_boundary_cell(grid_, c)) v.push_back(c);
    return make_set(0, false, v);
}
IndexSet Runtime::interiorCells() const {
    std::vector<int> v;
    for (int c = 0; c < grid_.nc; ++c)
        if (!is_boundary_cell(grid_, c)) v.push_back(c);
    return make_set(0, false, v);
}
IndexSet Runtime::firstCell(const IndexSet& faces) const {
    std::vector<int> v(faces.size());
    for (int i = 0; i < faces.size(); ++i) v[i] = grid_.face_cells[2 * faces.idx[i]];
    return make_set(0, false, v);
}
IndexSet Runtime::secondCell(const IndexSet& faces) const {
    std::vector<int> v(faces.size());
    for (int i = 0; i < faces.size(); ++i) v[i] = grid_.face_cells[2 * faces.idx[i] + 1];
    return make_set(0, false, v);
}
std::vector<uint8_t> Runtime::isEmpty(const IndexSet& s) const {
    std::vector<uint8_t> o(s.size());
    for (int i = 0; i < s.size(); ++i) o[i] = s.idx[i] < 0;
    return o;
}

ADB Runtime::norm(const IndexSet& s) const {
    ADB o;
    o.v.resize(s.size());
    for (int i = 0; i < s.size(); ++i)
        o.v[i] = s.codim == 0 ? grid_.cell_volumes[s.idx[i]] : grid_.face_areas[s.idx[i]];
    return o;
}
std::vector<double> Runtime::centroid(const IndexSet& s) const {
    const int d = grid_.dim;
    std::vector<double> o((size_t)s.size() * d);
    const std::vector<double>& src = s.codim == 0 ? grid_.cell_centroids : grid_.face_centroids;
    for (int i = 0; i < s.size(); ++i)
        for (int k = 0; k < d; ++k) o[(size_t)i * d + k] = src[(size_t)s.idx[i] * d + k];
    return o;
}
std::vector<double> Runtime::normal(const IndexSet& s) const {
    // EquelleRuntimeCPU.cpp:295-315: area-scaled normals renormalised to unit length.
    const int d = grid_.dim;
    std::vector<double> o((size_t)s.size() * d);
    for (int i = 0; i < s.size(); ++i) {
        double nn = 0;
        for (int k = 0; k < d; ++k) {
            const double x = grid_.face_normals[(size_t)s.idx[i] * d + k];
            nn += x * x;
        }
        nn = std::sqrt(nn);
        for (int k = 0; k < d; ++k) o[(size_t)i * d + k] = grid_.face_normals[(size_t)s.idx[i] * d + k] / nn;
    }
    return o;
}

ADB Runtime::gradient(const ADB& x) const { return adb_matmul(ops_.grad, x); }
ADB Runtime::negGradient(const ADB& x) const { return adb_matmul(ops_.ngrad, x); }
ADB Runtime::divergence(const ADB& x) const {
    if (x.size() == (int)ops_.internal_faces.size()) return interiorDivergence(x);
    return adb_matmul(ops_.fulldiv, x);
}
ADB Runtime::interiorDivergence(const ADB& x) const { return adb_matmul(ops_.div, x); }

ADB Runtime::operatorExtend(const ADB& x, const IndexSet& from, const IndexSet& to) const {
    if (x.size() != from.size()) fail("Extend: data and from_set sizes differ");
    return adb_scatter(x, subset_indices(to, from), to.size());
}
ADB Runtime::operatorExtend(double s, const IndexSet& to) const {
    ADB o;
    o.v.assign(to.size(), s);
    return o;
}
ADB Runtime::operatorOn(const ADB& x, const IndexSet& from, const IndexSet& to) const {
    if (x.size() != from.size()) fail("On: data and from_set sizes differ");
    return adb_gather(x, subset_indices(from, to));
}
IndexSet Runtime::operatorOn(const IndexSet& x, const IndexSet& from, const IndexSet& to) const {
    if (x.size() != from.size()) fail("On: data and from_set sizes differ");
    std::vector<int> pos = subset_indices(from, to);
    std::vector<int> v(pos.size());
    for (size_t i = 0; i < pos.size(); ++i) v[i] = x.idx[pos[i]];
    return make_set(x.codim, false, v);
}
ADB Runtime::trinaryIf(const std::vector<uint8_t>& p, const ADB& t, const ADB& f) const {
    return adb_select(p, t, f);
}
IndexSet Runtime::trinaryIf(const std::vector<uint8_t>& p, const IndexSet& t, const IndexSet& f) const {
    if ((int)p.size() != t.size() || (int)p.size() != f.size()) fail("The sets are not of the same size");
    std::vector<int> v(p.size());
    for (size_t i = 0; i < p.size(); ++i) v[i] = p[i] ? t.idx[i] : f.idx[i];
    return make_set(t.codim, false, v);
}

double Runtime::minReduce(const ADB& x) const { return *std::min_element(x.v.begin(), x.v.end()); }
double Runtime::maxReduce(const ADB& x) const { return *std::max_element(x.v.begin(), x.v.end()); }
double Runtime::sumReduce(const ADB& x) const {
    double s = 0;
    for (double v : x.v) s += v;
    return s;
}
double Runtime::prodReduce(const ADB& x) const {
    double s = 1;
    for (double v : x.v) s *= v;
    return s;
}
double Runtime::twoNorm(const ADB& x) const { return std::sqrt(dotv(x.v, x.v)); }

std::vector<double> Runtime::solveForUpdate(const ADB& residual, LinSolveReport* rep, int blocks) const {
    if (residual.J.empty()) fail("solveForUpdate: residual has no Jacobian");
    std::vector<double> du;
    LinSolveReport r;
    const int N = residual.size();
    if (blocks > 1) {
        // Block-ordered unknowns (k * m + c, the system of newtonSolveSystem) are solved in
        // interleaved order (B * c + k), which keeps the band of a stencil system narrow.
        if (N % blocks) fail("solveForUpdate: size not a multiple of the block count");
        const int m = N / blocks;
        std::vector<int> perm(N);
        for (int k = 0; k < blocks; ++k)
            for (int c = 0; c < m; ++c) perm[k * m + c] = blocks * c + k;
        const Csr& a = residual.J;
        std::vector<Triplet> t;
        for (int i = 0; i < N; ++i)
            for (int q = a.rp[i]; q < a.rp[i + 1]; ++q) t.push_back({perm[i], perm[a.ci[q]], a.v[q]});
        Csr pa = csr_from_triplets(N, N, t);
        std::vector<double> pb(N), px;
        for (int i = 0; i < N; ++i) pb[perm[i]] = residual.v[i];
        const bool direct = linear_solver == "direct" || (linear_solver == "auto" && N <= 20000);
        if (direct) r = solve_banded_direct(pa, pb, px);
        else r = solve_bicgstab_ilu0(pa, pb, px, solver_tol, solver_max_iter);
        if (r.converged) {
            du.resize(N);
            for (int i = 0; i < N; ++i) du[i] = px[perm[i]];
        }
    } else {
        const bool direct = linear_solver == "direct" || (linear_solver == "auto" && N <= 40000);
        if (direct) r = solve_banded_direct(residual.J, residual.v, du);
        else r = solve_bicgstab_ilu0(residual.J, residual.v, du, solver_tol, solver_max_iter);
    }
    if (!r.converged) fail("Linear solver convergence failure.");
    if (rep) *rep = r;
    return du;
}

ADB adb_subset(const ADB& u, int start, int len) {
    std::vector<int> rows(len);
    for (int i = 0; i < len; ++i) rows[i] = start + i;
    return adb_gather(u, rows);
}

ADB adb_superset(const ADB& a, int start, int total) {
    std::vector<int> rows(a.size());
    for (int i = 0; i < a.size(); ++i) rows[i] = start + i;
    return adb_scatter(a, rows, total);
}

}  // namespace eqo
