// Runtime core: device grid, HelperOps and grid built-ins (see runtime_core.hpp).
#include "runtime_core.hpp"

#include "tape.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>

#include "kernels.hpp"

namespace eqhip {

// ---------------------------------------------------------------- params
void ParamMap::parse(const std::string& text) {
    std::istringstream is(text);
    std::string tok;
    while (is >> tok) {
        const size_t e = tok.find('=');
        if (e != std::string::npos) kv[tok.substr(0, e)] = tok.substr(e + 1);
    }
}
double ParamMap::get_double(const std::string& k, double def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::strtod(it->second.c_str(), nullptr);
}
int ParamMap::get_int(const std::string& k, int def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : (int)std::strtol(it->second.c_str(), nullptr, 10);
}
bool ParamMap::get_bool(const std::string& k, bool def) const {
    auto it = kv.find(k);
    if (it == kv.end()) return def;
    return it->second == "true" || it->second == "1";
}
std::string ParamMap::get_string(const std::string& k, const std::string& def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : it->second;
}
std::string ParamMap::require(const std::string& k) const {
    auto it = kv.find(k);
    EQ_REQUIRE(it != kv.end(), "parameter " << k << " not found");
    return it->second;
}

// ---------------------------------------------------------------- grid
// Numbering of OPM's Cartesian GridManager: cells x-fastest; x-faces, then
// y-faces, then z-faces, each x-fastest; face_cells = (-axis cell, +axis cell)
// with -1 outside; normals area-scaled along +axis; cell_faces x-,x+,y-,y+,z-,z+.
GridHost build_cartesian(int dim, int nx, int ny, int nz, double dx, double dy, double dz) {
    EQ_REQUIRE(dim == 2 || dim == 3, "Cannot handle " << dim << " dimensions.");
    EQ_REQUIRE(nx > 0 && ny > 0 && (dim == 2 || nz > 0), "grid sizes must be positive");
    GridHost g;
    g.dim = dim;
    if (dim == 2) {
        nz = 1;
        dz = 1.0;
    }
    g.n[0] = nx;
    g.n[1] = ny;
    g.n[2] = nz;
    g.h[0] = dx;
    g.h[1] = dy;
    g.h[2] = dz;
    const long nc = (long)nx * ny * nz;
    const long fx = (long)(nx + 1) * ny * nz, fy = (long)nx * (ny + 1) * nz;
    const long fz = dim == 3 ? (long)nx * ny * (nz + 1) : 0;
    const long nf = fx + fy + fz;
    EQ_REQUIRE(nf < (1L << 31) && 2 * nf < (1L << 31), "grid too large for 32-bit indices");
    g.nc = (int)nc;
    g.nf = (int)nf;
    g.face_cells.assign(2 * nf, -1);
    g.face_areas.resize(nf);
    g.face_centroids.assign(dim * nf, 0.0);
    g.face_normals.assign(dim * nf, 0.0);
    const double area[3] = {dim == 3 ? dy * dz : dy, dim == 3 ? dx * dz : dx, dx * dy};
    auto C = [&](long i, long j, long k) { return (int)(i + nx * (j + (long)ny * k)); };
    // face blocks: axis a, extents (nx+da, ny+db, nz+dc)
    long base = 0;
    for (int a = 0; a < dim; ++a) {
        const long ex = nx + (a == 0), ey = ny + (a == 1), ez = nz + (a == 2);
        for (long k = 0; k < ez; ++k)
            for (long j = 0; j < ey; ++j)
                for (long i = 0; i < ex; ++i) {
                    const long f = base + i + ex * (j + ey * k);
                    long lo[3] = {i, j, k};
                    long hi[3] = {i, j, k};
                    lo[a] -= 1;
                    const long ext[3] = {nx, ny, nz};
                    g.face_cells[2 * f] = lo[a] >= 0 ? C(lo[0], lo[1], lo[2]) : -1;
                    g.face_cells[2 * f + 1] = hi[a] < ext[a] ? C(hi[0], hi[1], hi[2]) : -1;
                    g.face_areas[f] = area[a];
                    g.face_normals[dim * f + a] = area[a];
                    const double cc[3] = {(i + (a == 0 ? 0.0 : 0.5)) * dx, (j + (a == 1 ? 0.0 : 0.5)) * dy,
                                          (k + (a == 2 ? 0.0 : 0.5)) * dz};
                    for (int d = 0; d < dim; ++d) g.face_centroids[dim * f + d] = cc[d];
                }
        base += ex * ey * ez;
    }
    g.cell_volumes.assign(nc, dim == 3 ? dx * dy * dz : dx * dy);
    g.cell_centroids.resize(dim * nc);
    g.cell_facepos.resize(nc + 1);
    g.cell_faces.resize(2 * dim * nc);
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                const long c = C(i, j, k);
                g.cell_centroids[dim * c] = (i + 0.5) * dx;
                g.cell_centroids[dim * c + 1] = (j + 0.5) * dy;
                if (dim == 3) g.cell_centroids[dim * c + 2] = (k + 0.5) * dz;
                const long p = 2 * dim * c;
                g.cell_facepos[c] = (int)p;
                g.cell_faces[p + 0] = (int)(i + (nx + 1) * (j + (long)ny * k));
                g.cell_faces[p + 1] = (int)(i + 1 + (nx + 1) * (j + (long)ny * k));
                g.cell_faces[p + 2] = (int)(fx + i + (long)nx * (j + (long)(ny + 1) * k));
                g.cell_faces[p + 3] = (int)(fx + i + (long)nx * (j + 1 + (long)(ny + 1) * k));
                if (dim == 3) {
                    g.cell_faces[p + 4] = (int)(fx + fy + i + (long)nx * (j + (long)ny * k));
                    g.cell_faces[p + 5] = (int)(fx + fy + i + (long)nx * (j + (long)ny * (k + 1)));
                }
            }
    g.cell_facepos[nc] = (int)(2 * dim * nc);
    return g;
}

static IP compact_to_set(const DevPtr<uint8_t>& flags, int n, int want, int codim) {
    auto d = dev_alloc<int>(n > 0 ? n : 1);
    const int cnt = compact_flags(flags->p, n, want, d->p);
    return i_from_device(codim, d, cnt);
}

DeviceGrid::DeviceGrid(GridHost g) : host(std::move(g)) {
    dim = host.dim;
    nc = host.nc;
    nf = host.nf;
    face_cells = dev_from_host(host.face_cells.data(), host.face_cells.size());
    cell_facepos = dev_from_host(host.cell_facepos.data(), host.cell_facepos.size());
    cell_faces = dev_from_host(host.cell_faces.data(), host.cell_faces.size());
    cell_centroids = dev_from_host(host.cell_centroids.data(), host.cell_centroids.size());
    cell_volumes = dev_from_host(host.cell_volumes.data(), host.cell_volumes.size());
    face_centroids = dev_from_host(host.face_centroids.data(), host.face_centroids.size());
    face_areas = dev_from_host(host.face_areas.data(), host.face_areas.size());
    face_normals = dev_from_host(host.face_normals.data(), host.face_normals.size());
    // signed cell->face list for the divergence sweep (+f if the cell is the first cell)
    std::vector<int> sf(host.cell_faces.size());
    for (int c = 0; c < nc; ++c)
        for (int k = host.cell_facepos[c]; k < host.cell_facepos[c + 1]; ++k) {
            const int f = host.cell_faces[k];
            sf[k] = host.face_cells[2 * f] == c ? f : ~f;
        }
    signed_faces = dev_from_host(sf.data(), sf.size());
    // subsets: device flag kernels + ordered compaction (deviceGrid.cu:368-492)
    std::vector<int> all(std::max(nc, nf));
    for (size_t i = 0; i < all.size(); ++i) all[i] = (int)i;
    all_cells = make_index(0, true, std::vector<int>(all.begin(), all.begin() + nc));
    all_faces = make_index(1, true, std::vector<int>(all.begin(), all.begin() + nf));
    auto fflag = dev_alloc<uint8_t>(nf);
    launch_face_flags(face_cells->p, nf, fflag->p);
    interior_faces = compact_to_set(fflag, nf, 1, 1);
    boundary_faces = compact_to_set(fflag, nf, 0, 1);
    auto cflag = dev_alloc<uint8_t>(nc);
    launch_cell_boundary_flags(face_cells->p, cell_facepos->p, cell_faces->p, nc, cflag->p);
    boundary_cells = compact_to_set(cflag, nc, 1, 0);
    interior_cells = compact_to_set(cflag, nc, 0, 0);
    nif = interior_faces->size();
    // interior face cell pairs for the gradient sweep
    if_c1 = dev_alloc<int>(nif > 0 ? nif : 1);
    if_c2 = dev_alloc<int>(nif > 0 ? nif : 1);
    launch_gather_face_cells(face_cells->p, interior_faces->dev->p, nif, 0, if_c1->p);
    launch_gather_face_cells(face_cells->p, interior_faces->dev->p, nif, 1, if_c2->p);
    // HelperOps: grad = -ngrad (nif x nc), div = ngrad^T (nc x nif), fulldiv = fullngrad^T (nc x nf)
    {
        std::vector<int> rp(nif + 1), ci(2 * (size_t)nif);
        std::vector<double> v(2 * (size_t)nif);
        for (int i = 0; i < nif; ++i) {
            const int f = interior_faces->host[i];
            rp[i] = 2 * i;
            ci[2 * i] = host.face_cells[2 * f];
            ci[2 * i + 1] = host.face_cells[2 * f + 1];
            v[2 * i] = -1.0;
            v[2 * i + 1] = 1.0;
            if (ci[2 * i] > ci[2 * i + 1]) {
                std::swap(ci[2 * i], ci[2 * i + 1]);
                std::swap(v[2 * i], v[2 * i + 1]);
            }
        }
        rp[nif] = 2 * nif;
        grad = csr_from_host(nif, nc, rp.data(), ci.data(), v.data());
        for (double& x : v) x = -x;   // ngrad: +1 first cell, -1 second cell (same pattern)
        ngrad = csr_from_host(nif, nc, rp.data(), ci.data(), v.data());
    }
    auto transpose_signed = [&](bool interior_only) {
        // rows = cells; entries over the cell's faces in ascending face order:
        // +1 where the cell is the first cell, -1 where it is the second.
        std::vector<int> fidx(nf, -1);
        if (interior_only)
            for (int i = 0; i < nif; ++i) fidx[interior_faces->host[i]] = i;
        std::vector<int> rp(nc + 1, 0), ci;
        std::vector<double> v;
        ci.reserve(host.cell_faces.size());
        v.reserve(host.cell_faces.size());
        std::vector<std::pair<int, double>> row;
        for (int c = 0; c < nc; ++c) {
            row.clear();
            for (int k = host.cell_facepos[c]; k < host.cell_facepos[c + 1]; ++k) {
                const int f = host.cell_faces[k];
                const int col = interior_only ? fidx[f] : f;
                if (col < 0) continue;
                row.emplace_back(col, host.face_cells[2 * f] == c ? 1.0 : -1.0);
            }
            std::sort(row.begin(), row.end());
            for (auto& e : row) {
                ci.push_back(e.first);
                v.push_back(e.second);
            }
            rp[c + 1] = (int)ci.size();
        }
        return csr_from_host(nc, interior_only ? nif : nf, rp.data(), ci.data(), v.data());
    };
    div = transpose_signed(true);
    fulldiv = transpose_signed(false);
}

// ---------------------------------------------------------------- runtime
RuntimeCore::RuntimeCore(const ParamMap& p) : params_(p) {
    EQ_REQUIRE(!p.has("grid_filename"), "grid_filename: only Cartesian grids (grid_dim, nx, ny, nz, dx, dy, dz) "
                                        "are supported by this runtime");
    // createGridManager (EquelleRuntimeCUDA.cpp:29-60): defaults 6 x 1 (x 1), unit cells
    const int dim = p.get_int("grid_dim", 2);
    EQ_REQUIRE(dim == 2 || dim == 3, "Cannot handle " << dim << " dimensions.");
    grid_.reset(new DeviceGrid(build_cartesian(dim, p.get_int("nx", 6), p.get_int("ny", 1),
                                               dim == 3 ? p.get_int("nz", 1) : 1, p.get_double("dx", 1.0),
                                               p.get_double("dy", 1.0), dim == 3 ? p.get_double("dz", 1.0) : 1.0)));
    // solver options (EquelleRuntimeCUDA.cpp:70-79, linearSolver.cu:33-65)
    std::string s = p.get_string("solver", "BiCGStab");
    std::string sl = s;
    std::transform(sl.begin(), sl.end(), sl.begin(), ::tolower);
    EQ_REQUIRE(sl != "cpu", "solver=CPU (the host linear solver) is not part of this runtime; "
                            "use BiCGStab, CG or GMRes");
    solver.method = sl == "gmres" ? "gmres" : sl == "cg" ? "cg" : "bicgstab";
    EQ_REQUIRE(sl == "gmres" || sl == "cg" || sl == "bicgstab", "unknown solver " << s);
    std::string pc = p.get_string("preconditioner", p.get_string("precond", "diagonal"));
    std::transform(pc.begin(), pc.end(), pc.begin(), ::tolower);
    solver.precond = (pc == "diagonal" || pc == "jacobi") ? "jacobi" : pc;
    EQ_REQUIRE(solver.precond == "jacobi" || solver.precond == "ilu0" || solver.precond == "none" ||
                   solver.precond == "cpr",
               "unknown preconditioner " << pc);
    solver.max_iter = p.get_int("solver_max_iter", 1000);
    solver.tol = p.get_double("solver_tol", 1e-8);
    solver.precision = p.get_string("precision", "fp64");
    solver.gmres_restart = p.get_int("gmres_restart", 30);
    solver.inner_tol = p.get_double("inner_tol", 1e-4);
    solver.max_refine = p.get_int("max_refine", 30);
    stencil_solver = p.get_bool("stencil_solver", true);
    fuse_residual = p.get_bool("fuse_residual", true);
    forcing = p.get_double("forcing", forcing);
    eta_max = p.get_double("eta_max", eta_max);
    forcing0 = p.get_double("forcing0", forcing0);
    EQ_REQUIRE(forcing >= 0.0 && eta_max > 0.0 && eta_max < 1.0, "forcing must be >= 0 and eta_max in (0, 1)");
    verbose = p.get_int("verbose", 0);
    solver.verbose = verbose;
    max_iter = p.get_int("max_iter", 10);
    abs_res_tol = p.get_double("abs_res_tol", 1e-6);
    rel_res_tol = p.get_double("rel_res_tol", 0.0);
    output_to_file = p.get_bool("output_to_file", false);
}

IP RuntimeCore::face_cell(const IP& faces, int which) const {
    const auto key = std::make_pair(faces->id, which);
    auto it = face_cells_.find(key);
    if (it != face_cells_.end() && !it->second.first.expired()) return it->second.second;
    if (face_cells_.size() > 64)   // entries whose face set died can never hit again
        for (auto e = face_cells_.begin(); e != face_cells_.end();)
            e = e->second.first.expired() ? face_cells_.erase(e) : std::next(e);
    auto d = dev_alloc<int>(std::max(1, faces->size()));
    launch_gather_face_cells(grid_->face_cells->p, faces->dev->p, faces->size(), which, d->p);
    IP out = i_from_device(0, d, faces->size());
    face_cells_[key] = {faces, out};
    return out;
}
IP RuntimeCore::firstCell(const IP& faces) const { return face_cell(faces, 0); }
IP RuntimeCore::secondCell(const IP& faces) const { return face_cell(faces, 1); }
SP RuntimeCore::norm(const IP& s) const {
    auto d = dev_alloc<double>(std::max(1, s->size()));
    launch_gather_vec(s->codim == 0 ? grid_->cell_volumes->p : grid_->face_areas->p, s->dev->p, s->size(), 1, d->p);
    return s_from_device(d, s->size());
}
VP RuntimeCore::centroid(const IP& s) const {
    auto v = std::make_shared<VectorData>();
    v->n = s->size();
    v->dim = grid_->dim;
    v->val = dev_alloc<double>(std::max(1, s->size() * grid_->dim));
    launch_gather_vec(s->codim == 0 ? grid_->cell_centroids->p : grid_->face_centroids->p, s->dev->p, s->size(),
                      grid_->dim, v->val->p);
    return v;
}
VP RuntimeCore::normal(const IP& faces) const {
    // unit normals (EquelleRuntimeCPU.cpp:295-315)
    auto v = std::make_shared<VectorData>();
    v->n = faces->size();
    v->dim = grid_->dim;
    v->val = dev_alloc<double>(std::max(1, faces->size() * grid_->dim));
    launch_unit_normals(grid_->face_normals->p, faces->dev->p, faces->size(), grid_->dim, v->val->p);
    return v;
}

SP RuntimeCore::gradient(const SP& x) const {
    EQ_REQUIRE(x->n == grid_->nc, "Gradient need input defined on AllCells()");
    if (Tape* tp = tape_for(*x))
        EQ_TAPE_RETURN(tp, tp->node(T_MAT, tp->operand(x), -1, -1, tp->matrix(grid_->grad)), grid_->nif, x->ad,
                       gradient(x));
    // values: coalesced face sweep u[second] - u[first]; Jacobian: grad * J (cached pattern)
    auto v = dev_alloc<double>(std::max(1, grid_->nif));
    launch_gradient(x->val->p, grid_->if_c1->p, grid_->if_c2->p, v->p, grid_->nif);
    return s_matmul(grid_->grad, x, v);
}
SP RuntimeCore::gradient_matrix(const SP& x) const {
    EQ_REQUIRE(x->n == grid_->nc, "Gradient need input defined on AllCells()");
    return s_matmul(grid_->grad, x);
}
SP RuntimeCore::neg_gradient(const SP& x) const {
    EQ_REQUIRE(x->n == grid_->nc, "negGradient needs input defined on AllCells()");
    if (Tape* tp = tape_for(*x))
        EQ_TAPE_RETURN(tp, tp->node(T_MAT, tp->operand(x), -1, -1, tp->matrix(grid_->ngrad)), grid_->nif, x->ad,
                       neg_gradient(x));
    // values u[first] - u[second] (the gradient sweep with the face cells swapped);
    // Jacobian ngrad * J, the pattern of grad * J with negated values
    auto v = dev_alloc<double>(std::max(1, grid_->nif));
    launch_gradient(x->val->p, grid_->if_c2->p, grid_->if_c1->p, v->p, grid_->nif);
    return s_matmul(grid_->ngrad, x, v);
}
SP RuntimeCore::interior_divergence(const SP& x) const {
    EQ_REQUIRE(x->n == grid_->nif, "interiorDivergence needs input defined on InteriorFaces()");
    return divergence(x);   // div * x: the interior branch of divergence
}
SP RuntimeCore::divergence(const SP& x) const {
    const DeviceGrid& g = *grid_;
    EQ_REQUIRE(x->n == g.nf || x->n == g.nif, "Input for divergence has to be on AllFaces or on InteriorFaces()");
    if (Tape* tp = tape_for(*x))
        EQ_TAPE_RETURN(tp, tp->node(T_MAT, tp->operand(x), -1, -1, tp->matrix(x->n == g.nf ? g.fulldiv : g.div)), g.nc,
                       x->ad, divergence(x));
    // values: cell sweep over the signed face list (ascending face order per cell)
    const double* flux = x->val->p;
    DevPtr<double> full;
    if (x->n != g.nf) {
        PMP pm = position_map(g.all_faces, g.interior_faces);
        full = dev_alloc<double>(g.nf);
        launch_gather(x->val->p, pm->d_inv->p, full->p, g.nf);
        flux = full->p;
    }
    auto v = dev_alloc<double>(std::max(1, g.nc));
    launch_divergence(flux, g.cell_facepos->p, g.signed_faces->p, v->p, g.nc);
    return s_matmul(x->n == g.nf ? g.fulldiv : g.div, x, v);
}
SP RuntimeCore::divergence_matrix(const SP& x) const {
    const DeviceGrid& g = *grid_;
    EQ_REQUIRE(x->n == g.nf || x->n == g.nif, "Input for divergence has to be on AllFaces or on InteriorFaces()");
    return s_matmul(x->n == g.nf ? g.fulldiv : g.div, x);
}

IP RuntimeCore::range_set(int start, int len) const {
    auto key = std::make_pair(start, len);
    auto it = ranges_.find(key);
    if (it != ranges_.end()) return it->second;
    std::vector<int> idx(len);
    for (int i = 0; i < len; ++i) idx[i] = start + i;
    IP s = make_index(0, false, std::move(idx));
    ranges_[key] = s;
    return s;
}

// The pattern of J couples each cell only with itself (always present) and its Cartesian
// neighbours (W, E, S, N, B, T where they exist): the 5- or 7-point stencil of the grid.
bool RuntimeCore::stencil_pattern(const PatternP& P) const {
    auto it = stencil_patterns_.find(P->id);
    if (it != stencil_patterns_.end()) return it->second;
    const GridHost& g = grid_->host;
    const int nx = g.n[0], ny = g.n[1], nz = g.dim == 3 ? g.n[2] : 1, nxy = nx * ny;
    bool ok = P->rows == g.nc && P->cols == g.nc;
    for (int c = 0; ok && c < g.nc; ++c) {
        const int i = c % nx, j = (c / nx) % ny, k = c / nxy;
        bool diag = false;
        for (int e = P->rp()[c]; ok && e < P->rp()[c + 1]; ++e) {
            const int col = P->ci()[e];
            diag |= col == c;
            ok = col == c || (i > 0 && col == c - 1) || (i < nx - 1 && col == c + 1) || (j > 0 && col == c - nx) ||
                 (j < ny - 1 && col == c + nx) || (nz > 1 && k > 0 && col == c - nxy) ||
                 (nz > 1 && k < nz - 1 && col == c + nxy);
        }
        ok = ok && diag;
    }
    stencil_patterns_[P->id] = ok;
    return ok;
}

bool RuntimeCore::stencil_solves(const PatternP& P) const {
    return stencil_solver && solver.method == "bicgstab" && (solver.precond == "ilu0" || solver.precond == "jacobi") &&
           P && stencil_pattern(P);
}

void RuntimeCore::ensure_stencil() const {
    if (stencil_) return;
    FusedConfig fc;
    fc.precision = solver.precision;
    fc.precond = solver.precond;
    fc.solver_tol = solver.tol;
    fc.solver_max_iter = solver.max_iter;
    fc.inner_tol = solver.inner_tol;
    fc.max_refine = solver.max_refine;
    const GridHost& g = grid_->host;
    stencil_.reset(new HeatBatch(fc));
    stencil_slot_ = stencil_->add_linear(g.n[0], g.n[1], g.dim == 3 ? g.n[2] : 1);
}

bool RuntimeCore::stencil_sink(const PatternP& P, StencilSink* out) const {
    if (!stencil_solves(P)) return false;
    ensure_stencil();
    const HeatBatch::LinearTargets L = stencil_->linear_targets(stencil_slot_);
    out->F = L.F;
    for (int q = 0; q < 7; ++q) {
        out->K[q] = L.K[q];
        out->C64[q] = L.C64[q];
    }
    out->nx = L.nx;
    out->ny = L.ny;
    out->nz = L.nz;
    out->h = L.h;
    out->ncoef = L.ncoef;
    out->fp32 = L.fp32;
    out->token = ++sink_token_;
    return true;
}

double RuntimeCore::linear_tolerance(double rn, double r0) const {
    const double target = std::max(abs_res_tol, rel_res_tol > 0.0 ? rel_res_tol * r0 : 0.0);
    if (forcing <= 0.0 || target <= 0.0 || !(rn > 0.0) || !std::isfinite(rn)) return solver.tol;
    return std::min(eta_max, std::max(solver.tol, forcing * target / rn));
}

double RuntimeCore::first_tolerance(double rho) const {
    if (forcing <= 0.0 || forcing0 <= 0.0 || !(rho > 0.0) || !std::isfinite(rho)) return solver.tol;
    return std::min(eta_max, std::max(solver.tol, forcing0 * rho));
}

SP RuntimeCore::solve_for_update(const SP& residual, LinReport* rep, double tol) const {
    const double eta = tol > 0.0 ? tol : solver.tol;
    tape_read(*residual, "a linear solve");
    // BiCGStab with red-black ILU(0) or Jacobi on a Jacobian with the grid's stencil
    // pattern: the coefficients are gathered out of the CSR values (or were written in place
    // by a fused residual evaluation) and the solve runs on the stencil kernels (no column
    // indices in the iteration); other systems take the CSR / SELL-64 path.
    if (residual->jsink || stencil_solves(residual->pat)) {
        ensure_stencil();
        auto x = dev_alloc<double>(std::max(1, residual->n));
        LinearSolveStats st;
        if (residual->jsink) {
            EQ_REQUIRE(residual->jsink == sink_token_,
                       "the stencil solver's coefficients were overwritten after this residual was evaluated");
            st = stencil_->solve_linear_prepared(stencil_slot_, x->p, eta);
        } else {
            const Pattern& P = *residual->pat;
            st = stencil_->solve_linear(stencil_slot_, P.d_rp->p, P.d_ci->p, jac_values(*residual), residual->val->p,
                                        x->p, eta);
        }
        LinReport R;
        R.iterations = st.iterations;
        R.refinements = st.refinements;
        R.residual = st.relres;
        R.converged = st.relres <= eta;
        if (solver.verbose > 0)
            fprintf(stdout, "Linear solver (stencil %s/%s/%s): %d iterations, relative residual %g%s\n",
                    solver.method.c_str(), solver.precond.c_str(), solver.precision.c_str(), R.iterations,
                    R.residual, R.converged ? "" : " (not converged)");
        if (rep) *rep = R;
        return s_from_device(x, residual->n);
    }
    SolverConfig sc = solver;
    sc.tol = eta;
    if (solver.precond == "cpr") {
        // stacked [p; s] (newtonSolveSystem): stage 1 is one multigrid V-cycle on the
        // pressure block, the grid's stencil in the first nc unknowns
        const int np = grid_->nc;
        EQ_REQUIRE(residual->n == 2 * np, "preconditioner=cpr needs two stacked unknowns on AllCells() ("
                                              << residual->n << " unknowns for " << np << " cells)");
        if (!mg_) {
            FusedConfig fc;
            fc.precision = "mixed";   // the V-cycle in FP32: a preconditioner
            fc.cpt = 4;               // fixed launch shape (no tuning trials on this instance)
            fc.mg_scale = params_.get_double("mg_scale", fc.mg_scale);
            const GridHost& g = grid_->host;
            mg_.reset(new HeatBatch(fc));
            mg_slot_ = mg_->add_linear_multigrid(g.n[0], g.n[1], g.dim == 3 ? g.n[2] : 1);
        }
        CprStage st;
        st.np = np;
        st.setup = [this, np](const PatternP& P, const double* vals) {
            const PressureBlock& b = pressure_block(P, np);
            auto sv = dev_alloc<double>(std::max(1, b.pat->nnz));
            launch_gather(vals, b.map->p, sv->p, b.pat->nnz);
            mg_->set_linear_matrix(mg_slot_, b.pat->d_rp->p, b.pat->d_ci->p, sv->p);
        };
        st.apply = [this](const double* r, double* z) { mg_->vcycle(mg_slot_, r, z); };
        return linear_solve(sc, residual, rep, &st);
    }
    return linear_solve(sc, residual, rep);
}

// Rows and columns of the first np unknowns of P (host pass, once per pattern): the
// sub-pattern and each of its entries' position in P.
const RuntimeCore::PressureBlock& RuntimeCore::pressure_block(const PatternP& P, int np) const {
    auto it = pblocks_.find(P->id);
    if (it != pblocks_.end()) return it->second;
    const std::vector<int>&rp = P->rp(), &ci = P->ci();
    std::vector<int> srp(np + 1, 0), sci, map;
    for (int r = 0; r < np; ++r) {
        for (int k = rp[r]; k < rp[r + 1]; ++k)
            if (ci[k] < np) {
                sci.push_back(ci[k]);
                map.push_back(k);
            }
        srp[r + 1] = (int)sci.size();
    }
    PressureBlock b;
    b.pat = make_pattern(np, np, std::move(srp), std::move(sci));
    b.map = dev_from_host(map.data(), map.size());
    EQ_REQUIRE(stencil_pattern(b.pat), "preconditioner=cpr: the pressure block is not the grid's stencil");
    b.owner = P;
    return pblocks_[P->id] = b;
}

}  // namespace eqhip
