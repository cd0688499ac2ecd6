// Opm::GridManager for Cartesian grids, as used by the Equelle runtimes
// (EquelleRuntimeCUDA.cpp:29-60).  Builds the host UnstructuredGrid with the
// numbering of OPM's Cartesian grids: cells x-fastest; x-faces, then y-faces,
// then z-faces; face_cells = (-axis cell, +axis cell).
#ifndef EQUELLE_AMD_OPM_GRIDMANAGER_HPP
#define EQUELLE_AMD_OPM_GRIDMANAGER_HPP
#include <memory>
#include <string>
#include <vector>

#include <opm/grid/UnstructuredGrid.h>

namespace Opm {

class GridManager {
public:
    GridManager(int nx, int ny, double dx = 1.0, double dy = 1.0) { build(2, nx, ny, 1, dx, dy, 1.0); }
    GridManager(int nx, int ny, int nz, double dx = 1.0, double dy = 1.0, double dz = 1.0) {
        build(3, nx, ny, nz, dx, dy, dz);
    }
    const UnstructuredGrid* c_grid() const { return &g_; }

private:
    void build(int dim, int nx, int ny, int nz, double dx, double dy, double dz);
    UnstructuredGrid g_{};
    std::vector<int> face_cells_, facepos_, cell_faces_;
    std::vector<double> cc_, cv_, fc_, fa_, fn_;
};

inline void GridManager::build(int dim, int nx, int ny, int nz, double dx, double dy, double dz) {
    const long nc = (long)nx * ny * nz;
    const long fx = (long)(nx + 1) * ny * nz, fy = (long)nx * (ny + 1) * nz;
    const long fz = dim == 3 ? (long)nx * ny * (nz + 1) : 0;
    const long nf = fx + fy + fz;
    face_cells_.assign(2 * nf, -1);
    fa_.resize(nf);
    fc_.assign(dim * nf, 0.0);
    fn_.assign(dim * nf, 0.0);
    const double area[3] = {dim == 3 ? dy * dz : dy, dim == 3 ? dx * dz : dx, dx * dy};
    long base = 0;
    for (int a = 0; a < dim; ++a) {
        const long ex = nx + (a == 0), ey = ny + (a == 1), ez = nz + (a == 2);
        for (long k = 0; k < ez; ++k)
            for (long j = 0; j < ey; ++j)
                for (long i = 0; i < ex; ++i) {
                    const long f = base + i + ex * (j + ey * k);
                    long p[3] = {i, j, k};
                    const long ext[3] = {nx, ny, nz};
                    const long hi = p[a];
                    p[a] = hi - 1;
                    face_cells_[2 * f] = p[a] >= 0 ? (int)(p[0] + nx * (p[1] + (long)ny * p[2])) : -1;
                    p[a] = hi;
                    face_cells_[2 * f + 1] = hi < ext[a] ? (int)(p[0] + nx * (p[1] + (long)ny * p[2])) : -1;
                    fa_[f] = area[a];
                    fn_[dim * f + a] = area[a];
                    const double c[3] = {(i + (a == 0 ? 0.0 : 0.5)) * dx, (j + (a == 1 ? 0.0 : 0.5)) * dy,
                                         (k + (a == 2 ? 0.0 : 0.5)) * dz};
                    for (int d = 0; d < dim; ++d) fc_[dim * f + d] = c[d];
                }
        base += ex * ey * ez;
    }
    cv_.assign(nc, dim == 3 ? dx * dy * dz : dx * dy);
    cc_.resize(dim * nc);
    facepos_.resize(nc + 1);
    cell_faces_.resize(2 * dim * nc);
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                const long c = i + nx * (j + (long)ny * k);
                cc_[dim * c] = (i + 0.5) * dx;
                cc_[dim * c + 1] = (j + 0.5) * dy;
                if (dim == 3) cc_[dim * c + 2] = (k + 0.5) * dz;
                const long q = 2 * dim * c;
                facepos_[c] = (int)q;
                cell_faces_[q] = (int)(i + (nx + 1) * (j + (long)ny * k));
                cell_faces_[q + 1] = (int)(i + 1 + (nx + 1) * (j + (long)ny * k));
                cell_faces_[q + 2] = (int)(fx + i + (long)nx * (j + (long)(ny + 1) * k));
                cell_faces_[q + 3] = (int)(fx + i + (long)nx * (j + 1 + (long)(ny + 1) * k));
                if (dim == 3) {
                    cell_faces_[q + 4] = (int)(fx + fy + i + (long)nx * (j + (long)ny * k));
                    cell_faces_[q + 5] = (int)(fx + fy + i + (long)nx * (j + (long)ny * (k + 1)));
                }
            }
    facepos_[nc] = (int)(2 * dim * nc);
    g_.dimensions = dim;
    g_.number_of_cells = (int)nc;
    g_.number_of_faces = (int)nf;
    g_.number_of_nodes = (int)((nx + 1L) * (ny + 1L) * (dim == 3 ? nz + 1L : 1L));
    g_.face_cells = face_cells_.data();
    g_.cell_facepos = facepos_.data();
    g_.cell_faces = cell_faces_.data();
    g_.cell_centroids = cc_.data();
    g_.cell_volumes = cv_.data();
    g_.face_centroids = fc_.data();
    g_.face_areas = fa_.data();
    g_.face_normals = fn_.data();
    g_.cartdims[0] = nx;
    g_.cartdims[1] = ny;
    g_.cartdims[2] = dim == 3 ? nz : 1;
}

}  // namespace Opm
#endif
