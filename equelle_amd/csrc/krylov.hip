// Device Krylov solvers, SELL-64 SpMV, Jacobi and multicolour ILU(0); see krylov.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <unordered_map>

#include "engine.hpp"
#include "kernels.hpp"
#include "krylov.hpp"
#include "timing.hpp"

namespace eqhip {

namespace {
constexpr int BS = 256;
constexpr int W = 64;   // SELL slice height = wavefront size
constexpr int GM_MAX = 200;   // largest GMRES restart length
inline hipStream_t S() { return Device::get().stream(); }
__device__ inline long gid() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }
inline int vec_blocks(long n) {
    long b = (n + BS * 4 - 1) / (BS * 4);
    return (int)std::max(1L, std::min(b, 2048L));
}
#define KL(kernel, grid, ...)                                                  \
    do {                                                                       \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(BS), 0, S(), __VA_ARGS__); \
        EQ_HIP(hipGetLastError());                                             \
    } while (0)

// ---------------------------------------------------------------------------
// Solver structure cached per sparsity pattern
struct Layout {
    int n = 0, nslices = 0;
    long nnzp = 0;
    DevPtr<int> slice_ptr, col, src, diag;
    // multicolour ILU(0)
    int ncolors = 0;
    std::vector<int> color_off;        // host: rows of colour c are color_rows[color_off[c]..color_off[c+1])
    DevPtr<int> color_rows;
    // Triangular sweeps in SELL-64 per colour: the rows of colour c (colour order) form
    // slices of 64 that never cross a colour; slice s of the L part holds its rows' L
    // entries column-major from lsp[s] (row order within a row kept, so every sweep sums
    // in CSR order), padded with column -1; an L column of colour 0 is stored as -2 - col
    // (its z equals the right-hand side, so the sweeps read that).  Colour c's first slice
    // is cslice[c].
    std::vector<int> cslice;
    DevPtr<int> lsp, lcol, lsrc, usp, ucol, usrc;
    long lnnz = 0, unnz = 0;
    // ILU(0) factorisation as precomputed elimination steps: the row at colour position t
    // runs steps fsp[t]..fsp[t+1]; step s divides L entry fq[s] by the pivot at fdk[s]
    // (-1: none) and, unless the multiplier is zero, applies the updates fpp[s]..fpp[s+1]:
    // lu[fg[p]] -= l * lu[ff[p]].  Steps follow the colours of the eliminated rows, then
    // CSR order; updates the U entries of the pivot row in CSR order.
    DevPtr<int> fsp, fq, fdk, fpp, fg, ff;
    std::weak_ptr<const Pattern> pat;
};
using LayoutP = std::shared_ptr<const Layout>;

LayoutP get_layout(const PatternP& p) {
    static std::mutex mu;
    static std::unordered_map<uint64_t, LayoutP> cache;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(p->id);
        if (it != cache.end()) return it->second;
        for (auto i = cache.begin(); i != cache.end();) i = i->second->pat.expired() ? cache.erase(i) : std::next(i);
    }
    EQ_REQUIRE(p->rows == p->cols, "linear solve needs a square Jacobian (" << p->rows << "x" << p->cols << ")");
    auto L = std::make_shared<Layout>();
    const int n = p->rows;
    L->n = n;
    L->pat = p;
    L->nslices = (n + W - 1) / W;
    std::vector<int> sp(L->nslices + 1, 0);
    for (int s = 0; s < L->nslices; ++s) {
        int w = 0;
        for (int r = s * W; r < std::min(n, (s + 1) * W); ++r) w = std::max(w, p->rp()[r + 1] - p->rp()[r]);
        sp[s + 1] = sp[s] + w * W;
    }
    L->nnzp = sp[L->nslices];
    std::vector<int> col(L->nnzp), src(L->nnzp);
    for (int s = 0; s < L->nslices; ++s) {
        const int w = (sp[s + 1] - sp[s]) / W;
        for (int j = 0; j < w; ++j)
            for (int lane = 0; lane < W; ++lane) {
                const int r = s * W + lane;
                const long k = sp[s] + (long)j * W + lane;
                if (r < n && j < p->rp()[r + 1] - p->rp()[r]) {
                    col[k] = p->ci()[p->rp()[r] + j];
                    src[k] = p->rp()[r] + j;
                } else {
                    col[k] = r < n ? r : 0;
                    src[k] = -1;
                }
            }
    }
    std::vector<int> diag(n, -1);
    for (int r = 0; r < n; ++r)
        for (int k = p->rp()[r]; k < p->rp()[r + 1]; ++k)
            if (p->ci()[k] == r) diag[r] = k;
    // greedy colouring of the symmetrised graph in natural order
    std::vector<std::vector<int>> adjT(n);
    for (int r = 0; r < n; ++r)
        for (int k = p->rp()[r]; k < p->rp()[r + 1]; ++k)
            if (p->ci()[k] != r) adjT[p->ci()[k]].push_back(r);
    std::vector<int> color(n, -1);
    std::vector<int> mark(64, -1);
    int ncol = 0;
    for (int r = 0; r < n; ++r) {
        auto use = [&](int j) {
            if (color[j] >= 0) {
                if (color[j] >= (int)mark.size()) mark.resize(color[j] + 1, -1);
                mark[color[j]] = r;
            }
        };
        for (int k = p->rp()[r]; k < p->rp()[r + 1]; ++k)
            if (p->ci()[k] != r) use(p->ci()[k]);
        for (int j : adjT[r]) use(j);
        int c = 0;
        while (c < (int)mark.size() && mark[c] == r) ++c;
        color[r] = c;
        ncol = std::max(ncol, c + 1);
    }
    EQ_REQUIRE(ncol <= 255, "ILU(0) colouring needs more than 255 colours");
    L->ncolors = ncol;
    L->color_off.assign(ncol + 1, 0);
    for (int r = 0; r < n; ++r) L->color_off[color[r] + 1]++;
    for (int c = 0; c < ncol; ++c) L->color_off[c + 1] += L->color_off[c];
    std::vector<int> rows(n), fill(L->color_off.begin(), L->color_off.end() - 1);
    for (int r = 0; r < n; ++r) rows[fill[color[r]]++] = r;
    // per CSR entry: 0 L, 1 diagonal, 2 U, 3 same colour (dropped by ILU(0) here)
    std::vector<uint8_t> kind(p->nnz);
    for (int r = 0; r < n; ++r) {
        for (int k = p->rp()[r]; k < p->rp()[r + 1]; ++k) {
            const int c = p->ci()[k];
            kind[k] = c == r ? 1 : color[c] < color[r] ? 0 : color[c] > color[r] ? 2 : 3;
        }
    }
    // per-colour SELL-64 of the L (kind 0) and U (kind 2) entries
    {
        std::vector<int> slc(1, 0), lsp(1, 0), usp(1, 0), lcol, lsrc, ucol, usrc;
        L->cslice.assign(ncol + 1, 0);
        for (int c = 0; c < ncol; ++c) {
            const int b = L->color_off[c], e = L->color_off[c + 1];
            for (int s0 = b; s0 < e; s0 += W) {
                const int s1 = std::min(e, s0 + W);
                for (int part = 0; part < 2; ++part) {
                    const uint8_t want = part == 0 ? 0 : 2;
                    std::vector<int>& cols = part == 0 ? lcol : ucol;
                    std::vector<int>& srcs = part == 0 ? lsrc : usrc;
                    std::vector<int>& ptr = part == 0 ? lsp : usp;
                    int w = 0;
                    for (int t = s0; t < s1; ++t) {
                        int m = 0;
                        for (int k = p->rp()[rows[t]]; k < p->rp()[rows[t] + 1]; ++k) m += kind[k] == want;
                        w = std::max(w, m);
                    }
                    const size_t base = cols.size();
                    cols.resize(base + (size_t)w * W, -1);
                    srcs.resize(base + (size_t)w * W, -1);
                    for (int t = s0; t < s1; ++t) {
                        int j = 0;
                        for (int k = p->rp()[rows[t]]; k < p->rp()[rows[t] + 1]; ++k)
                            if (kind[k] == want) {
                                const int cc = p->ci()[k];
                                // an L column of colour 0 reads the right-hand side (-2 - col)
                                cols[base + (size_t)j * W + (t - s0)] = part == 0 && color[cc] == 0 ? -2 - cc : cc;
                                srcs[base + (size_t)j * W + (t - s0)] = k;
                                ++j;
                            }
                    }
                    ptr.push_back((int)cols.size());
                }
            }
            L->cslice[c + 1] = (int)lsp.size() - 1;
        }
        EQ_REQUIRE(lcol.size() < (size_t)INT32_MAX && ucol.size() < (size_t)INT32_MAX,
                   "ILU(0) sweep arrays exceed 2^31 entries");
        L->lnnz = (long)lcol.size();
        L->unnz = (long)ucol.size();
        for (auto* v : {&lcol, &lsrc, &ucol, &usrc})
            if (v->empty()) v->push_back(-1);   // keep every device array non-empty
        L->lsp = dev_from_host(lsp.data(), lsp.size());
        L->usp = dev_from_host(usp.data(), usp.size());
        L->lcol = dev_from_host(lcol.data(), lcol.size());
        L->lsrc = dev_from_host(lsrc.data(), lsrc.size());
        L->ucol = dev_from_host(ucol.data(), ucol.size());
        L->usrc = dev_from_host(usrc.data(), usrc.size());
    }
    // elimination steps of the factorisation (rows of colour 0 have none)
    {
        std::vector<int> fsp(1, 0), fq, fdk, fpp(1, 0), fg, ff;
        for (int t = 0; t < n; ++t) {
            const int i = rows[t];
            const int b = p->rp()[i], e = p->rp()[i + 1];
            for (int cc = 0; cc < color[i]; ++cc)
                for (int q = b; q < e; ++q) {
                    if (kind[q] != 0) continue;
                    const int k = p->ci()[q];
                    if (color[k] != cc) continue;
                    fq.push_back(q);
                    fdk.push_back(diag[k]);
                    for (int f = p->rp()[k]; f < p->rp()[k + 1]; ++f) {
                        if (kind[f] != 2) continue;
                        const int j = p->ci()[f];
                        for (int g = b; g < e; ++g)
                            if (p->ci()[g] == j) {
                                fg.push_back(g);
                                ff.push_back(f);
                                break;
                            }
                    }
                    EQ_REQUIRE(fg.size() < (size_t)INT32_MAX, "ILU(0) update list exceeds 2^31 entries");
                    fpp.push_back((int)fg.size());
                }
            fsp.push_back((int)fq.size());
        }
        for (auto* v : {&fq, &fdk, &fg, &ff})
            if (v->empty()) v->push_back(-1);   // keep every device array non-empty
        L->fsp = dev_from_host(fsp.data(), fsp.size());
        L->fq = dev_from_host(fq.data(), fq.size());
        L->fdk = dev_from_host(fdk.data(), fdk.size());
        L->fpp = dev_from_host(fpp.data(), fpp.size());
        L->fg = dev_from_host(fg.data(), fg.size());
        L->ff = dev_from_host(ff.data(), ff.size());
    }
    L->slice_ptr = dev_from_host(sp.data(), sp.size());
    L->col = dev_from_host(col.data(), col.size());
    L->src = dev_from_host(src.data(), src.size());
    L->diag = dev_from_host(diag.data(), diag.size());
    L->color_rows = dev_from_host(rows.data(), rows.size());
    std::lock_guard<std::mutex> lk(mu);
    cache[p->id] = L;
    return L;
}

// ---------------------------------------------------------------------------
// Kernels
template <class T>
__global__ void k_sell_vals(long nnzp, const int* __restrict__ src, const double* __restrict__ csr,
                            T* __restrict__ out) {
    long k = gid();
    if (k >= nnzp) return;
    const int s = src[k];
    out[k] = s >= 0 ? (T)csr[s] : (T)0;
}

// One wave64 per slice: lane = row within the slice, entries column-major.
template <class T>
__global__ __launch_bounds__(BS) void k_sell_spmv(int n, int nslices, const int* __restrict__ sp,
                                                  const int* __restrict__ col, const T* __restrict__ val,
                                                  const T* __restrict__ x, T* __restrict__ y,
                                                  const double* __restrict__ gate) {
    if (gate && gate[0] == 0.0) return;   // solve already finished (device flag)
    const long row = gid();
    const long slice = row / W;
    if (slice >= nslices) return;
    const int lane = (int)(row % W);
    const int off = sp[slice];
    const int w = (sp[slice + 1] - off) / W;
    T acc = 0;
    const int* c = col + off + lane;
    const T* v = val + off + lane;
    for (int j = 0; j < w; ++j) acc += v[j * W] * x[c[j * W]];
    if (row < n) y[row] = acc;
}

template <class Tin, class Tout>
__global__ void k_cast(long n, const Tin* __restrict__ in, double scale, Tout* __restrict__ out) {
    long i = gid();
    if (i < n) out[i] = (Tout)(in[i] * scale);
}

// x64 += scale * d
template <class T>
__global__ void k_axpy_mixed(long n, double scale, const T* __restrict__ d, double* __restrict__ x) {
    long i = gid();
    if (i < n) x[i] += scale * (double)d[i];
}

// r = b - y
__global__ void k_residual(long n, const double* __restrict__ b, const double* __restrict__ y,
                           double* __restrict__ r) {
    long i = gid();
    if (i < n) r[i] = b[i] - y[i];
}

template <class T>
__global__ void k_jacobi_setup(int n, const int* __restrict__ diag, const double* __restrict__ csr,
                               T* __restrict__ dinv) {
    long i = gid();
    if (i >= n) return;
    const int k = diag[i];
    const double d = k >= 0 ? csr[k] : 0.0;
    dinv[i] = (T)(d != 0.0 ? 1.0 / d : 1.0);
}

template <class T>
__global__ void k_scale(long n, const T* __restrict__ d, const T* __restrict__ a, T* __restrict__ out,
                        const double* __restrict__ gate) {
    if (gate && gate[0] == 0.0) return;
    long i = gid();
    if (i < n) out[i] = d ? d[i] * a[i] : a[i];
}

// ---- device scalars -------------------------------------------------------
// ACTIVE (1/0) gates every kernel of an iteration, so the host can queue iterations
// ahead and read the convergence state only every few iterations; IT counts the
// iterations that ran while active; BB = ||b||^2.
enum Slot { RHO = 0, RHO_OLD, ALPHA, OMEGA, BETA, RR, BREAK, RZ, ACTIVE, IT, BB, HN, NSLOT = 16 };
enum Act { ACT_STORE = 0, ACT_BICG_ALPHA, ACT_BICG_OMEGA, ACT_BICG_RHO, ACT_CG_ALPHA, ACT_CG_BETA, ACT_INIT };

__device__ inline double bred(double v) {
    __shared__ double lds[BS / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < BS / 64; ++i) r += lds[i];
    __syncthreads();
    return r;   // valid in thread 0
}

// up to 3 dot products, FP64 accumulation, partials[d*nb + block]
template <class T>
__global__ void k_dots(long n, int nd, const T* __restrict__ x0, const T* __restrict__ y0, const T* __restrict__ x1,
                       const T* __restrict__ y1, const T* __restrict__ x2, const T* __restrict__ y2,
                       double* __restrict__ part, const double* __restrict__ gate) {
    if (gate && gate[0] == 0.0) return;
    double a0 = 0, a1 = 0, a2 = 0;
    for (long i = gid(); i < n; i += (long)gridDim.x * BS) {
        a0 += (double)x0[i] * (double)y0[i];
        if (nd > 1) a1 += (double)x1[i] * (double)y1[i];
        if (nd > 2) a2 += (double)x2[i] * (double)y2[i];
    }
    a0 = bred(a0);
    if (threadIdx.x == 0) part[blockIdx.x] = a0;
    if (nd > 1) {
        a1 = bred(a1);
        if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = a1;
    }
    if (nd > 2) {
        a2 = bred(a2);
        if (threadIdx.x == 0) part[2 * gridDim.x + blockIdx.x] = a2;
    }
}

constexpr double kHuge = 1e30;   // largest Krylov step length accepted

// tol2, maxit: the ACT_BICG_RHO / ACT_CG_BETA / ACT_INIT actions update ACTIVE
// (continue while ||r||^2 > tol2 ||b||^2, no breakdown, IT < maxit).
__global__ void k_finalize(int act, int nd, const double* __restrict__ part, int nb, double* __restrict__ sc,
                           int slot, int gated, double tol2, int maxit) {
    if (gated && sc[ACTIVE] == 0.0) return;
    double d[3] = {0, 0, 0};
    for (int j = 0; j < nd; ++j) {
        double a = 0;
        for (int i = threadIdx.x; i < nb; i += BS) a += part[j * nb + i];
        a = bred(a);
        if (threadIdx.x == 0) d[j] = a;
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    switch (act) {
        case ACT_STORE:
            for (int j = 0; j < nd; ++j) sc[slot + j] = d[j];
            break;
        // A breakdown (zero, tiny or non-finite denominator) must not reach x: a bad alpha
        // stops the solve before the update (x keeps the last iterate), a bad omega
        // falls back to omega = 0 (x += alpha ph, r = s) and ends the solve after it.
        // kHuge bounds the step so the FP32 vectors cannot overflow either.
        case ACT_BICG_ALPHA: {
            const double a = sc[RHO] / d[0];
            if (d[0] == 0.0 || !isfinite(a) || fabs(a) > kHuge) {
                sc[BREAK] = 1;
                sc[ALPHA] = 0;
                sc[ACTIVE] = 0.0;
            } else sc[ALPHA] = a;
            break;
        }
        case ACT_BICG_OMEGA: {
            const double w = d[0] / d[1];
            if (d[1] == 0.0 || !isfinite(w) || fabs(w) > kHuge) { sc[OMEGA] = 0; sc[BREAK] = 2; }
            else sc[OMEGA] = w;
            break;
        }
        case ACT_BICG_RHO:
            sc[RHO_OLD] = sc[RHO];
            sc[RHO] = d[0];
            sc[RR] = d[1];
            sc[IT] += 1.0;
            if (d[0] == 0.0) sc[BREAK] = 3;
            sc[ACTIVE] = (d[1] > tol2 * sc[BB] && sc[BREAK] == 0.0 && sc[IT] < maxit && isfinite(d[1])) ? 1.0 : 0.0;
            break;
        case ACT_INIT:   // d[0] = rh.r (or r.z), d[1] = r.r at the start of a solve
            sc[RHO] = d[0];
            sc[RZ] = d[0];
            sc[RR] = d[1];
            sc[BB] = d[1];
            sc[IT] = 0.0;
            sc[BREAK] = 0.0;
            sc[ACTIVE] = (d[1] > 0.0 && isfinite(d[1]) && maxit > 0) ? 1.0 : 0.0;
            break;
        case ACT_CG_ALPHA: {
            const double a = sc[RZ] / d[0];
            if (d[0] == 0.0 || !isfinite(a) || fabs(a) > kHuge) {
                sc[ALPHA] = 0;
                sc[BREAK] = 4;
                sc[ACTIVE] = 0.0;
            } else sc[ALPHA] = a;
            break;
        }
        case ACT_CG_BETA:
            sc[BETA] = sc[RZ] != 0.0 ? d[0] / sc[RZ] : 0.0;
            sc[RZ] = d[0];
            sc[RR] = d[1];
            sc[IT] += 1.0;
            sc[ACTIVE] = (d[1] > tol2 * sc[BB] && sc[BREAK] == 0.0 && sc[IT] < maxit && isfinite(d[1])) ? 1.0 : 0.0;
            break;
    }
}

// ---- BiCGStab vector kernels ------------------------------------------------
// first: p = r ; else beta = (rho/rho_old)(alpha/omega), p = r + beta (p - omega v); ph = D p
template <class T>
__global__ void k_bicg_p(long n, int first, const T* __restrict__ r, const T* __restrict__ v, T* __restrict__ p,
                         T* __restrict__ ph, const T* __restrict__ dinv, const double* __restrict__ sc) {
    if (sc[ACTIVE] == 0.0) return;
    long i = gid();
    if (i >= n) return;
    T pi;
    if (first) pi = r[i];
    else {
        const double beta = (sc[RHO] / sc[RHO_OLD]) * (sc[ALPHA] / sc[OMEGA]);
        pi = r[i] + (T)beta * (p[i] - (T)sc[OMEGA] * v[i]);
    }
    p[i] = pi;
    if (ph) ph[i] = dinv ? dinv[i] * pi : pi;
}
// s = r - alpha v ; sh = D s
template <class T>
__global__ void k_bicg_s(long n, const T* __restrict__ r, const T* __restrict__ v, T* __restrict__ s,
                         T* __restrict__ sh, const T* __restrict__ dinv, const double* __restrict__ sc) {
    if (sc[ACTIVE] == 0.0) return;
    long i = gid();
    if (i >= n) return;
    const T si = r[i] - (T)sc[ALPHA] * v[i];
    s[i] = si;
    if (sh) sh[i] = dinv ? dinv[i] * si : si;
}
// x += alpha ph + omega sh ; r = s - omega t ; partial dots (rh.r, r.r)
template <class T>
__global__ void k_bicg_x(long n, T* __restrict__ x, const T* __restrict__ ph, const T* __restrict__ sh,
                         T* __restrict__ r, const T* __restrict__ s, const T* __restrict__ t,
                         const T* __restrict__ rh, const double* __restrict__ sc, double* __restrict__ part) {
    if (sc[ACTIVE] == 0.0) return;
    const T alpha = (T)sc[ALPHA], omega = (T)sc[OMEGA];
    double a0 = 0, a1 = 0;
    for (long i = gid(); i < n; i += (long)gridDim.x * BS) {
        x[i] += alpha * ph[i] + omega * sh[i];
        const T ri = s[i] - omega * t[i];
        r[i] = ri;
        a0 += (double)rh[i] * (double)ri;
        a1 += (double)ri * (double)ri;
    }
    a0 = bred(a0);
    if (threadIdx.x == 0) part[blockIdx.x] = a0;
    a1 = bred(a1);
    if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = a1;
}

// ---- CG vector kernels --------------------------------------------------------
template <class T>
__global__ void k_cg_update(long n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
                            const T* __restrict__ q, const double* __restrict__ sc) {
    if (sc[ACTIVE] == 0.0) return;
    long i = gid();
    if (i >= n) return;
    const T a = (T)sc[ALPHA];
    x[i] += a * p[i];
    r[i] -= a * q[i];
}
template <class T>
__global__ void k_cg_p(long n, int first, const T* __restrict__ z, T* __restrict__ p,
                       const double* __restrict__ sc) {
    if (sc[ACTIVE] == 0.0) return;
    long i = gid();
    if (i >= n) return;
    p[i] = first ? z[i] : z[i] + (T)sc[BETA] * p[i];
}

// ---- GMRES kernels ----------------------------------------------------------------
// The Hessenberg column, Givens rotations and the least-squares solve live in device
// memory (struct Gm), so an Arnoldi step needs no host round trip; every kernel is
// gated by gm->active and the host reads the state only at chunk boundaries.
struct Gm {
    double active;     // 1 while the cycle continues
    double rn;         // |g[j+1]|, the residual norm estimate
    double bn;         // ||b||
    double tol;
    double k;          // columns in the current cycle
    double total;      // Arnoldi steps over all cycles
    double hn;         // ||w|| after orthogonalisation
    double pad;
    double hcol[GM_MAX + 2], h2[GM_MAX + 2];
    double cs[GM_MAX], sn[GM_MAX], g[GM_MAX + 1], y[GM_MAX];
    double H[(GM_MAX + 1) * GM_MAX];
};
// partial dots of w against columns 0..k of V (V column-major, ld = n)
template <class T>
__global__ void k_vdots(long n, int k, const T* __restrict__ V, const T* __restrict__ w, double* __restrict__ part,
                        int nb, const Gm* __restrict__ gm) {
    if (gm->active == 0.0) return;
    const int col = blockIdx.y;
    double a = 0;
    const T* v = V + (long)col * n;
    for (long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)nb * BS) a += (double)v[i] * (double)w[i];
    a = bred(a);
    if (threadIdx.x == 0) part[(long)col * nb + blockIdx.x] = a;
}
// h2[col] = sum of the partials; pass 0 sets hcol, pass 1 adds (CGS2)
__global__ void k_vdots_final(int nb, const double* __restrict__ part, Gm* __restrict__ gm, int pass) {
    if (gm->active == 0.0) return;
    const int col = blockIdx.x;
    double a = 0;
    for (int i = threadIdx.x; i < nb; i += BS) a += part[(long)col * nb + i];
    a = bred(a);
    if (threadIdx.x == 0) {
        gm->h2[col] = a;
        gm->hcol[col] = pass == 0 ? a : gm->hcol[col] + a;
    }
}
// w -= sum_{j<=k} h2[j] V[:,j]
template <class T>
__global__ void k_vupdate(long n, int k, const T* __restrict__ V, const Gm* __restrict__ gm, T* __restrict__ w) {
    if (gm->active == 0.0) return;
    long i = gid();
    if (i >= n) return;
    double s = (double)w[i];
    for (int j = 0; j <= k; ++j) s -= gm->h2[j] * (double)V[(long)j * n + i];
    w[i] = (T)s;
}
// CGS2 pass 2's update, w -= sum_{c<=k} h2[c] V[:,c] (as k_vupdate), with the partial
// sums of ||w||^2 over a grid-stride loop of gridDim.x blocks
template <class T>
__global__ __launch_bounds__(BS) void k_vupdate_norm(long n, int k, const T* __restrict__ V, const Gm* __restrict__ gm,
                                                     T* __restrict__ w, double* __restrict__ part) {
    if (gm->active == 0.0) return;
    double a = 0.0;
    for (long i = gid(); i < n; i += (long)gridDim.x * BS) {
        double s = (double)w[i];
        for (int j = 0; j <= k; ++j) s -= gm->h2[j] * (double)V[(long)j * n + i];
        const T wt = (T)s;
        w[i] = wt;
        a += (double)wt * (double)wt;
    }
    a = bred(a);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
}
__device__ void gmres_step_body(int j, int m, const double* __restrict__ sc, Gm* __restrict__ gm, int maxit,
                                double* hc, const double* cs, const double* sn);
// hn = ||w|| from the partials of k_vupdate_norm (summed into slot HN), the new Givens
// rotation for column j, the residual estimate
__global__ __launch_bounds__(BS) void k_gmres_step_norm(int j, int m, const double* __restrict__ part, int nb,
                                                        double* __restrict__ sc, Gm* __restrict__ gm, int maxit) {
    if (gm->active == 0.0) return;
    // the column and the previous rotations staged in LDS by the whole block (one round
    // trip instead of a chain of dependent loads in the serial loop below)
    __shared__ double hc[GM_MAX + 2], cs[GM_MAX], sn[GM_MAX];
    for (int i = threadIdx.x; i <= j; i += BS) hc[i] = gm->hcol[i];
    for (int i = threadIdx.x; i < j; i += BS) {
        cs[i] = gm->cs[i];
        sn[i] = gm->sn[i];
    }
    double a = 0.0;
    for (int i = threadIdx.x; i < nb; i += BS) a += part[i];
    a = bred(a);   // its barriers also order the staging above
    if (threadIdx.x != 0) return;
    sc[HN] = a;
    gmres_step_body(j, m, sc, gm, maxit, hc, cs, sn);
}
__device__ void gmres_step_body(int j, int m, const double* __restrict__ sc, Gm* __restrict__ gm, int maxit,
                                double* hc, const double* cs, const double* sn) {
    const double hn = sqrt(fmax(sc[HN], 0.0));
    gm->hn = hn;
    hc[j + 1] = hn;
    for (int i = 0; i < j; ++i) {   // previous rotations
        const double t0 = cs[i] * hc[i] + sn[i] * hc[i + 1];
        hc[i + 1] = -sn[i] * hc[i] + cs[i] * hc[i + 1];
        hc[i] = t0;
    }
    const double den = hypot(hc[j], hc[j + 1]);
    gm->cs[j] = den > 0 ? hc[j] / den : 1.0;
    gm->sn[j] = den > 0 ? hc[j + 1] / den : 0.0;
    hc[j] = den;
    hc[j + 1] = 0;
    gm->g[j + 1] = -gm->sn[j] * gm->g[j];
    gm->g[j] = gm->cs[j] * gm->g[j];
    for (int i = 0; i <= j; ++i) gm->H[i * m + j] = hc[i];
    gm->rn = fabs(gm->g[j + 1]);
    gm->k = j + 1;
    gm->total += 1.0;
    if (gm->rn <= gm->tol * gm->bn || hn == 0.0 || !isfinite(gm->rn) || gm->total >= maxit) gm->active = 0.0;
}
// V[:,j+1] = w / hn; with z: also the next step's diagonal preconditioner, z = d V[:,j+1]
// (as k_scale computes it; d null: no preconditioner)
template <class T>
__global__ void k_gmres_next(long n, const T* __restrict__ w, T* __restrict__ vnext, const Gm* __restrict__ gm,
                             const T* __restrict__ d, T* __restrict__ z) {
    if (gm->active == 0.0) return;
    long i = gid();
    if (i >= n) return;
    const T v = (T)((double)w[i] / gm->hn);
    vnext[i] = v;
    if (z) z[i] = d ? d[i] * v : v;
}
// start of a cycle: beta = ||r|| (dot slot RR), V[:,0] = r / beta, g = beta e1
template <class T>
__global__ void k_gmres_start(long n, const T* __restrict__ r, T* __restrict__ v0, const double* __restrict__ sc,
                              Gm* __restrict__ gm, int maxit) {
    const double beta = sqrt(fmax(sc[RR], 0.0));
    const bool go = beta > gm->tol * gm->bn && isfinite(beta) && gm->total < maxit;
    long i = gid();
    if (go && i < n) v0[i] = (T)((double)r[i] / beta);
    if (i == 0) {
        gm->rn = beta;
        gm->active = go ? 1.0 : 0.0;
        gm->k = 0;
        for (int q = 0; q < GM_MAX + 1; ++q) gm->g[q] = 0.0;
        gm->g[0] = beta;
    }
}
// y = H^{-1} g over the k columns of the finished cycle: the triangle staged in LDS by the
// block when it fits (k <= GM_LDS), then back substitution by one thread
constexpr int GM_LDS = 64;
__global__ void k_gmres_solve(int m, Gm* __restrict__ gm) {
    const int k = (int)gm->k;
    __shared__ double H[GM_LDS * GM_LDS], y[GM_LDS], g[GM_LDS];
    const bool lds = k <= GM_LDS;
    if (lds) {
        for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
            const int i = e / k, q = e % k;
            if (q >= i) H[i * GM_LDS + q] = gm->H[i * m + q];
        }
        for (int i = threadIdx.x; i < k; i += blockDim.x) g[i] = gm->g[i];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (lds) {
        for (int i = k - 1; i >= 0; --i) {
            double s = g[i];
            for (int q = i + 1; q < k; ++q) s -= H[i * GM_LDS + q] * y[q];
            y[i] = H[i * GM_LDS + i] != 0 ? s / H[i * GM_LDS + i] : 0.0;
        }
        for (int i = 0; i < k; ++i) gm->y[i] = y[i];
        return;
    }
    for (int i = k - 1; i >= 0; --i) {
        double s = gm->g[i];
        for (int q = i + 1; q < k; ++q) s -= gm->H[i * m + q] * gm->y[q];
        gm->y[i] = gm->H[i * m + i] != 0 ? s / gm->H[i * m + i] : 0.0;
    }
}
// out = sum_{j<k} y[j] V[:,j]
template <class T>
__global__ void k_vcombine(long n, const T* __restrict__ V, const Gm* __restrict__ gm, T* __restrict__ out) {
    long i = gid();
    if (i >= n) return;
    const int k = (int)gm->k;
    double s = 0;
    for (int j = 0; j < k; ++j) s += gm->y[j] * (double)V[(long)j * n + i];
    out[i] = (T)s;
}
template <class T>
__global__ void k_scal(long n, double a, const T* __restrict__ x, T* __restrict__ y) {
    long i = gid();
    if (i < n) y[i] = (T)(a * (double)x[i]);
}
template <class T>
__global__ void k_add(long n, const T* __restrict__ a, T* __restrict__ x) {
    long i = gid();
    if (i < n) x[i] += a[i];
}
// r = b - w
template <class T>
__global__ void k_sub(long n, const T* __restrict__ b, const T* __restrict__ w, T* __restrict__ r) {
    long i = gid();
    if (i < n) r[i] = b[i] - w[i];
}
// CPR glue (skipped once the solve finished): x1 = [xp; 0], w = r - J x1, z[0:np] += xp
__global__ void k_cpr_embed(long n, long np, const double* __restrict__ xp, double* __restrict__ x1,
                            const double* __restrict__ gate) {
    long i = gid();
    if (i >= n || (gate && *gate == 0.0)) return;
    x1[i] = i < np ? xp[i] : 0.0;
}
__global__ void k_cpr_resid(long n, const double* __restrict__ r, const double* __restrict__ jx, double* __restrict__ w,
                            const double* __restrict__ gate) {
    long i = gid();
    if (i >= n || (gate && *gate == 0.0)) return;
    w[i] = r[i] - jx[i];
}
__global__ void k_cpr_add(long np, const double* __restrict__ xp, double* __restrict__ z,
                          const double* __restrict__ gate) {
    long i = gid();
    if (i >= np || (gate && *gate == 0.0)) return;
    z[i] += xp[i];
}

// ---- multicolour ILU(0) ----------------------------------------------------------
// ILU(0) rows of one colour (t = colour position, rows of earlier colours final):
// l = lu[q] / U(k,k) (0 without a pivot), then lu[g] -= l * lu[f] over the step's updates.
template <class T>
__global__ void k_ilu_factor(int cnt, const int* __restrict__ fsp, const int* __restrict__ fq,
                             const int* __restrict__ fdk, const int* __restrict__ fpp, const int* __restrict__ fg,
                             const int* __restrict__ ff, T* __restrict__ lu) {
    long t = gid();
    if (t >= cnt) return;
    for (int st = fsp[t]; st < fsp[t + 1]; ++st) {
        const int q = fq[st], dk = fdk[st];
        const T ukk = dk >= 0 ? lu[dk] : (T)0;
        const T l = ukk != (T)0 ? lu[q] / ukk : (T)0;
        lu[q] = l;
        if (l == (T)0) continue;
        for (int u = fpp[st]; u < fpp[st + 1]; ++u) lu[fg[u]] -= l * lu[ff[u]];
    }
}
// Triangular sweeps over the rows of one colour (t = position within the colour), from the
// per-colour SELL-64 arrays: lane t % 64 of slice t / 64, entries column-major, so the
// column and value loads of a wave are contiguous.  Entries of one row are visited in CSR
// order; column -1 marks padding, -2 - col an L column of colour 0, whose z is b (colour 0
// has no L entries, so its lower sweep is not run).  dg: the last colour, which has no U
// entries, divides by its pivot here and skips its upper sweep.
template <class T>
__global__ void k_ilu_lower(int cnt, const int* __restrict__ rows, const int* __restrict__ sp,
                            const int* __restrict__ col, const T* __restrict__ val, const T* __restrict__ b,
                            T* __restrict__ z, const T* __restrict__ dg, const double* __restrict__ gate) {
    if (gate && gate[0] == 0.0) return;
    long t = gid();
    if (t >= cnt) return;
    const int i = rows[t];
    const int s0 = sp[t / W], w = (sp[t / W + 1] - s0) / W;
    const int lane = (int)(t % W);
    T s = b[i];
    for (int j = 0; j < w; ++j) {
        const int k = s0 + j * W + lane;
        const int c = col[k];
        if (c >= 0) s -= val[k] * z[c];
        else if (c < -1) s -= val[k] * b[-2 - c];
    }
    if (dg) {
        const T d = dg[t];
        z[i] = d != (T)0 ? s / d : s;
    } else {
        z[i] = s;
    }
}
// b0: colour 0, whose lower sweep would have copied b into z, starts from b
template <class T>
__global__ void k_ilu_upper(int cnt, const int* __restrict__ rows, const int* __restrict__ sp,
                            const int* __restrict__ col, const T* __restrict__ val, const T* __restrict__ dg,
                            T* __restrict__ z, const T* __restrict__ b0, const double* __restrict__ gate) {
    if (gate && gate[0] == 0.0) return;
    long t = gid();
    if (t >= cnt) return;
    const int i = rows[t];
    const int s0 = sp[t / W], w = (sp[t / W + 1] - s0) / W;
    const int lane = (int)(t % W);
    T s = b0 ? b0[i] : z[i];
    for (int j = 0; j < w; ++j) {
        const int k = s0 + j * W + lane;
        const int c = col[k];
        if (c >= 0) s -= val[k] * z[c];
    }
    const T d = dg[t];
    z[i] = d != (T)0 ? s / d : s;
}
// Sweep values from the CSR factors: out[k] = lu[src[k]] (0 for padding)
template <class T>
__global__ void k_ilu_sweep_vals(long m, const int* __restrict__ src, const T* __restrict__ lu, T* __restrict__ out) {
    long k = gid();
    if (k >= m) return;
    const int s = src[k];
    out[k] = s >= 0 ? lu[s] : (T)0;
}
// Pivot per colour position: U(i,i), or 1 where row i has no stored diagonal
template <class T>
__global__ void k_ilu_pivots(int n, const int* __restrict__ rows, const int* __restrict__ diag,
                             const T* __restrict__ lu, T* __restrict__ dg) {
    long t = gid();
    if (t >= n) return;
    const int d = diag[rows[t]];
    dg[t] = d >= 0 ? lu[d] : (T)1;
}

inline unsigned nblk(long n) { return (unsigned)std::max(1L, (n + BS - 1) / BS); }

// ---------------------------------------------------------------------------
// Solver working set in precision T
template <class T>
struct System {
    int n = 0;
    LayoutP L;
    PatternP pat;
    DevPtr<T> sell;       // SELL values
    std::string precond;
    DevPtr<T> dinv;       // Jacobi
    DevPtr<T> lu;         // ILU(0) factors in CSR layout
    DevPtr<T> lval, uval, piv;   // the factors in the per-colour sweep layout, pivots by colour position
    DevPtr<double> sc, part;
    int nb = 1;
    const CprStage* cpr = nullptr;        // CPR (FP64): stage 1 on the pressure block, stage 2 ILU(0)
    DevPtr<double> cxp, cx1, cjx, cw;     // its work vectors

    void setup(const PatternP& p, const double* csr, const std::string& pc) {
        pat = p;
        n = p->rows;
        L = get_layout(p);
        sell = dev_alloc<T>(L->nnzp);
        KL(k_sell_vals<T>, nblk(L->nnzp), L->nnzp, L->src->p, csr, sell->p);
        precond = pc;
        if (pc == "jacobi") {
            dinv = dev_alloc<T>(n);
            KL(k_jacobi_setup<T>, nblk(n), n, L->diag->p, csr, dinv->p);
        } else if (pc == "ilu0") {
            lu = dev_alloc<T>(p->nnz);
            KL((k_cast<double, T>), nblk(p->nnz), (long)p->nnz, csr, 1.0, lu->p);
            for (int c = 1; c < L->ncolors; ++c) {
                const int cnt = L->color_off[c + 1] - L->color_off[c];
                if (cnt > 0)
                    KL(k_ilu_factor<T>, nblk(cnt), cnt, L->fsp->p + L->color_off[c], L->fq->p, L->fdk->p, L->fpp->p,
                       L->fg->p, L->ff->p, lu->p);
            }
            lval = dev_alloc<T>(std::max(1L, L->lnnz));
            uval = dev_alloc<T>(std::max(1L, L->unnz));
            piv = dev_alloc<T>(std::max(1, n));
            if (L->lnnz > 0) KL(k_ilu_sweep_vals<T>, nblk(L->lnnz), L->lnnz, L->lsrc->p, (const T*)lu->p, lval->p);
            if (L->unnz > 0) KL(k_ilu_sweep_vals<T>, nblk(L->unnz), L->unnz, L->usrc->p, (const T*)lu->p, uval->p);
            KL(k_ilu_pivots<T>, nblk(n), n, L->color_rows->p, L->diag->p, (const T*)lu->p, piv->p);
        }
        nb = vec_blocks(n);
        sc = dev_alloc<double>(NSLOT);
        part = dev_alloc<double>(3 * (size_t)nb + 8);
    }
    void spmv(const T* x, T* y, const double* gate = nullptr) const {
        TimedLaunch tl(sizeof(T) == 8 ? "spmv_fp64" : "spmv_fp32");
        KL(k_sell_spmv<T>, nblk((long)L->nslices * W), n, L->nslices, L->slice_ptr->p, L->col->p, sell->p, x, y,
           gate);
    }
    // z = M^{-1} r
    void apply_precond(const T* r, T* z, const double* gate = nullptr) const {
        if (cpr) {
            // stage 1: the pressure block's approximate inverse on the pressure rows
            cpr->apply((const double*)r, cxp->p);
            KL(k_cpr_embed, nblk(n), (long)n, (long)cpr->np, (const double*)cxp->p, cx1->p, gate);
            spmv((const T*)cx1->p, (T*)cjx->p, gate);
            KL(k_cpr_resid, nblk(n), (long)n, (const double*)r, (const double*)cjx->p, cw->p, gate);
            // stage 2: ILU(0) of the whole system on what stage 1 left, plus stage 1's part
            ilu_apply((const T*)cw->p, z, gate);
            KL(k_cpr_add, nblk(cpr->np), (long)cpr->np, (const double*)cxp->p, (double*)z, gate);
            return;
        }
        if (precond == "ilu0") {
            ilu_apply(r, z, gate);
        } else {
            KL(k_scale<T>, nblk(n), (long)n, precond == "jacobi" ? (const T*)dinv->p : (const T*)nullptr, r, z, gate);
        }
    }
    void ilu_apply(const T* r, T* z, const double* gate) const {
        {
            // colour 0 has no L entries (z = r there: read from r, no sweep); the last colour
            // has no U entries (its lower sweep divides by the pivots, no upper sweep)
            const int nc = L->ncolors;
            for (int c = 1; c < nc; ++c) {
                const int cnt = L->color_off[c + 1] - L->color_off[c];
                if (cnt > 0)
                    KL(k_ilu_lower<T>, nblk(cnt), cnt, L->color_rows->p + L->color_off[c], L->lsp->p + L->cslice[c],
                       L->lcol->p, (const T*)lval->p, r, z,
                       c == nc - 1 ? (const T*)piv->p + L->color_off[c] : (const T*)nullptr, gate);
            }
            for (int c = nc >= 2 ? nc - 2 : nc - 1; c >= 0; --c) {   // nc == 1: colour 0 alone
                const int cnt = L->color_off[c + 1] - L->color_off[c];
                if (cnt > 0)
                    KL(k_ilu_upper<T>, nblk(cnt), cnt, L->color_rows->p + L->color_off[c], L->usp->p + L->cslice[c],
                       L->ucol->p, (const T*)uval->p, (const T*)piv->p + L->color_off[c], z,
                       c == 0 ? r : (const T*)nullptr, gate);
            }
        }
    }
    const double* gate() const { return sc->p + ACTIVE; }
    // dot products + finalisation; gated ones are skipped once the solve finished
    void dots(int act, int nd, const T* x0, const T* y0, const T* x1 = nullptr, const T* y1 = nullptr,
              int slot = 0, bool gated = false, double tol2 = 0.0, int maxit = 0) const {
        KL(k_dots<T>, nb, (long)n, nd, x0, y0, x1, y1, (const T*)nullptr, (const T*)nullptr, part->p,
           gated ? gate() : (const double*)nullptr);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(BS), 0, S(), act, nd, part->p, nb, sc->p, slot, gated ? 1 : 0,
                           tol2, maxit);
        EQ_HIP(hipGetLastError());
    }
    double read_slot(int slot) const {
        double* h = Device::get().host_scratch();
        EQ_HIP(hipMemcpyAsync(h, sc->p + slot, sizeof(double), hipMemcpyDeviceToHost, S()));
        Device::get().sync();
        return h[0];
    }
    void read_slots(double* out, int count) const {
        double* h = Device::get().host_scratch();
        EQ_HIP(hipMemcpyAsync(h, sc->p, count * sizeof(double), hipMemcpyDeviceToHost, S()));
        Device::get().sync();
        for (int i = 0; i < count; ++i) out[i] = h[i];
    }
};

// Host view of a device flag without stalling the queue: the flag copied after chunk
// k is read once chunk k+1 has been queued (pinned ring + events).
struct FlagRing {
    static constexpr int NR = 4;
    hipEvent_t ev[NR] = {};
    double* host = nullptr;
    FlagRing() {
        host = Device::get().host_scratch() + 256;
        for (int i = 0; i < NR; ++i) EQ_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    }
    ~FlagRing() {
        for (int i = 0; i < NR; ++i)
            if (ev[i]) (void)hipEventDestroy(ev[i]);
    }
    void post(int k, const double* dflag) {
        EQ_HIP(hipMemcpyAsync(host + k % NR, dflag, sizeof(double), hipMemcpyDeviceToHost, S()));
        EQ_HIP(hipEventRecord(ev[k % NR], S()));
    }
    bool read(int k) {
        EQ_HIP(hipEventSynchronize(ev[k % NR]));
        return host[k % NR] != 0.0;
    }
};
constexpr int CHUNK = 4;   // iterations queued between flag reads

// BiCGStab, x0 = 0 (x zeroed here).  Relative tolerance on ||r|| / ||b||.  Iterations
// are queued in chunks; kernels of iterations past convergence are no-ops (gated).
template <class T>
int bicgstab(const System<T>& A, const T* b, T* x, double tol, int maxit, double* relres) {
    const long n = A.n;
    auto r = dev_alloc<T>(n), rh = dev_alloc<T>(n), p = dev_alloc<T>(n), v = dev_alloc<T>(n), s = dev_alloc<T>(n),
         t = dev_alloc<T>(n);
    const bool ilu = A.precond == "ilu0";
    auto ph = dev_alloc<T>(n), sh = dev_alloc<T>(n);
    const T* dinv = A.precond == "jacobi" ? A.dinv->p : nullptr;
    const double* g = A.gate();
    EQ_HIP(hipMemsetAsync(x, 0, n * sizeof(T), S()));
    EQ_HIP(hipMemsetAsync(A.sc->p, 0, NSLOT * sizeof(double), S()));
    EQ_HIP(hipMemcpyAsync(r->p, b, n * sizeof(T), hipMemcpyDeviceToDevice, S()));
    EQ_HIP(hipMemcpyAsync(rh->p, b, n * sizeof(T), hipMemcpyDeviceToDevice, S()));
    A.dots(ACT_INIT, 2, rh->p, r->p, r->p, r->p, 0, false, 0.0, maxit);   // rho = rh.r, rr = bb = r.r
    const double tol2 = tol * tol;
    FlagRing ring;
    int queued = 0;
    for (int k = 0; queued < maxit; ++k) {
        for (int q = 0; q < CHUNK && queued < maxit; ++q, ++queued) {
            KL(k_bicg_p<T>, nblk(n), n, queued == 0 ? 1 : 0, r->p, v->p, p->p, ilu ? (T*)nullptr : ph->p, dinv,
               A.sc->p);
            if (ilu) A.apply_precond(p->p, ph->p, g);
            A.spmv(ph->p, v->p, g);
            A.dots(ACT_BICG_ALPHA, 1, rh->p, v->p, nullptr, nullptr, 0, true);
            KL(k_bicg_s<T>, nblk(n), n, r->p, v->p, s->p, ilu ? (T*)nullptr : sh->p, dinv, A.sc->p);
            if (ilu) A.apply_precond(s->p, sh->p, g);
            A.spmv(sh->p, t->p, g);
            A.dots(ACT_BICG_OMEGA, 2, t->p, s->p, t->p, t->p, 0, true);
            KL(k_bicg_x<T>, A.nb, n, x, ph->p, sh->p, r->p, s->p, t->p, rh->p, A.sc->p, A.part->p);
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(BS), 0, S(), (int)ACT_BICG_RHO, 2, A.part->p, A.nb, A.sc->p,
                               0, 1, tol2, maxit);
            EQ_HIP(hipGetLastError());
        }
        ring.post(k, g);
        if (k > 0 && !ring.read(k - 1)) break;
    }
    double h[NSLOT];
    A.read_slots(h, NSLOT);
    const double bn = std::sqrt(std::max(h[BB], 0.0));
    if (relres) *relres = bn > 0 ? std::sqrt(std::max(h[RR], 0.0)) / bn : 0.0;
    return (int)h[IT];
}

template <class T>
int cg(const System<T>& A, const T* b, T* x, double tol, int maxit, double* relres) {
    const long n = A.n;
    auto r = dev_alloc<T>(n), z = dev_alloc<T>(n), p = dev_alloc<T>(n), q = dev_alloc<T>(n);
    const double* g = A.gate();
    EQ_HIP(hipMemsetAsync(x, 0, n * sizeof(T), S()));
    EQ_HIP(hipMemsetAsync(A.sc->p, 0, NSLOT * sizeof(double), S()));
    EQ_HIP(hipMemcpyAsync(r->p, b, n * sizeof(T), hipMemcpyDeviceToDevice, S()));
    A.apply_precond(r->p, z->p);
    A.dots(ACT_INIT, 2, r->p, z->p, r->p, r->p, 0, false, 0.0, maxit);   // rz, rr = bb
    const double tol2 = tol * tol;
    KL(k_cg_p<T>, nblk(n), n, 1, z->p, p->p, A.sc->p);
    FlagRing ring;
    int queued = 0;
    for (int k = 0; queued < maxit; ++k) {
        for (int c = 0; c < CHUNK && queued < maxit; ++c, ++queued) {
            A.spmv(p->p, q->p, g);
            A.dots(ACT_CG_ALPHA, 1, p->p, q->p, nullptr, nullptr, 0, true);
            KL(k_cg_update<T>, nblk(n), n, x, r->p, p->p, q->p, A.sc->p);
            A.apply_precond(r->p, z->p, g);
            A.dots(ACT_CG_BETA, 2, r->p, z->p, r->p, r->p, 0, true, tol2, maxit);
            KL(k_cg_p<T>, nblk(n), n, 0, z->p, p->p, A.sc->p);
        }
        ring.post(k, g);
        if (k > 0 && !ring.read(k - 1)) break;
    }
    double h[NSLOT];
    A.read_slots(h, NSLOT);
    const double bn = std::sqrt(std::max(h[BB], 0.0));
    if (relres) *relres = bn > 0 ? std::sqrt(std::max(h[RR], 0.0)) / bn : 0.0;
    return (int)h[IT];
}

// GMRES(m), right preconditioning, CGS2; x0 = 0.  The Arnoldi process, Givens
// rotations and the small least-squares solve run on the device (Gm); the host
// queues Arnoldi steps in chunks and reads the cycle's state only between chunks.
template <class T>
int gmres(const System<T>& A, const T* b, T* x, double tol, int maxit, int m, double* relres) {
    const long n = A.n;
    m = std::max(1, std::min(m, GM_MAX));
    auto V = dev_alloc<T>((size_t)n * (m + 1));
    auto w = dev_alloc<T>(n), z = dev_alloc<T>(n), r = dev_alloc<T>(n);
    auto gmb = dev_alloc<char>(sizeof(Gm));
    Gm* gm = (Gm*)gmb->p;
    const int nb = A.nb;
    auto part = dev_alloc<double>((size_t)nb * (m + 2));
    const unsigned nnb = std::min(nblk(n), 4096u);   // blocks of the update that sums ||w||^2
    auto npart = dev_alloc<double>(nnb);
    EQ_HIP(hipMemsetAsync(x, 0, n * sizeof(T), S()));
    EQ_HIP(hipMemsetAsync(gmb->p, 0, sizeof(Gm), S()));
    EQ_HIP(hipMemsetAsync(A.sc->p, 0, NSLOT * sizeof(double), S()));
    A.dots(ACT_STORE, 1, b, b, nullptr, nullptr, RR);
    const double bn = std::sqrt(std::max(A.read_slot(RR), 0.0));
    if (relres) *relres = 0.0;
    if (bn == 0.0 || maxit <= 0) return 0;
    const double hdr[4] = {1.0, bn, bn, tol};   // active, rn, bn, tol
    EQ_HIP(hipMemcpyAsync(gm, hdr, sizeof(hdr), hipMemcpyHostToDevice, S()));
    EQ_HIP(hipMemcpyAsync(r->p, b, n * sizeof(T), hipMemcpyDeviceToDevice, S()));
    const double* act = &gm->active;
    FlagRing ring;
    int posted = 0;
    Gm hg;
    // One Arnoldi step (column j): every kernel is gated by the device flag, so a step queued
    // past convergence is a no-op.
    // Jacobi or no preconditioner: each step's k_gmres_next also forms the next step's z
    const bool diag_pc = A.precond != "ilu0";
    const T* dpc = A.precond == "jacobi" ? (const T*)A.dinv->p : (const T*)nullptr;
    auto arnoldi = [&](int j) {
        T* vj = V->p + (size_t)j * n;
        if (j == 0 || !diag_pc) A.apply_precond(vj, z->p, act);
        A.spmv(z->p, w->p, act);
        for (int pass = 0; pass < 2; ++pass) {   // CGS2
            hipLaunchKernelGGL(k_vdots<T>, dim3(nb, j + 1), dim3(BS), 0, S(), n, j, (const T*)V->p, (const T*)w->p,
                               part->p, nb, (const Gm*)gm);
            EQ_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_vdots_final, dim3(j + 1), dim3(BS), 0, S(), nb, part->p, gm, pass);
            EQ_HIP(hipGetLastError());
            if (pass == 0) {
                KL(k_vupdate<T>, nblk(n), n, j, V->p, (const Gm*)gm, w->p);
            } else {   // the second update also sums ||w||^2; the Givens step finishes the sum
                KL(k_vupdate_norm<T>, nnb, n, j, (const T*)V->p, (const Gm*)gm, w->p, npart->p);
                hipLaunchKernelGGL(k_gmres_step_norm, dim3(1), dim3(BS), 0, S(), j, m, (const double*)npart->p,
                                   (int)nnb, A.sc->p, gm, maxit);
                EQ_HIP(hipGetLastError());
            }
        }
        KL(k_gmres_next<T>, nblk(n), n, w->p, V->p + (size_t)(j + 1) * n, (const Gm*)gm, dpc,
           diag_pc ? z->p : (T*)nullptr);
    };
    for (;;) {
        // cycle start: beta = ||r||, V0 = r / beta (decides whether to continue)
        A.dots(ACT_STORE, 1, r->p, r->p, nullptr, nullptr, RR);
        KL(k_gmres_start<T>, nblk(n), n, r->p, V->p, A.sc->p, gm, maxit);
        for (int j = 0; j < m;) {
            for (int c = 0; c < CHUNK && j < m; ++c, ++j) {
                arnoldi(j);
            }
            ring.post(posted, act);
            const bool more = posted == 0 || ring.read(posted - 1);
            ++posted;
            if (!more) break;
        }
        // x += M^{-1} V y over the cycle's columns
        hipLaunchKernelGGL(k_gmres_solve, dim3(1), dim3(BS), 0, S(), m, gm);
        EQ_HIP(hipGetLastError());
        KL(k_vcombine<T>, nblk(n), n, (const T*)V->p, (const Gm*)gm, w->p);
        A.apply_precond(w->p, z->p);
        KL(k_add<T>, nblk(n), n, z->p, x);
        EQ_HIP(hipMemcpyAsync(&hg, gm, offsetof(Gm, hcol), hipMemcpyDeviceToHost, S()));
        Device::get().sync();
        if (hg.rn <= tol * bn || !std::isfinite(hg.rn) || hg.total >= maxit || hg.k == 0) break;
        // true residual for the restart: r = b - A x
        A.spmv(x, w->p);
        KL(k_sub<T>, nblk(n), n, b, w->p, r->p);
        posted = 0;
    }
    if (relres) *relres = hg.rn / bn;
    return (int)hg.total;
}

template <class T>
int run_method(const SolverConfig& cfg, const System<T>& A, const T* b, T* x, double tol, int maxit, double* rr) {
    if (cfg.method == "bicgstab") return bicgstab<T>(A, b, x, tol, maxit, rr);
    if (cfg.method == "cg") return cg<T>(A, b, x, tol, maxit, rr);
    if (cfg.method == "gmres") return gmres<T>(A, b, x, tol, maxit, cfg.gmres_restart, rr);
    throw_error("unknown Krylov method " + cfg.method);
}

double dnorm(const double* a, long n) { return std::sqrt(reduce(RED_SUMSQ, a, n)); }

}  // namespace

// ===========================================================================
SP linear_solve(const SolverConfig& cfg, const SP& res, LinReport* rep, const CprStage* cpr) {
    EQ_REQUIRE(res->pat, "linear solve: the residual has no Jacobian");
    const PatternP& P = res->pat;
    const int n = res->n;
    EQ_REQUIRE(P->rows == n && P->cols == n, "linear solve: Jacobian is " << P->rows << "x" << P->cols
                                                                          << " for " << n << " unknowns");
    std::string pc = cfg.precond == "diagonal" ? "jacobi" : cfg.precond;
    LinReport R;
    auto x = dev_alloc<double>(n);
    if (cpr) {
        EQ_REQUIRE(cfg.precision == "fp64", "preconditioner=cpr runs in FP64 (precision=fp64)");
        EQ_REQUIRE(2 * cpr->np == n, "preconditioner=cpr: " << n << " unknowns are not two stacked halves");
        pc = "ilu0";
    }
    if (cfg.precision == "fp64") {
        System<double> A;
        A.setup(P, jac_values(*res), pc);
        if (cpr) {
            cpr->setup(P, jac_values(*res));
            A.cpr = cpr;
            A.cxp = dev_alloc<double>(std::max(1, cpr->np));
            A.cx1 = dev_alloc<double>(n);
            A.cjx = dev_alloc<double>(n);
            A.cw = dev_alloc<double>(n);
        }
        double rr = 0;
        R.iterations = run_method<double>(cfg, A, res->val->p, x->p, cfg.tol, cfg.max_iter, &rr);
        R.residual = rr;
        R.converged = rr <= cfg.tol;
    } else if (cfg.precision == "mixed") {
        System<float> A;
        A.setup(P, jac_values(*res), pc);
        auto r64 = dev_alloc<double>(n), y64 = dev_alloc<double>(n);
        auto r32 = dev_alloc<float>(n), d32 = dev_alloc<float>(n);
        EQ_HIP(hipMemsetAsync(x->p, 0, n * sizeof(double), S()));
        EQ_HIP(hipMemcpyAsync(r64->p, res->val->p, n * sizeof(double), hipMemcpyDeviceToDevice, S()));
        const double bn = dnorm(res->val->p, n);
        double rn = bn;
        int total = 0;
        while (bn > 0 && rn > cfg.tol * bn && R.refinements < cfg.max_refine && total < cfg.max_iter) {
            // FP32 solve of A d = r / ||r|| (scaled to unit norm), then x += ||r|| d
            KL((k_cast<double, float>), nblk(n), (long)n, r64->p, 1.0 / rn, r32->p);
            double rr = 0;
            total += run_method<float>(cfg, A, r32->p, d32->p, std::max(cfg.inner_tol, 1e-6),
                                       cfg.max_iter - total, &rr);
            KL(k_axpy_mixed<float>, nblk(n), (long)n, rn, d32->p, x->p);
            // FP64 defect
            launch_csr_spmv(n, P->d_rp->p, P->d_ci->p, jac_values(*res), x->p, y64->p);
            KL(k_residual, nblk(n), (long)n, res->val->p, y64->p, r64->p);
            const double rnew = dnorm(r64->p, n);
            ++R.refinements;
            if (!std::isfinite(rnew) || rnew >= rn) {   // stagnation
                rn = rnew;
                break;
            }
            rn = rnew;
        }
        R.iterations = total;
        R.residual = bn > 0 ? rn / bn : 0.0;
        R.converged = R.residual <= cfg.tol;
    } else {
        throw_error("unknown precision " + cfg.precision + " (fp64 | mixed)");
    }
    if (cfg.verbose > 0)
        fprintf(stdout, "Linear solver (%s/%s/%s): %d iterations, relative residual %g%s\n", cfg.method.c_str(),
                pc.c_str(), cfg.precision.c_str(), R.iterations, R.residual, R.converged ? "" : " (not converged)");
    if (rep) *rep = R;
    return s_from_device(x, n);
}

void sell_spmv_test(const SP& a, const double* xh, double* yh, bool fp32) {
    EQ_REQUIRE(a->pat, "sell_spmv_test: no Jacobian");
    const int n = a->n;
    auto x = dev_from_host(xh, n);
    auto y = dev_alloc<double>(n);
    if (fp32) {
        System<float> A;
        A.setup(a->pat, jac_values(*a), "none");
        auto x32 = dev_alloc<float>(n), y32 = dev_alloc<float>(n);
        KL((k_cast<double, float>), nblk(n), (long)n, x->p, 1.0, x32->p);
        A.spmv(x32->p, y32->p);
        KL((k_cast<float, double>), nblk(n), (long)n, y32->p, 1.0, y->p);
    } else {
        System<double> A;
        A.setup(a->pat, jac_values(*a), "none");
        A.spmv(x->p, y->p);
    }
    std::vector<double> h = dev_to_host(y->p, n);
    std::copy(h.begin(), h.end(), yh);
}

}  // namespace eqhip
