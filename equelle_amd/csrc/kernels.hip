// HIP kernels (gfx950, wave64) for the AD engine, grid operators and reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <hipcub/hipcub.hpp>

#include "core.hpp"
#include "kernels.hpp"

// Scratch of the allocation-free reduction (reduce_noalloc): the block partials of
// k_reduce_partial and the final value, NOALLOC_SCRATCH_DOUBLES doubles (16 KiB).  Each GPU
// holds its own copy for the life of the process.  Host code never passes this name to a
// kernel or a copy: it is a host-side stand-in, not a device address; the address comes
// from hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) double scratch[eqhip::NOALLOC_SCRATCH_DOUBLES];

namespace eqhip {

namespace {
constexpr int BS = 256;   // 4 wave64 per block

inline hipStream_t S() { return Device::get().stream(); }
inline unsigned blocks_for(long n, int per = BS) {
    long b = (n + per - 1) / per;
    return (unsigned)(b > 0 ? b : 1);
}
#define EQ_LAUNCH(kernel, grid, ...)                                          \
    do {                                                                       \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(BS), 0, S(), __VA_ARGS__); \
        EQ_HIP(hipGetLastError());                                             \
    } while (0)

__device__ inline long gid() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }

// ---------------------------------------------------------------------------
__global__ void k_row_of(const int* __restrict__ rp, int rows, int* __restrict__ row) {
    long r = gid();
    if (r >= rows) return;
    for (int k = rp[r]; k < rp[r + 1]; ++k) row[k] = (int)r;
}

// ---------------------------------------------------------------------------
// Symbolic passes (thread per row; rows hold ascending columns).
__device__ inline unsigned long long pmix(unsigned long long x) {   // pattern_mix
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
__global__ void k_pattern_hash(const int* __restrict__ rp, int nrp, const int* __restrict__ ci, int nnz,
                               unsigned long long* __restrict__ acc) {
    unsigned long long h = 0;
    const long n = (long)nrp + nnz, stride = (long)gridDim.x * blockDim.x;
    for (long k = gid(); k < n; k += stride)
        h += k < nrp ? pmix(((unsigned long long)k << 32) | (unsigned)rp[k])
                     : pmix(((unsigned long long)((k - nrp) | 0x80000000ull) << 32) | (unsigned)ci[k - nrp]);
    for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(acc, h);   // integer sum: order-independent
}
__global__ void k_int_differ(const int* __restrict__ a, const int* __restrict__ b, long n, int* __restrict__ flag) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = gid(); k < n; k += stride)
        if (a[k] != b[k]) *flag = 1;
}
__global__ void k_set_int(int* p, int v) { *p = v; }
__global__ void k_set_u64(unsigned long long* p) { *p = 0; }

__global__ void k_union_count(int rows, const int* __restrict__ arp, const int* __restrict__ aci,
                              const int* __restrict__ brp, const int* __restrict__ bci, int* __restrict__ cnt) {
    const long r = gid();
    if (r > rows) return;
    if (r == rows) {
        cnt[r] = 0;
        return;
    }
    int ka = arp[r], kb = brp[r], c = 0;
    const int ea = arp[r + 1], eb = brp[r + 1];
    while (ka < ea || kb < eb) {
        if (kb >= eb || (ka < ea && aci[ka] < bci[kb])) ++ka;
        else if (ka >= ea || bci[kb] < aci[ka]) ++kb;
        else ++ka, ++kb;
        ++c;
    }
    cnt[r] = c;
}
__global__ void k_union_fill(int rows, const int* __restrict__ arp, const int* __restrict__ aci,
                             const int* __restrict__ brp, const int* __restrict__ bci, const int* __restrict__ orp,
                             int* __restrict__ oci, int* __restrict__ ia, int* __restrict__ ib) {
    const long r = gid();
    if (r >= rows) return;
    int ka = arp[r], kb = brp[r], o = orp[r];
    const int ea = arp[r + 1], eb = brp[r + 1];
    while (ka < ea || kb < eb) {
        if (kb >= eb || (ka < ea && aci[ka] < bci[kb])) {
            oci[o] = aci[ka];
            ia[o] = ka++;
            ib[o] = -1;
        } else if (ka >= ea || bci[kb] < aci[ka]) {
            oci[o] = bci[kb];
            ia[o] = -1;
            ib[o] = kb++;
        } else {
            oci[o] = aci[ka];
            ia[o] = ka++;
            ib[o] = kb++;
        }
        ++o;
    }
}
__global__ void k_rows_count(int out_rows, const int* __restrict__ arp, const int* __restrict__ map,
                             int* __restrict__ cnt) {
    const long i = gid();
    if (i > out_rows) return;
    const int r = i < out_rows ? map[i] : -1;
    cnt[i] = r >= 0 ? arp[r + 1] - arp[r] : 0;
}
__global__ void k_rows_fill(int out_rows, const int* __restrict__ arp, const int* __restrict__ aci,
                            const int* __restrict__ map, const int* __restrict__ orp, int* __restrict__ oci,
                            int* __restrict__ src) {
    const long i = gid();
    if (i >= out_rows) return;
    const int r = map[i];
    if (r < 0) return;
    int o = orp[i];
    for (int k = arp[r]; k < arp[r + 1]; ++k, ++o) {
        oci[o] = aci[k];
        src[o] = k;
    }
}
// Structural product, row r: the contributions are the pairs (t, e) with t an entry of M's
// row r and e an entry of A's row mci[t], in (t, e) order; a column is new at its first
// contribution (A's rows have distinct columns, so earlier contributions of the same t
// never repeat it).
__global__ void k_spgemm_count(int rows, const int* __restrict__ mrp, const int* __restrict__ mci,
                               const int* __restrict__ arp, const int* __restrict__ aci, int cmax,
                               int* __restrict__ ucnt, int* __restrict__ ccnt, int* __restrict__ over) {
    const long r = gid();
    if (r > rows) return;
    if (r == rows) {
        ucnt[r] = ccnt[r] = 0;
        return;
    }
    int nc = 0;
    for (int t = mrp[r]; t < mrp[r + 1]; ++t) nc += arp[mci[t] + 1] - arp[mci[t]];
    ccnt[r] = nc;
    if (nc > cmax) {
        *over = 1;
        ucnt[r] = 0;
        return;
    }
    int u = 0;
    for (int t = mrp[r]; t < mrp[r + 1]; ++t) {
        const int k = mci[t];
        for (int e = arp[k]; e < arp[k + 1]; ++e) {
            const int col = aci[e];
            bool seen = false;
            for (int t2 = mrp[r]; t2 < t && !seen; ++t2) {
                const int k2 = mci[t2];
                for (int e2 = arp[k2]; e2 < arp[k2 + 1]; ++e2)
                    if (aci[e2] == col) {
                        seen = true;
                        break;
                    }
            }
            u += seen ? 0 : 1;
        }
    }
    ucnt[r] = u;
}
__global__ void k_spgemm_fill(int rows, const int* __restrict__ mrp, const int* __restrict__ mci,
                              const int* __restrict__ arp, const int* __restrict__ aci,
                              const int8_t* __restrict__ msign, const int* __restrict__ orp,
                              const int* __restrict__ cp, int* __restrict__ oci, int* __restrict__ seg,
                              int* __restrict__ mi, int* __restrict__ ai, int* __restrict__ sai) {
    const long r = gid();
    if (r >= rows) return;
    int o = orp[r], c = cp[r];
    long last = -1;   // columns emitted so far are <= last
    while (true) {
        int mn = 0x7fffffff;
        for (int t = mrp[r]; t < mrp[r + 1]; ++t) {
            const int k = mci[t];
            for (int e = arp[k]; e < arp[k + 1]; ++e) {
                const int col = aci[e];
                if (col > last && col < mn) mn = col;
            }
        }
        if (mn == 0x7fffffff) break;
        oci[o] = mn;
        seg[o] = c;
        for (int t = mrp[r]; t < mrp[r + 1]; ++t) {   // the column's contributions, in t order
            const int k = mci[t];
            for (int e = arp[k]; e < arp[k + 1]; ++e)
                if (aci[e] == mn) {
                    if (sai) {
                        sai[c] = msign[t] < 0 ? ~e : e;
                    } else {
                        mi[c] = t;
                        ai[c] = e;
                    }
                    ++c;
                    break;
                }
        }
        ++o;
        last = mn;
    }
}

__global__ void k_fill(double* out, double v, long n) {
    long i = gid();
    if (i < n) out[i] = v;
}
__global__ void k_fill_int(int* out, int v, long n) {
    long i = gid();
    if (i < n) out[i] = v;
}

__global__ void k_ew(int op, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out,
                     int n) {
    long i = gid();
    if (i >= n) return;
    const double x = a[i], y = b[i];
    double r;
    switch (op) {
        case EW_ADD: r = x + y; break;
        case EW_SUB: r = x - y; break;
        case EW_MUL: r = x * y; break;
        default: r = x / y; break;
    }
    out[i] = r;
}

__global__ void k_ew_scalar(int op, double s, const double* __restrict__ a, double* __restrict__ out, int n,
                            int scalar_left) {
    long i = gid();
    if (i >= n) return;
    const double x = a[i];
    out[i] = op == EW_MUL ? s * x : (scalar_left ? s / x : x / s);
}

__global__ void k_sqrt(const double* __restrict__ a, double* __restrict__ out, int n) {
    long i = gid();
    if (i < n) out[i] = sqrt(a[i]);
}

__global__ void k_coef(int kind, const double* __restrict__ a, const double* __restrict__ b, double s,
                       double* __restrict__ out, int n) {
    long i = gid();
    if (i >= n) return;
    double r;
    switch (kind) {
        case 0: r = 1.0 / b[i]; break;
        case 1: r = -(a[i] / (b[i] * b[i])); break;
        case 2: r = -(s / (a[i] * a[i])); break;
        case 3: r = 0.5 / a[i]; break;
        case 5: r = a[i] >= 0.0 ? 1.0 : -1.0; break;   // sign factor of norm(CollOfScalar)
        default: r = 1.0 / (b[i] * b[i]); break;
    }
    out[i] = r;
}

__device__ inline bool cmpf(int op, double a, double b) {
    switch (op) {
        case CMP_GT: return a > b;
        case CMP_GE: return a >= b;
        case CMP_LT: return a < b;
        case CMP_LE: return a <= b;
        case CMP_EQ: return a == b;
        default: return a != b;
    }
}

__global__ void k_compare(int op, const double* __restrict__ a, const double* __restrict__ b, double s,
                          int scalar_left, uint8_t* __restrict__ out, int n) {
    long i = gid();
    if (i >= n) return;
    const double x = a[i];
    const double y = b ? b[i] : s;
    out[i] = (b || !scalar_left) ? cmpf(op, x, y) : cmpf(op, s, x);
}

__global__ void k_select(const uint8_t* __restrict__ p, const double* __restrict__ a, const double* __restrict__ b,
                         double* __restrict__ out, int n) {
    long i = gid();
    if (i < n) out[i] = p[i] ? a[i] : b[i];
}
__global__ void k_select_int(const uint8_t* __restrict__ p, const int* __restrict__ a, const int* __restrict__ b,
                             int* __restrict__ out, int n) {
    long i = gid();
    if (i < n) out[i] = p[i] ? a[i] : b[i];
}
__global__ void k_is_empty(const int* __restrict__ idx, uint8_t* __restrict__ out, int n) {
    long i = gid();
    if (i < n) out[i] = idx[i] < 0;
}

__global__ void k_gather(const double* __restrict__ in, const int* __restrict__ map, double* __restrict__ out,
                         long n) {
    long i = gid();
    if (i >= n) return;
    const int m = map[i];
    out[i] = m >= 0 ? in[m] : 0.0;
}
__global__ void k_gather_int(const int* __restrict__ in, const int* __restrict__ map, int* __restrict__ out, int n) {
    long i = gid();
    if (i >= n) return;
    const int m = map[i];
    out[i] = m >= 0 ? in[m] : -1;
}

// ---------------------------------------------------------------------------
// Jacobian numeric kernels
__global__ void k_jlin(long nnz, const int* __restrict__ row, const double* __restrict__ a,
                       const int* __restrict__ ia, const double* __restrict__ A, const double* __restrict__ RA,
                       double alpha, const double* __restrict__ b, const int* __restrict__ ib,
                       const double* __restrict__ B, const double* __restrict__ RB, double beta,
                       const double* __restrict__ C, double* __restrict__ out) {
    long k = gid();
    if (k >= nnz) return;
    const bool need_row = A || B || C || RA || RB;
    const int r = need_row ? row[k] : 0;
    double s = 0.0;
    if (a) {
        const long j = ia ? ia[k] : k;
        if (j >= 0) s = alpha * (A ? A[r] : 1.0) * (RA ? RA[r] : 1.0) * a[j];
    }
    if (b) {
        const long j = ib ? ib[k] : k;
        if (j >= 0) s += beta * (B ? B[r] : 1.0) * (RB ? RB[r] : 1.0) * b[j];
    }
    out[k] = C ? C[r] * s : s;
}

__global__ void k_jselect(long nnz, const int* __restrict__ row, const uint8_t* __restrict__ p,
                          const double* __restrict__ a, const int* __restrict__ ia, double sa,
                          const double* __restrict__ RA, const double* __restrict__ b, const int* __restrict__ ib,
                          double sb, const double* __restrict__ RB, double* __restrict__ out) {
    long k = gid();
    if (k >= nnz) return;
    const int r = row[k];
    double v = 0.0;
    if (p[r]) {
        if (a) {
            const long j = ia ? ia[k] : k;
            if (j >= 0) v = sa * (RA ? RA[r] : 1.0) * a[j];
        }
    } else if (b) {
        const long j = ib ? ib[k] : k;
        if (j >= 0) v = sb * (RB ? RB[r] : 1.0) * b[j];
    }
    out[k] = v;
}

__global__ void k_spgemm_num(long nnz, const int* __restrict__ seg, const int* __restrict__ mi,
                             const int* __restrict__ ai, const double* __restrict__ mv, const double* __restrict__ av,
                             const int* __restrict__ arow, const double* __restrict__ ar, double scale,
                             double* __restrict__ out) {
    long k = gid();
    if (k >= nnz) return;
    double s = 0.0;
    if (ar) {
        for (int t = seg[k]; t < seg[k + 1]; ++t) s += mv[mi[t]] * (ar[arow[ai[t]]] * av[ai[t]]);
    } else {
        for (int t = seg[k]; t < seg[k + 1]; ++t) s += mv[mi[t]] * av[ai[t]];
    }
    out[k] = scale * s;
}

__global__ void k_spgemm_signed(long nnz, const int* __restrict__ seg, const int* __restrict__ sai,
                                const double* __restrict__ av, const int* __restrict__ arow,
                                const double* __restrict__ ar, double scale, double* __restrict__ out) {
    long k = gid();
    if (k >= nnz) return;
    double s = 0.0;
    for (int t = seg[k]; t < seg[k + 1]; ++t) {
        const int e = sai[t];
        const int j = e >= 0 ? e : ~e;
        const double v = ar ? ar[arow[j]] * av[j] : av[j];
        s += e >= 0 ? v : -v;
    }
    out[k] = scale * s;
}

__global__ void k_jscale(long nnz, const int* __restrict__ row, double s, const double* __restrict__ R,
                         const int* __restrict__ map, const double* __restrict__ v, double* __restrict__ out) {
    long k = gid();
    if (k >= nnz) return;
    out[k] = s * (R ? R[row[k]] : 1.0) * v[map ? map[k] : k];
}

__global__ void k_csr_spmv(int rows, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ v, const double* __restrict__ x, double* __restrict__ y) {
    long r = gid();
    if (r >= rows) return;
    double s = 0.0;
    for (int k = rp[r]; k < rp[r + 1]; ++k) s += v[k] * x[ci[k]];
    y[r] = s;
}

// ---------------------------------------------------------------------------
// Grid operators
__global__ void k_gradient(const double* __restrict__ u, const int* __restrict__ c1, const int* __restrict__ c2,
                           double* __restrict__ out, int nif) {
    long i = gid();
    if (i < nif) out[i] = u[c2[i]] - u[c1[i]];
}

// The block's cells hold one contiguous range of the signed face lists (2 dim faces per
// cell, at most 6 BS): staged through LDS by coalesced loads, then each cell sums its own.
__global__ void k_divergence(const double* __restrict__ flux, const int* __restrict__ facepos,
                             const int* __restrict__ sfaces, double* __restrict__ out, int nc) {
    __shared__ int lf[6 * BS];
    const long c0 = (long)blockIdx.x * BS, c = c0 + threadIdx.x;
    const int b = facepos[c0], e = facepos[min(c0 + BS, (long)nc)];
    const bool staged = e - b <= 6 * BS;   // block-uniform; always for a Cartesian grid
    if (staged)
        for (int k = b + (int)threadIdx.x; k < e; k += BS) lf[k - b] = sfaces[k];
    __syncthreads();
    if (c >= nc) return;
    double s = 0.0;
    for (int k = facepos[c]; k < facepos[c + 1]; ++k) {
        const int f = staged ? lf[k - b] : sfaces[k];
        s += f >= 0 ? flux[f] : -flux[~f];
    }
    out[c] = s;
}

__global__ void k_face_flags(const int* __restrict__ fc, int nf, uint8_t* __restrict__ interior) {
    long f = gid();
    if (f < nf) interior[f] = fc[2 * f] >= 0 && fc[2 * f + 1] >= 0;
}

__global__ void k_cell_boundary(const int* __restrict__ fc, const int* __restrict__ facepos,
                                const int* __restrict__ cf, int nc, uint8_t* __restrict__ bnd) {
    long c = gid();
    if (c >= nc) return;
    uint8_t b = 0;
    for (int k = facepos[c]; k < facepos[c + 1]; ++k) {
        const int f = cf[k];
        b |= (fc[2 * f] < 0) | (fc[2 * f + 1] < 0);
    }
    bnd[c] = b;
}

__global__ void k_gather_vec(const double* __restrict__ src, const int* __restrict__ idx, int n, int dim,
                             double* __restrict__ out) {
    long t = gid();
    if (t >= (long)n * dim) return;
    const long i = t / dim, d = t - i * dim;
    out[t] = src[(long)idx[i] * dim + d];
}

__global__ void k_gather_face_cells(const int* __restrict__ fc, const int* __restrict__ faces, int n, int which,
                                    int* __restrict__ out) {
    long i = gid();
    if (i < n) out[i] = fc[2 * (long)faces[i] + which];
}

__global__ void k_vec_norm(const double* __restrict__ v, int n, int dim, double* __restrict__ out) {
    long i = gid();
    if (i >= n) return;
    double s = 0.0;
    for (int d = 0; d < dim; ++d) {
        const double x = v[i * dim + d];
        s += x * x;
    }
    out[i] = sqrt(s);
}

__global__ void k_vec_dot(const double* __restrict__ a, const double* __restrict__ b, int n, int dim,
                          double* __restrict__ out) {
    long i = gid();
    if (i >= n) return;
    double s = a[i * dim] * b[i * dim];
    for (int d = 1; d < dim; ++d) s += a[i * dim + d] * b[i * dim + d];
    out[i] = s;
}

__global__ void k_vec_scale(const double* __restrict__ v, const double* __restrict__ s, double sc, int n, int dim,
                            int divide, double* __restrict__ out) {
    long t = gid();
    if (t >= (long)n * dim) return;
    const double f = s ? s[t / dim] : sc;
    out[t] = divide ? v[t] / f : v[t] * f;
}

__global__ void k_unit_normals(const double* __restrict__ nrm, const int* __restrict__ idx, int n, int dim,
                               double* __restrict__ out) {
    long i = gid();
    if (i >= n) return;
    const long f = idx[i];
    double s = 0.0;
    for (int d = 0; d < dim; ++d) s += nrm[f * dim + d] * nrm[f * dim + d];
    s = sqrt(s);
    for (int d = 0; d < dim; ++d) out[i * dim + d] = nrm[f * dim + d] / s;
}

__global__ void k_vec_col(const double* __restrict__ v, int n, int dim, int col, double* __restrict__ out) {
    long i = gid();
    if (i < n) out[i] = v[i * dim + col];
}

// ---------------------------------------------------------------------------
// Reductions: wave64 shuffles + LDS, fixed order.
__device__ inline double red_op(int op, double a, double b) {
    switch (op) {
        case RED_PROD: return a * b;
        case RED_MAX: return a > b ? a : b;
        case RED_MIN: return a < b ? a : b;
        default: return a + b;
    }
}
__device__ inline double red_init(int op) {
    switch (op) {
        case RED_PROD: return 1.0;
        case RED_MAX: return -DBL_MAX;
        case RED_MIN: return DBL_MAX;
        default: return 0.0;
    }
}

__device__ inline double block_reduce(int op, double v) {
    __shared__ double lds[BS / 64];
    for (int off = 32; off > 0; off >>= 1) v = red_op(op, v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = lds[0];
        for (int i = 1; i < BS / 64; ++i) s = red_op(op, s, lds[i]);
        lds[0] = s;
    }
    __syncthreads();
    const double r = lds[0];
    __syncthreads();
    return r;
}

__global__ void k_reduce_partial(int op, const double* __restrict__ a, const double* __restrict__ b, long n,
                                 double* __restrict__ part) {
    double v = red_init(op == RED_SUMSQ ? RED_SUM : op);
    const int rop = op == RED_SUMSQ ? RED_SUM : op;
    for (long i = gid(); i < n; i += (long)gridDim.x * BS) {
        double x = a[i];
        if (op == RED_SUMSQ) x = b ? x * b[i] : x * x;
        v = red_op(rop, v, x);
    }
    v = block_reduce(rop, v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void k_reduce_final(int op, const double* __restrict__ part, int np, double* __restrict__ out) {
    const int rop = op == RED_SUMSQ ? RED_SUM : op;
    double v = red_init(rop);
    for (int i = threadIdx.x; i < np; i += BS) v = red_op(rop, v, part[i]);
    v = block_reduce(rop, v);
    if (threadIdx.x == 0) out[0] = v;
}

inline int reduce_blocks(long n) {
    long b = (n + BS * 8 - 1) / (BS * 8);
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (int)b;
}

// ---------------------------------------------------------------------------
// Compaction
constexpr int CT = 4;   // elements per thread
__global__ void k_compact_count(const uint8_t* __restrict__ f, int n, int want, int* __restrict__ counts) {
    int c = 0;
    const long base = ((long)blockIdx.x * BS + threadIdx.x) * CT;
    for (int j = 0; j < CT; ++j) {
        const long i = base + j;
        if (i < n && (f[i] != 0) == (want != 0)) ++c;
    }
    __shared__ int lds[BS];
    lds[threadIdx.x] = c;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = lds[0];
}

__global__ void k_compact_scan(int* __restrict__ counts, int nb, int* __restrict__ total) {
    // single block exclusive scan over nb block counts (sequential chunks of BS)
    __shared__ int lds[BS];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += BS) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? counts[i] : 0;
        lds[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < BS; off <<= 1) {
            const int t = (int)threadIdx.x >= off ? lds[threadIdx.x - off] : 0;
            __syncthreads();
            lds[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) counts[i] = carry + lds[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == BS - 1) carry += lds[BS - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) total[0] = carry;
}

__global__ void k_compact_write(const uint8_t* __restrict__ f, int n, int want, const int* __restrict__ offs,
                                int* __restrict__ out) {
    int c = 0;
    const long base = ((long)blockIdx.x * BS + threadIdx.x) * CT;
    for (int j = 0; j < CT; ++j) {
        const long i = base + j;
        if (i < n && (f[i] != 0) == (want != 0)) ++c;
    }
    __shared__ int lds[BS];
    lds[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < BS; off <<= 1) {
        const int t = (int)threadIdx.x >= off ? lds[threadIdx.x - off] : 0;
        __syncthreads();
        lds[threadIdx.x] += t;
        __syncthreads();
    }
    int pos = offs[blockIdx.x] + lds[threadIdx.x] - c;
    for (int j = 0; j < CT; ++j) {
        const long i = base + j;
        if (i < n && (f[i] != 0) == (want != 0)) out[pos++] = (int)i;
    }
}
}  // namespace

// ===========================================================================
void launch_row_of(const int* rp, int rows, int* row) { EQ_LAUNCH(k_row_of, blocks_for(rows), rp, rows, row); }

uint64_t pattern_entries_hash(const int* rp, int nrp, const int* ci, int nnz) {
    auto acc = dev_alloc<unsigned long long>(1);
    EQ_LAUNCH(k_set_u64, 1, acc->p);
    const long n = (long)nrp + nnz;
    EQ_LAUNCH(k_pattern_hash, std::min<long>(blocks_for(n), 4096), rp, nrp, ci, nnz, acc->p);
    return (uint64_t)dev_to_host(acc->p, 1)[0];
}
bool int_arrays_differ(const int* a, const int* b, long n) {
    if (n <= 0) return false;
    auto flag = dev_alloc<int>(1);
    EQ_LAUNCH(k_set_int, 1, flag->p, 0);
    EQ_LAUNCH(k_int_differ, std::min<long>(blocks_for(n), 4096), a, b, n, flag->p);
    return dev_to_host(flag->p, 1)[0] != 0;
}
void exclusive_scan(const int* in, int* out, int n) {
    if (n <= 0) return;
    size_t bytes = 0;
    EQ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, S()));
    auto tmp = dev_alloc<char>(std::max<size_t>(bytes, 1));
    EQ_HIP(hipcub::DeviceScan::ExclusiveSum(tmp->p, bytes, in, out, n, S()));
}
void launch_union_count(int rows, const int* arp, const int* aci, const int* brp, const int* bci, int* cnt) {
    EQ_LAUNCH(k_union_count, blocks_for(rows + 1L), rows, arp, aci, brp, bci, cnt);
}
void launch_union_fill(int rows, const int* arp, const int* aci, const int* brp, const int* bci, const int* orp,
                       int* oci, int* ia, int* ib) {
    if (rows > 0) EQ_LAUNCH(k_union_fill, blocks_for(rows), rows, arp, aci, brp, bci, orp, oci, ia, ib);
}
void launch_rows_count(int out_rows, const int* arp, const int* map, int* cnt) {
    EQ_LAUNCH(k_rows_count, blocks_for(out_rows + 1L), out_rows, arp, map, cnt);
}
void launch_rows_fill(int out_rows, const int* arp, const int* aci, const int* map, const int* orp, int* oci,
                      int* src) {
    if (out_rows > 0) EQ_LAUNCH(k_rows_fill, blocks_for(out_rows), out_rows, arp, aci, map, orp, oci, src);
}
void launch_spgemm_count(int rows, const int* mrp, const int* mci, const int* arp, const int* aci, int cmax,
                         int* ucnt, int* ccnt, int* over) {
    EQ_LAUNCH(k_spgemm_count, blocks_for(rows + 1L), rows, mrp, mci, arp, aci, cmax, ucnt, ccnt, over);
}
void launch_spgemm_fill(int rows, const int* mrp, const int* mci, const int* arp, const int* aci,
                        const int8_t* msign, const int* orp, const int* cp, int* oci, int* seg, int* mi, int* ai,
                        int* sai) {
    if (rows > 0)
        EQ_LAUNCH(k_spgemm_fill, blocks_for(rows), rows, mrp, mci, arp, aci, msign, orp, cp, oci, seg, mi, ai, sai);
}
void launch_fill(double* out, double v, long n) {
    if (n > 0) EQ_LAUNCH(k_fill, blocks_for(n), out, v, n);
}
void launch_fill_int(int* out, int v, long n) {
    if (n > 0) EQ_LAUNCH(k_fill_int, blocks_for(n), out, v, n);
}
void launch_ew(int op, const double* a, const double* b, double* out, int n) {
    if (n > 0) EQ_LAUNCH(k_ew, blocks_for(n), op, a, b, out, n);
}
void launch_ew_scalar(int op, double s, const double* a, double* out, int n, int scalar_left) {
    if (n > 0) EQ_LAUNCH(k_ew_scalar, blocks_for(n), op, s, a, out, n, scalar_left);
}
void launch_sqrt(const double* a, double* out, int n) {
    if (n > 0) EQ_LAUNCH(k_sqrt, blocks_for(n), a, out, n);
}
void launch_coef(int kind, const double* a, const double* b, double s, double* out, int n) {
    if (n > 0) EQ_LAUNCH(k_coef, blocks_for(n), kind, a, b, s, out, n);
}
void launch_compare(int op, const double* a, const double* b, double s, int scalar_left, uint8_t* out, int n) {
    if (n > 0) EQ_LAUNCH(k_compare, blocks_for(n), op, a, b, s, scalar_left, out, n);
}
void launch_select(const uint8_t* p, const double* a, const double* b, double* out, int n) {
    if (n > 0) EQ_LAUNCH(k_select, blocks_for(n), p, a, b, out, n);
}
void launch_select_int(const uint8_t* p, const int* a, const int* b, int* out, int n) {
    if (n > 0) EQ_LAUNCH(k_select_int, blocks_for(n), p, a, b, out, n);
}
void launch_is_empty(const int* idx, uint8_t* out, int n) {
    if (n > 0) EQ_LAUNCH(k_is_empty, blocks_for(n), idx, out, n);
}
void launch_gather(const double* in, const int* map, double* out, long n) {
    if (n > 0) EQ_LAUNCH(k_gather, blocks_for(n), in, map, out, n);
}
void launch_gather_int(const int* in, const int* map, int* out, int n) {
    if (n > 0) EQ_LAUNCH(k_gather_int, blocks_for(n), in, map, out, n);
}
void launch_jlin(long nnz, const int* row, const double* a, const int* ia, const double* A, const double* RA,
                 double alpha, const double* b, const int* ib, const double* B, const double* RB, double beta,
                 const double* C, double* out) {
    if (nnz > 0) EQ_LAUNCH(k_jlin, blocks_for(nnz), nnz, row, a, ia, A, RA, alpha, b, ib, B, RB, beta, C, out);
}
void launch_jselect(long nnz, const int* row, const uint8_t* p, const double* a, const int* ia, double sa,
                    const double* RA, const double* b, const int* ib, double sb, const double* RB, double* out) {
    if (nnz > 0) EQ_LAUNCH(k_jselect, blocks_for(nnz), nnz, row, p, a, ia, sa, RA, b, ib, sb, RB, out);
}
void launch_spgemm_num(long nnz, const int* seg, const int* mi, const int* ai, const double* mv, const double* av,
                       const int* arow, const double* ar, double scale, double* out) {
    if (nnz > 0) EQ_LAUNCH(k_spgemm_num, blocks_for(nnz), nnz, seg, mi, ai, mv, av, arow, ar, scale, out);
}
void launch_spgemm_signed(long nnz, const int* seg, const int* sai, const double* av, const int* arow,
                          const double* ar, double scale, double* out) {
    if (nnz > 0) EQ_LAUNCH(k_spgemm_signed, blocks_for(nnz), nnz, seg, sai, av, arow, ar, scale, out);
}
void launch_jscale(long nnz, const int* row, double s, const double* R, const int* map, const double* v,
                   double* out) {
    if (nnz > 0) EQ_LAUNCH(k_jscale, blocks_for(nnz), nnz, row, s, R, map, v, out);
}
void launch_csr_spmv(int rows, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    if (rows > 0) EQ_LAUNCH(k_csr_spmv, blocks_for(rows), rows, rp, ci, v, x, y);
}
void launch_gradient(const double* u, const int* c1, const int* c2, double* out, int nif) {
    if (nif > 0) EQ_LAUNCH(k_gradient, blocks_for(nif), u, c1, c2, out, nif);
}
void launch_divergence(const double* flux, const int* facepos, const int* sfaces, double* out, int nc) {
    if (nc > 0) EQ_LAUNCH(k_divergence, blocks_for(nc), flux, facepos, sfaces, out, nc);
}
void launch_face_flags(const int* face_cells, int nf, uint8_t* interior) {
    if (nf > 0) EQ_LAUNCH(k_face_flags, blocks_for(nf), face_cells, nf, interior);
}
void launch_cell_boundary_flags(const int* face_cells, const int* facepos, const int* cell_faces, int nc,
                                uint8_t* boundary) {
    if (nc > 0) EQ_LAUNCH(k_cell_boundary, blocks_for(nc), face_cells, facepos, cell_faces, nc, boundary);
}
void launch_gather_vec(const double* src, const int* idx, int n, int dim, double* out) {
    if (n > 0) EQ_LAUNCH(k_gather_vec, blocks_for((long)n * dim), src, idx, n, dim, out);
}
void launch_gather_face_cells(const int* face_cells, const int* faces, int n, int which, int* out) {
    if (n > 0) EQ_LAUNCH(k_gather_face_cells, blocks_for(n), face_cells, faces, n, which, out);
}
void launch_vec_norm(const double* v, int n, int dim, double* out) {
    if (n > 0) EQ_LAUNCH(k_vec_norm, blocks_for(n), v, n, dim, out);
}
void launch_vec_dot(const double* a, const double* b, int n, int dim, double* out) {
    if (n > 0) EQ_LAUNCH(k_vec_dot, blocks_for(n), a, b, n, dim, out);
}
void launch_vec_ew(int op, const double* a, const double* b, int count, double* out) {
    launch_ew(op, a, b, out, count);
}
void launch_vec_scale(const double* v, const double* s, double sc, int n, int dim, int divide, double* out) {
    if (n > 0) EQ_LAUNCH(k_vec_scale, blocks_for((long)n * dim), v, s, sc, n, dim, divide, out);
}
void launch_unit_normals(const double* normals, const int* idx, int n, int dim, double* out) {
    if (n > 0) EQ_LAUNCH(k_unit_normals, blocks_for(n), normals, idx, n, dim, out);
}
void launch_vec_col(const double* v, int n, int dim, int col, double* out) {
    if (n > 0) EQ_LAUNCH(k_vec_col, blocks_for(n), v, n, dim, col, out);
}

double reduce(int op, const double* a, long n) {
    Device& dev = Device::get();
    const int nb = reduce_blocks(n);
    auto part = dev_alloc<double>(nb + 1);
    EQ_LAUNCH(k_reduce_partial, nb, op, a, (const double*)nullptr, n, part->p);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(BS), 0, S(), op, part->p, nb, part->p + nb);
    EQ_HIP(hipGetLastError());
    double* h = dev.host_scratch();
    EQ_HIP(hipMemcpyAsync(h, part->p + nb, sizeof(double), hipMemcpyDeviceToHost, S()));
    dev.sync();
    return h[0];
}

// Device address of `scratch` on each device, looked up once per device with that device
// current (hipGetSymbolAddress; each device has its own copy at its own address).
namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};
}  // namespace

double* noalloc_scratch() {
    int dev = 0;
    EQ_HIP(hipGetDevice(&dev));
    EQ_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device " << dev << " out of range for the no-alloc scratch");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        EQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        EQ_REQUIRE(p != nullptr, "hipGetSymbolAddress returned a null address for the no-alloc scratch");
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    return g_scratch_dev[dev];
}

void noalloc_scratch_init() { (void)noalloc_scratch(); }

// The kernels and launch shape of reduce(), with the partials in the device's scratch array
// instead of a pool block.  Every partial and the final slot are written before they are
// read, so nothing needs clearing between calls.
double reduce_noalloc(int op, const double* a, long n) {
    EQ_REQUIRE(n >= 0, "negative collection size " << n);
    const int nb = reduce_blocks(n);
    EQ_REQUIRE(nb + 1 <= NOALLOC_SCRATCH_DOUBLES, "reduction needs " << (nb + 1) << " scratch doubles, the no-alloc "
                                                  "scratch holds " << NOALLOC_SCRATCH_DOUBLES);
    double* part = noalloc_scratch();
    EQ_LAUNCH(k_reduce_partial, nb, op, a, (const double*)nullptr, n, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(BS), 0, S(), op, (const double*)part, nb, part + nb);
    EQ_HIP(hipGetLastError());
    Device& dev = Device::get();
    double* h = dev.host_scratch();   // pinned at device set-up
    EQ_HIP(hipMemcpyAsync(h, part + nb, sizeof(double), hipMemcpyDeviceToHost, S()));
    dev.sync();
    return h[0];
}

void launch_dot_to(const double* a, const double* b, long n, double* d_out) {
    const int nb = reduce_blocks(n);
    auto part = dev_alloc<double>(nb);
    EQ_LAUNCH(k_reduce_partial, nb, (int)RED_SUMSQ, a, b, n, part->p);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(BS), 0, S(), (int)RED_SUMSQ, part->p, nb, d_out);
    EQ_HIP(hipGetLastError());
}

int compact_flags(const uint8_t* flags, int n, int want, int* out) {
    if (n <= 0) return 0;
    const int nb = (int)((n + BS * CT - 1) / (BS * CT));
    auto counts = dev_alloc<int>(nb + 1);
    EQ_LAUNCH(k_compact_count, nb, flags, n, want, counts->p);
    hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(BS), 0, S(), counts->p, nb, counts->p + nb);
    EQ_HIP(hipGetLastError());
    EQ_LAUNCH(k_compact_write, nb, flags, n, want, counts->p, out);
    int total = 0;
    EQ_HIP(hipMemcpyAsync(&total, counts->p + nb, sizeof(int), hipMemcpyDeviceToHost, S()));
    Device::get().sync();
    return total;
}

}  // namespace eqhip
