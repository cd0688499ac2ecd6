// Fused nonlinear-heat engine (see heat_fused.hpp).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <tuple>
#include <type_traits>

#include "heat_fused.hpp"
#include "kernels.hpp"
#include "timing.hpp"

namespace eqhip {

namespace {
constexpr int BS = 256;   // cells per block (4 wave64)
inline hipStream_t S() { return Device::get().stream(); }

enum Slot { S_RHO = 0, S_RHO_OLD, S_ALPHA, S_OMEGA, S_RR, S_BB, S_BREAK, S_IT, S_FN, S_RN, S_BN, S_REF, S_ITT, NSLOT = 16 };
enum Vec { V_R = 0, V_RH, V_P, V_V, V_S, V_T, V_PH, V_SH, V_X, NVEC };
enum Act { A_INIT = 0, A_ALPHA, A_OMEGA, A_RHO, A_STORE };
enum Pc { PC_NONE = 0, PC_JACOBI = 1, PC_ILU = 2 };

enum Kind { K_HEAT = 0, K_TWOPHASE = 1 };

struct Desc {
    int n, nx, ny, nz;
    int xcd;                      // XCD-contiguous block order (same for all of a table)
    int kind;                     // K_HEAT | K_TWOPHASE
    int corey, pad;
    double rnx, rnxy;             // 1 / nx, 1 / (nx ny): cell -> (i, j, k) without integer division
    double dt;
    // heat: k(u) = k (1 + u^2), x / y face geometry factors, cell volume
    double k, cv, gx, gy, vol;
    double *u, *u0;
    // two-phase: transmissibility of the +x / +y / +z face of each cell, pore volume,
    // source rate, pressure, saturation and its value at the start of the step
    double mw_inv, mo_inv;        // 1 / watervisc, 1 / oilvisc
    const double *tx, *ty, *tz, *pv, *src;
    double *p, *sw, *sw0;
    double *F, *du, *r64;
    double *C, *W, *E, *S, *N, *B, *T;   // FP64 Jacobian stencil coefficients (B, T: 3D only)
    void *Cm, *Wm, *Em, *Sm, *Nm, *Bm, *Tm; // Krylov-precision copies (alias the FP64 ones in fp64 mode)
    void* dinv;                   // Jacobi: 1/C ; ILU: 1/pivot (black) or 1/C (red)
    void* v[NVEC];                // Krylov vectors
    double* sc;                   // NSLOT scalars
    int* act;                     // [0] Krylov active, [1] refinement active, [2] two-phase stage
    unsigned long long* ts;       // kernel timing only (else null): per block {start, end}
};

// What the last block of an instance does with the summed partials.
struct Fin {
    int act, nd, maxit, slot;
    double tol2;
};

template <int NT = BS>
__device__ inline void bred_store(double v, double* out) {
    __shared__ double lds[NT / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < NT / 64; ++i) s += lds[i];
        *out = s;
    }
    __syncthreads();
}
constexpr int FBS = 1024;   // finaliser block: one per instance, partials read 4+ per thread

// Launch shape.  Every cell kernel is instantiated for CPT (cells per thread) in
// {1, 2, 4, 8}: a block covers CPT*BS consecutive cells of one instance and thread t
// handles cells base + t, base + BS + t, ... (coalesced).  The shape of a batch
// membership is chosen by measurement (see Impl::shape_for).  Reductions are made
// shape-independent: a block writes one partial per group of BS cells (group m =
// cells [m BS, (m+1) BS) of the block, thread t holds cell t of each group) with the
// same tree for every shape, so results are bit-identical for any CPT and block order.
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2), so
// workgroup w runs on XCD w % 8.  With xcd set, the logical block (position in cell
// order) of workgroup w is chosen so that each XCD processes one contiguous range of
// blocks: stencil neighbours (rows above and below) then hit the same L2.  Bijective
// for any block count; partials are stored at the logical index.
constexpr int NXCD = 8;
__device__ inline int xcd_block(int w, int nblocks) {
    const int q = nblocks / NXCD, r = nblocks % NXCD, x = w % NXCD;
    return x * q + min(x, r) + w / NXCD;
}
#define LOCATE()                                                                                    \
    const int lb = D[0].xcd ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;         \
    if (D[0].ts && threadIdx.x == 0) D[0].ts[2 * lb] = wall_clock64();                              \
    const int2 bi = blk[lb];                                                                        \
    const Desc& d = D[bi.x];                                                                        \
    const int base = bi.y * (CPT * BS) + (int)threadIdx.x
// End stamp of a timed launch: the latest wave of the block (the kernel's execution
// span is min(start) .. max(end) over blocks, on the constant-rate wall clock).
#define KTIME_END()                                                                                 \
    if (D[0].ts && (threadIdx.x & 63) == 0) atomicMax(D[0].ts + 2 * lb + 1, wall_clock64())
#define FOR_CELLS(c) for (int c = base, c##_end = min(d.n, bi.y * (CPT * BS) + CPT * BS); c < c##_end; c += BS)
#define FOR_GROUPS(m, c) for (int m = 0, c = base; m < CPT; ++m, c += BS)

// One partial per group: v[m] is this thread's value for group m (0 when it has none).
// Within a wave the CPT vectors are reduced by recursive halving (reduce-scatter):
// at offset o the lanes with bit o clear keep the lower half of the groups and add
// their partner's (lane ^ o) values, the others keep the upper half; once one group
// per lane is left the usual butterfly finishes.  Every group is summed over the same
// lane pairs in the same order as a plain shuffle tree (addition is commutative), so
// the partial is bit-identical for every CPT.  Group g ends in lane g * 64 / CPT.
template <int CPT>
__device__ inline void group_store(const double (&v)[CPT], double* out) {
    __shared__ double lds[CPT][BS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double a[CPT];
    for (int m = 0; m < CPT; ++m) a[m] = v[m];
    int o = 32;
#pragma unroll
    for (int g = CPT; g > 1; g >>= 1, o >>= 1) {
        const bool up = (lane & o) != 0;
        const int h = g >> 1;
#pragma unroll
        for (int q = 0; q < h; ++q) {
            const double mine = up ? a[h + q] : a[q];
            const double send = up ? a[q] : a[h + q];
            a[q] = mine + __shfl_xor(send, o, 64);
        }
    }
    double x = a[0];
    for (; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    constexpr int SEG = 64 / CPT;
    if (lane % SEG == 0) lds[lane / SEG][w] = x;
    __syncthreads();
    if ((int)threadIdx.x < CPT) {
        double s = 0.0;
        for (int i = 0; i < BS / 64; ++i) s += lds[threadIdx.x][i];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// (i, j, k) of cell c = i + nx (j + ny k): double multiplies by 1/(nx ny) and 1/nx, each
// with one correction step (the product is within one ulp, so the floor is off by at
// most one).
__device__ inline void div_fix(int c, double rd, int dv, int& q, int& r) {
    int qq = (int)((double)c * rd);
    int rr = c - qq * dv;
    if (rr < 0) {
        --qq;
        rr += dv;
    } else if (rr >= dv) {
        ++qq;
        rr -= dv;
    }
    q = qq;
    r = rr;
}
__device__ inline void cell_ijk(const Desc& d, int c, int& i, int& j, int& k) {
    int rem = c;
    k = 0;
    if (d.nz > 1) div_fix(c, d.rnxy, d.nx * d.ny, k, rem);
    div_fix(rem, d.rnx, d.nx, j, i);
}

// Per-instance finalisation, run by one whole block: sums the instance's block
// partials in a fixed order (deterministic), then thread 0 applies the action to
// the instance's scalars.
__device__ void fin_body(const Desc& d, int2 r, const double* part, int nb, const Fin& f) {
    __shared__ double res[2];
    // partials were written by the previous kernel (visible at this kernel's start);
    // four independent sums per thread keep four loads in flight
    for (int q = 0; q < f.nd; ++q) {
        const double* __restrict__ pp = part + (long)q * nb;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int b = r.x + (int)threadIdx.x;
        for (; b + 3 * FBS < r.y; b += 4 * FBS) {
            a0 += pp[b];
            a1 += pp[b + FBS];
            a2 += pp[b + 2 * FBS];
            a3 += pp[b + 3 * FBS];
        }
        for (; b < r.y; b += FBS) a0 += pp[b];
        bred_store<FBS>((a0 + a1) + (a2 + a3), &res[q]);
    }
    if (threadIdx.x != 0) return;
    const int act = f.act, maxit = f.maxit, slot = f.slot;
    const double tol2 = f.tol2;
    double* sc = d.sc;
    switch (act) {
        case A_INIT: {   // start of a Krylov solve for instances under refinement
            if (!d.act[1]) {
                d.act[0] = 0;
                break;
            }
            sc[S_RHO] = res[0];
            sc[S_RR] = res[1];
            sc[S_BB] = res[1];
            sc[S_IT] = 0.0;
            sc[S_BREAK] = 0.0;
            sc[S_ALPHA] = sc[S_OMEGA] = sc[S_RHO_OLD] = 1.0;
            d.act[0] = res[1] > 0.0 && isfinite(res[1]);
            break;
        }
        case A_ALPHA:
            if (!d.act[0]) break;
            if (res[0] == 0.0 || !isfinite(res[0])) {
                sc[S_BREAK] = 1.0;
                d.act[0] = 0;
            } else sc[S_ALPHA] = sc[S_RHO] / res[0];
            break;
        case A_OMEGA:
            if (!d.act[0]) break;
            if (res[1] == 0.0 || !isfinite(res[1])) {
                sc[S_OMEGA] = 0.0;   // s already solves the system; finish with x += alpha ph
            } else sc[S_OMEGA] = res[0] / res[1];
            break;
        case A_RHO: {
            if (!d.act[0]) break;
            sc[S_RHO_OLD] = sc[S_RHO];
            sc[S_RHO] = res[0];
            sc[S_RR] = res[1];
            sc[S_IT] += 1.0;
            sc[S_ITT] += 1.0;
            const bool ok = isfinite(res[1]) && sc[S_OMEGA] != 0.0;
            if (!ok) sc[S_BREAK] = 2.0;
            d.act[0] = ok && res[1] > tol2 * sc[S_BB] && sc[S_IT] < maxit && res[0] != 0.0;
            break;
        }
        default:
            sc[slot] = res[0];
            break;
    }
}


__global__ __launch_bounds__(FBS) void k_fin(const Desc* __restrict__ D, const int2* __restrict__ rng,
                                            const double* __restrict__ part, int nb, Fin f) {
    fin_body(D[blockIdx.x], rng[blockIdx.x], part, nb, f);
}

// ---------------------------------------------------------------------------
// Residual + Jacobian assembly: one pass over u, u0; writes F, C/W/E/S/N
// (FP64 and, when mixed, FP32) and per-block partial ||F||^2.
// Two-phase mobilities (twophase.equelle computeWaterMob / computeOilMob):
// corey 1: krw = s, kro = 1 - s; corey 2: krw = s^2, kro = (1 - s)^2.
__device__ inline double mob_w(const Desc& d, double s) { return (d.corey == 2 ? s * s : s) * d.mw_inv; }
__device__ inline double mob_o(const Desc& d, double s) {
    const double so = 1.0 - s;
    return (d.corey == 2 ? so * so : so) * d.mo_inv;
}
__device__ inline double total_mob(const Desc& d, double s) { return mob_w(d, s) + mob_o(d, s); }
// fractional flow f = mw / (mw + mo) and its derivative
__device__ inline void frac_flow(const Desc& d, double s, double& f, double& df) {
    const double mw = mob_w(d, s), mo = mob_o(d, s);
    const double dmw = (d.corey == 2 ? 2.0 * s : 1.0) * d.mw_inv;
    const double dmo = -(d.corey == 2 ? 2.0 * (1.0 - s) : 1.0) * d.mo_inv;
    const double t = mw + mo;
    f = mw / t;
    df = (dmw * mo - mw * dmo) / (t * t);
}
// Total flux through the face first -> second (computeTotalFlux): trans * lambda_up *
// (p_first - p_second), lambda upwinded on the sign of -grad p.
__device__ inline double face_flux(const Desc& d, double trans, int first, int second) {
    const double ng = d.p[first] - d.p[second];
    const double lam = ng >= 0.0 ? total_mob(d, d.sw0[first]) : total_mob(d, d.sw0[second]);
    return (trans * lam) * ng;
}

// Two-phase pressure residual and Jacobian at cell c (computePressureResidual):
// F = sum_faces T lambda_up (p_c - p_nb) - source, J_c,nb = -T lambda_up.
__device__ inline double tp_pressure(const Desc& d, int c, int i, int j, int k, double* co) {
    const int nx = d.nx, nxy = d.nx * d.ny;
    const double pc = d.p[c];
    double F = -d.src[c];
    co[0] = 0.0;
    auto face = [&](int nb, double trans, bool c_first, double& cm) {
        const double pn = d.p[nb];
        const double ng = c_first ? pc - pn : pn - pc;   // p_first - p_second
        const double lam = ng >= 0.0 ? total_mob(d, d.sw0[c_first ? c : nb]) : total_mob(d, d.sw0[c_first ? nb : c]);
        const double tl = trans * lam;
        F += tl * (pc - pn);
        co[0] += tl;
        cm = -tl;
    };
    if (i > 0) face(c - 1, d.tx[c - 1], false, co[1]);
    if (i < d.nx - 1) face(c + 1, d.tx[c], true, co[2]);
    if (j > 0) face(c - nx, d.ty[c - nx], false, co[3]);
    if (j < d.ny - 1) face(c + nx, d.ty[c], true, co[4]);
    if (k > 0) face(c - nxy, d.tz[c - nxy], false, co[5]);
    if (k < d.nz - 1) face(c + nxy, d.tz[c], true, co[6]);
    return F;
}

// Two-phase transport residual and Jacobian at cell c (computeTransportResidual):
// F = sw - sw0 + dt/pv (sum_faces +-f_up flux - q), q = max(src,0) + min(src,0) f(sw),
// f upwinded on the sign of the total flux (fixed from the pressure solve).
__device__ inline double tp_transport(const Desc& d, int c, int i, int j, int k, double* co) {
    const int nx = d.nx, nxy = d.nx * d.ny;
    const double swc = d.sw[c];
    double fc, dfc;
    frac_flow(d, swc, fc, dfc);
    const double a = d.dt / d.pv[c];
    double div = 0.0, ddiag = 0.0;
    // face with `other`: c_first -> flux counts +, else -
    auto face = [&](int nb, double trans, bool c_first, double& cm) {
        const double flux = c_first ? face_flux(d, trans, c, nb) : face_flux(d, trans, nb, c);
        const double sgn = c_first ? 1.0 : -1.0;
        const bool up_c = (flux >= 0.0) == c_first;   // the first cell is upstream when flux >= 0
        cm = 0.0;
        if (up_c) {
            div += sgn * fc * flux;
            ddiag += sgn * dfc * flux;
        } else {
            double fn, dfn;
            frac_flow(d, d.sw[nb], fn, dfn);
            div += sgn * fn * flux;
            cm = a * sgn * dfn * flux;
        }
    };
    if (i > 0) face(c - 1, d.tx[c - 1], false, co[1]);
    if (i < d.nx - 1) face(c + 1, d.tx[c], true, co[2]);
    if (j > 0) face(c - nx, d.ty[c - nx], false, co[3]);
    if (j < d.ny - 1) face(c + nx, d.ty[c], true, co[4]);
    if (k > 0) face(c - nxy, d.tz[c - nxy], false, co[5]);
    if (k < d.nz - 1) face(c + nxy, d.tz[c], true, co[6]);
    const double src = d.src[c];
    const double insrc = src > 0.0 ? src : 0.0, outsrc = src < 0.0 ? src : 0.0;
    const double q = insrc + outsrc * fc;
    co[0] = 1.0 + a * (ddiag - outsrc * dfc);
    return (swc - d.sw0[c]) + a * (div - q);
}

template <bool MIXED, int CPT>
__global__ __launch_bounds__(BS) void k_assemble(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                 double* __restrict__ part) {
    LOCATE();
    double f2[CPT] = {};
    if (d.kind == K_TWOPHASE) {
        const int stage = d.act[2];
        FOR_GROUPS(m, c) {
            if (c >= d.n) continue;
            int i, j, k;
            cell_ijk(d, c, i, j, k);
            double co[7] = {0, 0, 0, 0, 0, 0, 0};
            const double F = stage == 0 ? tp_pressure(d, c, i, j, k, co) : tp_transport(d, c, i, j, k, co);
            d.F[c] = F;
            double* dst[7] = {d.C, d.W, d.E, d.S, d.N, d.B, d.T};
            for (int q = 0; q < (d.nz > 1 ? 7 : 5); ++q) dst[q][c] = co[q];
            if (MIXED) {
                float* dm[7] = {(float*)d.Cm, (float*)d.Wm, (float*)d.Em, (float*)d.Sm,
                                (float*)d.Nm, (float*)d.Bm, (float*)d.Tm};
                for (int q = 0; q < (d.nz > 1 ? 7 : 5); ++q) dm[q][c] = (float)co[q];
            }
            f2[m] = F * F;
        }
        group_store<CPT>(f2, part + lb * CPT);
        KTIME_END();
        return;
    }
    const int nx = d.nx, ny = d.ny;
    const double k = d.k, coef = d.dt / (d.cv * d.vol), gx = d.gx, gy = d.gy;
    const double* __restrict__ u = d.u;
    FOR_GROUPS(m, c) {
        if (c >= d.n) continue;
        int i, j, kk;
        cell_ijk(d, c, i, j, kk);
        const double uc = u[c];
        const double kc = k * (1.0 + uc * uc), dkc = 2.0 * k * uc;
        double div = 0.0, diag = 1.0, cw = 0.0, ce = 0.0, cs = 0.0, cn = 0.0;
        // faces in ascending face-id order: west (x-), east (x+), south (y-), north (y+)
        auto face = [&](double um, double g, double& cm) {
            const double km = k * (1.0 + um * um);
            const double kf = 0.5 * (kc + km);
            const double diff = uc - um;
            div += g * kf * diff;
            diag += coef * g * (0.5 * dkc * diff + kf);
            cm = coef * g * (0.5 * (2.0 * k * um) * diff - kf);
        };
        if (i > 0) face(u[c - 1], gx, cw);
        if (i < nx - 1) face(u[c + 1], gx, ce);
        if (j > 0) face(u[c - nx], gy, cs);
        if (j < ny - 1) face(u[c + nx], gy, cn);
        const double F = (uc - d.u0[c]) + coef * div;
        d.F[c] = F;
        d.C[c] = diag;
        d.W[c] = cw;
        d.E[c] = ce;
        d.S[c] = cs;
        d.N[c] = cn;
        if (MIXED) {
            ((float*)d.Cm)[c] = (float)diag;
            ((float*)d.Wm)[c] = (float)cw;
            ((float*)d.Em)[c] = (float)ce;
            ((float*)d.Sm)[c] = (float)cs;
            ((float*)d.Nm)[c] = (float)cn;
        }
        f2[m] = F * F;
    }
    group_store<CPT>(f2, part + lb * CPT);
    KTIME_END();
}

template <class T>
__device__ inline T stencil_apply(const Desc& d, const T* __restrict__ x, int c, int i, int j, int k) {
    const int nx = d.nx;
    T acc = ((const T*)d.Cm)[c] * x[c];
    if (i > 0) acc += ((const T*)d.Wm)[c] * x[c - 1];
    if (i < nx - 1) acc += ((const T*)d.Em)[c] * x[c + 1];
    if (j > 0) acc += ((const T*)d.Sm)[c] * x[c - nx];
    if (j < d.ny - 1) acc += ((const T*)d.Nm)[c] * x[c + nx];
    if (d.nz > 1) {
        const int nxy = nx * d.ny;
        if (k > 0) acc += ((const T*)d.Bm)[c] * x[c - nxy];
        if (k < d.nz - 1) acc += ((const T*)d.Tm)[c] * x[c + nxy];
    }
    return acc;
}

// y = A x with fused partial dot products: dot0 = V[a0].y, dot1 = V[a1].y (a1 < 0: y.y).
// Skipped for inactive instances (partials 0).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_spmv_dot(const Desc* __restrict__ D, const int2* __restrict__ blk, int ix,
                                                 int iy, int nd, int a0, int a1, double* __restrict__ part, int ng) {
    LOCATE();
    double s0[CPT] = {}, s1[CPT] = {};
    if (d.act[0]) {
        const T* __restrict__ x = (const T*)d.v[ix];
        T* __restrict__ y = (T*)d.v[iy];
        const T* __restrict__ va = (const T*)d.v[a0];
        const T* __restrict__ vb = a1 >= 0 ? (const T*)d.v[a1] : nullptr;
        FOR_GROUPS(m, c) {
            if (c >= d.n) continue;
            int i, j, k;
            cell_ijk(d, c, i, j, k);
            const T yc = stencil_apply<T>(d, x, c, i, j, k);
            y[c] = yc;
            s0[m] = (double)va[c] * (double)yc;
            if (nd > 1) s1[m] = (vb ? (double)vb[c] : (double)yc) * (double)yc;
        }
    }
    group_store<CPT>(s0, part + lb * CPT);
    if (nd > 1) group_store<CPT>(s1, part + ng + lb * CPT);
    KTIME_END();
}

// FP64 defect for the instances under refinement: r64 = F - J du.
template <int CPT>
__global__ __launch_bounds__(BS) void k_defect(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                               double* __restrict__ part) {
    LOCATE();
    double r2[CPT] = {};
    if (d.act[1]) {
        const int nx = d.nx;
        const double* __restrict__ x = d.du;
        FOR_GROUPS(m, c) {
            if (c >= d.n) continue;
            int i, j, k;
            cell_ijk(d, c, i, j, k);
            double acc = d.C[c] * x[c];
            if (i > 0) acc += d.W[c] * x[c - 1];
            if (i < nx - 1) acc += d.E[c] * x[c + 1];
            if (j > 0) acc += d.S[c] * x[c - nx];
            if (j < d.ny - 1) acc += d.N[c] * x[c + nx];
            if (d.nz > 1) {
                const int nxy = nx * d.ny;
                if (k > 0) acc += d.B[c] * x[c - nxy];
                if (k < d.nz - 1) acc += d.T[c] * x[c + nxy];
            }
            const double r = d.F[c] - acc;
            d.r64[c] = r;
            r2[m] = r * r;
        }
    }
    group_store<CPT>(r2, part + lb * CPT);
    KTIME_END();
}

// Preconditioner set-up: Jacobi 1/C, or red-black ILU(0) pivots.  Red = (i+j+k) even
// cells come first in the ordering; red rows have no L part, a black row's L part is its
// red neighbours and its pivot is C_b - sum_m A_bm A_mb / C_m (fill-in dropped).
// dinv holds 1/C for red cells and 1/pivot for black cells.
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_precond_setup(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                      int mode) {
    LOCATE();
    const T* Cm = (const T*)d.Cm;
    const T *Wm = (const T*)d.Wm, *Em = (const T*)d.Em, *Sm = (const T*)d.Sm, *Nm = (const T*)d.Nm;
    const T *Bm = (const T*)d.Bm, *Tm = (const T*)d.Tm;
    const int nx = d.nx, nxy = d.nx * d.ny;
    FOR_CELLS(c) {
        T piv = Cm[c];
        if (mode == PC_ILU) {
            int i, j, k;
            cell_ijk(d, c, i, j, k);
            if ((i + j + k) & 1) {
                if (i > 0) piv -= Wm[c] * Em[c - 1] / Cm[c - 1];
                if (i < nx - 1) piv -= Em[c] * Wm[c + 1] / Cm[c + 1];
                if (j > 0) piv -= Sm[c] * Nm[c - nx] / Cm[c - nx];
                if (j < d.ny - 1) piv -= Nm[c] * Sm[c + nx] / Cm[c + nx];
                if (d.nz > 1) {
                    if (k > 0) piv -= Bm[c] * Tm[c - nxy] / Cm[c - nxy];
                    if (k < d.nz - 1) piv -= Tm[c] * Bm[c + nxy] / Cm[c + nxy];
                }
            }
        }
        ((T*)d.dinv)[c] = piv != (T)0 ? (T)1 / piv : (T)1;
    }
}

// z = M^{-1} r for red-black ILU(0) in two passes: a black cell solves
// z_b = (r_b - sum_m A_bm r_m / C_m) / piv_b, then a red cell solves
// z_r = (r_r - sum_b A_rb z_b) / C_r.  1/C_m of a red neighbour is its dinv, so only
// multiplications remain.
template <class T>
__device__ inline T black_value(const Desc& d, const T* __restrict__ r, int b, int i, int j, int k) {
    const int nx = d.nx, ny = d.ny;
    const T* dv = (const T*)d.dinv;
    const T *Wm = (const T*)d.Wm, *Em = (const T*)d.Em, *Sm = (const T*)d.Sm, *Nm = (const T*)d.Nm;
    T s = r[b];
    if (i > 0) s -= Wm[b] * (r[b - 1] * dv[b - 1]);
    if (i < nx - 1) s -= Em[b] * (r[b + 1] * dv[b + 1]);
    if (j > 0) s -= Sm[b] * (r[b - nx] * dv[b - nx]);
    if (j < ny - 1) s -= Nm[b] * (r[b + nx] * dv[b + nx]);
    if (d.nz > 1) {
        const int nxy = nx * ny;
        if (k > 0) s -= ((const T*)d.Bm)[b] * (r[b - nxy] * dv[b - nxy]);
        if (k < d.nz - 1) s -= ((const T*)d.Tm)[b] * (r[b + nxy] * dv[b + nxy]);
    }
    return s * dv[b];
}
// Pass 1 of the apply: z of the black cells (their L part only touches red cells,
// whose z = r / C before the U sweep).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_ilu_black(const Desc* __restrict__ D, const int2* __restrict__ blk, int ir,
                                                  int iz) {
    LOCATE();
    if (d.act[0]) {
        const T* __restrict__ r = (const T*)d.v[ir];
        T* __restrict__ zo = (T*)d.v[iz];
        FOR_CELLS(c) {
            int i, j, k;
            cell_ijk(d, c, i, j, k);
            if ((i + j + k) & 1) zo[c] = black_value<T>(d, r, c, i, j, k);
        }
    }
    KTIME_END();
}
// Pass 2: z of the red cells from the black z of pass 1 (the U part of a red row).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_ilu_red(const Desc* __restrict__ D, const int2* __restrict__ blk, int ir,
                                                int iz) {
    LOCATE();
    if (d.act[0]) {
        const int nx = d.nx, ny = d.ny;
        const T* __restrict__ r = (const T*)d.v[ir];
        T* __restrict__ z = (T*)d.v[iz];
        FOR_CELLS(c) {
            int i, j, k;
            cell_ijk(d, c, i, j, k);
            if ((i + j + k) & 1) continue;
            T s = r[c];
            if (i > 0) s -= ((const T*)d.Wm)[c] * z[c - 1];
            if (i < nx - 1) s -= ((const T*)d.Em)[c] * z[c + 1];
            if (j > 0) s -= ((const T*)d.Sm)[c] * z[c - nx];
            if (j < ny - 1) s -= ((const T*)d.Nm)[c] * z[c + nx];
            if (d.nz > 1) {
                const int nxy = nx * ny;
                if (k > 0) s -= ((const T*)d.Bm)[c] * z[c - nxy];
                if (k < d.nz - 1) s -= ((const T*)d.Tm)[c] * z[c + nxy];
            }
            z[c] = s * ((const T*)d.dinv)[c];
        }
    }
    KTIME_END();
}

// Krylov start: r = b (already in r), rh = r, x = 0, p = v = 0; partial dots (rh.r, r.r).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_kinit(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                              double* __restrict__ part, int ng) {
    LOCATE();
    double rr[CPT] = {};
    if (d.act[1]) {
        FOR_GROUPS(m, c) {
            if (c >= d.n) continue;
            const T rc = ((T*)d.v[V_R])[c];
            ((T*)d.v[V_RH])[c] = rc;
            ((T*)d.v[V_X])[c] = 0;
            ((T*)d.v[V_P])[c] = 0;
            ((T*)d.v[V_V])[c] = 0;
            rr[m] = (double)rc * (double)rc;
        }
    }
    group_store<CPT>(rr, part + lb * CPT);
    group_store<CPT>(rr, part + ng + lb * CPT);
}

// p = r + beta (p - omega v) (first iteration: p = r); Jacobi / none: ph = M p
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_b1(const Desc* __restrict__ D, const int2* __restrict__ blk, int pc) {
    LOCATE();
    if (!d.act[0]) return;
    const double* sc = d.sc;
    T* p = (T*)d.v[V_P];
    const bool first = sc[S_IT] == 0.0;
    const T beta = first ? (T)0 : (T)((sc[S_RHO] / sc[S_RHO_OLD]) * (sc[S_ALPHA] / sc[S_OMEGA]));
    const T omega = (T)sc[S_OMEGA];
    FOR_CELLS(c) {
        const T r = ((const T*)d.v[V_R])[c];
        const T pv = first ? r : r + beta * (p[c] - omega * ((const T*)d.v[V_V])[c]);
        p[c] = pv;
        if (pc != PC_ILU) ((T*)d.v[V_PH])[c] = pc == PC_JACOBI ? ((const T*)d.dinv)[c] * pv : pv;
    }
}
// s = r - alpha v ; Jacobi / none: sh = M s
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_b2(const Desc* __restrict__ D, const int2* __restrict__ blk, int pc) {
    LOCATE();
    if (!d.act[0]) return;
    const T alpha = (T)d.sc[S_ALPHA];
    FOR_CELLS(c) {
        const T s = ((const T*)d.v[V_R])[c] - alpha * ((const T*)d.v[V_V])[c];
        ((T*)d.v[V_S])[c] = s;
        if (pc != PC_ILU) ((T*)d.v[V_SH])[c] = pc == PC_JACOBI ? ((const T*)d.dinv)[c] * s : s;
    }
}
// x += alpha ph + omega sh ; r = s - omega t ; partial dots (rh.r, r.r)
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_b3(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                           double* __restrict__ part, int ng) {
    LOCATE();
    double a0[CPT] = {}, a1[CPT] = {};
    if (d.act[0]) {
        const T alpha = (T)d.sc[S_ALPHA], omega = (T)d.sc[S_OMEGA];
        T* x = (T*)d.v[V_X];
        FOR_GROUPS(m, c) {
            if (c >= d.n) continue;
            x[c] += alpha * ((const T*)d.v[V_PH])[c] + omega * ((const T*)d.v[V_SH])[c];
            const T r = ((const T*)d.v[V_S])[c] - omega * ((const T*)d.v[V_T])[c];
            ((T*)d.v[V_R])[c] = r;
            a0[m] = (double)((const T*)d.v[V_RH])[c] * (double)r;
            a1[m] = (double)r * (double)r;
        }
    }
    group_store<CPT>(a0, part + lb * CPT);
    group_store<CPT>(a1, part + ng + lb * CPT);
}

__global__ void k_count_active(const Desc* __restrict__ D, int ninst, int which, int* __restrict__ out) {
    __shared__ int lds[BS];
    int s = 0;
    for (int i = threadIdx.x; i < ninst; i += BS) s += D[i].act[which] ? 1 : 0;
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = lds[0];
}

// refinement bookkeeping (one thread per instance):
//   mode 0: start  -> act[1] = ||F|| > 0; S_BN = S_RN = ||F||; counters reset
//   mode 1: after a defect: S_RN = sqrt(S_FN); act[1] = S_RN > tol*S_BN && S_REF < max_refine
//           and the defect decreased (stagnation stops the refinement)
__global__ void k_refine_flags(const Desc* __restrict__ D, int ninst, int mode, double tol, int max_refine) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    const Desc& d = D[i];
    double* sc = d.sc;
    if (mode == 0) {
        sc[S_BN] = sqrt(sc[S_FN]);
        sc[S_RN] = sc[S_BN];
        sc[S_REF] = 0.0;
        sc[S_ITT] = 0.0;
        d.act[1] = sc[S_BN] > 0.0 && isfinite(sc[S_BN]);
    } else {
        if (!d.act[1]) return;
        const double rn = sqrt(sc[S_FN]);   // defect norm stored by k_fin(A_STORE, S_FN)
        const bool progress = rn < sc[S_RN];
        sc[S_RN] = rn;
        sc[S_REF] += 1.0;
        d.act[1] = isfinite(rn) && progress && rn > tol * sc[S_BN] && sc[S_REF] < max_refine;
    }
}

// r (Krylov precision) = r64 / ||r64||  (mixed) or = F (fp64 mode: r64 == F)
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_load_rhs(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                 int scaled) {
    LOCATE();
    if (!d.act[1]) return;
    const double scale = scaled ? 1.0 / d.sc[S_RN] : 1.0;
    FOR_CELLS(c)((T*)d.v[V_R])[c] = (T)(d.r64[c] * scale);
}
// du += ||r64|| x  (mixed) or du = x (fp64)
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_accum(const Desc* __restrict__ D, const int2* __restrict__ blk, int scaled) {
    LOCATE();
    if (!d.act[1]) return;
    const double rn = d.sc[S_RN];
    FOR_CELLS(c) {
        const double x = (double)((const T*)d.v[V_X])[c];
        if (scaled) d.du[c] += rn * x;
        else d.du[c] = x;
    }
}
template <int CPT>
__global__ __launch_bounds__(BS) void k_zero_du(const Desc* __restrict__ D, const int2* __restrict__ blk) {
    LOCATE();
    FOR_CELLS(c) {
        d.du[c] = 0.0;
        d.r64[c] = d.F[c];
    }
}
// u -= du for the Newton unknown: heat u; two-phase p (stage 0) or sw (stage 1)
template <int CPT>
__global__ __launch_bounds__(BS) void k_newton_update(const Desc* __restrict__ D, const int2* __restrict__ blk) {
    LOCATE();
    double* x = d.kind == K_TWOPHASE ? (d.act[2] == 0 ? d.p : d.sw) : d.u;
    FOR_CELLS(c) x[c] -= d.du[c];
}
__global__ void k_gather_slot(const Desc* __restrict__ D, int ninst, int slot, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ninst) out[i] = D[i].sc[slot];
}
#undef KTIME_END
#undef FOR_GROUPS
#undef FOR_CELLS
#undef LOCATE

// ---------------------------------------------------------------------------
struct Inst {
    int id = 0;
    int kind = K_HEAT;
    int nx = 0, ny = 0, nz = 1, n = 0;
    double k = 0, cv = 1, dx = 1, dy = 1, dt = 0;
    // two-phase
    double mw_inv = 0, mo_inv = 0;
    int corey = 1;
    int stage = 0;                 // 0 pressure, 1 transport
    double r0_stage = 0;           // ||F|| at the start of the current Newton solve
    DevPtr<double> tp;             // tx, ty, tz, pv, src (n each)
    DevPtr<double> p, sw, sw0;
    DevPtr<double> u, u0, F, du, r64, C, W, E, S, N, B, T, sc;
    DevPtr<char> coef32, dinv, vecs;
    DevPtr<int> act;
    enum State { IDLE, START, SOLVE, DONE } state = IDLE;
    StepStats st;
    int ncoef() const { return nz > 1 ? 7 : 5; }
};

// Launch shape of a table: cells per thread and XCD-contiguous block order.
struct Shape {
    int cpt = 4;
    int xcd = 1;
    bool operator<(const Shape& o) const { return cpt != o.cpt ? cpt < o.cpt : xcd < o.xcd; }
};
constexpr int CPT_MAX = 8;

// Calls f(std::integral_constant<int, CPT>) for the table's cells-per-thread.
template <class F>
void by_cpt(int cpt, F&& f) {
    switch (cpt) {
        case 1: f(std::integral_constant<int, 1>()); break;
        case 2: f(std::integral_constant<int, 2>()); break;
        case 4: f(std::integral_constant<int, 4>()); break;
        case 8: f(std::integral_constant<int, 8>()); break;
        default: throw std::runtime_error("unsupported cells per thread " + std::to_string(cpt));
    }
}

struct Table {
    std::vector<Inst*> inst;
    DevPtr<Desc> desc;
    DevPtr<int2> blk, rng;   // rng: each instance's partials (groups of BS cells)
    int nb = 0;              // blocks
    int ng = 0;              // groups = nb * cpt
    Shape shape;
    DevPtr<double> part;
    uint64_t version = 0;
    int solves = 0;   // Krylov solves run on this membership (graphs are captured from the second)
    // captured BiCGStab iterations, keyed by (tolerance, max iterations)
    std::map<std::pair<double, int>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
    ~Table() {
        for (auto& g : graphs) {
            (void)hipGraphExecDestroy(g.second.second);
            (void)hipGraphDestroy(g.second.first);
        }
    }
};
using TableP = std::shared_ptr<Table>;
}  // namespace

struct HeatBatch::Impl {
    FusedConfig cfg;
    bool mixed = true;
    int pc = PC_ILU;
    std::map<int, std::unique_ptr<Inst>> insts;
    int next_id = 1;
    uint64_t version = 1;   // bumped whenever a descriptor field changes
    std::map<std::vector<int>, TableP> tcache;
    bool use_graph = true;
    int graph_iters = 2;    // BiCGStab iterations per graph launch
    static constexpr int NRING = 4;
    DevPtr<int> count;
    int* host_count = nullptr;
    hipEvent_t ev[NRING] = {};

    size_t ks() const { return mixed ? sizeof(float) : sizeof(double); }

    Desc desc_of(Inst& I) const {
        Desc d{};
        d.n = I.n;
        d.nx = I.nx;
        d.ny = I.ny;
        d.nz = I.nz;
        d.kind = I.kind;
        d.corey = I.corey;
        d.rnx = 1.0 / I.nx;
        d.rnxy = 1.0 / ((double)I.nx * I.ny);
        d.dt = I.dt;
        if (I.kind == K_HEAT) {
            d.k = I.k;
            d.cv = I.cv;
            d.gx = I.dy / I.dx;   // x-faces: area dy over centroid distance dx
            d.gy = I.dx / I.dy;
            d.vol = I.dx * I.dy;
            d.u = I.u->p;
            d.u0 = I.u0->p;
        } else {
            d.mw_inv = I.mw_inv;
            d.mo_inv = I.mo_inv;
            const double* t = I.tp->p;
            d.tx = t;
            d.ty = t + I.n;
            d.tz = t + 2L * I.n;
            d.pv = t + 3L * I.n;
            d.src = t + 4L * I.n;
            d.p = I.p->p;
            d.sw = I.sw->p;
            d.sw0 = I.sw0->p;
        }
        d.F = I.F->p;
        d.du = I.du->p;
        d.r64 = I.r64->p;
        d.C = I.C->p;
        d.W = I.W->p;
        d.E = I.E->p;
        d.S = I.S->p;
        d.N = I.N->p;
        d.B = I.B ? I.B->p : nullptr;
        d.T = I.T ? I.T->p : nullptr;
        if (mixed) {
            float* f = (float*)I.coef32->p;
            d.Cm = f;
            d.Wm = f + I.n;
            d.Em = f + 2L * I.n;
            d.Sm = f + 3L * I.n;
            d.Nm = f + 4L * I.n;
            d.Bm = I.nz > 1 ? f + 5L * I.n : nullptr;
            d.Tm = I.nz > 1 ? f + 6L * I.n : nullptr;
        } else {
            d.Cm = d.C;
            d.Wm = d.W;
            d.Em = d.E;
            d.Sm = d.S;
            d.Nm = d.N;
            d.Bm = d.B;
            d.Tm = d.T;
        }
        d.dinv = I.dinv->p;
        for (int v = 0; v < NVEC; ++v) d.v[v] = I.vecs->p + (size_t)v * I.n * ks();
        d.sc = I.sc->p;
        d.act = I.act->p;
        return d;
    }

    TableP build_table(const std::vector<Inst*>& list, Shape sh, unsigned long long* ts = nullptr) {
        TableP tp = std::make_shared<Table>();
        Table& t = *tp;
        t.version = version;
        t.inst = list;
        t.shape = sh;
        const int cpb = sh.cpt * BS;
        std::vector<Desc> ds;
        std::vector<int2> blk, rng;
        for (size_t q = 0; q < list.size(); ++q) {
            Desc d = desc_of(*list[q]);
            d.xcd = sh.xcd;
            d.ts = ts;
            ds.push_back(d);
            const int nbk = (list[q]->n + cpb - 1) / cpb;
            rng.push_back(make_int2((int)blk.size() * sh.cpt, ((int)blk.size() + nbk) * sh.cpt));
            for (int b = 0; b < nbk; ++b) blk.push_back(make_int2((int)q, b));
        }
        t.nb = (int)blk.size();
        t.ng = t.nb * sh.cpt;
        t.desc = dev_from_host(ds.data(), ds.size());
        t.blk = dev_from_host(blk.data(), blk.size());
        t.rng = dev_from_host(rng.data(), rng.size());
        t.part = dev_alloc<double>(2 * (size_t)t.ng + 2);
        return tp;
    }
    TableP table(const std::vector<Inst*>& list) {
        std::vector<int> key;
        for (Inst* I : list) key.push_back(I->id);
        auto it = tcache.find(key);
        if (it != tcache.end() && it->second->version == version) return it->second;
        if (tcache.size() > 64) tcache.clear();
        TableP tp = build_table(list, shape_for(list));
        tcache[key] = tp;
        return tp;
    }

    // ---- launch-shape tuner -------------------------------------------------------
    // A membership's signature is (log2 total cells, log2 instance count, log2 largest
    // instance).  The first time a signature is seen, each shape (cells per thread 1, 2,
    // 4, 8 x XCD order on/off) is timed on that membership with HIP events around the
    // Krylov iteration's stencil kernels (SpMV with its dots, and the preconditioner's
    // passes), and the fastest is kept for the signature.  Results do not depend on the
    // shape (shape-independent reductions), so tuning changes time only.
    std::map<std::tuple<int, int, int>, Shape> tuned;
    Shape last_shape;
    static int ilog2(long v) {
        int r = 0;
        while (v > 1) v >>= 1, ++r;
        return r;
    }
    Shape shape_for(const std::vector<Inst*>& list) {
        if (cfg.cpt > 0) {
            Shape sh;
            sh.cpt = cfg.cpt;
            sh.xcd = cfg.xcd;
            return last_shape = sh;
        }
        long tot = 0, mx = 0;
        for (Inst* I : list) tot += I->n, mx = std::max<long>(mx, I->n);
        const auto sig = std::make_tuple(ilog2(tot), ilog2((long)list.size()), ilog2(mx));
        auto it = tuned.find(sig);
        if (it != tuned.end()) return last_shape = it->second;
        Shape best;
        if (cfg.tune) {
            double best_ms = 1e300;
            for (int cpt : {1, 2, 4, 8})
                for (int xcd : {0, 1}) {
                    Shape sh;
                    sh.cpt = cpt;
                    sh.xcd = xcd;
                    const double ms = time_shape(list, sh);
                    if (ms < best_ms) best_ms = ms, best = sh;
                }
        }
        tuned[sig] = best;
        return last_shape = best;
    }
    double time_shape(const std::vector<Inst*>& list, Shape sh) {
        TableP tp = build_table(list, sh);
        Table& t = *tp;
        // every instance active so the kernels do their full work (a solve resets the flags)
        const int one[2] = {1, 1};
        for (Inst* I : list) EQ_HIP(hipMemcpyAsync(I->act->p, one, sizeof(one), hipMemcpyHostToDevice, S()));
        // a Krylov round (SpMV with dots, preconditioner) and a Newton round; a Newton
        // iteration runs about eight Krylov rounds
        auto krylov_round = [&] {
            if (mixed) spmv_dot<float>(t, V_PH, V_V, 1, (int)V_RH, -1, F(A_STORE, 1, 0, 0, NSLOT - 1), false);
            else spmv_dot<double>(t, V_PH, V_V, 1, (int)V_RH, -1, F(A_STORE, 1, 0, 0, NSLOT - 1), false);
            if (mixed) precond<float>(t, V_P, V_PH);
            else precond<double>(t, V_P, V_PH);
        };
        // (the FP64 defect stands for the Newton-level stencil kernels; it only writes
        // the refinement residual, which every solve re-initialises)
        auto newton_round = [&] {
            by_cpt(t.shape.cpt, [&](auto cpt_) {
                constexpr int CP = decltype(cpt_)::value;
                hipLaunchKernelGGL((k_defect<CP>), dim3((unsigned)t.nb), dim3(BS), 0, S(), t.desc->p, t.blk->p,
                                   t.part->p);
                EQ_HIP(hipGetLastError());
            });
        };
        hipEvent_t e0, e1;
        EQ_HIP(hipEventCreate(&e0));
        EQ_HIP(hipEventCreate(&e1));
        auto timed = [&](auto&& fn, int reps) {
            fn();   // warm
            EQ_HIP(hipEventRecord(e0, S()));
            for (int r = 0; r < reps; ++r) fn();
            EQ_HIP(hipEventRecord(e1, S()));
            EQ_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            EQ_HIP(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / reps;
        };
        const double kr = timed(krylov_round, 3);
        const double nr = timed(newton_round, 2);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return 8.0 * kr + nr;
    }

    int count_active(const Table& t, int which) {
        hipLaunchKernelGGL(k_count_active, dim3(1), dim3(BS), 0, S(), t.desc->p, (int)t.inst.size(), which, count->p);
        EQ_HIP(hipGetLastError());
        EQ_HIP(hipMemcpyAsync(host_count, count->p, sizeof(int), hipMemcpyDeviceToHost, S()));
        Device::get().sync();
        return *host_count;
    }
    // Finalisation of a reduction: one block per instance (kernel boundaries make the
    // partials of all blocks visible; no device-wide fences inside the kernels).
    void fin(const Table& t, const Fin& f) {
        hipLaunchKernelGGL(k_fin, dim3((unsigned)t.inst.size()), dim3(FBS), 0, S(), t.desc->p, t.rng->p, t.part->p,
                           t.ng, f);
        EQ_HIP(hipGetLastError());
    }
    static Fin F(int act, int nd, double tol2 = 0, int maxit = 0, int slot = 0) {
        Fin f;
        f.act = act;
        f.nd = nd;
        f.tol2 = tol2;
        f.maxit = maxit;
        f.slot = slot;
        return f;
    }
// Launch over every block of the table; the kernel is named with CP, the table's
// cells per thread, e.g. HL((k_b1<T, CP>), pc).
#define HL(kernel, ...)                                                                                    \
    by_cpt(t.shape.cpt, [&](auto cpt_) {                                                                   \
        constexpr int CP = decltype(cpt_)::value;                                                          \
        hipLaunchKernelGGL(kernel, dim3((unsigned)t.nb), dim3(BS), 0, S(), t.desc->p, t.blk->p, ##__VA_ARGS__); \
        EQ_HIP(hipGetLastError());                                                                         \
    })

    void assemble(const Table& t) {
        {
            TimedLaunch tl("assemble");
            if (mixed) HL((k_assemble<true, CP>), t.part->p);
            else HL((k_assemble<false, CP>), t.part->p);
        }
        fin(t, F(A_STORE, 1, 0, 0, S_FN));
    }

    template <class T>
    void spmv_dot(const Table& t, int ix, int iy, int nd, int a0, int a1, const Fin& f, bool with_fin = true) {
        {
            TimedLaunch tl(sizeof(T) == 4 ? "stencil_spmv_fp32" : "stencil_spmv_fp64");
            HL((k_spmv_dot<T, CP>), ix, iy, nd, a0, a1, t.part->p, t.ng);
        }
        if (with_fin) fin(t, f);
    }
    template <class T>
    void precond(const Table& t, int ir, int iz) {
        if (pc == PC_ILU) {
            HL((k_ilu_black<T, CP>), ir, iz);
            HL((k_ilu_red<T, CP>), ir, iz);
        }
    }

    // BiCGStab for the instances whose refinement flag is set; rhs in V_R.
    template <class T>
    void bicg_iteration(const Table& t, double tol, int maxit) {
        HL((k_b1<T, CP>), pc);
        precond<T>(t, V_P, V_PH);
        spmv_dot<T>(t, V_PH, V_V, 1, (int)V_RH, -1, F(A_ALPHA, 1));   // v = A ph; alpha = rho / rh.v
        HL((k_b2<T, CP>), pc);
        precond<T>(t, V_S, V_SH);
        spmv_dot<T>(t, V_SH, V_T, 2, (int)V_S, -1, F(A_OMEGA, 2));    // t = A sh; omega = t.s / t.t
        HL((k_b3<T, CP>), t.part->p, t.ng);
        fin(t, F(A_RHO, 2, tol * tol, maxit));
    }

    // Device -> host count of active instances without stalling the queue: the count
    // of launch k is read after launch k+1 has been queued (pinned ring + events).
    void count_async(const Table& t, int which, int slot) {
        hipLaunchKernelGGL(k_count_active, dim3(1), dim3(BS), 0, S(), t.desc->p, (int)t.inst.size(), which,
                           count->p + slot);
        EQ_HIP(hipGetLastError());
        EQ_HIP(hipMemcpyAsync(host_count + slot, count->p + slot, sizeof(int), hipMemcpyDeviceToHost, S()));
        EQ_HIP(hipEventRecord(ev[slot], S()));
    }
    int count_wait(int slot) {
        EQ_HIP(hipEventSynchronize(ev[slot]));
        return host_count[slot];
    }

    template <class T>
    void bicgstab(Table& t, double tol, int maxit) {
        HL((k_kinit<T, CP>), t.part->p, t.ng);
        fin(t, F(A_INIT, 2));
        // a membership seen once (request churn) runs eagerly; capture pays off from its second solve
        if (use_graph && !timing_enabled() && t.solves++ > 0) {
            // graph_iters iterations per launch; instances that converge inside a
            // launch are masked by their device flag, so extra iterations are no-ops.
            auto key = std::make_pair(tol, maxit);
            auto g = t.graphs.find(key);
            if (g == t.graphs.end()) {
                hipGraph_t graph;
                hipGraphExec_t exec;
                EQ_HIP(hipStreamBeginCapture(S(), hipStreamCaptureModeThreadLocal));
                for (int q = 0; q < graph_iters; ++q) bicg_iteration<T>(t, tol, maxit);
                EQ_HIP(hipStreamEndCapture(S(), &graph));
                EQ_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                g = t.graphs.emplace(key, std::make_pair(graph, exec)).first;
            }
            // pipelined: launch k+1 is queued before the active count of launch k is read
            int k = 0;
            for (int it = 0; it < maxit; it += graph_iters, ++k) {
                EQ_HIP(hipGraphLaunch(g->second.second, S()));
                count_async(t, 0, k % NRING);
                if (k > 0 && !count_wait((k - 1) % NRING)) break;
            }
            Device::get().sync();
            return;
        }
        if (!count_active(t, 0)) return;
        for (int it = 0; it < maxit; ++it) {
            bicg_iteration<T>(t, tol, maxit);
            if (!count_active(t, 0)) break;
        }
    }

    template <class T>
    void linear_solve(Table& t) {
        // du = 0, r64 = F; refinement flags from ||F||
        HL((k_zero_du<CP>));
        const unsigned ni = (unsigned)t.inst.size();
        hipLaunchKernelGGL(k_refine_flags, dim3((ni + 63) / 64), dim3(64), 0, S(), t.desc->p, (int)ni, 0,
                           cfg.solver_tol, cfg.max_refine);
        if (!mixed) {
            HL((k_load_rhs<T, CP>), 0);
            precond_setup<T>(t);
            bicgstab<T>(t, cfg.solver_tol, cfg.solver_max_iter);
            HL((k_accum<T, CP>), 0);
            return;
        }
        precond_setup<T>(t);
        for (int ref = 0; ref < cfg.max_refine; ++ref) {
            if (ref > 0 && !count_active(t, 1)) break;   // all start active (||F|| > 0)
            HL((k_load_rhs<T, CP>), 1);
            bicgstab<T>(t, cfg.inner_tol, cfg.solver_max_iter);
            HL((k_accum<T, CP>), 1);
            {
                TimedLaunch tl("defect_fp64");
                HL((k_defect<CP>), t.part->p);
            }
            fin(t, F(A_STORE, 1, 0, 0, S_FN));
            hipLaunchKernelGGL(k_refine_flags, dim3((ni + 63) / 64), dim3(64), 0, S(), t.desc->p, (int)ni, 1,
                               cfg.solver_tol, cfg.max_refine);
        }
    }
    template <class T>
    void precond_setup(const Table& t) {
        if (pc != PC_NONE) HL((k_precond_setup<T, CP>), pc);
    }
#undef HL

    // per-instance scalars of several slots in one device-to-host copy: [slot][instance]
    std::vector<double> gather(const Table& t, std::initializer_list<int> slots) {
        const int ni = (int)t.inst.size();
        auto d = dev_alloc<double>((size_t)ni * slots.size());
        int q = 0;
        for (int slot : slots) {
            hipLaunchKernelGGL(k_gather_slot, dim3((ni + 63) / 64), dim3(64), 0, S(), t.desc->p, ni, slot,
                               d->p + (size_t)q * ni);
            EQ_HIP(hipGetLastError());
            ++q;
        }
        return dev_to_host(d->p, (size_t)ni * slots.size());
    }
};

// ===========================================================================
HeatBatch::HeatBatch(const FusedConfig& cfg) : p_(new Impl()) {
    p_->cfg = cfg;
    EQ_REQUIRE(cfg.precision == "fp64" || cfg.precision == "mixed", "precision must be fp64 or mixed");
    p_->mixed = cfg.precision == "mixed";
    const std::string pc = cfg.precond == "diagonal" ? "jacobi" : cfg.precond;
    EQ_REQUIRE(pc == "jacobi" || pc == "ilu0" || pc == "none", "preconditioner must be jacobi, ilu0 or none");
    p_->pc = pc == "jacobi" ? PC_JACOBI : pc == "ilu0" ? PC_ILU : PC_NONE;
    p_->use_graph = cfg.graph;
    p_->count = dev_alloc<int>(Impl::NRING);
    EQ_HIP(hipHostMalloc((void**)&p_->host_count, Impl::NRING * sizeof(int), hipHostMallocDefault));
    for (int i = 0; i < Impl::NRING; ++i) EQ_HIP(hipEventCreateWithFlags(&p_->ev[i], hipEventDisableTiming));
}

HeatBatch::~HeatBatch() {
    Device::get().sync();
    for (int i = 0; i < Impl::NRING; ++i)
        if (p_->ev[i]) (void)hipEventDestroy(p_->ev[i]);
    if (p_->host_count) (void)hipHostFree(p_->host_count);
    delete p_;
}

size_t HeatBatch::estimate_bytes(int n, const std::string& precision, bool twophase) {
    const size_t ks = precision == "mixed" ? 4 : 8;
    const bool mixed = precision == "mixed";
    if (!twophase)   // u, u0, F, du, r64, 5 coefficients (+ FP32 copies), dinv, Krylov vectors
        return (size_t)n * (10 * 8 + (mixed ? 5 * 4 : 0) + ks * (1 + NVEC)) + NSLOT * 8 + 64;
    // tx, ty, tz, pv, src, p, sw, sw0, F, du, r64, 7 coefficients (+ FP32), dinv, Krylov vectors
    return (size_t)n * (18 * 8 + (mixed ? 7 * 4 : 0) + ks * (1 + NVEC)) + NSLOT * 8 + 64;
}

// Newton / Krylov storage common to both instance kinds.
static void alloc_solver(Inst& I, bool mixed, size_t ks) {
    const int n = I.n;
    I.F = dev_alloc<double>(n);
    I.du = dev_alloc<double>(n);
    I.r64 = dev_alloc<double>(n);
    I.C = dev_alloc<double>(n);
    I.W = dev_alloc<double>(n);
    I.E = dev_alloc<double>(n);
    I.S = dev_alloc<double>(n);
    I.N = dev_alloc<double>(n);
    if (I.nz > 1) {
        I.B = dev_alloc<double>(n);
        I.T = dev_alloc<double>(n);
    }
    I.sc = dev_alloc<double>(NSLOT);
    I.act = dev_alloc<int>(4);
    EQ_HIP(hipMemsetAsync(I.sc->p, 0, NSLOT * sizeof(double), S()));
    EQ_HIP(hipMemsetAsync(I.act->p, 0, 4 * sizeof(int), S()));
    if (mixed) I.coef32 = dev_alloc<char>((size_t)I.ncoef() * n * sizeof(float));
    I.dinv = dev_alloc<char>((size_t)n * ks);
    I.vecs = dev_alloc<char>((size_t)NVEC * n * ks);
}

int HeatBatch::add(int nx, int ny, double k, double cv, const double* u_host, double dx, double dy) {
    EQ_REQUIRE(nx > 0 && ny > 0, "instance sizes must be positive");
    EQ_REQUIRE((long)nx * ny < (1L << 31) - CPT_MAX * BS, "instance too large");
    auto I = std::make_unique<Inst>();
    I->id = p_->next_id++;
    I->nx = nx;
    I->ny = ny;
    I->n = nx * ny;
    I->k = k;
    I->cv = cv;
    I->dx = dx;
    I->dy = dy;
    I->u = dev_from_host(u_host, I->n);
    I->u0 = dev_alloc<double>(I->n);
    alloc_solver(*I, p_->mixed, p_->ks());
    const int id = I->id;
    p_->insts[id] = std::move(I);
    ++p_->version;
    return id;
}

int HeatBatch::add_twophase(const TwoPhaseSpec& sp, const double* perm, const double* poro, const double* sw_host,
                            int nsrc, const int* src_cells, const double* src_vals) {
    const int nx = sp.nx, ny = sp.ny, nz = sp.nz;
    EQ_REQUIRE(nx > 0 && ny > 0 && nz > 0, "instance sizes must be positive");
    EQ_REQUIRE((long)nx * ny * nz < (1L << 31) - CPT_MAX * BS, "instance too large");
    EQ_REQUIRE(sp.corey == 1 || sp.corey == 2, "corey must be 1 or 2");
    EQ_REQUIRE(sp.watervisc > 0 && sp.oilvisc > 0, "viscosities must be positive");
    auto I = std::make_unique<Inst>();
    I->id = p_->next_id++;
    I->kind = K_TWOPHASE;
    I->nx = nx;
    I->ny = ny;
    I->nz = nz;
    const int n = nx * ny * nz;
    I->n = n;
    I->dx = sp.dx;
    I->dy = sp.dy;
    I->corey = sp.corey;
    I->mw_inv = 1.0 / sp.watervisc;
    I->mo_inv = 1.0 / sp.oilvisc;
    // transmissibilities (computeTransmissibilities): half-transmissibility of a cell
    // towards a face = area * perm * |n . d| / |d|^2 with d the cell-to-face centroid
    // vector (|d| = h/2 on a Cartesian grid), harmonic sum of the two halves
    const double dz = nz > 1 ? sp.dz : 1.0;
    const double ax = sp.dy * dz, ay = sp.dx * dz, az = sp.dx * sp.dy;
    const double vol = sp.dx * sp.dy * dz;
    std::vector<double> h((size_t)5 * n, 0.0);
    auto half = [](double a, double perm_c, double hh) { return a * perm_c * (0.5 * hh) / (0.25 * hh * hh); };
    auto harm = [](double h1, double h2) { return 1.0 / (1.0 / h1 + 1.0 / h2); };
    for (int c = 0; c < n; ++c) {
        const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
        if (i < nx - 1) h[c] = harm(half(ax, perm[c], sp.dx), half(ax, perm[c + 1], sp.dx));
        if (j < ny - 1) h[n + c] = harm(half(ay, perm[c], sp.dy), half(ay, perm[c + nx], sp.dy));
        if (k < nz - 1) h[2L * n + c] = harm(half(az, perm[c], dz), half(az, perm[c + nx * ny], dz));
        const double ph = poro[c] < sp.min_poro ? sp.min_poro : poro[c];
        h[3L * n + c] = ph * vol;
    }
    for (int q = 0; q < nsrc; ++q) {
        EQ_REQUIRE(src_cells[q] >= 0 && src_cells[q] < n, "source cell " << src_cells[q] << " outside the grid");
        EQ_REQUIRE(q == 0 || src_cells[q] > src_cells[q - 1], "source cells must be ascending");
        h[4L * n + src_cells[q]] = src_vals[q];
    }
    I->tp = dev_from_host(h.data(), h.size());
    I->p = dev_alloc<double>(n);
    launch_fill(I->p->p, 0.0, n);
    I->sw = dev_from_host(sw_host, n);
    I->sw0 = dev_alloc<double>(n);
    alloc_solver(*I, p_->mixed, p_->ks());
    const int id = I->id;
    p_->insts[id] = std::move(I);
    ++p_->version;
    return id;
}

static Inst& get_inst(HeatBatch::Impl* p, int slot) {
    auto it = p->insts.find(slot);
    EQ_REQUIRE(it != p->insts.end(), "unknown batch slot " << slot);
    return *it->second;
}

void HeatBatch::remove(int slot) {
    get_inst(p_, slot);
    p_->tcache.clear();
    p_->insts.erase(slot);
    ++p_->version;
}
int HeatBatch::size_of(int slot) const { return get_inst(p_, slot).n; }
bool HeatBatch::in_flight(int slot) const {
    const Inst& I = get_inst(p_, slot);
    return I.state == Inst::START || I.state == Inst::SOLVE;
}
const StepStats& HeatBatch::stats(int slot) const { return get_inst(p_, slot).st; }
size_t HeatBatch::device_bytes(int slot) const {
    const Inst& I = get_inst(p_, slot);
    return estimate_bytes(I.n, p_->cfg.precision, I.kind == K_TWOPHASE);
}

static const int kStage[2] = {0, 1};

void HeatBatch::begin_step(int slot, double dt) {
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.state != Inst::START && I.state != Inst::SOLVE, "slot " << slot << " already has a step in flight");
    if (I.dt != dt) ++p_->version;
    I.dt = dt;
    if (I.kind == K_HEAT) {
        EQ_HIP(hipMemcpyAsync(I.u0->p, I.u->p, (size_t)I.n * sizeof(double), hipMemcpyDeviceToDevice, S()));
    } else {
        // sw0 = sw; pressure solve first, from the previous step's pressure
        EQ_HIP(hipMemcpyAsync(I.sw0->p, I.sw->p, (size_t)I.n * sizeof(double), hipMemcpyDeviceToDevice, S()));
        EQ_HIP(hipMemcpyAsync(I.act->p + 2, &kStage[0], sizeof(int), hipMemcpyHostToDevice, S()));
        I.stage = 0;
    }
    I.st = StepStats();
    I.state = Inst::START;
}

std::vector<int> HeatBatch::iterate() {
    Impl& P = *p_;
    std::vector<Inst*> solve, asmb;
    for (auto& kv : P.insts) {
        if (kv.second->state == Inst::SOLVE) solve.push_back(kv.second.get());
        if (kv.second->state == Inst::START || kv.second->state == Inst::SOLVE) asmb.push_back(kv.second.get());
    }
    if (!solve.empty()) {
        TableP tp = P.table(solve);
        Table& t = *tp;
        if (P.mixed) P.linear_solve<float>(t);
        else P.linear_solve<double>(t);
        by_cpt(t.shape.cpt, [&](auto cpt_) {
            constexpr int CP = decltype(cpt_)::value;
            hipLaunchKernelGGL((k_newton_update<CP>), dim3((unsigned)t.nb), dim3(BS), 0, S(), t.desc->p, t.blk->p);
            EQ_HIP(hipGetLastError());
        });
    }
    std::vector<int> done;
    if (asmb.empty()) return done;
    TableP tp = P.table(asmb);
    Table& t = *tp;
    P.assemble(t);
    // one device-to-host copy per Newton iteration: ||F||^2 and the solve statistics
    const size_t na = asmb.size();
    std::vector<double> g = P.gather(t, {S_FN, S_ITT, S_REF, S_RN});
    const FusedConfig& c = P.cfg;
    for (size_t q = 0; q < na; ++q) {
        Inst& I = *asmb[q];
        if (I.state == Inst::SOLVE) {   // it was solved and updated in this call
            I.st.newton += 1;
            I.st.krylov += (int)g[na + q];
            I.st.refinements += (int)g[2 * na + q];
            if (I.kind == K_TWOPHASE && I.stage == 0) {
                I.st.newton_p += 1;
                I.st.krylov_p += (int)g[na + q];
            }
            if (!std::isfinite(g[3 * na + q])) I.st.failed = true;
        }
        const double fn = std::sqrt(g[q]);
        if (I.state == Inst::START) {
            I.r0_stage = fn;
            if (I.kind == K_HEAT || I.stage == 0) I.st.r0 = fn;
            I.st.newton_stage = 0;
        } else {
            I.st.newton_stage += 1;
        }
        I.st.rn = fn;
        int stop = 0;
        if (fn <= c.abs_res_tol) stop = 1;
        else if (c.rel_res_tol > 0 && fn <= c.rel_res_tol * I.r0_stage) stop = 2;
        const bool bad = !std::isfinite(fn) || I.st.failed;
        const bool out = I.st.newton_stage >= c.max_iter;
        if (I.kind == K_TWOPHASE && I.stage == 0 && !bad && (stop || out)) {
            // pressure solved: the saturation solve starts from sw = 0.5 (twophase.equelle:107)
            I.st.stop_p = stop;
            I.stage = 1;
            EQ_HIP(hipMemcpyAsync(I.act->p + 2, &kStage[1], sizeof(int), hipMemcpyHostToDevice, S()));
            launch_fill(I.sw->p, 0.5, I.n);
            I.state = Inst::START;
            continue;
        }
        if (stop || bad || out) {
            I.st.stop = bad ? 0 : stop;
            if (I.kind == K_TWOPHASE && I.st.stop_p == 0) I.st.stop = 0;   // pressure did not converge
            I.st.failed = bad;
            I.state = Inst::DONE;
            done.push_back(I.id);
        } else {
            I.state = Inst::SOLVE;
        }
    }
    return done;
}

void HeatBatch::step(const std::vector<int>& slots, double dt) {
    for (int s : slots) begin_step(s, dt);
    int guard = 0;
    for (;;) {
        bool any = false;
        for (int s : slots) any |= in_flight(s);
        if (!any) break;
        iterate();
        EQ_REQUIRE(++guard < 10 * (p_->cfg.max_iter + 2), "batch step did not terminate");
    }
}

LaunchShape HeatBatch::launch_shape() const {
    LaunchShape r;
    r.cpt = p_->last_shape.cpt;
    r.xcd = p_->last_shape.xcd;
    r.signatures = (int)p_->tuned.size();
    return r;
}

double HeatBatch::time_kernel(const std::string& name, int reps, double* event_ms) {
    Impl& P = *p_;
    EQ_REQUIRE(reps > 0, "reps must be positive");
    std::vector<Inst*> all;
    for (auto& kv : P.insts) all.push_back(kv.second.get());
    EQ_REQUIRE(!all.empty(), "empty batch");
    // a table of the whole membership, in its tuned shape, whose launches stamp the
    // wall clock (block start, wave ends) into ts
    const Shape sh = P.shape_for(all);
    const int cpb = sh.cpt * BS;
    long nb = 0;
    for (Inst* I : all) nb += (I->n + cpb - 1) / cpb;
    auto ts = dev_alloc<unsigned long long>(2 * (size_t)nb);
    TableP tp = P.build_table(all, sh, ts->p);
    Table& t = *tp;
    EQ_REQUIRE(t.nb == nb, "timing table mismatch");
    // every instance active, so the kernels do their full work (the next solve resets the flags)
    for (Inst* I : all) {
        const int one[2] = {1, 1};
        EQ_HIP(hipMemcpyAsync(I->act->p, one, sizeof(one), hipMemcpyHostToDevice, S()));
    }
    Device::get().sync();
    int khz = 0;
    EQ_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
    EQ_REQUIRE(khz > 0, "no wall-clock rate");
    std::function<void()> launch;
    hipEvent_t e0 = nullptr, e1 = nullptr;
#define TL(kernel, ...)                                                                                          \
    launch = [&] {                                                                                               \
        by_cpt(t.shape.cpt, [&](auto cpt_) {                                                                     \
            constexpr int CP = decltype(cpt_)::value;                                                            \
            hipExtLaunchKernelGGL(kernel, dim3((unsigned)t.nb), dim3(BS), 0, S(), e0, e1, 0, t.desc->p, t.blk->p, \
                                  ##__VA_ARGS__);                                                                \
        });                                                                                                      \
    }
    if (name == "stencil_spmv") {
        if (P.mixed) TL((k_spmv_dot<float, CP>), (int)V_PH, (int)V_V, 1, (int)V_RH, -1, t.part->p, t.ng);
        else TL((k_spmv_dot<double, CP>), (int)V_PH, (int)V_V, 1, (int)V_RH, -1, t.part->p, t.ng);
    } else if (name == "assemble") {
        if (P.mixed) TL((k_assemble<true, CP>), t.part->p);
        else TL((k_assemble<false, CP>), t.part->p);
    } else if (name == "ilu_black") {
        if (P.mixed) TL((k_ilu_black<float, CP>), (int)V_P, (int)V_PH);
        else TL((k_ilu_black<double, CP>), (int)V_P, (int)V_PH);
    } else if (name == "ilu_red") {
        if (P.mixed) TL((k_ilu_red<float, CP>), (int)V_P, (int)V_PH);
        else TL((k_ilu_red<double, CP>), (int)V_P, (int)V_PH);
    } else if (name == "defect") {
        TL((k_defect<CP>), t.part->p);
    } else {
        EQ_REQUIRE(false, "unknown kernel " << name << " (stencil_spmv, assemble, ilu_black, ilu_red, defect)");
    }
#undef TL
    EQ_HIP(hipEventCreate(&e0));
    EQ_HIP(hipEventCreate(&e1));
    launch();   // warm
    EQ_HIP(hipGetLastError());
    std::vector<unsigned long long> h(2 * (size_t)nb);
    double total = 0.0, ev_total = 0.0;
    for (int r = 0; r < reps; ++r) {
        EQ_HIP(hipMemsetAsync(ts->p, 0, h.size() * sizeof(unsigned long long), S()));
        launch();
        EQ_HIP(hipMemcpyAsync(h.data(), ts->p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, S()));
        EQ_HIP(hipEventSynchronize(e1));
        Device::get().sync();
        float ms = 0.f;
        EQ_HIP(hipEventElapsedTime(&ms, e0, e1));
        ev_total += ms;
        unsigned long long lo = ~0ull, hi = 0;
        for (long b = 0; b < nb; ++b) lo = std::min(lo, h[2 * b]), hi = std::max(hi, h[2 * b + 1]);
        EQ_REQUIRE(hi > lo, "kernel clock stamps missing");
        total += (double)(hi - lo) / (double)khz;   // ms
    }
    EQ_HIP(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (event_ms) *event_ms = ev_total / reps;
    return total / reps;
}

void HeatBatch::get_state(int slot, double* u) const {
    Inst& I = get_inst(p_, slot);
    std::vector<double> h = dev_to_host(I.kind == K_HEAT ? I.u->p : I.sw->p, I.n);
    std::copy(h.begin(), h.end(), u);
}
void HeatBatch::get_twophase_state(int slot, double* p, double* sw) const {
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_TWOPHASE, "slot " << slot << " is not a two-phase instance");
    if (p) {
        std::vector<double> h = dev_to_host(I.p->p, I.n);
        std::copy(h.begin(), h.end(), p);
    }
    if (sw) {
        std::vector<double> h = dev_to_host(I.sw->p, I.n);
        std::copy(h.begin(), h.end(), sw);
    }
}
void HeatBatch::set_state(int slot, const double* u) {
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_HEAT, "set_state: slot " << slot << " is not a heat instance");
    I.u = dev_from_host(u, I.n);
    ++p_->version;
}

// CSR with ascending columns: bottom, south, west, diagonal, east, north, top (existing
// neighbours only; stored zeros kept, as in the AD engine's union patterns).
static void csr_from_stencil(const Inst& I, std::vector<int>& rp, std::vector<int>& ci, std::vector<double>& v) {
    std::vector<double> C = dev_to_host(I.C->p, I.n), W = dev_to_host(I.W->p, I.n), E = dev_to_host(I.E->p, I.n),
                        Sx = dev_to_host(I.S->p, I.n), N = dev_to_host(I.N->p, I.n), Bt, Tp;
    if (I.nz > 1) {
        Bt = dev_to_host(I.B->p, I.n);
        Tp = dev_to_host(I.T->p, I.n);
    }
    const int nx = I.nx, ny = I.ny, nz = I.nz, nxy = nx * ny;
    rp.assign(I.n + 1, 0);
    ci.clear();
    v.clear();
    for (int c = 0; c < I.n; ++c) {
        const int i = c % nx, j = (c / nx) % ny, k = c / nxy;
        if (k > 0) ci.push_back(c - nxy), v.push_back(Bt[c]);
        if (j > 0) ci.push_back(c - nx), v.push_back(Sx[c]);
        if (i > 0) ci.push_back(c - 1), v.push_back(W[c]);
        ci.push_back(c), v.push_back(C[c]);
        if (i < nx - 1) ci.push_back(c + 1), v.push_back(E[c]);
        if (j < ny - 1) ci.push_back(c + nx), v.push_back(N[c]);
        if (k < nz - 1) ci.push_back(c + nxy), v.push_back(Tp[c]);
        rp[c + 1] = (int)ci.size();
    }
}

void HeatBatch::export_residual(int slot, const double* u, const double* u0, double dt, std::vector<double>& F,
                                std::vector<int>& rp, std::vector<int>& ci, std::vector<double>& v) {
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_HEAT, "slot " << slot << " is not a heat instance");
    I.u = dev_from_host(u, I.n);
    I.u0 = dev_from_host(u0, I.n);
    I.dt = dt;
    ++p_->version;
    TableP tp = p_->table({&I});
    p_->assemble(*tp);
    F = dev_to_host(I.F->p, I.n);
    csr_from_stencil(I, rp, ci, v);
}

void HeatBatch::export_twophase_residual(int slot, int stage, const double* p, const double* sw0, const double* sw,
                                         double dt, std::vector<double>& F, std::vector<int>& rp,
                                         std::vector<int>& ci, std::vector<double>& v) {
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_TWOPHASE, "slot " << slot << " is not a two-phase instance");
    EQ_REQUIRE(stage == 0 || stage == 1, "stage must be 0 (pressure) or 1 (transport)");
    EQ_HIP(hipMemcpy(I.p->p, p, (size_t)I.n * sizeof(double), hipMemcpyHostToDevice));
    EQ_HIP(hipMemcpy(I.sw0->p, sw0, (size_t)I.n * sizeof(double), hipMemcpyHostToDevice));
    EQ_HIP(hipMemcpy(I.sw->p, sw, (size_t)I.n * sizeof(double), hipMemcpyHostToDevice));
    EQ_HIP(hipMemcpy(I.act->p + 2, &kStage[stage], sizeof(int), hipMemcpyHostToDevice));
    I.stage = stage;
    I.dt = dt;
    ++p_->version;
    TableP tp = p_->table({&I});
    p_->assemble(*tp);
    F = dev_to_host(I.F->p, I.n);
    csr_from_stencil(I, rp, ci, v);
}

}  // namespace eqhip
