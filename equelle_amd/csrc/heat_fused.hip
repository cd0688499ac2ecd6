// Fused nonlinear-heat engine (see heat_fused.hpp).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
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

// per-instance scalars; S_TOL2: the square of the instance's inner (mixed-precision) Krylov tolerance
// S_MEAN: the mean of the Krylov right-hand side over the cells (singular systems only)
// S_F0: ||F||^2 at the first Newton iteration of the step (stage); 0 until then
// S_LTOL: the linear tolerance of the current Newton iteration; S_ITOL: the square of the
// instance's base inner (mixed-precision) tolerance; S_NK: Newton iterations of the step
// (stage) so far; S_R1, S_R1B: ||F_1|| / ||F_0|| of the latest step (two-phase: of its
// pressure / transport stage), 0 until measured
// S_XUPD: the four-ranks-per-lane BiCGStab's pending solution update (set when the
// iteration's second finaliser ran for the instance, cleared by the next first finaliser)
// S_U0D: heat, 0 from the start of a step until its first assembly has copied u0 = u (the
// assembly's finaliser sets it); next to S_F0, which the step start clears with it
enum Slot { S_RHO = 0, S_RHO_OLD, S_ALPHA, S_OMEGA, S_RR, S_BB, S_BREAK, S_IT, S_FN, S_RN, S_BN, S_REF, S_ITT, S_TOL2,
            S_MEAN, S_BETA, S_CONT, S_F0, S_U0D, S_LTOL, S_ITOL, S_NK, S_R1, S_R1B, S_XUPD, NSLOT = 32 };
// Krylov vectors; V_Y is the red-rank scratch of the reduced operator (ilu0 only), V_Z the
// V-cycle's z = M^-1 r (two-phase pressure by CG with multigrid)
enum Vec { V_R = 0, V_RH, V_P, V_V, V_S, V_T, V_PH, V_SH, V_X, V_Y, V_Z, NVEC };
enum Act { A_INIT = 0, A_ALPHA, A_OMEGA, A_RHO, A_STORE, A_MEAN, A_CG_ALPHA, A_CG_RHO, A_CG_START, A_REFINE };
// Krylov methods of a solve (bit mask): BiCGStab, and CG for symmetric systems (the
// two-phase pressure stage in the reduced form)
enum Meth { M_BICG = 1, M_CG = 2 };
enum Pc { PC_NONE = 0, PC_JACOBI = 1, PC_ILU = 2 };
// Krylov form and right preconditioner of one instance: the full system with no
// preconditioner, Jacobi, or red-black ILU(0) applied in two passes; or the red-black
// reduced (Schur complement) system with its diagonal (= red-black ILU(0)).
enum Pm { PM_NONE = 0, PM_JACOBI = 1, PM_ILU = 2, PM_SCHUR = 3 };

// K_LINEAR: a linear system J x = b whose Jacobian has the pattern of a Cartesian 5- or
// 7-point stencil, gathered from CSR (the drop-in AD engine's Newton step); only the
// Krylov solve and its refinement run on it (no assembly, no Newton update).
enum Kind { K_HEAT = 0, K_TWOPHASE = 1, K_LINEAR = 2 };

struct Desc {
    int n, nx, ny, nz;
    int nxr;                      // real row length; heat rows may be padded to nx (see Inst::nxr)
    int xcd;                      // XCD-contiguous block order (same for all of a table)
    int kind;                     // K_HEAT | K_TWOPHASE | K_LINEAR
    int corey;
    int gbase, ngrp;              // the instance's partials: groups [gbase, gbase + ngrp) of BS cells
    int h;                        // pair ranks (n + 1) / 2: cells per colour, rounded up
    int pcm;                      // Krylov form / preconditioner (Pm)
    int ns;                       // Krylov slots per pair rank: 1 (reduced system) or 2 (full system)
    int ngrpk;                    // Krylov partials: groups [gbase, gbase + ngrpk) of BS pairs
    int vec;                      // Krylov kernels in the four-ranks-per-lane form (2D heat, nx % 8 == 0)
    int tile;                     // vec, nx % 512 == 0: the reduced operator in 16-row tiles (k_schur_tile)
    int cgp;                      // two-phase, reduced form: the pressure stage (symmetric) is solved by CG
    int mgp;                      // two-phase, full form: pressure by CG with the multigrid V-cycle
    int mgl;                      // multigrid: levels below this one (0 on the coarsest level)
    int mgt;                      // multigrid: the level belongs to the instance's workgroup tail (k_mg_tail)
    void* minv;                   // multigrid, coarsest level: dense (A + s 1 1^T)^-1, cell order
    double mgs;                   // multigrid: factor of the coarse-grid correction (z += mgs P z_c)
    double rnx, rnxy;             // 1 / nx, 1 / (nx ny): cell -> (i, j, k) without integer division
    double dt;
    // heat: k(u) = k (1 + u^2), x / y face geometry factors, cell volume
    double k, cv, gx, gy, vol;
    double hcgx, hcgy;            // dt gx / (cv vol), dt gy / (cv vol): set with dt, once per instance
    double *u, *u0;
    // two-phase: transmissibility of the +x / +y / +z face of each cell, pore volume,
    // source rate, pressure, saturation and its value at the start of the step
    double mw_inv, mo_inv;        // 1 / watervisc, 1 / oilvisc
    const double *tx, *ty, *tz, *pv, *src;
    double *p, *sw, *sw0;
    double *F, *du, *r64;
    // FP64 Jacobian stencil coefficients in cell order (B, T: 3D only), stored only for
    // two-phase and linear-system instances in mixed precision (the refinement defect reads
    // them); a heat instance recomputes its FP64 rows from u.  Null otherwise.
    double *C, *W, *E, *S, *N, *B, *T;
    // Krylov-precision coefficients C, W, E, S, N, B, T in red-black rank order (2 h each)
    void* K[7];
    void* dinv;                   // 2 h: red 1/C_r; black 1/pivot (ilu0), 1/C_b (jacobi), 1 (none)
    void* v[NVEC];                // Krylov vectors
    double* sc;                   // NSLOT scalars
    int* act;                     // [0] Krylov active, [1] refinement active, [2] two-phase stage
    // Batch-wide gate of the work that follows a Krylov solve (the solution update, the
    // refinement defect and flags, the Newton update, the assembly after it, its norms):
    // 1 except between a speculative solve whose blind iterations left an instance
    // unconverged and the host's catch-up, when that work is a no-op.
    const int* gate;
};
__device__ inline bool gated_off(const Desc& d) { return *d.gate == 0; }

// The Krylov method of an instance's current solve: CG for the two-phase pressure stage
// (a symmetric M-matrix: both cells of a face see the same upwinded mobility), in the
// reduced form with the Schur complement's diagonal or in the full form with the
// multigrid V-cycle, BiCGStab otherwise.  Kernels and finaliser actions of one method skip the
// instances of the other, so both can run in one batch.
__device__ inline int meth_of(const Desc& d) { return (d.cgp || d.mgp) && d.act[2] == 0 ? M_CG : M_BICG; }

// What the last block of an instance does with the summed partials.
struct Fin {
    int act, nd, maxit, slot;
    int kry;                      // partials of a Krylov kernel (groups of pairs) or of a cell kernel
    double tol2;
    // a one-instance table's gather in the same launch (k_gather_slots' stores and flag):
    // gout non-null -> after the action, gout[q] = sc[gs[q]] (q = 0..5), then *gflag = gseq
    double* gout;
    int* gflag;
    int gs[6];
    int gseq;
};
// The gather of a one-instance table from a finaliser's thread 0 (see k_gather_slots).
__device__ inline void fin_gather(const Desc& d, const Fin& f) {
    if (!gated_off(d))
        for (int q = 0; q < 6; ++q) f.gout[q] = d.sc[f.gs[q]];
    __threadfence_system();
    __hip_atomic_store(f.gflag, f.gseq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

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
// handles cells base + t, base + BS + t, ... (coalesced).  Each kernel class (assembly,
// defect, SpMV, preconditioner, vector updates) has its own CPT, chosen by measurement
// (see Impl::shape_for).  Reductions are shape-independent: a block writes one partial
// per group of BS cells (group m = cells [m BS, (m+1) BS) of the block, thread t holds
// cell t of each group) with the same tree for every shape, into the instance's range of
// groups [gbase, gbase + ceil(n / BS)), which does not depend on CPT either; results are
// bit-identical for any CPT and block order.
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2), so
// workgroup w runs on XCD w % 8.  With xcd set, the logical block (position in cell
// order) of workgroup w is chosen so that each XCD processes one contiguous range of
// blocks: stencil neighbours (rows above and below) then hit the same L2.  Bijective
// for any block count.
constexpr int NXCD = 8;
__device__ inline int xcd_block(int w, int nblocks) {
    const int q = nblocks / NXCD, r = nblocks % NXCD, x = w % NXCD;
    return x * q + min(x, r) + w / NXCD;
}
#define LOCATE()                                                                                    \
    const int lb = D[0].xcd ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;         \
    const int2 bi = blk[lb];                                                                        \
    const Desc& d = D[bi.x];                                                                        \
    const int base = bi.y * (CPT * BS) + (int)threadIdx.x
// This block's partials of the instance (pointer, number of its groups that hold cells).
#define PARTS(p) (p) + d.gbase + bi.y * CPT, d.ngrp - bi.y * CPT
#define PARTSK(p) (p) + d.gbase + bi.y * CPT, d.ngrpk - bi.y * CPT
#define FOR_CELLS(c) for (int c = base, c##_end = min(d.n, bi.y * (CPT * BS) + CPT * BS); c < c##_end; c += BS)
#define FOR_PAIRS(p) for (int p = base, p##_end = min(d.h, bi.y * (CPT * BS) + CPT * BS); p < p##_end; p += BS)
#define FOR_GROUPS(m, c) for (int m = 0, c = base; m < CPT; ++m, c += BS)

// One partial per group: v[m] is this thread's value for group m (0 when it has none).
// Within a wave the CPT vectors are reduced by recursive halving (reduce-scatter):
// at offset o the lanes with bit o clear keep the lower half of the groups and add
// their partner's (lane ^ o) values, the others keep the upper half; once one group
// per lane is left the usual butterfly finishes.  Every group is summed over the same
// lane pairs in the same order as a plain shuffle tree (addition is commutative), so
// the partial is bit-identical for every CPT.  Group g ends in lane g * 64 / CPT.
// Only the first `lim` groups (those holding cells of the instance) are stored.
template <int CPT>
__device__ inline void group_store(const double (&v)[CPT], double* out, int lim) {
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
    if ((int)threadIdx.x < CPT && (int)threadIdx.x < lim) {
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
// the instance's scalars.  ND: the dot products (partial sets) the kernel can sum, 2 or 5;
// each set is summed in the same order for either ND.
template <int ND>
__device__ void fin_body(const Desc& d, const double* part, int nb, const Fin& f) {
    if (!f.kry && gated_off(d)) {   // norms of gated cell kernels (assembly, defect)
        if (f.gout && threadIdx.x == 0) fin_gather(d, f);
        return;
    }
    const int2 r = make_int2(d.gbase, d.gbase + (f.kry ? d.ngrpk : d.ngrp));
    __shared__ double res[5];
    __shared__ double lds[ND][FBS / 64];
    // the scalars the action reads, loaded by thread 0 before the partials so both round
    // trips overlap (only this block writes the instance's scalars and flags)
    int a0 = 0, a1 = 0, mth = M_BICG;
    double p_rho = 0.0, p_omega = 0.0, p_it = 0.0, p_itt = 0.0, p_tol2 = 0.0, p_bb = 0.0;
    if (threadIdx.x == 0) {
        a0 = d.act[0];
        a1 = d.act[1];
        mth = meth_of(d);
        p_rho = d.sc[S_RHO];
        p_omega = d.sc[S_OMEGA];
        p_it = d.sc[S_IT];
        p_itt = d.sc[S_ITT];
        p_tol2 = d.sc[S_TOL2];
        p_bb = d.sc[S_BB];
    }
    // partials were written by the previous kernel (visible at this kernel's start); the
    // sets' partials are loaded together before any is added (eight per thread and set for
    // ND = 2, two rounds of four for ND = 5), and each set adds a chunk of eight partials
    // as ((v0 + v1) + (v2 + v3)) + ((v4 + v5) + (v6 + v7)) either way
    constexpr int U = ND > 2 ? 4 : 8;
    double s[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) s[q] = 0.0;
    for (int base = r.x; base < r.y; base += 8 * FBS) {
        double c[ND][8 / U];
#pragma unroll
        for (int half = 0; half < 8 / U; ++half) {
            double v[ND][U];
#pragma unroll
            for (int q = 0; q < ND; ++q)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int b = base + (int)threadIdx.x + (half * U + u) * FBS;
                    v[q][u] = (q < f.nd && b < r.y) ? part[(long)q * nb + b] : 0.0;
                }
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                if (U == 8)
                    c[q][half] = ((v[q][0] + v[q][1]) + (v[q][2 % U] + v[q][3 % U])) +
                                 ((v[q][4 % U] + v[q][5 % U]) + (v[q][6 % U] + v[q][7 % U]));
                else
                    c[q][half] = (v[q][0] + v[q][1]) + (v[q][2 % U] + v[q][3 % U]);
            }
        }
#pragma unroll
        for (int q = 0; q < ND; ++q) s[q] += U == 8 ? c[q][0] : c[q][0] + c[q][(8 / U) - 1];
    }
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        for (int off = 32; off > 0; off >>= 1) s[q] += __shfl_down(s[q], off, 64);
        if ((threadIdx.x & 63) == 0) lds[q][threadIdx.x >> 6] = s[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 5; ++q) {
            double a = 0.0;
            if (q < ND)
                for (int i = 0; i < FBS / 64; ++i) a += lds[q][i];
            res[q] = a;
        }
    }
    if (threadIdx.x != 0) return;
    const int act = f.act, maxit = f.maxit, slot = f.slot;
    const double tol2 = f.tol2;
    double* sc = d.sc;
    switch (act) {
        case A_INIT: {   // start of a Krylov solve for instances under refinement
            if (!a1) {
                d.act[0] = 0;
                break;
            }
            sc[S_RHO] = res[0];
            sc[S_RR] = res[1];
            sc[S_BB] = res[1];
            sc[S_IT] = 0.0;
            sc[S_BREAK] = 0.0;
            sc[S_XUPD] = 0.0;
            sc[S_ALPHA] = sc[S_OMEGA] = sc[S_RHO_OLD] = 1.0;
            d.act[0] = res[1] > 0.0 && isfinite(res[1]);
            break;
        }
        case A_ALPHA:
            sc[S_XUPD] = 0.0;   // the previous iteration's update has been applied
            if (!a0 || mth != M_BICG) break;
            if (res[0] == 0.0 || !isfinite(res[0])) {
                sc[S_BREAK] = 1.0;
                d.act[0] = 0;
            } else sc[S_ALPHA] = p_rho / res[0];
            break;
        case A_OMEGA: {
            if (!a0 || mth != M_BICG) break;
            const double omega = (res[1] == 0.0 || !isfinite(res[1])) ? 0.0   // s already solves the system
                                                                       : res[0] / res[1];
            sc[S_OMEGA] = omega;
            if (!d.vec) break;
            // four-ranks-per-lane form: the end of the iteration without a third reduction.
            // From t.s, t.t, rh.t, rh.s, s.s of the stored s and t: the next rho = rh.r and
            // ||r||^2 of r = s - omega t, as rh.s - omega rh.t and s.s - 2 omega t.s + omega^2
            // t.t (clamped at 0); k_b31 then forms x, r and the next search direction.
            const double rho = res[3] - omega * res[2];
            const double rr = fmax(0.0, res[4] - 2.0 * omega * res[0] + omega * omega * res[1]);
            sc[S_RHO_OLD] = p_rho;
            sc[S_RHO] = rho;
            sc[S_RR] = rr;
            const double it = p_it + 1.0;
            sc[S_IT] = it;
            sc[S_ITT] = p_itt + 1.0;
            sc[S_XUPD] = 1.0;
            const bool ok = isfinite(rr) && isfinite(rho) && omega != 0.0;
            if (!ok) sc[S_BREAK] = 2.0;
            d.act[0] = ok && rr > (tol2 < 0.0 ? p_tol2 : tol2) * p_bb && it < maxit && rho != 0.0;
            break;
        }
        case A_RHO: {
            if (!a0 || mth != M_BICG || d.vec) break;   // (the four-ranks-per-lane form: in A_OMEGA)
            sc[S_RHO_OLD] = p_rho;
            sc[S_RHO] = res[0];
            sc[S_RR] = res[1];
            const double it = p_it + 1.0;
            sc[S_IT] = it;
            sc[S_ITT] = p_itt + 1.0;
            const bool ok = isfinite(res[1]) && p_omega != 0.0;
            if (!ok) sc[S_BREAK] = 2.0;
            // tol2 < 0: the instance's own tolerance (S_TOL2)
            d.act[0] = ok && res[1] > (tol2 < 0.0 ? p_tol2 : tol2) * p_bb && it < maxit && res[0] != 0.0;
            break;
        }
        case A_CG_ALPHA:   // alpha = rho / p.q; a non-positive curvature ends the solve
            if (!a0 || mth != M_CG) break;
            if (!(res[0] > 0.0) || !isfinite(res[0])) {
                sc[S_BREAK] = 1.0;
                d.act[0] = 0;
            } else sc[S_ALPHA] = p_rho / res[0];
            break;
        case A_CG_RHO: {   // rho' = r.z, beta = rho' / rho, ||r||^2; convergence as A_RHO
            if (!a0 || mth != M_CG) break;
            const double beta = res[0] / p_rho;
            sc[S_BETA] = beta;
            sc[S_RHO] = res[0];
            sc[S_RR] = res[1];
            const double it = p_it + 1.0;
            sc[S_IT] = it;
            sc[S_ITT] = p_itt + 1.0;
            const bool ok = isfinite(res[1]) && isfinite(beta);
            if (!ok) sc[S_BREAK] = 2.0;
            d.act[0] = ok && res[1] > (tol2 < 0.0 ? p_tol2 : tol2) * p_bb && it < maxit && res[0] != 0.0;
            break;
        }
        case A_CG_START:   // V-cycle instances: rho = r.z of the first z, first direction p = z
            if (!a0 || mth != M_CG || !d.mgp) break;
            sc[S_RHO] = res[0];
            sc[S_BETA] = 0.0;
            if (!(res[0] > 0.0) || !isfinite(res[0])) {
                sc[S_BREAK] = 1.0;
                d.act[0] = 0;
            }
            break;
        case A_MEAN:   // Krylov right-hand side sum -> its mean over the cells (singular systems)
            sc[S_MEAN] = res[0] / (double)(d.ns == 1 ? d.n / 2 : d.n);   // black cells or all cells
            break;
        case A_REFINE: {   // the refinement defect's ||r||^2, then k_refine_flags' mode 1 (maxit: max_refine)
            sc[S_FN] = res[0];
            if (!a1) break;
            const double rn = sqrt(res[0]);
            const bool progress = rn < sc[S_RN];
            sc[S_RN] = rn;
            sc[S_REF] += 1.0;
            d.act[1] = isfinite(rn) && progress && rn > sc[S_LTOL] * sc[S_BN] && sc[S_REF] < maxit;
            sc[S_CONT] = d.act[1] ? 1.0 : 0.0;   // another round before the Newton update
            break;
        }
        default:
            sc[slot] = res[0];
            if (slot == S_FN && !f.kry) sc[S_U0D] = 1.0;   // an assembly has run: u0 is in place
            break;
    }
    if (f.gout) fin_gather(d, f);
}


template <int ND>
__global__ __launch_bounds__(FBS) void k_fin(const Desc* __restrict__ D, const double* __restrict__ part, int nb,
                                            Fin f) {
    fin_body<ND>(D[blockIdx.x], part, nb, f);
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

// Heat residual and Jacobian row at cell (i, j) from u at the cell (uc) and its four
// neighbours (values of absent neighbours are not used): F = u - u0 + dt/(cv vol) div
// with div = sum_faces g k_f (u_c - u_m), k_f = (k(u_c) + k(u_m)) / 2, k(u) = k (1 + u^2);
// co = {C, W, E, S, N} = dF/du at the cell and the neighbours (0 where absent).  Faces in
// ascending face-id order: west (x-), east (x+), south (y-), north (y+).  The refinement
// defect recomputes rows with this same function, so its FP64 Jacobian is bit-identical
// to the one the assembly derives its Krylov-precision coefficients from.
__device__ inline double heat_row(const Desc& d, int i, int j, double uc, double uw, double ue, double us,
                                  double un, double u0c, double (&co)[5]) {
    if (i >= d.nxr) {   // padding cell (rows padded to the Krylov layout): identity row, F = 0
        co[0] = 1.0;
        co[1] = co[2] = co[3] = co[4] = 0.0;
        return 0.0;
    }
    // per face (cg = dt g / (cv vol), the face's geometry factor g = area / centroid distance):
    // k_f = (k(u_c) + k(u_m)) / 2, flux term cg k_f (u_c - u_m), dF/du_c += cg (k u_c (u_c - u_m)
    // + k_f), dF/du_m = cg (k u_m (u_c - u_m) - k_f); about ten FP64 operations per face
    const double k = d.k;
    const double kfc = 0.5 * k * (2.0 + uc * uc);   // k_f less the neighbour's k u_m^2 / 2
    const double kuc = k * uc;
    double flux = 0.0, diag = 1.0;
    co[1] = co[2] = co[3] = co[4] = 0.0;
    auto face = [&](double um, double cg, double& cm) {
        const double diff = uc - um;
        const double kum = k * um;
        const double kf = (0.5 * kum) * um + kfc;
        flux += (cg * kf) * diff;
        diag += cg * (kuc * diff + kf);
        cm = cg * (kum * diff - kf);
    };
    if (i > 0) face(uw, d.hcgx, co[1]);
    if (i < d.nxr - 1) face(ue, d.hcgx, co[2]);
    if (j > 0) face(us, d.hcgy, co[3]);
    if (j < d.ny - 1) face(un, d.hcgy, co[4]);
    co[0] = diag;
    return (uc - u0c) + flux;
}

// u at cell c and its four neighbours (the cell's own value where a neighbour is absent,
// so every load is in bounds and issued unconditionally).
struct Nb5 {
    double c, w, e, s, n;
};
__device__ inline Nb5 load_nb5(const double* __restrict__ u, int c, int i, int j, int nx, int ny) {
    Nb5 r;
    r.c = u[c];
    r.w = u[i > 0 ? c - 1 : c];
    r.e = u[i < nx - 1 ? c + 1 : c];
    r.s = u[j > 0 ? c - nx : c];
    r.n = u[j < ny - 1 ? c + nx : c];
    return r;
}

// Four-ranks-per-lane form (Desc::vec: 2D heat, reduced system, nx % 8 == 0).  Lane l of
// wave w holds the four consecutive pair ranks p0 .. p0 + 3, p0 = block start + g BS + 4 l,
// of each group g = w, w + BS/64, ... < CPT of its block.  A row holds nx / 2 pair ranks, a
// multiple of 4, so the four ranks lie in one row and every four-rank load is aligned (one
// 16-byte load in FP32, two in FP64).  Rows and columns of the red-black layout then come
// out of the loaded vectors: in row j the cells of the pairs p0 .. p0 + 3 are i = 2 q ..
// 2 q + 7 (q = p0 - j nx / 2), so a cell's west and east neighbours (the other colour) are
// ranks p - 1 and p when its column is even, p and p + 1 when it is odd, and its south and
// north neighbours are ranks p -+ nx / 2.  Coefficients of absent neighbours are zero (the
// assembly writes 0 there), so an in-bounds value of the same vector stands in for them.
// A group's partial sum is the lane sums ((e0 + e1) + e2) + e3 reduced by a butterfly over
// the wave: the same order for every CPT and block order.
template <class T>
struct V4 {
    T e[4];
};
__device__ inline V4<float> ld4(const float* __restrict__ p) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    return {{v.x, v.y, v.z, v.w}};
}
__device__ inline V4<double> ld4(const double* __restrict__ p) {
    const double2 a = *reinterpret_cast<const double2*>(p), b = *reinterpret_cast<const double2*>(p + 2);
    return {{a.x, a.y, b.x, b.y}};
}
__device__ inline void st4(float* __restrict__ p, const V4<float>& v) {
    *reinterpret_cast<float4*>(p) = make_float4(v.e[0], v.e[1], v.e[2], v.e[3]);
}
__device__ inline void st4(double* __restrict__ p, const V4<double>& v) {
    *reinterpret_cast<double2*>(p) = make_double2(v.e[0], v.e[1]);
    *reinterpret_cast<double2*>(p + 2) = make_double2(v.e[2], v.e[3]);
}
#define FOR_VGROUPS(g, p0)                                                                          \
    for (int g = (int)threadIdx.x >> 6, p0 = bi.y * (CPT * BS) + g * BS + 4 * ((int)threadIdx.x & 63); \
         g < CPT; g += BS / 64, p0 += (BS / 64) * BS)
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Group g's partial (lane 0 stores it when the group holds ranks of the instance; cells:
// groups of cells, for the cell-order kernels).
__device__ inline void vgroup_store(const Desc& d, int2 bi, int cpt, int g, double v, double* __restrict__ part,
                                    bool cells = false) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0 && bi.y * cpt + g < (cells ? d.ngrp : d.ngrpk)) part[d.gbase + bi.y * cpt + g] = v;
}
template <class T>
__device__ inline T dot4(const V4<T>& a, const V4<T>& b) {
    return ((a.e[0] * b.e[0] + a.e[1] * b.e[1]) + a.e[2] * b.e[2]) + a.e[3] * b.e[3];
}
template <class T>
__device__ inline double dot4d(const V4<T>& a, const V4<T>& b) {
    return (((double)a.e[0] * (double)b.e[0] + (double)a.e[1] * (double)b.e[1]) + (double)a.e[2] * (double)b.e[2]) +
           (double)a.e[3] * (double)b.e[3];
}
// Eight consecutive FP64 values (cells 2 p0 .. 2 p0 + 7 of a row, 64-byte aligned).
__device__ inline void ld8(const double* __restrict__ p, double (&v)[8]) {
    const V4<double> a = ld4(p), b = ld4(p + 4);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        v[m] = a.e[m];
        v[4 + m] = b.e[m];
    }
}
// W x_w + E x_e + S x_s + N x_n as one fixed chain of fused multiply-adds, so the kernels
// that form this sum (the one-pass and two-pass operators, the red recovery) round it the
// same way wherever the compiler would otherwise contract it.
__device__ inline float stencil4(float kw, float xw, float ke, float xe, float ks, float xs, float kn, float xn) {
    return __builtin_fmaf(kn, xn, __builtin_fmaf(ks, xs, __builtin_fmaf(ke, xe, kw * xw)));
}
__device__ inline double stencil4(double kw, double xw, double ke, double xe, double ks, double xs, double kn,
                                  double xn) {
    return __builtin_fma(kn, xn, __builtin_fma(ks, xs, __builtin_fma(ke, xe, kw * xw)));
}
// Row j and the row's pair offset q0 of rank p0 (2 p0 is the first cell of the pair).
__device__ inline void vrow(const Desc& d, int p0, int& j, int& q0) {
    int i0;
    div_fix(2 * p0, d.rnx, d.nx, j, i0);
    q0 = i0 >> 1;
}
// The off-diagonal sum W + E + S + N (coefficients kq at the ranks p0.., x the other
// colour's vector) for the cells of colour col of row j.
template <class T>
__device__ inline V4<T> voffdiag(const Desc& d, const T* const* kq, int p0, int j, int q0, int col,
                                 const T* __restrict__ x) {
    const int hr = d.nx >> 1;
    const V4<T> xc = ld4(x + p0);
    const bool even = ((j + col) & 1) == 0;   // the cells of colour col have even columns i = 2 q
    const T e = x[even ? (q0 > 0 ? p0 - 1 : p0) : (q0 + 4 < hr ? p0 + 4 : p0)];
    const V4<T> xs = ld4(x + (j > 0 ? p0 - hr : p0)), xn = ld4(x + (j < d.ny - 1 ? p0 + hr : p0));
    const V4<T> kw = ld4(kq[1] + p0), ke = ld4(kq[2] + p0), ks = ld4(kq[3] + p0), kn = ld4(kq[4] + p0);
    V4<T> r;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const T xw = even ? (m == 0 ? e : xc.e[m - 1]) : xc.e[m];
        const T xe = even ? xc.e[m] : (m == 3 ? e : xc.e[m + 1]);
        r.e[m] = stencil4(kw.e[m], xw, ke.e[m], xe, ks.e[m], xs.e[m], kn.e[m], xn.e[m]);
    }
    return r;
}

// Residual + Jacobian assembly and per-group partial ||F||^2.
//   heat:      reads u (5-point), u0; writes F and the 5 coefficients in the Krylov
//              precision (FP32 mixed, FP64 fp64), red-black rank order
//   two-phase: writes F, the Krylov-precision coefficients and (MIXED) the FP64 ones in
//              cell order for the refinement defect
//   norm_only (heat instances): ||F||^2 alone, nothing stored - the check after a Newton
//              iteration predicted to be the step's last (F and the coefficients are only
//              read by a further iteration, which then assembles them first)
template <bool MIXED, int CPT>
__global__ __launch_bounds__(BS) void k_assemble(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                 double* __restrict__ part, int norm_only) {
    LOCATE();
    if (gated_off(d)) return;
    double f2[CPT] = {};
    if (d.kind == K_LINEAR) {   // coefficients come from k_gather_stencil
        group_store<CPT>(f2, PARTS(part));
        return;
    }
    if (d.kind == K_TWOPHASE) {
        const int stage = d.act[2];
        FOR_GROUPS(m, c) {
            if (c >= d.n) continue;
            int i, j, k;
            cell_ijk(d, c, i, j, k);
            double co[7] = {0, 0, 0, 0, 0, 0, 0};
            const double F = stage == 0 ? tp_pressure(d, c, i, j, k, co) : tp_transport(d, c, i, j, k, co);
            d.F[c] = F;
            const int nq = d.nz > 1 ? 7 : 5;
            if (MIXED) {
                double* dst[7] = {d.C, d.W, d.E, d.S, d.N, d.B, d.T};
                for (int q = 0; q < nq; ++q) dst[q][c] = co[q];
            }
            using K = typename std::conditional<MIXED, float, double>::type;
            const int sl = ((i + j + k) & 1) * d.h + (c >> 1);
            for (int q = 0; q < nq; ++q) ((K*)d.K[q])[sl] = (K)co[q];
            f2[m] = F * F;
        }
        group_store<CPT>(f2, PARTS(part));
        return;
    }
    using K = typename std::conditional<MIXED, float, double>::type;
    const int nx = d.nx, ny = d.ny;
    const double* __restrict__ u = d.u;
    const double* __restrict__ u0 = d.u0;
    double* __restrict__ Fo = d.F;
    K* __restrict__ co_out[5] = {(K*)d.K[0], (K*)d.K[1], (K*)d.K[2], (K*)d.K[3], (K*)d.K[4]};
    // the step's first assembly: u0 = u, copied here (S_U0D, set by this launch's finaliser)
    const bool cpy = d.sc[S_U0D] == 0.0;
    double* __restrict__ u0w = d.u0;
    FOR_GROUPS(m, c) {
        if (c >= d.n) continue;
        int i, j, kk;
        cell_ijk(d, c, i, j, kk);
        const Nb5 x = load_nb5(u, c, i, j, nx, ny);
        double co[5];
        const double F = heat_row(d, i, j, x.c, x.w, x.e, x.s, x.n, cpy ? x.c : u0[c], co);
        f2[m] = F * F;
        if (cpy) u0w[c] = x.c;
        if (norm_only) continue;
        Fo[c] = F;
        const int sl = ((i + j) & 1) * d.h + (c >> 1);
#pragma unroll
        for (int q = 0; q < 5; ++q) co_out[q][sl] = (K)co[q];
    }
    group_store<CPT>(f2, PARTS(part));
}

// FP64 heat rows for export (test hook): F and the 5 coefficients into the given arrays.
__global__ void k_heat_rows64(const Desc* __restrict__ D, double* __restrict__ F, double* __restrict__ co) {
    const Desc& d = D[0];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.n) return;
    int i, j, kk;
    cell_ijk(d, c, i, j, kk);
    const Nb5 x = load_nb5(d.u, c, i, j, d.nx, d.ny);
    double r[5];
    F[c] = heat_row(d, i, j, x.c, x.w, x.e, x.s, x.n, d.u0[c], r);
    for (int q = 0; q < 5; ++q) co[(size_t)q * d.n + c] = r[q];
}

// Linear-system instances: the right-hand side and the stencil coefficients of a CSR
// Jacobian whose pattern is a Cartesian 5- or 7-point stencil of the instance's grid
// (checked on the host once per pattern): each entry of row c goes to the coefficient of
// the neighbour it couples (C, W, E, S, N, B, T; zero where the row has no entry), in the
// Krylov precision in red-black rank order and (MIXED) in FP64 cell order for the
// refinement defect; F = rhs; per-group partials of ||F||^2.  Reads 12 B per entry and
// 12 B per row, writes 8 + (MIXED: 8 +) ks bytes per coefficient and row.
template <bool MIXED, int CPT>
__global__ __launch_bounds__(BS) void k_gather_stencil(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                       double* __restrict__ part, const int* __restrict__ rp,
                                                       const int* __restrict__ ci, const double* __restrict__ val,
                                                       const double* __restrict__ rhs) {
    LOCATE();
    using K = typename std::conditional<MIXED, float, double>::type;
    double f2[CPT] = {};
    const int nx = d.nx, ny = d.ny, nz = d.nz, nxy = nx * ny, nq = nz > 1 ? 7 : 5;
    // the rows of a group of BS consecutive cells are one contiguous range of at most 7 BS
    // entries (stencil pattern): staged through LDS by coalesced loads
    __shared__ int lci[7 * BS];
    __shared__ double lval[7 * BS];
    FOR_GROUPS(m, c) {
        const int g0 = c - (int)threadIdx.x, g1 = min(d.n, g0 + BS);
        if (g0 >= d.n) break;   // block-uniform
        const int eb = rp[g0], ee = rp[g1];
        const bool staged = ee - eb <= 7 * BS;   // always, for a stencil pattern
        __syncthreads();
        if (staged)
            for (int e = eb + (int)threadIdx.x; e < ee; e += BS) {
                lci[e - eb] = ci[e];
                lval[e - eb] = val[e];
            }
        __syncthreads();
        if (c >= d.n) continue;
        int i, j, k;
        cell_ijk(d, c, i, j, k);
        double co[7] = {0, 0, 0, 0, 0, 0, 0};
        const int e0 = rp[c], e1 = rp[c + 1];
        for (int e = e0; e < e1; ++e) {
            const int col = staged ? lci[e - eb] : ci[e];
            const double v = staged ? lval[e - eb] : val[e];
            const int q = col == c                              ? 0
                          : (i > 0 && col == c - 1)             ? 1
                          : (i < nx - 1 && col == c + 1)        ? 2
                          : (j > 0 && col == c - nx)            ? 3
                          : (j < ny - 1 && col == c + nx)       ? 4
                          : (nz > 1 && k > 0 && col == c - nxy) ? 5
                          : (nz > 1 && k < nz - 1 && col == c + nxy) ? 6 : -1;
            if (q >= 0) co[q] = v;
        }
        const double F = rhs ? rhs[c] : 0.0;
        d.F[c] = F;
        if (MIXED) {
            double* dst[7] = {d.C, d.W, d.E, d.S, d.N, d.B, d.T};
            for (int q = 0; q < nq; ++q) dst[q][c] = co[q];
        }
        const int sl = ((i + j + k) & 1) * d.h + (c >> 1);
        for (int q = 0; q < nq; ++q) ((K*)d.K[q])[sl] = (K)co[q];
        f2[m] = F * F;
    }
    group_store<CPT>(f2, PARTS(part));
}

// One instance's vector in cell order (FP64) to red-black rank order in the Krylov
// precision (full form: both slots of every pair), and back.
template <class T>
__global__ void k_cells_to_rb(const Desc* __restrict__ D, const double* __restrict__ src, int vec) {
    const Desc& d = D[0];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.n) return;
    int i, j, k;
    cell_ijk(d, c, i, j, k);
    ((T*)d.v[vec])[((i + j + k) & 1) * d.h + (c >> 1)] = (T)src[c];
}
template <class T>
__global__ void k_rb_to_cells(const Desc* __restrict__ D, int vec, double* __restrict__ dst) {
    const Desc& d = D[0];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.n) return;
    int i, j, k;
    cell_ijk(d, c, i, j, k);
    dst[c] = (double)((const T*)d.v[vec])[((i + j + k) & 1) * d.h + (c >> 1)];
}

// Linear-system instances whose right-hand side and coefficients were written in place (the
// drop-in path's fused residual kernel): per-group partials of ||F||^2, as the gather stores them.
template <int CPT>
__global__ __launch_bounds__(BS) void k_linear_fnorm(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                     double* __restrict__ part) {
    LOCATE();
    double f2[CPT] = {};
    FOR_GROUPS(m, c) {
        if (c >= d.n) continue;
        const double F = d.F[c];
        f2[m] = F * F;
    }
    group_store<CPT>(f2, PARTS(part));
}

// FP64 defect for the instances under refinement: r64 = F - J du.  Heat instances
// recompute their FP64 Jacobian rows from u (heat_row; reads u, du 5-point and F, writes
// r64: 32 B per cell); two-phase instances read their stored FP64 coefficients.
template <int CPT>
__global__ __launch_bounds__(BS) void k_defect(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                               double* __restrict__ part) {
    LOCATE();
    if (gated_off(d)) return;
    if (d.vec && d.kind == K_HEAT) {   // heat, nx % 8 == 0: four consecutive cells of one row per lane
        const bool on = d.act[1] != 0;
        const int nx = d.nx, ny = d.ny;
        const double* __restrict__ u = d.u;
        const double* __restrict__ x = d.du;
        const double* __restrict__ Fi = d.F;
        double* __restrict__ ro = d.r64;
        FOR_VGROUPS(g, c0) {
            double r2 = 0.0;
            if (on && c0 < d.n) {
                int j, i0;
                div_fix(c0, d.rnx, nx, j, i0);
                // absent neighbours: an in-bounds stand-in that heat_row and the sum below ignore
                const int w = i0 > 0 ? c0 - 1 : c0, e = i0 + 4 < nx ? c0 + 4 : c0;
                const int so = j > 0 ? c0 - nx : c0, no = j < ny - 1 ? c0 + nx : c0;
                const V4<double> uc = ld4(u + c0), us = ld4(u + so), un = ld4(u + no);
                const V4<double> xc = ld4(x + c0), xs = ld4(x + so), xn = ld4(x + no);
                const V4<double> Fc = ld4(Fi + c0);
                const double uw0 = u[w], ue3 = u[e], xw0 = x[w], xe3 = x[e];
                V4<double> r;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int i = i0 + m;
                    const double uw = m == 0 ? uw0 : uc.e[m > 0 ? m - 1 : 0], ue = m == 3 ? ue3 : uc.e[m < 3 ? m + 1 : 3];
                    const double xw = m == 0 ? xw0 : xc.e[m > 0 ? m - 1 : 0], xe = m == 3 ? xe3 : xc.e[m < 3 ? m + 1 : 3];
                    double co[5];
                    (void)heat_row(d, i, j, uc.e[m], uw, ue, us.e[m], un.e[m], 0.0, co);
                    double acc = co[0] * xc.e[m];
                    if (i > 0) acc += co[1] * xw;
                    if (i < nx - 1) acc += co[2] * xe;
                    if (j > 0) acc += co[3] * xs.e[m];
                    if (j < ny - 1) acc += co[4] * xn.e[m];
                    r.e[m] = Fc.e[m] - acc;
                }
                st4(ro + c0, r);
                r2 = ((r.e[0] * r.e[0] + r.e[1] * r.e[1]) + r.e[2] * r.e[2]) + r.e[3] * r.e[3];
            }
            vgroup_store(d, bi, CPT, g, r2, part, true);
        }
        return;
    }
    double r2[CPT] = {};
    if (d.act[1]) {
        const int nx = d.nx, ny = d.ny;
        const double* __restrict__ x = d.du;
        const double* __restrict__ Fi = d.F;
        double* __restrict__ ro = d.r64;
        if (d.kind == K_HEAT) {
            const double* __restrict__ u = d.u;
            FOR_GROUPS(m, c) {
                if (c >= d.n) continue;
                int i, j, kk;
                cell_ijk(d, c, i, j, kk);
                const Nb5 uu = load_nb5(u, c, i, j, nx, ny);
                const Nb5 xx = load_nb5(x, c, i, j, nx, ny);
                const double Fc = Fi[c];
                double co[5];
                (void)heat_row(d, i, j, uu.c, uu.w, uu.e, uu.s, uu.n, 0.0, co);
                double acc = co[0] * xx.c;
                if (i > 0) acc += co[1] * xx.w;
                if (i < nx - 1) acc += co[2] * xx.e;
                if (j > 0) acc += co[3] * xx.s;
                if (j < ny - 1) acc += co[4] * xx.n;
                const double r = Fc - acc;
                ro[c] = r;
                r2[m] = r * r;
            }
        } else if (d.C) {   // two-phase, linear system: cell-order FP64 coefficients (mixed precision only)
            FOR_GROUPS(m, c) {
                if (c >= d.n) continue;
                int i, j, k;
                cell_ijk(d, c, i, j, k);
                double acc = d.C[c] * x[c];
                if (i > 0) acc += d.W[c] * x[c - 1];
                if (i < nx - 1) acc += d.E[c] * x[c + 1];
                if (j > 0) acc += d.S[c] * x[c - nx];
                if (j < ny - 1) acc += d.N[c] * x[c + nx];
                if (d.nz > 1) {
                    const int nxy = nx * ny;
                    if (k > 0) acc += d.B[c] * x[c - nxy];
                    if (k < d.nz - 1) acc += d.T[c] * x[c + nxy];
                }
                const double r = Fi[c] - acc;
                ro[c] = r;
                r2[m] = r * r;
            }
        }
    }
    group_store<CPT>(r2, PARTS(part));
}

// ---------------------------------------------------------------------------
// Red-black Krylov space.
//
// Cell c has colour (i + j + k) & 1 (red 0, black 1) and rank c >> 1 within its colour:
// every pair of cells (2m, 2m + 1) holds one cell of each colour (neighbours in a row
// differ in colour; a pair wraps to the next row or plane only when the row or plane
// length is odd, which flips the colour once more), so each colour's ranks are dense in
// [0, h), h = (n + 1) / 2.  The Krylov-precision Jacobian coefficients K[q] and dinv are
// stored red ranks first, black ranks at offset h; a stencil neighbour of a cell has the
// other colour and rank (c +- 1) >> 1, (c +- nx) >> 1, (c +- nx ny) >> 1.
//
// Preconditioner ilu0: with red cells ordered first, ILU(0) of A = [C_r A_rb; A_br C_b]
// is exact on the red block and drops the black-black fill, M = (P + L) P^-1 (P + U)
// with P = diag(C_r, P_b), P_b = C_b - diag(A_br C_r^-1 A_rb).  The split-preconditioned
// operator (P + L)^-1 A (P + U)^-1 P is the identity on red cells and S P_b^-1 on black
// cells, S = C_b - A_br C_r^-1 A_rb the Schur complement, and P_b = diag(S).  So the
// solve runs BiCGStab on the black cells only: S x_b = g_b, g_b = b_b - A_br C_r^-1 b_r,
// right-preconditioned by 1/diag(S) (the ILU(0) black pivots), then x_r = C_r^-1 (b_r -
// A_rb x_b).  The red residual of the full system is zero by construction and the black
// one is the reduced system's, so the stopping test (relative to ||g||) is the full
// system's.  S is applied matrix-free in two passes: y_r = C_r^-1 A_rb v_b over red
// cells, then S v = C_b v_b - A_br y_r over black cells.  Krylov vectors have one slot
// per black rank (ns = 1).
// Full system (PM_ILU, PM_JACOBI, PM_NONE): BiCGStab on A in the same storage, two slots
// per pair (red rank p at p, black rank p at h + p; ns = 2), y = A x in one pass; PM_ILU
// applies M^-1 in two passes (black: z_b = (r_b - A_br C_r^-1 r_r) / P_b; red: z_r =
// (r_r - A_rb z_b) / C_r).  The reduced form moves fewer bytes per iteration; the full
// form is kept for systems where FP32 rounding in the reduced operator costs
// convergence (the singular two-phase pressure system).  The form is chosen per
// instance kind, and instances of both forms can share a batch: each operator kernel
// skips the instances of the other form.
//
// Every Krylov kernel runs one thread per pair rank p in [0, h); its partials are groups
// of BS pairs, [gbase, gbase + ngrpk).

// The cell of colour col and rank p with its (i, j, k), or -1 if the instance has none.
__device__ inline int cell_of(const Desc& d, int p, int col, int& i, int& j, int& k) {
    int c = 2 * p;
    if (c >= d.n) return -1;
    cell_ijk(d, c, i, j, k);
    if (((i + j + k) & 1) != col) {
        if (++c >= d.n) return -1;
        cell_ijk(d, c, i, j, k);
    }
    return c;
}

// sum over the existing neighbours of cell c (W, E, S, N, B, T order) of the row's
// coefficient (slot s of K[1..6]) times x at the neighbour's rank.
template <class T>
__device__ inline T offdiag(const Desc& d, int s, int c, int i, int j, int k, const T* __restrict__ x) {
    const int nx = d.nx, ny = d.ny;
    T acc = (T)0;
    if (i > 0) acc += ((const T*)d.K[1])[s] * x[(c - 1) >> 1];
    if (i < nx - 1) acc += ((const T*)d.K[2])[s] * x[(c + 1) >> 1];
    if (j > 0) acc += ((const T*)d.K[3])[s] * x[(c - nx) >> 1];
    if (j < ny - 1) acc += ((const T*)d.K[4])[s] * x[(c + nx) >> 1];
    if (d.nz > 1) {
        const int nxy = nx * ny;
        if (k > 0) acc += ((const T*)d.K[5])[s] * x[(c - nxy) >> 1];
        if (k < d.nz - 1) acc += ((const T*)d.K[6])[s] * x[(c + nxy) >> 1];
    }
    return acc;
}

// Reduced operator, pass 1 (red ranks): y_r = C_r^-1 (A_rb v_b), v = V[iv] (black slots).
// Strided form (the four-ranks-per-lane form applies the operator in one pass, k_schur_vec).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_schur_red(const Desc* __restrict__ D, const int2* __restrict__ blk, int iv,
                                                  int meth) {
    LOCATE();
    if (d.pcm != PM_SCHUR || d.vec || !d.act[0] || !(meth & meth_of(d))) return;
    const T* __restrict__ v = (const T*)d.v[iv];
    const T* __restrict__ dv = (const T*)d.dinv;
    T* __restrict__ y = (T*)d.v[V_Y];
    FOR_PAIRS(p) {
        int i, j, k;
        const int c = cell_of(d, p, 0, i, j, k);
        y[p] = c < 0 ? (T)0 : offdiag<T>(d, p, c, i, j, k, v) * dv[p];
    }
}
// Reduced operator, pass 2 (black ranks): w = S v = C_b v_b - A_br y_r, into V[iw], with
// fused partial dot products dot0 = V[a0].w, dot1 = V[a1].w (a1 < 0: w.w).  Strided form.
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_schur_black(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                    int iv, int iw, int nd, int a0, int a1,
                                                    double* __restrict__ part, int ng, int meth) {
    LOCATE();
    // block-uniform: the full-system kernel, the one-pass kernel, or the other method's
    // launch, writes these partials
    if (d.pcm != PM_SCHUR || d.vec || !(meth & meth_of(d))) return;
    double s0[CPT] = {}, s1[CPT] = {};
    if (d.act[0]) {
        const int h = d.h;
        const T* __restrict__ v = (const T*)d.v[iv];
        const T* __restrict__ y = (const T*)d.v[V_Y];
        const T* __restrict__ Cb = (const T*)d.K[0] + h;
        T* __restrict__ w = (T*)d.v[iw];
        const T* __restrict__ va = (const T*)d.v[a0];
        const T* __restrict__ vb = a1 >= 0 ? (const T*)d.v[a1] : nullptr;
        FOR_GROUPS(m, p) {
            if (p >= h) continue;
            int i, j, k;
            const int c = cell_of(d, p, 1, i, j, k);
            const T wc = c < 0 ? (T)0 : Cb[p] * v[p] - offdiag<T>(d, h + p, c, i, j, k, y);
            w[p] = wc;
            s0[m] = (double)va[p] * (double)wc;
            if (nd > 1) s1[m] = (vb ? (double)vb[p] : (double)wc) * (double)wc;
        }
    }
    group_store<CPT>(s0, PARTSK(part));
    if (nd > 1) group_store<CPT>(s1, PARTSK(part + ng));
}
// Reduced operator in one pass, four-ranks-per-lane form: w = S v = C_b v_b - A_br y_r
// with y_r = C_r^-1 (A_rb v_b) formed where it is used.  A lane's four black cells of row
// j need y_r of the four red cells of rows j - 1, j and j + 1 at the same pair ranks and of
// one more red cell of row j (west of the first or east of the last); each is computed
// exactly as a red-rank lane of the two-pass form computes it (voffdiag, then times 1/C_r),
// so w is bit-identical to the two-pass result.  No y_r is stored: of the two passes' 16
// arrays per rank (the red pass: 4 coefficients, 1/C_r, v, y written; the black pass: C_b,
// 4 coefficients, v, y, the dot operand, w written) 13 remain, neighbour rows coming from
// L2.  Fused partial dot products: dot0 = V[a0].w, dot1 = w.w; nd = 5: also dot2 = rh.w,
// dot3 = rh.V[a0], dot4 = V[a0].V[a0] (the BiCGStab iteration's end, A_OMEGA).
template <class T>
__device__ inline T red_y1(const Desc& d, const T* const* kq, const T* __restrict__ dvr, const T* __restrict__ v,
                           int pe, int j, bool red_even) {
    // the red cell at rank pe of row j (pair qe): W / E neighbours are ranks pe - 1 / pe for
    // a red cell at an even column, pe / pe + 1 at an odd one (both exist: pe is interior
    // to the row whenever it is asked for)
    const int hr = d.nx >> 1;
    const T xw = v[red_even ? pe - 1 : pe], xe = v[red_even ? pe : pe + 1];
    const T xs = v[j > 0 ? pe - hr : pe], xn = v[j < d.ny - 1 ? pe + hr : pe];
    const T r = stencil4(kq[1][pe], xw, kq[2][pe], xe, kq[3][pe], xs, kq[4][pe], xn);
    return r * dvr[pe];
}
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_schur_vec(const Desc* __restrict__ D, const int2* __restrict__ blk, int iv,
                                                  int iw, int nd, int a0, double* __restrict__ part, int ng,
                                                  int meth) {
    LOCATE();
    (void)base;
    if (d.pcm != PM_SCHUR || !d.vec || d.tile || !(meth & meth_of(d))) return;   // block-uniform
    const int h = d.h, hr = d.nx >> 1, ny = d.ny;
    const T* __restrict__ v = (const T*)d.v[iv];
    const T* __restrict__ dvr = (const T*)d.dinv;
    const T* __restrict__ Cb = (const T*)d.K[0] + h;
    T* __restrict__ w = (T*)d.v[iw];
    const T* __restrict__ va = (const T*)d.v[a0];
    const T* kr[5] = {nullptr, (const T*)d.K[1], (const T*)d.K[2], (const T*)d.K[3], (const T*)d.K[4]};
    const T* kb[5] = {nullptr, (const T*)d.K[1] + h, (const T*)d.K[2] + h, (const T*)d.K[3] + h,
                      (const T*)d.K[4] + h};
    const bool on = d.act[0] != 0;
    FOR_VGROUPS(g, p0) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
        if (on && p0 < h) {
            int j, q0;
            vrow(d, p0, j, q0);
            auto red4 = [&](int pr, int jr) {
                const V4<T> o = voffdiag<T>(d, kr, pr, jr, q0, 0, v);
                const V4<T> dd = ld4(dvr + pr);
                V4<T> y;
#pragma unroll
                for (int m = 0; m < 4; ++m) y.e[m] = o.e[m] * dd.e[m];
                return y;
            };
            const V4<T> zero{{(T)0, (T)0, (T)0, (T)0}};
            const V4<T> yc = red4(p0, j);
            const V4<T> ys = j > 0 ? red4(p0 - hr, j - 1) : zero;
            const V4<T> yn = j < ny - 1 ? red4(p0 + hr, j + 1) : zero;
            // black cells of row j at even columns (j odd): west neighbour of the first is the
            // red cell at rank p0 - 1; otherwise the east neighbour of the last, rank p0 + 4
            const bool even = (j & 1) != 0;
            const T ye = even ? (q0 > 0 ? red_y1<T>(d, kr, dvr, v, p0 - 1, j, false) : (T)0)
                              : (q0 + 4 < hr ? red_y1<T>(d, kr, dvr, v, p0 + 4, j, true) : (T)0);
            const V4<T> kw = ld4(kb[1] + p0), ke = ld4(kb[2] + p0), ks = ld4(kb[3] + p0), kn = ld4(kb[4] + p0);
            const V4<T> cb = ld4(Cb + p0), vv = ld4(v + p0), aa = ld4(va + p0);
            V4<T> wc;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const T xw = even ? (m == 0 ? ye : yc.e[m - 1]) : yc.e[m];
                const T xe = even ? yc.e[m] : (m == 3 ? ye : yc.e[m + 1]);
                const T o = stencil4(kw.e[m], xw, ke.e[m], xe, ks.e[m], ys.e[m], kn.e[m], yn.e[m]);
                wc.e[m] = cb.e[m] * vv.e[m] - o;
            }
            st4(w + p0, wc);
            s0 = dot4d(aa, wc);
            if (nd > 1) s1 = dot4d(wc, wc);
            if (nd > 2) {
                const V4<T> rh = ld4((const T*)d.v[V_RH] + p0);
                s2 = dot4d(rh, wc);
                s3 = dot4d(rh, aa);
                s4 = dot4d(aa, aa);
            }
        }
        vgroup_store(d, bi, CPT, g, s0, part);
        if (nd > 1) vgroup_store(d, bi, CPT, g, s1, part + ng);
        if (nd > 2) {
            vgroup_store(d, bi, CPT, g, s2, part + 2 * (long)ng);
            vgroup_store(d, bi, CPT, g, s3, part + 3 * (long)ng);
            vgroup_store(d, bi, CPT, g, s4, part + 4 * (long)ng);
        }
    }
}
// The reduced operator in tiles (Desc::tile: four ranks per lane, nx % 512 == 0, so a row
// holds a whole number of 256-rank groups).  Block b of an instance takes tile b: the 256
// pair ranks of group column b % G (G = nx / 512 groups per row) in the TILE_ROWS rows from
// TILE_ROWS (b / G).  Phase 1: the block's threads form y_r (k_schur_vec's red unknowns,
// the same arithmetic) of the TILE_ROWS + 2 red rows the tile touches, four ranks each,
// into LDS, and TILE_ROWS threads the red cell just outside the tile in each of its rows
// (west or east, by the row's colours).  Phase 2: wave w forms black row
// TILE_ROWS (b / G) + w from LDS with the fused dots, one partial per 256-rank group as
// k_schur_vec stores it; w is bit-identical to k_schur_vec's.  Each red unknown is formed
// 1.125 times per tile instead of 3 times, and the neighbour values come from LDS instead
// of L2.
constexpr int TILE_ROW = 264;   // LDS values per tile row: [3] west cell, [4, 260) the tile's ranks, [260] east cell
constexpr int TILE_ROWS = 16;   // black rows per tile: one wave each
constexpr int TBS = 64 * TILE_ROWS;
// FP32: at most 64 VGPRs, so eight waves per SIMD fit: two 16-wave blocks per CU, and a
// 2048 x 2048 grid's 512 tiles run in one round on 256 CUs.
template <class T>
__global__ __launch_bounds__(TBS, sizeof(T) == 4 ? 8 : 1) void k_schur_tile(const Desc* __restrict__ D,
                                                                           const int2* __restrict__ blk, int iv,
                                                                           int iw, int nd, int a0,
                                                                           double* __restrict__ part, int ng,
                                                                           int meth) {
    const int lb = D[0].xcd ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
    const int2 bi = blk[lb];
    const Desc& d = D[bi.x];
    if (d.pcm != PM_SCHUR || !d.tile || !(meth & meth_of(d))) return;   // block-uniform
    __shared__ __align__(16) T ylds[TILE_ROWS + 2][TILE_ROW];
    const int hr = d.nx >> 1, ny = d.ny, G = hr >> 8;
    const int band = bi.y / G, col = bi.y - band * G, j0 = TILE_ROWS * band, qt = col * 256;
    const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int j = j0 + w, q0 = qt + 4 * lane, p0 = j * hr + q0;
    if (!d.act[0]) {   // block-uniform: zero partials, as the other operator kernels store them
        if (lane == 0 && j < ny) {
            const long gi = d.gbase + (p0 >> 8);
            for (int q = 0; q < nd && q < 5; ++q) part[q * (long)ng + gi] = 0.0;
        }
        return;
    }
    const T* __restrict__ v = (const T*)d.v[iv];
    const T* __restrict__ dvr = (const T*)d.dinv;
    const T* kr[5] = {nullptr, (const T*)d.K[1], (const T*)d.K[2], (const T*)d.K[3], (const T*)d.K[4]};
    const V4<T> zero{{(T)0, (T)0, (T)0, (T)0}};
    // phase 2's diagonal, operand and dot operand, loaded first so their latency overlaps
    // phase 1 (the four coefficients wait until after it: with them the kernel needs more
    // than 64 VGPRs)
    const int h = d.h, pl = j < ny ? p0 : q0;   // (rows below the grid: an in-bounds stand-in)
    const V4<T> cb = ld4((const T*)d.K[0] + h + pl), vv = ld4(v + pl), aa = ld4((const T*)d.v[a0] + pl);
    // phase 1: red rows j0 - 1 .. j0 + TILE_ROWS (window rows 0 .. TILE_ROWS + 1; zero
    // outside the grid)
    for (int c = (int)threadIdx.x; c < (TILE_ROWS + 2) * 64; c += TBS) {
        const int r = c >> 6, l = c & 63, jr = j0 - 1 + r, qc = qt + 4 * l;
        V4<T> y = zero;
        if (jr >= 0 && jr < ny) {
            const int pr = jr * hr + qc;
            const V4<T> o = voffdiag<T>(d, kr, pr, jr, qc, 0, v);
            const V4<T> dd = ld4(dvr + pr);
#pragma unroll
            for (int m = 0; m < 4; ++m) y.e[m] = o.e[m] * dd.e[m];
        }
        *reinterpret_cast<V4<T>*>(&ylds[r][4 + 4 * l]) = y;
    }
    if (threadIdx.x < TILE_ROWS) {   // the red cells outside the tile in black row j0 + threadIdx.x
        const int jb = j0 + (int)threadIdx.x;
        T yw = (T)0, ye = (T)0;
        if (jb < ny) {
            const int pb = jb * hr + qt;
            if ((jb & 1) != 0) {   // black cells at even columns: the west neighbour of the first
                if (qt > 0) yw = red_y1<T>(d, kr, dvr, v, pb - 1, jb, false);
            } else if (qt + 256 < hr) {   // odd columns: the east neighbour of the last
                ye = red_y1<T>(d, kr, dvr, v, pb + 256, jb, true);
            }
        }
        ylds[1 + threadIdx.x][3] = yw;
        ylds[1 + threadIdx.x][4 + 256] = ye;
    }
    __syncthreads();
    if (j >= ny) return;   // wave-uniform: the tile's rows below the grid
    // phase 2: black row j (window row w + 1)
    const V4<T> kw = ld4((const T*)d.K[1] + h + pl), ke = ld4((const T*)d.K[2] + h + pl);
    const V4<T> ks = ld4((const T*)d.K[3] + h + pl), kn = ld4((const T*)d.K[4] + h + pl);
    const int rj = w + 1, o0 = 4 + 4 * lane;
    const V4<T> yc = *reinterpret_cast<const V4<T>*>(&ylds[rj][o0]);
    const V4<T> ys = *reinterpret_cast<const V4<T>*>(&ylds[rj - 1][o0]);
    const V4<T> yn = *reinterpret_cast<const V4<T>*>(&ylds[rj + 1][o0]);
    const bool even = (j & 1) != 0;   // black cells at the even columns of row j
    const T ye = even ? ylds[rj][o0 - 1] : ylds[rj][o0 + 4];
    V4<T> wc;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const T xw = even ? (m == 0 ? ye : yc.e[m - 1]) : yc.e[m];
        const T xe = even ? yc.e[m] : (m == 3 ? ye : yc.e[m + 1]);
        const T o = stencil4(kw.e[m], xw, ke.e[m], xe, ks.e[m], ys.e[m], kn.e[m], yn.e[m]);
        wc.e[m] = cb.e[m] * vv.e[m] - o;
    }
    st4((T*)d.v[iw] + p0, wc);
    double s0 = dot4d(aa, wc), s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
    if (nd > 1) s1 = dot4d(wc, wc);
    if (nd > 2) {
        const V4<T> rh = ld4((const T*)d.v[V_RH] + p0);
        s2 = dot4d(rh, wc);
        s3 = dot4d(rh, aa);
        s4 = dot4d(aa, aa);
    }
    const long gi = d.gbase + (p0 >> 8);
    s0 = wave_sum(s0);
    if (nd > 1) s1 = wave_sum(s1);
    if (nd > 2) {
        s2 = wave_sum(s2);
        s3 = wave_sum(s3);
        s4 = wave_sum(s4);
    }
    if (lane == 0) {
        part[gi] = s0;
        if (nd > 1) part[ng + gi] = s1;
        if (nd > 2) {
            part[2 * (long)ng + gi] = s2;
            part[3 * (long)ng + gi] = s3;
            part[4 * (long)ng + gi] = s4;
        }
    }
}
// Full operator (jacobi / none): y = A x over both slots of each pair, fused dots as above.
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_spmv_rb(const Desc* __restrict__ D, const int2* __restrict__ blk, int ix,
                                                int iy, int nd, int a0, int a1, double* __restrict__ part, int ng,
                                                int meth) {
    LOCATE();
    // block-uniform: the reduced-system kernel, or the other method's launch, writes these partials
    if (d.ns != 2 || !(meth & meth_of(d))) return;
    double s0[CPT] = {}, s1[CPT] = {};
    if (d.act[0]) {
        const int h = d.h;
        const T* __restrict__ x = (const T*)d.v[ix];
        T* __restrict__ y = (T*)d.v[iy];
        const T* __restrict__ K0 = (const T*)d.K[0];
        const T* __restrict__ va = (const T*)d.v[a0];
        const T* __restrict__ vb = a1 >= 0 ? (const T*)d.v[a1] : nullptr;
        FOR_GROUPS(m, p) {
            if (p >= h) continue;
#pragma unroll
            for (int col = 0; col < 2; ++col) {
                const int s = col * h + p;
                int i, j, k;
                const int c = cell_of(d, p, col, i, j, k);
                const T yc = c < 0 ? (T)0 : K0[s] * x[s] + offdiag<T>(d, s, c, i, j, k, x + (1 - col) * h);
                y[s] = yc;
                s0[m] += (double)va[s] * (double)yc;
                if (nd > 1) s1[m] += (vb ? (double)vb[s] : (double)yc) * (double)yc;
            }
        }
    }
    group_store<CPT>(s0, PARTSK(part));
    if (nd > 1) group_store<CPT>(s1, PARTSK(part + ng));
}

// Preconditioner set-up (pairs): dinv[p] = 1 / C_r of red rank p (1 for PM_NONE);
// dinv[h + p] of black rank p = 1 / P_b (PM_ILU, PM_SCHUR: C_b - sum_m A_bm A_mb / C_m over
// the red neighbours m, the ILU(0) pivot = diag of the Schur complement), 1 / C_b
// (PM_JACOBI) or 1 (PM_NONE); 0 for ranks without a cell.
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_precond_setup(const Desc* __restrict__ D, const int2* __restrict__ blk) {
    LOCATE();
    if (d.vec) return;   // block-uniform (the four-ranks-per-lane form: k_setup_vec)
    const int h = d.h, nx = d.nx, ny = d.ny, nxy = d.nx * d.ny, mode = d.pcm;
    const T* K[7];
    for (int q = 0; q < 7; ++q) K[q] = (const T*)d.K[q];
    T* __restrict__ dv = (T*)d.dinv;
    FOR_PAIRS(p) {
        int i, j, k;
        int c = cell_of(d, p, 0, i, j, k);
        if (c < 0) dv[p] = (T)0;
        else {
            const T cr = K[0][p];
            dv[p] = mode == PM_NONE ? (T)1 : cr != (T)0 ? (T)1 / cr : (T)1;
        }
        c = cell_of(d, p, 1, i, j, k);
        if (c < 0) {
            dv[h + p] = (T)0;
            continue;
        }
        const int s = h + p;
        T piv = K[0][s];
        if (mode == PM_ILU || mode == PM_SCHUR) {
            // A_bm of the black row times A_mb of the red neighbour's row (opposite direction)
            if (i > 0) piv -= K[1][s] * K[2][(c - 1) >> 1] / K[0][(c - 1) >> 1];
            if (i < nx - 1) piv -= K[2][s] * K[1][(c + 1) >> 1] / K[0][(c + 1) >> 1];
            if (j > 0) piv -= K[3][s] * K[4][(c - nx) >> 1] / K[0][(c - nx) >> 1];
            if (j < ny - 1) piv -= K[4][s] * K[3][(c + nx) >> 1] / K[0][(c + nx) >> 1];
            if (d.nz > 1) {
                if (k > 0) piv -= K[5][s] * K[6][(c - nxy) >> 1] / K[0][(c - nxy) >> 1];
                if (k < d.nz - 1) piv -= K[6][s] * K[5][(c + nxy) >> 1] / K[0][(c + nxy) >> 1];
            }
        }
        dv[s] = mode == PM_NONE ? (T)1 : piv != (T)0 ? (T)1 / piv : (T)1;
    }
}

// Full-system red-black ILU(0) apply z = M^-1 r (PM_ILU instances), pass 1 over black
// ranks: z_b = (r_b - sum_m A_bm r_m / C_m) / P_b (1 / C_m of a red neighbour is its dinv).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_ilu_black(const Desc* __restrict__ D, const int2* __restrict__ blk, int ir,
                                                  int iz) {
    LOCATE();
    if (d.pcm != PM_ILU || !d.act[0] || meth_of(d) != M_BICG) return;
    const int h = d.h, nx = d.nx, ny = d.ny, nxy = d.nx * d.ny;
    const T* __restrict__ r = (const T*)d.v[ir];
    const T* __restrict__ dv = (const T*)d.dinv;
    T* __restrict__ z = (T*)d.v[iz];
    FOR_PAIRS(p) {
        int i, j, k;
        const int c = cell_of(d, p, 1, i, j, k);
        const int s = h + p;
        if (c < 0) {
            z[s] = (T)0;
            continue;
        }
        T acc = (T)0;
        if (i > 0) acc += ((const T*)d.K[1])[s] * (r[(c - 1) >> 1] * dv[(c - 1) >> 1]);
        if (i < nx - 1) acc += ((const T*)d.K[2])[s] * (r[(c + 1) >> 1] * dv[(c + 1) >> 1]);
        if (j > 0) acc += ((const T*)d.K[3])[s] * (r[(c - nx) >> 1] * dv[(c - nx) >> 1]);
        if (j < ny - 1) acc += ((const T*)d.K[4])[s] * (r[(c + nx) >> 1] * dv[(c + nx) >> 1]);
        if (d.nz > 1) {
            if (k > 0) acc += ((const T*)d.K[5])[s] * (r[(c - nxy) >> 1] * dv[(c - nxy) >> 1]);
            if (k < d.nz - 1) acc += ((const T*)d.K[6])[s] * (r[(c + nxy) >> 1] * dv[(c + nxy) >> 1]);
        }
        z[s] = (r[s] - acc) * dv[s];
    }
}
// Pass 2 over red ranks: z_r = (r_r - A_rb z_b) / C_r.
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_ilu_red(const Desc* __restrict__ D, const int2* __restrict__ blk, int ir,
                                                int iz) {
    LOCATE();
    if (d.pcm != PM_ILU || !d.act[0] || meth_of(d) != M_BICG) return;
    const T* __restrict__ r = (const T*)d.v[ir];
    const T* __restrict__ dv = (const T*)d.dinv;
    T* __restrict__ z = (T*)d.v[iz];
    FOR_PAIRS(p) {
        int i, j, k;
        const int c = cell_of(d, p, 0, i, j, k);
        z[p] = c < 0 ? (T)0 : (r[p] - offdiag<T>(d, p, c, i, j, k, z + d.h)) * dv[p];
    }
}

// The two-phase pressure Jacobian (no-flow boundary, balanced sources) is singular with
// the constants as null space (its row sums are zero).  Rounded to FP32 its row sums are
// no longer exactly zero, and a right-hand side component along the constants is then
// amplified without bound by the Krylov solve.  For these instances the right-hand side
// is projected onto the range before the solve: its mean over the cells is subtracted
// (Krylov start below).  The FP64 refinement defect is computed in full, so the
// projection only keeps the inner solves well posed.
__device__ inline bool singular(const Desc& d) { return d.kind == K_TWOPHASE && d.act[2] == 0; }
// Partial sums of the Krylov right-hand side over the slots (for the mean above).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_rhs_sum(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                double* __restrict__ part) {
    LOCATE();
    double a[CPT] = {};
    if (d.act[1] && singular(d)) {
        FOR_GROUPS(m, p) {
            if (p >= d.h) continue;
            for (int w = 0; w < d.ns; ++w) a[m] += (double)((const T*)d.v[V_R])[p + w * d.h];
        }
    }
    group_store<CPT>(a, PARTSK(part));
}

// The right preconditioner's diagonal for slot s: reduced system -> 1 / P_b of black rank
// s; full system -> dinv of slot s (jacobi) or 1 (none, dinv = 1).
template <class T>
__device__ inline const T* slot_dinv(const Desc& d) {
    return (const T*)d.dinv + (d.ns == 1 ? d.h : 0);
}

// Krylov start: r = b (already in r; minus its mean for singular systems), x = 0.
// BiCGStab: rh = r, p = v = 0; partial dots (rh.r, r.r).  CG with the diagonal: p = z =
// M^-1 r; partial dots (r.z, r.r).  CG with the V-cycle: p = 0 (z and r.z follow the
// first V-cycle); partial dots (0, r.r).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_kinit(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                              double* __restrict__ part, int ng) {
    LOCATE();
    if (d.vec) return;   // block-uniform (k_setup_vec starts these solves)
    double rr[CPT] = {}, rz[CPT] = {};
    const bool cg = meth_of(d) == M_CG;
    if (d.act[1] && cg) {
        const bool proj = singular(d);
        const T mean = proj ? (T)d.sc[S_MEAN] : (T)0;
        const T* __restrict__ dv = slot_dinv<T>(d);
        FOR_GROUPS(m, p) {
            if (p >= d.h) continue;
            for (int w = 0; w < d.ns; ++w) {
                const int s = p + w * d.h;
                T rc = ((T*)d.v[V_R])[s];
                if (proj) {
                    int i, j, k;
                    if (cell_of(d, p, d.ns == 1 ? 1 : w, i, j, k) >= 0) rc -= mean;
                    ((T*)d.v[V_R])[s] = rc;
                }
                const T z = d.mgp ? (T)0 : dv[s] * rc;
                ((T*)d.v[V_P])[s] = z;
                ((T*)d.v[V_X])[s] = 0;
                rz[m] += (double)rc * (double)z;
                rr[m] += (double)rc * (double)rc;
            }
        }
    } else if (d.act[1]) {
        const bool proj = singular(d);
        const T mean = proj ? (T)d.sc[S_MEAN] : (T)0;
        FOR_GROUPS(m, p) {
            if (p >= d.h) continue;
            for (int w = 0; w < d.ns; ++w) {
                const int s = p + w * d.h;
                T rc = ((T*)d.v[V_R])[s];
                if (proj) {   // project onto the range: subtract the mean over the slots holding a cell
                    int i, j, k;
                    if (cell_of(d, p, d.ns == 1 ? 1 : w, i, j, k) >= 0) rc -= mean;
                    ((T*)d.v[V_R])[s] = rc;
                }
                ((T*)d.v[V_RH])[s] = rc;
                ((T*)d.v[V_X])[s] = 0;
                ((T*)d.v[V_P])[s] = 0;
                ((T*)d.v[V_V])[s] = 0;
                rr[m] += (double)rc * (double)rc;
            }
        }
    }
    group_store<CPT>(cg ? rz : rr, PARTSK(part));
    group_store<CPT>(rr, PARTSK(part + ng));
}

// CG (q = A p in V_V): x += alpha p, r -= alpha q; with the diagonal, partial dots (r.z,
// r.r) with z = M^-1 r (with the V-cycle k_cg_dots writes them once z is formed).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_cg2(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                            double* __restrict__ part, int ng) {
    LOCATE();
    if (meth_of(d) != M_CG) return;   // block-uniform; the BiCGStab kernel writes these partials
    double a0[CPT] = {}, a1[CPT] = {};
    if (d.act[0]) {
        const T alpha = (T)d.sc[S_ALPHA];
        T* __restrict__ x = (T*)d.v[V_X];
        T* __restrict__ r = (T*)d.v[V_R];
        const T* __restrict__ pp = (const T*)d.v[V_P];
        const T* __restrict__ q = (const T*)d.v[V_V];
        const T* __restrict__ dv = slot_dinv<T>(d);
        FOR_GROUPS(m, p) {
            if (p >= d.h) continue;
            for (int w = 0; w < d.ns; ++w) {
                const int s = p + w * d.h;
                x[s] += alpha * pp[s];
                const T rn = r[s] - alpha * q[s];
                r[s] = rn;
                if (!d.mgp) {
                    a0[m] += (double)rn * (double)(dv[s] * rn);
                    a1[m] += (double)rn * (double)rn;
                }
            }
        }
    }
    group_store<CPT>(a0, PARTSK(part));
    group_store<CPT>(a1, PARTSK(part + ng));
}
// CG with the V-cycle on an instance whose fine level is in the workgroup tail (the whole
// V-cycle runs in k_mg_tail): partial dots (r.z, r.r), z in V_Z.
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_cg_dots(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                double* __restrict__ part, int ng) {
    LOCATE();
    // block-uniform; k_cg2 wrote the others' partials, the fine level's last half sweep
    // those of instances whose fine level is above the tail
    if (!d.mgp || !d.mgt || meth_of(d) != M_CG) return;
    double a0[CPT] = {}, a1[CPT] = {};
    if (d.act[0]) {
        const T* __restrict__ r = (const T*)d.v[V_R];
        const T* __restrict__ z = (const T*)d.v[V_Z];
        FOR_GROUPS(m, p) {
            if (p >= d.h) continue;
            for (int w = 0; w < d.ns; ++w) {
                const int s = p + w * d.h;
                a0[m] += (double)r[s] * (double)z[s];
                a1[m] += (double)r[s] * (double)r[s];
            }
        }
    }
    group_store<CPT>(a0, PARTSK(part));
    group_store<CPT>(a1, PARTSK(part + ng));
}
// CG direction: p = z + beta p, z = M^-1 r (the diagonal, or the V-cycle's z in V_Z);
// only_mg: the V-cycle instances only (their first direction, beta = 0).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_cg3(const Desc* __restrict__ D, const int2* __restrict__ blk, int only_mg) {
    LOCATE();
    if (!d.act[0] || meth_of(d) != M_CG || (only_mg && !d.mgp)) return;
    const T beta = (T)d.sc[S_BETA];
    T* __restrict__ pp = (T*)d.v[V_P];
    const T* __restrict__ r = (const T*)d.v[V_R];
    const T* __restrict__ z = (const T*)d.v[V_Z];
    const T* __restrict__ dv = slot_dinv<T>(d);
    FOR_PAIRS(p) {
        for (int w = 0; w < d.ns; ++w) {
            const int s = p + w * d.h;
            pp[s] = (d.mgp ? z[s] : dv[s] * r[s]) + beta * pp[s];
        }
    }
}

// p = r + beta (p - omega v) (first iteration: p = r); ph = M^-1 p (diagonal M; PM_ILU:
// by the ILU passes).  Strided forms (the four-ranks-per-lane form: k_kinit and k_b31).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_b1(const Desc* __restrict__ D, const int2* __restrict__ blk) {
    LOCATE();
    if (d.vec || !d.act[0] || meth_of(d) != M_BICG) return;
    const double* sc = d.sc;
    T* p = (T*)d.v[V_P];
    const bool first = sc[S_IT] == 0.0;
    const T beta = first ? (T)0 : (T)((sc[S_RHO] / sc[S_RHO_OLD]) * (sc[S_ALPHA] / sc[S_OMEGA]));
    const T omega = (T)sc[S_OMEGA];
    const T* dv = slot_dinv<T>(d);
    FOR_PAIRS(q) {
        for (int w = 0; w < d.ns; ++w) {
            const int s = q + w * d.h;
            const T r = ((const T*)d.v[V_R])[s];
            const T pv = first ? r : r + beta * (p[s] - omega * ((const T*)d.v[V_V])[s]);
            p[s] = pv;
            if (d.pcm != PM_ILU) ((T*)d.v[V_PH])[s] = dv[s] * pv;
        }
    }
}
// s = r - alpha v ; sh = M^-1 s (diagonal M; PM_ILU: by the ILU passes)
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_b2(const Desc* __restrict__ D, const int2* __restrict__ blk) {
    LOCATE();
    if (!d.act[0] || meth_of(d) != M_BICG) return;
    const T alpha = (T)d.sc[S_ALPHA];
    const T* dv = slot_dinv<T>(d);
    if (d.vec) {   // the columns carry the preconditioner (k_setup_vec): no sh
        FOR_VGROUPS(g, p0) {
            if (p0 >= d.h) continue;
            const V4<T> r = ld4((const T*)d.v[V_R] + p0), vv = ld4((const T*)d.v[V_V] + p0);
            V4<T> sv;
#pragma unroll
            for (int m = 0; m < 4; ++m) sv.e[m] = r.e[m] - alpha * vv.e[m];
            st4((T*)d.v[V_S] + p0, sv);
        }
        return;
    }
    FOR_PAIRS(q) {
        for (int w = 0; w < d.ns; ++w) {
            const int s = q + w * d.h;
            const T sv = ((const T*)d.v[V_R])[s] - alpha * ((const T*)d.v[V_V])[s];
            ((T*)d.v[V_S])[s] = sv;
            if (d.pcm != PM_ILU) ((T*)d.v[V_SH])[s] = dv[s] * sv;
        }
    }
}
// x += alpha ph + omega sh ; r = s - omega t ; partial dots (rh.r, r.r).  Strided forms.
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_b3(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                           double* __restrict__ part, int ng) {
    LOCATE();
    // block-uniform; the CG kernels write these partials, and the four-ranks-per-lane form
    // ends its iterations in k_b31 (no partials)
    if (meth_of(d) != M_BICG || d.vec) return;
    double a0[CPT] = {}, a1[CPT] = {};
    if (d.act[0]) {
        const T alpha = (T)d.sc[S_ALPHA], omega = (T)d.sc[S_OMEGA];
        T* x = (T*)d.v[V_X];
        FOR_GROUPS(m, q) {
            if (q >= d.h) continue;
            for (int w = 0; w < d.ns; ++w) {
                const int s = q + w * d.h;
                x[s] += alpha * ((const T*)d.v[V_PH])[s] + omega * ((const T*)d.v[V_SH])[s];
                const T r = ((const T*)d.v[V_S])[s] - omega * ((const T*)d.v[V_T])[s];
                ((T*)d.v[V_R])[s] = r;
                a0[m] += (double)((const T*)d.v[V_RH])[s] * (double)r;
                a1[m] += (double)r * (double)r;
            }
        }
    }
    group_store<CPT>(a0, PARTSK(part));
    group_store<CPT>(a1, PARTSK(part + ng));
}

// Four-ranks-per-lane form, the end of a BiCGStab iteration and the start of the next in
// one pass (rho, ||r||^2 and beta come from the second finaliser, A_OMEGA), on the column-
// scaled system (no preconditioner vectors: p and s are the operator's operands):
//   y += alpha p + omega s ; r = s - omega t ;
//   while the solve goes on: p = r + beta (p - omega v).
// S_XUPD says the iteration's update is pending (the instance was active when it started).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_b31(const Desc* __restrict__ D, const int2* __restrict__ blk) {
    LOCATE();
    (void)base;
    if (!d.vec || meth_of(d) != M_BICG || d.sc[S_XUPD] == 0.0) return;
    const double* sc = d.sc;
    const bool more = d.act[0] != 0;
    const T alpha = (T)sc[S_ALPHA], omega = (T)sc[S_OMEGA];
    const T beta = more ? (T)((sc[S_RHO] / sc[S_RHO_OLD]) * (sc[S_ALPHA] / sc[S_OMEGA])) : (T)0;
    T* __restrict__ x = (T*)d.v[V_X];
    T* __restrict__ r = (T*)d.v[V_R];
    T* __restrict__ p = (T*)d.v[V_P];
    FOR_VGROUPS(g, p0) {
        if (p0 >= d.h) continue;
        const V4<T> xo = ld4(x + p0), po = ld4(p + p0);
        const V4<T> sv = ld4((const T*)d.v[V_S] + p0), tv = ld4((const T*)d.v[V_T] + p0);
        V4<T> xn, rn;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            xn.e[m] = xo.e[m] + (alpha * po.e[m] + omega * sv.e[m]);
            rn.e[m] = sv.e[m] - omega * tv.e[m];
        }
        st4(x + p0, xn);
        st4(r + p0, rn);
        if (more) {
            const V4<T> vv = ld4((const T*)d.v[V_V] + p0);
            V4<T> pn;
#pragma unroll
            for (int m = 0; m < 4; ++m) pn.e[m] = rn.e[m] + beta * (po.e[m] - omega * vv.e[m]);
            st4(p + p0, pn);
        }
    }
}

// ---------------------------------------------------------------------------
// Aggregation multigrid for the two-phase pressure stage (Desc::mgp: full red-black
// system, preconditioned CG).  Level l + 1 aggregates the 2 x 2 x 2 cells of level l:
// cell (i, j, k) belongs to aggregate (i/2, j/2, k/2), so a coarse grid has
// ceil(n/2) cells per axis.  Prolongation P is piecewise constant and the coarse
// operator is the Galerkin product A_{l+1} = P^T A_l P, again a 7-point stencil: two
// neighbouring aggregates are coupled by the sum of the fine couplings across their
// interface, and the diagonal is the sum of the aggregate's fine block, so zero row
// sums (the pressure null space) carry over.  Every level is stored like the fine one:
// red-black rank order, coefficients K[q], right-hand side in V_R, correction in V_Z.
// One V-cycle computes z = M r from zero.  On each level except the coarsest: a red,
// then a black Gauss-Seidel half sweep, and the residual summed over each aggregate as
// the next level's right-hand side.  On the coarsest level (at most MG_COARSE cells):
// z = (A + s 1 1^T)^-1 b with a dense inverse formed at set-up; s 1 1^T (s = mean
// diagonal / n) removes the constant null space of the singular pressure operator, and
// for b with zero sum the result is the solution with zero sum.  Going up: z += s P
// z_coarse (s = Desc::mgs, the over-correction factor, 0 < s < 2), then a black and a
// red half sweep.  Pre- and post-smoothing are adjoint and the coarse solve is symmetric,
// so M is symmetric positive definite and CG applies.
// Desc::mgl counts the levels below a level (0 on the coarsest); the kernels of one
// level skip the instances whose coarsest level it is.
constexpr int MG_COARSE = 64;
constexpr int MG_BS = 256;   // threads of the dense coarsest-level kernels

constexpr int MG_TAIL = 4096;   // levels of at most this many cells run in one workgroup (k_mg_tail)
constexpr int TAIL_BS = 1024;

__device__ inline bool mg_on(const Desc& d) { return d.mgp && d.act[0] && meth_of(d) == M_CG; }
// red-black slot of cell c = (i, j, k) of a level
__device__ inline int rb_slot(const Desc& d, int c, int i, int j, int k) { return ((i + j + k) & 1) * d.h + (c >> 1); }

// Per-element bodies, shared by the one-level kernels and the workgroup tail.
// Gauss-Seidel update of the colour-col cell of pair p: x = (b - offdiag x) / C (zero: x
// is zero before the sweep, so the neighbours are not read); a slot without a cell gets 0.
template <class T>
__device__ inline T mg_smooth_at(const Desc& d, int p, int col, int zero) {
    const int h = d.h;
    T* __restrict__ x = (T*)d.v[V_Z];
    int i, j, k;
    const int c = cell_of(d, p, col, i, j, k);
    const int s = col * h + p;
    T xs = (T)0;
    if (c >= 0) {
        const T off = zero ? (T)0 : offdiag<T>(d, s, c, i, j, k, x + (1 - col) * h);
        xs = (((const T*)d.v[V_R])[s] - off) / ((const T*)d.K[0])[s];
    }
    x[s] = xs;
    return xs;
}
// Red residual of pair p after a red-black pass (the black residual is zero then):
// t_r = b_r - C_r x_r - A_rb x_b into V_T.
template <class T>
__device__ inline void mg_resid_at(const Desc& d, int p) {
    const T* __restrict__ x = (const T*)d.v[V_Z];
    int i, j, k;
    const int c = cell_of(d, p, 0, i, j, k);
    ((T*)d.v[V_T])[p] = c < 0 ? (T)0
                              : ((const T*)d.v[V_R])[p] - (((const T*)d.K[0])[p] * x[p] + offdiag<T>(d, p, c, i, j, k, x + d.h));
}
// Coarse cell cc: b_c = the sum of its aggregate's red residuals (fine level f).
template <class T>
__device__ inline void mg_restrict_at(const Desc& f, const Desc& c, int cc) {
    int ic, jc, kc;
    cell_ijk(c, cc, ic, jc, kc);
    const T* __restrict__ t = (const T*)f.v[V_T];
    T acc = (T)0;
    for (int k = 2 * kc; k < min(2 * kc + 2, f.nz); ++k)
        for (int j = 2 * jc; j < min(2 * jc + 2, f.ny); ++j)
            for (int i = 2 * ic + ((2 * ic + j + k) & 1); i < min(2 * ic + 2, f.nx); i += 2)   // the red cell of the row
                acc += t[(i + f.nx * (j + f.ny * k)) >> 1];
    ((T*)c.v[V_R])[rb_slot(c, cc, ic, jc, kc)] = acc;
}
// Fine pair p: z_f += z_c of each cell's aggregate.
template <class T>
__device__ inline void mg_prolong_at(const Desc& f, const Desc& c, int p) {
    T* __restrict__ xf = (T*)f.v[V_Z];
    const T* __restrict__ xc = (const T*)c.v[V_Z];
    for (int col = 0; col < 2; ++col) {
        int i, j, k;
        if (cell_of(f, p, col, i, j, k) < 0) continue;
        const int ic = i >> 1, jc = j >> 1, kc = k >> 1;
        xf[col * f.h + p] += (T)f.mgs * xc[rb_slot(c, ic + c.nx * (jc + c.ny * kc), ic, jc, kc)];
    }
}

// offdiag of a black cell whose red neighbours are read as z_r + s z_c of the neighbour's
// aggregate (level cd): the prolongation applied on the fly.  Going up, the red values
// after z += s P z_c are read only by the black half sweep that follows (the red half sweep
// after it overwrites them, and the black sweep overwrites the black ones), so the black
// sweep computes them itself: the same sums in the same order as offdiag on the
// prolonged vector.
template <class T>
__device__ inline T offdiag_prolonged(const Desc& d, const Desc& cd, int s, int c, int i, int j, int k,
                                      const T* __restrict__ x, const T* __restrict__ xc) {
    const int nx = d.nx, ny = d.ny;
    auto z = [&](int nb, int ni, int nj, int nk) -> T {
        const int ic = ni >> 1, jc = nj >> 1, kc = nk >> 1;
        return x[nb >> 1] + (T)d.mgs * xc[rb_slot(cd, ic + cd.nx * (jc + cd.ny * kc), ic, jc, kc)];
    };
    T acc = (T)0;
    if (i > 0) acc += ((const T*)d.K[1])[s] * z(c - 1, i - 1, j, k);
    if (i < nx - 1) acc += ((const T*)d.K[2])[s] * z(c + 1, i + 1, j, k);
    if (j > 0) acc += ((const T*)d.K[3])[s] * z(c - nx, i, j - 1, k);
    if (j < ny - 1) acc += ((const T*)d.K[4])[s] * z(c + nx, i, j + 1, k);
    if (d.nz > 1) {
        const int nxy = nx * ny;
        if (k > 0) acc += ((const T*)d.K[5])[s] * z(c - nxy, i, j, k - 1);
        if (k < d.nz - 1) acc += ((const T*)d.K[6])[s] * z(c + nxy, i, j, k + 1);
    }
    return acc;
}

// One-level kernels for the levels above an instance's tail (Desc::mgt = 0).
// Gauss-Seidel half sweep over the cells of colour col; dots (the fine level's last half
// sweep, red): also the CG partial dots (r.z, r.r) over both slots of each pair (r = V_R,
// z = V_Z, the black half final already).
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_mg_smooth(const Desc* __restrict__ D, const int2* __restrict__ blk, int col,
                                                  int zero, int dots, double* __restrict__ part, int ng) {
    LOCATE();
    // block-uniform; k_cg_dots covers the instances whose fine level is in the tail
    if (d.mgt || !d.mgp || meth_of(d) != M_CG) return;
    double a0[CPT] = {}, a1[CPT] = {};
    if (d.act[0]) {
        const int h = d.h;
        const T* __restrict__ r = (const T*)d.v[V_R];
        const T* __restrict__ z = (const T*)d.v[V_Z];
        FOR_GROUPS(m, p) {
            if (p >= h) continue;
            const T xs = mg_smooth_at<T>(d, p, col, zero);
            if (dots) {
                const T b = r[col * h + p];
                const int o = (1 - col) * h + p;
                a0[m] = (double)b * (double)xs + (double)r[o] * (double)z[o];
                a1[m] = (double)b * (double)b + (double)r[o] * (double)r[o];
            }
        }
    }
    if (dots) {
        group_store<CPT>(a0, PARTSK(part));
        group_store<CPT>(a1, PARTSK(part + ng));
    }
}
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_mg_resid(const Desc* __restrict__ D, const int2* __restrict__ blk) {
    LOCATE();
    if (d.mgt || !mg_on(d)) return;
    FOR_PAIRS(p) mg_resid_at<T>(d, p);
}

// Fine level F and coarse level Cd of one instance (thread per cell or pair of a level).
#define MG_LOCATE(N)                                                                               \
    const int2 bi = blk[blockIdx.x];                                                               \
    const Desc& f = F[bi.x];                                                                       \
    const Desc& c = Cd[bi.x];                                                                      \
    const int t_ = bi.y * BS + (int)threadIdx.x;                                                   \
    if (t_ >= (N)) return

// Galerkin coarse operator (thread per coarse cell), before the solve of a pressure
// Newton iteration (refinement flag set).
template <class T>
__global__ __launch_bounds__(BS) void k_mg_galerkin(const Desc* __restrict__ F, const Desc* __restrict__ Cd,
                                                    const int2* __restrict__ blk) {
    MG_LOCATE(c.n);
    if (!c.mgp || !c.act[1] || meth_of(c) != M_CG) return;
    int ic, jc, kc;
    cell_ijk(c, t_, ic, jc, kc);
    T co[7] = {0, 0, 0, 0, 0, 0, 0};
    const int nxf = f.nx, nyf = f.ny, nzf = f.nz;
    for (int k = 2 * kc; k < min(2 * kc + 2, nzf); ++k)
        for (int j = 2 * jc; j < min(2 * jc + 2, nyf); ++j)
            for (int i = 2 * ic; i < min(2 * ic + 2, nxf); ++i) {
                const int fc = i + nxf * (j + nyf * k);
                const int s = rb_slot(f, fc, i, j, k);
                co[0] += ((const T*)f.K[0])[s];
                // a coupling to a cell of the same aggregate goes to the diagonal, otherwise
                // to the direction of the neighbouring aggregate
                if (i > 0) co[(i & 1) ? 0 : 1] += ((const T*)f.K[1])[s];
                if (i < nxf - 1) co[(i & 1) ? 2 : 0] += ((const T*)f.K[2])[s];
                if (j > 0) co[(j & 1) ? 0 : 3] += ((const T*)f.K[3])[s];
                if (j < nyf - 1) co[(j & 1) ? 4 : 0] += ((const T*)f.K[4])[s];
                if (nzf > 1) {
                    if (k > 0) co[(k & 1) ? 0 : 5] += ((const T*)f.K[5])[s];
                    if (k < nzf - 1) co[(k & 1) ? 6 : 0] += ((const T*)f.K[6])[s];
                }
            }
    const int sc = rb_slot(c, t_, ic, jc, kc);
    for (int q = 0; q < 7; ++q) ((T*)c.K[q])[sc] = co[q];
}

// Restriction (thread per coarse cell), for fine levels above the tail.
template <class T>
__global__ __launch_bounds__(BS) void k_mg_restrict(const Desc* __restrict__ F, const Desc* __restrict__ Cd,
                                                    const int2* __restrict__ blk) {
    MG_LOCATE(c.n);
    if (f.mgt || !mg_on(c)) return;
    mg_restrict_at<T>(f, c, t_);
}

// Prolongation and the black half sweep after it (thread per fine pair), for fine levels
// above the tail: z_b = (b_b - A_br (z_r + P z_c)) / C_b (offdiag_prolonged); z_r keeps
// its pre-prolongation value until the red half sweep overwrites it.
template <class T>
__global__ __launch_bounds__(BS) void k_mg_prolong_black(const Desc* __restrict__ F, const Desc* __restrict__ Cd,
                                                         const int2* __restrict__ blk) {
    MG_LOCATE(f.h);
    if (f.mgt || !mg_on(c)) return;
    T* __restrict__ x = (T*)f.v[V_Z];
    int i, j, k;
    const int cell = cell_of(f, t_, 1, i, j, k);
    const int s = f.h + t_;
    T xs = (T)0;
    if (cell >= 0) {
        const T off = offdiag_prolonged<T>(f, c, s, cell, i, j, k, x, (const T*)c.v[V_Z]);
        xs = (((const T*)f.v[V_R])[s] - off) / ((const T*)f.K[0])[s];
    }
    x[s] = xs;
}
#undef MG_LOCATE

// Coarsest level set-up (one block per instance): the dense inverse of A + s 1 1^T by
// Gauss-Jordan elimination in FP64 (symmetric positive definite: no pivoting), stored in
// the Krylov precision.
template <class T>
__global__ __launch_bounds__(MG_BS) void k_mg_dense_setup(const Desc* __restrict__ Cd) {
    const Desc& c = Cd[blockIdx.x];
    if (!c.mgp || !c.act[1] || meth_of(c) != M_CG) return;
    __shared__ double A[MG_COARSE][MG_COARSE];
    __shared__ double col[MG_COARSE];
    __shared__ double sdiag;
    const int n = c.n, tid = threadIdx.x;
    // diagonal mean -> s
    if (tid == 0) {
        double sum = 0.0;
        for (int q = 0; q < n; ++q) {
            int i, j, k;
            cell_ijk(c, q, i, j, k);
            sum += (double)((const T*)c.K[0])[rb_slot(c, q, i, j, k)];
        }
        sdiag = sum / ((double)n * (double)n);
    }
    __syncthreads();
    const double sh = sdiag;
    for (int e = tid; e < n * n; e += MG_BS) A[e / n][e % n] = sh;
    __syncthreads();
    for (int q = tid; q < n; q += MG_BS) {
        int i, j, k;
        cell_ijk(c, q, i, j, k);
        const int s = rb_slot(c, q, i, j, k);
        A[q][q] += (double)((const T*)c.K[0])[s];
        const int nxy = c.nx * c.ny;
        if (i > 0) A[q][q - 1] += (double)((const T*)c.K[1])[s];
        if (i < c.nx - 1) A[q][q + 1] += (double)((const T*)c.K[2])[s];
        if (j > 0) A[q][q - c.nx] += (double)((const T*)c.K[3])[s];
        if (j < c.ny - 1) A[q][q + c.nx] += (double)((const T*)c.K[4])[s];
        if (c.nz > 1) {
            if (k > 0) A[q][q - nxy] += (double)((const T*)c.K[5])[s];
            if (k < c.nz - 1) A[q][q + nxy] += (double)((const T*)c.K[6])[s];
        }
    }
    __syncthreads();
    // in-place Gauss-Jordan inversion
    for (int m = 0; m < n; ++m) {
        const double piv = 1.0 / A[m][m];
        for (int i = tid; i < n; i += MG_BS) col[i] = A[i][m];
        __syncthreads();
        for (int j = tid; j < n; j += MG_BS) A[m][j] = (j == m ? 1.0 : A[m][j]) * piv;
        __syncthreads();
        for (int e = tid; e < n * n; e += MG_BS) {
            const int i = e / n, j = e % n;
            if (i != m) A[i][j] = (j == m ? 0.0 : A[i][j]) - col[i] * A[m][j];
        }
        __syncthreads();
    }
    T* __restrict__ inv = (T*)c.minv;
    for (int e = tid; e < n * n; e += MG_BS) inv[e] = (T)A[e / n][e % n];
}

// The tail of the V-cycle (one workgroup per instance): levels [first, first + count) of
// the instance, top first (its levels of at most MG_TAIL cells; the last one is the
// coarsest, at most MG_COARSE cells).  The top level's right-hand side is in place; the
// workgroup runs the descent, the dense coarsest solve and the ascent, with a workgroup
// barrier between stages (the data of these levels stays in L2).  On return the top
// level's V_Z holds its correction.
template <class T>
__global__ __launch_bounds__(TAIL_BS) void k_mg_tail(const Desc* __restrict__ L, const int2* __restrict__ range) {
    const int2 rg = range[blockIdx.x];
    if (!mg_on(L[rg.x])) return;   // uniform over the workgroup
    const int tid = threadIdx.x, top = rg.x, last = rg.x + rg.y - 1;
    for (int q = top; q < last; ++q) {
        const Desc& d = L[q];
        for (int p = tid; p < d.h; p += TAIL_BS) mg_smooth_at<T>(d, p, 0, 1);
        __syncthreads();
        for (int p = tid; p < d.h; p += TAIL_BS) mg_smooth_at<T>(d, p, 1, 0);
        __syncthreads();
        for (int p = tid; p < d.h; p += TAIL_BS) mg_resid_at<T>(d, p);
        __syncthreads();
        const Desc& c = L[q + 1];
        for (int cc = tid; cc < c.n; cc += TAIL_BS) mg_restrict_at<T>(d, c, cc);
        __syncthreads();
    }
    {   // coarsest: z = inverse * b
        const Desc& c = L[last];
        __shared__ double bb[MG_COARSE];
        const int n = c.n;
        const T* __restrict__ b = (const T*)c.v[V_R];
        T* __restrict__ x = (T*)c.v[V_Z];
        for (int q = tid; q < n; q += TAIL_BS) {
            int i, j, k;
            cell_ijk(c, q, i, j, k);
            bb[q] = (double)b[rb_slot(c, q, i, j, k)];
        }
        __syncthreads();
        const T* __restrict__ inv = (const T*)c.minv;
        for (int q = tid; q < n; q += TAIL_BS) {
            double acc = 0.0;
            for (int e = 0; e < n; ++e) acc += (double)inv[q * n + e] * bb[e];
            int i, j, k;
            cell_ijk(c, q, i, j, k);
            x[rb_slot(c, q, i, j, k)] = (T)acc;
        }
        if ((n & 1) && tid == 0) {   // the last pair holds one cell; the other colour's slot stays 0
            int i, j, k;
            cell_ijk(c, n - 1, i, j, k);
            x[(1 - ((i + j + k) & 1)) * c.h + (n >> 1)] = (T)0;
        }
        __syncthreads();
    }
    for (int q = last - 1; q >= top; --q) {
        const Desc& d = L[q];
        for (int p = tid; p < d.h; p += TAIL_BS) mg_prolong_at<T>(d, L[q + 1], p);
        __syncthreads();
        for (int p = tid; p < d.h; p += TAIL_BS) mg_smooth_at<T>(d, p, 1, 0);
        __syncthreads();
        for (int p = tid; p < d.h; p += TAIL_BS) mg_smooth_at<T>(d, p, 0, 0);
        __syncthreads();
    }
}

// Copies of the step start (u0 = u, sw0 = sw) of many instances in one launch: job q
// copies n doubles and owns the blocks from b0 on (MC_EPB elements each); bj maps a block
// to its job.
struct McJob {
    double* dst;
    const double* src;
    long n, b0;
};
constexpr int MC_EPT = 8;
constexpr long MC_EPB = (long)BS * MC_EPT;
__global__ __launch_bounds__(BS) void k_multi_copy(const McJob* __restrict__ jobs, const int* __restrict__ bj) {
    const McJob j = jobs[bj[blockIdx.x]];
    const long base = ((long)blockIdx.x - j.b0) * MC_EPB + threadIdx.x;
#pragma unroll
    for (int e = 0; e < MC_EPT; ++e) {
        const long i = base + (long)e * BS;
        if (i < j.n) j.dst[i] = j.src[i];
    }
}

// out[0] = number of instances with flag `which` set, out[1] = the largest Krylov
// iteration count of the current solve (S_IT) over the instances.
__global__ void k_count_active(const Desc* __restrict__ D, int ninst, int which, int* __restrict__ out,
                               int* __restrict__ gate) {
    __shared__ int lds[2][BS];
    int s = 0, mx = 0;
    for (int i = threadIdx.x; i < ninst; i += BS) {
        s += D[i].act[which] ? 1 : 0;
        mx = max(mx, (int)D[i].sc[S_IT]);
    }
    lds[0][threadIdx.x] = s;
    lds[1][threadIdx.x] = mx;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            lds[0][threadIdx.x] += lds[0][threadIdx.x + w];
            lds[1][threadIdx.x] = max(lds[1][threadIdx.x], lds[1][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = lds[0][0];
        out[1] = lds[1][0];
        if (gate) *gate = lds[0][0] == 0 ? 1 : 0;   // the post-solve work runs once every instance converged
    }
}

// refinement bookkeeping (one thread per instance):
//   mode 0: start  -> act[1] = ||F|| > 0; S_BN = S_RN = ||F||; counters reset (skipped while
//           S_CONT marks a refinement that continues from the previous round)
//   (after a defect, the finaliser's A_REFINE: S_RN = sqrt(S_FN); act[1] = S_RN > S_LTOL
//   S_BN && S_REF < max_refine and the defect decreased (stagnation stops the refinement);
//   S_CONT = act[1]: the Newton update waits for the next round)
struct Forcing {
    double rel, abs, forcing, eta_max;
    double forcing0;   // first iteration of a step: forcing0 * (||F_1|| / ||F_0|| of the latest step)
    int mixed;
};
__global__ void k_refine_flags(const Desc* __restrict__ D, int ninst, int mode, double tol, int max_refine,
                               Forcing fc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    const Desc& d = D[i];
    double* sc = d.sc;
    if (mode == 0) {
        if (sc[S_CONT] != 0.0) return;   // refinement of the current Newton iteration continues
        sc[S_BN] = sqrt(sc[S_FN]);
        sc[S_RN] = sc[S_BN];
        sc[S_REF] = 0.0;
        sc[S_ITT] = 0.0;
        // (a Newton instance whose ||F|| meets abs_res_tol solves nothing: a step start that
        // runs its first solve before the host has checked ||F_0|| leaves u as it is)
        d.act[1] = sc[S_BN] > 0.0 && isfinite(sc[S_BN]) && (d.kind == K_LINEAR || sc[S_BN] > fc.abs);
        // this Newton iteration's linear tolerance (inexact Newton; FusedConfig::forcing)
        double eta = tol;
        if (fc.forcing > 0.0 && d.kind != K_LINEAR) {
            const int r1 = d.kind == K_TWOPHASE && d.act[2] == 1 ? S_R1B : S_R1;
            if (sc[S_F0] <= 0.0) {   // the step's (stage's) first iteration
                sc[S_F0] = sc[S_FN];
                sc[S_NK] = 0.0;
                // no tighter than the nonlinear error the iteration will leave, as the latest
                // step's first iteration left it
                if (fc.forcing0 > 0.0 && sc[r1] > 0.0) eta = fmin(fc.eta_max, fmax(tol, fc.forcing0 * sc[r1]));
            } else {
                sc[S_NK] += 1.0;
                if (sc[S_NK] == 1.0 && sc[S_F0] > 0.0) sc[r1] = sqrt(sc[S_FN] / sc[S_F0]);
                const double fn = sc[S_BN], target = fmax(fc.abs, fc.rel > 0.0 ? fc.rel * sqrt(sc[S_F0]) : 0.0);
                if (fn > 0.0 && target > 0.0 && isfinite(fn))
                    eta = fmin(fc.eta_max, fmax(tol, fc.forcing * target / fn));
            }
        }
        sc[S_LTOL] = eta;
        sc[S_TOL2] = fc.mixed ? fmax(sc[S_ITOL], eta * eta) : eta * eta;
    }
}

// Right-hand side of the Krylov solve from the refinement residual: F in the first round
// (du = 0), r64 = F - J du after it; scaled by 1 / ||.|| in mixed precision,
// b = (T)(src * scale).  Reduced system -> g_b = b_b - A_br (C_r^-1 b_r) on black ranks;
// full system -> b in both slots of each pair (0 where a rank has no cell).
template <class T>
__device__ inline T rhs_value(const double* __restrict__ src, int c, double scale) {
    return (T)(src[c] * scale);
}
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_load_rhs(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                 int scaled, int first) {
    LOCATE();
    if (!d.act[1] || d.vec) return;   // (the four-ranks-per-lane form: k_setup_vec)
    const double scale = scaled ? 1.0 / d.sc[S_RN] : 1.0;
    const bool fst = first < 0 ? d.sc[S_REF] == 0.0 : first != 0;   // first < 0: the instance's own round
    const double* __restrict__ src = fst ? d.F : d.r64;
    const int h = d.h, nx = d.nx, ny = d.ny, nxy = d.nx * d.ny;
    T* __restrict__ r = (T*)d.v[V_R];
    const T* __restrict__ dv = (const T*)d.dinv;
    FOR_PAIRS(p) {
        int i, j, k;
        if (d.ns == 2) {
            int c = cell_of(d, p, 0, i, j, k);
            r[p] = c < 0 ? (T)0 : rhs_value<T>(src, c, scale);
            c = cell_of(d, p, 1, i, j, k);
            r[h + p] = c < 0 ? (T)0 : rhs_value<T>(src, c, scale);
            continue;
        }
        const int c = cell_of(d, p, 1, i, j, k);
        if (c < 0) {
            r[p] = (T)0;
            continue;
        }
        const int s = h + p;
        const T* K[7];
        for (int q = 1; q < 7; ++q) K[q] = (const T*)d.K[q];
        T acc = (T)0;
        if (i > 0) acc += K[1][s] * (rhs_value<T>(src, c - 1, scale) * dv[(c - 1) >> 1]);
        if (i < nx - 1) acc += K[2][s] * (rhs_value<T>(src, c + 1, scale) * dv[(c + 1) >> 1]);
        if (j > 0) acc += K[3][s] * (rhs_value<T>(src, c - nx, scale) * dv[(c - nx) >> 1]);
        if (j < ny - 1) acc += K[4][s] * (rhs_value<T>(src, c + nx, scale) * dv[(c + nx) >> 1]);
        if (d.nz > 1) {
            if (k > 0) acc += K[5][s] * (rhs_value<T>(src, c - nxy, scale) * dv[(c - nxy) >> 1]);
            if (k < d.nz - 1) acc += K[6][s] * (rhs_value<T>(src, c + nxy, scale) * dv[(c + nxy) >> 1]);
        }
        r[p] = rhs_value<T>(src, c, scale) - acc;
    }
}
// Four-ranks-per-lane form, the start of a solve in one pass (k_precond_setup, k_load_rhs
// and k_kinit of the strided forms): a lane takes the four red and the four black cells of
// pair ranks p0 .. p0 + 3 of row j and forms
//   1 / C_r of its red cells and, in the first refinement round of a Newton iteration
//     (scale_cols), 1 / P_b of its black ones, P_b = C_b - sum over the red neighbours m of
//     A_bm A_mb / C_m (the red-black ILU(0) pivot, the neighbours' 1 / C_m formed from
//     their C_m as their own lanes form them), and then the black columns scaled by 1 / P_b
//     in place: C_b and the red coefficients A_mb.  The Krylov solve then runs on S D
//     (D = diag(1 / P_b)), the right-preconditioned reduced system, without a
//     preconditioner vector: its iterate y gives x_b = D y (k_accum);
//   g_b = b_b - sum_m A_bm (b_m / C_m), b = (T)(src scale) (src = F in the first refinement
//     round, r64 after it; scale = 1 / ||src|| in mixed precision; A_bm is the black row's,
//     which the scaling leaves alone);
//   r = rh = p = g, y = 0, and the partial dots (r.r, r.r) for A_INIT.
template <class T>
__device__ inline T inv_or_one(T c) {
    return c != (T)0 ? (T)1 / c : (T)1;
}
template <class T, int CPT>
__global__ __launch_bounds__(BS) void k_setup_vec(const Desc* __restrict__ D, const int2* __restrict__ blk,
                                                  double* __restrict__ part, int ng, int scaled, int first,
                                                  int scale_cols) {
    LOCATE();
    (void)base;
    if (!d.vec) return;   // block-uniform
    // the first refinement round of a Newton iteration (the coefficients as assembled):
    // pivots, then the columns scaled; later rounds reuse both
    const bool cols = scale_cols && d.sc[S_REF] == 0.0;
    const bool on = d.act[1] != 0;
    const double scale = scaled ? 1.0 / d.sc[S_RN] : 1.0;
    const bool fst = first < 0 ? d.sc[S_REF] == 0.0 : first != 0;   // first < 0: the instance's own round
    const double* __restrict__ src = fst ? d.F : d.r64;
    const int h = d.h, nx = d.nx, ny = d.ny, hr = nx >> 1;
    const T* __restrict__ K0 = (const T*)d.K[0];
    const T* __restrict__ K1 = (const T*)d.K[1];
    const T* __restrict__ K2 = (const T*)d.K[2];
    const T* __restrict__ K3 = (const T*)d.K[3];
    const T* __restrict__ K4 = (const T*)d.K[4];
    T* __restrict__ dv = (T*)d.dinv;
    T* __restrict__ K0w = (T*)d.K[0];
    T* __restrict__ K1w = (T*)d.K[1];
    T* __restrict__ K2w = (T*)d.K[2];
    T* __restrict__ K3w = (T*)d.K[3];
    T* __restrict__ K4w = (T*)d.K[4];
    FOR_VGROUPS(g, p0) {
        double rr = 0.0;
        if (p0 < h) {
            int j, q0;
            vrow(d, p0, j, q0);
            const bool evb = (j & 1) != 0;   // black cells at the even columns of row j
            // red cells of row j at ranks p0 .. p0 + 3, and the one just outside (west of the
            // first black cell when evb, east of the last otherwise; absent at the row's end)
            const V4<T> cr = ld4(K0 + p0), kwr = ld4(K1 + p0), ker = ld4(K2 + p0);
            const int pe = evb ? p0 - 1 : p0 + 4;
            const bool has_e = evb ? q0 > 0 : q0 + 4 < hr;
            const T cre = has_e ? K0[pe] : (T)1, kwe = has_e ? K1[pe] : (T)0, kee = has_e ? K2[pe] : (T)0;
            V4<T> dr;
#pragma unroll
            for (int m = 0; m < 4; ++m) dr.e[m] = inv_or_one(cr.e[m]);
            st4(dv + p0, dr);
            // red cells of rows j - 1 and j + 1 at the same ranks (their N / S coefficients
            // couple them to row j's black cells)
            const V4<T> one{{(T)1, (T)1, (T)1, (T)1}}, zero{{(T)0, (T)0, (T)0, (T)0}};
            const V4<T> crs = j > 0 ? ld4(K0 + p0 - hr) : one, kns = j > 0 ? ld4(K4 + p0 - hr) : zero;
            const V4<T> crn = j < ny - 1 ? ld4(K0 + p0 + hr) : one, ksn = j < ny - 1 ? ld4(K3 + p0 + hr) : zero;
            const V4<T> cb = ld4(K0 + h + p0), kw = ld4(K1 + h + p0), ke = ld4(K2 + h + p0);
            const V4<T> ks = ld4(K3 + h + p0), kn = ld4(K4 + h + p0);
            V4<T> db;
            if (cols) {
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    // the black cell's red W / E neighbours: (C, the coefficient towards the black cell)
                    const T cW = evb ? (m == 0 ? cre : cr.e[m > 0 ? m - 1 : 0]) : cr.e[m];
                    const T aW = evb ? (m == 0 ? kee : ker.e[m > 0 ? m - 1 : 0]) : ker.e[m];
                    const T cE = evb ? cr.e[m] : (m == 3 ? cre : cr.e[m < 3 ? m + 1 : 3]);
                    const T aE = evb ? kwr.e[m] : (m == 3 ? kwe : kwr.e[m < 3 ? m + 1 : 3]);
                    T piv = cb.e[m];
                    piv -= kw.e[m] * aW / cW;
                    piv -= ke.e[m] * aE / cE;
                    piv -= ks.e[m] * kns.e[m] / crs.e[m];
                    piv -= kn.e[m] * ksn.e[m] / crn.e[m];
                    db.e[m] = inv_or_one(piv);
                }
                st4(dv + h + p0, db);
                // right preconditioning by column scaling: the columns of the black cells (C_b
                // and the four red coefficients towards each black cell) times 1 / P_b.  Each
                // of those red coefficients couples its red cell to one black cell only, whose
                // lane alone reads it (for the pivot above) and writes it here.
                V4<T> cbs, kwo, keo, kso, kno;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    cbs.e[m] = cb.e[m] * db.e[m];
                    kso.e[m] = kns.e[m] * db.e[m];   // red S neighbour's N coefficient
                    kno.e[m] = ksn.e[m] * db.e[m];   // red N neighbour's S coefficient
                }
                st4(K0w + h + p0, cbs);
                if (j > 0) st4(K4w + p0 - hr, kso);
                if (j < ny - 1) st4(K3w + p0 + hr, kno);
                // row j.  evb: red rank p0 + m (an odd column) couples by its W coefficient to
                // black m and by its E coefficient to black m + 1, so this lane owns K1 of red
                // p0 .. p0 + 3 and K2 of red p0 - 1 .. p0 + 2.  Otherwise red p0 + m (an even
                // column) couples by its E coefficient to black m and by its W one to black
                // m - 1: this lane owns K2 of red p0 .. p0 + 3 and K1 of red p0 + 1 .. p0 + 4.
                // (Each lane read these values above; the ones it loaded but does not own are
                // not used.)
                if (evb) {
#pragma unroll
                    for (int m = 0; m < 4; ++m) kwo.e[m] = kwr.e[m] * db.e[m];
                    st4(K1w + p0, kwo);
                    if (has_e) K2w[p0 - 1] = kee * db.e[0];
                    K2w[p0] = ker.e[0] * db.e[1];
                    K2w[p0 + 1] = ker.e[1] * db.e[2];
                    K2w[p0 + 2] = ker.e[2] * db.e[3];
                } else {
#pragma unroll
                    for (int m = 0; m < 4; ++m) keo.e[m] = ker.e[m] * db.e[m];
                    st4(K2w + p0, keo);
                    K1w[p0 + 1] = kwr.e[1] * db.e[0];
                    K1w[p0 + 2] = kwr.e[2] * db.e[1];
                    K1w[p0 + 3] = kwr.e[3] * db.e[2];
                    if (has_e) K1w[p0 + 4] = kwe * db.e[3];
                }
            }
            if (on) {
                // right-hand side: b at the row's eight cells, the red cells above and below,
                // and the outside red cell
                const int c0 = 2 * p0;
                double sc[8], ss[8], sn[8];
                ld8(src + c0, sc);
                ld8(src + (j > 0 ? c0 - nx : c0), ss);
                ld8(src + (j < ny - 1 ? c0 + nx : c0), sn);
                const T be = has_e ? (T)(src[evb ? c0 - 1 : c0 + 8] * scale) : (T)0;
                V4<T> gv;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int bc = evb ? 2 * m : 2 * m + 1;   // the black cell's place among the eight
                    const T bW = evb ? (m == 0 ? be : (T)(sc[bc > 0 ? bc - 1 : 0] * scale)) : (T)(sc[bc - 1] * scale);
                    const T bE = evb ? (T)(sc[bc + 1] * scale) : (m == 3 ? be : (T)(sc[bc < 7 ? bc + 1 : 7] * scale));
                    const T cW = evb ? (m == 0 ? cre : cr.e[m > 0 ? m - 1 : 0]) : cr.e[m];
                    const T cE = evb ? cr.e[m] : (m == 3 ? cre : cr.e[m < 3 ? m + 1 : 3]);
                    const T yW = bW * inv_or_one(cW), yE = bE * inv_or_one(cE);
                    const T yS = (T)(ss[bc] * scale) * inv_or_one(crs.e[m]);
                    const T yN = (T)(sn[bc] * scale) * inv_or_one(crn.e[m]);
                    gv.e[m] = (T)(sc[bc] * scale) - stencil4(kw.e[m], yW, ke.e[m], yE, ks.e[m], yS, kn.e[m], yN);
                }
                st4((T*)d.v[V_R] + p0, gv);
                st4((T*)d.v[V_RH] + p0, gv);
                st4((T*)d.v[V_P] + p0, gv);
                st4((T*)d.v[V_X] + p0, zero);
                rr = dot4d(gv, gv);
            }
        }
        vgroup_store(d, bi, CPT, g, rr, part);
        vgroup_store(d, bi, CPT, g, rr, part + ng);
    }
}

// du (+)= ||.|| x (mixed) or du = x (fp64), in cell order; the first round writes du
// (every instance of the launch: 0 where the refinement is inactive), later rounds add.
// Reduced system: the red unknowns are recovered first, x_r = C_r^-1 (b_r - A_rb x_b).
// A solve whose last residual is not finite (breakdown, overflow) contributes nothing:
// the refinement defect then shows no progress and the refinement stops.
// (FP32 up to four ranks per lane: eight waves per SIMD, as k_schur_tile.)
template <class T, int CPT>
__global__ __launch_bounds__(BS, sizeof(T) == 4 && CPT <= 4 ? 8 : 1) void k_accum(const Desc* __restrict__ D,
                                                                                  const int2* __restrict__ blk,
                                                                                  int scaled, int first) {
    LOCATE();
    if (gated_off(d)) return;
    const bool first_round = first < 0 ? d.sc[S_REF] == 0.0 : first != 0;   // first < 0: the instance's own round
    const bool ok = d.act[1] && isfinite(d.sc[S_RR]);
    if (!ok && !first_round) return;
    const double rn = d.sc[S_RN];
    const double scale = scaled ? 1.0 / rn : 1.0;
    const double* __restrict__ src = first_round ? d.F : d.r64;
    const int h = d.h;
    const T* __restrict__ x = (const T*)d.v[V_X];
    const T* __restrict__ dv = (const T*)d.dinv;
    double* __restrict__ du = d.du;
    auto put = [&](int c, T xc) {
        const double v = scaled ? rn * (double)xc : (double)xc;
        if (first_round) du[c] = ok ? v : 0.0;
        else du[c] += v;
    };
    if (d.vec) {   // heat, reduced form: lanes of four pairs (cells 2 p0 .. 2 p0 + 7 of row j)
        const T* kq[5] = {nullptr, (const T*)d.K[1], (const T*)d.K[2], (const T*)d.K[3], (const T*)d.K[4]};
        FOR_VGROUPS(g, p0) {
            if (p0 >= h) continue;
            const int c0 = 2 * p0;
            double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (ok) {
                int j, q0;
                vrow(d, p0, j, q0);
                const bool evr = (j & 1) == 0;   // red cells at the even columns of the row
                // red unknowns recovered: x_r = C_r^-1 (b_r - A_rb x_b)
                // the Krylov iterate is y of the column-scaled system (k_setup_vec): the red
                // coefficients towards black cells are scaled, so A_rb x_b = (scaled A_rb) y,
                // and x_b = y / P_b
                const V4<T> o = voffdiag<T>(d, kq, p0, j, q0, 0, x);
                const V4<T> dc = ld4(dv + p0), yb = ld4(x + p0), db = ld4(dv + h + p0);
                double sc[8];
                ld8(src + c0, sc);
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int rc = evr ? 2 * m : 2 * m + 1;
                    const T xr = ((T)(sc[rc] * scale) - o.e[m]) * dc.e[m];
                    const T xb = yb.e[m] * db.e[m];
                    v[rc] = scaled ? rn * (double)xr : (double)xr;
                    v[rc ^ 1] = scaled ? rn * (double)xb : (double)xb;
                }
            }
            V4<double> a{{v[0], v[1], v[2], v[3]}}, b{{v[4], v[5], v[6], v[7]}};
            if (!first_round) {
                const V4<double> a0 = ld4(du + c0), b0 = ld4(du + c0 + 4);
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    a.e[m] += a0.e[m];
                    b.e[m] += b0.e[m];
                }
            }
            st4(du + c0, a);
            st4(du + c0 + 4, b);
        }
        return;
    }
    FOR_PAIRS(p) {
        int i, j, k;
        const int cr = cell_of(d, p, 0, i, j, k);
        if (cr >= 0)
            put(cr, !ok ? (T)0
                        : d.ns == 2 ? x[p]
                                    : (rhs_value<T>(src, cr, scale) - offdiag<T>(d, p, cr, i, j, k, x)) * dv[p]);
        const int cb = cell_of(d, p, 1, i, j, k);
        if (cb >= 0) put(cb, ok ? x[(d.ns == 2 ? h : 0) + p] : (T)0);
    }
}
// u -= du for the Newton unknown: heat u; two-phase p (stage 0) or sw (stage 1)
template <int CPT>
__global__ __launch_bounds__(BS) void k_newton_update(const Desc* __restrict__ D, const int2* __restrict__ blk) {
    LOCATE();
    if (gated_off(d) || d.sc[S_CONT] != 0.0) return;   // the refinement continues: no update yet
    if (d.vec) {   // heat, four consecutive cells per lane
        FOR_VGROUPS(g, c0) {
            if (c0 >= d.n) continue;
            V4<double> a = ld4(d.u + c0);
            const V4<double> b = ld4(d.du + c0);
#pragma unroll
            for (int m = 0; m < 4; ++m) a.e[m] -= b.e[m];
            st4(d.u + c0, a);
        }
        return;
    }
    double* x = d.kind == K_TWOPHASE ? (d.act[2] == 0 ? d.p : d.sw) : d.u;
    FOR_CELLS(c) x[c] -= d.du[c];
}
// out[q * ninst + i] = scalar slots.x, .y, .z, .w, s4, s5 (q = 0..5) of instance i, into
// pinned host memory; one block, which then stores seq into *flag (pinned too) after a
// system-scope fence, so the host can wait for the values by watching the flag.
constexpr int GATHER_BS = 1024;
__global__ __launch_bounds__(GATHER_BS) void k_gather_slots(const Desc* __restrict__ D, int ninst, int4 slots,
                                                             int s4, int s5, double* __restrict__ out,
                                                             int* __restrict__ flag, int seq) {
    for (int i = (int)threadIdx.x; i < ninst; i += (int)blockDim.x) {
        if (gated_off(D[i])) continue;
        const double* sc = D[i].sc;
        out[i] = sc[slots.x];
        out[ninst + i] = sc[slots.y];
        out[2 * ninst + i] = sc[slots.z];
        out[3 * ninst + i] = sc[slots.w];
        out[4 * ninst + i] = sc[s4];
        out[5 * ninst + i] = sc[s5];
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}
#undef FOR_GROUPS
#undef FOR_CELLS
#undef LOCATE
#undef PARTS
#undef PARTSK
#undef FOR_PAIRS

// ---------------------------------------------------------------------------
struct Inst {
    int id = 0;
    int kind = K_HEAT;
    int nx = 0, ny = 0, nz = 1, n = 0;
    // heat instances in the four-ranks-per-lane form store rows padded to a multiple of 8
    // cells (nx = the padded stride, nxr = the real row length; 0: not padded).  A padding
    // cell has an identity row and zero right-hand side and is coupled to nothing, so its
    // Krylov components stay exactly zero; the state crosses the boundary without padding.
    int nxr = 0;
    double k = 0, cv = 1, dx = 1, dy = 1, dt = 0;
    int real_nx() const { return nxr ? nxr : nx; }
    int real_n() const { return real_nx() * ny * nz; }
    // two-phase
    double mw_inv = 0, mo_inv = 0;
    int corey = 1;
    int stage = 0;                 // 0 pressure, 1 transport
    double r0_stage = 0;           // ||F|| at the start of the current Newton solve
    int max_newton = 0;            // per-slot Newton limit per solve (0: the batch's max_iter)
    int round = 0;                 // refinement round of the current Newton iteration (host count)
    int newton_prev = 0;           // Newton iterations of the instance's previous step (0: none yet)
    bool u0_pending = false;       // heat: u0 = u is left to the step's first assembly (S_U0D = 0)
    DevPtr<double> tp;             // tx, ty, tz, pv, src (n each)
    DevPtr<double> p, sw, sw0;
    DevPtr<double> u, u0, F, du, r64, C, W, E, S, N, B, T, sc;
    DevPtr<char> kcoef, dinv, vecs;   // Krylov-precision coefficients, preconditioner, vectors
    DevPtr<int> act;
    enum State { IDLE, START, SOLVE, DONE } state = IDLE;
    StepStats st;
    int pcm = PM_SCHUR;            // Krylov form / preconditioner (storage is sized for it)
    // two-phase in the full form (multigrid pressure stage): the transport stage runs in the
    // reduced form (the full form's storage holds it)
    bool split = false;
    int pcm_now() const { return split && stage == 1 ? PM_SCHUR : pcm; }
    // multigrid hierarchy (two-phase pressure by CG with the V-cycle): coarse levels 1, 2,
    // ... (coefficients 7 x 2 h, vectors V_R, V_Z, V_T 2 h each, Krylov precision) and the
    // coarsest level's dense inverse (the fine level's when it is already small)
    struct Level {
        int nx = 1, ny = 1, nz = 1, n = 0;
        DevPtr<char> K, v;
        int h() const { return (n + 1) / 2; }
    };
    bool mg_on = false;
    std::vector<Level> mg;
    DevPtr<char> minv;
    int mg_tail = 0;               // first level (0 = fine) of at most MG_TAIL cells: k_mg_tail runs it and below
    int ncoef() const { return nz > 1 ? 7 : 5; }
    int h() const { return (n + 1) / 2; }
    int ns() const { return pcm == PM_SCHUR ? 1 : 2; }
};

// Kernel classes; each has its own cells per thread: assembly, refinement defect, the
// Krylov operator, the cell-order <-> red-black conversions (preconditioner set-up,
// right-hand side, solution update) and the Krylov vector updates.
enum Cls { C_ASM = 0, C_DEF, C_SPMV, C_PC, C_VEC, NCLS };

// Launch shape of a table: cells per thread per kernel class and the XCD-contiguous
// block order (one for the whole table).
struct Shape {
    int cpt[NCLS] = {4, 4, 4, 4, 4};
    int xcd = 1;
};
constexpr int CPT_MAX = 8;
inline int cpt_index(int cpt) { return cpt == 1 ? 0 : cpt == 2 ? 1 : cpt == 4 ? 2 : 3; }

// Calls f(std::integral_constant<int, CPT>) for a cells-per-thread value.
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
    DevPtr<int2> blk[4];     // cell kernels: block -> (instance, local block) for 1, 2, 4, 8 cells per thread
    int nb[4] = {};          // blocks of each map
    DevPtr<int2> pblk[4];    // pair kernels (Krylov space): the same over each instance's h pair ranks
    int pnb[4] = {};
    DevPtr<int2> tblk;       // tiled operator: block -> (instance, tile) (k_schur_tile)
    int tnb = 0;
    int ng = 0;              // partial groups over all instances (an instance has ceil(n / BS))
    bool has_schur = false, has_full = false, has_ilu = false;   // Krylov forms present
    bool has_tp = false;                                         // two-phase instances present
    // instances whose Krylov kernels run in the four-ranks-per-lane form, and the others
    bool has_vec = false, has_nonvec = false;
    bool has_schur_strided = false;   // reduced-form instances in the strided form
    bool has_tile = false;            // four-ranks-per-lane instances with the tiled operator
    bool has_vec_onepass = false;     // four-ranks-per-lane instances with the one-pass operator
    // Multigrid plan over the instances with a hierarchy; index l = 1 .. levels - 1: the
    // instances that have level l, their level l - 1 (fine) and level l (coarse)
    // descriptors, blocks over the coarse cells, the fine pairs and the coarse pairs.
    struct Mg {
        int levels = 0;
        std::vector<DevPtr<Desc>> fine, coarse;
        std::vector<DevPtr<int2>> cblk, fpblk, cpblk;
        std::vector<int> ncb, nfpb, ncpb;
        DevPtr<Desc> coarsest;   // each instance's coarsest level
        int ncoarsest = 0;
        // workgroup tails: each instance's levels from its tail top down (flat), and per
        // instance the (first, count) range
        DevPtr<Desc> tail;
        DevPtr<int2> trange;
        int ntail = 0;
        std::vector<char> sep;   // level l has one-level launches (some instance has it above its tail)
        bool tail0 = false;      // an instance whose fine level is in its tail (dots by k_cg_dots)
    };
    std::unique_ptr<Mg> mg;
    Shape shape;
    DevPtr<double> part;
    uint64_t version = 0;
    int solves = 0;   // Krylov solves run on this membership (graphs are captured from the second)
    // captured BiCGStab iteration, and the iterations the last solve needed, keyed by
    // (tolerance, max iterations)
    // keyed by (tolerance, iteration cap, Krylov methods present); the predictions also by
    // the position of the solve in its time step (Newton iteration, refinement round), whose
    // iteration counts repeat from step to step
    std::map<std::tuple<double, int, int>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
    std::map<std::tuple<double, int, int, int>, int> pred;
    // whole speculative Newton iterations, keyed by (blind Krylov iterations, methods + 8 x
    // norm-only assembly + 16 x step start, Krylov precision, the pinned gather buffer)
    std::map<std::tuple<int, int, int, uintptr_t>, std::pair<hipGraph_t, hipGraphExec_t>> newton_graphs;
    ~Table() {
        // a replay's last kernel may still be retiring after the host saw its gather flag:
        // no executable graph is destroyed before the stream has drained
        if (!graphs.empty() || !newton_graphs.empty()) (void)hipStreamSynchronize(S());
        for (auto* m : {&graphs})
            for (auto& g : *m) {
                (void)hipGraphExecDestroy(g.second.second);
                (void)hipGraphDestroy(g.second.first);
            }
        for (auto& g : newton_graphs) {
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
    static constexpr int NRING = 4;
    // NRING x {active count, max Krylov iterations} and the gathered per-instance scalars,
    // in pinned host memory the kernels store into directly (visible once the stream or the
    // event has completed): a host check costs the kernel and the wait, no copy launch
    int* host_count = nullptr;   // NRING ring slots, then the speculative solve's slot (SPEC)
    static constexpr int SPEC = NRING;
    hipEvent_t ev[NRING] = {};
    // Speculative solve: after the predicted number of blind Krylov iterations, the work that
    // follows the solve is queued at once, gated on device (gate = every instance converged,
    // set by the count after the blind iterations); the host resolves it at the next
    // synchronisation: a short prediction means the gated work was a no-op, and the solve is
    // finished with checked iterations and the work queued again (same results either way).
    DevPtr<int> gate;        // 1 int, batch-wide; 1 outside an unresolved speculative solve
    bool allow_spec = false;   // set by iterate() around the solve it can resolve
    struct Spec {
        bool on = false;
        Table* t = nullptr;   // alive: iterate() holds the solve's table until resolved
        std::tuple<double, int, int, int> pkey;
        double tol = 0;
        int maxit = 0, meth = 0, k = 0;
    } spec;
    void open_gate() {
        const int one = 1;
        EQ_HIP(hipMemcpyAsync(gate->p, &one, sizeof(int), hipMemcpyHostToDevice, S()));
        Device::get().sync();
    }
    double* gpin = nullptr;
    size_t gcap = 0;
    int* gflag = nullptr;   // pinned: the sequence number of the latest gather (k_gather_slots)
    int gseq = 0;           // the latest direct gather's number (1 .. 2^30)
    // the number every gather captured in a graph stores (the flag is cleared before a replay)
    static constexpr int GSEQ_GRAPH = -7;
    // step-start copies queued by begin_step, issued together before the next device work
    // that reads or replaces those fields (flush_copies)
    std::vector<McJob> pend;
    DevPtr<char> mc_dev;
    // pinned host image of the job table (mc_cap bytes): the upload reads it asynchronously,
    // so the next flush waits on mc_ev (recorded after the upload) before rewriting it
    char* mc_host = nullptr;
    size_t mc_cap = 0;
    hipEvent_t mc_ev = nullptr;

    size_t ks() const { return mixed ? sizeof(float) : sizeof(double); }

    void flush_copies() {
        if (pend.empty()) return;
        if (pend.size() == 1) {
            EQ_HIP(hipMemcpyAsync(pend[0].dst, pend[0].src, (size_t)pend[0].n * sizeof(double),
                                  hipMemcpyDeviceToDevice, S()));
            pend.clear();
            return;
        }
        std::vector<int> bj;
        for (size_t q = 0; q < pend.size(); ++q) {
            pend[q].b0 = (long)bj.size();
            const long nb = (pend[q].n + MC_EPB - 1) / MC_EPB;
            bj.insert(bj.end(), (size_t)nb, (int)q);
        }
        const size_t jb = pend.size() * sizeof(McJob), bb = bj.size() * sizeof(int), hb = jb + bb;
        EQ_HIP(hipEventSynchronize(mc_ev));   // the previous upload has read mc_host
        if (mc_cap < hb) {
            if (mc_host) EQ_HIP(hipHostFree(mc_host));
            mc_host = nullptr;
            mc_cap = std::max(hb, (size_t)65536);
            EQ_HIP(hipHostMalloc((void**)&mc_host, mc_cap, hipHostMallocDefault));
        }
        std::memcpy(mc_host, pend.data(), jb);
        std::memcpy(mc_host + jb, bj.data(), bb);
        if (!mc_dev || mc_dev->n < hb) mc_dev = dev_alloc<char>(std::max(hb, (size_t)65536));
        EQ_HIP(hipMemcpyAsync(mc_dev->p, mc_host, hb, hipMemcpyHostToDevice, S()));
        EQ_HIP(hipEventRecord(mc_ev, S()));
        if (!bj.empty()) {
            hipLaunchKernelGGL(k_multi_copy, dim3((unsigned)bj.size()), dim3(BS), 0, S(), (const McJob*)mc_dev->p,
                               (const int*)(mc_dev->p + jb));
            EQ_HIP(hipGetLastError());
        }
        pend.clear();
    }

    // two-phase instances in the reduced form solve their pressure stage by CG with the
    // Schur complement's diagonal (those in the full form with the V-cycle: Inst::mg_on)
    bool cg_capable(const Inst& I) const {
        return cfg.twophase_pressure_cg && I.kind == K_TWOPHASE && I.pcm == PM_SCHUR;
    }
    // Level l >= 1 of an instance's hierarchy (the fields the multigrid kernels read).
    Desc level_desc(Inst& I, int l) const {
        Desc d = desc_of(I);
        const Inst::Level& L = I.mg[l - 1];
        d.n = L.n;
        d.nx = L.nx;
        d.ny = L.ny;
        d.nz = L.nz;
        d.rnx = 1.0 / L.nx;
        d.rnxy = 1.0 / ((double)L.nx * L.ny);
        d.xcd = 0;
        const size_t h = L.h();
        d.h = (int)h;
        d.ns = 2;
        d.vec = 0;
        for (int q = 0; q < 7; ++q) d.K[q] = L.K->p + q * 2 * h * ks();
        for (int v = 0; v < NVEC; ++v) d.v[v] = nullptr;
        d.v[V_R] = L.v->p;
        d.v[V_Z] = L.v->p + 2 * h * ks();
        d.v[V_T] = L.v->p + 4 * h * ks();
        d.mgl = (int)I.mg.size() - l;
        d.mgt = l >= I.mg_tail;
        d.minv = d.mgl == 0 ? (void*)I.minv->p : nullptr;
        return d;
    }
    void build_mg(Table& t) {
        std::vector<Inst*> mi;
        for (Inst* I : t.inst)
            if (I->mg_on) mi.push_back(I);
        if (mi.empty()) return;
        auto m = std::make_unique<Table::Mg>();
        for (Inst* I : mi) m->levels = std::max(m->levels, (int)I->mg.size() + 1);
        m->sep.assign(m->levels, 0);
        std::vector<Desc> td;
        std::vector<int2> tr;
        for (Inst* I : mi) {
            for (int l = 0; l < I->mg_tail; ++l) m->sep[l] = 1;
            m->tail0 |= I->mg_tail == 0;
            const int first = (int)td.size();
            for (int l = I->mg_tail; l <= (int)I->mg.size(); ++l) td.push_back(l == 0 ? desc_of(*I) : level_desc(*I, l));
            tr.push_back(make_int2(first, (int)td.size() - first));
        }
        m->tail = dev_from_host(td.data(), td.size());
        m->trange = dev_from_host(tr.data(), tr.size());
        m->ntail = (int)tr.size();
        m->fine.resize(m->levels);
        m->coarse.resize(m->levels);
        m->cblk.resize(m->levels);
        m->fpblk.resize(m->levels);
        m->cpblk.resize(m->levels);
        m->ncb.assign(m->levels, 0);
        m->nfpb.assign(m->levels, 0);
        m->ncpb.assign(m->levels, 0);
        auto blocks = [](std::vector<int2>& b, int q, int count) {
            for (int k = 0; k < (count + BS - 1) / BS; ++k) b.push_back(make_int2(q, k));
        };
        for (int l = 1; l < m->levels; ++l) {
            std::vector<Desc> fd, cd;
            std::vector<int2> cb, fpb, cpb;
            for (Inst* I : mi) {
                if ((int)I->mg.size() < l) continue;
                const int q = (int)fd.size();
                fd.push_back(l == 1 ? desc_of(*I) : level_desc(*I, l - 1));
                fd.back().xcd = 0;
                cd.push_back(level_desc(*I, l));
                blocks(cb, q, cd.back().n);
                blocks(fpb, q, fd.back().h);
                blocks(cpb, q, cd.back().h);
            }
            m->fine[l] = dev_from_host(fd.data(), fd.size());
            m->coarse[l] = dev_from_host(cd.data(), cd.size());
            m->ncb[l] = (int)cb.size();
            m->nfpb[l] = (int)fpb.size();
            m->ncpb[l] = (int)cpb.size();
            m->cblk[l] = dev_from_host(cb.data(), cb.size());
            m->fpblk[l] = dev_from_host(fpb.data(), fpb.size());
            m->cpblk[l] = dev_from_host(cpb.data(), cpb.size());
        }
        std::vector<Desc> cs;
        for (Inst* I : mi) cs.push_back(I->mg.empty() ? desc_of(*I) : level_desc(*I, (int)I->mg.size()));
        m->coarsest = dev_from_host(cs.data(), cs.size());
        m->ncoarsest = (int)cs.size();
        t.mg = std::move(m);
    }
    Desc desc_of(Inst& I) const {
        Desc d{};
        d.n = I.n;
        d.nx = I.nx;
        d.nxr = I.real_nx();
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
            d.hcgx = I.dt / (I.cv * d.vol) * d.gx;
            d.hcgy = I.dt / (I.cv * d.vol) * d.gy;
            d.u = I.u->p;
            d.u0 = I.u0->p;
        } else if (I.kind == K_TWOPHASE) {
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
        auto p64 = [](const DevPtr<double>& a) { return a ? a->p : nullptr; };
        d.C = p64(I.C);
        d.W = p64(I.W);
        d.E = p64(I.E);
        d.S = p64(I.S);
        d.N = p64(I.N);
        d.B = p64(I.B);
        d.T = p64(I.T);
        const size_t h = I.h();
        d.h = (int)h;
        d.pcm = I.pcm_now();
        d.ns = d.pcm == PM_SCHUR ? 1 : 2;
        d.vec = cfg.pair_vectors && (I.kind == K_HEAT || (I.kind == K_LINEAR && I.nz == 1)) && I.pcm == PM_SCHUR &&
                I.nx % 8 == 0;
        d.tile = d.vec && cfg.operator_tiles && I.nx % 512 == 0;
        d.cgp = cg_capable(I);
        d.mgp = I.mg_on;
        d.mgs = cfg.mg_scale;
        d.mgl = (int)I.mg.size();
        d.mgt = I.mg_on && I.mg_tail == 0;
        d.minv = I.mg_on && I.mg.empty() ? (void*)I.minv->p : nullptr;
        for (int q = 0; q < 7; ++q) d.K[q] = q < I.ncoef() ? I.kcoef->p + q * 2 * h * ks() : nullptr;
        d.dinv = I.dinv->p;
        for (int v = 0; v < NVEC; ++v) d.v[v] = I.vecs->p + (size_t)v * d.ns * h * ks();
        d.sc = I.sc->p;
        d.act = I.act->p;
        d.gate = gate->p;
        return d;
    }

    TableP build_table(const std::vector<Inst*>& list, const Shape& sh) {
        TableP tp = std::make_shared<Table>();
        Table& t = *tp;
        t.version = version;
        t.inst = list;
        t.shape = sh;
        std::vector<Desc> ds;
        int g = 0;
        for (size_t q = 0; q < list.size(); ++q) {
            Desc d = desc_of(*list[q]);
            d.xcd = sh.xcd;
            d.gbase = g;
            d.ngrp = (list[q]->n + BS - 1) / BS;
            d.ngrpk = (list[q]->h() + BS - 1) / BS;
            g += d.ngrp;
            ds.push_back(d);
            t.has_schur |= d.pcm == PM_SCHUR;
            t.has_full |= d.pcm != PM_SCHUR;
            t.has_ilu |= d.pcm == PM_ILU;
            t.has_tp |= d.kind == K_TWOPHASE;
            t.has_vec |= d.vec != 0;
            t.has_nonvec |= d.vec == 0;
            t.has_schur_strided |= d.pcm == PM_SCHUR && d.vec == 0;
            t.has_tile |= d.tile != 0;
            t.has_vec_onepass |= d.vec && !d.tile;
        }
        t.ng = g;
        for (int ci = 0; ci < 4; ++ci) {
            const int cpb = (1 << ci) * BS;
            std::vector<int2> blk, pblk;
            for (size_t q = 0; q < list.size(); ++q) {
                const int nbk = (list[q]->n + cpb - 1) / cpb, npk = (list[q]->h() + cpb - 1) / cpb;
                for (int b = 0; b < nbk; ++b) blk.push_back(make_int2((int)q, b));
                for (int b = 0; b < npk; ++b) pblk.push_back(make_int2((int)q, b));
            }
            t.nb[ci] = (int)blk.size();
            t.blk[ci] = dev_from_host(blk.data(), blk.size());
            t.pnb[ci] = (int)pblk.size();
            t.pblk[ci] = dev_from_host(pblk.data(), pblk.size());
        }
        {   // tiles of TILE_ROWS rows x 256 pair ranks (k_schur_tile)
            std::vector<int2> tb;
            for (size_t q = 0; q < list.size(); ++q) {
                if (!ds[q].tile) continue;
                const int tiles = (ds[q].nx / 512) * ((ds[q].ny + TILE_ROWS - 1) / TILE_ROWS);
                for (int b = 0; b < tiles; ++b) tb.push_back(make_int2((int)q, b));
            }
            t.tnb = (int)tb.size();
            if (!tb.empty()) t.tblk = dev_from_host(tb.data(), tb.size());
        }
        t.desc = dev_from_host(ds.data(), ds.size());
        t.part = dev_alloc<double>(5 * (size_t)t.ng + 2);   // up to five dot products per launch
        build_mg(t);
        return tp;
    }
    TableP table(const std::vector<Inst*>& list) {
        std::vector<int> key;   // the instances and, for split ones, the stage (their Krylov form)
        for (Inst* I : list) key.push_back(2 * I->id + (I->split && I->stage == 1 ? 1 : 0));
        auto it = tcache.find(key);
        if (it != tcache.end() && it->second->version == version) return it->second;
        if (tcache.size() > 64) tcache.clear();
        TableP tp = build_table(list, shape_for(list));
        tcache[key] = tp;
        return tp;
    }

    // ---- launch-shape tuner -------------------------------------------------------
    // A membership's signature is (log2 total cells, log2 instance count, log2 largest
    // instance).  The first time a signature is seen, every kernel class is timed on that
    // membership with HIP events at 1, 2, 4 and 8 cells per thread, with the XCD block
    // order off and on; each class keeps its fastest cells per thread, and the block
    // order with the smaller weighted total (weights: launches of the class per Newton
    // iteration, about 1 assembly, 2 defects, 16 operator applies, 2 conversion rounds
    // and 8 vector-update rounds) is kept for the table.  Results do not depend on the shape
    // (shape-independent reductions), so tuning changes time only.
    std::map<std::tuple<int, int, int>, Shape> tuned;
    struct Measured {   // a membership tuned by measurement
        long cells;       // its total cells
        int lmax;         // log2 of its largest instance
        Shape shape;
    };
    std::vector<Measured> measured;
    Shape last_shape;
    static int ilog2(long v) {
        int r = 0;
        while (v > 1) v >>= 1, ++r;
        return r;
    }
    Shape shape_for(const std::vector<Inst*>& list) {
        if (cfg.cpt > 0 || cfg.cpt_fixed[0] > 0) {
            Shape sh;
            for (int c = 0; c < NCLS; ++c) sh.cpt[c] = cfg.cpt > 0 ? cfg.cpt : cfg.cpt_fixed[c];
            sh.xcd = cfg.xcd;
            return last_shape = sh;
        }
        long tot = 0, mx = 0;
        for (Inst* I : list) tot += I->n, mx = std::max<long>(mx, I->n);
        const auto sig = std::make_tuple(ilog2(tot), ilog2((long)list.size()), ilog2(mx));
        auto it = tuned.find(sig);
        if (it != tuned.end()) return last_shape = it->second;
        // a membership no larger (in cells) than one already measured whose largest
        // instance is of the same size class (the instances still iterating late in a step,
        // which come and go with their convergence) takes the shape of the smallest such
        // measured membership instead of a tuning run of its own; results do not depend on
        // the shape
        const Measured* ref = nullptr;
        for (const Measured& m : measured)
            if (m.cells >= tot && m.lmax == ilog2(mx) && (!ref || m.cells < ref->cells)) ref = &m;
        Shape best;
        if (ref) {
            best = ref->shape;
        } else {
            if (cfg.tune) best = tune_shape(list);
            measured.push_back({tot, ilog2(mx), best});
        }
        tuned[sig] = best;
        return last_shape = best;
    }
    // One trial launch of a kernel class at the table's shape for that class.
    void trial(Table& t, int cls) {
        switch (cls) {
            case C_ASM: launch_assemble(t); break;
            case C_DEF: launch_defect(t); break;
            case C_SPMV:   // the operator and the ILU passes
                if (mixed) {
                    ilu_apply<float>(t, V_P, V_PH);
                    launch_op<float>(t, V_PH, V_V, 1, (int)V_RH, -1);
                } else {
                    ilu_apply<double>(t, V_P, V_PH);
                    launch_op<double>(t, V_PH, V_V, 1, (int)V_RH, -1);
                }
                break;
            case C_PC:
                if (mixed) conversions<float>(t);
                else conversions<double>(t);
                break;
            default:   // the vector updates of both Krylov methods (each skips the other's instances)
                if (mixed) {
                    vec_round<float>(t);
                    cg_round<float>(t);
                } else {
                    vec_round<double>(t);
                    cg_round<double>(t);
                }
                break;
        }
    }
    Shape tune_shape(const std::vector<Inst*>& list) {
        // every instance active so the kernels do their full work.  Trial launches only
        // write what the next refinement round re-initialises (Krylov vectors, partials) or
        // rewrite F, the coefficients and the refinement residual from the current u and du
        // (unchanged values); du itself, which an instance between refinement rounds keeps,
        // is not written.
        const int one[2] = {1, 1};
        for (Inst* I : list) {
            EQ_HIP(hipMemcpyAsync(I->act->p, one, sizeof(one), hipMemcpyHostToDevice, S()));
            launch_fill(I->sc->p + S_XUPD, 1.0, 1);   // k_b31 does its full work (the next solve clears it)
        }
        static const double weight[NCLS] = {1.0, 2.0, 16.0, 2.0, 8.0};
        hipEvent_t e0, e1;
        EQ_HIP(hipEventCreate(&e0));
        EQ_HIP(hipEventCreate(&e1));
        Shape best;
        double best_total = 1e300;
        for (int xcd : {0, 1}) {
            Shape sh;
            sh.xcd = xcd;
            TableP tp = build_table(list, sh);
            Table& t = *tp;
            double total = 0.0;
            for (int cls = 0; cls < NCLS; ++cls) {
                if (cls == C_DEF && !mixed) continue;   // no refinement defect in fp64 mode
                // the shapes in turn, three rounds of four launches each; a shape's time is
                // its fastest round (clock and neighbour noise only ever add time)
                static const int cpts[4] = {1, 2, 4, 8};
                double t_min[4] = {1e300, 1e300, 1e300, 1e300};
                for (int round = 0; round < 3; ++round)
                    for (int q = 0; q < 4; ++q) {
                        t.shape.cpt[cls] = cpts[q];
                        trial(t, cls);   // warm
                        EQ_HIP(hipEventRecord(e0, S()));
                        for (int r = 0; r < 4; ++r) trial(t, cls);
                        EQ_HIP(hipEventRecord(e1, S()));
                        EQ_HIP(hipEventSynchronize(e1));
                        float ms = 0.f;
                        EQ_HIP(hipEventElapsedTime(&ms, e0, e1));
                        t_min[q] = std::min(t_min[q], (double)ms);
                    }
                double best_ms = 1e300;
                for (int q = 0; q < 4; ++q)
                    if (t_min[q] < best_ms) best_ms = t_min[q], sh.cpt[cls] = cpts[q];
                total += weight[cls] * best_ms;
            }
            if (total < best_total) best_total = total, best = sh;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return best;
    }

    // out[0] = active instances (flag `which`), out[1] = max Krylov iterations, synchronously
    // (gate: the post-solve gate is set from the count as well)
    int count_active(const Table& t, int which, int* maxit = nullptr, int* gate_dev = nullptr) {
        hipLaunchKernelGGL(k_count_active, dim3(1), dim3(BS), 0, S(), t.desc->p, (int)t.inst.size(), which, host_count,
                           gate_dev);
        EQ_HIP(hipGetLastError());
        Device::get().sync();
        if (maxit) *maxit = host_count[1];
        return host_count[0];
    }
    // The same without stalling the queue: the result of ring slot k is read later.
    void count_async(const Table& t, int which, int slot) {
        hipLaunchKernelGGL(k_count_active, dim3(1), dim3(BS), 0, S(), t.desc->p, (int)t.inst.size(), which,
                           host_count + 2 * slot, (int*)nullptr);
        EQ_HIP(hipGetLastError());
        EQ_HIP(hipEventRecord(ev[slot], S()));
    }
    int count_wait(int slot, int* maxit = nullptr) {
        EQ_HIP(hipEventSynchronize(ev[slot]));
        if (maxit) *maxit = host_count[2 * slot + 1];
        return host_count[2 * slot];
    }
    // Finalisation of a reduction: one block per instance (kernel boundaries make the
    // partials of all blocks visible; no device-wide fences inside the kernels).
    void fin(const Table& t, const Fin& f) {
        if (f.nd > 2)
            hipLaunchKernelGGL(k_fin<5>, dim3((unsigned)t.inst.size()), dim3(FBS), 0, S(), t.desc->p, t.part->p, t.ng,
                               f);
        else
            hipLaunchKernelGGL(k_fin<2>, dim3((unsigned)t.inst.size()), dim3(FBS), 0, S(), t.desc->p, t.part->p, t.ng,
                               f);
        EQ_HIP(hipGetLastError());
    }
    // finaliser of a Krylov reduction (partials over pairs) ...
    static Fin F(int act, int nd, double tol2 = 0, int maxit = 0, int slot = 0) {
        Fin f{};
        f.act = act;
        f.nd = nd;
        f.kry = 1;
        f.tol2 = tol2;
        f.maxit = maxit;
        f.slot = slot;
        return f;
    }
    // ... and of a cell kernel (assembly, defect): partials over cells, stored in `slot`
    static Fin Fcell(int slot) {
        Fin f = F(A_STORE, 1, 0, 0, slot);
        f.kry = 0;
        return f;
    }
// Launch over every block of the table at the cells per thread of kernel class `cls`;
// the kernel is named with CP, e.g. HL(C_ASM, (k_assemble<true, CP>), t.part->p).
#define HL(cls, kernel, ...)                                                                                \
    by_cpt(t.shape.cpt[cls], [&](auto cpt_) {                                                               \
        constexpr int CP = decltype(cpt_)::value;                                                           \
        const int ci_ = cpt_index(CP);                                                                      \
        hipLaunchKernelGGL(kernel, dim3((unsigned)t.nb[ci_]), dim3(BS), 0, S(), t.desc->p, t.blk[ci_]->p,   \
                           ##__VA_ARGS__);                                                                  \
        EQ_HIP(hipGetLastError());                                                                          \
    })
// The same over the pair ranks of the Krylov space.
#define HLP(cls, kernel, ...)                                                                               \
    by_cpt(t.shape.cpt[cls], [&](auto cpt_) {                                                               \
        constexpr int CP = decltype(cpt_)::value;                                                           \
        const int ci_ = cpt_index(CP);                                                                      \
        hipLaunchKernelGGL(kernel, dim3((unsigned)t.pnb[ci_]), dim3(BS), 0, S(), t.desc->p, t.pblk[ci_]->p, \
                           ##__VA_ARGS__);                                                                  \
        EQ_HIP(hipGetLastError());                                                                          \
    })

    void launch_assemble(const Table& t, int norm_only = 0) {
        TimedLaunch tl("assemble");
        if (mixed) HL(C_ASM, (k_assemble<true, CP>), t.part->p, norm_only);
        else HL(C_ASM, (k_assemble<false, CP>), t.part->p, norm_only);
    }
    void assemble(const Table& t, int norm_only = 0) {
        launch_assemble(t, norm_only);
        fin(t, Fcell(S_FN));
    }
    void launch_defect(const Table& t) {
        TimedLaunch tl("defect_fp64");
        HL(C_DEF, (k_defect<CP>), t.part->p);
    }
    // w = A v (full-system instances) and S v (reduced-system instances) with the fused dots
    // (reduced-system instances of the methods in meth only).  The four-ranks-per-lane
    // instances solve the column-scaled system: their operand for ph (sh) is p (s).
    template <class T>
    void launch_op(const Table& t, int iv, int iw, int nd, int a0, int a1, int meth = M_BICG | M_CG) {
        TimedLaunch tl(sizeof(T) == 4 ? "operator_fp32" : "operator_fp64");
        if (t.has_schur_strided) {
            HLP(C_SPMV, (k_schur_red<T, CP>), iv, meth);
            HLP(C_SPMV, (k_schur_black<T, CP>), iv, iw, nd, a0, a1, t.part->p, t.ng, meth);
        }
        const int ivv = iv == V_PH ? (int)V_P : iv == V_SH ? (int)V_S : iv;
        if (t.has_vec_onepass) HLP(C_SPMV, (k_schur_vec<T, CP>), ivv, iw, nd, a0, t.part->p, t.ng, meth);
        if (t.has_tile) {
            hipLaunchKernelGGL((k_schur_tile<T>), dim3((unsigned)t.tnb), dim3(TBS), 0, S(), t.desc->p, t.tblk->p, ivv,
                               iw, nd, a0, t.part->p, t.ng, meth);
            EQ_HIP(hipGetLastError());
        }
        if (t.has_full) HLP(C_SPMV, (k_spmv_rb<T, CP>), iv, iw, nd, a0, a1, t.part->p, t.ng, meth);
    }
    // Multigrid: coarse operators and coarsest inverses of the instances about to solve a
    // pressure Newton iteration by CG with the V-cycle.
    template <class T>
    void mg_setup(const Table& t) {
        if (!t.mg) return;
        const Table::Mg& m = *t.mg;
        for (int l = 1; l < m.levels; ++l) {
            hipLaunchKernelGGL((k_mg_galerkin<T>), dim3((unsigned)m.ncb[l]), dim3(BS), 0, S(), m.fine[l]->p,
                               m.coarse[l]->p, m.cblk[l]->p);
            EQ_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL((k_mg_dense_setup<T>), dim3((unsigned)m.ncoarsest), dim3(MG_BS), 0, S(), m.coarsest->p);
        EQ_HIP(hipGetLastError());
    }
    // One V-cycle, z = M r (V_R -> V_Z), for the V-cycle instances of the table, followed
    // by the CG partial dots (r.z, r.r) of every V-cycle instance.
    template <class T>
    void vcycle_dots(const Table& t) {
        const Table::Mg& m = *t.mg;
        auto smooth = [&](int l, int col, int zero, int dots) {
            if (l == 0) {
                HLP(C_PC, (k_mg_smooth<T, CP>), col, zero, dots, t.part->p, t.ng);
            } else {
                hipLaunchKernelGGL((k_mg_smooth<T, 1>), dim3((unsigned)m.ncpb[l]), dim3(BS), 0, S(), m.coarse[l]->p,
                                   m.cpblk[l]->p, col, zero, 0, (double*)nullptr, 0);
                EQ_HIP(hipGetLastError());
            }
        };
        for (int l = 0; l + 1 < m.levels && m.sep[l]; ++l) {
            smooth(l, 0, 1, 0);
            smooth(l, 1, 0, 0);
            if (l == 0) {
                HLP(C_PC, (k_mg_resid<T, CP>));
            } else {
                hipLaunchKernelGGL((k_mg_resid<T, 1>), dim3((unsigned)m.ncpb[l]), dim3(BS), 0, S(), m.coarse[l]->p,
                                   m.cpblk[l]->p);
                EQ_HIP(hipGetLastError());
            }
            hipLaunchKernelGGL((k_mg_restrict<T>), dim3((unsigned)m.ncb[l + 1]), dim3(BS), 0, S(), m.fine[l + 1]->p,
                               m.coarse[l + 1]->p, m.cblk[l + 1]->p);
            EQ_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL((k_mg_tail<T>), dim3((unsigned)m.ntail), dim3(TAIL_BS), 0, S(), m.tail->p, m.trange->p);
        EQ_HIP(hipGetLastError());
        for (int l = m.levels - 2; l >= 0; --l) {
            if (!m.sep[l]) continue;
            hipLaunchKernelGGL((k_mg_prolong_black<T>), dim3((unsigned)m.nfpb[l + 1]), dim3(BS), 0, S(),
                               m.fine[l + 1]->p, m.coarse[l + 1]->p, m.fpblk[l + 1]->p);
            EQ_HIP(hipGetLastError());
            smooth(l, 0, 0, l == 0);
        }
        if (m.tail0) HLP(C_VEC, (k_cg_dots<T, CP>), t.part->p, t.ng);
    }
    // z = M^-1 r for the full-system ILU(0) instances
    template <class T>
    void ilu_apply(const Table& t, int ir, int iz) {
        if (!t.has_ilu) return;
        HLP(C_SPMV, (k_ilu_black<T, CP>), ir, iz);
        HLP(C_SPMV, (k_ilu_red<T, CP>), ir, iz);
    }
    template <class T>
    void spmv_dot(const Table& t, int ix, int iy, int nd, int a0, int a1, const Fin& f, int meth) {
        launch_op<T>(t, ix, iy, nd, a0, a1, meth);
        fin(t, f);
    }
    // the conversion class is timed on the right-hand side kernel alone: k_accum adds into
    // du, which must survive between the refinement rounds of an instance
    template <class T>
    void conversions(const Table& t) {
        if (t.has_nonvec) HLP(C_PC, (k_load_rhs<T, CP>), 1, 0);
        if (t.has_vec) HLP(C_PC, (k_setup_vec<T, CP>), t.part->p, t.ng, 1, 0, 0);   // (no column scaling)
    }
    template <class T>
    void vec_round(const Table& t) {
        if (t.has_nonvec) HLP(C_VEC, (k_b1<T, CP>));
        HLP(C_VEC, (k_b2<T, CP>));
        if (t.has_nonvec) HLP(C_VEC, (k_b3<T, CP>), t.part->p, t.ng);
        if (t.has_vec) HLP(C_VEC, (k_b31<T, CP>));
    }
    template <class T>
    void cg_round(const Table& t) {
        HLP(C_VEC, (k_cg2<T, CP>), t.part->p, t.ng);
        HLP(C_VEC, (k_cg3<T, CP>), 0);
    }

    // One Krylov iteration for the instances whose refinement flag is set (rhs in V_R):
    // BiCGStab (right preconditioner: ph = M^-1 p and sh = M^-1 s are formed in k_b1 /
    // k_b2) for the M_BICG instances, preconditioned CG for the M_CG ones; meth = the
    // methods present in the solve.  The four-ranks-per-lane instances take their first
    // direction from k_kinit, and the second operator's five dot products give them the
    // iteration's end in A_OMEGA: k_b31 updates x and r and forms the next direction (two
    // reductions per iteration instead of three).
    template <class T>
    void krylov_iteration(const Table& t, double tol, int maxit, int meth) {
        const double tol2 = tol < 0.0 ? -1.0 : tol * tol;
        if (meth & M_BICG) {
            const int nd2 = t.has_vec ? 5 : 2;
            if (t.has_nonvec) {
                HLP(C_VEC, (k_b1<T, CP>));
                ilu_apply<T>(t, V_P, V_PH);
            }
            spmv_dot<T>(t, V_PH, V_V, 1, (int)V_RH, -1, F(A_ALPHA, 1), M_BICG);   // v = A ph; alpha = rho / rh.v
            HLP(C_VEC, (k_b2<T, CP>));
            ilu_apply<T>(t, V_S, V_SH);
            // t = A sh; omega = t.s / t.t (four ranks per lane: rho, ||r|| and the convergence too)
            spmv_dot<T>(t, V_SH, V_T, nd2, (int)V_S, -1, F(A_OMEGA, nd2, tol2, maxit), M_BICG);
            if (t.has_nonvec) {
                HLP(C_VEC, (k_b3<T, CP>), t.part->p, t.ng);
                fin(t, F(A_RHO, 2, tol2, maxit));
            }
            if (t.has_vec) HLP(C_VEC, (k_b31<T, CP>));
        }
        if (meth & M_CG) {
            spmv_dot<T>(t, V_P, V_V, 1, (int)V_P, -1, F(A_CG_ALPHA, 1), M_CG);    // q = A p; alpha = rho / p.q
            HLP(C_VEC, (k_cg2<T, CP>), t.part->p, t.ng);                         // x, r; r.z, r.r
            if (t.mg) vcycle_dots<T>(t);                                         // z = V-cycle(r); r.z, r.r
            fin(t, F(A_CG_RHO, 2, tol2, maxit));
            HLP(C_VEC, (k_cg3<T, CP>), 0);                                       // p = z + beta p
        }
    }
    // The methods of the instances of a solve (host-side stage of each instance).
    int methods(const Table& t) const {
        int m = 0;
        for (const Inst* I : t.inst) m |= (cg_capable(*I) || I->mg_on) && I->stage == 0 ? M_CG : M_BICG;
        return m;
    }

    // One BiCGStab solve.  From the second solve on a membership, one iteration is
    // captured in a hipGraph and replayed: the iterations the previous solve with the
    // same tolerance needed, minus one, are queued without looking at the device; then
    // one synchronous check of the active count; further iterations (if any) are
    // pipelined, launch k+1 being queued before the count after launch k is read.
    // Instances that converge earlier are masked by their device flag, so extra
    // iterations are no-ops.
    // The prediction key of a solve: (tolerance, iteration cap, methods, the solve's position
    // in the step: the latest instance's Newton iteration and refinement round).
    std::tuple<double, int, int, int> pred_key(const Table& t, double tol, int maxit) const {
        int at = 0;
        for (const Inst* I : t.inst) at = std::max(at, 16 * I->st.newton_stage + std::min(I->round, 15));
        return std::make_tuple(tol, maxit, methods(t), at);
    }
    // The start of a Krylov solve: right-hand side mean (two-phase pressure), r, rh, x and
    // the first direction, the V-cycle instances' first z.
    template <class T>
    void krylov_start(const Table& t, int meth) {
        if (t.has_tp) {   // two-phase pressure: mean of the right-hand side
            HLP(C_VEC, (k_rhs_sum<T, CP>), t.part->p);
            fin(t, F(A_MEAN, 1));
        }
        if (t.has_vec) {   // the four-ranks-per-lane instances' pivots, column scaling, right-hand side, start
            if (mixed) HLP(C_PC, (k_setup_vec<T, CP>), t.part->p, t.ng, 1, -1, 1);
            else HLP(C_PC, (k_setup_vec<T, CP>), t.part->p, t.ng, 0, 1, 1);
        }
        if (t.has_nonvec) HLP(C_VEC, (k_kinit<T, CP>), t.part->p, t.ng);
        fin(t, F(A_INIT, 2));
        if (t.mg && (meth & M_CG)) {   // V-cycle instances: first z, rho = r.z, p = z
            vcycle_dots<T>(t);
            fin(t, F(A_CG_START, 1));
            HLP(C_VEC, (k_cg3<T, CP>), 1);
        }
    }
    template <class T>
    void bicgstab(Table& t, double tol, int maxit) {
        const int meth = methods(t);
        const auto pkey = pred_key(t, tol, maxit);
        krylov_start<T>(t, meth);
        int iters = 0;
        if (use_graph && !timing_enabled() && t.solves++ > 0) {
            // graphs of one and of two iterations (the blind launches go in pairs: one
            // inter-graph gap per two iterations)
            auto graph_of = [&](int reps) { return graph_exec<T>(t, tol, maxit, meth, reps); };
            const hipGraphExec_t exec = graph_of(1);
            const int pred = t.pred[pkey];
            if (allow_spec && pred >= 1) {
                // speculative: the predicted iterations, then the count (into the SPEC slot
                // and the gate) without waiting; iterate() resolves it (finish_spec)
                const int nb = std::min(maxit, pred);
                int k = 0;
                if (nb >= 2) {
                    const hipGraphExec_t exec2 = graph_of(2);
                    for (; k + 2 <= nb; k += 2) EQ_HIP(hipGraphLaunch(exec2, S()));
                }
                for (; k < nb; ++k) EQ_HIP(hipGraphLaunch(exec, S()));
                hipLaunchKernelGGL(k_count_active, dim3(1), dim3(BS), 0, S(), t.desc->p, (int)t.inst.size(), 0,
                                   host_count + 2 * SPEC, gate->p);
                EQ_HIP(hipGetLastError());
                spec.on = true;
                spec.t = &t;
                spec.pkey = pkey;
                spec.tol = tol;
                spec.maxit = maxit;
                spec.meth = meth;
                spec.k = k;
                return;
            }
            const int blind = std::min(maxit, std::max(0, pred - 1));
            int k = 0;
            if (blind >= 2) {
                const hipGraphExec_t exec2 = graph_of(2);
                for (; k + 2 <= blind; k += 2) EQ_HIP(hipGraphLaunch(exec2, S()));
            }
            for (; k < blind; ++k) EQ_HIP(hipGraphLaunch(exec, S()));
            bool done = k >= maxit;
            if (done) {   // every allowed iteration was queued blind: read the count once
                count_active(t, 0, &iters);
            } else {
                EQ_HIP(hipGraphLaunch(exec, S()));
                ++k;
                done = count_active(t, 0, &iters) == 0;
            }
            int pending = -1;
            while (!done && k < maxit) {
                EQ_HIP(hipGraphLaunch(exec, S()));
                count_async(t, 0, k % NRING);
                if (pending >= 0 && count_wait(pending % NRING, &iters) == 0) done = true;
                pending = k++;
            }
            if (!done && pending >= 0) count_wait(pending % NRING, &iters);
            Device::get().sync();
        } else {
            if (count_active(t, 0)) {
                for (int it = 0; it < maxit; ++it) {
                    krylov_iteration<T>(t, tol, maxit, meth);
                    if (!count_active(t, 0, &iters)) break;
                }
            }
        }
        t.pred[pkey] = iters;
    }

    // The captured graph of `reps` Krylov iterations for (tolerance, iteration cap, methods).
    template <class T>
    hipGraphExec_t graph_exec(Table& t, double tol, int maxit, int meth, int reps) {
        const auto gk = std::make_tuple(tol, maxit, meth + 8 * reps);
        auto g = t.graphs.find(gk);
        if (g == t.graphs.end()) {
            hipGraph_t graph;
            hipGraphExec_t exec;
            EQ_HIP(hipStreamBeginCapture(S(), hipStreamCaptureModeThreadLocal));
            for (int r = 0; r < reps; ++r) krylov_iteration<T>(t, tol, maxit, meth);
            EQ_HIP(hipStreamEndCapture(S(), &graph));
            EQ_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            g = t.graphs.emplace(gk, std::make_pair(graph, exec)).first;
        }
        return g->second.second;
    }
    // Resolves a speculative solve after the host synchronisation that follows its gated
    // work: with every instance converged after the blind iterations (the count in the SPEC
    // slot), that work ran; otherwise the solve continues with checked iterations, the gate
    // is reopened, and the caller queues the gated work again (returns true).
    template <class T>
    bool finish_spec() {
        Table& t = *spec.t;
        spec.on = false;
        int iters = host_count[2 * SPEC + 1];
        if (host_count[2 * SPEC] == 0) {
            t.pred[spec.pkey] = iters;
            return false;
        }
        const hipGraphExec_t exec = graph_exec<T>(t, spec.tol, spec.maxit, spec.meth, 1);
        int k = spec.k, pending = -1;
        bool done = false;
        while (!done && k < spec.maxit) {
            EQ_HIP(hipGraphLaunch(exec, S()));
            count_async(t, 0, k % NRING);
            if (pending >= 0 && count_wait(pending % NRING, &iters) == 0) done = true;
            pending = k++;
        }
        if (!done && pending >= 0) count_wait(pending % NRING, &iters);
        open_gate();
        t.pred[spec.pkey] = iters;
        return true;
    }

    // Linear solve of one Newton iteration: flags and preconditioner, right-hand side, the
    // Krylov solve, then (linear_tail) the solution and, in mixed precision, the FP64 defect
    // and refinement flags.  One refinement round per call: an instance whose refinement
    // continues (S_CONT) skips the Newton update and runs its next round in the next call, so
    // no host check is needed between rounds; the first round of an instance reads F and
    // writes du (its own round counter decides).
    template <class T>
    void linear_solve(Table& t) {
        solve_front<T>(t);
        // per-instance tolerance (S_TOL2: the iteration's eta; mixed: the inner tolerance)
        bicgstab<T>(t, -1.0, cfg.solver_max_iter);
        linear_tail<T>(t);
    }
    // refinement flags from ||F|| (the first round reads F and writes du: no clearing pass),
    // the preconditioner and the Krylov right-hand side
    template <class T>
    void solve_front(const Table& t) {
        const unsigned ni = (unsigned)t.inst.size();
        hipLaunchKernelGGL(k_refine_flags, dim3((ni + 63) / 64), dim3(64), 0, S(), t.desc->p, (int)ni, 0,
                           tol_call > 0.0 ? tol_call : cfg.solver_tol, cfg.max_refine, forcing());
        // (the four-ranks-per-lane instances' pivots and right-hand side: k_setup_vec)
        if (t.has_nonvec) precond_setup<T>(t);
        mg_setup<T>(t);
        if (!t.has_nonvec) return;
        if (!mixed) HLP(C_PC, (k_load_rhs<T, CP>), 0, 1);
        else HLP(C_PC, (k_load_rhs<T, CP>), 1, -1);
    }
    // A whole speculative Newton iteration of a membership whose instances all solve, as
    // one graph: the solve's front, the Krylov start, nb blind iterations, the count (SPEC
    // slot and gate), the solution, defect and refinement flags, the Newton update, the
    // assembly at the new state and the gathered scalars (g) - the launches of the separate
    // path, with one graph launch and one wait.  False (nothing launched) when the solve
    // has no prediction yet or is not speculative; the caller then takes the separate path.
    template <class T>
    bool newton_graph_iteration(Table& t, std::vector<double>& g, int norm_only, int start = 0) {
        if (!allow_spec || !use_graph || timing_enabled() || t.solves == 0 || tol_call > 0.0) return false;
        const int maxit = cfg.solver_max_iter, meth = methods(t), ni = (int)t.inst.size();
        const auto pkey = pred_key(t, -1.0, maxit);
        const auto it = t.pred.find(pkey);
        if (it == t.pred.end() || it->second < 1) return false;
        const int nb = std::min(maxit, it->second);
        ensure_gpin(ni);
        // the whole iteration's launches (captured once, or issued directly)
        auto launches = [&](int seq) {
            if (start) {   // the step's first assembly (u0 = u) ahead of its first solve
                launch_assemble(t, 0);
                fin(t, Fcell(S_FN));
            }
            solve_front<T>(t);
            krylov_start<T>(t, meth);
            for (int k = 0; k < nb; ++k) krylov_iteration<T>(t, -1.0, maxit, meth);
            hipLaunchKernelGGL(k_count_active, dim3(1), dim3(BS), 0, S(), t.desc->p, ni, 0, host_count + 2 * SPEC,
                               gate->p);
            EQ_HIP(hipGetLastError());
            linear_tail<T>(t);
            newton_update(t);
            launch_assemble(t, norm_only);
            if (ni == 1) {   // one instance: the norm's finaliser also stores the gathered scalars
                Fin f = Fcell(S_FN);
                const int gs[6] = {S_FN, S_ITT, S_REF, S_RN, S_CONT, S_BN};
                f.gout = gpin;
                f.gflag = gflag;
                for (int q = 0; q < 6; ++q) f.gs[q] = gs[q];
                f.gseq = seq;
                fin(t, f);
            } else {
                fin(t, Fcell(S_FN));
                launch_gather(t, make_int4(S_FN, S_ITT, S_REF, S_RN), S_CONT, S_BN, seq);
            }
        };
        int wait_seq = GSEQ_GRAPH;
        if (cfg.newton_graph) {
            const auto gk = std::make_tuple(nb, meth + 8 * norm_only + 16 * start, (int)sizeof(T), (uintptr_t)gpin);
            auto gi = t.newton_graphs.find(gk);
            if (gi == t.newton_graphs.end()) {
                hipGraph_t graph;
                hipGraphExec_t exec;
                EQ_HIP(hipStreamBeginCapture(S(), hipStreamCaptureModeThreadLocal));
                launches(GSEQ_GRAPH);
                EQ_HIP(hipStreamEndCapture(S(), &graph));
                EQ_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                gi = t.newton_graphs.emplace(gk, std::make_pair(graph, exec)).first;
            }
            __atomic_store_n(gflag, 0, __ATOMIC_RELEASE);   // every earlier gather has been waited for
            EQ_HIP(hipGraphLaunch(gi->second.second, S()));
        } else {   // the same launches one by one: the first starts at once on an idle stream
            gseq = gseq >= (1 << 30) ? 1 : gseq + 1;
            wait_seq = gseq;
            launches(gseq);
        }
        ++t.solves;
        spec.on = true;
        spec.t = &t;
        spec.pkey = pkey;
        spec.tol = -1.0;
        spec.maxit = maxit;
        spec.meth = meth;
        spec.k = nb;
        wait_gather(wait_seq);
        g.assign(gpin, gpin + 6 * (size_t)ni);
        return true;
    }
    template <class T>
    void linear_tail(Table& t) {
        if (!mixed) {
            HLP(C_PC, (k_accum<T, CP>), 0, 1);
            return;
        }
        const unsigned ni = (unsigned)t.inst.size();
        HLP(C_PC, (k_accum<T, CP>), 1, -1);
        launch_defect(t);
        // ||r64||^2 and the refinement flags in the finaliser (k_refine_flags' mode 1)
        Fin f = Fcell(S_FN);
        f.act = A_REFINE;
        f.maxit = cfg.max_refine;
        fin(t, f);
        (void)ni;
    }
    double tol_call = 0.0;   // > 0: the tolerance of the current solve_linear call (linear-system instances)
    Forcing forcing() const {
        Forcing f;
        f.rel = cfg.rel_res_tol;
        f.abs = cfg.abs_res_tol;
        f.forcing = cfg.forcing;
        f.eta_max = cfg.eta_max;
        f.forcing0 = cfg.forcing0;
        f.mixed = mixed ? 1 : 0;
        return f;
    }
    template <class T>
    void precond_setup(const Table& t) {
        HLP(C_PC, (k_precond_setup<T, CP>));
    }
    // Krylov form of a new instance: reduced system for ilu0 where the kind's option asks
    // for it, else the full system with the batch's preconditioner.
    int form_of(int kind) const {
        if (pc != PC_ILU) return pc == PC_JACOBI ? PM_JACOBI : PM_NONE;
        return cfg.reduced_form(kind == K_TWOPHASE) ? PM_SCHUR : PM_ILU;
    }
    void newton_update(const Table& t) { HL(C_VEC, (k_newton_update<CP>)); }

    // Linear-system instance: right-hand side and coefficients from a CSR Jacobian, ||F||^2
    // into S_FN; a refinement left over from an earlier solve is cleared (S_CONT).
    void gather_linear(const Table& t, const int* rp, const int* ci, const double* val, const double* rhs) {
        if (mixed) HL(C_ASM, (k_gather_stencil<true, CP>), t.part->p, rp, ci, val, rhs);
        else HL(C_ASM, (k_gather_stencil<false, CP>), t.part->p, rp, ci, val, rhs);
        fin(t, Fcell(S_FN));
        for (const Inst* I : t.inst) launch_fill(I->sc->p + S_CONT, 0.0, 1);
    }
    // The same with F and the coefficients already in place: ||F||^2 only.
    void prepared_linear(const Table& t) {
        HL(C_ASM, (k_linear_fnorm<CP>), t.part->p);
        fin(t, Fcell(S_FN));
        for (const Inst* I : t.inst) launch_fill(I->sc->p + S_CONT, 0.0, 1);
    }
    // J du = F for linear-system instances: one solve in FP64, or FP32 solves with FP64
    // refinement rounds (one per linear_solve call) until every instance's refinement has
    // stopped; du holds the solution.
    void solve_linear_table(Table& t) {
        if (!mixed) {
            linear_solve<double>(t);
            return;
        }
        // the round keys the Krylov iteration predictions (the counts repeat from call to call)
        for (int round = 0; round <= cfg.max_refine; ++round) {
            for (Inst* I : t.inst) I->round = round;
            linear_solve<float>(t);
            if (count_active(t, 1) == 0) break;
        }
        for (Inst* I : t.inst) I->round = 0;
    }
#undef HL
#undef HLP

    // per-instance scalars of five slots in one launch, stored into pinned host memory:
    // result[q * ninst + i] = scalar slot q of instance i
    // six slots per instance (the sixth: S_BN, ||F|| at the start of the latest solve)
    std::vector<double> gather(const Table& t, int4 slots, int s4, int s5 = S_BN) {
        const int ni = (int)t.inst.size();
        ensure_gpin(ni);
        gseq = gseq >= (1 << 30) ? 1 : gseq + 1;
        launch_gather(t, slots, s4, s5, gseq);
        wait_gather(gseq);
        return std::vector<double>(gpin, gpin + 6 * (size_t)ni);
    }
    // Waits for the latest gather by watching its flag (the values are in place once it
    // shows the gather's number), with a stream query every so often: a stream that has
    // finished without the flag (an error) is reported by synchronising it.
    void wait_gather(int seq) {
        for (long it = 1;; ++it) {
            if (__atomic_load_n(gflag, __ATOMIC_ACQUIRE) == seq) return;
            if ((it & 255) == 0) {
                const hipError_t e = hipStreamQuery(S());
                if (e == hipErrorNotReady) continue;
                EQ_HIP(e);
                if (__atomic_load_n(gflag, __ATOMIC_ACQUIRE) == seq) return;
                Device::get().sync();
                EQ_REQUIRE(__atomic_load_n(gflag, __ATOMIC_ACQUIRE) == seq, "gathered scalars missing");
                return;
            }
        }
    }
    void ensure_gpin(int ni) {
        if (gcap >= 6 * (size_t)ni) return;
        Device::get().sync();
        if (gpin) EQ_HIP(hipHostFree(gpin));
        gpin = nullptr;
        gcap = 6 * (size_t)ni;
        EQ_HIP(hipHostMalloc((void**)&gpin, gcap * sizeof(double), hipHostMallocCoherent));
    }
    void launch_gather(const Table& t, int4 slots, int s4, int s5, int seq) {
        const int ni = (int)t.inst.size();
        const int nt = std::min(GATHER_BS, std::max(64, (ni + 63) / 64 * 64));
        hipLaunchKernelGGL(k_gather_slots, dim3(1), dim3(nt), 0, S(), t.desc->p, ni, slots, s4, s5, gpin, gflag,
                           seq);
        EQ_HIP(hipGetLastError());
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
    EQ_REQUIRE(cfg.krylov_form_heat == "reduced" || cfg.krylov_form_heat == "full" || cfg.krylov_form_heat == "auto",
               "krylov_form_heat must be reduced, full or auto");
    EQ_REQUIRE(cfg.krylov_form_twophase == "reduced" || cfg.krylov_form_twophase == "full" ||
                   cfg.krylov_form_twophase == "auto",
               "krylov_form_twophase must be reduced, full or auto");
    p_->use_graph = cfg.graph;
    EQ_HIP(hipHostMalloc((void**)&p_->host_count, 2 * (Impl::NRING + 1) * sizeof(int), hipHostMallocCoherent));
    EQ_HIP(hipHostMalloc((void**)&p_->gflag, sizeof(int), hipHostMallocCoherent));
    *p_->gflag = 0;
    p_->gate = dev_alloc<int>(1);
    const int one = 1;
    EQ_HIP(hipMemcpy(p_->gate->p, &one, sizeof(int), hipMemcpyHostToDevice));
    for (int i = 0; i < Impl::NRING; ++i) EQ_HIP(hipEventCreateWithFlags(&p_->ev[i], hipEventDisableTiming));
    EQ_HIP(hipEventCreateWithFlags(&p_->mc_ev, hipEventDisableTiming));
}

HeatBatch::~HeatBatch() {
    Device::get().sync();
    for (int i = 0; i < Impl::NRING; ++i)
        if (p_->ev[i]) (void)hipEventDestroy(p_->ev[i]);
    if (p_->host_count) (void)hipHostFree(p_->host_count);
    if (p_->gpin) (void)hipHostFree(p_->gpin);
    if (p_->gflag) (void)hipHostFree(p_->gflag);
    if (p_->mc_ev) (void)hipEventDestroy(p_->mc_ev);
    if (p_->mc_host) (void)hipHostFree(p_->mc_host);
    delete p_;
}

size_t HeatBatch::estimate_bytes(int n, const std::string& precision, bool twophase, bool reduced) {
    const bool mixed = precision == "mixed";
    const size_t ks = mixed ? 4 : 8;
    const size_t h = ((size_t)n + 1) / 2;
    const size_t ncoef = twophase ? 7 : 5;
    const size_t slots = reduced ? h : 2 * h;
    // F, du, r64; Krylov-precision coefficients and dinv (2 h each); Krylov vectors
    size_t b = (size_t)n * 3 * 8 + (ncoef + 1) * 2 * h * ks + (size_t)NVEC * slots * ks;
    if (!twophase) b += (size_t)n * 2 * 8;   // u, u0
    else b += (size_t)n * 8 * 8 + (mixed ? (size_t)n * ncoef * 8 : 0);   // tx ty tz pv src p sw sw0 (+ FP64 J)
    // multigrid levels (full form): 7 coefficients and 2 vectors in 2 h slots, n / 8 + n / 64 + ...
    // cells, and the coarsest inverse
    if (twophase && !reduced) b += ((size_t)n / 7 + 64) * 9 * ks + (size_t)MG_COARSE * MG_COARSE * ks;
    return b + NSLOT * 8 + 64;
}

// Newton / Krylov storage common to every instance kind.  FP64 Jacobian coefficients in
// cell order are stored only for two-phase and linear-system instances in mixed precision
// (their refinement defect reads them); the Krylov-precision coefficients, dinv and the Krylov
// vectors are in red-black rank order.
static void alloc_solver(Inst& I, bool mixed, size_t ks, double inner_tol) {
    const int n = I.n;
    const size_t h = I.h();
    I.F = dev_alloc<double>(n);
    I.du = dev_alloc<double>(n);
    I.r64 = dev_alloc<double>(n);
    if ((I.kind == K_TWOPHASE || I.kind == K_LINEAR) && mixed) {
        I.C = dev_alloc<double>(n);
        I.W = dev_alloc<double>(n);
        I.E = dev_alloc<double>(n);
        I.S = dev_alloc<double>(n);
        I.N = dev_alloc<double>(n);
        if (I.nz > 1) {
            I.B = dev_alloc<double>(n);
            I.T = dev_alloc<double>(n);
        }
    }
    I.sc = dev_alloc<double>(NSLOT);
    I.act = dev_alloc<int>(4);
    EQ_HIP(hipMemsetAsync(I.sc->p, 0, NSLOT * sizeof(double), S()));
    EQ_HIP(hipMemsetAsync(I.act->p, 0, 4 * sizeof(int), S()));
    launch_fill(I.sc->p + S_TOL2, inner_tol * inner_tol, 1);
    launch_fill(I.sc->p + S_ITOL, inner_tol * inner_tol, 1);
    I.kcoef = dev_alloc<char>((size_t)I.ncoef() * 2 * h * ks);
    I.dinv = dev_alloc<char>(2 * h * ks);
    I.vecs = dev_alloc<char>((size_t)NVEC * I.ns() * h * ks);
}

// Row stride of a heat instance: padded to a multiple of 8 cells where its Krylov kernels
// run four pair ranks per lane (reduced form, pair_vectors).
static int heat_stride(int nx, const FusedConfig& cfg) {
    return cfg.pair_vectors && cfg.reduced_form(false) ? (nx + 7) / 8 * 8 : nx;
}
// Cell-order values of the real cells -> the instance's (padded) layout, zeros in the padding.
static std::vector<double> pad_rows(const Inst& I, const double* v) {
    std::vector<double> out((size_t)I.n, 0.0);
    const int nr = I.real_nx(), rows = I.ny * I.nz;
    for (int r = 0; r < rows; ++r) std::copy(v + (size_t)r * nr, v + (size_t)r * nr + nr, out.begin() + (size_t)r * I.nx);
    return out;
}
// The instance's layout -> the real cells (cnt arrays of I.n values each, back to back).
static std::vector<double> unpad_rows(const Inst& I, const std::vector<double>& v, int cnt = 1) {
    if (!I.nxr) return v;
    const int nr = I.nxr, rows = I.ny * I.nz;
    std::vector<double> out((size_t)cnt * I.real_n());
    for (int q = 0; q < cnt; ++q)
        for (int r = 0; r < rows; ++r)
            std::copy(v.begin() + (size_t)q * I.n + (size_t)r * I.nx, v.begin() + (size_t)q * I.n + (size_t)r * I.nx + nr,
                      out.begin() + (size_t)q * I.real_n() + (size_t)r * nr);
    return out;
}

// Multigrid hierarchy of a two-phase instance: halve every axis longer than one cell
// (rounding up) until at most MG_COARSE cells are left.
static void alloc_mg(Inst& I, size_t ks) {
    I.mg_on = true;
    int nx = I.nx, ny = I.ny, nz = I.nz, n = I.n;
    while (n > MG_COARSE && (nx > 1 || ny > 1 || nz > 1)) {
        nx = (nx + 1) / 2;
        ny = (ny + 1) / 2;
        nz = (nz + 1) / 2;
        n = nx * ny * nz;
        Inst::Level L;
        L.nx = nx;
        L.ny = ny;
        L.nz = nz;
        L.n = n;
        L.K = dev_alloc<char>((size_t)7 * 2 * L.h() * ks);
        L.v = dev_alloc<char>((size_t)3 * 2 * L.h() * ks);   // V_R, V_Z, V_T
        I.mg.push_back(std::move(L));
    }
    EQ_REQUIRE(n <= MG_COARSE, "multigrid coarsest level of " << n << " cells");
    I.minv = dev_alloc<char>((size_t)n * n * ks);
    I.mg_tail = 0;
    while (I.mg_tail < (int)I.mg.size() && (I.mg_tail == 0 ? I.n : I.mg[I.mg_tail - 1].n) > MG_TAIL) ++I.mg_tail;
}

int HeatBatch::add(int nx, int ny, double k, double cv, const double* u_host, double dx, double dy) {
    EQ_REQUIRE(nx > 0 && ny > 0, "instance sizes must be positive");
    EQ_REQUIRE((long)nx * ny < (1L << 31) - CPT_MAX * BS, "instance too large");
    auto I = std::make_unique<Inst>();
    I->id = p_->next_id++;
    I->pcm = p_->form_of(K_HEAT);
    const int nxp = heat_stride(nx, p_->cfg);
    I->nx = nxp;
    I->nxr = nxp != nx ? nx : 0;
    I->ny = ny;
    I->n = nxp * ny;
    I->k = k;
    I->cv = cv;
    I->dx = dx;
    I->dy = dy;
    I->u = dev_from_host(pad_rows(*I, u_host).data(), I->n);
    I->u0 = dev_alloc<double>(I->n);
    alloc_solver(*I, p_->mixed, p_->ks(), p_->cfg.inner_tol);
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
    I->pcm = p_->form_of(K_TWOPHASE);
    alloc_solver(*I, p_->mixed, p_->ks(), p_->cfg.inner_tol_twophase);
    if (p_->cfg.twophase_pressure_cg && p_->cfg.twophase_pressure_mg && I->pcm == PM_ILU) alloc_mg(*I, p_->ks());
    I->split = p_->cfg.twophase_split_form && I->mg_on;
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
    p_->flush_copies();   // a queued copy may name this instance's fields
    get_inst(p_, slot);
    p_->tcache.clear();
    p_->insts.erase(slot);
    ++p_->version;
}
int HeatBatch::add_linear_multigrid(int nx, int ny, int nz) {
    EQ_REQUIRE(nx > 0 && ny > 0 && nz > 0, "instance sizes must be positive");
    EQ_REQUIRE((long)nx * ny * nz < (1L << 31) - CPT_MAX * BS, "instance too large");
    auto I = std::make_unique<Inst>();
    I->id = p_->next_id++;
    I->kind = K_LINEAR;
    I->nx = nx;
    I->ny = ny;
    I->nz = nz;
    I->n = nx * ny * nz;
    I->pcm = PM_ILU;   // full red-black form: the V-cycle works on both colours
    alloc_solver(*I, p_->mixed, p_->ks(), p_->cfg.inner_tol);
    alloc_mg(*I, p_->ks());
    const int id = I->id;
    p_->insts[id] = std::move(I);
    ++p_->version;
    return id;
}

void HeatBatch::set_linear_matrix(int slot, const int* rp, const int* ci, const double* val) {
    Impl& P = *p_;
    P.flush_copies();
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_LINEAR && I.mg_on, "slot " << slot << " is not a multigrid linear-system instance");
    TableP tp = P.table({&I});
    Table& t = *tp;
    P.gather_linear(t, rp, ci, val, nullptr);
    // the set-up kernels (Galerkin levels, coarsest inverse) run for instances whose solve
    // is about to start: act[1] set, act[2] = 0 (a CG stage)
    launch_fill_int(I.act->p + 1, 1, 1);
    if (P.mixed) P.mg_setup<float>(t);
    else P.mg_setup<double>(t);
}

void HeatBatch::vcycle(int slot, const double* r, double* z) {
    Impl& P = *p_;
    P.flush_copies();
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_LINEAR && I.mg_on, "slot " << slot << " is not a multigrid linear-system instance");
    TableP tp = P.table({&I});
    Table& t = *tp;
    const unsigned nb = (unsigned)((I.n + BS - 1) / BS);
    launch_fill_int(I.act->p, 1, 1);   // the V-cycle kernels run for active CG instances
    if (P.mixed) {
        hipLaunchKernelGGL(k_cells_to_rb<float>, dim3(nb), dim3(BS), 0, S(), t.desc->p, r, (int)V_R);
        EQ_HIP(hipGetLastError());
        P.vcycle_dots<float>(t);
        hipLaunchKernelGGL(k_rb_to_cells<float>, dim3(nb), dim3(BS), 0, S(), t.desc->p, (int)V_Z, z);
    } else {
        hipLaunchKernelGGL(k_cells_to_rb<double>, dim3(nb), dim3(BS), 0, S(), t.desc->p, r, (int)V_R);
        EQ_HIP(hipGetLastError());
        P.vcycle_dots<double>(t);
        hipLaunchKernelGGL(k_rb_to_cells<double>, dim3(nb), dim3(BS), 0, S(), t.desc->p, (int)V_Z, z);
    }
    EQ_HIP(hipGetLastError());
}

int HeatBatch::add_linear(int nx, int ny, int nz) {
    EQ_REQUIRE(nx > 0 && ny > 0 && nz > 0, "instance sizes must be positive");
    EQ_REQUIRE((long)nx * ny * nz < (1L << 31) - CPT_MAX * BS, "instance too large");
    auto I = std::make_unique<Inst>();
    I->id = p_->next_id++;
    I->kind = K_LINEAR;
    I->nx = nx;
    I->ny = ny;
    I->nz = nz;
    I->n = nx * ny * nz;
    I->pcm = p_->form_of(K_LINEAR);
    alloc_solver(*I, p_->mixed, p_->ks(), p_->cfg.inner_tol);
    const int id = I->id;
    p_->insts[id] = std::move(I);
    ++p_->version;
    return id;
}

HeatBatch::LinearTargets HeatBatch::linear_targets(int slot) {
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_LINEAR, "slot " << slot << " is not a linear-system instance");
    LinearTargets L;
    L.F = I.F->p;
    const size_t ks = p_->ks(), h = I.h();
    for (int q = 0; q < I.ncoef(); ++q) L.K[q] = I.kcoef->p + q * 2 * h * ks;
    if (p_->mixed) {
        const DevPtr<double>* c64[7] = {&I.C, &I.W, &I.E, &I.S, &I.N, &I.B, &I.T};
        for (int q = 0; q < I.ncoef(); ++q) L.C64[q] = (*c64[q])->p;
    }
    L.nx = I.nx;
    L.ny = I.ny;
    L.nz = I.nz;
    L.h = (int)h;
    L.ncoef = I.ncoef();
    L.fp32 = p_->mixed;
    return L;
}

LinearSolveStats HeatBatch::solve_linear(int slot, const int* rp, const int* ci, const double* val, const double* rhs,
                                         double* x, double tol) {
    return solve_linear_impl(slot, rp, ci, val, rhs, x, tol);
}

LinearSolveStats HeatBatch::solve_linear_prepared(int slot, double* x, double tol) {
    return solve_linear_impl(slot, nullptr, nullptr, nullptr, nullptr, x, tol);
}

LinearSolveStats HeatBatch::solve_linear_impl(int slot, const int* rp, const int* ci, const double* val,
                                              const double* rhs, double* x, double tol) {
    Impl& P = *p_;
    P.flush_copies();
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_LINEAR, "slot " << slot << " is not a linear-system instance");
    TableP tp = P.table({&I});
    Table& t = *tp;
    if (rp)
        P.gather_linear(t, rp, ci, val, rhs);
    else
        P.prepared_linear(t);
    P.tol_call = tol;
    try {
        P.solve_linear_table(t);
    } catch (...) {
        P.tol_call = 0.0;
        throw;
    }
    P.tol_call = 0.0;
    // S_ITT: Krylov iterations of the solve; S_REF: refinement rounds; S_RN / S_BN: FP64
    // defect over ||F|| (mixed); S_RR / S_BB: the Krylov residual over its right-hand side
    std::vector<double> g = P.gather(t, make_int4(S_ITT, S_REF, S_RN, S_BN), S_RR);
    std::vector<double> bb = P.gather(t, make_int4(S_BB, S_BB, S_BB, S_BB), S_BB);
    LinearSolveStats st;
    st.iterations = (int)g[0];
    st.refinements = (int)g[1];
    st.relres = P.mixed ? (g[3] > 0 ? g[2] / g[3] : 0.0) : (bb[0] > 0 ? std::sqrt(std::max(g[4], 0.0) / bb[0]) : 0.0);
    if (g[3] == 0.0) st.relres = 0.0;   // zero right-hand side: du = 0
    EQ_HIP(hipMemcpyAsync(x, I.du->p, (size_t)I.n * sizeof(double), hipMemcpyDeviceToDevice, S()));
    return st;
}

void HeatBatch::set_max_newton(int slot, int max_newton) {
    EQ_REQUIRE(max_newton >= 0, "max_newton must be >= 0, got " << max_newton);
    get_inst(p_, slot).max_newton = max_newton;
}
int HeatBatch::size_of(int slot) const { return get_inst(p_, slot).real_n(); }
int HeatBatch::heat_cells(int nx, int ny, const FusedConfig& cfg) { return heat_stride(nx, cfg) * ny; }
bool HeatBatch::in_flight(int slot) const {
    const Inst& I = get_inst(p_, slot);
    return I.state == Inst::START || I.state == Inst::SOLVE;
}
const StepStats& HeatBatch::stats(int slot) const { return get_inst(p_, slot).st; }
size_t HeatBatch::device_bytes(int slot) const {
    const Inst& I = get_inst(p_, slot);
    return estimate_bytes(I.n, p_->cfg.precision, I.kind == K_TWOPHASE, I.pcm == PM_SCHUR);
}

static const int kStage[2] = {0, 1};

void HeatBatch::begin_step(int slot, double dt) {
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.state != Inst::START && I.state != Inst::SOLVE, "slot " << slot << " already has a step in flight");
    if (I.dt != dt) ++p_->version;
    I.dt = dt;
    if (I.kind == K_HEAT) {
        I.u0_pending = true;   // u0 = u in the step's first assembly (S_U0D = 0, cleared with S_F0 below)
    } else {
        // sw0 = sw; pressure solve first, from the previous step's pressure
        p_->pend.push_back({I.sw0->p, I.sw->p, (long)I.n, 0});
        EQ_HIP(hipMemcpyAsync(I.act->p + 2, &kStage[0], sizeof(int), hipMemcpyHostToDevice, S()));
        I.stage = 0;
    }
    I.st = StepStats();
    I.state = Inst::START;
    // S_F0: the step's first Newton iteration sets it; S_U0D (heat): its first assembly
    launch_fill(I.sc->p + S_F0, 0.0, I.kind == K_HEAT ? 2 : 1);
}

std::vector<int> HeatBatch::iterate() {
    Impl& P = *p_;
    if (P.spec.on) {   // left unresolved by an error in the previous call
        P.spec.on = false;
        P.open_gate();
    }
    P.flush_copies();
    std::vector<Inst*> solve, asmb;
    for (auto& kv : P.insts) {
        if (kv.second->state == Inst::SOLVE) solve.push_back(kv.second.get());
        if (kv.second->state == Inst::START || kv.second->state == Inst::SOLVE) asmb.push_back(kv.second.get());
    }
    TableP tps;
    // one device-to-host copy per Newton iteration: ||F||^2, the solve statistics and
    // whether the refinement continues (then u was not updated and F is unchanged)
    std::vector<double> g;
    bool whole = false;   // the iteration ran as one graph (every instance solves)
    // Every instance solves, is a heat instance, and its previous step converged after as
    // many Newton iterations as this one will have done: the assembly after the update
    // computes ||F|| alone (if some instance goes on after all, the full assembly follows).
    int norm_only = !solve.empty() && solve.size() == asmb.size();
    for (const Inst* I : solve)
        if (I->kind != K_HEAT || I->newton_prev <= 0 || I->st.newton_stage + 1 < I->newton_prev) norm_only = 0;
    // A step start of heat instances whose first solve the previous steps predict: the
    // first assembly and the first Newton iteration in one graph; the host reads ||F_0||
    // (S_BN) afterwards, and an instance that had already converged (||F_0|| <=
    // abs_res_tol) found its solve a no-op on the device.
    bool merged = false;
    if (solve.empty() && !asmb.empty() && P.cfg.speculate && P.cfg.max_iter >= 1) {
        bool ok = true;
        int no = 1;   // norm-only assembly: each previous step took one Newton iteration
        for (const Inst* I : asmb) {
            ok &= I->state == Inst::START && I->kind == K_HEAT;
            if (I->newton_prev != 1) no = 0;
        }
        if (ok) {
            tps = P.table(asmb);
            P.allow_spec = true;
            merged = P.mixed ? P.newton_graph_iteration<float>(*tps, g, no, 1)
                             : P.newton_graph_iteration<double>(*tps, g, no, 1);
            P.allow_spec = false;
            if (merged) whole = true, norm_only = no;
        }
    }
    if (!solve.empty()) {
        tps = P.table(solve);
        Table& t = *tps;
        P.allow_spec = P.cfg.speculate;
        if (solve.size() == asmb.size())
            whole = P.mixed ? P.newton_graph_iteration<float>(t, g, norm_only)
                            : P.newton_graph_iteration<double>(t, g, norm_only);
        if (!whole) {
            if (P.mixed) P.linear_solve<float>(t);
            else P.linear_solve<double>(t);
        }
        P.allow_spec = false;
        if (!whole) P.newton_update(t);
    }
    std::vector<int> done;
    if (asmb.empty()) return done;
    TableP tp = whole ? tps : P.table(asmb);
    Table& t = *tp;
    const size_t na = asmb.size();
    if (!whole) {
        P.assemble(t, norm_only);
        g = P.gather(t, make_int4(S_FN, S_ITT, S_REF, S_RN), S_CONT);
    }
    for (Inst* I : asmb) I->u0_pending = false;   // the assembly copied it (S_U0D)
    if (P.spec.on && (P.mixed ? P.finish_spec<float>() : P.finish_spec<double>())) {
        // the blind iterations were too few: the work after the solve was gated off and is
        // queued again now that the solve has finished
        if (P.mixed) P.linear_tail<float>(*tps);
        else P.linear_tail<double>(*tps);
        P.newton_update(*tps);
        P.assemble(t, norm_only);
        g = P.gather(t, make_int4(S_FN, S_ITT, S_REF, S_RN), S_CONT);
    }
    const FusedConfig& c = P.cfg;
    auto process = [&](size_t q, const std::vector<double>& g) {
        Inst& I = *asmb[q];
        if (I.state == Inst::SOLVE && g[4 * na + q] != 0.0) {   // next refinement round
            ++I.round;
            return;
        }
        I.round = 0;
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
        const int mi = I.max_newton > 0 ? std::min(I.max_newton, c.max_iter) : c.max_iter;
        const bool out = I.st.newton_stage >= mi;
        if (I.kind == K_TWOPHASE && I.stage == 0 && !bad && (stop || out)) {
            // pressure solved: the saturation solve starts from sw = 0.5 (twophase.equelle:107)
            I.st.stop_p = stop;
            I.stage = 1;
            EQ_HIP(hipMemcpyAsync(I.act->p + 2, &kStage[1], sizeof(int), hipMemcpyHostToDevice, S()));
            launch_fill(I.sw->p, 0.5, I.n);
            launch_fill(I.sc->p + S_F0, 0.0, 1);   // the transport stage's own first iteration
            I.state = Inst::START;
            return;
        }
        if (stop || bad || out) {
            I.st.stop = bad ? 0 : stop;
            if (I.kind == K_TWOPHASE && I.st.stop_p == 0) I.st.stop = 0;   // pressure did not converge
            I.st.failed = bad;
            I.state = Inst::DONE;
            I.newton_prev = I.st.newton_stage;
            done.push_back(I.id);
        } else {
            I.state = Inst::SOLVE;
        }
    };
    if (merged) {   // the step start as its own assembly's check would have seen it, then the solve
        std::vector<double> g0(6 * na, 0.0);
        for (size_t q = 0; q < na; ++q) g0[q] = g[5 * na + q] * g[5 * na + q];
        for (size_t q = 0; q < na; ++q) process(q, g0);
        for (size_t q = 0; q < na; ++q)
            if (asmb[q]->state == Inst::SOLVE) process(q, g);
    } else {
        for (size_t q = 0; q < na; ++q) process(q, g);
    }
    if (norm_only) {   // an instance goes on: its next solve reads F and the coefficients
        bool more = false;
        for (const Inst* I : asmb) more |= I->state == Inst::SOLVE;
        if (more) P.assemble(t);
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
        iterate();   // a Newton iteration takes one call per refinement round
        EQ_REQUIRE(++guard < 4 * (p_->cfg.max_iter + 2) * (p_->cfg.max_refine + 2), "batch step did not terminate");
    }
}

LaunchShape HeatBatch::launch_shape() const {
    LaunchShape r;
    r.cpt = p_->last_shape.cpt[C_SPMV];
    r.xcd = p_->last_shape.xcd;
    r.signatures = (int)p_->measured.size();   // memberships tuned by measurement
    for (int c = 0; c < NCLS; ++c) r.cpt_class[c] = p_->last_shape.cpt[c];
    return r;
}

double HeatBatch::time_kernel(const std::string& name, int reps, double* event_ms) {
    Impl& P = *p_;
    P.flush_copies();
    EQ_REQUIRE(reps > 0, "reps must be positive");
    std::vector<Inst*> all;
    for (auto& kv : P.insts) all.push_back(kv.second.get());
    EQ_REQUIRE(!all.empty(), "empty batch");
    // the table of the whole membership in its tuned shape
    TableP tp = P.table(all);
    Table& t = *tp;
    // every instance active, so the kernels do their full work (the next solve resets the flags)
    for (Inst* I : all) {
        const int one[2] = {1, 1};
        EQ_HIP(hipMemcpyAsync(I->act->p, one, sizeof(one), hipMemcpyHostToDevice, S()));
        launch_fill(I->sc->p + S_XUPD, 1.0, 1);
    }
    Device::get().sync();
    std::function<void(hipEvent_t, hipEvent_t)> launch;
#define TL(cls, kernel, ...)                                                                                    \
    launch = [&](hipEvent_t s0, hipEvent_t s1) {                                                                \
        by_cpt(t.shape.cpt[cls], [&](auto cpt_) {                                                               \
            constexpr int CP = decltype(cpt_)::value;                                                           \
            const int ci_ = cpt_index(CP);                                                                      \
            hipExtLaunchKernelGGL(kernel, dim3((unsigned)t.nb[ci_]), dim3(BS), 0, S(), s0, s1, 0, t.desc->p,    \
                                  t.blk[ci_]->p, ##__VA_ARGS__);                                                \
        });                                                                                                     \
    }
#define TLP(cls, kernel, ...)                                                                                   \
    launch = [&](hipEvent_t s0, hipEvent_t s1) {                                                                \
        by_cpt(t.shape.cpt[cls], [&](auto cpt_) {                                                               \
            constexpr int CP = decltype(cpt_)::value;                                                           \
            const int ci_ = cpt_index(CP);                                                                      \
            hipExtLaunchKernelGGL(kernel, dim3((unsigned)t.pnb[ci_]), dim3(BS), 0, S(), s0, s1, 0, t.desc->p,   \
                                  t.pblk[ci_]->p, ##__VA_ARGS__);                                               \
        });                                                                                                     \
    }
    bool red = false, full = false, ilu = false;
    for (Inst* I : all) red |= I->pcm == PM_SCHUR, full |= I->pcm != PM_SCHUR, ilu |= I->pcm == PM_ILU;
    bool vec = false, tile = false;
    for (Inst* I : all) vec |= P.desc_of(*I).vec != 0, tile |= P.desc_of(*I).tile != 0;
    if (name == "operator" && tile) {
        launch = [&](hipEvent_t s0, hipEvent_t s1) {
            if (P.mixed)
                hipExtLaunchKernelGGL((k_schur_tile<float>), dim3((unsigned)t.tnb), dim3(TBS), 0, S(), s0, s1, 0,
                                      t.desc->p, t.tblk->p, (int)V_P, (int)V_V, 1, (int)V_RH, t.part->p, t.ng,
                                      (int)(M_BICG | M_CG));
            else
                hipExtLaunchKernelGGL((k_schur_tile<double>), dim3((unsigned)t.tnb), dim3(TBS), 0, S(), s0, s1, 0,
                                      t.desc->p, t.tblk->p, (int)V_P, (int)V_V, 1, (int)V_RH, t.part->p, t.ng,
                                      (int)(M_BICG | M_CG));
        };
    } else if (name == "operator" && vec) {
        if (P.mixed)
            TLP(C_SPMV, (k_schur_vec<float, CP>), (int)V_P, (int)V_V, 1, (int)V_RH, t.part->p, t.ng,
                (int)(M_BICG | M_CG));
        else
            TLP(C_SPMV, (k_schur_vec<double, CP>), (int)V_P, (int)V_V, 1, (int)V_RH, t.part->p, t.ng,
                (int)(M_BICG | M_CG));
    } else if (name == "operator_red" && red) {
        if (P.mixed) TLP(C_SPMV, (k_schur_red<float, CP>), (int)V_PH, (int)(M_BICG | M_CG));
        else TLP(C_SPMV, (k_schur_red<double, CP>), (int)V_PH, (int)(M_BICG | M_CG));
    } else if (name == "operator_black" && red) {
        if (P.mixed)
            TLP(C_SPMV, (k_schur_black<float, CP>), (int)V_PH, (int)V_V, 1, (int)V_RH, -1, t.part->p, t.ng,
                (int)(M_BICG | M_CG));
        else
            TLP(C_SPMV, (k_schur_black<double, CP>), (int)V_PH, (int)V_V, 1, (int)V_RH, -1, t.part->p, t.ng,
                (int)(M_BICG | M_CG));
    } else if (name == "ilu_black" && ilu) {
        if (P.mixed) TLP(C_SPMV, (k_ilu_black<float, CP>), (int)V_P, (int)V_PH);
        else TLP(C_SPMV, (k_ilu_black<double, CP>), (int)V_P, (int)V_PH);
    } else if (name == "ilu_red" && ilu) {
        if (P.mixed) TLP(C_SPMV, (k_ilu_red<float, CP>), (int)V_P, (int)V_PH);
        else TLP(C_SPMV, (k_ilu_red<double, CP>), (int)V_P, (int)V_PH);
    } else if (name == "spmv" && full) {
        if (P.mixed)
            TLP(C_SPMV, (k_spmv_rb<float, CP>), (int)V_PH, (int)V_V, 1, (int)V_RH, -1, t.part->p, t.ng,
                (int)(M_BICG | M_CG));
        else
            TLP(C_SPMV, (k_spmv_rb<double, CP>), (int)V_PH, (int)V_V, 1, (int)V_RH, -1, t.part->p, t.ng,
                (int)(M_BICG | M_CG));
    } else if (name == "assemble") {
        if (P.mixed) TL(C_ASM, (k_assemble<true, CP>), t.part->p, 0);
        else TL(C_ASM, (k_assemble<false, CP>), t.part->p, 0);
    } else if (name == "defect" && P.mixed) {   // the refinement defect exists in mixed precision only
        TL(C_DEF, (k_defect<CP>), t.part->p);
    } else if (name == "b1") {
        if (P.mixed) TLP(C_VEC, (k_b1<float, CP>));
        else TLP(C_VEC, (k_b1<double, CP>));
    } else if (name == "b2") {
        if (P.mixed) TLP(C_VEC, (k_b2<float, CP>));
        else TLP(C_VEC, (k_b2<double, CP>));
    } else if (name == "b3") {
        if (P.mixed) TLP(C_VEC, (k_b3<float, CP>), t.part->p, t.ng);
        else TLP(C_VEC, (k_b3<double, CP>), t.part->p, t.ng);
    } else if (name == "b31") {
        if (P.mixed) TLP(C_VEC, (k_b31<float, CP>));
        else TLP(C_VEC, (k_b31<double, CP>));
    } else if (name == "load_rhs") {
        if (P.mixed) TLP(C_PC, (k_load_rhs<float, CP>), 1, 0);
        else TLP(C_PC, (k_load_rhs<double, CP>), 1, 0);
    } else if (name == "setup" && vec) {
        if (P.mixed) TLP(C_PC, (k_setup_vec<float, CP>), t.part->p, t.ng, 1, 0, 0);
        else TLP(C_PC, (k_setup_vec<double, CP>), t.part->p, t.ng, 1, 0, 0);
    } else {
        EQ_REQUIRE(false, "unknown kernel " << name << " (operator [reduced system, four ranks per lane], "
                                                       "operator_red, operator_black [reduced system], spmv, "
                                                       "ilu_black, ilu_red [full system], assemble, defect [mixed], "
                                                       "b1, b2, b3 [strided form], b31 [four ranks per lane], "
                                                       "load_rhs [strided form], setup [four ranks per lane])");
    }
#undef TL
#undef TLP
    hipEvent_t e0 = nullptr, e1 = nullptr;
    EQ_HIP(hipEventCreate(&e0));
    EQ_HIP(hipEventCreate(&e1));
    launch(nullptr, nullptr);   // warm
    EQ_HIP(hipGetLastError());
    // each launch between its own start and stop events (hipExtLaunchKernel)
    double total = 0.0;
    for (int r = 0; r < reps; ++r) {
        launch(e0, e1);
        EQ_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        EQ_HIP(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    // the same number of launches back to back, one event pair around all of them
    EQ_HIP(hipEventRecord(e0, S()));
    for (int r = 0; r < reps; ++r) launch(nullptr, nullptr);
    EQ_HIP(hipEventRecord(e1, S()));
    EQ_HIP(hipEventSynchronize(e1));
    float batch_ms = 0.f;
    EQ_HIP(hipEventElapsedTime(&batch_ms, e0, e1));
    EQ_HIP(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (event_ms) *event_ms = batch_ms / reps;
    return total / reps;
}

void HeatBatch::get_state(int slot, double* u) const {
    p_->flush_copies();
    Inst& I = get_inst(p_, slot);
    std::vector<double> h = unpad_rows(I, dev_to_host(I.kind == K_HEAT ? I.u->p : I.sw->p, I.n));
    std::copy(h.begin(), h.end(), u);
}
void HeatBatch::get_twophase_state(int slot, double* p, double* sw) const {
    p_->flush_copies();
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
// State uploads copy into the instance's existing buffers: descriptors, tables and
// captured graphs keep their pointers, so nothing is rebuilt.  A queued step-start copy
// (u0 = u, sw0 = sw) is issued first, so an upload inside a step changes the Newton
// start, not the state the step began from.
// A step-start copy left to the first assembly (heat) is done now.
static void settle_u0(Inst& I) {
    if (!I.u0_pending) return;
    EQ_HIP(hipMemcpyAsync(I.u0->p, I.u->p, (size_t)I.n * sizeof(double), hipMemcpyDeviceToDevice, S()));
    launch_fill(I.sc->p + S_U0D, 1.0, 1);
    I.u0_pending = false;
}
static void upload(double* dst, const std::vector<double>& h) {
    EQ_HIP(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, S()));
    Device::get().sync();   // pageable source
}
void HeatBatch::set_state(int slot, const double* u) {
    p_->flush_copies();
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_HEAT, "set_state: slot " << slot << " is not a heat instance");
    settle_u0(I);
    upload(I.u->p, pad_rows(I, u));
}
void HeatBatch::set_twophase_state(int slot, const double* p, const double* sw) {
    p_->flush_copies();
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_TWOPHASE, "slot " << slot << " is not a two-phase instance");
    if (p) upload(I.p->p, std::vector<double>(p, p + I.n));
    if (sw) upload(I.sw->p, std::vector<double>(sw, sw + I.n));
}

// CSR with ascending columns: bottom, south, west, diagonal, east, north, top (existing
// neighbours only; stored zeros kept, as in the AD engine's union patterns).  co holds the
// FP64 coefficients C, W, E, S, N (, B, T), n each.
static void csr_from_stencil(const Inst& I, const std::vector<double>& co, std::vector<int>& rp,
                             std::vector<int>& ci, std::vector<double>& v) {
    const size_t n = I.real_n();   // co: the real cells' coefficients
    const double *C = co.data(), *W = C + n, *E = C + 2 * n, *Sx = C + 3 * n, *N = C + 4 * n;
    const double* Bt = I.nz > 1 ? C + 5 * n : nullptr;
    const double* Tp = I.nz > 1 ? C + 6 * n : nullptr;
    const int nx = I.real_nx(), ny = I.ny, nz = I.nz, nxy = nx * ny;
    rp.assign(n + 1, 0);
    ci.clear();
    v.clear();
    for (int c = 0; c < (int)n; ++c) {
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
// The FP64 coefficients of an instance in cell order, n each: the stored cell-order ones
// (two-phase, mixed precision) or the Krylov ones (fp64 mode) moved out of red-black order.
static std::vector<double> stored_coefs(const Inst& I, bool mixed) {
    std::vector<double> co;
    if (I.C) {
        for (const DevPtr<double>* a : {&I.C, &I.W, &I.E, &I.S, &I.N, &I.B, &I.T}) {
            if (!*a) continue;
            std::vector<double> h = dev_to_host((*a)->p, I.n);
            co.insert(co.end(), h.begin(), h.end());
        }
        return co;
    }
    EQ_REQUIRE(!mixed, "no FP64 coefficients stored");
    const size_t h = I.h(), n = I.n;
    std::vector<double> rb = dev_to_host((const double*)I.kcoef->p, I.ncoef() * 2 * h);
    co.resize(I.ncoef() * n);
    for (int q = 0; q < I.ncoef(); ++q)
        for (size_t c = 0; c < n; ++c) {
            const size_t i = c % I.nx, j = (c / I.nx) % I.ny, k = c / ((size_t)I.nx * I.ny);
            co[q * n + c] = rb[q * 2 * h + ((i + j + k) & 1) * h + (c >> 1)];
        }
    return co;
}

void HeatBatch::export_residual(int slot, const double* u, const double* u0, double dt, std::vector<double>& F,
                                std::vector<int>& rp, std::vector<int>& ci, std::vector<double>& v) {
    p_->flush_copies();
    Inst& I = get_inst(p_, slot);
    EQ_REQUIRE(I.kind == K_HEAT, "slot " << slot << " is not a heat instance");
    I.u = dev_from_host(pad_rows(I, u).data(), I.n);
    I.u0 = dev_from_host(pad_rows(I, u0).data(), I.n);
    I.u0_pending = false;
    launch_fill(I.sc->p + S_U0D, 1.0, 1);   // the given u0, not a step start
    I.dt = dt;
    ++p_->version;
    TableP tp = p_->table({&I});
    p_->assemble(*tp);
    const std::vector<double> Fp = dev_to_host(I.F->p, I.n);
    std::vector<double> co;
    if (!p_->mixed) {
        co = stored_coefs(I, false);
    } else {   // mixed precision: the FP64 rows the refinement defect applies
        auto f = dev_alloc<double>(I.n);
        auto c = dev_alloc<double>(5 * (size_t)I.n);
        hipLaunchKernelGGL(k_heat_rows64, dim3((unsigned)((I.n + BS - 1) / BS)), dim3(BS), 0, S(), tp->desc->p,
                           f->p, c->p);
        EQ_HIP(hipGetLastError());
        co = dev_to_host(c->p, 5 * (size_t)I.n);
        EQ_REQUIRE(dev_to_host(f->p, I.n) == Fp, "FP64 row recomputation differs from the assembly");
    }
    F = unpad_rows(I, Fp);
    csr_from_stencil(I, unpad_rows(I, co, 5), rp, ci, v);
}

void HeatBatch::export_twophase_residual(int slot, int stage, const double* p, const double* sw0, const double* sw,
                                         double dt, std::vector<double>& F, std::vector<int>& rp,
                                         std::vector<int>& ci, std::vector<double>& v) {
    p_->flush_copies();
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
    csr_from_stencil(I, stored_coefs(I, p_->mixed), rp, ci, v);
}

}  // namespace eqhip
