// Host launchers for the HIP kernels of the AD engine, the grid operators and
// the reductions.  All launches go to Device::get().stream().
#pragma once
#include <cstdint>

namespace eqhip {

enum EwOp { EW_ADD = 0, EW_SUB = 1, EW_MUL = 2, EW_DIV = 3 };
enum CmpOp { CMP_GT = 0, CMP_GE = 1, CMP_LT = 2, CMP_LE = 3, CMP_EQ = 4, CMP_NE = 5 };
enum RedOp { RED_SUM = 0, RED_PROD = 1, RED_MAX = 2, RED_MIN = 3, RED_SUMSQ = 4 };

void launch_row_of(const int* rp, int rows, int* row);
void launch_fill(double* out, double v, long n);
void launch_fill_int(int* out, int v, long n);
void launch_ew(int op, const double* a, const double* b, double* out, int n);
// op EW_MUL: s*a ; EW_DIV: scalar_left ? s/a : a/s
void launch_ew_scalar(int op, double s, const double* a, double* out, int n, int scalar_left);
void launch_sqrt(const double* a, double* out, int n);
// Per-row coefficient helpers for the AD product / quotient rules.
//   kind 0: out = 1/b            kind 1: out = -a/(b*b)
//   kind 2: out = -s/(a*a)       kind 3: out = 0.5/a
//   kind 4: out = 1/(b*b)
void launch_coef(int kind, const double* a, const double* b, double s, double* out, int n);
void launch_compare(int op, const double* a, const double* b, double s, int scalar_left, uint8_t* out, int n);
void launch_select(const uint8_t* p, const double* a, const double* b, double* out, int n);
void launch_select_int(const uint8_t* p, const int* a, const int* b, int* out, int n);
void launch_is_empty(const int* idx, uint8_t* out, int n);
// out[i] = map[i] >= 0 ? in[map[i]] : 0
void launch_gather(const double* in, const int* map, double* out, long n);
void launch_gather_int(const int* in, const int* map, int* out, int n);

// Jacobian numeric kernels (one thread per output entry).
//   out[k] = C(r) * ( alpha*A(r)*RA(r)*a[ia[k]] + beta*B(r)*RB(r)*b[ib[k]] )
// with r = row[k]; a/b may be null (term absent); ia/ib null => identity map;
// ia/ib entries -1 => term absent for that entry; A, RA, B, RB, C null => 1.
void launch_jlin(long nnz, const int* row, const double* a, const int* ia, const double* A, const double* RA,
                 double alpha, const double* b, const int* ib, const double* B, const double* RB, double beta,
                 const double* C, double* out);
// out[k] = p[r] ? sa*RA(r)*a[ia[k]] : sb*RB(r)*b[ib[k]]  (missing => 0)
void launch_jselect(long nnz, const int* row, const uint8_t* p, const double* a, const int* ia, double sa,
                    const double* RA, const double* b, const int* ib, double sb, const double* RB, double* out);
// out[k] = scale * sum_{t in [seg[k], seg[k+1])} mv[mi[t]] * ar(arow[ai[t]]) * av[ai[t]]
// (ar: row factors of the operand, indexed by the row of its entry; null => 1)
void launch_spgemm_num(long nnz, const int* seg, const int* mi, const int* ai, const double* mv, const double* av,
                       const int* arow, const double* ar, double scale, double* out);
// the same for a matrix of +-1 entries: sai[t] = operand entry, or ~entry for a -1
void launch_spgemm_signed(long nnz, const int* seg, const int* sai, const double* av, const int* arow,
                          const double* ar, double scale, double* out);
// out[k] = s * R(row[k]) * v[map[k]]  (R null => 1, map null => k)
void launch_jscale(long nnz, const int* row, double s, const double* R, const int* map, const double* v,
                   double* out);

// ---- symbolic passes on the device (CSR patterns with ascending columns per row) ----
// Sum over the entries of the content hash of a pattern (rows + 1 row pointers at
// positions 0.., nnz column indices at positions 2^31 + k; see pattern_mix).
uint64_t pattern_entries_hash(const int* rp, int nrp, const int* ci, int nnz);
// Whether two device int arrays of n entries differ anywhere.
bool int_arrays_differ(const int* a, const int* b, long n);
// out[0..n) = exclusive prefix sums of in[0..n) (out[0] = 0).
void exclusive_scan(const int* in, int* out, int n);
// Union of two patterns: entries per row (cnt[rows] = 0), then the columns and, per output
// entry, the entry of each operand (-1 where absent).
void launch_union_count(int rows, const int* arp, const int* aci, const int* brp, const int* bci, int* cnt);
void launch_union_fill(int rows, const int* arp, const int* aci, const int* brp, const int* bci, const int* orp,
                       int* oci, int* ia, int* ib);
// Row gather: output row i = input row map[i] (-1: empty): entries per row, then the
// columns and each entry's source entry.
void launch_rows_count(int out_rows, const int* arp, const int* map, int* cnt);
void launch_rows_fill(int out_rows, const int* arp, const int* aci, const int* map, const int* orp, int* oci,
                      int* src);
// Structural product M * A: per row the distinct output columns (ucnt) and the
// contributions (ccnt; ucnt[rows] = ccnt[rows] = 0); *over = 1 when a row has more than
// cmax contributions.  The fill writes the columns ascending and, per output entry, its
// contributions in matrix-entry order: (mi, ai) pairs, or sai (ai, or ~ai where msign < 0)
// when msign is given.
void launch_spgemm_count(int rows, const int* mrp, const int* mci, const int* arp, const int* aci, int cmax,
                         int* ucnt, int* ccnt, int* over);
void launch_spgemm_fill(int rows, const int* mrp, const int* mci, const int* arp, const int* aci,
                        const int8_t* msign, const int* orp, const int* cp, int* oci, int* seg, int* mi, int* ai,
                        int* sai);

// y = A x, one thread per row, entries summed in column order.
void launch_csr_spmv(int rows, const int* rp, const int* ci, const double* v, const double* x, double* y);

// Grid operators on values.
// gradient: out[i] = u[c2[i]] - u[c1[i]]
void launch_gradient(const double* u, const int* c1, const int* c2, double* out, int nif);
// divergence: out[c] = sum_k sgn * flux[f] over the cell's faces; signed_faces[k] = f (cell is
// first cell => +) or ~f (cell is second cell => -).
void launch_divergence(const double* flux, const int* facepos, const int* signed_faces, double* out, int nc);
// face_cells -> interior / boundary flags; cells -> boundary flag
void launch_face_flags(const int* face_cells, int nf, uint8_t* interior);
void launch_cell_boundary_flags(const int* face_cells, const int* facepos, const int* cell_faces, int nc,
                                uint8_t* boundary);
// geometry gathers: out[i*dim + d] = src[idx[i]*dim + d]
void launch_gather_vec(const double* src, const int* idx, int n, int dim, double* out);
void launch_gather_face_cells(const int* face_cells, const int* faces, int n, int which, int* out);
void launch_vec_norm(const double* v, int n, int dim, double* out);
void launch_vec_dot(const double* a, const double* b, int n, int dim, double* out);
void launch_vec_ew(int op, const double* a, const double* b, int count, double* out);
void launch_vec_scale(const double* v, const double* s, double sc, int n, int dim, int divide, double* out);
void launch_unit_normals(const double* normals, const int* idx, int n, int dim, double* out);
void launch_vec_col(const double* v, int n, int dim, int col, double* out);

// Deterministic reductions (fixed grid and order for a given n).
double reduce(int op, const double* a, long n);
// The same reduction with its partials in a fixed device array instead of a pool block:
// no device memory is allocated or freed.  NOALLOC_SCRATCH_DOUBLES bounds the partials
// (reduce uses at most 1024 blocks, so every n fits).  Calls overlapping on one GPU race.
constexpr int NOALLOC_SCRATCH_DOUBLES = 2048;
double reduce_noalloc(int op, const double* a, long n);
// Device address of that array on the current device (looked up once per device).
double* noalloc_scratch();
// Looks the address up for the current device (called at device set-up).
void noalloc_scratch_init();
// Deterministic dot product into a device scalar (no host sync).
void launch_dot_to(const double* a, const double* b, long n, double* d_out);

// Stream compaction of flagged indices: writes the ascending indices i with
// flag[i] != 0 (flag == want) into out, returns the count.
int compact_flags(const uint8_t* flags, int n, int want, int* out);

}  // namespace eqhip
