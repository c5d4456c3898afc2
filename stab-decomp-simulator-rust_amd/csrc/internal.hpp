// internal.hpp -- device-buffer descriptors and kernel launchers shared by
// kernels.hip and runtime.hip.  Not part of the public C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace nx {

// Program op codes.  0..10 are the Clifford gates in the numbering of the
// public header (H X Y Z S Sdg SqrtX SqrtXdg CX CZ Swap).
enum : int {
  OP_H = 0, OP_X = 1, OP_Y = 2, OP_Z = 3, OP_S = 4, OP_SDG = 5, OP_SQRTX = 6, OP_SQRTXDG = 7,
  OP_CX = 8, OP_CZ = 9, OP_SWAP = 10,
  OP_PROJ = 20,  // a = qubit, b = outcome
  OP_DISC = 21,  // a = qubit
};

// per-term error codes reported by k_program (the reference panics or
// returns an error at these points)
enum ErrCode : int {
  DEV_OK = 0,
  DEV_ERR_SAME_VEC_MINUS = 1,  // resolve_superposition.rs:120-122 panic
  DEV_ERR_CANNOT_DISCARD = 2,  // discard.rs:116 CannotDiscardQubit
  DEV_ERR_G_ROW = 3,           // discard.rs:140 unreachable
  DEV_ERR_G_COL = 4,           // discard.rs:155 unreachable
};

struct ProgOp {
  int32_t code, a, b, c;
};

// Term tables of chi CH-forms on n qubits, W = ceil(n/64) words per row,
// rows padded to R = 64*W.  G/F/M: [chi][R][W] u64; g: [chi][R] gamma/2;
// v, s: [chi][W]; ph: [chi] phase_factor k mod 8.
struct FormsDev {
  uint64_t *G, *F, *M;
  uint8_t* g;
  uint64_t *v, *s;
  uint8_t* ph;
  int64_t chi;
  int n, W;
};

// Quadratic-form rewrite of each term (see k_qform).
struct QFormDev {
  uint64_t *Q, *B, *C;       // [chi][R][W]
  uint64_t* CT;              // [chi][R][W] C transposed: row j = constraints containing x_j
  uint8_t* a;                // [chi][R]
  uint64_t *alo, *ahi, *rhs; // [chi][W]
  int32_t *nc, *k;           // [chi]
  int32_t* p0;               // [chi] (int32: read with scalar loads)
  int4* meta;                // [chi] {nc, k, p0, 0}: one scalar load per term in k_pairs
  int64_t chi;
  int n, W;
};

// Coefficients: exact Scalars when scalar != 0, packed one int32 per term
// (cs = -1 for Scalar::Zero, else k mod 8 | (r + CS_ROFF) << 3), else the
// complex128 values c.
enum : int { CS_ROFF = 1024 };
struct CoefDev {
  int scalar;
  const int32_t* cs;
  const double2* c;
};
__host__ __device__ inline int cs_k(int32_t v) { return v & 7; }
__host__ __device__ inline int cs_r(int32_t v) { return (v >> 3) - CS_ROFF; }

// Ket-side variants of a term set, bit-sliced: variant k of term t is the
// term's quadratic form with
//   a_r  -> a_r + 2 * bit k of hi[(t*R + r)*VW + k/64]      (row r)
//   rhs_c -> rhs_c ^ bit k of rhs[(t*R + c)*VW + k/64]       (constraint c)
//   p0   -> p0 + sum_b 2^b * bit k of ph[(t*3 + b)*VW + k/64] (mod 8)
// and the same B, constraint rows and magnitude.  A Pauli image P phi_t is
// such a variant: a Pauli left-multiplication changes only s, gamma and the
// phase of the CH-form, so G, F, M and v (hence B and the constraint rows)
// stay those of phi_t.
struct VarDev {
  const uint64_t* hi;
  const uint64_t* rhs;
  const uint64_t* ph;
};

// One pair job: the Hermitian sum over kets f in [0, nfixed) of Y and bras
// o in [f, nother) (triangle) or [0, nother) of X, for every variant of
// the ket; the job owns waves [wave0, wave0 + nwaves) of the launch.
// row_off[f] is the first unit (J consecutive bras) of ket row f.
struct PairJobDev {
  QFormDev X, Y;
  VarDev V;
  CoefDev cx, cy;
  const int64_t* row_off;
  int64_t nfixed, nother;
  int32_t triangle, nvar;
  int64_t wave0, nwaves;
  int64_t out_base;  // variant k -> out[out_base + k]
  int64_t base_out;  // >= 0: the base (ket itself, identity) -> out[base_out]
  // unit constraints (sampling): x_{uq[u]} = ub[u*(VW+1)] for the base, the
  // variants in masks ub[u*(VW+1) + 1 .. VW] take the other value; they are
  // substituted into every term by launch_qproject_var (the pair kernel only
  // reads the projected forms, so nunit > 0 requires termvar)
  int32_t nunit;
  const int32_t* uq;
  const uint64_t* ub;
  // termvar != 0 (with units): X = Y are forms projected onto the window
  // outcomes by launch_qproject_var and V holds their variant tables
  int32_t termvar;
  // split_ab != 0: the variant sums are written as their dyadic part A at
  // out[out_base + k] and sqrt(2) part B at out[out_base + nvar + k]
  int32_t split_ab;
  // variables already substituted in X and Y (forms projected by
  // launch_qproject or launch_qproject_var): they take no part in the
  // elimination
  uint64_t removed[4];
  // reduction segments seg0 .. seg0 + nseg - 1 (RED_SEG waves each)
  int64_t seg0;
  int32_t nseg;
  // non-null (termvar windows, VW = 1): every pair (ket f, bra o >= f) also
  // writes its exact per-variant values as a record of FOLD_REC words at
  // rec[FOLD_REC * (f chi - f (f - 1) / 2 + o - f)]: the vanishing-variant
  // mask, the three Z8 phase planes, then (c0 + dv) & 7 | mag << 8, so that
  // variant k's value is zeta^((c0 + dv + D_k) & 7) 2^(-mag/2)
  uint64_t* rec;
};

// words per pair record (see PairJobDev::rec) and node slots per window in
// the reference-order fold (nodes of relative depth 1..6: 2 + 4 + .. + 64)
enum : int { FOLD_REC = 5, FOLD_NODES = 128 };

// waves per first-stage reduction segment
enum : int { RED_SEG = 128 };

// FP64 tables computed on the host with the same libm calls as the
// reference (cos/sin of k*pi/4, powf(2, -r/2)).
enum : int { TAB_COS = 0, TAB_SIN = 8, TAB_RMAX = 1024, TAB_POW = 16 + TAB_RMAX, TAB_SIZE = 16 + 2 * TAB_RMAX + 1 };

hipError_t launch_init_zero(const FormsDev& out, hipStream_t st);
hipError_t launch_program(const FormsDev& in, const FormsDev& out, const ProgOp* prog, int nprog, int4* status,
                          hipStream_t st);
hipError_t launch_kron(const FormsDev& A, const FormsDev& B, const FormsDev& O, hipStream_t st);
hipError_t launch_gather(const FormsDev& in, const FormsDev& O, const int64_t* idx, hipStream_t st);
// tab (one word per row only): also the amplitude tables of every term,
// amp_table_bytes(1) bytes each
hipError_t launch_qform(const FormsDev& in, const QFormDev& q, hipStream_t st, void* tab = nullptr);
// bitstrings per thread of the amplitude kernel
inline int amp_strings_per_thread(int W) { return W == 1 ? 3 : (W == 2 ? 4 : 2); }
// number of term chunks (blockIdx.y) of an amplitude launch over nb strings
// and chi terms: the launch fills whole rounds of the resident blocks, with
// at least AMP_MIN_TERMS terms per chunk and at most 64 chunks
int amp_chunks(int W, int64_t nb, int64_t chi);
// chunk c covers terms [c chi / nchunks, (c + 1) chi / nchunks); nchunks > 1:
// partial must hold nchunks * nb values
// per-term nibble tables of the amplitude kernel: amp_table_bytes(W) bytes
// per term, written for all chi terms of q by launch_amp_tables and read by
// launch_amplitude
size_t amp_table_bytes(int W);
hipError_t launch_amp_tables(const QFormDev& q, void* tab, hipStream_t st);
hipError_t launch_amplitude(const QFormDev& q, const void* tab, const double2* coef, const uint64_t* bits, int64_t nb,
                            int nchunks, const double* tabs, double2* out, double2* partial, uint8_t* per_term,
                            int enumerate, double gre, double gim, hipStream_t st);
hipError_t launch_amp_reduce(const double2* partial, int64_t nb, int nchunks, double2* out, double gre, double gim,
                             hipStream_t st);
// Sampling tails (tails.hip): job j holds the completions x = base | y of
// a path on its last d <= 16 qubits qs (bit i of y -> qubit qs[i]); for
// each term t of q, k_tail_restrict writes a record of TAIL_REC words to
// recs (the term restricted to the d tail variables), k_tail_eval sums the
// terms' exact values into the components al[(c nv + y) 4 + m] of term
// chunk c < nchunks (blocks block0 .. block0 + cols * nchunks - 1, cols =
// ceil(nv / 1024)), k_tail_norms the chunks in order into the leaf norms
// ab[y] = A, ab[nv + y] = B (|psi(x)|^2 / |g|^2 = A + sqrt2 B).  term0 and
// str0: the job's first term (restrict grid) and completion (norms grid)
// in the launch.  All jobs of one launch have row word width W <= 2.
constexpr int TAIL_REC = 68;
struct TailJobDev {
  QFormDev q;
  const int32_t* cs;
  const uint64_t* base;
  const int32_t* qs;
  uint32_t* recs;
  double* al;
  double* ab;
  int32_t d, nchunks;
  int64_t nv, term0, block0, str0;
};
// term chunks of a tail of nv completions over chi terms: enough blocks
// (completion columns of 1024 x chunks) for about 512 per tail, at least 16
// terms per chunk, at most 64 chunks
inline int tail_chunks(int64_t chi, int64_t nv) {
  const int64_t cols = (nv + 1023) / 1024;
  const int64_t want = (512 + cols - 1) / cols;
  return (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(want, chi / 16), 64));
}
hipError_t launch_tails(int W, const TailJobDev* jobs, int njobs, int64_t total_terms, int64_t total_blocks,
                        int64_t total_strings, hipStream_t st);
// Pair sums over units of J = 8 consecutive other indices o of a fixed
// index f (triangle: o in [f, nother), nfixed == nother, rows from the
// closed form below; else o in [0, nother)), four waves per block, nblocks
// blocks; per block partial[block] = {A_re, A_im, B_re, B_im}: the sum is
// A + sqrt(2) B, accumulated exactly (powers of two) for Scalar
// coefficients, B = 0 for complex128 coefficients.  With ev0 and ev1 the
// launch itself carries the two timing events (start and end of the
// kernel, no separate event records on the stream).
hipError_t launch_pairs(const QFormDev& X, const QFormDev& Y, int64_t nfixed, int64_t nother, int triangle, int J,
                        int mode, const CoefDev& cx, const CoefDev& cy, const double* tabs, double4* partial,
                        int nblocks, uint32_t* exact_out, int64_t exact_ld, hipStream_t st,
                        hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// sum of n double4 partials in a fixed order (one block) into out (if not
// null); with norm_out, sqrt((A_re + sqrt2 B_re) g2) into norm_out[0]
hipError_t launch_reduce(const double4* partial, int64_t n, double4* out, hipStream_t st,
                         double* norm_out = nullptr, double g2 = 1.0);

// Work units of the pair triangle {(f, o) : f <= o < chi} cut into units of
// J consecutive o: S(M) = sum_{m=1..M} ceil(m / J) = J q (q + 1) / 2 +
// r (q + 1) for M = q J + r, so row f starts at unit S(chi) - S(chi - f)
// and the triangle holds S(chi) units.
__host__ __device__ inline int64_t tri_units_upto(int64_t M, int J) {
  const int64_t q = M / J, r = M % J;
  return (int64_t)J * q * (q + 1) / 2 + r * (q + 1);
}

// Pauli variants of every term of `in`: variant k applies the ops
// vops[voff[k] .. voff[k+1]) (q << 2 | op, op 1 = X, 2 = Y, 3 = Z) in order.
hipError_t launch_pauli_var(const FormsDev& in, const int32_t* voff, const int32_t* vops, int nvar, int VW,
                            uint64_t* vhi, uint64_t* vrhs, uint64_t* vph, hipStream_t st);
// waves [wave_base, wave_base + nwaves) of a launch plan whose jobs all have
// word width W; wave_job[w] = job of wave w; partial holds 2 (64 * VW + 1)
// doubles per wave of the plan: the real Hermitian sums per variant and the
// base, then their sqrt(2) parts (exact sums of Scalar-coefficient states
// are A + sqrt(2) B with A, B sums of powers of two).  units: 0 Pauli
// variants, 1 unit constraints, 2 unit constraints and pair records (VW = 1)
hipError_t launch_pairs_var(int W, int VW, int units, const PairJobDev* jobs, const int32_t* wave_job,
                            int64_t wave_base, int64_t nwaves, int J, const double* tabs, double* partial,
                            hipStream_t st);
// waves of k_pairs_var<W, VW, units> resident on the current device at once
// (occupancy x CUs x 4 waves per 256-thread block); 0 for an unknown shape
int64_t pairs_var_resident_waves(int W, int VW, int units);
// out[jobs[j].out_base + k] = sum over the job's waves of partial (k < nvar),
// and the base sum to out[jobs[j].base_out] when that is >= 0: first per
// segment of RED_SEG waves (seg_job[s] = job of segment s) into seg_part,
// then per job over its segments, both in a fixed order
hipError_t launch_reduce_var(const PairJobDev* jobs, int njobs, int VW, const double* partial, const int32_t* seg_job,
                             int64_t nseg, double* seg_part, double* out, hipStream_t st);
// atomicMin into *dmin (entering as ncoord + 1) of the smallest D such that
// some pair value <X_o| Z^t |X_f> (o >= f) with t != 0 supported on the
// first D coordinates is nonzero; coord[r] = coordinate of qubit r or -1;
// removed (W words, or null) = variables substituted already;
// VW = ceil(ncoord / 64); offsets and J as for launch_pairs (triangle)
hipError_t launch_pairs_sym(const QFormDev& X, const int64_t* offsets, int J, int VW, const int32_t* coord,
                            const uint64_t* removed, int* dmin, int nblocks, hipStream_t st);
// out = in with x_{uq[u]} = ub[u] substituted into every term's form, u <
// nunit (the fields the pair kernels read: B, C, a, rhs, p0, nc, k)
// projection onto the fixed outcomes fq/fb (as launch_qproject), then onto
// a window with one variant per outcome (tables in vout: hi and rhs
// [chi][64W][VW], ph [chi][3][VW]); uq, ub as the unit constraints of a job
hipError_t launch_qproject_var(const QFormDev& in, const QFormDev& out, const VarDev& vout, int VW,
                               const int32_t* fq, const int32_t* fb, int nfix, const int32_t* uq,
                               const uint64_t* ub, int nunit, hipStream_t st);
// k_qproject_var for several states in one launch (one row word width W and
// VW per launch): job j projects the terms of jobs[j].in as launch_qproject_var
// does with its own fixed and look-ahead outcomes, in blocks jobs[j].block0
// .. jobs[j + 1].block0 - 1 (four terms per block); nblocks blocks in all
struct QProjJobDev {
  QFormDev in, out;
  VarDev vout;
  const int32_t* fq;
  const int32_t* fb;
  const int32_t* uq;
  const uint64_t* ub;
  int32_t nfix, nunit;
  int64_t block0;
};
hipError_t launch_qproject_var_jobs(int W, int VW, const QProjJobDev* jobs, int njobs, int64_t nblocks,
                                    hipStream_t st);
hipError_t launch_qproject(const QFormDev& in, const QFormDev& out, const int32_t* uq, const int32_t* ub, int nunit,
                           hipStream_t st);
// Coefficient backward (k_grad_var): for the Pauli variants V of the terms
// of X (up to 64 VW strings, weights wts) and coefficient values c, per
// tile of GT x GT pairs the row and column partials of y = G c, G =
// sum_k w_k G_k; per wave (nblocks * 4 waves) the 64 VW per-string
// Hermitian sums in partial.  tile_row and tile_col hold grad_tiles(chi) *
// grad_tile_size() values each.
int64_t grad_tiles(int64_t chi);
int grad_tile_size();
hipError_t launch_grad_var(const QFormDev& X, const VarDev& V, int VW, const double2* c, int nvar, const double* wts,
                           const double* tabs, int nblocks, double2* tile_row, double2* tile_col, double* partial,
                           hipStream_t st);
// y (+)= scale * (G c) from the tile partials, per index in a fixed order
hipError_t launch_grad_reduce(const double2* tile_row, const double2* tile_col, int64_t chi, double scale,
                              int accumulate, double2* y, hipStream_t st);
// vals[k] = scale * sum over the nwaves per-wave partials of string k
// (partial[w * stride + k]), k < nvar
hipError_t launch_grad_vals(const double* partial, int64_t nwaves, int stride, int nvar, double scale, double* vals,
                            hipStream_t st);
// *loss = sum_k wts[k] vals[k], k < n, in order
hipError_t launch_grad_loss(const double* vals, const double* wts, int n, double* loss, hipStream_t st);

// Reference-order norms of the nodes of a sampling window (fold.hip).
// k_fold_diag: D[f][u] = number of non-deterministic projections of term f
// on the path to node u (255: the term does not survive), from the records
// of the diagonal pairs.  k_fold_terms: the FP64 value the reference adds
// for pair p (reference order) at node u, for pairs [p0, p0 + np); per
// 256-pair chunk c the nonzero ones, in pair order, at t[u * ld + 256 c ...]
// (ld: a multiple of 256, at least np) and their count at cnt[u * nch + c]
// (nch = ceil(np / 256)).  k_fold_sum: sums[u] += those values in pair
// order, one wave per node.  err (device int) is set to 1 when a node value
// is not of the form zeta^K 2^(-R/2) or 0.
hipError_t launch_fold_diag(const uint64_t* rec, int64_t chi, int d, uint8_t* D, int* err, hipStream_t st);
hipError_t launch_fold_terms(const uint64_t* rec, int64_t chi, int64_t p0, int64_t np, int64_t ld, int d,
                             const uint8_t* D, const int32_t* cs, const double* tabs, double* t, int32_t* cnt,
                             int* err, hipStream_t st);
hipError_t launch_fold_sum(const double* t, const int32_t* cnt, int64_t np, int64_t ld, int nnodes, double* sums,
                           hipStream_t st);

}  // namespace nx
