/*
 * necstar_hip.h -- C-ABI of the MI355X (gfx950) stabilizer-decomposition
 * hot path: the drop-in boundary under the Python module `necstar`.
 *
 * One nx_state is one QuantumState of the reference
 * (crates/necstar/src/state/mod.rs:67-707): chi CH-form terms (device
 * resident, bit-packed u64 rows), their coefficients and a global factor.
 * Every function returns NX_OK (0) or an error code; the message of the last
 * error on the calling thread is nx_last_error().  Messages reproduce the
 * reference's Display strings (crates/necstar/src/error/mod.rs:10-91,
 * crates/stabilizer-ch-form-rust/src/error/mod.rs:7-67) so the binding layer
 * can raise ValueError(message) exactly as python/src/state.rs does.
 *
 * Gate codes (op): 0 H, 1 X, 2 Y, 3 Z, 4 S, 5 Sdg, 6 SqrtX, 7 SqrtXdg,
 * 8 CX, 9 CZ, 10 Swap, 11 T, 12 Tdg, 13 CCX.  A gate list is int32
 * [ngates][4] = (op, q0, q1, q2).  Bitstrings are little-endian (qubit j is
 * bit j) packed into `words` u64 per string.  Pauli ops: 0 I, 1 X, 2 Y, 3 Z.
 * Seeds are 32 bytes (python/src/utils.rs:4-39 packs ints into them); NULL
 * means "from entropy".
 */
#ifndef NECSTAR_HIP_H
#define NECSTAR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nx_state nx_state;
typedef struct nx_batch nx_batch;

enum {
  NX_OK = 0,
  NX_E_VALUE = 1,    /* reference Error -> PyValueError                     */
  NX_E_TYPE = 2,     /* seed type error -> PyTypeError                      */
  NX_E_DEVICE = 3,   /* HIP runtime failure / no usable gfx950 device       */
  NX_E_INTERNAL = 4  /* a reference panic!/unreachable! state               */
};

/* Error message of the last failed call on this thread. */
const char *nx_last_error(void);

/* Selects the HIP device used by later calls on this thread (default 0)
 * and checks it is a gfx950.  No reference counterpart. */
int nx_init(int device);

/* use != 0: later work of the calling thread on its current device goes to
 * `stream` (a hipStream_t; NULL is the null stream) instead of the
 * library's own non-blocking stream, so that it is ordered with the
 * caller's other work on that stream without events; use == 0: back to the
 * library's stream.  Work already enqueued is waited for first.  Batch
 * graphs are still captured on the library's stream.  No reference
 * counterpart. */
int nx_set_stream(void *stream, int use);

/* QuantumState::from_circuit -> StabDecompCompiler::compile
 * (crates/necstar/src/state/mod.rs:98-102, state/compiler/mod.rs:44-140). */
int nx_state_compile(int num_qubits, int64_t ngates, const int32_t *gates, nx_state **out);

/* Clone / free (PyO3 value semantics, python/src/state.rs:12-15). */
int nx_state_clone(const nx_state *s, nx_state **out);
void nx_state_free(nx_state *s);

/* num_qubits / stabilizer_rank getters (state/mod.rs:686-696). */
int nx_state_num_qubits(const nx_state *s, int32_t *out);
int nx_state_stabilizer_rank(const nx_state *s, int64_t *out);

/* QuantumState::apply_gate / apply_gates (state/mod.rs:529-671,
 * stabilizer_decomposed_state/gates.rs:7-37). */
int nx_state_apply_gates(nx_state *s, int64_t ngates, const int32_t *gates);

/* QuantumState::norm (state/mod.rs:702-706, norm.rs:8-35). */
int nx_state_norm(const nx_state *s, double *out);
int nx_state_norm_squared(const nx_state *s, double *out);

/* nx_state_norm into device memory d_out[0] (8 bytes), asynchronously: the
 * same ordering against `stream` as nx_state_amplitudes_async, and the same
 * value as nx_state_norm.  No reference counterpart (for device-resident
 * pipelines; norm() itself returns to the host). */
int nx_state_norm_async(const nx_state *s, double *d_out, void *stream);

/* nx_state_norm that allocates nothing: no reference counterpart; for
 * callers that cannot allocate (tight serving loops, harnesses that forbid
 * hipMalloc).  Every device buffer of the call (quadratic forms, coefficient
 * copies, pair-unit offsets, partial sums, result) lives in one file-scope
 * device array of 64 MiB per GPU, held until the process ends: up to about
 * 30,000 terms at n <= 64 or 7,900 at n <= 128; larger states fail with
 * NX_E_VALUE.  Same kernels and summation order as nx_state_norm (equal
 * results).  Calls on one GPU share that array: overlapping calls on one GPU
 * (several threads) race.  Call nx_init on the device first (it sets the
 * device up once, which allocates). */
int nx_state_norm_noalloc(const nx_state *s, double *out);

/* QuantumState::exp_value (state/mod.rs:309-313, exp_value.rs:7-61).
 * dense != 0: ops[len] little-endian (PauliString::Dense);
 * dense == 0: sparse terms (ops[i], qubits[i]) (PauliString::Sparse). */
int nx_state_exp_value(const nx_state *s, int dense, int32_t len, const int32_t *ops, const int32_t *qubits,
                       double *out);

/* QuantumState::inner_product <a|b> (state/mod.rs:165-172,
 * stabilizer_decomposed_state/inner_product.rs:9-19). */
int nx_state_inner_product(const nx_state *a, const nx_state *b, double *re, double *im);

/* <x|psi> for nbits bitstrings (host buffers).  out: complex128 [nbits]
 * interleaved (re, im).  per_term (optional, may be NULL): uint8
 * [chi][nbits], 0xFF for a zero amplitude else the phase k (mod 8) of
 * <x|phi_i> before omega; the exponent r of term i is per_term_r[i]
 * (optional).  Same arithmetic as to_statevector (statevector.rs:12-24,
 * form/amplitude.rs:28-44) per string.  An API addition: the reference
 * reaches amplitudes only through to_statevector / inner_product. */
int nx_state_amplitudes(const nx_state *s, int64_t nbits, const uint64_t *bits, int32_t words, double *out,
                        uint8_t *per_term, int32_t *per_term_r);
/* Same with device pointers (inputs already resident in HBM). */
int nx_state_amplitudes_device(const nx_state *s, int64_t nbits, const uint64_t *d_bits, int32_t words,
                               double *d_out);

/* The same, asynchronously: the pass is enqueued on the library's stream
 * after the work already enqueued on `stream` (a hipStream_t, e.g. PyTorch's
 * current stream; NULL is the null stream), and later work on `stream` is
 * ordered after it; the call returns without waiting.  Results are also
 * complete when any later call that returns host data (norm, exp_value, ...)
 * returns.  No reference counterpart. */
int nx_state_amplitudes_async(const nx_state *s, int64_t nbits, const uint64_t *d_bits, int32_t words,
                              double *d_out, void *stream);

/* QuantumState::to_statevector (statevector.rs:12-24): out complex128 [2^n]. */
int nx_state_statevector(const nx_state *s, double *out);

/* project_normalized / project_unnormalized (state/mod.rs:319-351,
 * projection.rs:10-57). */
int nx_state_project(nx_state *s, int32_t qubit, int32_t outcome, int32_t normalized);

/* QuantumState::discard (state/mod.rs:361-365, discard.rs:5-12). */
int nx_state_discard(nx_state *s, int32_t qubit);

/* measure / measure_all (state/mod.rs:190-227, measurement.rs:10-103). */
int nx_state_measure(nx_state *s, int32_t nq, const int32_t *qargs, const uint8_t *seed32, uint8_t *outcomes);

/* sample (state/mod.rs:247-284, sampling.rs:13-189): distinct outcomes in
 * DFS order, outcome k as bits little-endian in 2 u64 words (bit i <->
 * qargs[i]) and its count.  Capacity max_out; *nout receives the count,
 * also when the call fails because max_out is too small (call again with
 * that capacity and the same seed).  Every node probability is the
 * reference's FP64 value: norm_squared() of the projected state summed over
 * its surviving terms in the reference's pair order (norm.rs:8-30,
 * projection.rs:35-51), so a seed gives the reference's outcomes.  Holds
 * one 40-byte record per term pair: at most 23,169 terms. */
int nx_state_sample(const nx_state *s, int32_t nq, const int32_t *qargs, uint64_t shots, const uint8_t *seed32,
                    int64_t max_out, uint64_t *out_bits, uint64_t *out_counts, int64_t *nout);

/* sample with exact node probabilities (arguments and results as
 * nx_state_sample): exact sums of the pair values, leading levels whose
 * marginal is certified uniform drawn at p = 1/2 without norms, and the
 * last <= 14 levels from exact amplitudes when every qubit is sampled.  The
 * same distribution; a seed gives the reference's outcomes wherever its
 * FP64 sums are exact (e.g. dyadic probabilities), and may differ where
 * they round (a tie p = 1/2 the reference computes as 1/2 +- 1 ulp takes
 * the other branch of Binomial).  No reference counterpart; no term
 * limit. */
int nx_state_sample_exact(const nx_state *s, int32_t nq, const int32_t *qargs, uint64_t shots,
                          const uint8_t *seed32, int64_t max_out, uint64_t *out_bits, uint64_t *out_counts,
                          int64_t *nout);

/* QuantumCircuit::random_clifford (python/src/circuit.rs:73-79,
 * crates/necstar/src/circuit/random_clifford.rs:20-27,
 * crates/stabilizer-ch-form-rust/src/circuit/random_clifford.rs:7-218):
 * the gate list (op, q0, q1, 0) of a uniformly random Clifford on n <= 1024
 * qubits, drawn from the seed as the reference draws it.  At most
 * 2 n^2 + 3 n gates; capacity max_gates, *ngates receives the count. */
int nx_random_clifford(int32_t n, const uint8_t *seed32, int64_t max_gates, int32_t *gates, int64_t *ngates);

/* Parity export of term i in the reference's layout: G, F, M as n*n bytes
 * (row-major), gamma n bytes (k mod 8), v, s n bytes, phase 1 byte
 * (form/mod.rs:11-22).  No reference counterpart (tables are private). */
int nx_state_export_term(const nx_state *s, int64_t term, uint8_t *G, uint8_t *F, uint8_t *M, uint8_t *gamma,
                         uint8_t *v, uint8_t *sv, uint8_t *phase);

/* Coefficients: exact Scalar (zero, k, r) arrays, or complex128 values.
 * set_coefficients switches the state to Complex64 coefficients, the
 * generic StabilizerDecomposedState<Complex64> path (coefficient.rs:6-42). */
int nx_state_get_scalar_coefficients(const nx_state *s, uint8_t *zero, uint8_t *k, int64_t *r);
int nx_state_get_coefficients(const nx_state *s, double *re_im);
int nx_state_set_coefficients(nx_state *s, const double *re_im);
int nx_state_global_factor(const nx_state *s, double *re, double *im);

/* norm_squared() of the state projected (projection.rs:10-31, unnormalised)
 * onto outcome fixed_b[i] of qubit fixed_q[i] for every i, and onto each of
 * the 2^nlook outcomes of qubits look_q: out[k] for outcome (k >> i) & 1 on
 * look_q[i].  nlook <= 7.  One device pass for all 2^nlook outcomes (the
 * projections enter the pair sums as constraints); sample() is built on it.
 * No reference counterpart (the reference clones, projects and evaluates
 * norm() once per outcome). */
int nx_state_projection_norms(const nx_state *s, int32_t nfixed, const int32_t *fixed_q, const int32_t *fixed_b,
                              int32_t nlook, const int32_t *look_q, double *out);

/* The largest D such that the Z-basis marginal of qubits qargs[0..D) is
 * provably uniform (every outcome has probability 2^-D): no pair value
 * <phi_i| Z^t |phi_j> with t != 0 on those qubits is nonzero, decided for
 * all 2^D - 1 strings at once by one symbolic pass over the pairs (D <= 128).
 * sample() draws p = 1/2 without norms for that many levels.  No reference
 * counterpart (sampling.rs:80-189 evaluates every node's norms). */
int nx_state_uniform_prefix(const nx_state *s, int32_t nq, const int32_t *qargs, int32_t *out);

/* The same for the marginal conditioned on outcome fixed_b[i] of qubit
 * fixed_q[i] (the pairs of the projected terms).  No reference counterpart. */
int nx_state_uniform_prefix_given(const nx_state *s, int32_t nfixed, const int32_t *fixed_q, const int32_t *fixed_b,
                                  int32_t nq, const int32_t *qargs, int32_t *out);

/* |<x|psi>|^2 / |g|^2 = A[y] + sqrt(2) B[y] for the 2^d completions x =
 * base | y of `base` (W words, zero on the qs qubits) on qubits qs[0..d)
 * (bit i of y -> qubit qs[i]), d <= 16, n <= 128, Scalar coefficients:
 * the exact leaf norms of a sampling tail (sampling with exact=True).
 * Parity aid; the reference evaluates amplitudes one at a time
 * (amplitude.rs:28-44, statevector.rs:12-24). */
int nx_state_completion_norms(const nx_state *s, const uint64_t *base, int32_t d, const int32_t *qs, double *A,
                              double *B);

/* Exact pair scalars <phi_i|phi_j> for all i, j < chi (chi*chi uint32:
 * 0xFFFFFFFF zero, else k | r << 8).  Parity aid; the reference computes
 * the same values in StabilizerCHForm::inner_product (inner_product.rs:14-48). */
int nx_state_pair_table(const nx_state *s, uint32_t *out);
/* The same for the terms of two states: out[i * chi_b + j] = <a_i|b_j>, the
 * pairs of QuantumState::inner_product (stabilizer_decomposed_state/
 * inner_product.rs:9-19).  Parity aid. */
int nx_state_cross_pair_table(const nx_state *a, const nx_state *b, uint32_t *out);

/* exp_value for npauli dense Pauli strings ops[npauli][n] (little-endian,
 * 0 I, 1 X, 2 Y, 3 Z): out[k] as exp_value.rs:7-61 computes it. */
int nx_state_exp_values(const nx_state *s, int32_t npauli, const int32_t *ops, double *out);

/* Coefficient backward of L = sum_k w_k exp_value(P_k) with respect to the
 * decomposition coefficients: grad = dL/d(conj c) * 2 = 2 |g|^2 sum_k w_k G_k c,
 * (G_k)_{ji} = <phi_j|P_k phi_i>.  paulis: npauli dense strings ops[npauli][n].
 * grad: complex128 [chi] (the PyTorch convention for a real loss of complex
 * leaves); *loss receives L.  One pass over the pairs per 128 strings
 * computes L and the gradient together (O(chi^2 / 16) device memory, no
 * chi x chi matrix).  No reference counterpart (an addition for
 * differentiation; the forward is exp_value.rs:7-61). */
int nx_state_grad_exp_values(const nx_state *s, int32_t npauli, const int32_t *ops, const double *weights,
                             double *loss, double *grad);
/* The same with device memory: coefficient values d_coef (complex128
 * [chi]; NULL: the state's own), L into d_loss[0], the gradient into
 * d_grad (complex128 [chi]); weights stay host memory.  Ordered against
 * `stream` like nx_state_amplitudes_async, no host wait.  No reference
 * counterpart (used by the torch.autograd.Function). */
int nx_state_grad_exp_values_device(const nx_state *s, int32_t npauli, const int32_t *ops, const double *weights,
                                    const double *d_coef, double *d_loss, double *d_grad, void *stream);

/* Batches (many circuits of uneven n and chi in one launch plan).  State b
 * (copied at creation; later changes to it are not seen) gets its norm
 * (norm.rs:32-35) and the exp_value (exp_value.rs:7-61) of npauli[b] dense
 * Pauli strings; ops concatenates the strings of all states (npauli[b]
 * strings of length n_b each).  nx_batch_run writes, for state b in order,
 * norm_b then exp_value_b,1 .. exp_value_b,npauli[b].  All pairs of all
 * states and all strings of a state run in one persistent pair launch per
 * row-word width.  flags: NX_BATCH_AUTOTUNE times a grid of launch shapes
 * (pair-chunk length, wave count) once and keeps the fastest (cached per
 * shape class); NX_BATCH_GRAPH captures the launches into a HIP graph that
 * every nx_batch_run replays.  The reference has no batch API: this is the
 * batched form of calling QuantumState::norm and exp_value per circuit. */
enum { NX_BATCH_AUTOTUNE = 1, NX_BATCH_GRAPH = 2 };
int nx_batch_create(int32_t nstates, const nx_state *const *states, const int32_t *npauli, const int32_t *ops,
                    int32_t flags, nx_batch **out);
int nx_batch_num_outputs(const nx_batch *b, int64_t *out);
/* pair-chunk length, total waves and variant words (1: up to 64 strings per
 * job, 2: up to 128) of the batch's launch plan */
int nx_batch_shape(const nx_batch *b, int32_t *J, int64_t *waves, int32_t *variant_words);
int nx_batch_run(nx_batch *b, double *out);
void nx_batch_free(nx_batch *b);

/* Sampling many states at once: for state b, the outcomes of
 * nx_state_sample_exact(states[b], nq[b], qargs_b, shots[b], seed_b) -- the
 * same draws, state by state -- where qargs concatenates the nq[b] qubit
 * lists and seeds the 32-byte seeds (NULL: fresh entropy per state).  Each
 * state's DFS keeps its own RNG order; the windows that all states wait for
 * at one time are evaluated together (one projection launch and one pair
 * launch plan per round).  nout[b] receives state b's number of distinct
 * outcomes; its outcomes follow those of states 0..b-1 in out_bits (two
 * words each, as nx_state_sample) and out_counts.  When the sum exceeds
 * max_out, nout is still filled and the call fails (retry with the sum).
 * stats (6 int64, or NULL): rounds, window passes, exact tails, then the
 * microseconds spent evaluating rounds, waiting for the DFS threads, and
 * (of the first) waiting for the rounds' launches to finish.  The
 * reference has no batch API: this is the batched form of calling
 * QuantumState::sample (sampling.rs:80-189) per circuit. */
int nx_batch_sample_exact(int32_t nstates, const nx_state *const *states, const int32_t *nq, const int32_t *qargs,
                          const uint64_t *shots, const uint8_t *seeds, int64_t max_out, uint64_t *out_bits,
                          uint64_t *out_counts, int64_t *nout, int64_t *stats);

/* Kernel timing with HIP events recorded around each launch on the stream
 * the library enqueues on.  kernel: "k_amplitude", "k_amp_tables", "k_amp_reduce",
 * "k_tail", "k_pairs", "k_reduce", "k_pairs_var", "k_pairs_sym",
 * "k_grad_var", "k_qform", "k_qproject_var", "k_fold", "k_program",
 * "batch_pairs" (a whole nx_batch_run launch plan).  Enabling clears the totals.  No
 * reference counterpart (measurement aid). */
int nx_profile_enable(int on);
/* Restricts the timing to the comma-separated kernel names (NULL or "":
 * every kernel): each timed launch adds two event records to the stream. */
int nx_profile_only(const char *kernels);
/* kernel "allocations": *count = device-buffer requests to the library's
 * allocator plus page-locked host allocations so far (*total_ms = 0),
 * whether or not profiling is enabled. */
int nx_profile_get(const char *kernel, double *total_ms, int64_t *count);

/* Drops the derived per-term device data (the quadratic forms) so the
 * next call recomputes it; the device copies of the coefficients stay
 * resident until the coefficients change.  No reference counterpart. */
int nx_state_drop_cache(nx_state *s);

#ifdef __cplusplus
}
#endif
#endif
