/*
 * necstar_oracle.h -- CPU oracle for the stabilizer-decomposition hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference algorithms of mutekichi/stab-decomp-simulator-rust ("NECSTAR"),
 * kept in the reference's own data layout (one byte per matrix entry, like
 * ndarray::Array2<bool>) and op order.  It is used by tests/ as the parity
 * checker, by __graft_entry__.smoke() as the checker and by bench.py's
 * cpu_baseline leg as the timed CPU baseline.  The product library
 * (libnecstar_hip.so) never links, loads or calls it.
 *
 * Every function cites the reference file it restates (paths relative to
 * the reference repository root).
 */
#ifndef NECSTAR_ORACLE_H
#define NECSTAR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gate op codes shared with the product C-ABI (include/necstar_hip.h). */
enum {
  ORC_G_H = 0, ORC_G_X = 1, ORC_G_Y = 2, ORC_G_Z = 3, ORC_G_S = 4,
  ORC_G_SDG = 5, ORC_G_SQRTX = 6, ORC_G_SQRTXDG = 7, ORC_G_CX = 8,
  ORC_G_CZ = 9, ORC_G_SWAP = 10, ORC_G_T = 11, ORC_G_TDG = 12,
  ORC_G_CCX = 13
};

/* Error codes (one per reference error variant that the path can raise). */
enum {
  ORC_OK = 0,
  ORC_E_QUBIT_OOB = 1,          /* QubitIndexOutOfBounds(q, n)            */
  ORC_E_INVALID_NUM_QUBITS = 2, /* InvalidNumQubits(n)                    */
  ORC_E_DUPLICATE_QUBITS = 3,   /* ch-form DuplicateQubitIndices(q)       */
  ORC_E_IMPOSSIBLE_PROJ = 4,    /* ImpossibleProjection{q, desired}       */
  ORC_E_QUBIT_COUNT = 5,        /* QubitCountMismatch                     */
  ORC_E_CANNOT_DISCARD = 6,     /* CannotDiscardQubit(q)                  */
  ORC_E_SV_TOO_LARGE = 7,       /* StatevectorTooLarge(n)                 */
  ORC_E_NON_CLIFFORD = 8,       /* NonCliffordGateApplication(name)       */
  ORC_E_SAMPLING_TOO_MANY = 9,  /* SamplingTooManyQubits                  */
  ORC_E_DUPLICATE_QARG = 10,    /* necstar DuplicateQubitIndex(q)         */
  ORC_E_PAULI_LENGTH = 11,      /* InvalidPauliStringLength               */
  ORC_E_EMPTY_QARGS = 12,       /* EmptyQubitIndices                      */
  ORC_E_GATE_NOT_SUPPORTED = 13,/* compiler GateNotSupported(name)        */
  ORC_E_NOT_IMPLEMENTED = 14,   /* NotImplemented(msg)                    */
  ORC_E_INTERNAL = 15           /* reference panics (unreachable!/panic!) */
};

typedef struct orc_chform orc_chform;
typedef struct orc_state orc_state;

const char *orc_last_error(void);

/* ---- single CH-form (stabilizer-ch-form-rust/src/form) ---- */
orc_chform *orc_chf_new(int n, int *err);
void orc_chf_free(orc_chform *f);
orc_chform *orc_chf_clone(const orc_chform *f);
int orc_chf_num_qubits(const orc_chform *f);
int orc_chf_apply_gate(orc_chform *f, int op, int q0, int q1);
int orc_chf_project(orc_chform *f, int q, int outcome, int *deterministic);
int orc_chf_discard(orc_chform *f, int q);
orc_chform *orc_chf_kron(const orc_chform *a, const orc_chform *b, int *err);
/* <0..0|phi> without omega: writes zero flag, phase k (mod 8), r */
int orc_chf_amplitude_at_zero(const orc_chform *f, int *zero, int *k, int *r);
int orc_chf_amplitude(const orc_chform *f, const uint8_t *bits, int *zero,
                      int *k, int *r);
/* <a|b>: exact scalar (zero,k,r) before omega, and the Complex64 result */
int orc_chf_inner_product(const orc_chform *a, const orc_chform *b, int *zero,
                          int *k, int *r, double *re, double *im);
/* table export: G,F,M n*n bytes row-major, gamma n bytes (k mod 8),
 * v,s n bytes, phase 1 byte */
void orc_chf_export(const orc_chform *f, uint8_t *G, uint8_t *F, uint8_t *M,
                    uint8_t *gamma, uint8_t *v, uint8_t *s, uint8_t *phase);
int orc_chf_statevector(const orc_chform *f, double *out /* 2*2^n */);

/* ---- chi-term state (crates/necstar/src/state) ---- */
/* gates: ngates x 4 ints (op, q0, q1, q2) */
orc_state *orc_compile(int n, int ngates, const int *gates, int *err);
void orc_state_free(orc_state *s);
orc_state *orc_state_clone(const orc_state *s);
int orc_state_num_qubits(const orc_state *s);
int orc_state_rank(const orc_state *s);
int orc_state_is_complex(const orc_state *s);
const orc_chform *orc_state_term(const orc_state *s, int i);
/* Scalar coefficient of term i (zero flag, phase k mod 8, r) */
void orc_state_coeff_scalar(const orc_state *s, int i, int *zero, int *k,
                            int64_t *r);
/* Complex value of coefficient i (Scalar::to_complex for Scalar states) */
void orc_state_coeff_complex(const orc_state *s, int i, double *re,
                             double *im);
/* switch the state to Complex64 coefficients (the generic
 * StabilizerDecomposedState<Complex64> path) with the given values */
void orc_state_set_complex_coeffs(orc_state *s, const double *re_im);
void orc_state_global_factor(const orc_state *s, double *re, double *im);

int orc_state_apply_gate(orc_state *s, int op, int q0, int q1, int q2);
int orc_state_norm_squared(const orc_state *s, double *out);
int orc_state_norm(const orc_state *s, double *out);
/* Pauli: dense=1 -> ops[len] little-endian (0=I,1=X,2=Y,3=Z);
 *        dense=0 -> sparse terms (ops[i], qubits[i]) */
int orc_state_exp_value(const orc_state *s, int dense, int len, const int *ops,
                        const int *qubits, double *out);
int orc_state_inner_product(const orc_state *a, const orc_state *b,
                            double *re, double *im);
int orc_state_project_unnormalized(orc_state *s, int q, int outcome);
int orc_state_project_normalized(orc_state *s, int q, int outcome);
int orc_state_discard(orc_state *s, int q);
int orc_state_statevector(const orc_state *s, double *out);
/* amplitude <x|psi> summed over terms (sequential term order, the same
 * operations as to_statevector), per-term exact scalars optional */
int orc_state_amplitude(const orc_state *s, const uint8_t *bits, double *re,
                        double *im, int *zero_t, int *k_t, int *r_t);
/* seed: 32 bytes */
int orc_state_measure(orc_state *s, int nq, const int *qargs,
                      const uint8_t *seed, uint8_t *outcomes);
/* sample: outcomes as (bits little-endian in 2 u64 words, count).
 * out_bits capacity: max_out*2 u64, out_counts capacity max_out.
 * returns number of distinct outcomes via *nout. */
int orc_state_sample(const orc_state *s, int nq, const int *qargs,
                     uint64_t shots, const uint8_t *seed, int max_out,
                     uint64_t *out_bits, uint64_t *out_counts, int *nout);

/* ---- RNG restatement (rand 0.8 StdRng = ChaCha12, rand_distr Binomial) ---- */
typedef struct orc_rng orc_rng;
orc_rng *orc_rng_new(const uint8_t *seed32);
void orc_rng_free(orc_rng *r);
uint32_t orc_rng_next_u32(orc_rng *r);
uint64_t orc_rng_next_u64(orc_rng *r);
double orc_rng_gen_f64(orc_rng *r);
uint64_t orc_rng_binomial(orc_rng *r, uint64_t n, double p);
/* raw ChaCha block (rounds = 8/12/20) for known-answer tests */
void orc_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream,
                      int rounds, uint32_t out[16]);

#ifdef __cplusplus
}
#endif
#endif
