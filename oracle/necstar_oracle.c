/*
 * necstar_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of the reference hot path of
 * mutekichi/stab-decomp-simulator-rust.  Data layout follows the reference
 * (one byte per F/G/M entry, like ndarray::Array2<bool>; gamma and the phase
 * factor as k mod 8), and so does the order of every operation, including
 * the per-call clones and the normalize-op list that the reference recomputes
 * for every pair.  That makes it both the parity checker for the HIP product
 * and a fair single-core timing proxy for the reference.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this code.  Built with -ffp-contract=off so FP64 matches Rust (no FMA).
 */
#include "necstar_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* errors (stabilizer-ch-form-rust/src/error/mod.rs, necstar/src/error/mod.rs) */
static __thread char g_err[512];

static int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}
const char *orc_last_error(void) { return g_err; }

static int err_oob(int q, int n) {
  return set_err(ORC_E_QUBIT_OOB, "Qubit index %d is out of bounds for %d qubits.", q, n);
}
static int err_dup(int q) {
  return set_err(ORC_E_DUPLICATE_QUBITS,
                 "Two-qubit gate requires two different qubit indices, but got the same: %d.", q);
}
/* ch-form variant has no space before the second ket
 * (stabilizer-ch-form-rust/src/error/mod.rs ImpossibleProjection) */
static int err_impossible_chf(int q, int desired) {
  return set_err(ORC_E_IMPOSSIBLE_PROJ,
                 "Impossible projection on qubit %d: cannot project determined state |%d> onto|%d>.",
                 q, desired ? 0 : 1, desired ? 1 : 0);
}
static int err_impossible_state(int q, int desired) {
  return set_err(ORC_E_IMPOSSIBLE_PROJ,
                 "Impossible projection on qubit %d: cannot project determined state |%d> onto |%d>.",
                 q, desired ? 0 : 1, desired ? 1 : 0);
}
static int err_internal(const char *what) { return set_err(ORC_E_INTERNAL, "%s", what); }

/* ------------------------------------------------------------------------ */
/* FP64 helpers: num-complex 0.4 arithmetic, no FMA                          */
typedef struct { double re, im; } cplx;
static inline cplx c_mul(cplx a, cplx b) {
  cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static inline cplx c_add(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx c_conj(cplx a) { cplx r = {a.re, -a.im}; return r; }
static inline cplx c_scale(cplx a, double s) { cplx r = {a.re * s, a.im * s}; return r; }

/* PhaseFactor::to_complex (form/types/phase_factor.rs:19-22) */
static cplx phase_to_complex(int k) {
  double angle = (double)k * M_PI_4;
  cplx r = {cos(angle), sin(angle)};
  return r;
}
/* Scalar::to_complex (form/types/scalar.rs:17-27, state/types/scalar.rs:28-38) */
static cplx scalar_to_complex(int zero, int k, int64_t r) {
  if (zero) { cplx z = {0.0, 0.0}; return z; }
  double norm = pow(2.0, -((double)r) / 2.0);
  return c_scale(phase_to_complex(k), norm);
}

/* ------------------------------------------------------------------------ */
/* CH-form (stabilizer-ch-form-rust/src/form/mod.rs:11-22)                    */
struct orc_chform {
  int n;
  uint8_t *G, *F, *M; /* n*n, row-major */
  uint8_t *gamma;     /* k mod 8 per row */
  uint8_t *v, *s;
  cplx omega;
  uint8_t phase;      /* phase_factor k mod 8 */
};

#define IDX(f, i, j) ((size_t)(i) * (size_t)(f)->n + (size_t)(j))

static orc_chform *chf_alloc(int n) {
  orc_chform *f = (orc_chform *)calloc(1, sizeof(orc_chform));
  size_t nn = (size_t)n * (size_t)n;
  f->n = n;
  f->G = (uint8_t *)calloc(nn ? nn : 1, 1);
  f->F = (uint8_t *)calloc(nn ? nn : 1, 1);
  f->M = (uint8_t *)calloc(nn ? nn : 1, 1);
  f->gamma = (uint8_t *)calloc(n ? n : 1, 1);
  f->v = (uint8_t *)calloc(n ? n : 1, 1);
  f->s = (uint8_t *)calloc(n ? n : 1, 1);
  f->omega.re = 1.0;
  f->omega.im = 0.0;
  f->phase = 0;
  return f;
}

/* StabilizerCHForm::new (form/mod.rs:44-65) */
orc_chform *orc_chf_new(int n, int *err) {
  if (n <= 0) {
    *err = set_err(ORC_E_INVALID_NUM_QUBITS,
                   "Number of qubits must be greater than zero, got %d.", n);
    return NULL;
  }
  orc_chform *f = chf_alloc(n);
  for (int i = 0; i < n; i++) {
    f->G[IDX(f, i, i)] = 1;
    f->F[IDX(f, i, i)] = 1;
  }
  *err = 0;
  return f;
}

void orc_chf_free(orc_chform *f) {
  if (!f) return;
  free(f->G); free(f->F); free(f->M); free(f->gamma); free(f->v); free(f->s);
  free(f);
}

orc_chform *orc_chf_clone(const orc_chform *f) {
  orc_chform *c = chf_alloc(f->n);
  size_t nn = (size_t)f->n * f->n;
  memcpy(c->G, f->G, nn); memcpy(c->F, f->F, nn); memcpy(c->M, f->M, nn);
  memcpy(c->gamma, f->gamma, f->n); memcpy(c->v, f->v, f->n); memcpy(c->s, f->s, f->n);
  c->omega = f->omega;
  c->phase = f->phase;
  return c;
}

int orc_chf_num_qubits(const orc_chform *f) { return f->n; }

/* matrix_operations.rs:7-29: xor_rows(m, target, source) m[t,:] ^= m[s,:];
 * xor_columns(m, target, source) m[:,t] ^= m[:,s] */
static void xor_rows(orc_chform *f, uint8_t *m, int t, int s) {
  for (int j = 0; j < f->n; j++) m[IDX(f, t, j)] ^= m[IDX(f, s, j)];
}
static void xor_cols(orc_chform *f, uint8_t *m, int t, int s) {
  for (int i = 0; i < f->n; i++) m[IDX(f, i, t)] ^= m[IDX(f, i, s)];
}

static int check1(const orc_chform *f, int q) {
  if (q < 0 || q >= f->n) return err_oob(q, f->n);
  return 0;
}
static int check2(const orc_chform *f, int a, int b) {
  if (a < 0 || a >= f->n) return err_oob(a, f->n);
  if (b < 0 || b >= f->n) return err_oob(b, f->n);
  if (a == b) return err_dup(a);
  return 0;
}

/* left_multiplication/cx.rs:13-51 */
static int left_cx(orc_chform *f, int c, int t) {
  int e = check2(f, c, t);
  if (e) return e;
  int dot = 0;
  for (int j = 0; j < f->n; j++) dot ^= f->M[IDX(f, c, j)] & f->F[IDX(f, t, j)];
  int gc = f->gamma[c], gt = f->gamma[t];
  int g = (gc + gt) % 8;
  if (dot) g = (g + 4) % 8;
  f->gamma[c] = (uint8_t)g;
  xor_rows(f, f->G, t, c);
  xor_rows(f, f->F, c, t);
  xor_rows(f, f->M, c, t);
  return 0;
}

/* left_multiplication/cz.rs:12-35 */
static int left_cz(orc_chform *f, int q1, int q2) {
  int e = check2(f, q1, q2);
  if (e) return e;
  for (int j = 0; j < f->n; j++) f->M[IDX(f, q1, j)] ^= f->G[IDX(f, q2, j)];
  for (int j = 0; j < f->n; j++) f->M[IDX(f, q2, j)] ^= f->G[IDX(f, q1, j)];
  return 0;
}

/* left_multiplication/s.rs:13-33 */
static int left_s(orc_chform *f, int q) {
  int e = check1(f, q);
  if (e) return e;
  for (int j = 0; j < f->n; j++) f->M[IDX(f, q, j)] ^= f->G[IDX(f, q, j)];
  f->gamma[q] = (uint8_t)((f->gamma[q] + 6) % 8); /* MINUS_I */
  return 0;
}
static int left_sdg(orc_chform *f, int q) {
  int e = check1(f, q);
  if (e) return e;
  for (int j = 0; j < f->n; j++) f->M[IDX(f, q, j)] ^= f->G[IDX(f, q, j)];
  f->gamma[q] = (uint8_t)((f->gamma[q] + 2) % 8); /* PLUS_I */
  return 0;
}

/* left_multiplication/z.rs:7-14 */
static int left_z(orc_chform *f, int q) {
  int e = check1(f, q);
  if (e) return e;
  f->gamma[q] = (uint8_t)((f->gamma[q] + 4) % 8);
  return 0;
}

/* left_multiplication/x.rs:12-48 */
static int left_x(orc_chform *f, int q) {
  int e = check1(f, q);
  if (e) return e;
  int beta = 0;
  for (int j = 0; j < f->n; j++) {
    int m = f->M[IDX(f, q, j)], fr = f->F[IDX(f, q, j)], v = f->v[j], s = f->s[j];
    int t1 = m & !v & s;
    int t23 = fr & v & (m ^ s);
    beta ^= t1 ^ t23;
  }
  if (beta) f->phase = (uint8_t)((f->phase + 4) % 8);
  for (int j = 0; j < f->n; j++) {
    int fr = f->F[IDX(f, q, j)], m = f->M[IDX(f, q, j)], v = f->v[j];
    f->s[j] ^= (uint8_t)((fr & !v) ^ (m & v));
  }
  f->phase = (uint8_t)((f->phase + f->gamma[q]) % 8);
  return 0;
}

/* left_multiplication/y.rs:4-11 */
static int left_y(orc_chform *f, int q) {
  int e = left_z(f, q);
  if (e) return e;
  e = left_x(f, q);
  if (e) return e;
  f->phase = (uint8_t)((f->phase + 2) % 8);
  return 0;
}

/* right_multiplication/cx.rs:7-26 */
static int right_cx(orc_chform *f, int c, int t) {
  int e = check2(f, c, t);
  if (e) return e;
  xor_cols(f, f->G, c, t);
  xor_cols(f, f->F, t, c);
  xor_cols(f, f->M, c, t);
  return 0;
}

/* right_multiplication/cz.rs:8-35 */
static int right_cz(orc_chform *f, int q1, int q2) {
  int e = check2(f, q1, q2);
  if (e) return e;
  for (int p = 0; p < f->n; p++) f->M[IDX(f, p, q1)] ^= f->F[IDX(f, p, q2)];
  for (int p = 0; p < f->n; p++) f->M[IDX(f, p, q2)] ^= f->F[IDX(f, p, q1)];
  for (int p = 0; p < f->n; p++)
    if (f->F[IDX(f, p, q1)] && f->F[IDX(f, p, q2)]) f->gamma[p] = (uint8_t)((f->gamma[p] + 4) % 8);
  return 0;
}

/* right_multiplication/s.rs:8-24 */
static int right_s(orc_chform *f, int q) {
  int e = check1(f, q);
  if (e) return e;
  for (int p = 0; p < f->n; p++) f->M[IDX(f, p, q)] ^= f->F[IDX(f, p, q)];
  for (int p = 0; p < f->n; p++)
    if (f->F[IDX(f, p, q)]) f->gamma[p] = (uint8_t)((f->gamma[p] + 6) % 8);
  return 0;
}

/* resolve_superposition.rs:6-176.  t and u are copies owned by the caller. */
static int resolve_superposition(orc_chform *f, const uint8_t *t, const uint8_t *u, int delta) {
  int n = f->n, same = 1;
  for (int i = 0; i < n; i++)
    if (t[i] != u[i]) { same = 0; break; }
  if (same) { /* handle_same_vecs_case (:115-133) */
    switch (delta) {
      case 0: memcpy(f->s, t, n); return 0;
      case 4: return err_internal("Inconsistent state: superposition with -1 coefficient.");
      case 2: memcpy(f->s, t, n); f->phase = (uint8_t)((f->phase + 1) % 8); return 0;
      case 6: memcpy(f->s, t, n); f->phase = (uint8_t)((f->phase + 7) % 8); return 0;
      default: return err_internal("unreachable delta in resolve_superposition");
    }
  }
  /* get_differing_indices (:135-153) */
  int *d0 = (int *)malloc(sizeof(int) * n), *d1 = (int *)malloc(sizeof(int) * n);
  int n0 = 0, n1 = 0;
  for (int i = 0; i < n; i++)
    if (t[i] != u[i]) {
      if (f->v[i]) d1[n1++] = i; else d0[n0++] = i;
    }
  /* apply_basis_transform_circuit (:155-176) */
  int pivot, e = 0;
  if (n0 == 0) {
    pivot = d1[0];
    for (int k = 1; k < n1 && !e; k++) e = right_cx(f, d1[k], pivot);
  } else {
    pivot = d0[0];
    for (int k = 1; k < n0 && !e; k++) e = right_cx(f, pivot, d0[k]);
    for (int k = 0; k < n1 && !e; k++) e = right_cz(f, pivot, d1[k]);
  }
  free(d0); free(d1);
  if (e) return e;
  if (t[pivot]) {
    memcpy(f->s, u, n);
    switch (delta) {
      case 0: f->s[pivot] = 0; f->v[pivot] ^= 1; break;
      case 4: f->s[pivot] = 1; f->v[pivot] ^= 1; f->phase = (uint8_t)((f->phase + 4) % 8); break;
      case 2:
        if (f->v[pivot]) {
          f->phase = (uint8_t)((f->phase + 1) % 8); f->s[pivot] = 0;
          if ((e = right_s(f, pivot))) return e;
        } else {
          f->s[pivot] = 1; f->v[pivot] = 1;
          if ((e = right_s(f, pivot))) return e;
          f->phase = (uint8_t)((f->phase + 2) % 8);
        }
        break;
      case 6:
        if (f->v[pivot]) {
          f->phase = (uint8_t)((f->phase + 7) % 8); f->s[pivot] = 1;
          if ((e = right_s(f, pivot))) return e;
        } else {
          f->s[pivot] = 0; f->v[pivot] = 1;
          if ((e = right_s(f, pivot))) return e;
          f->phase = (uint8_t)((f->phase + 6) % 8);
        }
        break;
      default: return err_internal("unreachable delta in resolve_superposition");
    }
  } else {
    memcpy(f->s, t, n);
    switch (delta) {
      case 0: f->s[pivot] = 0; f->v[pivot] ^= 1; break;
      case 4: f->s[pivot] = 1; f->v[pivot] ^= 1; break;
      case 2:
        if (f->v[pivot]) {
          f->phase = (uint8_t)((f->phase + 1) % 8); f->s[pivot] = 1;
          if ((e = right_s(f, pivot))) return e;
        } else {
          f->s[pivot] = 0; f->v[pivot] = 1;
          if ((e = right_s(f, pivot))) return e;
        }
        break;
      case 6:
        if (f->v[pivot]) {
          f->phase = (uint8_t)((f->phase + 7) % 8); f->s[pivot] = 0;
          if ((e = right_s(f, pivot))) return e;
        } else {
          f->s[pivot] = 1; f->v[pivot] = 1;
          if ((e = right_s(f, pivot))) return e;
        }
        break;
      default: return err_internal("unreachable delta in resolve_superposition");
    }
  }
  return 0;
}

/* left_multiplication/h.rs:13-65 */
static int left_h(orc_chform *f, int q) {
  int e = check1(f, q);
  if (e) return e;
  int n = f->n;
  uint8_t *t = (uint8_t *)malloc(n), *u = (uint8_t *)malloc(n);
  int alpha = 0, beta = 0;
  for (int j = 0; j < n; j++) {
    int g = f->G[IDX(f, q, j)], fr = f->F[IDX(f, q, j)], m = f->M[IDX(f, q, j)];
    int v = f->v[j], s = f->s[j];
    t[j] = (uint8_t)(s ^ (g & v));
    u[j] = (uint8_t)(s ^ (fr & !v) ^ (m & v));
    alpha ^= g & !v & s;
    beta ^= (m & !v & s) ^ (fr & v & (m ^ s));
  }
  int delta = (alpha ^ beta) ? (f->gamma[q] + 4) % 8 : f->gamma[q];
  if (alpha) f->phase = (uint8_t)((f->phase + 4) % 8);
  e = resolve_superposition(f, t, u, delta);
  free(t); free(u);
  return e;
}

/* left_multiplication/swap.rs:7-24 */
static int left_swap(orc_chform *f, int a, int b) {
  int e = check2(f, a, b);
  if (e) return e;
  if ((e = left_cx(f, a, b))) return e;
  if ((e = left_cx(f, b, a))) return e;
  return left_cx(f, a, b);
}

/* left_multiplication/sqrt_x.rs:6-16 */
static int left_sqrt_x(orc_chform *f, int q, int dg) {
  int e = left_h(f, q);
  if (e) return e;
  e = dg ? left_sdg(f, q) : left_s(f, q);
  if (e) return e;
  return left_h(f, q);
}

/* gate_application.rs:162-177 (apply_gate); op codes from the header */
int orc_chf_apply_gate(orc_chform *f, int op, int q0, int q1) {
  switch (op) {
    case ORC_G_H: return left_h(f, q0);
    case ORC_G_X: return left_x(f, q0);
    case ORC_G_Y: return left_y(f, q0);
    case ORC_G_Z: return left_z(f, q0);
    case ORC_G_S: return left_s(f, q0);
    case ORC_G_SDG: return left_sdg(f, q0);
    case ORC_G_SQRTX: return left_sqrt_x(f, q0, 0);
    case ORC_G_SQRTXDG: return left_sqrt_x(f, q0, 1);
    case ORC_G_CX: return left_cx(f, q0, q1);
    case ORC_G_CZ: return left_cz(f, q0, q1);
    case ORC_G_SWAP: return left_swap(f, q0, q1);
    default: return err_internal("non-Clifford op passed to a CH-form");
  }
}

/* amplitude.rs:10-22 */
int orc_chf_amplitude_at_zero(const orc_chform *f, int *zero, int *k, int *r) {
  for (int j = 0; j < f->n; j++)
    if (!f->v[j] && f->s[j]) { *zero = 1; *k = 0; *r = 0; return 0; }
  int w = 0;
  for (int j = 0; j < f->n; j++) w += f->v[j];
  *zero = 0; *k = f->phase; *r = w;
  return 0;
}

/* amplitude.rs:28-44 (clone, X on every set bit in ascending order) */
int orc_chf_amplitude(const orc_chform *f, const uint8_t *bits, int *zero, int *k, int *r) {
  orc_chform *c = orc_chf_clone(f);
  int e = 0;
  for (int i = 0; i < f->n && !e; i++)
    if (bits[i]) e = left_x(c, i);
  if (!e) e = orc_chf_amplitude_at_zero(c, zero, k, r);
  orc_chf_free(c);
  return e;
}

/* get_qubit_state.rs:6-31 ; returns 0 determined / 1 superposition */
static int qubit_state(const orc_chform *f, int q, int *value) {
  int sup = 0, par = 0;
  for (int j = 0; j < f->n; j++) {
    if (f->G[IDX(f, q, j)] && f->v[j]) sup = 1;
    if (f->G[IDX(f, q, j)] && f->s[j]) par ^= 1;
  }
  *value = par;
  return sup;
}

/* project.rs:33-76 */
int orc_chf_project(orc_chform *f, int q, int outcome, int *deterministic) {
  int e = check1(f, q);
  if (e) return e;
  int value;
  int sup = qubit_state(f, q, &value);
  if (!sup) {
    if (value != (outcome ? 1 : 0)) return err_impossible_chf(q, outcome);
    *deterministic = 1;
    return 0;
  }
  int n = f->n, alpha = 0;
  uint8_t *t = (uint8_t *)malloc(n), *s = (uint8_t *)malloc(n);
  for (int j = 0; j < n; j++) {
    int g = f->G[IDX(f, q, j)];
    t[j] = (uint8_t)((g & f->v[j]) ^ f->s[j]);
    alpha ^= g & !f->v[j] & f->s[j];
    s[j] = f->s[j];
  }
  int delta = (alpha ^ (outcome ? 1 : 0)) ? 4 : 0;
  e = resolve_superposition(f, s, t, delta);
  free(t); free(s);
  if (e) return e;
  *deterministic = 0;
  return 0;
}

/* discard.rs:98-132 */
static int set_s_v_to_false(orc_chform *f, int q) {
  if (!f->v[q] && !f->s[q]) return 0;
  int clean = -1, e;
  for (int i = 0; i < f->n; i++)
    if (i != q && !f->v[i] && !f->s[i]) { clean = i; break; }
  if (clean < 0) {
    int c1 = -1, c2 = -1;
    for (int i = 0; i < f->n; i++)
      if (!f->v[i] && f->s[i]) {
        if (c1 < 0) c1 = i;
        else { c2 = i; break; }
      }
    if (c1 >= 0 && c2 >= 0) {
      if ((e = right_cx(f, c1, c2))) return e;
      f->s[c2] = 0;
      clean = c2;
    }
  }
  if (clean < 0)
    return set_err(ORC_E_CANNOT_DISCARD, "Cannot discard qubit %d because it is not in a proper state", q);
  if (clean == q) return 0;
  if ((e = right_cx(f, clean, q))) return e;
  if ((e = right_cx(f, q, clean))) return e;
  if ((e = right_cx(f, clean, q))) return e;
  uint8_t tv = f->v[q]; f->v[q] = f->v[clean]; f->v[clean] = tv;
  uint8_t ts = f->s[q]; f->s[q] = f->s[clean]; f->s[clean] = ts;
  return 0;
}

/* discard.rs:135-161 */
static int transform_g(orc_chform *f, int q) {
  int e;
  if (!f->G[IDX(f, q, q)]) {
    int pivot = -1;
    for (int i = 0; i < f->n; i++)
      if (i != q && f->G[IDX(f, q, i)]) { pivot = i; break; }
    if (pivot < 0) return err_internal("Cannot zero G row due to G matrix state.");
    if ((e = right_cx(f, q, pivot))) return e;
  }
  for (int i = 0; i < f->n; i++)
    if (i != q && f->G[IDX(f, i, q)])
      if ((e = left_cx(f, q, i))) return e;
  for (int i = 0; i < f->n; i++)
    if (i != q && f->G[IDX(f, q, i)]) {
      if (f->v[i]) return err_internal("Cannot zero G column due to v vector state.");
      if ((e = right_cx(f, i, q))) return e;
    }
  return 0;
}

/* discard.rs:164-181 */
static int transform_m(orc_chform *f, int q) {
  int e;
  for (int i = 0; i < f->n; i++)
    if (i != q && f->M[IDX(f, i, q)])
      if ((e = left_cx(f, q, i))) return e;
  if (f->M[IDX(f, q, q)])
    if ((e = left_sdg(f, q))) return e;
  for (int i = 0; i < f->n; i++)
    if (i != q && f->M[IDX(f, q, i)])
      if ((e = right_cz(f, q, i))) return e;
  return 0;
}

/* discard.rs:64-96 */
static void remove_row_col(int n_old, uint8_t *m, int q) {
  int ni = 0;
  for (int i = 0; i < n_old; i++) {
    if (i == q) continue;
    int nj = 0;
    for (int j = 0; j < n_old; j++) {
      if (j == q) continue;
      m[(size_t)ni * (n_old - 1) + nj] = m[(size_t)i * n_old + j];
      nj++;
    }
    ni++;
  }
}
static void remove_elem(int n_old, uint8_t *v, int q) {
  for (int i = q; i < n_old - 1; i++) v[i] = v[i + 1];
}

/* discard.rs:21-44 */
int orc_chf_discard(orc_chform *f, int q) {
  int e = check1(f, q);
  if (e) return e;
  if ((e = set_s_v_to_false(f, q))) return e;
  if ((e = transform_g(f, q))) return e;
  if ((e = transform_m(f, q))) return e;
  int n_old = f->n;
  remove_row_col(n_old, f->G, q);
  remove_row_col(n_old, f->F, q);
  remove_row_col(n_old, f->M, q);
  remove_elem(n_old, f->gamma, q);
  remove_elem(n_old, f->v, q);
  remove_elem(n_old, f->s, q);
  f->n = n_old - 1;
  return 0;
}

/* kron.rs:12-59 */
orc_chform *orc_chf_kron(const orc_chform *a, const orc_chform *b, int *err) {
  int n = a->n + b->n;
  orc_chform *r = orc_chf_new(n, err);
  if (!r) return NULL;
  for (int i = 0; i < n; i++) r->G[IDX(r, i, i)] = r->F[IDX(r, i, i)] = 0;
  for (int i = 0; i < a->n; i++)
    for (int j = 0; j < a->n; j++) {
      r->G[IDX(r, i, j)] = a->G[IDX(a, i, j)];
      r->F[IDX(r, i, j)] = a->F[IDX(a, i, j)];
      r->M[IDX(r, i, j)] = a->M[IDX(a, i, j)];
    }
  for (int i = 0; i < b->n; i++)
    for (int j = 0; j < b->n; j++) {
      r->G[IDX(r, a->n + i, a->n + j)] = b->G[IDX(b, i, j)];
      r->F[IDX(r, a->n + i, a->n + j)] = b->F[IDX(b, i, j)];
      r->M[IDX(r, a->n + i, a->n + j)] = b->M[IDX(b, i, j)];
    }
  memcpy(r->gamma, a->gamma, a->n); memcpy(r->gamma + a->n, b->gamma, b->n);
  memcpy(r->v, a->v, a->n); memcpy(r->v + a->n, b->v, b->n);
  memcpy(r->s, a->s, a->n); memcpy(r->s + a->n, b->s, b->n);
  r->omega = c_mul(a->omega, b->omega);
  r->phase = (uint8_t)((a->phase + b->phase) % 8);
  *err = 0;
  return r;
}

/* inner_product.rs:52-123 op list; ops encoded (kind, a, b) */
enum { OP_H, OP_SDG, OP_X, OP_CX, OP_CZ };
typedef struct { int kind, a, b; } iop;

static int apply_iop(orc_chform *f, iop o) {
  switch (o.kind) {
    case OP_H: return left_h(f, o.a);
    case OP_SDG: return left_sdg(f, o.a);
    case OP_X: return left_x(f, o.a);
    case OP_CX: return left_cx(f, o.a, o.b);
    default: return left_cz(f, o.a, o.b);
  }
}

static int normalize_ops(const orc_chform *self, iop **ops_out, int *nops, int *phase) {
  orc_chform *c = orc_chf_clone(self);
  int n = c->n, cap = 64, cnt = 0, e = 0;
  iop *ops = (iop *)malloc(sizeof(iop) * cap);
#define PUSH(K, A, B)                                                   \
  do {                                                                  \
    if (cnt == cap) { cap *= 2; ops = (iop *)realloc(ops, sizeof(iop) * cap); } \
    ops[cnt].kind = (K); ops[cnt].a = (A); ops[cnt].b = (B);            \
    if ((e = apply_iop(c, ops[cnt]))) goto done;                        \
    cnt++;                                                              \
  } while (0)
  for (int j = 0; j < n; j++) {
    int pr = j;
    if (!c->G[IDX(c, j, j)]) {
      int k = -1;
      for (int kk = j + 1; kk < n; kk++)
        if (c->G[IDX(c, kk, j)]) { k = kk; break; }
      if (k < 0) { e = err_internal("G matrix is not full rank, invalid stabilizer state."); goto done; }
      pr = k;
    }
    if (pr != j) {
      PUSH(OP_CX, pr, j);
      PUSH(OP_CX, j, pr);
      PUSH(OP_CX, pr, j);
    }
    for (int i = 0; i < n; i++)
      if (i != j && c->G[IDX(c, i, j)]) PUSH(OP_CX, j, i);
  }
  for (int r = 0; r < n; r++)
    for (int col = r + 1; col < n; col++)
      if (c->M[IDX(c, r, col)]) PUSH(OP_CZ, r, col);
  for (int q = 0; q < n; q++)
    if (c->M[IDX(c, q, q)]) PUSH(OP_SDG, q, 0);
  for (int i = 0; i < n; i++)
    if (c->v[i]) PUSH(OP_H, i, 0);
  for (int i = 0; i < n; i++)
    if (c->s[i]) PUSH(OP_X, i, 0);
#undef PUSH
done:
  *phase = c->phase;
  orc_chf_free(c);
  if (e) { free(ops); return e; }
  *ops_out = ops;
  *nops = cnt;
  return 0;
}

/* inner_product.rs:14-48 */
int orc_chf_inner_product(const orc_chform *a, const orc_chform *b, int *zero, int *k, int *r,
                          double *re, double *im) {
  if (a->n != b->n)
    return set_err(ORC_E_QUBIT_COUNT,
                   "The qubit counts of the two states must match for calculating inner product, got %d and %d.",
                   a->n, b->n);
  iop *ops; int nops, phase, e;
  if ((e = normalize_ops(a, &ops, &nops, &phase))) return e;
  orc_chform *t = orc_chf_clone(b);
  for (int i = 0; i < nops && !e; i++) e = apply_iop(t, ops[i]);
  free(ops);
  if (e) { orc_chf_free(t); return e; }
  int z, kk, rr;
  orc_chf_amplitude_at_zero(t, &z, &kk, &rr);
  orc_chf_free(t);
  /* res * phase.conjugated() */
  if (!z) kk = (kk + (8 - phase) % 8) % 8;
  *zero = z; *k = z ? 0 : kk; *r = z ? 0 : rr;
  cplx val = scalar_to_complex(z, kk, rr);
  cplx res = c_mul(c_mul(a->omega, b->omega), val);
  *re = res.re; *im = res.im;
  return 0;
}

void orc_chf_export(const orc_chform *f, uint8_t *G, uint8_t *F, uint8_t *M, uint8_t *gamma,
                    uint8_t *v, uint8_t *s, uint8_t *phase) {
  size_t nn = (size_t)f->n * f->n;
  memcpy(G, f->G, nn); memcpy(F, f->F, nn); memcpy(M, f->M, nn);
  memcpy(gamma, f->gamma, f->n); memcpy(v, f->v, f->n); memcpy(s, f->s, f->n);
  *phase = f->phase;
}

/* form/statevector.rs:12-27 */
int orc_chf_statevector(const orc_chform *f, double *out) {
  size_t dim = (size_t)1 << f->n;
  uint8_t *bits = (uint8_t *)malloc(f->n);
  for (size_t i = 0; i < dim; i++) {
    for (int j = 0; j < f->n; j++) bits[j] = (uint8_t)((i >> j) & 1);
    int z, k, r;
    orc_chf_amplitude(f, bits, &z, &k, &r);
    cplx a = c_mul(scalar_to_complex(z, k, r), f->omega);
    out[2 * i] = a.re; out[2 * i + 1] = a.im;
  }
  free(bits);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* chi-term state (necstar/src/state/stabilizer_decomposed_state/mod.rs:18-73)*/
struct orc_state {
  int n, chi, cap;
  orc_chform **forms;
  int is_complex;
  uint8_t *cz, *ck; /* Scalar coefficients: zero flag, phase k */
  int64_t *cr;
  cplx *cc;         /* Complex64 coefficients */
  cplx g;           /* global_factor */
};

static orc_state *state_new(int n) {
  orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
  s->n = n;
  s->cap = 4;
  s->forms = (orc_chform **)calloc(s->cap, sizeof(orc_chform *));
  s->cz = (uint8_t *)calloc(s->cap, 1);
  s->ck = (uint8_t *)calloc(s->cap, 1);
  s->cr = (int64_t *)calloc(s->cap, sizeof(int64_t));
  s->cc = (cplx *)calloc(s->cap, sizeof(cplx));
  s->g.re = 1.0; s->g.im = 0.0;
  return s;
}
static void state_push(orc_state *s, orc_chform *f, int z, int k, int64_t r, cplx c) {
  if (s->chi == s->cap) {
    s->cap *= 2;
    s->forms = (orc_chform **)realloc(s->forms, sizeof(orc_chform *) * s->cap);
    s->cz = (uint8_t *)realloc(s->cz, s->cap);
    s->ck = (uint8_t *)realloc(s->ck, s->cap);
    s->cr = (int64_t *)realloc(s->cr, sizeof(int64_t) * s->cap);
    s->cc = (cplx *)realloc(s->cc, sizeof(cplx) * s->cap);
  }
  s->forms[s->chi] = f;
  s->cz[s->chi] = (uint8_t)z; s->ck[s->chi] = (uint8_t)k; s->cr[s->chi] = r; s->cc[s->chi] = c;
  s->chi++;
}
void orc_state_free(orc_state *s) {
  if (!s) return;
  for (int i = 0; i < s->chi; i++) orc_chf_free(s->forms[i]);
  free(s->forms); free(s->cz); free(s->ck); free(s->cr); free(s->cc);
  free(s);
}
orc_state *orc_state_clone(const orc_state *s) {
  orc_state *c = state_new(s->n);
  c->is_complex = s->is_complex;
  for (int i = 0; i < s->chi; i++)
    state_push(c, orc_chf_clone(s->forms[i]), s->cz[i], s->ck[i], s->cr[i], s->cc[i]);
  c->g = s->g;
  return c;
}
int orc_state_num_qubits(const orc_state *s) { return s->n; }
int orc_state_rank(const orc_state *s) { return s->chi; }
int orc_state_is_complex(const orc_state *s) { return s->is_complex; }
const orc_chform *orc_state_term(const orc_state *s, int i) { return s->forms[i]; }
void orc_state_coeff_scalar(const orc_state *s, int i, int *zero, int *k, int64_t *r) {
  *zero = s->cz[i]; *k = s->ck[i]; *r = s->cr[i];
}
/* Into<Complex64> for the coefficient */
static cplx coeff_value(const orc_state *s, int i) {
  if (s->is_complex) return s->cc[i];
  return scalar_to_complex(s->cz[i], s->ck[i], s->cr[i]);
}
void orc_state_coeff_complex(const orc_state *s, int i, double *re, double *im) {
  cplx c = coeff_value(s, i);
  *re = c.re; *im = c.im;
}
void orc_state_set_complex_coeffs(orc_state *s, const double *re_im) {
  s->is_complex = 1;
  for (int i = 0; i < s->chi; i++) { s->cc[i].re = re_im[2 * i]; s->cc[i].im = re_im[2 * i + 1]; }
}
void orc_state_global_factor(const orc_state *s, double *re, double *im) { *re = s->g.re; *im = s->g.im; }

/* (c_i.conj() * c_j).into(): exact for Scalar (state/types/scalar.rs:63-74,
 * :131-145), num-complex product for Complex64 (coefficient.rs:10-15) */
static cplx coeff_conj_mul(const orc_state *s, int i, const orc_state *t, int j) {
  if (s->is_complex || t->is_complex) {
    return c_mul(c_conj(coeff_value(s, i)), coeff_value(t, j));
  }
  if (s->cz[i] || t->cz[j]) { cplx z = {0.0, 0.0}; return z; }
  int k = ((8 - s->ck[i]) % 8 + t->ck[j]) % 8;
  return scalar_to_complex(0, k, s->cr[i] + t->cr[j]);
}

/* Amplify (state/types/scalar.rs:108-119, coefficient.rs:28-34) */
static void coeff_amplify(orc_state *s, int i, int factor) {
  if (s->is_complex) {
    double scale = pow(2.0, (double)factor / 2.0);
    s->cc[i] = c_scale(s->cc[i], scale);
  } else if (!s->cz[i]) {
    s->cr[i] -= factor;
  }
}

/* ------------------------------------------------------------------------ */
/* magic states (necstar/src/state/magic_states/cat_state.rs, t_state.rs)    */
typedef struct { int z, k; int64_t r; } scal;

static int apply_list(orc_chform *f, int cnt, const int *ops) {
  int e = 0;
  for (int i = 0; i < cnt && !e; i++) e = orc_chf_apply_gate(f, ops[3 * i], ops[3 * i + 1], ops[3 * i + 2]);
  return e;
}

/* cat_state.rs zero_minus_i_one_state */
static orc_chform *zero_minus_i_one(int m, int *err) {
  orc_chform *f = orc_chf_new(m, err);
  if (!f) return NULL;
  left_h(f, 0);
  for (int i = 1; i < m; i++) left_cx(f, 0, i);
  left_sdg(f, 0);
  return f;
}
/* cat_state.rs even_parity_state */
static orc_chform *even_parity(int m, int *err) {
  orc_chform *f = orc_chf_new(m, err);
  if (!f) return NULL;
  for (int i = 0; i < m - 1; i++) left_h(f, i);
  for (int i = 0; i < m - 1; i++) left_cx(f, i, m - 1);
  return f;
}
/* cat_state.rs even_parity_phase_flipped_state */
static orc_chform *even_parity_flipped(int m, int *err) {
  orc_chform *f = even_parity(m, err);
  if (!f) return NULL;
  for (int i = 0; i < m; i++)
    for (int j = i + 1; j < m; j++) left_cz(f, i, j);
  return f;
}

static orc_state *scalar_state(int n) { return state_new(n); }
static void push_scal(orc_state *s, orc_chform *f, int k, int64_t r) {
  cplx z = {0, 0};
  state_push(s, f, 0, k, r, z);
}

static orc_state *cat_state(int m, int *err);

/* stabilizer_decomposed_state/kron.rs:7-22 */
static orc_state *state_kron(const orc_state *a, const orc_state *b, int *err) {
  orc_state *r = scalar_state(a->n + b->n);
  for (int i = 0; i < a->chi; i++)
    for (int j = 0; j < b->chi; j++) {
      orc_chform *f = orc_chf_kron(a->forms[i], b->forms[j], err);
      if (!f) { orc_state_free(r); return NULL; }
      int z = a->cz[i] || b->cz[j];
      cplx zc = {0, 0};
      state_push(r, f, z, z ? 0 : (a->ck[i] + b->ck[j]) % 8, z ? 0 : a->cr[i] + b->cr[j], zc);
    }
  *err = 0;
  return r;
}

/* cat_state.rs construct_cat_state and helpers */
static orc_state *cat_state(int m, int *err) {
  orc_state *s;
  switch (m) {
    case 1: {
      s = scalar_state(1);
      push_scal(s, orc_chf_new(1, err), 0, 0);
      return s;
    }
    case 2: {
      s = scalar_state(2);
      orc_chform *f = orc_chf_new(2, err);
      int ops[] = {ORC_G_H, 0, 0, ORC_G_CX, 0, 1, ORC_G_S, 1, 0};
      apply_list(f, 3, ops);
      push_scal(s, f, 0, 0);
      return s;
    }
    case 4: {
      s = scalar_state(4);
      push_scal(s, zero_minus_i_one(4, err), 7, 1);
      push_scal(s, even_parity(4, err), 2, 0);
      return s;
    }
    case 6: {
      s = scalar_state(6);
      push_scal(s, zero_minus_i_one(6, err), 0, 2);
      push_scal(s, even_parity(6, err), 3, 1);
      push_scal(s, even_parity_flipped(6, err), 5, 1);
      return s;
    }
    case 3:
    case 5: {
      s = cat_state(m + 1, err);
      int nq = s->n;
      for (int i = 0; i < s->chi; i++) { /* reduce_cat_state */
        int det;
        if ((*err = orc_chf_project(s->forms[i], nq - 1, 0, &det))) { orc_state_free(s); return NULL; }
        if ((*err = orc_chf_discard(s->forms[i], nq - 1))) { orc_state_free(s); return NULL; }
      }
      s->n -= 1;
      return s;
    }
    default: {
      if (m <= 0) {
        *err = set_err(ORC_E_INVALID_NUM_QUBITS, "Number of qubits must be greater than zero, got %d.", m);
        return NULL;
      }
      orc_state *a = cat_state(m - 4, err);
      if (!a) return NULL;
      orc_state *b = cat_state(6, err);
      s = state_kron(a, b, err);
      orc_state_free(a); orc_state_free(b);
      if (!s) return NULL;
      int q0 = m - 5, q1 = m - 4;
      for (int i = 0; i < s->chi; i++) { /* project_ch_form_onto_cat_state */
        orc_chform *f = s->forms[i];
        int det;
        left_sdg(f, q0);
        left_cx(f, q0, q1);
        left_h(f, q0);
        if ((*err = orc_chf_project(f, q0, 0, &det)) || (*err = orc_chf_project(f, q1, 0, &det)) ||
            (*err = orc_chf_discard(f, q1)) || (*err = orc_chf_discard(f, q0))) {
          orc_state_free(s);
          return NULL;
        }
      }
      s->n -= 2;
      return s;
    }
  }
}

/* t_state.rs construct_t_tensor_state */
static orc_state *t_tensor_state(int t, int *err) {
  orc_state *cat = cat_state(t, err);
  if (!cat) return NULL;
  orc_state *s = scalar_state(cat->n);
  for (int i = 0; i < cat->chi; i++)
    push_scal(s, orc_chf_clone(cat->forms[i]), cat->ck[i], cat->cr[i] + 1);
  for (int i = 0; i < cat->chi; i++) {
    orc_chform *f = orc_chf_clone(cat->forms[i]);
    left_x(f, 0);
    left_s(f, 0);
    push_scal(s, f, (cat->ck[i] + 7) % 8, cat->cr[i] + 1);
  }
  orc_state_free(cat);
  *err = 0;
  return s;
}

static const char *gate_name(int op) {
  static const char *names[] = {"H", "X", "Y", "Z", "S", "Sdg", "SqrtX", "SqrtXdg",
                                "CX", "CZ", "Swap", "T", "Tdg", "CCX"};
  return (op >= 0 && op <= 13) ? names[op] : "?";
}

/* compiler/mod.rs:44-140 */
orc_state *orc_compile(int n, int ngates, const int *gates, int *err) {
  int t = 0, nops = 0;
  int *ops = (int *)malloc(sizeof(int) * 3 * (2 * (size_t)ngates + 1));
  for (int g = 0; g < ngates; g++) {
    int op = gates[4 * g], q0 = gates[4 * g + 1], q1 = gates[4 * g + 2];
    if (op <= ORC_G_SWAP) {
      ops[3 * nops] = op; ops[3 * nops + 1] = q0; ops[3 * nops + 2] = q1; nops++;
    } else if (op == ORC_G_T || op == ORC_G_TDG) {
      int anc = n + t;
      ops[3 * nops] = ORC_G_CX; ops[3 * nops + 1] = q0; ops[3 * nops + 2] = anc; nops++;
      if (op == ORC_G_TDG) { ops[3 * nops] = ORC_G_SDG; ops[3 * nops + 1] = q0; ops[3 * nops + 2] = 0; nops++; }
      t++;
    } else {
      free(ops);
      *err = set_err(ORC_E_GATE_NOT_SUPPORTED, "Gate %s is not supported.", gate_name(op));
      return NULL;
    }
  }
  if (t == 0) {
    /* from_clifford_circuit(...).unwrap(): the reference panics on error */
    orc_chform *f = orc_chf_new(n, err);
    if (!f) { free(ops); return NULL; }
    if ((*err = apply_list(f, nops, ops))) { orc_chf_free(f); free(ops); return NULL; }
    orc_state *s = scalar_state(n);
    push_scal(s, f, 0, 0);
    free(ops);
    return s;
  }
  orc_state *ts = t_tensor_state(t, err);
  if (!ts) { free(ops); return NULL; }
  orc_state *out = scalar_state(n);
  for (int i = 0; i < ts->chi; i++) {
    orc_chform *base = orc_chf_new(n, err);
    if (!base) { orc_state_free(ts); orc_state_free(out); free(ops); return NULL; }
    orc_chform *full = orc_chf_kron(base, ts->forms[i], err);
    orc_chf_free(base);
    if ((*err = apply_list(full, nops, ops))) {
      orc_chf_free(full); orc_state_free(ts); orc_state_free(out); free(ops);
      return NULL;
    }
    int ok = 1, ndet = 0;
    for (int q = n + t - 1; q >= n; q--) {
      int det;
      if (orc_chf_project(full, q, 0, &det)) { ok = 0; break; }
      if (det) ndet++;
    }
    if (ok) {
      for (int q = n + t - 1; q >= n; q--) {
        if ((*err = orc_chf_discard(full, q))) {
          orc_chf_free(full); orc_state_free(ts); orc_state_free(out); free(ops);
          return NULL;
        }
      }
      push_scal(out, full, ts->ck[i], ts->cr[i] - ndet); /* amplify(#det) */
    } else {
      orc_chf_free(full);
    }
  }
  orc_state_free(ts);
  free(ops);
  *err = 0;
  return out;
}

/* ------------------------------------------------------------------------ */
/* chi-term state operations                                                 */

/* stabilizer_decomposed_state/gates.rs:7-116 */
int orc_state_apply_gate(orc_state *s, int op, int q0, int q1, int q2) {
  (void)q2;
  if (op == ORC_G_T || op == ORC_G_TDG || op == ORC_G_CCX)
    return set_err(ORC_E_NON_CLIFFORD, "Direct application of non-Clifford gate %s is not supported.",
                   gate_name(op));
  for (int i = 0; i < s->chi; i++) {
    int e = orc_chf_apply_gate(s->forms[i], op, q0, q1);
    if (e) return e;
  }
  return 0;
}

/* norm.rs:8-30 */
int orc_state_norm_squared(const orc_state *s, double *out) {
  cplx sum = {0.0, 0.0};
  for (int i = 0; i < s->chi; i++) {
    int z, k, r, e;
    double re, im;
    if ((e = orc_chf_inner_product(s->forms[i], s->forms[i], &z, &k, &r, &re, &im))) return e;
    cplx ip = {re, im};
    sum = c_add(sum, c_mul(coeff_conj_mul(s, i, s, i), ip));
    for (int j = i + 1; j < s->chi; j++) {
      if ((e = orc_chf_inner_product(s->forms[i], s->forms[j], &z, &k, &r, &re, &im))) return e;
      cplx ipo = {re, im};
      cplx term = c_mul(coeff_conj_mul(s, i, s, j), ipo);
      sum = c_add(sum, c_add(term, c_conj(term)));
    }
  }
  *out = sum.re * (s->g.re * s->g.re + s->g.im * s->g.im);
  return 0;
}
int orc_state_norm(const orc_state *s, double *out) {
  double n2;
  int e = orc_state_norm_squared(s, &n2);
  if (e) return e;
  *out = sqrt(n2);
  return 0;
}

/* gate_application.rs:186-210 apply_pauli (dense in qubit order, sparse in
 * term order); 0=I,1=X,2=Y,3=Z */
static int apply_pauli(orc_chform *f, int dense, int len, const int *ops, const int *qubits) {
  int e = 0;
  for (int i = 0; i < len && !e; i++) {
    int q = dense ? i : qubits[i];
    switch (ops[i]) {
      case 1: e = left_x(f, q); break;
      case 2: e = left_y(f, q); break;
      case 3: e = left_z(f, q); break;
      default: break;
    }
  }
  return e;
}

/* exp_value.rs:7-61 */
int orc_state_exp_value(const orc_state *s, int dense, int len, const int *ops, const int *qubits,
                        double *out) {
  if (dense) {
    if (len != s->n)
      return set_err(ORC_E_PAULI_LENGTH, "Invalid Pauli string length: expected %d, found %d.", s->n, len);
  } else {
    int mx = 0;
    for (int i = 0; i < len; i++)
      if (qubits[i] > mx) mx = qubits[i];
    if (mx >= s->n)
      return set_err(ORC_E_PAULI_LENGTH, "Invalid Pauli string length: expected %d, found %d.", s->n,
                     mx + 1);
  }
  cplx ev = {0.0, 0.0};
  for (int i = 0; i < s->chi; i++) {
    orc_chform *evolved = orc_chf_clone(s->forms[i]);
    int e = apply_pauli(evolved, dense, len, ops, qubits);
    if (e) { orc_chf_free(evolved); return e; }
    int z, k, r;
    double re, im;
    if ((e = orc_chf_inner_product(s->forms[i], evolved, &z, &k, &r, &re, &im))) {
      orc_chf_free(evolved); return e;
    }
    cplx ipd = {re, im};
    ev = c_add(ev, c_mul(coeff_conj_mul(s, i, s, i), ipd));
    for (int j = i + 1; j < s->chi; j++) {
      if ((e = orc_chf_inner_product(s->forms[j], evolved, &z, &k, &r, &re, &im))) {
        orc_chf_free(evolved); return e;
      }
      cplx ipo = {re, im};
      cplx term = c_mul(coeff_conj_mul(s, j, s, i), ipo);
      ev = c_add(ev, c_add(term, c_conj(term)));
    }
    orc_chf_free(evolved);
  }
  *out = ev.re * (s->g.re * s->g.re + s->g.im * s->g.im);
  return 0;
}

/* stabilizer_decomposed_state/inner_product.rs:9-19 */
int orc_state_inner_product(const orc_state *a, const orc_state *b, double *re, double *im) {
  cplx res = {0.0, 0.0};
  for (int i = 0; i < a->chi; i++)
    for (int j = 0; j < b->chi; j++) {
      int z, k, r, e;
      double pre, pim;
      if ((e = orc_chf_inner_product(a->forms[i], b->forms[j], &z, &k, &r, &pre, &pim))) return e;
      cplx ip = {pre, pim};
      res = c_add(res, c_mul(coeff_conj_mul(a, i, b, j), ip));
    }
  res = c_mul(c_mul(res, c_conj(a->g)), b->g);
  *re = res.re; *im = res.im;
  return 0;
}

/* projection.rs:27-57 */
int orc_state_project_unnormalized(orc_state *s, int q, int outcome) {
  if (q < 0 || q >= s->n) return err_oob(q, s->n);
  int w = 0;
  for (int i = 0; i < s->chi; i++) {
    int det;
    int e = orc_chf_project(s->forms[i], q, outcome, &det);
    if (e) { orc_chf_free(s->forms[i]); continue; }
    s->forms[w] = s->forms[i];
    s->cz[w] = s->cz[i]; s->ck[w] = s->ck[i]; s->cr[w] = s->cr[i]; s->cc[w] = s->cc[i];
    if (!det) coeff_amplify(s, w, -1);
    w++;
  }
  s->chi = w;
  return 0;
}

/* projection.rs:10-21 */
int orc_state_project_normalized(orc_state *s, int q, int outcome) {
  int e = orc_state_project_unnormalized(s, q, outcome);
  if (e) return e;
  double nrm;
  if ((e = orc_state_norm(s, &nrm))) return e;
  if (fabs(nrm) < 1e-12 || isnan(nrm)) return err_impossible_state(q, outcome);
  s->g.re = s->g.re / nrm;
  s->g.im = s->g.im / nrm;
  return 0;
}

/* stabilizer_decomposed_state/discard.rs:5-12 */
int orc_state_discard(orc_state *s, int q) {
  for (int i = 0; i < s->chi; i++) {
    int e = orc_chf_discard(s->forms[i], q);
    if (e) return e;
  }
  s->n -= 1;
  return 0;
}

/* stabilizer_decomposed_state/statevector.rs:12-24 */
int orc_state_statevector(const orc_state *s, double *out) {
  if (s->n > 28)
    return set_err(ORC_E_SV_TOO_LARGE, "Calculating the statevector for a state with %d qubits is not feasible.",
                   s->n);
  size_t dim = (size_t)1 << s->n;
  for (size_t x = 0; x < 2 * dim; x++) out[x] = 0.0;
  double *tmp = (double *)malloc(sizeof(double) * 2 * dim);
  for (int i = 0; i < s->chi; i++) {
    orc_chf_statevector(s->forms[i], tmp);
    cplx c = coeff_value(s, i);
    for (size_t x = 0; x < dim; x++) {
      cplx a = {tmp[2 * x], tmp[2 * x + 1]};
      cplx acc = {out[2 * x], out[2 * x + 1]};
      acc = c_add(acc, c_mul(a, c));
      out[2 * x] = acc.re; out[2 * x + 1] = acc.im;
    }
  }
  for (size_t x = 0; x < dim; x++) {
    cplx acc = {out[2 * x], out[2 * x + 1]};
    acc = c_mul(acc, s->g);
    out[2 * x] = acc.re; out[2 * x + 1] = acc.im;
  }
  free(tmp);
  return 0;
}

/* one amplitude <x|psi>: the to_statevector arithmetic for a single x */
int orc_state_amplitude(const orc_state *s, const uint8_t *bits, double *re, double *im, int *zero_t,
                        int *k_t, int *r_t) {
  cplx acc = {0.0, 0.0};
  for (int i = 0; i < s->chi; i++) {
    int z, k, r;
    orc_chf_amplitude(s->forms[i], bits, &z, &k, &r);
    if (zero_t) { zero_t[i] = z; k_t[i] = k; r_t[i] = r; }
    cplx a = c_mul(scalar_to_complex(z, k, r), s->forms[i]->omega);
    acc = c_add(acc, c_mul(a, coeff_value(s, i)));
  }
  acc = c_mul(acc, s->g);
  *re = acc.re; *im = acc.im;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* RNG: rand 0.8.5 StdRng (= rand_chacha 0.3 ChaCha12Rng over rand_core 0.6
 * BlockRng with a 4-block buffer), Standard f64, UniformFloat<f64>, and
 * rand_distr 0.4.3 Binomial (BINV / BTPE).  Third-party algorithms, restated
 * from their published descriptions; the stream is not pinned by any
 * reference fixture. */
static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
#define QR(a, b, c, d)                               \
  a += b; d ^= a; d = rotl32(d, 16);                 \
  c += d; b ^= c; b = rotl32(b, 12);                 \
  a += b; d ^= a; d = rotl32(d, 8);                  \
  c += d; b ^= c; b = rotl32(b, 7);

void orc_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds,
                      uint32_t out[16]) {
  uint32_t st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                     key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                     (uint32_t)counter, (uint32_t)(counter >> 32), (uint32_t)stream,
                     (uint32_t)(stream >> 32)};
  uint32_t x[16];
  memcpy(x, st, sizeof st);
  for (int i = 0; i < rounds; i += 2) {
    QR(x[0], x[4], x[8], x[12]);
    QR(x[1], x[5], x[9], x[13]);
    QR(x[2], x[6], x[10], x[14]);
    QR(x[3], x[7], x[11], x[15]);
    QR(x[0], x[5], x[10], x[15]);
    QR(x[1], x[6], x[11], x[12]);
    QR(x[2], x[7], x[8], x[13]);
    QR(x[3], x[4], x[9], x[14]);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + st[i];
}

struct orc_rng {
  uint32_t key[8];
  uint64_t counter;
  uint32_t buf[64];
  int index;
};

orc_rng *orc_rng_new(const uint8_t *seed) {
  orc_rng *r = (orc_rng *)calloc(1, sizeof(orc_rng));
  for (int i = 0; i < 8; i++)
    r->key[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) |
                ((uint32_t)seed[4 * i + 2] << 16) | ((uint32_t)seed[4 * i + 3] << 24);
  r->counter = 0;
  r->index = 64;
  return r;
}
void orc_rng_free(orc_rng *r) { free(r); }
static void rng_refill(orc_rng *r) {
  for (int b = 0; b < 4; b++) orc_chacha_block(r->key, r->counter + (uint64_t)b, 0, 12, r->buf + 16 * b);
  r->counter += 4;
  r->index = 0;
}
uint32_t orc_rng_next_u32(orc_rng *r) {
  if (r->index >= 64) rng_refill(r);
  return r->buf[r->index++];
}
uint64_t orc_rng_next_u64(orc_rng *r) {
  uint64_t lo, hi;
  if (r->index < 63) {
    lo = r->buf[r->index]; hi = r->buf[r->index + 1]; r->index += 2;
  } else if (r->index >= 64) {
    rng_refill(r);
    lo = r->buf[0]; hi = r->buf[1]; r->index = 2;
  } else {
    lo = r->buf[63];
    rng_refill(r);
    hi = r->buf[0]; r->index = 1;
  }
  return (hi << 32) | lo;
}
double orc_rng_gen_f64(orc_rng *r) {
  uint64_t v = orc_rng_next_u64(r) >> 11;
  return (double)v * (1.0 / 9007199254740992.0);
}

/* UniformFloat<f64>::new / sample */
typedef struct { double low, scale; } unif;
static double f64_from_bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t f64_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }
static unif unif_new(double low, double high) {
  double max_rand = f64_from_bits(((uint64_t)1023 << 52) | (UINT64_MAX >> 12)) - 1.0;
  double scale = high - low;
  for (;;) {
    if (!(scale * max_rand + low >= high)) break;
    scale = f64_from_bits(f64_bits(scale) - 1);
  }
  unif u = {low, scale};
  return u;
}
static double unif_sample(const unif *u, orc_rng *r) {
  uint64_t v = orc_rng_next_u64(r) >> 12;
  double v12 = f64_from_bits(((uint64_t)1023 << 52) | v);
  double v01 = v12 - 1.0;
  return v01 * u->scale + u->low;
}

static int64_t f64_to_i64(double x) {
  if (isnan(x)) return 0;
  if (x >= 9.2233720368547758e18) return INT64_MAX;
  if (x <= -9.2233720368547758e18) return INT64_MIN;
  return (int64_t)x;
}

/* compiler-rt __powidf2 (what f64::powi lowers to) */
static double powi_f64(double a, int b) {
  int recip = b < 0;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

static double stirling(double a) {
  double a2 = a * a;
  return (13860. - (462. - (132. - (99. - 140. / a2) / a2) / a2) / a2) / a / 166320.;
}

/* rand_distr 0.4.3 Binomial::sample (caller validated 0 <= p <= 1) */
uint64_t orc_rng_binomial(orc_rng *rng, uint64_t nn, double pp) {
  if (pp == 0.0) return 0;
  if (pp == 1.0) return nn;
  double p = pp <= 0.5 ? pp : 1.0 - pp;
  uint64_t result;
  double q = 1. - p;
  const double BINV_THRESHOLD = 10.;
  if ((double)nn * p < BINV_THRESHOLD && nn <= (uint64_t)INT32_MAX) {
    double s = p / q;
    double a = ((double)(nn + 1)) * s;
    double r = powi_f64(q, (int)nn);
    double u = orc_rng_gen_f64(rng);
    uint64_t x = 0;
    while (u > r) {
      u -= r;
      x += 1;
      r *= a / (double)x - s;
    }
    result = x;
  } else {
    const int64_t SQUEEZE_THRESHOLD = 20;
    double n = (double)nn;
    double np = n * p;
    double npq = np * q;
    double f_m = np + p;
    int64_t m = f64_to_i64(f_m);
    double p1 = floor(2.195 * sqrt(npq) - 4.6 * q) + 0.5;
    double x_m = (double)m + 0.5;
    double x_l = x_m - p1;
    double x_r = x_m + p1;
    double c = 0.134 + 20.5 / (15.3 + (double)m);
    double p2 = p1 * (1. + 2. * c);
#define LAMBDA(a) ((a) * (1. + 0.5 * (a)))
    double lambda_l = LAMBDA((f_m - x_l) / (f_m - x_l * p));
    double lambda_r = LAMBDA((x_r - f_m) / (x_r * q));
#undef LAMBDA
    double p3 = p2 + c / lambda_l;
    double p4 = p3 + c / lambda_r;
    int64_t y;
    unif gen_u = unif_new(0., p4);
    unif gen_v = unif_new(0., 1.);
    for (;;) {
      double u = unif_sample(&gen_u, rng);
      double v = unif_sample(&gen_v, rng);
      if (!(u > p1)) {
        y = f64_to_i64(x_m - p1 * v + u);
        break;
      }
      if (!(u > p2)) {
        double x = x_l + (u - p1) / c;
        v = v * c + 1.0 - fabs(x - x_m) / p1;
        if (v > 1.) continue;
        y = f64_to_i64(x);
      } else if (!(u > p3)) {
        y = f64_to_i64(x_l + log(v) / lambda_l);
        if (y < 0) continue;
        v *= (u - p2) * lambda_l;
      } else {
        y = f64_to_i64(x_r - log(v) / lambda_r);
        if (y > 0 && (uint64_t)y > nn) continue;
        v *= (u - p3) * lambda_r;
      }
      int64_t kk = llabs(y - m);
      if (!(kk > SQUEEZE_THRESHOLD && (double)kk < 0.5 * npq - 1.)) {
        double s = p / q;
        double a = s * (n + 1.);
        double f = 1.0;
        if (m < y) {
          int64_t i = m;
          for (;;) {
            i += 1;
            f *= a / (double)i - s;
            if (i == y) break;
          }
        } else if (m > y) {
          int64_t i = y;
          for (;;) {
            i += 1;
            f /= a / (double)i - s;
            if (i == m) break;
          }
        }
        if (v > f) continue;
        break;
      }
      double k = (double)kk;
      double rho = (k / npq) * ((k * (k / 3. + 0.625) + 1. / 6.) / npq + 0.5);
      double t = -0.5 * k * k / npq;
      double alpha = log(v);
      if (alpha < t - rho) break;
      if (alpha > t + rho) continue;
      double x1 = (double)(y + 1);
      double f1 = (double)(m + 1);
      double z = (double)(f64_to_i64(n) + 1 - m);
      double w = (double)(f64_to_i64(n) - y + 1);
      if (alpha > x_m * log(f1 / x1) + (n - (double)m + 0.5) * log(z / w) +
                      (double)(y - m) * log(w * p / (x1 * q)) + stirling(f1) + stirling(z) -
                      stirling(x1) - stirling(w))
        continue;
      break;
    }
    result = (uint64_t)y;
  }
  if (p != pp) return nn - result;
  return result;
}

/* ------------------------------------------------------------------------ */
/* measurement.rs / sampling.rs                                              */

/* stabilizer_decomposed_state/mod.rs:44-67 */
static int validate_qargs(const orc_state *s, int nq, const int *qargs) {
  if (nq == 0) return set_err(ORC_E_EMPTY_QARGS, "Qubit index list is empty.");
  for (int i = 0; i < nq; i++)
    if (qargs[i] < 0 || qargs[i] >= s->n) return err_oob(qargs[i], s->n);
  if (nq > 1) {
    int *sorted = (int *)malloc(sizeof(int) * nq);
    memcpy(sorted, qargs, sizeof(int) * nq);
    for (int i = 1; i < nq; i++) { /* insertion sort */
      int v = sorted[i], j = i - 1;
      while (j >= 0 && sorted[j] > v) { sorted[j + 1] = sorted[j]; j--; }
      sorted[j + 1] = v;
    }
    for (int i = 0; i + 1 < nq; i++)
      if (sorted[i] == sorted[i + 1]) {
        int d = sorted[i];
        free(sorted);
        return set_err(ORC_E_DUPLICATE_QARG, "Duplicate qubit index found: %d.", d);
      }
    free(sorted);
  }
  return 0;
}

/* measurement.rs:34-103 */
static int measure_single(orc_state *s, int q, orc_rng *rng, int *outcome) {
  orc_state *zero = orc_state_clone(s), *one = orc_state_clone(s);
  int e;
  orc_state_project_unnormalized(zero, q, 0);
  orc_state_project_unnormalized(one, q, 1);
  double p0, p1;
  if ((e = orc_state_norm_squared(zero, &p0)) || (e = orc_state_norm_squared(one, &p1))) {
    orc_state_free(zero); orc_state_free(one); return e;
  }
  double total = p0 + p1;
  if (fabs(total) < 1e-12) {
    orc_state_free(zero); orc_state_free(one);
    return set_err(ORC_E_NOT_IMPLEMENTED, "Not implemented: Measurement on a zero-norm state is not possible.");
  }
  double pz = p0 / total;
  double sample = orc_rng_gen_f64(rng);
  int res;
  double nsq;
  orc_state *chosen;
  if (sample < pz) { res = 0; nsq = p0; chosen = zero; orc_state_free(one); }
  else { res = 1; nsq = p1; chosen = one; orc_state_free(zero); }
  /* *self = chosen */
  for (int i = 0; i < s->chi; i++) orc_chf_free(s->forms[i]);
  free(s->forms); free(s->cz); free(s->ck); free(s->cr); free(s->cc);
  *s = *chosen;
  free(chosen);
  double nrm = sqrt(nsq);
  if (fabs(nrm) < 1e-12)
    return set_err(ORC_E_NOT_IMPLEMENTED, "Not implemented: Measurement resulted in a zero-norm state, which is not physical.");
  cplx f = {1.0 / nrm, 0.0};
  s->g = c_mul(s->g, f);
  *outcome = res;
  return 0;
}

int orc_state_measure(orc_state *s, int nq, const int *qargs, const uint8_t *seed, uint8_t *outcomes) {
  int e = validate_qargs(s, nq, qargs);
  if (e) return e;
  orc_rng *rng = orc_rng_new(seed);
  for (int i = 0; i < nq; i++) {
    int o = 0;
    if ((e = measure_single(s, qargs[i], rng, &o))) { orc_rng_free(rng); return e; }
    outcomes[i] = (uint8_t)o;
  }
  orc_rng_free(rng);
  return 0;
}

typedef struct {
  const int *q, *bit;
  int len;
  int max_out, nout;
  uint64_t *bits, *counts;
  orc_rng *rng;
} sample_ctx;

/* sampling.rs:80-189 */
static int recursive_sample(const orc_state *s, sample_ctx *cx, int idx, uint64_t shots, uint64_t lo,
                            uint64_t hi) {
  if (shots == 0) return 0;
  if (idx == cx->len) {
    if (cx->nout >= cx->max_out) return err_internal("sample output buffer too small");
    cx->bits[2 * cx->nout] = lo; cx->bits[2 * cx->nout + 1] = hi;
    cx->counts[cx->nout] = shots;
    cx->nout++;
    return 0;
  }
  int q = cx->q[idx], b = cx->bit[idx], e;
  orc_state *zero = orc_state_clone(s), *one = orc_state_clone(s);
  orc_state_project_unnormalized(zero, q, 0);
  orc_state_project_unnormalized(one, q, 1);
  double nz1, nz2, no;
  if ((e = orc_state_norm_squared(zero, &nz1)) || (e = orc_state_norm_squared(zero, &nz2)) ||
      (e = orc_state_norm_squared(one, &no))) {
    orc_state_free(zero); orc_state_free(one); return e;
  }
  double pz = nz1 / (nz2 + no);
  if (!isnan(pz)) pz = pz < 0.0 ? 0.0 : (pz > 1.0 ? 1.0 : pz);
  if (!(pz >= 0.0)) {
    orc_state_free(zero); orc_state_free(one);
    return set_err(ORC_E_NOT_IMPLEMENTED, "p < 0 or is NaN in binomial distribution");
  }
  uint64_t nzeros = orc_rng_binomial(cx->rng, shots, pz);
  uint64_t nones = shots - nzeros;
  if (nzeros > 0) {
    if ((e = orc_state_discard(zero, q)) || (e = recursive_sample(zero, cx, idx + 1, nzeros, lo, hi))) {
      orc_state_free(zero); orc_state_free(one); return e;
    }
  }
  if (nones > 0) {
    uint64_t nlo = lo, nhi = hi;
    if (b < 64) nlo |= (uint64_t)1 << b; else nhi |= (uint64_t)1 << (b - 64);
    if ((e = orc_state_apply_gate(one, ORC_G_X, q, 0, 0)) || (e = orc_state_discard(one, q)) ||
        (e = recursive_sample(one, cx, idx + 1, nones, nlo, nhi))) {
      orc_state_free(zero); orc_state_free(one); return e;
    }
  }
  orc_state_free(zero); orc_state_free(one);
  return 0;
}

/* sampling.rs:13-77 */
int orc_state_sample(const orc_state *s, int nq, const int *qargs, uint64_t shots, const uint8_t *seed,
                     int max_out, uint64_t *out_bits, uint64_t *out_counts, int *nout) {
  int e = validate_qargs(s, nq, qargs);
  if (e) return e;
  if (nq > 128) return set_err(ORC_E_SAMPLING_TOO_MANY, "Sampling more than 128 qubits is not supported.");
  int *qs = (int *)malloc(sizeof(int) * nq), *bs = (int *)malloc(sizeof(int) * nq);
  for (int i = 0; i < nq; i++) { qs[i] = qargs[i]; bs[i] = i; }
  for (int i = 1; i < nq; i++) { /* stable sort, descending qarg */
    int vq = qs[i], vb = bs[i], j = i - 1;
    while (j >= 0 && qs[j] < vq) { qs[j + 1] = qs[j]; bs[j + 1] = bs[j]; j--; }
    qs[j + 1] = vq; bs[j + 1] = vb;
  }
  sample_ctx cx = {qs, bs, nq, max_out, 0, out_bits, out_counts, orc_rng_new(seed)};
  e = recursive_sample(s, &cx, 0, shots, 0, 0);
  orc_rng_free(cx.rng);
  free(qs); free(bs);
  *nout = cx.nout;
  return e;
}
