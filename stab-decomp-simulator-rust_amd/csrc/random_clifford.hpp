// random_clifford.hpp -- uniformly random Clifford circuits (host only).
//
// Follows crates/stabilizer-ch-form-rust/src/circuit/random_clifford.rs:7-218:
// Bravyi & Maslov, "Hadamard-free circuits expose the structure of the
// Clifford group" (arXiv:2003.09412v2), U = F1 * H * S * F2, with H and the
// permutation S drawn from the quantum Mallows distribution (Algorithm 1,
// Lemma 6) and the Hadamard-free layers F1, F2 filled under rules C1-C5 of
// Theorem 1.  The draws are consumed in the reference's order from rand
// 0.8.5's StdRng, so a seed gives the reference's gate list:
//   rng.gen::<f64>()        -> gen_f64 (53 high bits of next_u64)
//   rng.gen_range(0..=1u8)  -> next_u32() >> 31 (UniformInt<u8> widens to
//                              u32; for range 2 the modulus zone rejects
//                              nothing, and the result is the high word of
//                              v * 2)
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

#include "rng.hpp"

namespace nx {

struct CliffordDraw {
  std::vector<uint8_t> h, pz, px;
  std::vector<int64_t> s;
  // n x n row-major bit matrices; gamma symmetric, delta unit lower triangular
  std::vector<uint8_t> gamma1, delta1, gamma2, delta2;
};

// Rust's saturating `f64 as isize`
static inline int64_t rust_f64_to_isize(double x) {
  if (std::isnan(x)) return 0;
  if (x >= 9223372036854775807.0) return INT64_MAX;
  if (x <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)x;
}

// 4.0f64.powi(m): exact powers of two, +inf from m = 512 on
static inline double four_pow(int64_t m) { return m >= 512 ? INFINITY : std::ldexp(1.0, (int)(2 * m)); }

// random_clifford.rs:7-46
static void sample_quantum_mallows(int64_t n, ChaCha12Rng& rng, std::vector<uint8_t>& h, std::vector<int64_t>& s) {
  h.assign(n, 0);
  s.assign(n, 0);
  std::vector<int64_t> avail(n);
  for (int64_t i = 0; i < n; ++i) avail[i] = i;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t m = n - i;
    const double r = rng.gen_f64();
    // inverse transform for P(a) ~ 2^-a over a in [1, 2m]
    const double scaled = r * (four_pow(m) - 1.0);
    const double val = scaled + 1.0;
    const double ceil_log = std::ceil(std::log2(val));
    const int64_t cl = rust_f64_to_isize(ceil_log);
    int64_t a = cl >= 0 ? (2 * m + 1) - cl : INT64_MAX;  // cl < 0 cannot occur (val >= 1)
    if (a <= 0) a = 1;
    uint64_t k1;
    if (a <= m) {
      h[i] = 1;
      k1 = (uint64_t)a;
    } else {
      h[i] = 0;
      k1 = (uint64_t)(2 * m - a + 1);
    }
    const uint64_t k0 = k1 - 1;  // usize arithmetic: 0 - 1 wraps and takes the pop branch
    if (k0 < (uint64_t)avail.size()) {
      s[i] = avail[k0];
      avail.erase(avail.begin() + (int64_t)k0);
    } else {
      s[i] = avail.back();
      avail.pop_back();
    }
  }
}

static inline uint8_t draw_bit(ChaCha12Rng& rng) { return (uint8_t)(rng.next_u32() >> 31); }

// random_clifford.rs:63-119
static CliffordDraw generate_clifford_params(int64_t n, ChaCha12Rng& rng) {
  CliffordDraw d;
  sample_quantum_mallows(n, rng, d.h, d.s);
  const size_t nn = (size_t)(n * n);
  d.gamma1.assign(nn, 0);
  d.gamma2.assign(nn, 0);
  d.delta1.assign(nn, 0);
  d.delta2.assign(nn, 0);
  for (int64_t i = 0; i < n; ++i) d.delta1[i * n + i] = d.delta2[i * n + i] = 1;
  d.pz.resize(n);
  d.px.resize(n);
  for (int64_t i = 0; i < n; ++i) d.pz[i] = draw_bit(rng);
  for (int64_t i = 0; i < n; ++i) d.px[i] = draw_bit(rng);
  for (int64_t i = 0; i < n; ++i) {
    d.gamma2[i * n + i] = draw_bit(rng);
    if (d.h[i]) d.gamma1[i * n + i] = draw_bit(rng);
  }
  for (int64_t i = 1; i < n; ++i)
    for (int64_t j = 0; j < i; ++j) {
      const uint8_t g2 = draw_bit(rng);
      d.gamma2[i * n + j] = d.gamma2[j * n + i] = g2;
      d.delta2[i * n + j] = draw_bit(rng);
      const bool hi = d.h[i] != 0, hj = d.h[j] != 0;
      const int64_t si = d.s[i], sj = d.s[j];
      // C1, C2: CZ(i, j) of F1 is free
      if (((si < sj || hj) && hi) || (hj && sj < si)) {
        const uint8_t b = draw_bit(rng);
        d.gamma1[i * n + j] = d.gamma1[j * n + i] = b;
      }
      // C3 - C5: CNOT entry (i, j) of F1 is free
      if ((si < sj || hj) && (si > sj || !hi) && (hj || !hi)) d.delta1[i * n + j] = draw_bit(rng);
    }
  return d;
}

struct GateOut {
  std::vector<int32_t> g;  // [ngates][4] = (op, q0, q1, 0) in the public gate codes
  void add(int op, int64_t a, int64_t b = 0) {
    g.push_back(op);
    g.push_back((int32_t)a);
    g.push_back((int32_t)b);
    g.push_back(0);
  }
};

// random_clifford.rs:122-165: CNOT layer, then S and CZ, then the Paulis
static void hadamard_free_layer(GateOut& out, int64_t n, const std::vector<uint8_t>& gamma,
                                const std::vector<uint8_t>& delta, const uint8_t* pz, const uint8_t* px) {
  enum { H = 0, X = 1, Y = 2, Z = 3, S = 4, CX = 8, CZ = 9 };
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = j + 1; i < n; ++i)
      if (delta[i * n + j]) out.add(CX, j, i);
  for (int64_t i = 0; i < n; ++i) {
    if (gamma[i * n + i]) out.add(S, i);
    for (int64_t j = 0; j < i; ++j)
      if (gamma[i * n + j]) out.add(CZ, i, j);
  }
  if (pz && px)
    for (int64_t i = 0; i < n; ++i) {
      if (pz[i] && px[i]) out.add(Y, i);
      else if (pz[i]) out.add(Z, i);
      else if (px[i]) out.add(X, i);
    }
}

// random_clifford.rs:168-182: SWAPs that realise the permutation
static void permutation_layer(GateOut& out, const std::vector<int64_t>& s) {
  const int64_t n = (int64_t)s.size();
  std::vector<int64_t> p(n);
  for (int64_t i = 0; i < n; ++i) p[i] = i;
  for (int64_t i = 0; i < n; ++i) {
    if (p[i] == s[i]) continue;
    int64_t j = 0;
    while (p[j] != s[i]) ++j;
    out.add(10, i, j);
    std::swap(p[i], p[j]);
  }
}

// random_clifford.rs:189-218; upper bound on the gate count: 2 n^2 + 3 n
static std::vector<int32_t> random_clifford_gates(int64_t n, ChaCha12Rng& rng) {
  GateOut out;
  if (n == 0) return out.g;
  const CliffordDraw d = generate_clifford_params(n, rng);
  hadamard_free_layer(out, n, d.gamma2, d.delta2, d.pz.data(), d.px.data());
  permutation_layer(out, d.s);
  for (int64_t i = 0; i < n; ++i)
    if (d.h[i]) out.add(0, i);
  hadamard_free_layer(out, n, d.gamma1, d.delta1, nullptr, nullptr);
  return out.g;
}

}  // namespace nx
