// device_form.hpp -- CH-form tables held by one wavefront (gfx950, wave64).
//
// Layout in registers: the term's n x n GF(2) matrices G, F, M are split by
// rows: lane l holds rows l, l+64, ... (slot sl = row / 64), each row as W
// u64 words (bit j of word w = column 64*w + j).  gamma is per row (k/2 mod
// 4, since the reference only ever stores even k).  v, s and the phase
// factor are wave-uniform.  Left multiplications are row operations: the
// source row is broadcast with v_readlane and only the owning lane writes.
// Right multiplications are column operations: every lane edits bits of its
// own rows.  All control flow depends on wave-uniform values only.
//
// Gate rules restate stabilizer-ch-form-rust/src/form/ (file:line cited per
// function); the resolve-superposition fan-outs are applied as one batched
// column update, which is equal to the reference's sequential loop because
// none of the sequential steps reads a column written by an earlier step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.hpp"

namespace nx {

typedef uint64_t u64;
typedef uint32_t u32;

__device__ __forceinline__ u64 rl64(u64 x, int lane) {
  u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)x, lane);
  u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(x >> 32), lane);
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ int rl32(int x, int lane) { return __builtin_amdgcn_readlane(x, lane); }
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
// wave-uniform copy (scalar registers) of a value every lane holds equally
__device__ __forceinline__ u64 uni64(u64 x) {
  const u32 lo = (u32)__builtin_amdgcn_readfirstlane((int)(u32)x);
  const u32 hi = (u32)__builtin_amdgcn_readfirstlane((int)(u32)(x >> 32));
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ int par64(u64 x) { return __builtin_popcountll(x) & 1; }
__device__ __forceinline__ u32 lo32(u64 x) { return (u32)x; }
__device__ __forceinline__ u32 hi32(u64 x) { return (u32)(x >> 32); }
// a ^ b ^ c in one gfx950 three-input bitwise operation (truth table 0x96)
__device__ __forceinline__ u32 xor3(u32 a, u32 b, u32 c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96); }
// x ^ (y & m) for a lane mask m (all ones or zero), one three-input
// operation (truth table 0x78) per dword
// (the builtin returns int: each half is taken as u32 before widening)
__device__ __forceinline__ u64 xor_and(u64 x, u64 y, u32 m) {
  const u32 lo = (u32)__builtin_amdgcn_bitop3_b32((u32)x, (u32)y, m, 0x78);
  const u32 hi = (u32)__builtin_amdgcn_bitop3_b32((u32)(x >> 32), (u32)(y >> 32), m, 0x78);
  return ((u64)hi << 32) | lo;
}
// -(bit b of x) as a lane mask
__device__ __forceinline__ u32 bitmask(u64 x, int b) { return (u32) - (int)((x >> b) & 1ull); }
__device__ __forceinline__ int ffs64(u64 x) { return __builtin_ctzll(x); }

// word of variable c in a row of W words (always word 0 when W = 1, which
// the compiler cannot infer from c < 64 by itself)
template <int W>
__device__ __forceinline__ int wordof(int c) { return W == 1 ? 0 : (c >> 6); }

// Element w of a small register array.  For W >= 3 a chain of selects over
// the elements is folded by the compiler into a load at a run-time index,
// which sends the whole array (and the CH-form holding it) to scratch
// memory; masks keep every element in registers.
template <int W>
__device__ __forceinline__ u64 wsel(const u64 (&x)[W], int w) {
  if constexpr (W >= 3) {
    u64 r = 0ull;
#pragma unroll
    for (int k = 0; k < W; ++k) r |= x[k] & (0ull - (u64)(w == k));
    return r;
  } else {
    u64 r = x[0];
#pragma unroll
    for (int k = 1; k < W; ++k)
      if (w == k) r = x[k];
    return r;
  }
}
template <int W>
__device__ __forceinline__ int getbit(const u64 (&x)[W], int c) {
  return (int)((wsel<W>(x, wordof<W>(c)) >> (c & 63)) & 1ull);
}
// The word updates below touch every word with a mask that is zero except
// in word c/64: a conditional store x[k] would be folded by the compiler into
// a store at the run-time index c/64, which sends the array to scratch.
template <int W>
__device__ __forceinline__ void flipbit(u64 (&x)[W], int c, int b) {
  const u64 m = ((u64)(b & 1)) << (c & 63);
#pragma unroll
  for (int k = 0; k < W; ++k) x[k] ^= (wordof<W>(c) == k) ? m : 0ull;
}
template <int W>
__device__ __forceinline__ void setbitv(u64 (&x)[W], int c, int b) {
  const u64 clr = ~(1ull << (c & 63)), set = ((u64)(b & 1)) << (c & 63);
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const bool here = wordof<W>(c) == k;
    x[k] = (x[k] & (here ? clr : ~0ull)) | (here ? set : 0ull);
  }
}
template <int W>
__device__ __forceinline__ bool anyv(const u64 (&x)[W]) {
  u64 a = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) a |= x[k];
  return a != 0;
}
template <int W>
__device__ __forceinline__ int ffsv(const u64 (&x)[W]) {  // x != 0
  int r = -1;
#pragma unroll
  for (int k = W - 1; k >= 0; --k)
    if (x[k]) r = 64 * k + ffs64(x[k]);
  return r;
}
template <int W>
__device__ __forceinline__ int parand(const u64 (&a)[W], const u64 (&b)[W]) {
  int p = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) p += __builtin_popcountll(a[k] & b[k]);
  return p & 1;
}

template <int W>
struct Form {
  u64 G[W][W], F[W][W], M[W][W];  // [slot][word]
  int g[W];                       // gamma / 2 (mod 4) of the row in each slot
  u64 v[W], s[W];                 // uniform
  int ph;                         // phase_factor k mod 8, uniform
  int n;                          // uniform
  int lane;
};

// broadcast row r of X (uniform r) to all lanes
template <int W>
__device__ __forceinline__ void bcast(const u64 (&X)[W][W], int r, u64 (&out)[W]) {
  const int sl = wordof<W>(r), ln = r & 63;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    u64 val;
    if constexpr (W >= 3) {  // masks, as in wsel
      val = 0ull;
#pragma unroll
      for (int k = 0; k < W; ++k) val |= X[k][w] & (0ull - (u64)(sl == k));
    } else {
      val = X[0][w];
#pragma unroll
      for (int k = 1; k < W; ++k)
        if (sl == k) val = X[k][w];
    }
    out[w] = rl64(val, ln);
  }
}
template <int W>
__device__ __forceinline__ int bcast_g(const Form<W>& f, int r) {
  const int sl = wordof<W>(r);
  int val;
  if constexpr (W >= 3) {  // masks, as in wsel
    val = 0;
#pragma unroll
    for (int k = 0; k < W; ++k) val |= f.g[k] & -(int)(sl == k);
  } else {
    val = f.g[0];
#pragma unroll
    for (int k = 1; k < W; ++k)
      if (sl == k) val = f.g[k];
  }
  return rl32(val, r & 63);
}
template <int W>
__device__ __forceinline__ bool owns(const Form<W>& f, int r, int sl) {
  return f.lane == (r & 63) && sl == wordof<W>(r);
}

// ---- left multiplications (row operations) -------------------------------

// left_multiplication/cx.rs:13-51
template <int W>
__device__ __forceinline__ void left_cx(Form<W>& f, int c, int t) {
  u64 Gc[W], Ft[W], Mt[W];
  bcast<W>(f.G, c, Gc);
  bcast<W>(f.F, t, Ft);
  bcast<W>(f.M, t, Mt);
  const int gt = bcast_g<W>(f, t);
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    if (owns<W>(f, c, sl)) {
      int dot = parand<W>(f.M[sl], Ft);
      f.g[sl] = (f.g[sl] + gt + 2 * dot) & 3;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        f.F[sl][w] ^= Ft[w];
        f.M[sl][w] ^= Mt[w];
      }
    }
    if (owns<W>(f, t, sl)) {
#pragma unroll
      for (int w = 0; w < W; ++w) f.G[sl][w] ^= Gc[w];
    }
  }
}

// left_multiplication/cz.rs:12-35
template <int W>
__device__ __forceinline__ void left_cz(Form<W>& f, int a, int b) {
  u64 Ga[W], Gb[W];
  bcast<W>(f.G, a, Ga);
  bcast<W>(f.G, b, Gb);
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    if (owns<W>(f, a, sl)) {
#pragma unroll
      for (int w = 0; w < W; ++w) f.M[sl][w] ^= Gb[w];
    }
    if (owns<W>(f, b, sl)) {
#pragma unroll
      for (int w = 0; w < W; ++w) f.M[sl][w] ^= Ga[w];
    }
  }
}

// left_multiplication/s.rs:13-33 (add = 3 for S: gamma *= -i; 1 for Sdg)
template <int W>
__device__ __forceinline__ void left_s_like(Form<W>& f, int q, int add) {
#pragma unroll
  for (int sl = 0; sl < W; ++sl)
    if (owns<W>(f, q, sl)) {
#pragma unroll
      for (int w = 0; w < W; ++w) f.M[sl][w] ^= f.G[sl][w];
      f.g[sl] = (f.g[sl] + add) & 3;
    }
}

// left_multiplication/z.rs:7-14
template <int W>
__device__ __forceinline__ void left_z(Form<W>& f, int q) {
#pragma unroll
  for (int sl = 0; sl < W; ++sl)
    if (owns<W>(f, q, sl)) f.g[sl] = (f.g[sl] + 2) & 3;
}

// left_multiplication/x.rs:12-48
template <int W>
__device__ __forceinline__ void left_x(Form<W>& f, int q) {
  u64 Fq[W], Mq[W];
  bcast<W>(f.F, q, Fq);
  bcast<W>(f.M, q, Mq);
  const int gq = bcast_g<W>(f, q);
  int beta = 0;
#pragma unroll
  for (int w = 0; w < W; ++w)
    beta += __builtin_popcountll((Mq[w] & ~f.v[w] & f.s[w]) ^ (Fq[w] & f.v[w] & (Mq[w] ^ f.s[w])));
  f.ph = (f.ph + 4 * (beta & 1) + 2 * gq) & 7;
#pragma unroll
  for (int w = 0; w < W; ++w) f.s[w] ^= (Fq[w] & ~f.v[w]) ^ (Mq[w] & f.v[w]);
}

// left_multiplication/y.rs:4-11
template <int W>
__device__ __forceinline__ void left_y(Form<W>& f, int q) {
  left_z<W>(f, q);
  left_x<W>(f, q);
  f.ph = (f.ph + 2) & 7;
}

// ---- right multiplications (column operations on every row) --------------

// right_multiplication/s.rs:8-24
template <int W>
__device__ __forceinline__ void right_s(Form<W>& f, int q) {
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    int fq = getbit<W>(f.F[sl], q);
    flipbit<W>(f.M[sl], q, fq);
    if (fq) f.g[sl] = (f.g[sl] + 3) & 3;
  }
}

// right_multiplication/cx.rs:7-26 (one gate)
template <int W>
__device__ __forceinline__ void right_cx(Form<W>& f, int c, int t) {
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    flipbit<W>(f.G[sl], c, getbit<W>(f.G[sl], t));
    flipbit<W>(f.F[sl], t, getbit<W>(f.F[sl], c));
    flipbit<W>(f.M[sl], c, getbit<W>(f.M[sl], t));
  }
}

// right_multiplication/cz.rs:8-35 for every i in D (D excludes p): the
// sequential loop of resolve_superposition.rs:171-173 / discard.rs:174-178
template <int W>
__device__ __forceinline__ void right_cz_fan(Form<W>& f, int p, const u64 (&D)[W]) {
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    int fp = getbit<W>(f.F[sl], p);
    int par = parand<W>(f.F[sl], D);
    flipbit<W>(f.M[sl], p, par);
    if (fp) {
#pragma unroll
      for (int w = 0; w < W; ++w) f.M[sl][w] ^= D[w];
      if (par) f.g[sl] = (f.g[sl] + 2) & 3;
    }
  }
}

// right_cx(p, i) for every i in D (resolve_superposition.rs:167-170)
template <int W>
__device__ __forceinline__ void right_cx_fan_ctrl(Form<W>& f, int p, const u64 (&D)[W]) {
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    flipbit<W>(f.G[sl], p, parand<W>(f.G[sl], D));
    if (getbit<W>(f.F[sl], p)) {
#pragma unroll
      for (int w = 0; w < W; ++w) f.F[sl][w] ^= D[w];
    }
    flipbit<W>(f.M[sl], p, parand<W>(f.M[sl], D));
  }
}

// right_cx(i, p) for every i in D (resolve_superposition.rs:160-165,
// discard.rs:152-158)
template <int W>
__device__ __forceinline__ void right_cx_fan_targ(Form<W>& f, int p, const u64 (&D)[W]) {
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    if (getbit<W>(f.G[sl], p)) {
#pragma unroll
      for (int w = 0; w < W; ++w) f.G[sl][w] ^= D[w];
    }
    flipbit<W>(f.F[sl], p, parand<W>(f.F[sl], D));
    if (getbit<W>(f.M[sl], p)) {
#pragma unroll
      for (int w = 0; w < W; ++w) f.M[sl][w] ^= D[w];
    }
  }
}

// resolve_superposition.rs:6-113 ; delta in units of i (0:+1 1:+i 2:-1 3:-i)
template <int W>
__device__ __forceinline__ int resolve(Form<W>& f, const u64 (&t)[W], const u64 (&u)[W], int delta) {
  bool same = true;
#pragma unroll
  for (int w = 0; w < W; ++w) same = same && (t[w] == u[w]);
  if (same) {  // handle_same_vecs_case (:115-133)
    if (delta == 2) return DEV_ERR_SAME_VEC_MINUS;
#pragma unroll
    for (int w = 0; w < W; ++w) f.s[w] = t[w];
    if (delta == 1) f.ph = (f.ph + 1) & 7;
    if (delta == 3) f.ph = (f.ph + 7) & 7;
    return DEV_OK;
  }
  u64 d0[W], d1[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    u64 d = t[w] ^ u[w];
    d0[w] = d & ~f.v[w];
    d1[w] = d & f.v[w];
  }
  int p;
  if (anyv<W>(d0)) {
    p = ffsv<W>(d0);
    flipbit<W>(d0, p, 1);
    right_cx_fan_ctrl<W>(f, p, d0);
    if (anyv<W>(d1)) right_cz_fan<W>(f, p, d1);
  } else {
    p = ffsv<W>(d1);
    flipbit<W>(d1, p, 1);
    right_cx_fan_targ<W>(f, p, d1);
  }
  const int tp = getbit<W>(t, p);
  const int vp = getbit<W>(f.v, p);
#pragma unroll
  for (int w = 0; w < W; ++w) f.s[w] = tp ? u[w] : t[w];
  if (tp) {
    if (delta == 0) { setbitv<W>(f.s, p, 0); flipbit<W>(f.v, p, 1); }
    else if (delta == 2) { setbitv<W>(f.s, p, 1); flipbit<W>(f.v, p, 1); f.ph = (f.ph + 4) & 7; }
    else if (delta == 1) {
      if (vp) { f.ph = (f.ph + 1) & 7; setbitv<W>(f.s, p, 0); right_s<W>(f, p); }
      else { setbitv<W>(f.s, p, 1); setbitv<W>(f.v, p, 1); right_s<W>(f, p); f.ph = (f.ph + 2) & 7; }
    } else {
      if (vp) { f.ph = (f.ph + 7) & 7; setbitv<W>(f.s, p, 1); right_s<W>(f, p); }
      else { setbitv<W>(f.s, p, 0); setbitv<W>(f.v, p, 1); right_s<W>(f, p); f.ph = (f.ph + 6) & 7; }
    }
  } else {
    if (delta == 0) { setbitv<W>(f.s, p, 0); flipbit<W>(f.v, p, 1); }
    else if (delta == 2) { setbitv<W>(f.s, p, 1); flipbit<W>(f.v, p, 1); }
    else if (delta == 1) {
      if (vp) { f.ph = (f.ph + 1) & 7; setbitv<W>(f.s, p, 1); right_s<W>(f, p); }
      else { setbitv<W>(f.s, p, 0); setbitv<W>(f.v, p, 1); right_s<W>(f, p); }
    } else {
      if (vp) { f.ph = (f.ph + 7) & 7; setbitv<W>(f.s, p, 0); right_s<W>(f, p); }
      else { setbitv<W>(f.s, p, 1); setbitv<W>(f.v, p, 1); right_s<W>(f, p); }
    }
  }
  return DEV_OK;
}

// left_multiplication/h.rs:13-65
template <int W>
__device__ __forceinline__ int left_h(Form<W>& f, int q) {
  u64 Gq[W], Fq[W], Mq[W], t[W], u[W];
  bcast<W>(f.G, q, Gq);
  bcast<W>(f.F, q, Fq);
  bcast<W>(f.M, q, Mq);
  const int gq = bcast_g<W>(f, q);
  int alpha = 0, beta = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const u64 v = f.v[w], s = f.s[w];
    t[w] = s ^ (Gq[w] & v);
    u[w] = s ^ (Fq[w] & ~v) ^ (Mq[w] & v);
    alpha += __builtin_popcountll(Gq[w] & ~v & s);
    beta += __builtin_popcountll((Mq[w] & ~v & s) ^ (Fq[w] & v & (Mq[w] ^ s)));
  }
  alpha &= 1;
  beta &= 1;
  const int delta = (gq + 2 * (alpha ^ beta)) & 3;
  if (alpha) f.ph = (f.ph + 4) & 7;
  return resolve<W>(f, t, u, delta);
}

// project.rs:33-76 ; returns error, sets *dead (impossible) / *det
template <int W>
__device__ __forceinline__ int project(Form<W>& f, int q, int outcome, int* dead, int* det) {
  u64 Gq[W], gv[W];
  bcast<W>(f.G, q, Gq);
#pragma unroll
  for (int w = 0; w < W; ++w) gv[w] = Gq[w] & f.v[w];
  if (!anyv<W>(gv)) {
    const int value = parand<W>(Gq, f.s);
    if (value != outcome) *dead = 1;
    else *det = 1;
    return DEV_OK;
  }
  u64 t[W], scopy[W];
  int alpha = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    t[w] = gv[w] ^ f.s[w];
    scopy[w] = f.s[w];
    alpha += __builtin_popcountll(Gq[w] & ~f.v[w] & f.s[w]);
  }
  const int delta = ((alpha & 1) ^ outcome) ? 2 : 0;
  *det = 0;
  return resolve<W>(f, scopy, t, delta);
}

// rows of X (per slot) whose bit c is set, as a W-word row mask
template <int W>
__device__ __forceinline__ void column_mask(const u64 (&X)[W][W], int c, u64 (&out)[W]) {
#pragma unroll
  for (int sl = 0; sl < W; ++sl) out[sl] = __ballot(getbit<W>(X[sl], c));
}

template <int W>
__device__ __forceinline__ void remove_bit(u64 (&x)[W], int q) {
  const int wq = q >> 6, bq = q & 63;
  const u64 low = bq ? ((1ull << bq) - 1ull) : 0ull;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const u64 nxt = (w + 1 < W) ? (x[w + 1] & 1ull) : 0ull;
    if (w == wq) x[w] = (x[w] & low) | ((x[w] >> 1) & ~low) | (nxt << 63);
    else if (w > wq) x[w] = (x[w] >> 1) | (nxt << 63);
  }
}

// rows r >= q move down by one (row r <- row r+1)
template <int W>
__device__ __forceinline__ void remove_row(Form<W>& f, u64 (&X)[W][W], int q) {
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const bool mv = (64 * sl + f.lane) >= q;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      u64 x = X[sl][w];
      u32 lo = (u32)__shfl_down((int)(u32)x, 1);
      u32 hi = (u32)__shfl_down((int)(u32)(x >> 32), 1);
      u64 nx = ((u64)hi << 32) | lo;
      u64 wrap = 0;
      if (sl + 1 < W) wrap = rl64(X[(sl + 1 < W) ? sl + 1 : sl][w], 0);
      if (f.lane == 63) nx = wrap;
      if (mv) X[sl][w] = nx;
    }
  }
}

// discard.rs:21-182
template <int W>
__device__ __forceinline__ int discard(Form<W>& f, int q) {
  u64 valid[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    int lo = 64 * w, hi = f.n - lo;
    valid[w] = hi >= 64 ? ~0ull : (hi <= 0 ? 0ull : ((1ull << hi) - 1ull));
  }
  // set_s_v_to_false (:98-132)
  if (getbit<W>(f.v, q) || getbit<W>(f.s, q)) {
    u64 cm[W];
#pragma unroll
    for (int w = 0; w < W; ++w) cm[w] = ~f.v[w] & ~f.s[w] & valid[w];
    flipbit<W>(cm, q, getbit<W>(cm, q));  // clear bit q
    int clean = anyv<W>(cm) ? ffsv<W>(cm) : -1;
    if (clean < 0) {
      u64 cand[W];
      int cnt = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        cand[w] = ~f.v[w] & f.s[w] & valid[w];
        cnt += __builtin_popcountll(cand[w]);
      }
      if (cnt >= 2) {
        int c1 = ffsv<W>(cand);
        flipbit<W>(cand, c1, 1);
        int c2 = ffsv<W>(cand);
        right_cx<W>(f, c1, c2);
        setbitv<W>(f.s, c2, 0);
        clean = c2;
      }
    }
    if (clean < 0) return DEV_ERR_CANNOT_DISCARD;
    if (clean != q) {
      right_cx<W>(f, clean, q);
      right_cx<W>(f, q, clean);
      right_cx<W>(f, clean, q);
      int vq = getbit<W>(f.v, q), vc = getbit<W>(f.v, clean);
      int sq = getbit<W>(f.s, q), sc = getbit<W>(f.s, clean);
      setbitv<W>(f.v, q, vc); setbitv<W>(f.v, clean, vq);
      setbitv<W>(f.s, q, sc); setbitv<W>(f.s, clean, sq);
    }
  }
  // transform_g (:135-161)
  u64 Gq[W];
  bcast<W>(f.G, q, Gq);
  if (!getbit<W>(Gq, q)) {
    u64 cand[W];
#pragma unroll
    for (int w = 0; w < W; ++w) cand[w] = Gq[w] & valid[w];
    if (!anyv<W>(cand)) return DEV_ERR_G_ROW;
    right_cx<W>(f, q, ffsv<W>(cand));
  }
  u64 cm[W];
  column_mask<W>(f.G, q, cm);
  flipbit<W>(cm, q, getbit<W>(cm, q));
  while (anyv<W>(cm)) {
    int i = uni(ffsv<W>(cm));
    flipbit<W>(cm, i, 1);
    left_cx<W>(f, q, i);
  }
  bcast<W>(f.G, q, Gq);
  flipbit<W>(Gq, q, 1);  // D = G[q,:] without q (G[q][q] == 1 here)
  {
    u64 dv[W];
#pragma unroll
    for (int w = 0; w < W; ++w) dv[w] = Gq[w] & f.v[w];
    if (anyv<W>(dv)) return DEV_ERR_G_COL;
  }
  if (anyv<W>(Gq)) right_cx_fan_targ<W>(f, q, Gq);
  // transform_m (:164-181)
  column_mask<W>(f.M, q, cm);
  flipbit<W>(cm, q, getbit<W>(cm, q));
  while (anyv<W>(cm)) {
    int i = uni(ffsv<W>(cm));
    flipbit<W>(cm, i, 1);
    left_cx<W>(f, q, i);
  }
  u64 Mq[W];
  bcast<W>(f.M, q, Mq);
  if (getbit<W>(Mq, q)) {
    left_s_like<W>(f, q, 1);
    bcast<W>(f.M, q, Mq);
  }
  flipbit<W>(Mq, q, getbit<W>(Mq, q));
  if (anyv<W>(Mq)) right_cz_fan<W>(f, q, Mq);
  // remove row and column q (:33-41, :64-96)
  remove_row<W>(f, f.G, q);
  remove_row<W>(f, f.F, q);
  remove_row<W>(f, f.M, q);
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const bool mv = (64 * sl + f.lane) >= q;
    int nx = __shfl_down(f.g[sl], 1);
    int wrap = 0;
    if (sl + 1 < W) wrap = rl32(f.g[(sl + 1 < W) ? sl + 1 : sl], 0);
    if (f.lane == 63) nx = wrap;
    if (mv) f.g[sl] = nx;
  }
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    remove_bit<W>(f.G[sl], q);
    remove_bit<W>(f.F[sl], q);
    remove_bit<W>(f.M[sl], q);
  }
  remove_bit<W>(f.v, q);
  remove_bit<W>(f.s, q);
  f.n -= 1;
  return DEV_OK;
}

}  // namespace nx
