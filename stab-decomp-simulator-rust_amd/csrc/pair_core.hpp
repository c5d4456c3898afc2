// pair_core.hpp -- per-term table loads and the GF(2) exponential-sum
// elimination of one (bra, ket) pair, shared by the pair kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "device_form.hpp"
#include "internal.hpp"

namespace nx {

__device__ __forceinline__ size_t rowidx(int64_t t, int R, int r, int W, int w) {
  return ((size_t)t * R + r) * W + w;
}

template <int W>
__device__ __forceinline__ void load_form(Form<W>& f, const FormsDev& in, int64_t t, int lane) {
  const int R = 64 * in.W;
  f.lane = lane;
  f.n = in.n;
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const int r = 64 * sl + lane;
    const bool okr = sl < in.W;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const bool ok = okr && w < in.W;
      f.G[sl][w] = ok ? in.G[rowidx(t, R, r, in.W, w)] : 0ull;
      f.F[sl][w] = ok ? in.F[rowidx(t, R, r, in.W, w)] : 0ull;
      f.M[sl][w] = ok ? in.M[rowidx(t, R, r, in.W, w)] : 0ull;
    }
    f.g[sl] = okr ? (int)in.g[(size_t)t * R + r] : 0;
  }
#pragma unroll
  for (int w = 0; w < W; ++w) {
    f.v[w] = w < in.W ? in.v[(size_t)t * in.W + w] : 0ull;
    f.s[w] = w < in.W ? in.s[(size_t)t * in.W + w] : 0ull;
  }
  f.ph = in.ph[t];
}

template <int W>
__device__ __forceinline__ void store_form(const Form<W>& f, const FormsDev& out, int64_t t) {
  const int R = 64 * out.W;
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    if (sl >= out.W) continue;
    const int r = 64 * sl + f.lane;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (w >= out.W) continue;
      out.G[rowidx(t, R, r, out.W, w)] = f.G[sl][w];
      out.F[rowidx(t, R, r, out.W, w)] = f.F[sl][w];
      out.M[rowidx(t, R, r, out.W, w)] = f.M[sl][w];
    }
    out.g[(size_t)t * R + r] = (uint8_t)f.g[sl];
  }
  if (f.lane < out.W) {
    out.v[(size_t)t * out.W + f.lane] = wsel<W>(f.v, f.lane);
    out.s[(size_t)t * out.W + f.lane] = wsel<W>(f.s, f.lane);
  }
  if (f.lane == 0) out.ph[t] = (uint8_t)f.ph;
}


// ---------------------------------------------------------------------------
// Pair kernel.
template <int W>
struct PairRegs {
  u64 B[W][W];
  int a[W];
  u64 ca[W][W], cb[W][W];
  u64 rhsa[W], rhsb[W];
  u64 rem[W];
  int c0, mag;
  int lane;
  u64 nodiag;  // ~(1 << lane)
};

template <int W>
__device__ __forceinline__ void clearbit(u64 (&x)[W], int pw, int pb) {
  const u64 m = ~(1ull << pb);
#pragma unroll
  for (int w = 0; w < W; ++w) x[w] &= (w == pw) ? m : ~0ull;
}

template <int W>
__device__ __forceinline__ int bcast_int(const int (&x)[W], int r) {
  const int sl = wordof<W>(r);
  int val;
  if constexpr (W >= 3) {  // masks, as in wsel
    val = 0;
#pragma unroll
    for (int k = 0; k < W; ++k) val |= x[k] & -(int)(sl == k);
  } else {
    val = x[0];
#pragma unroll
    for (int k = 1; k < W; ++k)
      if (sl == k) val = x[k];
  }
  return rl32(val, r & 63);
}

// Constraint rows a substitution must update: none (Gauss phase), the bra
// and ket rows (while bra constraints are taken; they come first), or the
// ket rows only (bra rows are never read again once all are taken).
enum : int { CONS_NONE = 0, CONS_AB = 1, CONS_B = 2 };

// substitute x_p = c ^ <r, x> (r excludes p) into the form and into the
// constraint rows that CONS names.  Rows of eliminated variables and
// padding rows may hold junk: they are never broadcast again, and every
// column of an eliminated variable is cleared in all rows.
template <int W, int CONS>
__device__ __forceinline__ void subst(PairRegs<W>& P, int p, const u64 (&r)[W], int c) {
  const int ap = bcast_int<W>(P.a, p);
  u64 Bp[W];
  bcast<W>(P.B, p, Bp);
  const int t2 = 2 * ap * c;
  P.c0 = (P.c0 + t2) & 7;
  // a_l += ap (1 - 2c) r_l + 2 (c + r_l) B_pl ;  B_l ^= r_l ((ap odd) r ^ Bp) ^ B_pl r
  const int lin = ap - t2;  // ap (1 - 2c); only its value mod 4 is used
  const int pw = wordof<W>(p), pb = p & 63;
  const u32 cm = c ? ~0u : 0u;
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    // lane masks: mr where r_l, mb where B_pl
    const u32 mr = bitmask(r[sl], P.lane);
    const u32 mb = bitmask(wsel<W>(P.B[sl], pw), pb);
    // (c ^ r_l) & B_pl ? 2 : 0
    const u32 two = __builtin_amdgcn_bitop3_b32(mr ^ cm, mb, 2u, 0x80);
    P.a[sl] = (int)(((u32)P.a[sl] + ((u32)lin & mr) + two) & 3u);
    // B_l ^= (t & mr) ^ (r & mb), t = ((ap odd) r) ^ B_p; then clear this
    // lane's diagonal bit and column p
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const u64 t = ((ap & 1) ? r[w] : 0ull) ^ Bp[w];
      const u64 keep = ((w == sl) ? P.nodiag : ~0ull) & ((w == pw) ? ~(1ull << pb) : ~0ull);
      P.B[sl][w] = xor_and(xor_and(P.B[sl][w], t, mr), r[w], mb) & keep;
    }
    if (CONS == CONS_AB) {
      const u32 ma = bitmask(wsel<W>(P.ca[sl], pw), pb);
#pragma unroll
      for (int w = 0; w < W; ++w) P.ca[sl][w] = xor_and(P.ca[sl][w], r[w], ma) & ((w == pw) ? ~(1ull << pb) : ~0ull);
      P.rhsa[sl] ^= __ballot((ma & cm) != 0u);
    }
    if (CONS != CONS_NONE) {
      const u32 mb2 = bitmask(wsel<W>(P.cb[sl], pw), pb);
#pragma unroll
      for (int w = 0; w < W; ++w) P.cb[sl][w] = xor_and(P.cb[sl][w], r[w], mb2) & ((w == pw) ? ~(1ull << pb) : ~0ull);
      P.rhsb[sl] ^= __ballot((mb2 & cm) != 0u);
    }
  }
  clearbit<W>(P.rem, pw, pb);
}

// returns 1 if a constraint is inconsistent (the sum is zero).  RREF: the
// bra rows are in reduced echelon form on their lowest variables (as the
// quadratic-form kernel writes them): no bra row holds another's lowest
// variable, so a bra substitution leaves the other bra rows as they are and
// none of them becomes empty.
template <int W, bool USEB, bool RREF = false>
__device__ __forceinline__ int take_constraint(PairRegs<W>& P, int c) {
  u64 row[W];
  bcast<W>(USEB ? P.cb : P.ca, c, row);
  const int rhs = USEB ? getbit<W>(P.rhsb, c) : getbit<W>(P.rhsa, c);
  if (!USEB && RREF) {
    const int p = ffsv<W>(row);
    flipbit<W>(row, p, 1);
    subst<W, CONS_B>(P, p, row, rhs);
    return 0;
  }
  if (!anyv<W>(row)) return rhs;
  const int p = ffsv<W>(row);
  flipbit<W>(row, p, 1);
  subst<W, USEB ? CONS_B : CONS_AB>(P, p, row, rhs);
  return 0;
}

// keeps the wave-uniform elimination state in scalar registers (the
// compiler otherwise carries some of it in vector registers through the
// loops below)
template <int W>
__device__ __forceinline__ void uniform_state(PairRegs<W>& P) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    P.rem[w] = uni64(P.rem[w]);
    P.rhsa[w] = uni64(P.rhsa[w]);
    P.rhsb[w] = uni64(P.rhsb[w]);
  }
  P.c0 = uni(P.c0);
  P.mag = uni(P.mag);
}

// exact <bra|ket>: returns zero flag, K (mod 8), r.  RREF: the bra rows
// are in reduced echelon form (take_constraint).
template <int W, bool RREF = false>
__device__ __forceinline__ void pair_eval(PairRegs<W>& P, int nca, int ncb, int* zero, int* K, int* R) {
  uniform_state<W>(P);
  for (int c = 0; c < nca; ++c) {
    if (RREF) {
      take_constraint<W, false, true>(P, c);
    } else if (uni(take_constraint<W, false>(P, c))) {
      *zero = 1;
      return;
    }
    uniform_state<W>(P);
  }
  for (int c = 0; c < ncb; ++c) {
    if (uni(take_constraint<W, true>(P, c))) { *zero = 1; return; }
    uniform_state<W>(P);
  }
  // One exit: the loop runs while coupled variables remain and no factor
  // has vanished (a structured loop keeps the scalar control flow short).
  int zf = 0;
  for (;;) {
    // Variables with no coupling left are final: in this phase a variable's
    // coefficient and row change only through its coupling to the pivot, so
    // each sums out alone to 2 (a = 0), sqrt2 zeta^{+-1} (a odd) or 0 (a = 2),
    // all of them at once from ballots of their coefficients.
    u64 nz[W];
    bool more = false, iso_any = false;
#pragma unroll
    for (int sl = 0; sl < W; ++sl) {
      bool rownz = false;
#pragma unroll
      for (int w = 0; w < W; ++w) rownz = rownz || (P.B[sl][w] != 0ull);
      nz[sl] = uni64(__ballot(rownz)) & P.rem[sl];
      more = more || (nz[sl] != 0ull);
      iso_any = iso_any || (nz[sl] != P.rem[sl]);
    }
    if (iso_any) {
      int n0 = 0, n1 = 0, n3 = 0;
#pragma unroll
      for (int sl = 0; sl < W; ++sl) {
        const u64 iso = P.rem[sl] & ~nz[sl];
        zf |= (uni64(__ballot(P.a[sl] == 2)) & iso) != 0ull;
        n0 += __builtin_popcountll(uni64(__ballot(P.a[sl] == 0)) & iso);
        n1 += __builtin_popcountll(uni64(__ballot(P.a[sl] == 1)) & iso);
        n3 += __builtin_popcountll(uni64(__ballot(P.a[sl] == 3)) & iso);
        P.rem[sl] = nz[sl];
      }
      P.mag -= 2 * n0 + n1 + n3;
      P.c0 = (P.c0 + n1 + 7 * n3) & 7;
    }
    if (zf || !more) break;
    const int p = uni(ffsv<W>(nz));
    const int pw = wordof<W>(p), pb = p & 63;
    clearbit<W>(P.rem, pw, pb);
    u64 Bp[W];
    bcast<W>(P.B, p, Bp);
    const int ap = bcast_int<W>(P.a, p);
    if (ap & 1) {
      // sum over x_p: sqrt2 zeta^s i^{-s lambda}, lambda = <B_p, x>, s = +-1
      if constexpr (W == 1) {
        // Further odd variables q with no coupling to p nor to each other are
        // summed out in the same step: no such sum changes another's row or
        // coefficient (B_pq = 0), so the updates commute, every row and
        // coefficient they read is the one before the step, and they apply at
        // once (up to eight variables per step, the set S).  Per variable q:
        // mag -1, c0 + 1 (a_q = 1) or 7 (a_q = 3); per row l coupled to q:
        // a_l + 3 (a_q = 1) or 1 (a_q = 3), B_l ^= B_q.
        u64 cand = uni64(__ballot(P.a[0] & 1)) & nz[0] & ~Bp[0] & ~(1ull << pb);
        if (cand) {
          const u64 B0 = P.B[0][0];
          const int a0 = P.a[0];
          u64 Bn = xor_and(B0, Bp[0], bitmask(B0, pb)), S = 1ull << pb;
#pragma unroll
          for (int k = 0; k < 7; ++k) {
            const int q = ffs64(cand);
            const u64 Bq = rl64(B0, q);
            Bn = xor_and(Bn, Bq, bitmask(B0, q));
            const u64 bq = 1ull << q;
            S |= bq;
            cand &= ~(Bq | bq);
            if (!cand) break;
          }
          const u64 S1 = S & uni64(__ballot(a0 == 1));
          const int ns = __builtin_popcountll(S), n1 = __builtin_popcountll(S1);
          P.mag -= ns;
          P.c0 = (P.c0 + 7 * ns - 6 * n1) & 7;
          const u64 nb = B0 & S;
          P.a[0] = (a0 + 2 * __builtin_popcountll(nb & S1) + __builtin_popcountll(nb)) & 3;
          P.rem[0] &= ~S;
          P.B[0][0] = Bn & P.nodiag & ~S;
          uniform_state<W>(P);
          continue;
        }
      }
      const int add = (ap == 1) ? 3 : 1;
      P.mag -= 1;
      P.c0 = (P.c0 + (ap == 1 ? 1 : 7)) & 7;
#pragma unroll
      for (int sl = 0; sl < W; ++sl) {
        // mb = all ones where B_pl (bit p of this lane's row)
        const u32 mb = bitmask(wsel<W>(P.B[sl], pw), pb);
        P.a[sl] = (P.a[sl] + (int)((u32)add & mb)) & 3;
        // B_l = (B_l ^ (B_p & mb)) & ~e_l & ~e_p
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const u64 keep = ((w == sl) ? P.nodiag : ~0ull) & ((w == pw) ? ~(1ull << pb) : ~0ull);
          P.B[sl][w] = xor_and(P.B[sl][w], Bp[w], mb) & keep;
        }
      }
      uniform_state<W>(P);
    } else {
      // sum over x_p: 2 [lambda = a_p / 2]; then substitute for x_q
      P.mag -= 2;
#pragma unroll
      for (int sl = 0; sl < W; ++sl) clearbit<W>(P.B[sl], pw, pb);
      const int qv = ffsv<W>(Bp);
      flipbit<W>(Bp, qv, 1);
      subst<W, CONS_NONE>(P, qv, Bp, ap >> 1);
      uniform_state<W>(P);
    }
  }
  *zero = zf;
  *K = P.c0;
  *R = P.mag;
}

// Gauss-Jordan on a term's nc constraint rows (row c in lane c & 63 of
// slot c >> 6), each nonzero row solved for its lowest variable: afterwards
// that variable appears in no other row and stays the row's lowest (a row
// only ever gains variables above a later pivot, which exceeds its own);
// empty rows (dependent rows of a projected form) stay empty.  The steps
// depend on C alone, so every kernel that starts from the same rows takes
// the same steps; also(c, hit) applies step c (row l ^= row c for the lanes
// with hit[slot of l]) to what the caller keeps per row.
template <int W, typename Also>
__device__ __forceinline__ void rref_rows(u64 (&C)[W][W], int nc, int lane, Also&& also) {
  for (int c = 0; c < nc; ++c) {
    u64 row[W];
    bcast<W>(C, c, row);
    if (!anyv<W>(row)) continue;
    const int p = ffsv<W>(row);
    const int pw = wordof<W>(p), pb = p & 63;
    bool hit[W];
#pragma unroll
    for (int sl = 0; sl < W; ++sl) {
      hit[sl] = (64 * sl + lane != c) && ((wsel<W>(C[sl], pw) >> pb) & 1ull);
#pragma unroll
      for (int w = 0; w < W; ++w) C[sl][w] ^= hit[sl] ? row[w] : 0ull;
    }
    also(c, hit);
  }
}

// rows of term t (word offsets fit 32 bits: chi * 64W * W < 2^32 is checked
// where the forms are allocated); RHS: also the constraint right-hand sides
template <int W, bool RHS = true>
__device__ __forceinline__ void load_qrows(const QFormDev& q, int64_t t, int lane, u64 (&B)[W][W], int (&a)[W],
                                           u64 (&C)[W][W], u64 (&rhs)[W]) {
  const uint32_t R = 64 * W;
  const uint32_t tw = (uint32_t)t * (R * W);
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const uint32_t r = 64 * sl + lane;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      B[sl][w] = q.B[tw + r * W + w];
      C[sl][w] = q.C[tw + r * W + w];
    }
    a[sl] = q.a[(uint32_t)t * R + r];
  }
  if constexpr (RHS) {
#pragma unroll
    for (int w = 0; w < W; ++w) rhs[w] = uni64(q.rhs[(uint32_t)t * W + w]);
  }
}

}  // namespace nx
