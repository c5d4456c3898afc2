// variants.hip -- pair sums over many ket variants at once (gfx950, wave64).
//
//  k_pauli_var  one wavefront per term: the bit-sliced differences between
//               the quadratic form of P_k phi_t and that of phi_t for up to
//               64*VW Pauli strings P_k
//  k_pairs_var  one wavefront per (bra, ket) pair, as k_pairs, but the
//               elimination carries every ket variant along: B, the pivot
//               order, the parity of each linear coefficient and the
//               constraint rows are the same for all variants (a variant
//               only adds even amounts to the linear coefficients and flips
//               constraint right-hand sides), so the GF(2) work is done
//               once and the variant-dependent quantities (high bit of each
//               linear coefficient, right-hand sides, Z8 phase, zero flag)
//               are u64 masks with one bit per variant.  Lane k then turns
//               variant k's exact value into FP64 and accumulates it.
//  k_reduce_var per job and variant, the ordered sum of the wave partials
//  k_grad_var   the coefficient backward: weighted pair values reduced over
//               the strings and accumulated into y = G c per tile of pairs
//               (k_grad_reduce, k_grad_vals, k_grad_loss finish it)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_form.hpp"
#include "internal.hpp"
#include "pair_core.hpp"

namespace nx {

// b += x per variant, b and x in Z8 as three bit planes
__device__ __forceinline__ void z8_add(u64& b0, u64& b1, u64& b2, u64 x0, u64 x1, u64 x2) {
  const u64 c0 = b0 & x0;
  b0 ^= x0;
  const u64 t = b1 ^ x1;
  const u64 c1 = (b1 & x1) | (c0 & t);
  b1 = t ^ c0;
  b2 ^= x2 ^ c1;
}

// Exact real part of w * zeta^k * 2^{-s/2} (w = 1 or 2): cos(k pi/4) is 0,
// +-1 or +-2^{-1/2}, so the value is +-w 2^{-t/2} with t = s or s + 1; it
// is added to A when t is even and as +-w 2^{-(t+1)/2} to B when t is odd,
// so that the sum is A + sqrt(2) B.  Each addend is a power of two, so A and
// B are exact while their exponent span fits a double's 53 bits.
// Written without branches (lanes differ in k and s): the addend not taken
// is +0.0, and x + 0.0 == x for every x the sums reach (they start at +0.0
// and no addend is a zero, so neither becomes -0.0).
__device__ __forceinline__ void add_exact(int k, int s, double w, double& A, double& B) {
  k &= 7;
  const double sg = (k == 2 || k == 6) ? 0.0 : ((k == 0 || k == 1 || k == 7) ? w : -w);
  const int t = (k & 1) ? s + 1 : s;
  const double val = ldexp(sg, -((t + 1) >> 1));  // (t + 1) >> 1 == t / 2 for even t
  A = A + ((t & 1) ? 0.0 : val);
  B = B + ((t & 1) ? val : 0.0);
}

// OR of x over the 64 lanes (all active), as a wave-uniform value.  DPP
// steps, no LDS: quad permutations [1,0,3,2] and [2,3,0,1], the half-row and
// row mirrors leave every lane of a 16-lane row with the row's OR; lane 15's
// value broadcast into rows 1 and 3, then lane 31's into rows 2 and 3 (the
// rows not written receive 0) leave the total in lane 63.
__device__ __forceinline__ u32 wave_or32(u32 x) {
  x |= (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
  x |= (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
  x |= (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false);  // row_half_mirror
  x |= (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false);  // row_mirror
  x |= (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15 -> rows 1, 3
  x |= (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31 -> rows 2, 3
  return (u32)__builtin_amdgcn_readlane((int)x, 63);
}
__device__ __forceinline__ u64 wave_or64(u64 x) {
  return ((u64)wave_or32((u32)(x >> 32)) << 32) | wave_or32((u32)x);
}

template <int W, int VW>
struct VarRegs {
  u64 hi[W][VW];               // row r: variants whose a_r carries an extra 2
  u64 ra[W][VW];               // bra constraint row c: variants whose rhs is flipped (unit mode)
  u64 rb[W][VW];               // ket constraint row c: variants whose rhs is flipped
  u64 d0[VW], d1[VW], d2[VW];  // uniform: Z8 phase offset of each variant, plus dv
  int dv;                      // uniform: Z8 offset shared by all variants (kept out of the planes)
  u64 z[VW];                   // uniform: variants whose sum is already zero
  int bz;                      // uniform: the base (identity) sum is zero or not wanted
};

template <int W, int VW>
__device__ __forceinline__ void bcast_v(const u64 (&X)[W][VW], int r, u64 (&out)[VW]) {
  const int sl = wordof<W>(r), ln = r & 63;
#pragma unroll
  for (int v = 0; v < VW; ++v) {
    u64 val;
    if constexpr (W >= 3) {  // masks, as in wsel
      val = 0ull;
#pragma unroll
      for (int k = 0; k < W; ++k) val |= X[k][v] & (0ull - (u64)(sl == k));
    } else {
      val = X[0][v];
#pragma unroll
      for (int k = 1; k < W; ++k)
        if (sl == k) val = X[k][v];
    }
    out[v] = rl64(val, ln);
  }
}

template <int W, int VW>
__device__ __forceinline__ bool all_zero(const VarRegs<W, VW>& V) {
  u64 a = V.bz ? ~0ull : 0ull;
#pragma unroll
  for (int v = 0; v < VW; ++v) a &= V.z[v];
  return uni(a == ~0ull) != 0;
}

// keeps the wave-uniform elimination state in scalar registers
template <int W, int VW>
__device__ __forceinline__ void uniform_state(PairRegs<W>& P, VarRegs<W, VW>& V) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    P.rem[w] = uni64(P.rem[w]);
    P.rhsa[w] = uni64(P.rhsa[w]);
    P.rhsb[w] = uni64(P.rhsb[w]);
  }
  P.c0 = uni(P.c0);
  P.mag = uni(P.mag);
  V.bz = uni(V.bz);
  V.dv = uni(V.dv);
#pragma unroll
  for (int v = 0; v < VW; ++v) {
    V.z[v] = uni64(V.z[v]);
    V.d0[v] = uni64(V.d0[v]);
    V.d1[v] = uni64(V.d1[v]);
    V.d2[v] = uni64(V.d2[v]);
  }
}

// subst<W, CONS> of pair_core.hpp for every variant: x_p = c_k ^ <r, x>
// with c_k = c ^ bit k of E.  Base quantities (a, B, constraint rows, base
// rhs) are updated as in subst with c; the variant masks pick up the
// differences: a_l changes by r_l (c_k ? -a_p,k : a_p,k) + 2 ((c_k ^ r_l) &
// B_pl), which differs from the base change by 2 (r_l (h_p,k ^ (a_p odd &
// e_k)) ^ (B_pl & e_k)); the phase gains 2 (a_p,k c_k mod 4) per variant and
// 2 (a_p c mod 4) for the base.  Bra constraints carry no variant flips
// (E = 0 while they are taken).  PH = false skips the phase planes and dv.
template <int W, int VW, int CONS, bool UNITS, bool PH = true>
__device__ __forceinline__ void subst_v(PairRegs<W>& P, VarRegs<W, VW>& V, int p, const u64 (&r)[W], int c,
                                        const u64 (&E)[VW]) {
  const int ap = bcast_int<W>(P.a, p);
  u64 H[VW];
  bcast_v<W, VW>(V.hi, p, H);
  u64 Bp[W];
  bcast<W>(P.B, p, Bp);
  const int t2 = 2 * ap * c;
  P.c0 = (P.c0 + t2) & 7;
  u64 L[VW];
#pragma unroll
  for (int v = 0; v < VW; ++v) {
    if (PH) {
      // variant phase minus base phase, 2 (a_p,k c_k - a_p c) mod 8, as the
      // full variant term (bit-sliced) plus 8 - 2 a_p c (in dv, below)
      const u64 cm = c ? ~E[v] : E[v];                  // variants with c_k = 1
      const u64 lo = (ap & 1) ? cm : 0ull;              // (a_p,k c_k) mod 2
      const u64 hi = cm & ((ap & 2) ? ~H[v] : H[v]);    // high bit of a_p,k c_k
      z8_add(V.d0[v], V.d1[v], V.d2[v], 0ull, lo, hi);
    }
    L[v] = H[v] ^ ((ap & 1) ? E[v] : 0ull);
  }
  if (PH) V.dv = (V.dv + 8 - t2) & 7;
  const int lin = ap - t2;  // ap (1 - 2c); only its value mod 4 is used
  const int pw = wordof<W>(p), pb = p & 63;
  const u32 cm = c ? ~0u : 0u;
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    // lane masks: mr where r_l, mb where B_pl
    const u32 mr = bitmask(r[sl], P.lane);
    const u32 mb = bitmask(wsel<W>(P.B[sl], pw), pb);
    const u32 two = __builtin_amdgcn_bitop3_b32(mr ^ cm, mb, 2u, 0x80);  // (c ^ r_l) & B_pl ? 2 : 0
    P.a[sl] = (int)(((u32)P.a[sl] + ((u32)lin & mr) + two) & 3u);
#pragma unroll
    for (int v = 0; v < VW; ++v) V.hi[sl][v] = xor_and(xor_and(V.hi[sl][v], L[v], mr), E[v], mb);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const u64 t = ((ap & 1) ? r[w] : 0ull) ^ Bp[w];
      const u64 keep = ((w == sl) ? P.nodiag : ~0ull) & ((w == pw) ? ~(1ull << pb) : ~0ull);
      P.B[sl][w] = xor_and(xor_and(P.B[sl][w], t, mr), r[w], mb) & keep;
    }
    if (CONS == CONS_AB) {
      const int ha = (int)((wsel<W>(P.ca[sl], pw) >> pb) & 1ull);
      const u32 ma = (u32)-ha;
#pragma unroll
      for (int w = 0; w < W; ++w) P.ca[sl][w] = xor_and(P.ca[sl][w], r[w], ma) & ((w == pw) ? ~(1ull << pb) : ~0ull);
      P.rhsa[sl] ^= __ballot(ha & c);
      if (UNITS) {
#pragma unroll
        for (int v = 0; v < VW; ++v) V.ra[sl][v] = xor_and(V.ra[sl][v], E[v], ma);
      }
    }
    if (CONS != CONS_NONE) {
      const int hb = (int)((wsel<W>(P.cb[sl], pw) >> pb) & 1ull);
      const u32 mb2 = (u32)-hb;
#pragma unroll
      for (int w = 0; w < W; ++w) P.cb[sl][w] = xor_and(P.cb[sl][w], r[w], mb2) & ((w == pw) ? ~(1ull << pb) : ~0ull);
      P.rhsb[sl] ^= __ballot(hb & c);
      if (CONS == CONS_B || UNITS) {
#pragma unroll
        for (int v = 0; v < VW; ++v) V.rb[sl][v] = xor_and(V.rb[sl][v], E[v], mb2);
      }
    }
  }
  clearbit<W>(P.rem, pw, pb);
}

template <int W, int VW, bool USEB, bool UNITS>
__device__ __forceinline__ void take_constraint_v(PairRegs<W>& P, VarRegs<W, VW>& V, int c) {
  u64 row[W];
  bcast<W>(USEB ? P.cb : P.ca, c, row);
  const int rhs = USEB ? getbit<W>(P.rhsb, c) : getbit<W>(P.rhsa, c);
  u64 E[VW];
  if (USEB) {
    bcast_v<W, VW>(V.rb, c, E);
  } else if (UNITS) {
    bcast_v<W, VW>(V.ra, c, E);
  } else {
#pragma unroll
    for (int v = 0; v < VW; ++v) E[v] = 0ull;
  }
  if (!USEB) {
    // the bra rows are in reduced echelon form on their lowest variables
    // (as k_qform and k_qproject_var store them): no bra row holds another's
    // lowest variable, so a bra substitution leaves the other bra rows (and
    // their rhs flips) as they are.  Only projected rows can be empty.
    if (UNITS && !anyv<W>(row)) {
      // 0 = rhs_k: the sum vanishes for the variants with rhs_k = 1
      if (rhs) V.bz = 1;
#pragma unroll
      for (int v = 0; v < VW; ++v) V.z[v] |= rhs ? ~E[v] : E[v];
      return;
    }
    const int p = ffsv<W>(row);
    flipbit<W>(row, p, 1);
    subst_v<W, VW, CONS_B, UNITS>(P, V, p, row, rhs, E);
    return;
  }
  if (!anyv<W>(row)) {
    // 0 = rhs_k: the sum vanishes for the variants with rhs_k = 1
    if (rhs) V.bz = 1;
#pragma unroll
    for (int v = 0; v < VW; ++v) V.z[v] |= rhs ? ~E[v] : E[v];
    return;
  }
  const int p = ffsv<W>(row);
  flipbit<W>(row, p, 1);
  subst_v<W, VW, CONS_B, UNITS>(P, V, p, row, rhs, E);
}

// pair_eval of pair_core.hpp for every variant and for the base (the ket
// itself).  P.c0 follows the base; the planes hold each variant's offset.
// With UNITS both terms are forms projected by k_qproject_var onto the
// window outcomes (one variant per outcome, the unit constraints already
// substituted into each term), and the bra rows carry rhs flips too.
template <int W, int VW, bool UNITS>
__device__ __forceinline__ void pair_eval_v(PairRegs<W>& P, VarRegs<W, VW>& V, int nca, int ncb) {
  uniform_state<W, VW>(P, V);
  for (int c = 0; c < nca; ++c) {
    take_constraint_v<W, VW, false, UNITS>(P, V, c);
    uniform_state<W, VW>(P, V);
    if (all_zero<W, VW>(V)) return;
  }
  for (int c = 0; c < ncb; ++c) {
    take_constraint_v<W, VW, true, UNITS>(P, V, c);
    uniform_state<W, VW>(P, V);
    if (all_zero<W, VW>(V)) return;
  }
  uniform_state<W, VW>(P, V);
  while (true) {
    // Variables with no coupling left are final (as in pair_eval): the
    // base's factors come from ballots of the coefficients; each variant's
    // difference needs the variable's high-bit mask H, read per variable.
    u64 nz[W];
    bool more = false, iso_any = false;
#pragma unroll
    for (int sl = 0; sl < W; ++sl) {
      bool rownz = false;
#pragma unroll
      for (int w = 0; w < W; ++w) rownz = rownz || (P.B[sl][w] != 0ull);
      nz[sl] = uni64(__ballot(rownz)) & P.rem[sl];
      more = more || (nz[sl] != 0ull);
      iso_any = iso_any || ((P.rem[sl] & ~nz[sl]) != 0ull);
    }
    if (iso_any) {
      int ne = 0, n1 = 0, n3 = 0;
#pragma unroll
      for (int sl = 0; sl < W; ++sl) {
        const u64 iso = P.rem[sl] & ~nz[sl];
        const u64 m2 = uni64(__ballot(P.a[sl] == 2)) & iso;
        const u64 m1 = uni64(__ballot(P.a[sl] == 1)) & iso;
        const u64 m3 = uni64(__ballot(P.a[sl] == 3)) & iso;
        const u64 me = iso & ~(m1 | m3);
        if (m2) V.bz = 1;
        ne += __builtin_popcountll(me);
        n1 += __builtin_popcountll(m1);
        n3 += __builtin_popcountll(m3);
        // a_p,k = 2: 1 + i^2 = 0; a_p,k = 0: factor 2.  Each such lane holds
        // its own row's mask: one OR over the wave covers all of them.
        if (me) {
          const bool in_e = (me >> P.lane) & 1ull, two = (m2 >> P.lane) & 1ull;
#pragma unroll
          for (int v = 0; v < VW; ++v)
            V.z[v] |= wave_or64(in_e ? (two ? ~V.hi[sl][v] : V.hi[sl][v]) : 0ull);
        }
        // 1 + i^a: sqrt2 zeta (a = 1) or sqrt2 zeta^7 (a = 3); a variant
        // whose a_p differs from the base by 2 gets the other one: offset
        // 7 - 1 = 6 (base a_p = 1) or 1 - 7 = 2 (base a_p = 3)
        for (u64 m = m1 | m3; m; m &= m - 1) {
          const int q = 64 * sl + __builtin_ctzll(m);
          u64 H[VW];
          bcast_v<W, VW>(V.hi, q, H);
          const bool one = (m1 >> (q & 63)) & 1ull;
#pragma unroll
          for (int v = 0; v < VW; ++v) z8_add(V.d0[v], V.d1[v], V.d2[v], 0ull, H[v], one ? H[v] : 0ull);
        }
        P.rem[sl] = nz[sl];
      }
      P.mag -= 2 * ne + n1 + n3;
      P.c0 = (P.c0 + n1 + 7 * n3) & 7;
      uniform_state<W, VW>(P, V);
      if (all_zero<W, VW>(V)) return;
    }
    if (!more) break;
    const int p = uni(ffsv<W>(nz));
    const int pw = wordof<W>(p), pb = p & 63;
    clearbit<W>(P.rem, pw, pb);
    u64 Bp[W];
    bcast<W>(P.B, p, Bp);
    const int ap = bcast_int<W>(P.a, p);
    u64 H[VW];
    bcast_v<W, VW>(V.hi, p, H);
    if (ap & 1) {
      const int add = (ap == 1) ? 3 : 1;
      P.mag -= 1;
      P.c0 = (P.c0 + (ap == 1 ? 1 : 7)) & 7;
#pragma unroll
      for (int v = 0; v < VW; ++v) z8_add(V.d0[v], V.d1[v], V.d2[v], 0ull, H[v], ap == 1 ? H[v] : 0ull);
#pragma unroll
      for (int sl = 0; sl < W; ++sl) {
        const u32 mb = bitmask(wsel<W>(P.B[sl], pw), pb);
        P.a[sl] = (P.a[sl] + (int)((u32)add & mb)) & 3;
#pragma unroll
        for (int v = 0; v < VW; ++v) V.hi[sl][v] = xor_and(V.hi[sl][v], H[v], mb);
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const u64 keep = ((w == sl) ? P.nodiag : ~0ull) & ((w == pw) ? ~(1ull << pb) : ~0ull);
          P.B[sl][w] = xor_and(P.B[sl][w], Bp[w], mb) & keep;
        }
      }
      uniform_state<W, VW>(P, V);
    } else {
      P.mag -= 2;
#pragma unroll
      for (int sl = 0; sl < W; ++sl) clearbit<W>(P.B[sl], pw, pb);
      const int qv = ffsv<W>(Bp);
      flipbit<W>(Bp, qv, 1);
      subst_v<W, VW, CONS_NONE, UNITS>(P, V, qv, Bp, ap >> 1, H);
      uniform_state<W, VW>(P, V);
    }
  }
}

// ---------------------------------------------------------------------------
// Projection of every term's quadratic form onto x_{uq[u]} = ub[u], u <
// nunit in order: the substitutions pair_eval_v makes for unit constraints,
// applied to one form at a time.  A unit substitution x_p = c is affine in
// the form: a_l gains 2 c B_pl, the phase 2 a_p c, and every constraint
// holding x_p loses it and flips its right-hand side by c.  A pair of
// projected forms (with the substituted variables removed from the
// elimination) therefore carries what the pair of unprojected forms carries
// after the same substitutions, and the prefix of a sampling window costs
// chi instead of chi^2 substitutions.  Rows of substituted variables keep
// junk that is never read.
template <int W>
__global__ __launch_bounds__(256) void k_qproject(QFormDev in, QFormDev out, const int32_t* __restrict__ uq,
                                                  const int32_t* __restrict__ ub, int nunit) {
  const int64_t t = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  const int lane = threadIdx.x & 63;
  if (t >= in.chi) return;
  const int R = 64 * W;
  PairRegs<W> P;
  P.lane = lane;
  P.nodiag = ~(1ull << lane);
  load_qrows<W>(in, t, lane, P.B, P.a, P.ca, P.rhsa);
#pragma unroll
  for (int sl = 0; sl < W; ++sl)
#pragma unroll
    for (int w = 0; w < W; ++w) P.cb[sl][w] = 0ull;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    P.rhsb[w] = 0ull;
    P.rem[w] = 0ull;
  }
  P.c0 = uni(in.p0[t]);
  P.mag = 0;
  u64 r0[W];
#pragma unroll
  for (int w = 0; w < W; ++w) r0[w] = 0ull;
  for (int u = 0; u < nunit; ++u) {
    subst<W, CONS_AB>(P, uni(uq[u]), r0, uni(ub[u]));
    uniform_state<W>(P);
  }
  // back to reduced echelon form (as k_qproject_var_term), the rhs following
  rref_rows<W>(P.ca, uni(in.nc[t]), lane, [&](int c, const bool (&hit)[W]) {
    const bool rc = getbit<W>(P.rhsa, c) != 0;
#pragma unroll
    for (int sl = 0; sl < W; ++sl) {
      const u64 m = __ballot(hit[sl]);
      P.rhsa[sl] = uni64(P.rhsa[sl] ^ (rc ? m : 0ull));
    }
  });
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const int r = 64 * sl + lane;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      out.B[rowidx(t, R, r, W, w)] = P.B[sl][w];
      out.C[rowidx(t, R, r, W, w)] = P.ca[sl][w];
    }
    out.a[(size_t)t * R + r] = (uint8_t)P.a[sl];
  }
  if (lane == 0) {
#pragma unroll
    for (int w = 0; w < W; ++w) out.rhs[(size_t)t * W + w] = P.rhsa[w];
    out.p0[t] = P.c0;
    out.nc[t] = in.nc[t];
    out.k[t] = in.k[t];
  }
}

// ---------------------------------------------------------------------------
// Projection of every term onto the fixed outcomes x_{fq[u]} = fb[u], u <
// nfix, of a sampling path, then onto the look-ahead outcomes of its window,
// one variant per outcome: x_{uq[u]} = ub[u * (VW + 1)] for the base and the
// other value for the variants in the masks that follow, u < nunit in order,
// substituted into the term's own form as in pair_eval_v's unit
// constraints.  The result is the base form (out) and, per term, the
// variants' differences from it (vout): high bits of the linear
// coefficients (hi, per row), right-hand-side flips of the term's support
// rows (rhs, per row) and Z8 phase planes (ph, 3 x VW words).  A
// substitution is linear in the form, so the pair <Pi_k phi_o | Pi_k phi_f>
// of two projected terms is the elimination of the difference of the two,
// with variant differences hi_f ^ hi_o, the ket's and the bra's rhs flips and
// phase d_f - d_o: each pair of the window pass starts after the unit
// substitutions (Pi_k is a projector, so this is <phi_o| Pi_k |phi_f>).
// term t of one state (one wavefront)
template <int W, int VW>
__device__ __forceinline__ void qproject_var_term(const QFormDev& in, const QFormDev& out, const VarDev& vout,
                                                  const int32_t* __restrict__ fq, const int32_t* __restrict__ fb,
                                                  int nfix, const int32_t* __restrict__ uq,
                                                  const uint64_t* __restrict__ ub, int nunit, int64_t t, int lane) {
  const int R = 64 * W;
  PairRegs<W> P;
  VarRegs<W, VW> V;
  P.lane = lane;
  P.nodiag = ~(1ull << lane);
  load_qrows<W>(in, t, lane, P.B, P.a, P.ca, P.rhsa);
#pragma unroll
  for (int sl = 0; sl < W; ++sl)
#pragma unroll
    for (int w = 0; w < W; ++w) P.cb[sl][w] = 0ull;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    P.rhsb[w] = 0ull;
    P.rem[w] = 0ull;
  }
  P.c0 = uni(in.p0[t]);
  P.mag = 0;
#pragma unroll
  for (int sl = 0; sl < W; ++sl)
#pragma unroll
    for (int v = 0; v < VW; ++v) V.hi[sl][v] = V.ra[sl][v] = V.rb[sl][v] = 0ull;
#pragma unroll
  for (int v = 0; v < VW; ++v) V.d0[v] = V.d1[v] = V.d2[v] = V.z[v] = 0ull;
  V.bz = 0;
  V.dv = 0;
  u64 r0[W];
#pragma unroll
  for (int w = 0; w < W; ++w) r0[w] = 0ull;
  // the fixed outcomes first (k_qproject's substitutions: the variant
  // differences are still zero and stay so)
  for (int u = 0; u < nfix; ++u) {
    subst<W, CONS_AB>(P, uni(fq[u]), r0, uni(fb[u]));
    uniform_state<W>(P);
  }
  for (int u = 0; u < nunit; ++u) {
    const int q = uni(uq[u]);
    const int b = (int)uni64(ub[(size_t)u * (VW + 1)]);
    u64 E[VW];
#pragma unroll
    for (int v = 0; v < VW; ++v) E[v] = uni64(ub[(size_t)u * (VW + 1) + 1 + v]);
    subst_v<W, VW, CONS_AB, true>(P, V, q, r0, b, E);
    uniform_state<W, VW>(P, V);
  }
  // back to reduced echelon form: a substitution for x_q takes x_q out of
  // every row, which can empty a row or leave two rows with the same lowest
  // variable.  The base rhs and the variants' flips follow the row
  // operations (the window pass substitutes bra rows without updating the
  // other bra rows).
  rref_rows<W>(P.ca, uni(in.nc[t]), lane, [&](int c, const bool (&hit)[W]) {
    const bool rc = getbit<W>(P.rhsa, c) != 0;
    u64 E[VW];
    bcast_v<W, VW>(V.ra, c, E);
#pragma unroll
    for (int sl = 0; sl < W; ++sl) {
      const u64 m = __ballot(hit[sl]);
      P.rhsa[sl] = uni64(P.rhsa[sl] ^ (rc ? m : 0ull));
#pragma unroll
      for (int v = 0; v < VW; ++v) V.ra[sl][v] ^= hit[sl] ? E[v] : 0ull;
    }
  });
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const int r = 64 * sl + lane;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      out.B[rowidx(t, R, r, W, w)] = P.B[sl][w];
      out.C[rowidx(t, R, r, W, w)] = P.ca[sl][w];
    }
    out.a[(size_t)t * R + r] = (uint8_t)P.a[sl];
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      ((u64*)vout.hi)[((size_t)t * R + r) * VW + v] = V.hi[sl][v];
      ((u64*)vout.rhs)[((size_t)t * R + r) * VW + v] = V.ra[sl][v];
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int w = 0; w < W; ++w) out.rhs[(size_t)t * W + w] = P.rhsa[w];
    out.p0[t] = P.c0;
    out.nc[t] = in.nc[t];
    out.k[t] = in.k[t];
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      // the stored planes hold each variant's full offset: dv folded in
      z8_add(V.d0[v], V.d1[v], V.d2[v], (V.dv & 1) ? ~0ull : 0ull, (V.dv & 2) ? ~0ull : 0ull,
             (V.dv & 4) ? ~0ull : 0ull);
      ((u64*)vout.ph)[((size_t)t * 3 + 0) * VW + v] = V.d0[v];
      ((u64*)vout.ph)[((size_t)t * 3 + 1) * VW + v] = V.d1[v];
      ((u64*)vout.ph)[((size_t)t * 3 + 2) * VW + v] = V.d2[v];
    }
  }
}

template <int W, int VW>
__global__ __launch_bounds__(256) void k_qproject_var(QFormDev in, QFormDev out, VarDev vout,
                                                      const int32_t* __restrict__ fq,
                                                      const int32_t* __restrict__ fb, int nfix,
                                                      const int32_t* __restrict__ uq,
                                                      const uint64_t* __restrict__ ub, int nunit) {
  const int64_t t = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  if (t >= in.chi) return;
  qproject_var_term<W, VW>(in, out, vout, fq, fb, nfix, uq, ub, nunit, t, threadIdx.x & 63);
}

// The same for several states at once (sampling windows of many circuits):
// block b belongs to the job j with jobs[j].block0 <= b < jobs[j + 1].block0
// (found by a scalar binary search over the njobs jobs), four terms per block.
template <int W, int VW>
__global__ __launch_bounds__(256) void k_qproject_var_jobs(const QProjJobDev* __restrict__ jobs, int njobs) {
  const int64_t b = blockIdx.x;
  int lo = 0, hi = njobs;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (uni64((u64)jobs[mid].block0) <= (u64)b) lo = mid;
    else hi = mid;
  }
  const QProjJobDev& j = jobs[lo];
  const int64_t t = (b - (int64_t)uni64((u64)j.block0)) * 4 + uni((int)(threadIdx.x >> 6));
  if (t >= j.in.chi) return;
  qproject_var_term<W, VW>(j.in, j.out, j.vout, j.fq, j.fb, j.nfix, j.uq, j.ub, j.nunit, t, threadIdx.x & 63);
}

// ---------------------------------------------------------------------------
// Uniform Z-marginals.  Read the ket variants as symbols: variant k is Z on
// coordinate k of a qubit order, and bit k of a row's mask is the
// coefficient of t_k in a linear form over GF(2).  Every update of the
// masks above is linear in them (a substitution's constant for a variant
// is c ^ <H_p, t>, a rank-1 update adds H_p), so after the same elimination
// the row masks are the linear forms themselves.  Z^t never moves the
// support, so the constraint phase is the same for every t; after it,
// <bra| Z^t |ket> vanishes exactly when an uncoupled variable with even
// coefficient a_p + 2 <H_p, t> = 2 (mod 4), i.e. <H_p, t> != a_p / 2.  The
// pair is therefore nonzero exactly for the t with <H_p, t> = a_p / 2 at
// every uncoupled even variable p.  An uncoupled variable's row never
// changes again, so its lane keeps its equation: eqm marks those lanes
// (per slot), em the ones whose constant a_p / 2 is 1.
// Returns 1 if the pair vanishes for every t (an inconsistent constraint).
template <int W, int VW, bool USEB>
__device__ __forceinline__ int take_constraint_sym(PairRegs<W>& P, VarRegs<W, VW>& V, int c) {
  u64 row[W];
  bcast<W>(USEB ? P.cb : P.ca, c, row);
  const int rhs = USEB ? getbit<W>(P.rhsb, c) : getbit<W>(P.rhsa, c);
  if (!anyv<W>(row)) return rhs;
  const int p = ffsv<W>(row);
  flipbit<W>(row, p, 1);
  u64 E[VW];
#pragma unroll
  for (int v = 0; v < VW; ++v) E[v] = 0ull;
  // bra rows in reduced echelon form (every stored form is): a bra
  // substitution leaves the other bra rows as they are
  subst_v<W, VW, CONS_B, false, false>(P, V, p, row, rhs, E);
  return 0;
}

template <int W, int VW>
__device__ __forceinline__ int pair_eval_sym(PairRegs<W>& P, VarRegs<W, VW>& V, int nca, int ncb, u64 (&eqm)[W],
                                             u64 (&em)[W]) {
  uniform_state<W>(P);
  for (int c = 0; c < nca; ++c) {
    if (uni(take_constraint_sym<W, VW, false>(P, V, c))) return 1;
    uniform_state<W>(P);
  }
  for (int c = 0; c < ncb; ++c) {
    if (uni(take_constraint_sym<W, VW, true>(P, V, c))) return 1;
    uniform_state<W>(P);
  }
#pragma unroll
  for (int s = 0; s < W; ++s) eqm[s] = em[s] = 0ull;
  for (;;) {
    u64 nz[W];
    bool more = false;
#pragma unroll
    for (int sl = 0; sl < W; ++sl) {
      bool rownz = false;
#pragma unroll
      for (int w = 0; w < W; ++w) rownz = rownz || (P.B[sl][w] != 0ull);
      nz[sl] = uni64(__ballot(rownz)) & P.rem[sl];
      more = more || (nz[sl] != 0ull);
      const u64 iso = P.rem[sl] & ~nz[sl];
      const u64 odd = uni64(__ballot(P.a[sl] & 1));
      eqm[sl] |= iso & ~odd;
      em[sl] |= iso & uni64(__ballot(P.a[sl] == 2));
      P.rem[sl] = nz[sl];
    }
    if (!more) break;
    const int p = uni(ffsv<W>(nz));
    const int pw = wordof<W>(p), pb = p & 63;
    clearbit<W>(P.rem, pw, pb);
    u64 Bp[W];
    bcast<W>(P.B, p, Bp);
    const int ap = bcast_int<W>(P.a, p);
    u64 H[VW];
    bcast_v<W, VW>(V.hi, p, H);
    if (ap & 1) {
      const int add = (ap == 1) ? 3 : 1;
#pragma unroll
      for (int sl = 0; sl < W; ++sl) {
        const int bpl = (int)((wsel<W>(P.B[sl], pw) >> pb) & 1ull);
        const u64 mb = 0ull - (u64)bpl;
        P.a[sl] = (P.a[sl] + (bpl ? add : 0)) & 3;
#pragma unroll
        for (int v = 0; v < VW; ++v) V.hi[sl][v] ^= H[v] & mb;
#pragma unroll
        for (int w = 0; w < W; ++w) P.B[sl][w] ^= Bp[w] & mb;
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (w == sl) P.B[sl][w] &= ~(1ull << P.lane);
        clearbit<W>(P.B[sl], pw, pb);
      }
    } else {
#pragma unroll
      for (int sl = 0; sl < W; ++sl) clearbit<W>(P.B[sl], pw, pb);
      const int qv = ffsv<W>(Bp);
      flipbit<W>(Bp, qv, 1);
      subst_v<W, VW, CONS_NONE, false, false>(P, V, qv, Bp, ap >> 1, H);
    }
    uniform_state<W>(P);
  }
  return 0;
}

// Smallest D <= lim such that the equations (rows EQ on the lanes of valid,
// one per lane and slot, constants em) have a solution t != 0 supported on
// coordinates [0, D); -1 if there is none.  Columns are taken in order by a
// lane-parallel elimination over the rows: with all constants zero, the
// first column that is not a pivot (it lies in the span of the earlier
// columns) gives D; otherwise D is the first column after which every
// non-pivot row has a zero constant (the constants lie in the span of the
// columns so far).  EQ and em are overwritten.
template <int W, int VW>
__device__ __forceinline__ int first_solution(u64 (&EQ)[W][VW], const u64 (&valid)[W], u64 (&em)[W], int lim,
                                              int lane) {
  u64 used[W];
  u64 anye = 0ull;
#pragma unroll
  for (int s = 0; s < W; ++s) {
    used[s] = 0ull;
    em[s] &= valid[s];
    anye |= em[s];
  }
  const bool homog = anye == 0ull;
  for (int c = 0; c < lim; ++c) {
    const int cw = c >> 6, cb = c & 63;
    int ps = -1;
    u64 cand = 0ull;
    u64 hasm[W];
#pragma unroll
    for (int s = 0; s < W; ++s) {
      u64 colw = EQ[s][0];
#pragma unroll
      for (int v = 1; v < VW; ++v)
        if (cw == v) colw = EQ[s][v];
      hasm[s] = uni64(__ballot((int)((colw >> cb) & 1ull))) & valid[s];
      const u64 cs = hasm[s] & ~used[s];
      if (ps < 0 && cs != 0ull) {
        ps = s;
        cand = cs;
      }
    }
    if (ps < 0) {
      if (homog) return c + 1;
      continue;
    }
    const int pl = (int)__builtin_ctzll(cand);
    u64 piv[VW];
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      u64 x = EQ[0][v];
#pragma unroll
      for (int s = 1; s < W; ++s)
        if (ps == s) x = EQ[s][v];
      piv[v] = rl64(x, pl);
    }
    u64 pe = 0ull;
#pragma unroll
    for (int s = 0; s < W; ++s)
      if (ps == s) pe = (em[s] >> pl) & 1ull;
#pragma unroll
    for (int s = 0; s < W; ++s) {
      const u64 own = (ps == s) ? (1ull << pl) : 0ull;
      used[s] |= own;
      const u64 hit = hasm[s] & ~own;
      const bool mine = (hit >> lane) & 1ull;
#pragma unroll
      for (int v = 0; v < VW; ++v) EQ[s][v] ^= mine ? piv[v] : 0ull;
      if (pe) em[s] ^= hit;
    }
    if (!homog) {
      u64 pend = 0ull;
#pragma unroll
      for (int s = 0; s < W; ++s) pend |= em[s] & ~used[s];
      if (pend == 0ull) return c + 1;
    }
  }
  return -1;
}

// Over all pairs (bra o >= ket f) of one term set, the smallest D for which
// some <phi_o| Z^t |phi_f> with t != 0 supported on the first D coordinates
// is nonzero: atomicMin into *dmin (which enters as ncoord + 1).  coord[r]
// is the coordinate of qubit r (-1 when it is not tracked); variables in
// removed (W words, may be null) were substituted already (projected
// forms).  Units and offsets as in k_pairs (triangle, J bras per unit).
template <int W, int VW>
__global__ __launch_bounds__(256) void k_pairs_sym(QFormDev X, const int64_t* __restrict__ offsets, int64_t chi,
                                                   int J, const int32_t* __restrict__ coord,
                                                   const u64* __restrict__ removed, int* __restrict__ dmin) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  const int64_t nunits = offsets[chi];
  const int n = X.n;
  u64 valid[W];
  u64 h0[W][VW];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int hi = n - 64 * w;
    valid[w] = hi >= 64 ? ~0ull : (hi <= 0 ? 0ull : ((1ull << hi) - 1ull));
    if (removed) valid[w] &= ~uni64(removed[w]);
    const int r = 64 * w + lane;
    const int cr = r < n ? coord[r] : -1;
#pragma unroll
    for (int v = 0; v < VW; ++v) h0[w][v] = (cr >= 0 && (cr >> 6) == v) ? (1ull << (cr & 63)) : 0ull;
  }
  for (int64_t u = wave; u < nunits; u += nwaves) {
    int64_t lo_r = 0, hi_r = chi;
    while (hi_r - lo_r > 1) {
      const int64_t mid = (lo_r + hi_r) >> 1;
      if (offsets[mid] <= u) lo_r = mid;
      else hi_r = mid;
    }
    const int64_t fx = lo_r;
    const int64_t o_lo = fx + (u - offsets[fx]) * J;
    const int64_t o_hi = (o_lo + J < chi) ? o_lo + J : chi;
    const int ncf = uni(X.nc[fx]);
    for (int64_t o = o_lo; o < o_hi; ++o) {
      const int cap = uni(__hip_atomic_load(dmin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (cap <= 1) return;
      u64 Bf[W][W], Cf[W][W], rf[W];
      int af[W];
      load_qrows<W>(X, fx, lane, Bf, af, Cf, rf);
      u64 Bo[W][W], Co[W][W], ro[W];
      int ao[W];
      load_qrows<W>(X, o, lane, Bo, ao, Co, ro);
      const int nco = uni(X.nc[o]);
      PairRegs<W> P;
      VarRegs<W, VW> V;
      P.lane = lane;
      P.nodiag = ~(1ull << lane);
#pragma unroll
      for (int sl = 0; sl < W; ++sl) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
          P.B[sl][w] = Bf[sl][w] ^ Bo[sl][w];
          P.ca[sl][w] = Co[sl][w];
          P.cb[sl][w] = Cf[sl][w];
        }
        P.a[sl] = (af[sl] - ao[sl]) & 3;
#pragma unroll
        for (int v = 0; v < VW; ++v) {
          V.hi[sl][v] = h0[sl][v];
          V.ra[sl][v] = V.rb[sl][v] = 0ull;
        }
      }
#pragma unroll
      for (int w = 0; w < W; ++w) {
        P.rhsa[w] = ro[w];
        P.rhsb[w] = rf[w];
        P.rem[w] = valid[w];
      }
#pragma unroll
      for (int v = 0; v < VW; ++v) V.d0[v] = V.d1[v] = V.d2[v] = V.z[v] = 0ull;
      V.bz = 0;
      V.dv = 0;
      P.c0 = 0;
      P.mag = 0;
      u64 eqm[W], em[W];
      if (pair_eval_sym<W, VW>(P, V, nco, ncf, eqm, em)) continue;
      // only D < cap can lower the minimum
      const int d = first_solution<W, VW>(V.hi, eqm, em, cap - 1, lane);
      if (d > 0 && lane == 0) atomicMin(dmin, d);
    }
  }
}

// ---------------------------------------------------------------------------
// Pauli variants.  For P applied to phi_t (gate_application.rs:186-210 op
// order, left_multiplication/{x,y,z}.rs): Z_q adds 2 to gamma_q/2, X_q
// flips s by u_q = (F_q & ~v) ^ (M_q & v) and adds 4 beta + 2 gamma_q/2 to
// the phase, Y_q = Z_q, X_q, then +2.  With ds the total change of s, the
// quadratic form's a_j = gamma_j/2 + 2 (par(s & w_j) ^ c_j) gains
// 2 (zy_j ^ par(ds & w_j)), the constraint rhs (s at the positions with
// v = 0, in order) flips by ds there -- carried through the row reduction
// the stored rows went through -- and p0 = phase changes by dph.
template <int W, int VW>
__global__ __launch_bounds__(256) void k_pauli_var(FormsDev in, const int32_t* __restrict__ voff,
                                                   const int32_t* __restrict__ vops, int nvar,
                                                   u64* __restrict__ vhi, u64* __restrict__ vrhs,
                                                   u64* __restrict__ vph) {
  const int64_t t = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  const int lane = threadIdx.x & 63;
  if (t >= in.chi) return;
  Form<W> f;
  load_form<W>(f, in, t, lane);
  const int n = in.n, R = 64 * W;
  u64 Wm[W][W];
  int kpos[W];
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
#pragma unroll
    for (int w = 0; w < W; ++w) Wm[sl][w] = (f.M[sl][w] & ~f.v[w]) ^ (f.F[sl][w] & f.v[w]);
    kpos[sl] = -1;
  }
  // the constraint rows as the quadratic-form kernel forms them (column k
  // of F for the c-th k with v_k = 0), to reduce the rhs flips with them
  u64 C[W][W];
#pragma unroll
  for (int sl = 0; sl < W; ++sl)
#pragma unroll
    for (int w = 0; w < W; ++w) C[sl][w] = 0ull;
  int nc = 0;
  for (int k = 0; k < n; ++k) {
    if (getbit<W>(f.v, k)) continue;
    u64 col[W];
#pragma unroll
    for (int sw = 0; sw < W; ++sw) col[sw] = __ballot(getbit<W>(f.F[sw], k));
#pragma unroll
    for (int sl = 0; sl < W; ++sl)
      if (owns<W>(f, nc, sl)) {
        kpos[sl] = k;
#pragma unroll
        for (int w = 0; w < W; ++w) C[sl][w] = col[w];
      }
    ++nc;
  }
  u64 hacc[W][VW], racc[W][VW], p0[VW], p1[VW], p2[VW];
#pragma unroll
  for (int v = 0; v < VW; ++v) {
    p0[v] = p1[v] = p2[v] = 0ull;
#pragma unroll
    for (int sl = 0; sl < W; ++sl) hacc[sl][v] = racc[sl][v] = 0ull;
  }
  for (int k = 0; k < nvar; ++k) {
    u64 s[W];
#pragma unroll
    for (int w = 0; w < W; ++w) s[w] = f.s[w];
    int ph = f.ph;
    int zy[W];
#pragma unroll
    for (int sl = 0; sl < W; ++sl) zy[sl] = 0;
    const int e1 = voff[k + 1];
    for (int e = voff[k]; e < e1; ++e) {
      const int code = vops[e];
      const int q = code >> 2, op = code & 3;
      int gadd = 0;
      if (op == 2 || op == 3) {
#pragma unroll
        for (int sl = 0; sl < W; ++sl)
          if (owns<W>(f, q, sl)) zy[sl] ^= 1;
        gadd = 2;
      }
      if (op == 1 || op == 2) {
        u64 Fq[W], Mq[W];
        bcast<W>(f.F, q, Fq);
        bcast<W>(f.M, q, Mq);
        const int gq = (bcast_g<W>(f, q) + gadd) & 3;
        int beta = 0;
#pragma unroll
        for (int w = 0; w < W; ++w)
          beta += __builtin_popcountll((Mq[w] & ~f.v[w] & s[w]) ^ (Fq[w] & f.v[w] & (Mq[w] ^ s[w])));
        ph = (ph + 4 * (beta & 1) + 2 * gq) & 7;
#pragma unroll
        for (int w = 0; w < W; ++w) s[w] ^= (Fq[w] & ~f.v[w]) ^ (Mq[w] & f.v[w]);
      }
      if (op == 2) ph = (ph + 2) & 7;
    }
    u64 ds[W];
#pragma unroll
    for (int w = 0; w < W; ++w) ds[w] = s[w] ^ f.s[w];
    const int dph = (ph - f.ph) & 7;
    const int kv = k >> 6;
    const u64 bit = 1ull << (k & 63);
#pragma unroll
    for (int sl = 0; sl < W; ++sl) {
      const bool real_row = 64 * sl + lane < n;
      const int hb = real_row ? (zy[sl] ^ parand<W>(ds, Wm[sl])) : 0;
      const int rb = kpos[sl] >= 0 ? getbit<W>(ds, kpos[sl]) : 0;
#pragma unroll
      for (int v = 0; v < VW; ++v)
        if (v == kv) {
          if (hb) hacc[sl][v] |= bit;
          if (rb) racc[sl][v] |= bit;
        }
    }
#pragma unroll
    for (int v = 0; v < VW; ++v)
      if (v == kv) {
        if (dph & 1) p0[v] |= bit;
        if (dph & 2) p1[v] |= bit;
        if (dph & 4) p2[v] |= bit;
      }
  }
  // the flips of the stored (reduced) rows: each row is a sum of the rows
  // above, so its flip is the sum of theirs
  rref_rows<W>(C, nc, lane, [&](int c, const bool (&hit)[W]) {
    u64 E[VW];
    bcast_v<W, VW>(racc, c, E);
#pragma unroll
    for (int sl = 0; sl < W; ++sl)
#pragma unroll
      for (int v = 0; v < VW; ++v) racc[sl][v] ^= hit[sl] ? E[v] : 0ull;
  });
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const size_t r = (size_t)t * R + 64 * sl + lane;
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      vhi[r * VW + v] = hacc[sl][v];
      vrhs[r * VW + v] = racc[sl][v];
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      vph[((size_t)t * 3 + 0) * VW + v] = p0[v];
      vph[((size_t)t * 3 + 1) * VW + v] = p1[v];
      vph[((size_t)t * 3 + 2) * VW + v] = p2[v];
    }
  }
}

// ---------------------------------------------------------------------------
// Hermitian pair sums over ket variants.  Each wave belongs to one job
// (wave_job) and walks that job's units with the job's wave stride; lane k
// accumulates variant 64*v + k in acc[v], every lane the base in accb:
//   sum_f sum_{o >= f} (o == f ? 1 : 2) Re(conj(c_o) c_f <X_o | Y_f^(k)>)
// which is exp_value.rs:38-58 for Y_f^(k) = P_k phi_f (norm.rs:16-27 for
// the base).  Per wave, partial holds 64*VW variant sums then the base.
// REC: also write the pair records of jd.rec (window passes whose node
// probabilities are folded in the reference's order; VW = 1)
template <int W, int VW, bool UNITS, bool REC = false>
__global__ __launch_bounds__(256, (W == 1 && VW == 1 && !REC) ? 8 : 1) void k_pairs_var(const PairJobDev* __restrict__ jobs,
                                                   const int32_t* __restrict__ wave_job, int64_t wave_base,
                                                   int64_t nwaves_launch, int J, const double* __restrict__ tabs,
                                                   double* __restrict__ partial) {
  constexpr int NVS = 64 * VW + 1;
  const int lane = threadIdx.x & 63;
  const int64_t lw = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  if (lw >= nwaves_launch) return;
  const int64_t wave = wave_base + lw;
  const int jb = uni(wave_job[wave]);
  const PairJobDev& jd = jobs[jb];
  const QFormDev& X = jd.X;
  const QFormDev& Y = jd.Y;
  const int64_t nfixed = jd.nfixed, nother = jd.nother;
  const int64_t nunits = jd.row_off[nfixed];
  const int n = X.n, R = 64 * W;
  const int want_base = jd.base_out >= 0;
  u64 valid[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int hi = n - 64 * w;
    valid[w] = hi >= 64 ? ~0ull : (hi <= 0 ? 0ull : ((1ull << hi) - 1ull));
    valid[w] &= ~uni64(jd.removed[w]);
  }
  u64 invalid[VW];
#pragma unroll
  for (int v = 0; v < VW; ++v) {
    const int lo = jd.nvar - 64 * v;
    invalid[v] = lo >= 64 ? 0ull : (lo <= 0 ? ~0ull : ~((1ull << lo) - 1ull));
  }
  // acc: FP64 sums (complex coefficients), or the dyadic part A of the
  // exact sums (exact Scalar coefficients) with the sqrt(2) part in accB
  const bool exact = jd.cx.scalar && jd.cy.scalar;
  double acc[VW], accB[VW], accb = 0.0, accbB = 0.0;
#pragma unroll
  for (int v = 0; v < VW; ++v) acc[v] = accB[v] = 0.0;
  for (int64_t u = wave - jd.wave0; u < nunits; u += jd.nwaves) {
    int64_t lo_r = 0, hi_r = nfixed;
    while (hi_r - lo_r > 1) {
      const int64_t mid = (lo_r + hi_r) >> 1;
      if (jd.row_off[mid] <= u) lo_r = mid;
      else hi_r = mid;
    }
    const int64_t fx = lo_r;
    const int64_t base = jd.triangle ? fx : 0;
    const int64_t o_lo = base + (u - jd.row_off[fx]) * J;
    const int64_t o_hi = (o_lo + J < nother) ? o_lo + J : nother;
    const int ncf = uni(Y.nc[fx]), kf = uni(Y.k[fx]), pf = uni(Y.p0[fx]);
    // ket variant tables: Pauli strings (!UNITS), or window outcomes on both
    // sides (UNITS with termvar: k_qproject_var tables for X = Y)
    const bool tv = UNITS && jd.termvar;
    const bool ket_tab = !UNITS || tv;
    u64 vp0[VW], vp1[VW], vp2[VW];
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      vp0[v] = ket_tab ? uni64(jd.V.ph[((size_t)fx * 3 + 0) * VW + v]) : 0ull;
      vp1[v] = ket_tab ? uni64(jd.V.ph[((size_t)fx * 3 + 1) * VW + v]) : 0ull;
      vp2[v] = ket_tab ? uni64(jd.V.ph[((size_t)fx * 3 + 2) * VW + v]) : 0ull;
    }
    for (int64_t o = o_lo; o < o_hi; ++o) {
      // ket rows are re-read per pair (cache hits) to keep registers free
      u64 Bf[W][W], Cf[W][W], rf[W];
      int af[W];
      load_qrows<W>(Y, fx, lane, Bf, af, Cf, rf);
      u64 Bo[W][W], Co[W][W], ro[W];
      int ao[W];
      load_qrows<W>(X, o, lane, Bo, ao, Co, ro);
      const int nco = uni(X.nc[o]), ko = uni(X.k[o]), po = uni(X.p0[o]);
      PairRegs<W> P;
      VarRegs<W, VW> V;
      P.lane = lane;
      P.nodiag = ~(1ull << lane);
#pragma unroll
      for (int sl = 0; sl < W; ++sl) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
          P.B[sl][w] = Bf[sl][w] ^ Bo[sl][w];
          P.ca[sl][w] = Co[sl][w];
          P.cb[sl][w] = Cf[sl][w];
        }
        P.a[sl] = (af[sl] - ao[sl]) & 3;
        const size_t r = (size_t)fx * R + 64 * sl + lane;
        const size_t ro_ = (size_t)o * R + 64 * sl + lane;
#pragma unroll
        for (int v = 0; v < VW; ++v) {
          V.hi[sl][v] = ket_tab ? jd.V.hi[r * VW + v] : 0ull;
          V.rb[sl][v] = ket_tab ? jd.V.rhs[r * VW + v] : 0ull;
          V.ra[sl][v] = 0ull;
          if (tv) {
            V.hi[sl][v] ^= jd.V.hi[ro_ * VW + v];
            V.ra[sl][v] = jd.V.rhs[ro_ * VW + v];
          }
        }
      }
#pragma unroll
      for (int w = 0; w < W; ++w) {
        P.rhsa[w] = ro[w];
        P.rhsb[w] = rf[w];
        P.rem[w] = valid[w];
      }
#pragma unroll
      for (int v = 0; v < VW; ++v) {
        V.d0[v] = vp0[v];
        V.d1[v] = vp1[v];
        V.d2[v] = vp2[v];
        if (tv)  // minus the bra's phase planes: + ~d here, the + 1 (mod 8) in dv
          z8_add(V.d0[v], V.d1[v], V.d2[v], ~uni64(jd.V.ph[((size_t)o * 3 + 0) * VW + v]),
                 ~uni64(jd.V.ph[((size_t)o * 3 + 1) * VW + v]), ~uni64(jd.V.ph[((size_t)o * 3 + 2) * VW + v]));
        V.z[v] = invalid[v];
      }
      V.bz = !want_base;
      V.dv = tv ? 1 : 0;
      P.c0 = (pf - po) & 7;
      P.mag = kf + ko;
      pair_eval_v<W, VW, UNITS>(P, V, nco, ncf);
      if constexpr (REC && UNITS && VW == 1) {
        if (lane == 0) {
          // exact per-variant values, at the pair's position in the
          // reference's summation order (row f, then o = f .. chi - 1)
          uint64_t* rp = jd.rec + (size_t)FOLD_REC * (size_t)(fx * nother - fx * (fx - 1) / 2 + (o - fx));
          rp[0] = V.z[0];
          rp[1] = V.d0[0];
          rp[2] = V.d1[0];
          rp[3] = V.d2[0];
          rp[4] = (uint64_t)((P.c0 + V.dv) & 7) | ((uint64_t)(uint32_t)P.mag << 8);
        }
      }
      // (c_bra.conj() * c_ket).into(), shared by the variants
      double cr_, ci_;
      int kc = 0, rc = 0, czero = 0;  // exact coefficient product zeta^kc 2^{-rc/2}
      if (exact) {
        const int32_t sb = jd.cx.cs[o], sk = jd.cy.cs[fx];
        czero = (sb | sk) < 0;
        kc = (cs_k(sk) - cs_k(sb)) & 7;
        rc = cs_r(sb) + cs_r(sk);
        cr_ = 0.0;
        ci_ = 0.0;
      } else {
        const double2 b0 = jd.cx.c[o], k0 = jd.cy.c[fx];
        const double br = b0.x, bi = -b0.y;
        cr_ = br * k0.x - bi * k0.y;
        ci_ = br * k0.y + bi * k0.x;
      }
      const bool diag = (o == fx);
      const double nrm = tabs[TAB_POW + P.mag];
      const double w2 = diag ? 1.0 : 2.0;
      if (!V.bz) {
        const int K = P.c0;
        if (exact) {
          if (!czero) add_exact(kc + K, rc + P.mag, w2, accb, accbB);
        } else {
          const double ir = tabs[TAB_COS + K] * nrm, ii = tabs[TAB_SIN + K] * nrm;
          const double tr = cr_ * ir - ci_ * ii;
          accb = diag ? accb + tr : accb + (tr + tr);
        }
      }
#pragma unroll
      for (int v = 0; v < VW; ++v) {
        const int zk = (int)((V.z[v] >> lane) & 1ull);
        if (!zk) {
          const int D = (int)((V.d0[v] >> lane) & 1ull) | ((int)((V.d1[v] >> lane) & 1ull) << 1) |
                        ((int)((V.d2[v] >> lane) & 1ull) << 2);
          const int K = (P.c0 + V.dv + D) & 7;
          if (exact) {
            if (!czero) add_exact(kc + K, rc + P.mag, w2, acc[v], accB[v]);
          } else {
            const double ir = tabs[TAB_COS + K] * nrm, ii = tabs[TAB_SIN + K] * nrm;
            const double tr = cr_ * ir - ci_ * ii;
            acc[v] = diag ? acc[v] + tr : acc[v] + (tr + tr);
          }
        }
      }
    }
  }
  // per wave: the NVS sums (A parts), then their sqrt(2) parts (B)
#pragma unroll
  for (int v = 0; v < VW; ++v) {
    partial[(size_t)wave * 2 * NVS + 64 * v + lane] = acc[v];
    partial[(size_t)wave * 2 * NVS + NVS + 64 * v + lane] = accB[v];
  }
  if (lane == 0) {
    partial[(size_t)wave * 2 * NVS + 64 * VW] = accb;
    partial[(size_t)wave * 2 * NVS + NVS + 64 * VW] = accbB;
  }
}

// First stage: block b sums, for each of the 2 (64 VW + 1) values of a wave
// (the variant and base sums, A parts then sqrt(2) parts), the partials of
// the waves of segment b: four runs of consecutive waves (one per quarter of
// the block, each in wave order) whose totals are added in run order; a
// thread per value, so each wave's row is one coalesced read.
constexpr int RED_SEG_NT = 512;
template <int VW>
__global__ __launch_bounds__(RED_SEG_NT) void k_reduce_seg(const PairJobDev* __restrict__ jobs,
                                                           const int32_t* __restrict__ seg_job,
                                                           const double* __restrict__ partial,
                                                           double* __restrict__ seg_part) {
  constexpr int NV2 = 2 * (64 * VW + 1);
  constexpr int GT = 128, NG = RED_SEG_NT / GT;
  __shared__ double s_p[NG][NV2];
  const int64_t b = blockIdx.x;
  const PairJobDev& jd = jobs[seg_job[b]];
  const int64_t w0 = jd.wave0 + (b - jd.seg0) * RED_SEG;
  const int64_t wend = jd.wave0 + jd.nwaves;
  const int64_t w1 = (w0 + RED_SEG < wend) ? w0 + RED_SEG : wend;
  const int g = threadIdx.x / GT, tl = threadIdx.x % GT;
  constexpr int64_t PER = RED_SEG / NG;
  const int64_t r0 = w0 + g * PER;
  const int64_t r1 = (r0 + PER < w1) ? r0 + PER : w1;
  for (int t = tl; t < NV2; t += GT) {
    // loads issued 8 at a time, added in wave order
    double a = 0.0;
    int64_t w = r0;
    for (; w + 8 <= r1; w += 8) {
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = partial[(size_t)(w + k) * NV2 + t];
#pragma unroll
      for (int k = 0; k < 8; ++k) a = a + v[k];
    }
    for (; w < r1; ++w) a = a + partial[(size_t)w * NV2 + t];
    s_p[g][t] = a;
  }
  __syncthreads();
  if (g != 0) return;
  for (int t = tl; t < NV2; t += GT) {
    double a = s_p[0][t];
#pragma unroll
    for (int k = 1; k < NG; ++k) a = a + s_p[k][t];
    seg_part[(size_t)b * NV2 + t] = a;
  }
}

// Second stage: block j adds job j's segment sums, in four runs of
// consecutive segments (one per quarter of the block, each in segment order)
// whose totals are then added in run order; the value is A + sqrt(2) B (or A
// and B apart with split_ab).
constexpr int RED_VAR_NT = 1024;
template <int VW>
__global__ __launch_bounds__(RED_VAR_NT) void k_reduce_var(const PairJobDev* __restrict__ jobs,
                                                           const double* __restrict__ seg_part,
                                                           double* __restrict__ out) {
  constexpr int NVS = 64 * VW + 1;
  constexpr int NG = RED_VAR_NT / 256;
  constexpr double SQRT2 = 1.4142135623730951;  // nearest double to sqrt(2)
  __shared__ double s_a[NG][NVS], s_b[NG][NVS];
  const PairJobDev& jd = jobs[blockIdx.x];
  const int g = threadIdx.x >> 8, tl = threadIdx.x & 255;
  const int64_t per = (jd.nseg + NG - 1) / NG;
  const int64_t r0 = jd.seg0 + g * per;
  const int64_t s1 = (r0 + per < jd.seg0 + jd.nseg) ? r0 + per : jd.seg0 + jd.nseg;
  for (int v = tl; v < NVS; v += 256) {
    // loads issued 8 segments at a time, added in segment order
    double ta = 0.0, tb = 0.0;
    int64_t s = r0;
    for (; s + 8 <= s1; s += 8) {
      double va[8], vb[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        va[k] = seg_part[(size_t)(s + k) * 2 * NVS + v];
        vb[k] = seg_part[(size_t)(s + k) * 2 * NVS + NVS + v];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        ta = ta + va[k];
        tb = tb + vb[k];
      }
    }
    for (; s < s1; ++s) {
      ta = ta + seg_part[(size_t)s * 2 * NVS + v];
      tb = tb + seg_part[(size_t)s * 2 * NVS + NVS + v];
    }
    s_a[g][v] = ta;
    s_b[g][v] = tb;
  }
  __syncthreads();
  if (g != 0) return;
  for (int v = tl; v < NVS; v += 256) {
    double ta = s_a[0][v], tb = s_b[0][v];
#pragma unroll
    for (int k = 1; k < NG; ++k) {
      ta = ta + s_a[k][v];
      tb = tb + s_b[k][v];
    }
    const double t = ta + SQRT2 * tb;
    if (v < jd.nvar) {
      if (jd.split_ab) {
        out[jd.out_base + v] = ta;
        out[jd.out_base + jd.nvar + v] = tb;
      } else {
        out[jd.out_base + v] = t;
      }
    }
    if (v == 64 * VW && jd.base_out >= 0) out[jd.base_out] = t;
  }
}

// ---------------------------------------------------------------------------
// Coefficient backward.  For L = sum_k w_k exp_value(P_k) = |g|^2 c^H G c
// with G = sum_k w_k G_k, (G_k)_{ji} = <phi_j|P_k phi_i> (Hermitian), the
// gradient in PyTorch's convention is 2 |g|^2 y, y = G c.  The pairs (ket f,
// bra o >= f) run in tiles of GT kets x GT bras (block row bf, block column
// bo >= bf); per pair the elimination serves all strings (pair_eval_v), the
// lanes' string values reduce to g = G_{of} = sum_k w_k <phi_o|P_k phi_f>,
// and the tile's lanes i < GT accumulate y_o += g c_f (column o = o0 + i)
// and y_f += conj(g) c_o (row f = f0 + i, o != f).  Each tile writes its GT
// row and GT column partials; k_grad_reduce sums, for each index j, the
// partials of the tiles in its block row and block column in a fixed order.
// The same pass accumulates, per string, the Hermitian sums of
// exp_value.rs:38-58 (per-wave partials, summed by k_grad_vals).
constexpr int GT = 16;

// block row bf and column bo of tile t in the row-major order of the block
// triangle {bo >= bf} of nb x nb blocks: row bf starts at
// bf nb - bf (bf - 1) / 2
__device__ __forceinline__ int64_t gtile_start(int64_t bf, int64_t nb) { return bf * nb - bf * (bf - 1) / 2; }
__device__ __forceinline__ void gtile_of(int64_t t, int64_t nb, int64_t* bf, int64_t* bo) {
  // rows from the end: m = nb - bf is the smallest m with m (m + 1) / 2 >= T - t
  const int64_t T = nb * (nb + 1) / 2, r = T - t;
  int64_t m = (int64_t)ceil(0.5 * (sqrt(1.0 + 8.0 * (double)r) - 1.0));
  m = (int64_t)uni64((uint64_t)m);
  if (m > nb) m = nb;
  if (m < 1) m = 1;
  while (m > 1 && (m - 1) * m / 2 >= r) --m;
  while (m * (m + 1) / 2 < r) ++m;
  *bf = nb - m;
  *bo = *bf + (t - gtile_start(*bf, nb));
}

template <int W, int VW>
__global__ __launch_bounds__(256) void k_grad_var(QFormDev X, VarDev VT, const double2* __restrict__ c, int nvar,
                                                  const double* __restrict__ wts, const double* __restrict__ tabs,
                                                  int64_t ntiles, double2* __restrict__ tile_row,
                                                  double2* __restrict__ tile_col, double* __restrict__ partial) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  const int64_t chi = X.chi, nb = (chi + GT - 1) / GT;
  const int n = X.n, R = 64 * W;
  u64 valid[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int hi = n - 64 * w;
    valid[w] = hi >= 64 ? ~0ull : (hi <= 0 ? 0ull : ((1ull << hi) - 1ull));
  }
  u64 invalid[VW];
  double wk[VW], acc[VW];
#pragma unroll
  for (int v = 0; v < VW; ++v) {
    const int lo = nvar - 64 * v;
    invalid[v] = lo >= 64 ? 0ull : (lo <= 0 ? ~0ull : ~((1ull << lo) - 1ull));
    wk[v] = (64 * v + lane < nvar) ? wts[64 * v + lane] : 0.0;
    acc[v] = 0.0;
  }
  for (int64_t t = wave; t < ntiles; t += nwaves) {
    int64_t bf, bo;
    gtile_of(t, nb, &bf, &bo);
    const int64_t f0 = bf * GT, o0 = bo * GT;
    const int64_t f1 = (f0 + GT < chi) ? f0 + GT : chi, o1 = (o0 + GT < chi) ? o0 + GT : chi;
    double yr_re = 0.0, yr_im = 0.0, yc_re = 0.0, yc_im = 0.0;
    for (int64_t fx = f0; fx < f1; ++fx) {
      const int ncf = uni(X.nc[fx]), kf = uni(X.k[fx]), pf = uni(X.p0[fx]);
      u64 vp0[VW], vp1[VW], vp2[VW];
#pragma unroll
      for (int v = 0; v < VW; ++v) {
        vp0[v] = uni64(VT.ph[((size_t)fx * 3 + 0) * VW + v]);
        vp1[v] = uni64(VT.ph[((size_t)fx * 3 + 1) * VW + v]);
        vp2[v] = uni64(VT.ph[((size_t)fx * 3 + 2) * VW + v]);
      }
      const double2 cf = c[fx];
      for (int64_t o = (o0 > fx ? o0 : fx); o < o1; ++o) {
        u64 Bf[W][W], Cf[W][W], rf[W];
        int af[W];
        load_qrows<W>(X, fx, lane, Bf, af, Cf, rf);
        u64 Bo[W][W], Co[W][W], ro[W];
        int ao[W];
        load_qrows<W>(X, o, lane, Bo, ao, Co, ro);
        const int nco = uni(X.nc[o]), ko = uni(X.k[o]), po = uni(X.p0[o]);
        PairRegs<W> P;
        VarRegs<W, VW> V;
        P.lane = lane;
        P.nodiag = ~(1ull << lane);
#pragma unroll
        for (int sl = 0; sl < W; ++sl) {
#pragma unroll
          for (int w = 0; w < W; ++w) {
            P.B[sl][w] = Bf[sl][w] ^ Bo[sl][w];
            P.ca[sl][w] = Co[sl][w];
            P.cb[sl][w] = Cf[sl][w];
          }
          P.a[sl] = (af[sl] - ao[sl]) & 3;
          const size_t r = (size_t)fx * R + 64 * sl + lane;
#pragma unroll
          for (int v = 0; v < VW; ++v) {
            V.hi[sl][v] = VT.hi[r * VW + v];
            V.rb[sl][v] = VT.rhs[r * VW + v];
            V.ra[sl][v] = 0ull;
          }
        }
#pragma unroll
        for (int w = 0; w < W; ++w) {
          P.rhsa[w] = ro[w];
          P.rhsb[w] = rf[w];
          P.rem[w] = valid[w];
        }
#pragma unroll
        for (int v = 0; v < VW; ++v) {
          V.d0[v] = vp0[v];
          V.d1[v] = vp1[v];
          V.d2[v] = vp2[v];
          V.z[v] = invalid[v];
        }
        V.bz = 1;  // the identity string is not wanted here
        V.dv = 0;
        P.c0 = (pf - po) & 7;
        P.mag = kf + ko;
        pair_eval_v<W, VW, false>(P, V, nco, ncf);
        if (all_zero<W, VW>(V)) continue;  // every string's pair value vanishes
        // (c_bra.conj() * c_ket).into(), shared by the strings
        const double2 co = c[o];
        const double cr_ = co.x * cf.x + co.y * cf.y, ci_ = co.x * cf.y - co.y * cf.x;
        const bool diag = (o == fx);
        const double nrm = tabs[TAB_POW + P.mag];
        double gr = 0.0, gi = 0.0;
#pragma unroll
        for (int v = 0; v < VW; ++v) {
          if (!((V.z[v] >> lane) & 1ull)) {
            const int D = (int)((V.d0[v] >> lane) & 1ull) | ((int)((V.d1[v] >> lane) & 1ull) << 1) |
                          ((int)((V.d2[v] >> lane) & 1ull) << 2);
            const int K = (P.c0 + V.dv + D) & 7;
            const double ir = tabs[TAB_COS + K] * nrm, ii = tabs[TAB_SIN + K] * nrm;
            const double tr = cr_ * ir - ci_ * ii;
            acc[v] = diag ? acc[v] + tr : acc[v] + (tr + tr);
            gr = gr + wk[v] * ir;
            gi = gi + wk[v] * ii;
          }
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          gr = gr + __shfl_xor(gr, off);
          gi = gi + __shfl_xor(gi, off);
        }
        // y_o += g c_f ; y_f += conj(g) c_o off the diagonal
        if (lane == (int)(o - o0)) {
          yc_re = yc_re + (gr * cf.x - gi * cf.y);
          yc_im = yc_im + (gr * cf.y + gi * cf.x);
        }
        if (!diag && lane == (int)(fx - f0)) {
          yr_re = yr_re + (gr * co.x + gi * co.y);
          yr_im = yr_im + (gr * co.y - gi * co.x);
        }
      }
    }
    if (lane < GT) {
      tile_row[(size_t)t * GT + lane] = make_double2(yr_re, yr_im);
      tile_col[(size_t)t * GT + lane] = make_double2(yc_re, yc_im);
    }
  }
#pragma unroll
  for (int v = 0; v < VW; ++v) partial[(size_t)wave * 64 * VW + 64 * v + lane] = acc[v];
}

// y_j * scale for each index j: the column partials of the tiles (bf, j/GT)
// for bf = 0 .. j/GT, then the row partials of the tiles (j/GT, bo) for
// bo = j/GT .. nb - 1, in that order; accumulate != 0 adds to y
__global__ void k_grad_reduce(const double2* __restrict__ tile_row, const double2* __restrict__ tile_col, int64_t chi,
                              double scale, int accumulate, double2* __restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= chi) return;
  const int64_t nb = (chi + GT - 1) / GT, b = j / GT, i = j % GT;
  double re = 0.0, im = 0.0;
  for (int64_t bf = 0; bf <= b; ++bf) {
    const double2 v = tile_col[(size_t)(gtile_start(bf, nb) + (b - bf)) * GT + i];
    re = re + v.x;
    im = im + v.y;
  }
  const int64_t rb = gtile_start(b, nb);
  for (int64_t bo = b; bo < nb; ++bo) {
    const double2 v = tile_row[(size_t)(rb + (bo - b)) * GT + i];
    re = re + v.x;
    im = im + v.y;
  }
  re = re * scale;
  im = im * scale;
  if (accumulate) {
    const double2 old = y[j];
    re = old.x + re;
    im = old.y + im;
  }
  y[j] = make_double2(re, im);
}

// per string k < nvar: sum of the per-wave partials in wave order, times
// scale, to vals[k]; one block per string
__global__ __launch_bounds__(256) void k_grad_vals(const double* __restrict__ partial, int64_t nwaves, int stride,
                                                   double scale, double* __restrict__ vals) {
  __shared__ double sh[256];
  const int k = blockIdx.x;
  double a = 0.0;
  for (int64_t w = threadIdx.x; w < nwaves; w += 256) a = a + partial[(size_t)w * stride + k];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) vals[k] = sh[0] * scale;
}

// loss = sum_k w_k vals[k] in string order (one thread)
__global__ void k_grad_loss(const double* __restrict__ vals, const double* __restrict__ wts, int n,
                            double* __restrict__ loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double l = 0.0;
  for (int k = 0; k < n; ++k) l = l + wts[k] * vals[k];
  *loss = l;
}

// ---------------------------------------------------------------------------
static inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

#define NX_DISPATCH_WV(W_, VW_, ...)                                                         \
  switch ((W_) * 8 + (VW_)) {                                                               \
    case 9: { constexpr int WW = 1, VV = 1; __VA_ARGS__; } break;                           \
    case 10: { constexpr int WW = 1, VV = 2; __VA_ARGS__; } break;                          \
    case 17: { constexpr int WW = 2, VV = 1; __VA_ARGS__; } break;                          \
    case 18: { constexpr int WW = 2, VV = 2; __VA_ARGS__; } break;                          \
    case 25: { constexpr int WW = 3, VV = 1; __VA_ARGS__; } break;                          \
    case 26: { constexpr int WW = 3, VV = 2; __VA_ARGS__; } break;                          \
    case 33: { constexpr int WW = 4, VV = 1; __VA_ARGS__; } break;                          \
    case 34: { constexpr int WW = 4, VV = 2; __VA_ARGS__; } break;                          \
    default: return hipErrorInvalidValue;                                                   \
  }
#define NX_LAUNCH_PAIRS_VAR(UU)                                                                              \
  NX_DISPATCH_WV(W, VW,                                                                                      \
                 k_pairs_var<WW, VV, UU><<<dim3(nblk(nwaves, 4)), dim3(256), 0, st>>>(jobs, wave_job, wave_base, \
                                                                                      nwaves, J, tabs, partial))

#define NX_LAUNCH_PAIRS_VAR_REC(WW)                                                          \
  k_pairs_var<WW, 1, true, true><<<dim3(nblk(nwaves, 4)), dim3(256), 0, st>>>(jobs, wave_job, wave_base, \
                                                                             nwaves, J, tabs, partial)

hipError_t launch_pauli_var(const FormsDev& in, const int32_t* voff, const int32_t* vops, int nvar, int VW,
                            uint64_t* vhi, uint64_t* vrhs, uint64_t* vph, hipStream_t st) {
  if (in.chi == 0) return hipSuccess;
  if (nvar > 64 * VW) return hipErrorInvalidValue;
  NX_DISPATCH_WV(in.W, VW,
                 k_pauli_var<WW, VV><<<dim3(nblk(in.chi, 4)), dim3(256), 0, st>>>(in, voff, vops, nvar, vhi, vrhs,
                                                                                   vph));
  return hipGetLastError();
}

int64_t pairs_var_resident_waves(int W, int VW, int units) {
  if (W < 1 || W > 4 || VW < 1 || VW > 2) return 0;
  const void* f = nullptr;
#define NX_PV_FUNC(UU) NX_DISPATCH_WV(W, VW, f = (const void*)k_pairs_var<WW, VV, UU>)
  if (units) NX_PV_FUNC(true)
  else NX_PV_FUNC(false)
#undef NX_PV_FUNC
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, 256, 0) != hipSuccess) per = 1;
  return (int64_t)std::max(per, 1) * std::max(cus, 1) * 4;
}

hipError_t launch_pairs_var(int W, int VW, int units, const PairJobDev* jobs, const int32_t* wave_job,
                            int64_t wave_base, int64_t nwaves, int J, const double* tabs, double* partial,
                            hipStream_t st) {
  if (nwaves == 0) return hipSuccess;
  if (units == 2) {  // with pair records: one word of variants
    if (VW != 1) return hipErrorInvalidValue;
    switch (W) {
      case 1: NX_LAUNCH_PAIRS_VAR_REC(1); break;
      case 2: NX_LAUNCH_PAIRS_VAR_REC(2); break;
      case 3: NX_LAUNCH_PAIRS_VAR_REC(3); break;
      case 4: NX_LAUNCH_PAIRS_VAR_REC(4); break;
      default: return hipErrorInvalidValue;
    }
  } else if (units) {
    NX_LAUNCH_PAIRS_VAR(true)
  } else {
    NX_LAUNCH_PAIRS_VAR(false)
  }
  return hipGetLastError();
}

hipError_t launch_qproject_var(const QFormDev& in, const QFormDev& out, const VarDev& vout, int VW,
                               const int32_t* fq, const int32_t* fb, int nfix, const int32_t* uq,
                               const uint64_t* ub, int nunit, hipStream_t st) {
  if (in.chi == 0) return hipSuccess;
  const dim3 g(nblk(in.chi, 4)), b(256);
#define NX_QPV(WW, VV) k_qproject_var<WW, VV><<<g, b, 0, st>>>(in, out, vout, fq, fb, nfix, uq, ub, nunit)
  switch (in.W * 4 + VW) {
    case 1 * 4 + 1: NX_QPV(1, 1); break;
    case 1 * 4 + 2: NX_QPV(1, 2); break;
    case 2 * 4 + 1: NX_QPV(2, 1); break;
    case 2 * 4 + 2: NX_QPV(2, 2); break;
    case 3 * 4 + 1: NX_QPV(3, 1); break;
    case 3 * 4 + 2: NX_QPV(3, 2); break;
    case 4 * 4 + 1: NX_QPV(4, 1); break;
    case 4 * 4 + 2: NX_QPV(4, 2); break;
    default: return hipErrorInvalidValue;
  }
#undef NX_QPV
  return hipGetLastError();
}

hipError_t launch_qproject_var_jobs(int W, int VW, const QProjJobDev* jobs, int njobs, int64_t nblocks,
                                    hipStream_t st) {
  if (njobs <= 0 || nblocks <= 0) return hipSuccess;
  const dim3 g((unsigned)nblocks), b(256);
#define NX_QPVJ(WW, VV) k_qproject_var_jobs<WW, VV><<<g, b, 0, st>>>(jobs, njobs)
  switch (W * 4 + VW) {
    case 1 * 4 + 1: NX_QPVJ(1, 1); break;
    case 1 * 4 + 2: NX_QPVJ(1, 2); break;
    case 2 * 4 + 1: NX_QPVJ(2, 1); break;
    case 2 * 4 + 2: NX_QPVJ(2, 2); break;
    case 3 * 4 + 1: NX_QPVJ(3, 1); break;
    case 3 * 4 + 2: NX_QPVJ(3, 2); break;
    case 4 * 4 + 1: NX_QPVJ(4, 1); break;
    case 4 * 4 + 2: NX_QPVJ(4, 2); break;
    default: return hipErrorInvalidValue;
  }
#undef NX_QPVJ
  return hipGetLastError();
}

hipError_t launch_qproject(const QFormDev& in, const QFormDev& out, const int32_t* uq, const int32_t* ub, int nunit,
                           hipStream_t st) {
  if (in.chi == 0) return hipSuccess;
  const dim3 g(nblk(in.chi, 4)), b(256);
  switch (in.W) {
    case 1: k_qproject<1><<<g, b, 0, st>>>(in, out, uq, ub, nunit); break;
    case 2: k_qproject<2><<<g, b, 0, st>>>(in, out, uq, ub, nunit); break;
    case 3: k_qproject<3><<<g, b, 0, st>>>(in, out, uq, ub, nunit); break;
    case 4: k_qproject<4><<<g, b, 0, st>>>(in, out, uq, ub, nunit); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pairs_sym(const QFormDev& X, const int64_t* offsets, int J, int VW, const int32_t* coord,
                            const uint64_t* removed, int* dmin, int nblocks, hipStream_t st) {
  if (X.chi == 0) return hipSuccess;
  NX_DISPATCH_WV(X.W, VW,
                 k_pairs_sym<WW, VV><<<dim3(nblocks), dim3(256), 0, st>>>(X, offsets, X.chi, J, coord, removed,
                                                                          dmin));
  return hipGetLastError();
}

hipError_t launch_reduce_var(const PairJobDev* jobs, int njobs, int VW, const double* partial, const int32_t* seg_job,
                             int64_t nseg, double* seg_part, double* out, hipStream_t st) {
  if (njobs == 0 || nseg == 0) return hipSuccess;
  if (VW == 1) {
    k_reduce_seg<1><<<dim3((unsigned)nseg), dim3(RED_SEG_NT), 0, st>>>(jobs, seg_job, partial, seg_part);
    k_reduce_var<1><<<dim3(njobs), dim3(RED_VAR_NT), 0, st>>>(jobs, seg_part, out);
  } else if (VW == 2) {
    k_reduce_seg<2><<<dim3((unsigned)nseg), dim3(RED_SEG_NT), 0, st>>>(jobs, seg_job, partial, seg_part);
    k_reduce_var<2><<<dim3(njobs), dim3(RED_VAR_NT), 0, st>>>(jobs, seg_part, out);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int64_t grad_tiles(int64_t chi) {
  const int64_t nb = (chi + GT - 1) / GT;
  return nb * (nb + 1) / 2;
}

int grad_tile_size() { return GT; }

hipError_t launch_grad_var(const QFormDev& X, const VarDev& V, int VW, const double2* c, int nvar, const double* wts,
                           const double* tabs, int nblocks, double2* tile_row, double2* tile_col, double* partial,
                           hipStream_t st) {
  if (X.chi == 0) return hipSuccess;
  const int64_t nt = grad_tiles(X.chi);
  NX_DISPATCH_WV(X.W, VW,
                 k_grad_var<WW, VV><<<dim3(nblocks), dim3(256), 0, st>>>(X, V, c, nvar, wts, tabs, nt, tile_row,
                                                                         tile_col, partial));
  return hipGetLastError();
}

hipError_t launch_grad_reduce(const double2* tile_row, const double2* tile_col, int64_t chi, double scale,
                              int accumulate, double2* y, hipStream_t st) {
  if (chi == 0) return hipSuccess;
  k_grad_reduce<<<dim3(nblk(chi, 256)), dim3(256), 0, st>>>(tile_row, tile_col, chi, scale, accumulate, y);
  return hipGetLastError();
}

hipError_t launch_grad_vals(const double* partial, int64_t nwaves, int stride, int nvar, double scale, double* vals,
                            hipStream_t st) {
  if (nvar == 0) return hipSuccess;
  k_grad_vals<<<dim3(nvar), dim3(256), 0, st>>>(partial, nwaves, stride, scale, vals);
  return hipGetLastError();
}

hipError_t launch_grad_loss(const double* vals, const double* wts, int n, double* loss, hipStream_t st) {
  k_grad_loss<<<dim3(1), dim3(64), 0, st>>>(vals, wts, n, loss);
  return hipGetLastError();
}

}  // namespace nx
