// kernels.hip -- HIP kernels of the CH-form hot path for gfx950.
//
//  k_program   one wavefront per term: applies a list of Clifford gates,
//              Z-basis projections and qubit discards to every term
//              (compile, chi-state gates, projection filter, Pauli images)
//  k_kron      block-diagonal tensor product of two term sets
//  k_gather    copies a subset of terms (clone / stream compaction)
//  k_qform     one wavefront per term: rewrites the CH-form as an affine
//              support {x : C x = rhs} and a Z8 quadratic phase
//              p0 + 2 a.x + 4 sum_{l<m} Q_lm x_l x_m, magnitude 2^{-|v|/2}
//  k_amplitude one lane per bitstring, loops over terms: <x|phi_i> exact
//              (zero, phase, r) and the FP64 sum over coefficients
//  k_pairs     one wavefront per (bra, ket) pair: the exponential sum
//              <phi_a|phi_b> evaluated by GF(2) variable elimination (lane =
//              variable row, pivots broadcast with readlane), FP64 partial
//              sums per wavefront
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "device_form.hpp"
#include "internal.hpp"
#include "pair_core.hpp"

namespace nx {

// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(256) void k_init_zero(FormsDev out) {
  const int64_t t = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  const int lane = threadIdx.x & 63;
  if (t >= out.chi) return;
  Form<W> f;
  f.lane = lane;
  f.n = out.n;
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const int r = 64 * sl + lane;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const u64 diag = (r < out.n && (r >> 6) == w) ? (1ull << (r & 63)) : 0ull;
      f.G[sl][w] = diag;
      f.F[sl][w] = diag;
      f.M[sl][w] = 0ull;
    }
    f.g[sl] = 0;
  }
#pragma unroll
  for (int w = 0; w < W; ++w) { f.v[w] = 0ull; f.s[w] = 0ull; }
  f.ph = 0;
  store_form<W>(f, out, t);
}

// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(256) void k_program(FormsDev in, FormsDev out, const ProgOp* __restrict__ prog,
                                                 int nprog, int4* __restrict__ status) {
  const int64_t t = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  const int lane = threadIdx.x & 63;
  if (t >= in.chi) return;
  Form<W> f;
  load_form<W>(f, in, t, lane);
  int dead = 0, ndet = 0, nrnd = 0, err = 0;
  for (int i = 0; i < nprog; ++i) {
    const int code = uni(prog[i].code), a = uni(prog[i].a), b = uni(prog[i].b);
    switch (code) {
      case OP_H: err = left_h<W>(f, a); break;
      case OP_X: left_x<W>(f, a); break;
      case OP_Y: left_y<W>(f, a); break;
      case OP_Z: left_z<W>(f, a); break;
      case OP_S: left_s_like<W>(f, a, 3); break;
      case OP_SDG: left_s_like<W>(f, a, 1); break;
      case OP_SQRTX:
      case OP_SQRTXDG:
        err = left_h<W>(f, a);
        if (!err) {
          left_s_like<W>(f, a, code == OP_SQRTX ? 3 : 1);
          err = left_h<W>(f, a);
        }
        break;
      case OP_CX: left_cx<W>(f, a, b); break;
      case OP_CZ: left_cz<W>(f, a, b); break;
      case OP_SWAP:
        left_cx<W>(f, a, b);
        left_cx<W>(f, b, a);
        left_cx<W>(f, a, b);
        break;
      case OP_PROJ: {
        int d = 0, dt = 0;
        err = project<W>(f, a, b, &d, &dt);
        if (d) dead = 1;
        else if (dt) ++ndet;
        else ++nrnd;
        break;
      }
      case OP_DISC: err = discard<W>(f, a); break;
      default: err = 99; break;
    }
    if (dead || err) break;
  }
  store_form<W>(f, out, t);
  if (lane == 0) status[t] = make_int4(dead, ndet, nrnd, err);
}

// ---------------------------------------------------------------------------
// one thread per (output term, row)
__device__ __forceinline__ u64 shifted_word(const u64* row, int Win, int nshift, int w) {
  const int ws = nshift >> 6, bs = nshift & 63;
  const int i0 = w - ws, i1 = w - ws - 1;
  u64 r = 0;
  if (i0 >= 0 && i0 < Win) r |= row[i0] << bs;
  if (bs && i1 >= 0 && i1 < Win) r |= row[i1] >> (64 - bs);
  return r;
}

__global__ void k_kron(FormsDev A, FormsDev B, FormsDev O) {
  const int RO = 64 * O.W, RA = 64 * A.W, RB = 64 * B.W;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= O.chi * RO) return;
  const int64_t o = gid / RO;
  const int r = (int)(gid % RO);
  const int64_t ia = o / B.chi, jb = o % B.chi;
  const u64* srcs[3][2] = {{A.G, B.G}, {A.F, B.F}, {A.M, B.M}};
  u64* dsts[3] = {O.G, O.F, O.M};
  for (int m = 0; m < 3; ++m) {
    for (int w = 0; w < O.W; ++w) {
      u64 val = 0;
      if (r < A.n) {
        if (w < A.W) val = srcs[m][0][rowidx(ia, RA, r, A.W, w)];
      } else if (r < A.n + B.n) {
        val = shifted_word(srcs[m][1] + rowidx(jb, RB, r - A.n, B.W, 0), B.W, A.n, w);
      }
      dsts[m][rowidx(o, RO, r, O.W, w)] = val;
    }
  }
  uint8_t gv = 0;
  if (r < A.n) gv = A.g[(size_t)ia * RA + r];
  else if (r < A.n + B.n) gv = B.g[(size_t)jb * RB + (r - A.n)];
  O.g[(size_t)o * RO + r] = gv;
  if (r < O.W) {
    const int w = r;
    u64 av = w < A.W ? A.v[(size_t)ia * A.W + w] : 0ull;
    u64 as = w < A.W ? A.s[(size_t)ia * A.W + w] : 0ull;
    av |= shifted_word(B.v + (size_t)jb * B.W, B.W, A.n, w);
    as |= shifted_word(B.s + (size_t)jb * B.W, B.W, A.n, w);
    O.v[(size_t)o * O.W + w] = av;
    O.s[(size_t)o * O.W + w] = as;
  }
  if (r == 0) O.ph[o] = (uint8_t)((A.ph[ia] + B.ph[jb]) & 7);
}

// one thread per (output term, row): out[o] = in[idx[o]]
__global__ void k_gather(FormsDev in, FormsDev O, const int64_t* __restrict__ idx) {
  const int R = 64 * O.W;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= O.chi * R) return;
  const int64_t o = gid / R;
  const int r = (int)(gid % R);
  const int64_t src = idx ? idx[o] : o;
  for (int w = 0; w < O.W; ++w) {
    const size_t a = rowidx(src, R, r, O.W, w), b = rowidx(o, R, r, O.W, w);
    O.G[b] = in.G[a];
    O.F[b] = in.F[a];
    O.M[b] = in.M[a];
  }
  O.g[(size_t)o * R + r] = in.g[(size_t)src * R + r];
  if (r < O.W) {
    O.v[(size_t)o * O.W + r] = in.v[(size_t)src * O.W + r];
    O.s[(size_t)o * O.W + r] = in.s[(size_t)src * O.W + r];
  }
  if (r == 0) O.ph[o] = in.ph[src];
}

// ---------------------------------------------------------------------------
// Layout of the amplitude kernel's per-term tables (built with the
// quadratic forms at one word per row, else by k_amp_tables).
constexpr int AMP_NCD = 8;  // constraints tested row by row
constexpr int AMP_NT = 256;  // threads per block

// Q-table layout.  Q_l holds only bits m > l, so the entries of nibble
// position b (variables 4b .. 4b+3) are zero below bit 4b + 1, which lies in
// dword b >> 3.  An entry keeps the dwords from there on, the first rounded
// down to an even dword (8-byte aligned entries) except in the last word,
// whose upper half alone is kept for its last 8 positions.  For W = 1 that
// is 8 bytes for positions 0..7 and 4 bytes for 8..15; the 16 entries of a
// position then span 32 or 16 LDS banks once and a wave's lookups at one
// position never conflict.
template <int W>
struct QLay {
  static constexpr __host__ __device__ int d0(int b) {
    return ((b >> 3) == 2 * W - 1) ? (b >> 3) : ((b >> 3) & ~1);
  }
  static constexpr __host__ __device__ int nd(int b) { return 2 * W - d0(b); }
  // byte offset of the 16 entries of position b
  static constexpr __host__ __device__ int off(int b) {
    int o = 0;
    for (int g = 0; g < (b >> 3); ++g) o += 8 * 16 * 4 * nd(8 * g);
    return o + (b & 7) * 16 * 4 * nd(b);
  }
  static constexpr int BYTES = off(16 * W);
};

template <int W>
struct AmpLds {
  static constexpr int NB = 16 * W;  // nibble positions
  static constexpr int NE = NB * 16; // entries of one table
  static constexpr int NCD = AMP_NCD;
  // staging of one term for the table build (words): Q rows, CT rows
  // ([64W][W] each), C rows [NCD][W]
  static constexpr int R_CT = 64 * W * W, R_C = 2 * 64 * W * W, RAW = R_C + NCD * W;
  static constexpr int NPRE = (RAW + AMP_NT - 1) / AMP_NT;
  // one term's tables (bytes; the same image in device memory and in an LDS
  // buffer): TQ (packed), C rows [NCD][W] words, P[0..8] (double2, formed in
  // LDS from the coefficient), TC [NE][W] words (terms with more than NCD
  // constraints only)
  static constexpr int B_C = (QLay<W>::BYTES + 15) & ~15, B_PH = B_C + NCD * W * 8, B_CT = B_PH + 9 * 16,
                       BUF = B_CT + NE * W * 8;
  // 16-byte pieces a copy moves: V_HEAD before P (every term), then V_CT of
  // TC from piece V_CT0 on; at most NPF per thread, the first NPF_REG of
  // them held in registers while the current term is evaluated
  static constexpr int V_HEAD = B_PH / 16, V_CT = NE * W * 8 / 16, V_CT0 = B_CT / 16;
  static constexpr int NPF = (V_HEAD + V_CT + AMP_NT - 1) / AMP_NT;
  static constexpr int NPF_REG = NPF < 4 ? NPF : 4;
  static constexpr size_t BYTES = 2 * (size_t)BUF;         // evaluation: two buffers
  static constexpr size_t BUILD_BYTES = (size_t)RAW * 8 + BUF;  // table build: staging + one buffer
  // piece e of a copy -> its 16-byte index in the table image
  static __device__ __forceinline__ int piece(int e) { return e < V_HEAD ? e : e - V_HEAD + V_CT0; }
};

// ---------------------------------------------------------------------------
// Quadratic-form rewrite of each term.  With u_j = (F_j & ~v) ^ (M_j & v),
// w_j = (M_j & ~v) ^ (F_j & v), c_j = par(F_j & M_j & v), applying X to the
// set bits of x in ascending order (amplitude.rs:36-43) leaves
//   s' = s ^ XOR_{j in x} u_j
//   phase' = phase + sum_{j in x} (gamma_j + 4 beta_j),
//   beta_j = par(s_before_j & w_j) ^ c_j ,
// so <x|phi> = [ (s' & ~v) == 0 ] zeta^{p0 + 2 a.x + 4 sum_{l<m} Q_lm x_l x_m}
// 2^{-|v|/2} with a_j = gamma_j/2 + 2 (par(s & w_j) ^ c_j) mod 4 and
// Q_lm = par(u_l & w_m) for l < m.  The support constraints are the columns
// k of F with v_k = 0: par(F[:,k] & x) = s_k.
// LDS written by the lanes of a wave, then read by other lanes of it
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Transpose of a 64 x 64 bit matrix held one row per lane (bit k of lane
// i's word = entry (i, k)): six rounds of swapping the off-diagonal blocks
// of halving size with the partner lane.
__device__ __forceinline__ u64 transpose64(u64 x, int lane) {
  const u64 masks[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                        0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int j = 32 >> s;
    const u64 m = masks[s];
    const u32 ylo = (u32)__shfl_xor((int)lo32(x), j), yhi = (u32)__shfl_xor((int)hi32(x), j);
    const u64 y = ((u64)yhi << 32) | ylo;
    x = (lane & j) ? ((x & ~m) | ((y & ~m) >> j)) : ((x & m) | ((y & m) << j));
  }
  return x;
}

// tab != nullptr (W = 1 only): also the term's amplitude tables, the image
// k_amp_tables writes, at tab + t AmpLds<1>::BUF
template <int W>
__global__ __launch_bounds__(256) void k_qform(FormsDev in, QFormDev q, char* __restrict__ tab) {
  const int64_t t = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  const int lane = threadIdx.x & 63;
  if (t >= in.chi) return;
  Form<W> f;
  load_form<W>(f, in, t, lane);
  const int n = in.n, R = 64 * W;
  u64 U[W][W], Wm[W][W], Q[W][W];
  int a[W];
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    u64 fm[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      U[sl][w] = (f.F[sl][w] & ~f.v[w]) ^ (f.M[sl][w] & f.v[w]);
      Wm[sl][w] = (f.M[sl][w] & ~f.v[w]) ^ (f.F[sl][w] & f.v[w]);
      fm[w] = f.F[sl][w] & f.M[sl][w];
      Q[sl][w] = 0ull;
    }
    const int c = parand<W>(fm, f.v);
    const int L = parand<W>(f.s, Wm[sl]) ^ c;
    a[sl] = (f.g[sl] + 2 * L) & 3;
  }
  u64 Bm[W][W];
  // W = 1: per-wave rows and nibble tables in LDS
  __shared__ u64 s_wt[4][64], s_tab[4][16 * 16];
  const int wv = threadIdx.x >> 6;
  if constexpr (W == 1) {
    // Q_l = (XOR_{k in u_l} WT_k) restricted to m > l, WT = W transposed:
    // WT rows in LDS, combined four at a time into per-wave nibble tables
    // (16 lookups per lane instead of 64 parities)
    s_wt[wv][lane] = transpose64(Wm[0][0], lane);  // row k: bit m = bit k of w_m
    wave_lds_sync();
    // s_tab[16 b + v] = XOR of WT_{4b + i} for the set bits i of v
#pragma unroll
    for (int e = lane; e < 256; e += 64) {
      const int b = e >> 4, v = e & 15;
      u64 acc = 0ull;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if ((v >> i) & 1) acc ^= s_wt[wv][4 * b + i];
      s_tab[wv][e] = acc;
    }
    wave_lds_sync();
    u64 p = 0ull;
#pragma unroll
    for (int b = 0; b < 16; ++b) p ^= s_tab[wv][16 * b + (int)((U[0][0] >> (4 * b)) & 15ull)];
    // strict upper triangle (m > l); rows at or above n have u = 0
    Q[0][0] = (lane < 63) ? (p & (~0ull << (lane + 1))) : 0ull;
    // B = Q ^ Q^T
    Bm[0][0] = Q[0][0] ^ transpose64(Q[0][0], lane);
  } else {
    for (int m = 0; m < n; ++m) {
      u64 wm[W];
      bcast<W>(Wm, m, wm);
#pragma unroll
      for (int sl = 0; sl < W; ++sl) {
        const int row = 64 * sl + lane;
        const int bit = (row < m) ? parand<W>(U[sl], wm) : 0;
        flipbit<W>(Q[sl], m, bit);
      }
    }
    // B = Q ^ Q^T
#pragma unroll
    for (int sl = 0; sl < W; ++sl)
#pragma unroll
      for (int w = 0; w < W; ++w) Bm[sl][w] = Q[sl][w];
    for (int l = 0; l < n; ++l) {
#pragma unroll
      for (int sw = 0; sw < W; ++sw) {
        const u64 col = __ballot(getbit<W>(Q[sw], l));
#pragma unroll
        for (int sl = 0; sl < W; ++sl)
          if (owns<W>(f, l, sl)) Bm[sl][sw] |= col;
      }
    }
  }
  // constraints
  u64 C[W][W];
#pragma unroll
  for (int sl = 0; sl < W; ++sl)
#pragma unroll
    for (int w = 0; w < W; ++w) C[sl][w] = 0ull;
  u64 rhs[W];
#pragma unroll
  for (int w = 0; w < W; ++w) rhs[w] = 0ull;
  int nc = 0;
  for (int k = 0; k < n; ++k) {
    if (getbit<W>(f.v, k)) continue;
    u64 col[W];
#pragma unroll
    for (int sw = 0; sw < W; ++sw) col[sw] = __ballot(getbit<W>(f.F[sw], k));
#pragma unroll
    for (int sl = 0; sl < W; ++sl)
      if (owns<W>(f, nc, sl)) {
#pragma unroll
        for (int w = 0; w < W; ++w) C[sl][w] = col[w];
      }
    if (getbit<W>(f.s, k)) flipbit<W>(rhs, nc, 1);
    ++nc;
  }
  // reduced echelon form on the rows' lowest variables (same affine
  // subspace: the right-hand sides follow the row operations), so that the
  // pair kernels substitute a bra row without updating the other bra rows
  rref_rows<W>(C, nc, lane, [&](int c, const bool (&hit)[W]) {
    const bool rc = getbit<W>(rhs, c) != 0;
#pragma unroll
    for (int sl = 0; sl < W; ++sl) {
      const u64 m = __ballot(hit[sl]);
      rhs[sl] ^= rc ? m : 0ull;
    }
  });
  // C transposed: row j holds the constraints (bits over c) containing x_j
  u64 CT[W][W];
#pragma unroll
  for (int sl = 0; sl < W; ++sl)
#pragma unroll
    for (int w = 0; w < W; ++w) CT[sl][w] = 0ull;
  if constexpr (W == 1) {
    CT[0][0] = transpose64(C[0][0], lane);
  } else {
    for (int j = 0; j < n; ++j) {
#pragma unroll
      for (int sw = 0; sw < W; ++sw) {
        const u64 col = __ballot(getbit<W>(C[sw], j));
#pragma unroll
        for (int sl = 0; sl < W; ++sl)
          if (owns<W>(f, j, sl)) CT[sl][sw] = col;
      }
    }
  }
  if constexpr (W == 1) {
    if (tab) {
      // the amplitude tables (as amp_build forms them from the stored rows):
      // nibble tables of the Q rows, of the C^T rows when the term has more
      // than AMP_NCD constraints, and the first AMP_NCD constraint rows
      using L = AmpLds<1>;
      using Y = QLay<1>;
      wave_lds_sync();
      s_wt[wv][lane] = Q[0][0];
      s_tab[wv][lane] = CT[0][0];
      wave_lds_sync();
      char* img = tab + (size_t)t * L::BUF;
      const bool bigc = nc > L::NCD;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int e = lane + 64 * i, b = e >> 4, v = e & 15;
        u64 qv = 0ull, cv = 0ull;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((v >> k) & 1) {
            qv ^= s_wt[wv][4 * b + k];
            cv ^= s_tab[wv][4 * b + k];
          }
        // positions 0-7 keep both dwords, 8-15 the upper one
        uint32_t* dq = (uint32_t*)(img + Y::off(b)) + v * Y::nd(b);
        if (Y::d0(b) == 0) {
          dq[0] = lo32(qv);
          dq[1] = hi32(qv);
        } else {
          dq[0] = hi32(qv);
        }
        if (bigc) ((u64*)(img + L::B_CT))[e] = cv;
      }
      if (lane < L::NCD) ((u64*)(img + L::B_C))[lane] = C[0][0];
    }
  }
  // store
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const int r = 64 * sl + lane;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      q.Q[rowidx(t, R, r, W, w)] = Q[sl][w];
      q.B[rowidx(t, R, r, W, w)] = Bm[sl][w];
      q.C[rowidx(t, R, r, W, w)] = C[sl][w];
      q.CT[rowidx(t, R, r, W, w)] = CT[sl][w];
    }
    q.a[(size_t)t * R + r] = (uint8_t)a[sl];
  }
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const u64 lo = __ballot(a[sl] & 1), hi = __ballot((a[sl] >> 1) & 1);
    if (lane == 0) {
      q.alo[(size_t)t * W + sl] = lo;
      q.ahi[(size_t)t * W + sl] = hi;
      q.rhs[(size_t)t * W + sl] = rhs[sl];
    }
  }
  if (lane == 0) {
    int k = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) k += __builtin_popcountll(f.v[w]);
    q.nc[t] = nc;
    q.k[t] = k;
    q.p0[t] = f.ph;
    q.meta[t] = make_int4(nc, k, f.ph, 0);
  }
}

// ---------------------------------------------------------------------------
// Amplitudes.  Each thread evaluates XB bitstrings; blockIdx.y selects a
// chunk of terms [t0, t1).  Per term and string x:
//   support  C x = rhs, tested row by row (parity of C_k & x) when the term
//            has at most AMP_NCD constraints, else from a table of CT rows;
//   phase    p0 + 2 a.x + 4 sum_{l<m} Q_lm x_l x_m (mod 8) with a = alo + 2 ahi
//            = p0 + 2 popc(x & alo) + 4 par(x & (ahi ^ XOR_{l in x} Q_l)),
//            since 4 popc(x & ahi) = 4 par(x & ahi) (mod 8) and Q_l holds
//            only the bits m > l.
// XOR_{l in x} Q_l (and C x = XOR_{l in x} CT_l) are read from per-term
// nibble tables in LDS: for nibble position b (variables 4b .. 4b+3) and
// value v, TQ[b][v] = XOR_{k : v_k = 1} Q_{4b+k}, so a string costs one
// lookup per nibble.  An entry is W words; for W = 1 the 16 entries of one
// position fill the 32 LDS banks once, so the lookups of a wave at one
// position never conflict.  The nibble offsets of a thread's strings do not
// change over the term loop: for W = 1 they are computed once and the two
// table buffers are addressed with immediate offsets (the term loop is
// unrolled by two).  The tables of every term are built once per state by
// k_amp_tables into device memory; an evaluating block copies the next
// term's tables into its other LDS buffer while it evaluates the current
// one (one barrier per term).
// FP64: per term the table P[ph] = Scalar(ph, k).to_complex() * coefficient
// (the reference's to_statevector arithmetic: (cos, sin) * 2^{-k/2}, then the
// complex product) is formed once; P[8] = 0 is selected outside the support.
// Each string adds P[ph] in term order within the chunk; chunks are summed
// in chunk order by k_amp_reduce, which also applies the global factor.

template <int W>
__device__ __forceinline__ void amp_load_raw(const QFormDev& q, int64_t t, u64 (&pre)[AmpLds<W>::NPRE]) {
  using L = AmpLds<W>;
  const size_t base = (size_t)t * (64 * W) * W;  // row 0 of term t in [chi][R][W]
#pragma unroll
  for (int i = 0; i < L::NPRE; ++i) {
    const int e = threadIdx.x + AMP_NT * i;
    pre[i] = 0ull;
    if (e < L::R_CT) pre[i] = q.Q[base + e];
    else if (e < L::R_C) pre[i] = q.CT[base + (e - L::R_CT)];
    else if (e < L::RAW) pre[i] = q.C[base + (e - L::R_C)];
  }
}

template <int W>
__device__ __forceinline__ void amp_store_raw(u64* raw, const u64 (&pre)[AmpLds<W>::NPRE]) {
  using L = AmpLds<W>;
#pragma unroll
  for (int i = 0; i < L::NPRE; ++i) {
    const int e = threadIdx.x + AMP_NT * i;
    if (e < L::RAW) raw[e] = pre[i];
  }
}

// tables of one term from its staged rows (all but P)
template <int W>
__device__ __forceinline__ void amp_build(const u64* raw, char* buf, bool bigc) {
  using L = AmpLds<W>;
  using Y = QLay<W>;
#pragma unroll
  for (int i = 0; i < (L::NE + AMP_NT - 1) / AMP_NT; ++i) {
    const int e = threadIdx.x + AMP_NT * i;
    if (e < L::NE) {
      const int b = e >> 4, v = e & 15;
      u64 qv[W], cv[W];
#pragma unroll
      for (int ww = 0; ww < W; ++ww) {
        qv[ww] = 0ull;
        cv[ww] = 0ull;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((v >> k) & 1) {
            qv[ww] ^= raw[(4 * b + k) * W + ww];
            cv[ww] ^= raw[L::R_CT + (4 * b + k) * W + ww];
          }
      }
      const int d0 = Y::d0(b), nd = 2 * W - d0;
      uint32_t* dq = (uint32_t*)(buf + Y::off(b)) + v * nd;
#pragma unroll
      for (int d = 0; d < 2 * W; ++d)
        if (d >= d0) dq[d - d0] = (uint32_t)(qv[d >> 1] >> (32 * (d & 1)));
      if (bigc) {
        u64* dc = (u64*)(buf + L::B_CT) + e * W;
#pragma unroll
        for (int ww = 0; ww < W; ++ww) dc[ww] = cv[ww];
      }
    }
  }
  if (threadIdx.x < L::NCD * W) ((u64*)(buf + L::B_C))[threadIdx.x] = raw[L::R_C + threadIdx.x];
}

// P[ph] = (cos, sin)(ph pi / 4) * nrm * c for ph < 8, P[8] = 0; c and nrm =
// 2^{-k/2} of the term (threads 0..8)
__device__ __forceinline__ void amp_phase_table(char* buf_ph, double2 c, double nrm, const double* __restrict__ tabs) {
  if (threadIdx.x < 9) {
    double tr = 0.0, ti = 0.0;
    if (threadIdx.x < 8) {
      const double zr = tabs[TAB_COS + threadIdx.x] * nrm, zi = tabs[TAB_SIN + threadIdx.x] * nrm;
      tr = zr * c.x - zi * c.y;
      ti = zr * c.y + zi * c.x;
    }
    ((double2*)buf_ph)[threadIdx.x] = make_double2(tr, ti);
  }
}

// the tables of every term (one block per term, grid-strided) into tab:
// chi images of AmpLds<W>::BUF bytes (P and, for terms with at most NCD
// constraints, TC left unwritten)
template <int W>
__global__ __launch_bounds__(AMP_NT) void k_amp_tables(QFormDev q, uint4* __restrict__ tab) {
  using L = AmpLds<W>;
  extern __shared__ __align__(16) u64 amp_lds[];
  u64* raw = amp_lds;
  char* buf = (char*)(amp_lds + L::RAW);
  for (int64_t t = blockIdx.x; t < q.chi; t += gridDim.x) {
    u64 pre[L::NPRE];
    amp_load_raw<W>(q, t, pre);
    amp_store_raw<W>(raw, pre);
    const bool bigc = uni(q.nc[t]) > L::NCD;
    __syncthreads();
    amp_build<W>(raw, buf, bigc);
    __syncthreads();
    const int ncp = L::V_HEAD + (bigc ? L::V_CT : 0);
    uint4* dst = tab + (size_t)t * (L::BUF / 16);
#pragma unroll
    for (int i = 0; i < L::NPF; ++i) {
      const int e = threadIdx.x + AMP_NT * i;
      if (e < ncp) dst[L::piece(e)] = ((const uint4*)buf)[L::piece(e)];
    }
    __syncthreads();
  }
}

// one packed Q entry of ND dwords
template <int ND>
__device__ __forceinline__ void ld_entry(const char* p, uint32_t (&e)[2 * 4]) {
  if constexpr (ND == 1) {
    e[0] = *(const uint32_t*)p;
  } else if constexpr (ND == 2) {
    const uint2 a = *(const uint2*)p;
    e[0] = a.x;
    e[1] = a.y;
  } else {
#pragma unroll
    for (int i = 0; i < ND; i += 2) {
      const uint2 a = ((const uint2*)p)[i >> 1];
      e[i] = a.x;
      e[i + 1] = a.y;
    }
  }
}

// 16-byte copies of the amplitude tables (a plain vector type, kept in
// registers across the term evaluation)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int V>
struct IConst {
  static constexpr int value = V;
};

// one block of an amplitude launch: string column bx, term chunk by
template <int W, int XB>
__device__ __forceinline__ void amp_block(const QFormDev& q, const uint4* __restrict__ tab,
                                          const double2* __restrict__ coef,
                                          const u64* __restrict__ bits, int64_t nb, int64_t nchunks,
                                          const double* __restrict__ tabs, double2* __restrict__ out,
                                          uint8_t* __restrict__ per_term, int enumerate, int final_g, double gre,
                                          double gim, int64_t bx, int64_t by) {
  using L = AmpLds<W>;
  using Y = QLay<W>;
  constexpr int NB = L::NB;
  constexpr bool PREADDR = (W == 1);
  extern __shared__ __align__(16) u64 amp_lds[];
  char* bufs = (char*)amp_lds;  // two table buffers of L::BUF bytes
  const int64_t t0 = by * q.chi / nchunks;
  const int64_t t1 = (by + 1) * q.chi / nchunks;
  int64_t bidx[XB];
  bool active[XB];
  u64 x[XB][W];
#pragma unroll
  for (int j = 0; j < XB; ++j) {
    bidx[j] = (bx * XB + j) * blockDim.x + threadIdx.x;
    active[j] = bidx[j] < nb;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (enumerate) x[j][w] = (w == 0 && active[j]) ? (u64)bidx[j] : 0ull;
      else x[j][w] = active[j] ? bits[(size_t)bidx[j] * W + w] : 0ull;
    }
  }
  auto nib = [&](int j, int b) { return (uint32_t)((x[j][b >> 4] >> (4 * (b & 15))) & 15ull); };
  // the same from x_j ^ z with z = 0 varying per term (z = t >> 62 for a
  // term index t < 2^62): offsets derived from it are recomputed per term
  // instead of being hoisted out of the term loop into registers
  auto nib_local = [&](const u64(&xl)[W], int b) { return (uint32_t)((xl[b >> 4] >> (4 * (b & 15))) & 15ull); };
  auto opaque_x = [&](int j, u64 z, u64(&xl)[W]) {
#pragma unroll
    for (int w = 0; w < W; ++w) xl[w] = x[j][w] ^ z;
  };
  // byte offset of string j's Q entry at nibble position b (W == 1)
  uint32_t off[PREADDR ? XB : 1][PREADDR ? NB : 1];
  if constexpr (PREADDR) {
#pragma unroll
    for (int j = 0; j < XB; ++j)
#pragma unroll
      for (int b = 0; b < NB; ++b) off[j][b] = nib(j, b) * (uint32_t)(4 * Y::nd(b));
  }
  double are[XB], aim[XB];
#pragma unroll
  for (int j = 0; j < XB; ++j) { are[j] = 0.0; aim[j] = 0.0; }
  // table pieces of term t: fetched into registers (pieces past ncp load
  // piece 0 and are not stored), then stored into an LDS buffer, the
  // pieces past NPF_REG straight from device memory
  auto tab_fetch = [&](int64_t t, int ncp, u32x4(&pf)[L::NPF_REG]) {
    const u32x4* src = (const u32x4*)tab + (size_t)t * (L::BUF / 16);
#pragma unroll
    for (int i = 0; i < L::NPF_REG; ++i) {
      const int e = threadIdx.x + AMP_NT * i;
      pf[i] = src[e < ncp ? L::piece(e) : 0];
    }
  };
  auto tab_put = [&](char* buf, int64_t t, int ncp, const u32x4(&pf)[L::NPF_REG]) {
    const u32x4* src = (const u32x4*)tab + (size_t)t * (L::BUF / 16);
    u32x4* dst = (u32x4*)buf;
#pragma unroll
    for (int i = 0; i < L::NPF_REG; ++i) {
      const int e = threadIdx.x + AMP_NT * i;
      if (e < ncp) dst[L::piece(e)] = pf[i];
    }
#pragma unroll
    for (int i = L::NPF_REG; i < L::NPF; ++i) {
      const int e = threadIdx.x + AMP_NT * i;
      if (e < ncp) dst[L::piece(e)] = src[L::piece(e)];
    }
  };
  auto tab_pieces = [&](int nc) { return L::V_HEAD + (nc > L::NCD ? L::V_CT : 0); };
  if (t0 < t1) {
    const int ncp0 = tab_pieces(uni(q.nc[t0]));
    u32x4 pf0[L::NPF_REG];
    tab_fetch(t0, ncp0, pf0);
    tab_put(bufs, t0, ncp0, pf0);
    amp_phase_table(bufs + L::B_PH, coef[t0], tabs[TAB_POW + uni(q.k[t0])], tabs);
  }
  __syncthreads();

  // acc (dwords) = init ^ XOR_{l in x_j} Q_l from the packed Q table of
  // buffer BI, two positions at a time
  auto qlook = [&](auto bic, int j, u64 z, uint32_t(&acc)[2 * W]) {
    constexpr int BI = decltype(bic)::value;
    const char* tb = bufs + BI * L::BUF;
    u64 xl[W];
    if constexpr (!PREADDR) opaque_x(j, z, xl);
#pragma unroll
    for (int b = 0; b < NB; b += 2) {
      const int nd = Y::nd(b), d0 = Y::d0(b);
      uint32_t o0, o1;
      if constexpr (PREADDR) {
        o0 = off[j][b];
        o1 = off[j][b + 1];
      } else {
        o0 = nib_local(xl, b) * (uint32_t)(4 * nd);
        o1 = nib_local(xl, b + 1) * (uint32_t)(4 * nd);
      }
      const char* p0 = tb + Y::off(b) + o0;
      const char* p1 = tb + Y::off(b + 1) + o1;
      uint32_t e0[8], e1[8];
      if (nd == 1) {
        ld_entry<1>(p0, e0);
        ld_entry<1>(p1, e1);
      } else if (nd == 2) {
        ld_entry<2>(p0, e0);
        ld_entry<2>(p1, e1);
      } else if (nd == 4) {
        ld_entry<4>(p0, e0);
        ld_entry<4>(p1, e1);
      } else if (nd == 6) {
        ld_entry<6>(p0, e0);
        ld_entry<6>(p1, e1);
      } else {
        ld_entry<8>(p0, e0);
        ld_entry<8>(p1, e1);
      }
#pragma unroll
      for (int d = 0; d < 2 * W; ++d)
        if (d >= d0) acc[d] = xor3(acc[d], e0[(d - d0) & 7], e1[(d - d0) & 7]);
    }
  };
  // y = C x_j from the full CT table of buffer BI (terms with many constraints)
  auto ctlook = [&](auto bic, int j, u64 z, u64(&y)[W]) {
    constexpr int BI = decltype(bic)::value;
    const u64* tc = (const u64*)(bufs + BI * L::BUF + L::B_CT);
    u64 xl[W];
    opaque_x(j, z, xl);
#pragma unroll
    for (int w = 0; w < W; ++w) y[w] = 0ull;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const u64* e = tc + ((size_t)b * 16 + nib_local(xl, b)) * W;
#pragma unroll
      for (int w = 0; w < W; ++w) y[w] ^= e[w];
    }
  };

  // wave-uniform values of a term (scalar registers), read one term ahead
  struct TermU {
    int k, p0, nc;
    u64 alo[W], ahi[W], rhs[W];
  };
  auto load_u = [&](int64_t t, TermU& u) {
    u.k = uni(q.k[t]);
    u.p0 = uni(q.p0[t]);
    u.nc = uni(q.nc[t]);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      u.alo[w] = uni64(q.alo[(size_t)t * W + w]);
      u.ahi[w] = uni64(q.ahi[(size_t)t * W + w]);
      u.rhs[w] = uni64(q.rhs[(size_t)t * W + w]);
    }
  };
  TermU cur;
  if (t0 < t1) load_u(t0, cur);

  // evaluate term t from buffer BI, copy term t + 1 into the other buffer
  auto run = [&](auto bic, int64_t t) {
    constexpr int BI = decltype(bic)::value;
    const bool more = t + 1 < t1;
    double2 cn = make_double2(0.0, 0.0);
    TermU nxt = cur;
    u32x4 pf[L::NPF_REG];
    int ncp = 0;
    if (more) {  // in flight during the evaluation below
      load_u(t + 1, nxt);
      cn = coef[t + 1];
      ncp = tab_pieces(nxt.nc);
      tab_fetch(t + 1, ncp, pf);
    }
    const char* buf = bufs + BI * L::BUF;
    const int p0 = cur.p0, nc = cur.nc;
    const bool bigc = nc > L::NCD;
    const u64 z = (u64)(t >> 62);
    // support test: bad[j] & 1 is set outside the support
    uint32_t bad[XB];
#pragma unroll
    for (int j = 0; j < XB; ++j) bad[j] = 0u;
    if (!bigc) {
      const u64* crow = (const u64*)(buf + L::B_C);
      for (int kk = 0; kk < nc; ++kk) {
        u64 cr[W];
#pragma unroll
        for (int w = 0; w < W; ++w) cr[w] = crow[kk * W + w];
        const uint32_t rk = (uint32_t)(cur.rhs[0] >> kk) & 1u;
#pragma unroll
        for (int j = 0; j < XB; ++j) {
          uint32_t p = rk;
#pragma unroll
          for (int w = 0; w < W; ++w) {
            p += (uint32_t)__builtin_popcount(lo32(x[j][w]) & lo32(cr[w]));
            p += (uint32_t)__builtin_popcount(hi32(x[j][w]) & hi32(cr[w]));
          }
          bad[j] |= p;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < XB; ++j) {
      if (bigc) {
        u64 y[W];
        ctlook(bic, j, z, y);
#pragma unroll
        for (int w = 0; w < W; ++w) bad[j] |= (y[w] != cur.rhs[w]) ? 1u : 0u;
      }
      // acc = ahi ^ XOR_{l in x} Q_l
      uint32_t acc[2 * W];
#pragma unroll
      for (int w = 0; w < W; ++w) {
        acc[2 * w] = lo32(cur.ahi[w]);
        acc[2 * w + 1] = hi32(cur.ahi[w]);
      }
      qlook(bic, j, z, acc);
      uint32_t pa = 0u, pq = 0u;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        pa += (uint32_t)__builtin_popcountll(x[j][w] & cur.alo[w]);
        pq += (uint32_t)__builtin_popcount(lo32(x[j][w]) & acc[2 * w]);
        pq += (uint32_t)__builtin_popcount(hi32(x[j][w]) & acc[2 * w + 1]);
      }
      const int ph = (int)(((uint32_t)p0 + 2u * pa + 4u * pq) & 7u);
      const bool outside = (bad[j] & 1u) != 0u;
      if (per_term && active[j]) per_term[(size_t)t * nb + bidx[j]] = outside ? (uint8_t)0xFF : (uint8_t)ph;
      const double2 d = ((const double2*)(buf + L::B_PH))[outside ? 8 : ph];
      are[j] = are[j] + d.x;
      aim[j] = aim[j] + d.y;
    }
    if (more) {
      char* nbuf = bufs + (1 - BI) * L::BUF;
      tab_put(nbuf, t + 1, ncp, pf);
      amp_phase_table(nbuf + L::B_PH, cn, tabs[TAB_POW + nxt.k], tabs);
    }
    __syncthreads();
    cur = nxt;
  };

  for (int64_t t = t0; t < t1; t += 2) {
    run(IConst<0>{}, t);
    if (t + 1 < t1) run(IConst<1>{}, t + 1);
  }
#pragma unroll
  for (int j = 0; j < XB; ++j) {
    if (!active[j]) continue;
    if (final_g) {  // * global_factor (statevector.rs:23)
      out[bidx[j]] = make_double2(are[j] * gre - aim[j] * gim, are[j] * gim + aim[j] * gre);
    } else {
      out[(size_t)by * nb + bidx[j]] = make_double2(are[j], aim[j]);
    }
  }
}

template <int W, int XB>
__global__ __launch_bounds__(AMP_NT) void k_amplitude(QFormDev q, const uint4* __restrict__ tab,
                                                   const double2* __restrict__ coef,
                                                   const u64* __restrict__ bits, int64_t nb, int64_t nchunks,
                                                   const double* __restrict__ tabs, double2* __restrict__ out,
                                                   uint8_t* __restrict__ per_term, int enumerate, int final_g,
                                                   double gre, double gim) {
  amp_block<W, XB>(q, tab, coef, bits, nb, nchunks, tabs, out, per_term, enumerate, final_g, gre, gim, blockIdx.x,
                   blockIdx.y);
}

// sum of the per-chunk partial amplitudes in chunk order, then * global_factor
// (the loads of eight chunks, the last group's too, issued before their adds)
__global__ void k_amp_reduce(const double2* __restrict__ part, int64_t nb, int nchunks, double2* __restrict__ out,
                             double gre, double gim) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double r = 0.0, i = 0.0;
  for (int c = 0; c < nchunks; c += 8) {
    double2 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = c + k < nchunks ? part[(size_t)(c + k) * nb + b] : make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (c + k < nchunks) {
        r = r + v[k].x;
        i = i + v[k].y;
      }
    }
  }
  out[b] = make_double2(r * gre - i * gim, r * gim + i * gre);
}

// Coefficients: exact Scalars when cx.scalar != 0, else complex128 values.
// Work decomposition: fixed index f in [0, nfixed) pairs with the other
// indices o in [lo_f, nother), lo_f = f (triangle) or 0 (full), cut into
// units of J; the row of a unit comes from a closed form (unit_row).  mode
// bit 0: fixed index is the ket (else the bra); mode bit 1: Hermitian real
// sum (diagonal once, off-diagonal twice, norm.rs:16-27 and
// exp_value.rs:38-58) else the complex sum (inner_product.rs:12-17).
// 2^{-r/2}: the host table entry pow(2, -r/2), formed exactly as
// ldexp(pow(2, -1/2) or 1, -floor(r/2)) (equal to the table for every |r| <=
// TAB_RMAX; a test compares the two)
__device__ __forceinline__ double pow_half(int r) {
  return ldexp((r & 1) ? 0x1.6a09e667f3bcdp-1 : 1.0, -(r >> 1));
}

// Row f of unit u and the first unit of that row (wave-uniform).  Full
// rows hold U = ceil(nother / J) units each.  Triangle rows (nfixed ==
// nother = N) hold ceil((N - f) / J) units starting at S(N) - S(N - f), S =
// tri_units_upto: with t = S(N) - u, m = N - f is the smallest m with
// S(m) >= t.  S(m) = m (m + J) / (2J) + r (J - r) / (2J) (m = qJ + r), so the
// root of m (m + J) = 2 J t is at most one above m; the loops settle it.
// units of J = PAIR_UNIT pairs (a compile-time power of two: the unit
// arithmetic is shifts and masks)
constexpr int PAIR_UNIT = 8, PAIR_UNIT_LOG = 3;
__device__ __forceinline__ int64_t tri_units8(int64_t M) {
  const int64_t q = M >> PAIR_UNIT_LOG, r = M & (PAIR_UNIT - 1);
  return ((q * (q + 1)) << (PAIR_UNIT_LOG - 1)) + r * (q + 1);
}
__device__ __forceinline__ int64_t unit_row(int64_t u, int64_t nother, int triangle, int64_t* first) {
  if (!triangle) {
    const int64_t U = (nother + PAIR_UNIT - 1) >> PAIR_UNIT_LOG;
    int64_t f = (int64_t)((double)u / (double)U);
    f = (int64_t)uni64((uint64_t)f);
    while (f * U > u) --f;
    while ((f + 1) * U <= u) ++f;
    *first = f * U;
    return f;
  }
  const int64_t SN = tri_units8(nother);
  const int64_t t = SN - u;
  const double dJ = (double)PAIR_UNIT;
  int64_t m = (int64_t)ceil(0.5 * (sqrt(dJ * dJ + 8.0 * dJ * (double)t) - dJ));
  m = (int64_t)uni64((uint64_t)m);
  if (m > nother) m = nother;
  if (m < 1) m = 1;
  while (m > 1 && tri_units8(m - 1) >= t) --m;
  while (tri_units8(m) < t) ++m;
  *first = SN - tri_units8(m);
  return nother - m;
}

// Re of zeta^K 2^{-R/2} (times 2^x2) added exactly to A (dyadic part) or B
// (sqrt 2 part): cos(K pi/4) is +-1, +-sqrt2/2 or 0 and 2^{-R/2} is 2^{-R/2}
// or sqrt2 2^{-(R+1)/2}, so the term is +-2^e in A when K and R have the
// same parity, in B otherwise, with e = -ceil(R/2) - (K odd, R even).
__device__ __forceinline__ void add_re_exact(int K, int R, int x2, double& A, double& B) {
  K &= 7;
  if ((K & 3) == 2) return;  // cos(pi/2), cos(3pi/2)
  const bool neg = K >= 3 && K <= 5;
  const int kodd = K & 1, rodd = R & 1;
  const int e = -((R + 1) >> 1) - (kodd & (rodd ^ 1)) + x2;
  const double v = ldexp(neg ? -1.0 : 1.0, e);
  if (kodd == rodd) A = A + v;
  else B = B + v;
}

// The same, evaluated in vector registers (every lane alike) and without
// branches: a term that is off (a zero coefficient) or has cos = 0 adds 0.0,
// which leaves the exact sums unchanged.
__device__ __forceinline__ void add_re_exact_v(int K, int R, int x2, bool live, double& A, double& B) {
  K &= 7;
  const bool on = live && (K & 3) != 2;
  const bool neg = K >= 3 && K <= 5;
  const int kodd = K & 1, rodd = R & 1;
  const int e = -((R + 1) >> 1) - (kodd & (rodd ^ 1)) + x2;
  const double v = on ? ldexp(neg ? -1.0 : 1.0, e) : 0.0;
  const bool toA = kodd == rodd;
  A = A + (toA ? v : 0.0);
  B = B + (toA ? 0.0 : v);
}

// a wave-uniform value copied into a vector register, so that what is
// computed from it issues on the vector unit (the pair kernel is bound by
// scalar issue)
__device__ __forceinline__ int vreg(int x) {
  int r;
  asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
  return r;
}

// SAME: X and Y (and cx and cy) are the same state's forms and coefficients
// (norm, exp_value): one set of table pointers is kept in scalar registers,
// which leaves room for 8 waves per SIMD at W = 1.
// MODE: the mode argument when it is 0..3 at compile time (no per-pair role
// selects), or -1 to read it at run time.  CK: 1 (exact Scalar
// coefficients) or 0 (complex128) when known at compile time, with no
// exact_out; -1 to read both at run time.  X and Y are written by k_qform
// (constraint rows in reduced echelon form, and the meta records), not
// projected forms.
template <int W, bool SAME, int MODE, int CK>
__global__ __launch_bounds__(256, (SAME && W == 1) ? 8 : 1) void k_pairs(QFormDev X, QFormDev Y,
                                               int64_t nfixed, int64_t nother, int triangle, int J, int mode_rt,
                                               CoefDev cx, CoefDev cy, const double* __restrict__ tabs,
                                               double4* __restrict__ partial, uint32_t* __restrict__ exact_out,
                                               int64_t exact_ld) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  const int64_t nunits =
      triangle ? tri_units8(nother) : nfixed * ((nother + PAIR_UNIT - 1) >> PAIR_UNIT_LOG);
  const int n = X.n;
  u64 valid[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int hi = n - 64 * w;
    valid[w] = hi >= 64 ? ~0ull : (hi <= 0 ? 0ull : ((1ull << hi) - 1ull));
  }
  const int mode = MODE >= 0 ? MODE : mode_rt;
  const bool fixed_is_ket = mode & 1;
  const bool herm = (mode >> 1) & 1;
  const bool exact = CK >= 0 ? CK == 1 : cx.scalar != 0;
  if (CK >= 0) exact_out = nullptr;
  // cos, sin of k pi / 4 (the host tables) in LDS, for complex128
  // coefficients: read with a lane address instead of a scalar load
  __shared__ double s_cs[16];
  if (threadIdx.x < 16) s_cs[threadIdx.x] = tabs[TAB_COS + threadIdx.x];
  __syncthreads();
  double a_re = 0.0, a_im = 0.0, b_re = 0.0, b_im = 0.0;
  for (int64_t u = wave; u < nunits; u += nwaves) {
    int64_t first = 0;
    // term indices fit 32 bits (the forms' 32-bit row offsets already
    // require chi * 64W * W < 2^32)
    const int fx = (int)unit_row(u, nother, triangle, &first);
    const int base = triangle ? fx : 0;
    const int o_lo = base + ((int)(u - first) << PAIR_UNIT_LOG);
    const int o_hi = (o_lo + PAIR_UNIT < (int)nother) ? o_lo + PAIR_UNIT : (int)nother;
    const QFormDev& qf = (SAME || !fixed_is_ket) ? X : Y;
    const QFormDev& qo = (SAME || fixed_is_ket) ? X : Y;
    u64 Bf[W][W], Cf[W][W], rf[W];
    int af[W];
    load_qrows<W>(qf, fx, lane, Bf, af, Cf, rf);
    const int4 mf = qf.meta[fx];
    const int ncf = uni(mf.x), kf = uni(mf.y), pf = uni(mf.z);
    const CoefDev& cf = (SAME || !fixed_is_ket) ? cx : cy;
    const CoefDev& co = (SAME || fixed_is_ket) ? cx : cy;
    const bool scf = CK >= 0 ? CK == 1 : cf.scalar != 0, sco = CK >= 0 ? CK == 1 : co.scalar != 0;
    const int32_t csf = scf ? uni(cf.cs[fx]) : 0;
    const double2 cvf = scf ? make_double2(0.0, 0.0) : cf.c[fx];
    // the unit's other terms: lane j holds the record, right-hand sides and
    // packed coefficient of term o_lo + j (one vector load each per unit
    // instead of scalar loads and address arithmetic per pair)
    const int jl = lane < o_hi - o_lo ? lane : 0;
    const int4 mo_l = qo.meta[o_lo + jl];
    u64 ro_l[W];
#pragma unroll
    for (int w = 0; w < W; ++w) ro_l[w] = qo.rhs[(uint32_t)(o_lo + jl) * W + w];
    const int32_t cso_l = sco ? co.cs[o_lo + jl] : 0;
    for (int o = o_lo; o < o_hi; ++o) {
      const int jo = o - o_lo;
      u64 Bo[W][W], Co[W][W], ro[W];
      int ao[W];
      load_qrows<W, false>(qo, o, lane, Bo, ao, Co, ro);
#pragma unroll
      for (int w = 0; w < W; ++w) ro[w] = rl64(ro_l[w], jo);
      const int nco = rl32(mo_l.x, jo), ko = rl32(mo_l.y, jo), po = rl32(mo_l.z, jo);
      // the other term's coefficient, read now and used after the elimination
      const int32_t cso = sco ? rl32(cso_l, jo) : 0;
      const double2 cvo = sco ? make_double2(0.0, 0.0) : co.c[o];
      const int bra = fixed_is_ket ? o : fx;
      const int ket = fixed_is_ket ? fx : o;
      PairRegs<W> P;
      P.lane = lane;
      P.nodiag = ~(1ull << lane);
#pragma unroll
      for (int sl = 0; sl < W; ++sl) {
#pragma unroll
        for (int w = 0; w < W; ++w) P.B[sl][w] = Bf[sl][w] ^ Bo[sl][w];
        // a = a_ket - a_bra
        P.a[sl] = fixed_is_ket ? ((af[sl] - ao[sl]) & 3) : ((ao[sl] - af[sl]) & 3);
      }
      // constraints: bra rows in ca, ket rows in cb
#pragma unroll
      for (int sl = 0; sl < W; ++sl)
#pragma unroll
        for (int w = 0; w < W; ++w) {
          P.ca[sl][w] = fixed_is_ket ? Co[sl][w] : Cf[sl][w];
          P.cb[sl][w] = fixed_is_ket ? Cf[sl][w] : Co[sl][w];
        }
#pragma unroll
      for (int w = 0; w < W; ++w) {
        P.rhsa[w] = fixed_is_ket ? ro[w] : rf[w];
        P.rhsb[w] = fixed_is_ket ? rf[w] : ro[w];
        P.rem[w] = valid[w];
      }
      const int pbra = fixed_is_ket ? po : pf, pket = fixed_is_ket ? pf : po;
      P.c0 = (pket - pbra) & 7;
      P.mag = kf + ko;
      const int nca = fixed_is_ket ? nco : ncf, ncb = fixed_is_ket ? ncf : nco;
      int zero = 0, K = 0, Rr = 0;
      pair_eval<W, true>(P, nca, ncb, &zero, &K, &Rr);
      if (exact_out && lane == 0)
        exact_out[(size_t)bra * exact_ld + ket] = zero ? 0xFFFFFFFFu : ((uint32_t)K | ((uint32_t)Rr << 8));
      if (zero) continue;
      if (exact) {
        // c_bra* c_ket <bra|ket> = zeta^{K'} 2^{-R'/2}, summed exactly (a
        // Scalar::Zero coefficient, cs < 0, adds nothing)
        const int32_t sb = vreg(fixed_is_ket ? cso : csf), sk = vreg(fixed_is_ket ? csf : cso);
        const bool live = (sb | sk) >= 0;
        const int Kp = vreg(K) + cs_k(sk) - cs_k(sb);
        const int Rp = vreg(Rr) + cs_r(sb) + cs_r(sk);
        if (herm) {
          add_re_exact_v(Kp, Rp, vreg(bra) != ket ? 1 : 0, live, a_re, b_re);
        } else {
          add_re_exact_v(Kp, Rp, 0, live, a_re, b_re);
          add_re_exact_v(Kp - 2, Rp, 0, live, a_im, b_im);  // Im zeta^K = Re zeta^{K-2}
        }
      } else {
        // (c_bra.conj() * c_ket).into() * <bra|ket> as num-complex forms it
        // (inner_product.rs:44-47: omega_a * omega_b = 1)
        const double nrm = pow_half(Rr);
        const double ir = s_cs[K] * nrm, ii = s_cs[8 + K] * nrm;
        const double2 b0 = fixed_is_ket ? cvo : cvf, k0 = fixed_is_ket ? cvf : cvo;
        const double br = b0.x, bi = -b0.y;
        const double cr_ = br * k0.x - bi * k0.y, ci_ = br * k0.y + bi * k0.x;
        const double tr = cr_ * ir - ci_ * ii, ti = cr_ * ii + ci_ * ir;
        if (herm) {
          a_re = (bra == ket) ? a_re + tr : a_re + (tr + tr);
        } else {
          a_re = a_re + tr;
          a_im = a_im + ti;
        }
      }
    }
  }
  // the block's four waves, in wave order, to partial[block]
  __shared__ double4 s_part[4];
  if (lane == 0) s_part[threadIdx.x >> 6] = make_double4(a_re, a_im, b_re, b_im);
  __syncthreads();
  if (threadIdx.x == 0) {
    double4 r = s_part[0];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      r.x += s_part[w].x;
      r.y += s_part[w].y;
      r.z += s_part[w].z;
      r.w += s_part[w].w;
    }
    partial[blockIdx.x] = r;
  }
}

// single-block reduction of the per-block partial sums in a fixed order:
// thread i sums partial[i], partial[i + 1024], ... in index order (eight
// loads in flight per group), then a fixed tree over the threads.  With
// norm_out, also sqrt((A_re + sqrt2 B_re) g2) there (the norm, as the host
// forms it from the same sums)
__global__ __launch_bounds__(1024) void k_reduce(const double4* __restrict__ partial, int64_t n,
                                                 double4* __restrict__ out, double* __restrict__ norm_out,
                                                 double g2) {
  __shared__ double4 sh[1024];
  double4 r = make_double4(0.0, 0.0, 0.0, 0.0);
  // groups of eight loads (the last, partial group too) issued before their adds
  for (int64_t k = threadIdx.x; k < n; k += 8 * 1024) {
    double4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = k + u * 1024 < n ? partial[k + u * 1024] : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (k + u * 1024 < n) {
        r.x += v[u].x;
        r.y += v[u].y;
        r.z += v[u].z;
        r.w += v[u].w;
      }
    }
  }
  sh[threadIdx.x] = r;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double4 v = sh[threadIdx.x + s];
      double4 w = sh[threadIdx.x];
      w.x += v.x;
      w.y += v.y;
      w.z += v.z;
      w.w += v.w;
      sh[threadIdx.x] = w;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (out) out[0] = sh[0];
    if (norm_out) norm_out[0] = sqrt((sh[0].x + 1.4142135623730951 * sh[0].z) * g2);
  }
}

// ---------------------------------------------------------------------------
// launchers

#define NX_DISPATCH_W(W_, ...)                             \
  switch (W_) {                                            \
    case 1: { constexpr int WW = 1; __VA_ARGS__; } break;  \
    case 2: { constexpr int WW = 2; __VA_ARGS__; } break;  \
    case 3: { constexpr int WW = 3; __VA_ARGS__; } break;  \
    case 4: { constexpr int WW = 4; __VA_ARGS__; } break;  \
    default: return hipErrorInvalidValue;                  \
  }

static inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

hipError_t launch_init_zero(const FormsDev& out, hipStream_t st) {
  if (out.chi == 0) return hipSuccess;
  NX_DISPATCH_W(out.W, k_init_zero<WW><<<dim3(blocks_for(out.chi, 4)), dim3(256), 0, st>>>(out));
  return hipGetLastError();
}

hipError_t launch_program(const FormsDev& in, const FormsDev& out, const ProgOp* prog, int nprog, int4* status,
                          hipStream_t st) {
  if (in.chi == 0) return hipSuccess;
  NX_DISPATCH_W(in.W, k_program<WW><<<dim3(blocks_for(in.chi, 4)), dim3(256), 0, st>>>(in, out, prog, nprog, status));
  return hipGetLastError();
}

hipError_t launch_kron(const FormsDev& A, const FormsDev& B, const FormsDev& O, hipStream_t st) {
  const int64_t total = O.chi * 64 * O.W;
  if (total == 0) return hipSuccess;
  k_kron<<<dim3(blocks_for(total, 256)), dim3(256), 0, st>>>(A, B, O);
  return hipGetLastError();
}

hipError_t launch_gather(const FormsDev& in, const FormsDev& O, const int64_t* idx, hipStream_t st) {
  const int64_t total = O.chi * 64 * O.W;
  if (total == 0) return hipSuccess;
  k_gather<<<dim3(blocks_for(total, 256)), dim3(256), 0, st>>>(in, O, idx);
  return hipGetLastError();
}

hipError_t launch_qform(const FormsDev& in, const QFormDev& q, hipStream_t st, void* tab) {
  if (tab && in.W != 1) return hipErrorInvalidValue;
  if (in.chi == 0) return hipSuccess;
  NX_DISPATCH_W(in.W, k_qform<WW><<<dim3(blocks_for(in.chi, 4)), dim3(256), 0, st>>>(in, q, (char*)tab));
  return hipGetLastError();
}

// resident blocks of k_amplitude<W, XB> on the device (occupancy x CUs),
// queried once per instance
template <int W, int XB>
static int amp_resident() {
  static std::atomic<int> cached{0};
  int v = cached.load(std::memory_order_relaxed);
  if (v > 0) return v;
  const void* f = (const void*)k_amplitude<W, XB>;
  const size_t sh = AmpLds<W>::BYTES;
  if (sh > 64 * 1024) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, AMP_NT, sh) != hipSuccess) per = 1;
  v = std::max(per, 1) * std::max(cus, 1);
  cached.store(v, std::memory_order_relaxed);
  return v;
}

constexpr int AMP_ROUNDS = 2;      // rounds of resident blocks a chunked launch fills
constexpr int AMP_MIN_TERMS = 8;   // fewest terms per chunk

int amp_chunks(int W, int64_t nb, int64_t chi) {
  if (nb <= 0 || chi <= 1) return 1;
  int cap = 1;
#define NX_AMP_CAP(WW, XBV) cap = amp_resident<WW, XBV>();
  switch (W) {
    case 1: NX_AMP_CAP(1, 3) break;
    case 2: NX_AMP_CAP(2, 4) break;
    case 3: NX_AMP_CAP(3, 2) break;
    default: NX_AMP_CAP(4, 2) break;
  }
#undef NX_AMP_CAP
  const int64_t per_block = (int64_t)AMP_NT * amp_strings_per_thread(W);
  const int64_t cols = (nb + per_block - 1) / per_block;
  for (int m = AMP_ROUNDS; m >= 1; --m) {
    const int64_t nch = (int64_t)m * cap / cols;
    if (nch >= 2 && chi / nch >= AMP_MIN_TERMS) return (int)std::min<int64_t>(std::min<int64_t>(nch, chi), 64);
  }
  return 1;
}

size_t amp_table_bytes(int W) {
  switch (W) {
    case 1: return AmpLds<1>::BUF;
    case 2: return AmpLds<2>::BUF;
    case 3: return AmpLds<3>::BUF;
    default: return AmpLds<4>::BUF;
  }
}

hipError_t launch_amp_tables(const QFormDev& q, void* tab, hipStream_t st) {
  if (q.chi <= 0) return hipSuccess;
  const unsigned nblk = (unsigned)std::min<int64_t>(q.chi, 16384);
#define NX_TAB_LAUNCH(WW)                                                                                       \
  {                                                                                                              \
    const size_t sh = AmpLds<WW>::BUILD_BYTES;                                                                   \
    if (sh > 64 * 1024) {                                                                                        \
      hipError_t e = hipFuncSetAttribute((const void*)k_amp_tables<WW>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                         (int)sh);                                                               \
      if (e != hipSuccess) return e;                                                                             \
    }                                                                                                            \
    k_amp_tables<WW><<<dim3(nblk), dim3(AMP_NT), sh, st>>>(q, (uint4*)tab);                                      \
  }
  switch (q.W) {
    case 1: NX_TAB_LAUNCH(1) break;
    case 2: NX_TAB_LAUNCH(2) break;
    case 3: NX_TAB_LAUNCH(3) break;
    case 4: NX_TAB_LAUNCH(4) break;
    default: return hipErrorInvalidValue;
  }
#undef NX_TAB_LAUNCH
  return hipGetLastError();
}

hipError_t launch_amplitude(const QFormDev& q, const void* tab, const double2* coef, const u64* bits, int64_t nb,
                            int nchunks, const double* tabs, double2* out, double2* partial, uint8_t* per_term,
                            int enumerate, double gre, double gim, hipStream_t st) {
  if (nb == 0) return hipSuccess;
  const int final_g = nchunks == 1;
  double2* dst = final_g ? out : partial;
  const int xb = amp_strings_per_thread(q.W);
  const dim3 grid(blocks_for(nb, AMP_NT * xb), nchunks);
#define NX_AMP_LAUNCH(WW, XBV)                                                                                   \
  {                                                                                                               \
    const size_t sh = AmpLds<WW>::BYTES;                                                                          \
    if (sh > 64 * 1024) {                                                                                         \
      hipError_t e = hipFuncSetAttribute((const void*)k_amplitude<WW, XBV>,                                      \
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);                    \
      if (e != hipSuccess) return e;                                                                              \
    }                                                                                                             \
    k_amplitude<WW, XBV><<<grid, dim3(AMP_NT), sh, st>>>(q, (const uint4*)tab, coef, bits, nb, nchunks, tabs, dst, \
                                                         per_term, enumerate, final_g, gre, gim);                 \
  }
  switch (q.W) {
    case 1: NX_AMP_LAUNCH(1, 3) break;
    case 2: NX_AMP_LAUNCH(2, 4) break;
    case 3: NX_AMP_LAUNCH(3, 2) break;
    case 4: NX_AMP_LAUNCH(4, 2) break;
    default: return hipErrorInvalidValue;
  }
#undef NX_AMP_LAUNCH
  return hipGetLastError();
}

hipError_t launch_amp_reduce(const double2* partial, int64_t nb, int nchunks, double2* out, double gre, double gim,
                             hipStream_t st) {
  if (nb == 0 || nchunks <= 1) return hipSuccess;
  k_amp_reduce<<<dim3(blocks_for(nb, 256)), dim3(256), 0, st>>>(partial, nb, nchunks, out, gre, gim);
  return hipGetLastError();
}

hipError_t launch_pairs(const QFormDev& X, const QFormDev& Y, int64_t nfixed, int64_t nother, int triangle, int J,
                        int mode, const CoefDev& cx, const CoefDev& cy, const double* tabs, double4* partial,
                        int nblocks, uint32_t* exact_out, int64_t exact_ld, hipStream_t st, hipEvent_t ev0,
                        hipEvent_t ev1) {
  if ((triangle && nfixed != nother) || J != PAIR_UNIT) return hipErrorInvalidValue;
  const bool timed = ev0 && ev1;
  const bool same = X.B == Y.B && X.C == Y.C && X.a == Y.a && cx.scalar == cy.scalar && cx.cs == cy.cs &&
                    cx.c == cy.c;
  // the combinations the runtime uses: norm (same, fixed bra, Hermitian),
  // pair tables (same, complex), exp_value (fixed ket, Hermitian), inner
  // products (complex); any other mode is read at run time
#define NX_PAIRS_CK(SM, MD, CKV)                                                                              \
  NX_DISPATCH_W(X.W, if (timed) {                                                                             \
    hipExtLaunchKernelGGL(k_pairs<WW, SM, MD, CKV>, dim3(nblocks), dim3(256), 0, st, ev0, ev1, 0, X,          \
                          SM ? X : Y, nfixed, nother, triangle, J, mode, cx, SM ? cx : cy, tabs, partial,     \
                          exact_out, exact_ld);                                                               \
  } else {                                                                                                    \
    k_pairs<WW, SM, MD, CKV><<<dim3(nblocks), dim3(256), 0, st>>>(X, SM ? X : Y, nfixed, nother, triangle, J, \
                                                                 mode, cx, SM ? cx : cy, tabs, partial,       \
                                                                 exact_out, exact_ld);                        \
  })
#define NX_PAIRS(SM, MD) NX_PAIRS_CK(SM, MD, -1)
  // the norm: coefficient kind fixed at compile time
  if (same && mode == 2 && !exact_out && cx.scalar) {
    NX_PAIRS_CK(true, 2, 1);
  } else if (same && mode == 2 && !exact_out) {
    NX_PAIRS_CK(true, 2, 0);
  } else if (same && mode == 2) {
    NX_PAIRS(true, 2);
  } else if (same && mode == 0) {
    NX_PAIRS(true, 0);
  } else if (same) {
    NX_PAIRS(true, -1);
  } else if (mode == 3 && !exact_out && cx.scalar && cy.scalar) {  // exp_value: exact coefficients
    NX_PAIRS_CK(false, 3, 1);
  } else if (mode == 3 && !exact_out && !cx.scalar && !cy.scalar) {
    NX_PAIRS_CK(false, 3, 0);
  } else if (mode == 3) {
    NX_PAIRS(false, 3);
  } else if (mode == 0) {
    NX_PAIRS(false, 0);
  } else {
    NX_PAIRS(false, -1);
  }
#undef NX_PAIRS
#undef NX_PAIRS_CK
  return hipGetLastError();
}

hipError_t launch_reduce(const double4* partial, int64_t n, double4* out, hipStream_t st, double* norm_out,
                         double g2) {
  k_reduce<<<dim3(1), dim3(1024), 0, st>>>(partial, n, out, norm_out, g2);
  return hipGetLastError();
}

}  // namespace nx
