// tails.hip -- exact amplitudes of the 2^d completions of a sampling path.
//
// A sampling tail fixes every qubit but the last d <= 16 of the DFS order
// (values b) and needs |<x|psi>|^2 for the 2^d completions x = b | y.  With
// a term written as its quadratic form (the amplitude kernels' form:
// <x|phi> = [C x = rhs] zeta^{p0 + 2 a.x + 4 sum_{l<m} Q_lm x_l x_m}
// 2^{-k/2}, a = alo + 2 ahi, B = Q ^ Q^T), the fixed part folds into a form
// over the d tail variables y alone:
//   constant   p0' = p0 + 2 popc(b & alo) + 4 popc(b & ahi)
//                    + 4 sum_{l in b} popc(Q_l & b)                (mod 8)
//   linear     a'_m = a_m + 2 par(B_m & b)                         (mod 4)
//   quadratic  U_ij = B_{qs_i qs_j} for i < j
//   support    C' y = rhs ^ C b restricted to the tail columns, reduced to
//              echelon form (r <= d rows; a zero row with right-hand side 1
//              means the term vanishes on every completion)
// so a completion costs four nibble lookups instead of 16 W: one table
// entry per nibble position p and value v holds the XOR of U_{4p+m} over
// the set bits m of v (low 16 bits) and of the syndrome columns of the
// same variables (high 16 bits); y is supported iff the XORed syndrome
// equals the reduced right-hand sides, and its phase is
//   p0' + 2 popc(y & alo') + 4 par(y & (ahi' ^ XOR_{i in y} U_i))   (mod 8).
// k_tail_restrict builds one record per (tail, term) -- one wavefront per
// term -- and k_tail_eval adds Scalar(phase + coefficient) exactly into the
// four components of each completion (basis 1, zeta, zeta^2, zeta^3), term
// chunks summed in chunk order by k_tail_norms.  The
// per-(term, completion) values are those of amplitude_at_computational_basis
// (amplitude.rs:28-44), so the exact sums equal the full evaluation's.
#include <hip/hip_runtime.h>

#include "device_form.hpp"
#include "internal.hpp"

namespace nx {

// record words (TAIL_REC per term):
//   [0]       alo' | ahi' << 16
//   [1]       p0' | dead << 3 | target << 16
//   [2 + e]   table entry e = 16 p + v (p < 4 nibble positions, v < 16)
//   [66]      k (2^{-k/2}), [67] the term's coefficient (CoefDev::cs)
constexpr int TR_TAB = 2, TR_K = 66, TR_CS = 67;

// the job of a wave-uniform global index g with jobs[j].*first <= g <
// jobs[j + 1].*first (scalar loads)
template <int64_t TailJobDev::*F>
__device__ __forceinline__ int tail_job(const TailJobDev* jobs, int njobs, int64_t g) {
  int lo = 0, hi = njobs;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((int64_t)uni64((u64)(jobs[mid].*F)) <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// the same for an index that differs between lanes (the lanes of one wave
// may belong to different jobs)
__device__ __forceinline__ int tail_job_lane(const TailJobDev* jobs, int njobs, int64_t g) {
  int lo = 0, hi = njobs;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (jobs[mid].str0 <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// bits qs[0..d) of a row (W words), gathered into bit i (uniform qs)
template <int W>
__device__ __forceinline__ u32 gather_bits(const u64 (&row)[W], const int32_t* qs, int d) {
  u32 m = 0u;
  for (int i = 0; i < d; ++i) {
    const int q = uni(qs[i]);
    m |= (u32)((wsel<W>(row, q >> 6) >> (q & 63)) & 1ull) << i;
  }
  return m;
}

template <int W>
__global__ __launch_bounds__(256) void k_tail_restrict(const TailJobDev* __restrict__ jobs, int njobs,
                                                       int64_t total_terms) {
  const int64_t gw = (int64_t)blockIdx.x * 4 + uni((int)(threadIdx.x >> 6));  // uniform in the wave
  if (gw >= total_terms) return;
  const TailJobDev& J = jobs[tail_job<&TailJobDev::term0>(jobs, njobs, gw)];
  const int64_t t = gw - (int64_t)uni64((u64)J.term0);
  const QFormDev& q = J.q;
  const int lane = threadIdx.x & 63;
  const int d = uni(J.d);
  constexpr int R = 64 * W;
  u64 b[W], alo[W], ahi[W], rhs[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    b[w] = uni64(J.base[w]);
    alo[w] = uni64(q.alo[(size_t)t * W + w]);
    ahi[w] = uni64(q.ahi[(size_t)t * W + w]);
    rhs[w] = uni64(q.rhs[(size_t)t * W + w]);
  }
  const int nc = uni(q.nc[t]);
  // rows l = 64 sl + lane
  u64 Brow[W][W], Qrow[W][W], Crow[W][W];
#pragma unroll
  for (int sl = 0; sl < W; ++sl)
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const size_t o = ((size_t)t * R + 64 * sl + lane) * W + w;
      Brow[sl][w] = q.B[o];
      Qrow[sl][w] = q.Q[o];
      Crow[sl][w] = q.C[o];
    }
  // constant phase and the tail variables' linear coefficients
  int quad = 0;
  u64 wmask[W];
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const bool bl = (b[sl] >> lane) & 1ull;
    int cq = 0, cb = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      cq += __builtin_popcountll(Qrow[sl][w] & b[w]);
      cb += __builtin_popcountll(Brow[sl][w] & b[w]);
    }
    quad += __builtin_popcountll(uni64(__ballot(bl && (cq & 1))));
    wmask[sl] = uni64(__ballot(cb & 1));
  }
  int pb = 0, hb = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    pb += __builtin_popcountll(b[w] & alo[w]);
    hb += __builtin_popcountll(b[w] & ahi[w]);
  }
  const int p0 = (uni(q.p0[t]) + 2 * pb + 4 * hb + 4 * quad) & 7;
  u32 alo2 = 0u, ahi2 = 0u;
  for (int i = 0; i < d; ++i) {
    const int m = uni(J.qs[i]), mw = m >> 6, mb = m & 63;
    const int am = (int)((wsel<W>(alo, mw) >> mb) & 1ull) + 2 * (int)((wsel<W>(ahi, mw) >> mb) & 1ull);
    const int a2 = (am + 2 * (int)((wsel<W>(wmask, mw) >> mb) & 1ull)) & 3;
    alo2 |= (u32)(a2 & 1) << i;
    ahi2 |= (u32)(a2 >> 1) << i;
  }
  // U rows: lane i < d holds U_i (bits j > i)
  u32 cm[W];
#pragma unroll
  for (int sl = 0; sl < W; ++sl) cm[sl] = gather_bits<W>(Brow[sl], J.qs, d);
  const u32 dmask = d >= 32 ? ~0u : ((1u << d) - 1u);
  u32 Uv = 0u;
  for (int i = 0; i < d; ++i) {
    const int m = uni(J.qs[i]);
    u32 src = cm[0];
#pragma unroll
    for (int sl = 1; sl < W; ++sl)
      if ((m >> 6) == sl) src = cm[sl];
    const u32 Ui = (u32)rl32((int)src, m & 63) & dmask & ~((2u << i) - 1u);
    if (lane == i) Uv = Ui;
  }
  // support: constraint rows k < nc restricted to the tail columns, right-
  // hand sides moved by C b, reduced to echelon form over the d columns
  u32 cr[W], rr[W];
  bool act[W], piv[W];
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    const int k = 64 * sl + lane;
    act[sl] = k < nc;
    piv[sl] = false;
    int cb = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) cb += __builtin_popcountll(Crow[sl][w] & b[w]);
    cr[sl] = act[sl] ? gather_bits<W>(Crow[sl], J.qs, d) : 0u;
    rr[sl] = act[sl] ? (u32)((((wsel<W>(rhs, sl) >> lane) & 1ull) ^ (u64)(cb & 1))) : 0u;
  }
  for (int j = 0; j < d; ++j) {
    int psl = -1, plane = 0;
#pragma unroll
    for (int sl = 0; sl < W; ++sl) {
      const u64 cand = uni64(__ballot(act[sl] && !piv[sl] && ((cr[sl] >> j) & 1u)));
      if (psl < 0 && cand) {
        psl = sl;
        plane = ffs64(cand);
      }
    }
    if (psl < 0) continue;
    u32 prow = 0u, prr = 0u;
#pragma unroll
    for (int sl = 0; sl < W; ++sl)
      if (sl == psl) {
        prow = (u32)rl32((int)cr[sl], plane);
        prr = (u32)rl32((int)rr[sl], plane);
      }
#pragma unroll
    for (int sl = 0; sl < W; ++sl) {
      const bool me = sl == psl && lane == plane;
      if (me) piv[sl] = true;
      if (act[sl] && !me && ((cr[sl] >> j) & 1u)) {
        cr[sl] ^= prow;
        rr[sl] ^= prr;
      }
    }
  }
  bool dead = false;
#pragma unroll
  for (int sl = 0; sl < W; ++sl) dead = dead || uni64(__ballot(act[sl] && !piv[sl] && rr[sl])) != 0ull;
  // syndrome columns (lane i < d: bit o = bit i of the o-th pivot row in row
  // order) and the target right-hand sides
  u32 colv = 0u, target = 0u;
  int o = 0;
#pragma unroll
  for (int sl = 0; sl < W; ++sl) {
    u64 pm = uni64(__ballot(piv[sl]));
    while (pm) {
      const int pl = ffs64(pm);
      pm &= pm - 1ull;
      const u32 row = (u32)rl32((int)cr[sl], pl);
      target |= ((u32)rl32((int)rr[sl], pl) & 1u) << o;
      colv |= ((row >> (lane & 31)) & 1u & (lane < d ? 1u : 0u)) << o;
      ++o;
    }
  }
  // nibble tables: lane e = 16 p + v
  const u32 packed = lane < d ? (Uv | (colv << 16)) : 0u;
  const int p = lane >> 4, v = lane & 15;
  u32 ent = 0u;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const u32 src = (u32)__shfl((int)packed, 4 * p + m);
    if ((v >> m) & 1) ent ^= src;
  }
  uint32_t* rec = J.recs + (size_t)t * TAIL_REC;
  rec[TR_TAB + lane] = ent;
  if (lane == 0) {
    rec[0] = alo2 | (ahi2 << 16);
    rec[1] = (u32)p0 | (dead ? 8u : 0u) | (target << 16);
    rec[TR_K] = (u32)uni(q.k[t]);
    rec[TR_CS] = (u32)uni(J.cs[t]);
  }
}

// Exact zeta^K 2^{-R/2} into the four components of the basis 1, zeta,
// zeta^2, zeta^3 (zeta^{j+4} = -zeta^j).  For even R it is +-2^{-R/2} on
// component K mod 4; for odd R, sqrt2 = zeta + zeta^-1 splits it into
// 2^{-(R+1)/2} on zeta^{K+1} and zeta^{K-1}.  mag = 2^{-R/2} (even R) or
// 2^{-(R+1)/2} (odd R), the same for every completion of a term.  Each
// addend is a power of two, so the components are exact sums while their
// exponent span fits a double.
__device__ __forceinline__ void tail_add(int K, bool odd, double mag, double (&al)[4]) {
  const int ka = odd ? ((K + 1) & 7) : K, kb = (K + 7) & 7;
  const double v1 = (ka & 4) ? -mag : mag;
  const double v2 = odd ? ((kb & 4) ? -mag : mag) : 0.0;
  const int i1 = ka & 3, i2 = kb & 3;
#pragma unroll
  for (int j = 0; j < 4; ++j) al[j] = al[j] + ((j == i1) ? v1 : 0.0) + ((j == i2) ? v2 : 0.0);
}

constexpr int TAIL_XB = 4;   // completions per thread
constexpr int TAIL_TILE = 32;  // term records staged in LDS at a time

// block b of job j (jobs[j].block0 <= b): completion column (b - block0) %
// cols, term chunk (b - block0) / cols; al[(chunk nv + y) 4 + m]
__global__ __launch_bounds__(256) void k_tail_eval(const TailJobDev* __restrict__ jobs, int njobs) {
  __shared__ u32 s_rec[TAIL_TILE * TAIL_REC];
  const int64_t blk = blockIdx.x;
  const TailJobDev& J = jobs[tail_job<&TailJobDev::block0>(jobs, njobs, blk)];
  const int64_t nv = (int64_t)uni64((u64)J.nv);
  const int64_t per = 256 * TAIL_XB;
  const int64_t cols = (nv + per - 1) / per;
  const int64_t local = blk - (int64_t)uni64((u64)J.block0);
  const int64_t col = local % cols, chunk = local / cols;
  const int nch = uni(J.nchunks);
  const int64_t chi = J.q.chi;
  const int64_t t0 = chunk * chi / nch, t1 = (chunk + 1) * chi / nch;
  u32 y[TAIL_XB], off[TAIL_XB][4];
  bool active[TAIL_XB];
  double al[TAIL_XB][4];
#pragma unroll
  for (int x = 0; x < TAIL_XB; ++x) {
    const int64_t yi = (col * TAIL_XB + x) * 256 + threadIdx.x;
    active[x] = yi < nv;
    y[x] = active[x] ? (u32)yi : 0u;
#pragma unroll
    for (int p = 0; p < 4; ++p) off[x][p] = (u32)(TR_TAB + 16 * p + ((y[x] >> (4 * p)) & 15u));
#pragma unroll
    for (int m = 0; m < 4; ++m) al[x][m] = 0.0;
  }
  for (int64_t tb = t0; tb < t1; tb += TAIL_TILE) {
    const int nt = (int)((t1 - tb) < TAIL_TILE ? (t1 - tb) : TAIL_TILE);
    __syncthreads();
    for (int e = threadIdx.x; e < nt * TAIL_REC; e += 256) s_rec[e] = J.recs[(size_t)tb * TAIL_REC + e];
    __syncthreads();
    for (int tt = 0; tt < nt; ++tt) {
      const u32* r = s_rec + tt * TAIL_REC;
      const u32 w1 = r[1];
      const int32_t cs = (int32_t)r[TR_CS];
      if ((w1 & 8u) || cs < 0) continue;  // vanishes on every completion, or a Scalar::Zero coefficient
      const u32 w0 = r[0];
      const u32 alo = w0 & 0xFFFFu, ahi = w0 >> 16, target = w1 >> 16;
      const int p0 = (int)(w1 & 7u), Rr = (int)r[TR_K] + cs_r(cs), Kc = cs_k(cs);
      const bool odd = Rr & 1;
      const double mag = ldexp(1.0, -((Rr + (odd ? 1 : 0)) >> 1));
#pragma unroll
      for (int x = 0; x < TAIL_XB; ++x) {
        const u32 acc = xor3(r[off[x][0]], r[off[x][1]], r[off[x][2]]) ^ r[off[x][3]];
        const bool supp = (acc >> 16) == target;
        const int ph = (p0 + 2 * __builtin_popcount(y[x] & alo) +
                        4 * (__builtin_popcount(y[x] & ((ahi ^ acc) & 0xFFFFu)) & 1)) & 7;
        if (supp && active[x]) tail_add((ph + Kc) & 7, odd, mag, al[x]);
      }
    }
  }
#pragma unroll
  for (int x = 0; x < TAIL_XB; ++x) {
    if (!active[x]) continue;
#pragma unroll
    for (int m = 0; m < 4; ++m) J.al[((size_t)chunk * nv + y[x]) * 4 + m] = al[x][m];
  }
}

// the chunks of each completion's components summed in chunk order, then
// its leaf norm |psi|^2 = A + sqrt2 B (A = sum a_j^2, B = a0 a1 + a1 a2 + a2
// a3 - a0 a3, each from exact products in double-double, rounded once)
__global__ void k_tail_norms(const TailJobDev* __restrict__ jobs, int njobs, int64_t total) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const TailJobDev& J = jobs[tail_job_lane(jobs, njobs, g)];
  const int64_t k = g - J.str0;
  if (k >= J.nv) return;
  double a[4];
  for (int m = 0; m < 4; ++m) {
    double v = J.al[k * 4 + m];
    for (int c = 1; c < J.nchunks; ++c) v = v + J.al[((size_t)c * J.nv + k) * 4 + m];
    a[m] = v;
  }
  auto two_prod_sum = [](const double* x, const double* y, const double* s, int n) {
    double hi = 0.0, lo = 0.0;
    for (int i = 0; i < n; ++i) {
      const double pa = x[i] * s[i], pb = y[i];
      const double p = pa * pb, pe = fma(pa, pb, -p);
      const double t = hi + p, bb = t - hi;
      double e = (hi - (t - bb)) + (p - bb);
      e = e + (lo + pe);
      hi = t + e;
      lo = e - (hi - t);
    }
    return hi + lo;
  };
  const double xa[4] = {a[0], a[1], a[2], a[3]}, ya[4] = {a[0], a[1], a[2], a[3]}, one[4] = {1.0, 1.0, 1.0, 1.0};
  const double xb[4] = {a[0], a[1], a[2], a[0]}, yb[4] = {a[1], a[2], a[3], a[3]}, sb[4] = {1.0, 1.0, 1.0, -1.0};
  J.ab[k] = two_prod_sum(xa, ya, one, 4);
  J.ab[J.nv + k] = two_prod_sum(xb, yb, sb, 4);
}

hipError_t launch_tails(int W, const TailJobDev* jobs, int njobs, int64_t total_terms, int64_t total_blocks,
                        int64_t total_strings, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  if (total_terms > 0) {
    const dim3 g((unsigned)((total_terms + 3) / 4)), b(256);
    if (W == 1) k_tail_restrict<1><<<g, b, 0, st>>>(jobs, njobs, total_terms);
    else if (W == 2) k_tail_restrict<2><<<g, b, 0, st>>>(jobs, njobs, total_terms);
    else return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_tail_eval<<<dim3((unsigned)total_blocks), dim3(256), 0, st>>>(jobs, njobs);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  k_tail_norms<<<dim3((unsigned)((total_strings + 255) / 256)), dim3(256), 0, st>>>(jobs, njobs, total_strings);
  return hipGetLastError();
}

}  // namespace nx
