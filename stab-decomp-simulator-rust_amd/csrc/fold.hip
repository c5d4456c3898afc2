// fold.hip -- node norms of a sampling window in the reference's FP64
// summation order (gfx950, wave64).
//
// The reference takes a node's probability from norm_squared() of the state
// projected on the node's path (norm.rs:8-30, sampling.rs:151-157): a
// sequential FP64 sum over the surviving terms i <= j of
//   Re(conj(c_i') c_j' <phi_i'|phi_j'>)     (twice for i < j)
// where phi' = 2^(D/2) Pi phi is the projected CH-form (D = number of
// non-deterministic projections on the path, projection.rs:35-51) and
// c' = c amplified by 2^(-D/2).  The window pass (k_pairs_var with records)
// gives every pair's exact values <phi_o|Pi_leaf|phi_f> at the window's 64
// leaves; here
//  k_fold_diag   D of every term at every node, from the diagonal records
//  k_fold_terms  per (pair, node): the node value as an exact scalar (sum of
//                its leaves), then the FP64 term exactly as the reference
//                forms it: Scalar -> Complex64 conversions of the
//                coefficient product and of the inner product, their
//                num-complex product, doubled off the diagonal
//  k_fold_sum    per node (one wave each), the sequential sum of those terms
//                in pair order
// so a node's norm carries the reference's rounding bit for bit.
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace nx {

// node u of a window of depth d: relative depth j (1..d), outcome index k
// (bit i = outcome of the window's qubit i), u = 2^j - 2 + k
__device__ __forceinline__ void node_of(int u, int* j, int* k) {
  const int jj = 31 - __clz(u + 2);
  *j = jj;
  *k = u + 2 - (1 << jj);
}

// the leaves below node (j, k): variants v < 2^d with v = k mod 2^j, i.e.
// bit k of every block of 2^j bits
__device__ __forceinline__ uint64_t leaves_of(int d, int j, int k) {
  const uint64_t rep = j == 1   ? 0x5555555555555555ull
                       : j == 2 ? 0x1111111111111111ull
                       : j == 3 ? 0x0101010101010101ull
                       : j == 4 ? 0x0001000100010001ull
                       : j == 5 ? 0x0000000100000001ull
                                : 0x0000000000000001ull;
  const uint64_t low = d >= 6 ? ~0ull : ((1ull << (1 << d)) - 1ull);
  return (rep << k) & low;
}

// sum over the non-vanishing leaves of a record of zeta^((c + D_v) & 7), as
// integer components in the basis 1, zeta, zeta^2, zeta^3 (zeta^4 = -1)
__device__ __forceinline__ void node_components(const uint64_t* r, uint64_t leaves, int a[4]) {
  const uint64_t live = leaves & ~r[0];
  const int c = (int)(r[4] & 7ull);
  a[0] = a[1] = a[2] = a[3] = 0;
#pragma unroll
  for (int D = 0; D < 8; ++D) {
    const uint64_t m = live & ((D & 1) ? r[1] : ~r[1]) & ((D & 2) ? r[2] : ~r[2]) & ((D & 4) ? r[3] : ~r[3]);
    const int cnt = __popcll(m);
    const int K = (c + D) & 7;
    a[K & 3] += (K & 4) ? -cnt : cnt;
  }
}

__device__ __forceinline__ bool pow2(int x, int* s) {
  if (x <= 0 || (x & (x - 1))) return false;
  *s = __ffs(x) - 1;
  return true;
}

// a0 + a1 zeta + a2 zeta^2 + a3 zeta^3 = zeta^K 2^(h/2), or zero; false if
// the value is of neither form.  sqrt2 zeta^K = zeta^(K-1) + zeta^(K+1)
// occupies components {1, 3} (K even) or {0, 2} (K odd).
__device__ __forceinline__ bool z8_scalar(const int a[4], int* zero, int* K, int* h) {
  int nz = 0, i0 = -1, i1 = -1;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (a[i] != 0) {
      if (nz == 0) i0 = i;
      else i1 = i;
      ++nz;
    }
  *zero = (nz == 0);
  *K = 0;
  *h = 0;
  if (nz == 0) return true;
  int s = 0;
  if (nz == 1) {
    const int v = a[i0];
    if (!pow2(v < 0 ? -v : v, &s)) return false;
    *K = i0 + (v < 0 ? 4 : 0);
    *h = 2 * s;
    return true;
  }
  if (nz != 2 || i1 != i0 + 2) return false;
  const int x = a[i0], y = a[i1];
  if ((x < 0 ? -x : x) != (y < 0 ? -y : y) || !pow2(x < 0 ? -x : x, &s)) return false;
  if (i0 == 1) *K = (x > 0) ? (y < 0 ? 0 : 2) : (y > 0 ? 4 : 6);
  else *K = (x > 0) ? (y > 0 ? 1 : 7) : (y > 0 ? 3 : 5);
  *h = 2 * s + 1;
  return true;
}

// first pair of row f in the reference order (row f holds o = f .. chi - 1)
__device__ __forceinline__ int64_t row_start(int64_t f, int64_t chi) { return f * chi - f * (f - 1) / 2; }

__global__ __launch_bounds__(256) void k_fold_diag(const uint64_t* __restrict__ rec, int64_t chi, int d,
                                                   uint8_t* __restrict__ D, int* __restrict__ err) {
  const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t f = gid / FOLD_NODES;
  const int u = (int)(gid % FOLD_NODES);
  if (f >= chi) return;
  uint8_t out = 255;
  if (u < (2 << d) - 2) {
    int j, k;
    node_of(u, &j, &k);
    const uint64_t* r = rec + (size_t)FOLD_REC * (size_t)row_start(f, chi);
    int a[4], zero, K, h;
    node_components(r, leaves_of(d, j, k), a);
    const int mag = (int)(r[4] >> 8);
    // <phi_f|Pi_u|phi_f> = 2^-D: K = 0 and an even magnitude 2D
    if (!z8_scalar(a, &zero, &K, &h) || (!zero && (K != 0 || ((mag - h) & 1) || mag - h < 0 || mag - h > 508)))
      *err = 1;
    else if (!zero)
      out = (uint8_t)((mag - h) >> 1);
  }
  D[(size_t)f * FOLD_NODES + u] = out;
}

// The term of pair p at node u (0.0 when it vanishes): the node value as the
// exact sum of its leaves, then the FP64 product as the reference forms it
__device__ __forceinline__ double fold_term(const uint64_t* __restrict__ rec, int64_t chi, int64_t p, int u,
                                            uint64_t leaves, const uint8_t* __restrict__ D,
                                            const int32_t* __restrict__ cs, const double* __restrict__ tabs,
                                            int* __restrict__ err) {
  const uint64_t* r = rec + (size_t)FOLD_REC * (size_t)p;
  if ((leaves & ~r[0]) == 0ull) return 0.0;  // every leaf below the node vanishes
  // row f of pair p: the largest f with row_start(f) <= p
  const double b = 2.0 * (double)chi + 1.0;
  int64_t f = (int64_t)((b - sqrt(b * b - 8.0 * (double)p)) * 0.5);
  if (f < 0) f = 0;
  if (f > chi - 1) f = chi - 1;
  while (f > 0 && row_start(f, chi) > p) --f;
  while (f + 1 < chi && row_start(f + 1, chi) <= p) ++f;
  const int64_t o = f + (p - row_start(f, chi));
  const int32_t sf = cs[f], so = cs[o];
  if ((sf | so) < 0) return 0.0;
  int a[4], zero, K, h;
  node_components(r, leaves, a);
  if (!z8_scalar(a, &zero, &K, &h)) {
    *err = 1;
    return 0.0;
  }
  if (zero) return 0.0;
  const int Df = D[(size_t)f * FOLD_NODES + u], Do = D[(size_t)o * FOLD_NODES + u];
  const int mag = (int)(r[4] >> 8);
  // reference pair (i, j) = (f, o): conj(c_f') c_o' and <phi_f'|phi_o'> =
  // conj(<phi_o|Pi|phi_f>) 2^((Df + Do)/2)
  const int kc = (cs_k(so) - cs_k(sf)) & 7;
  const int rc = cs_r(sf) + cs_r(so) + Df + Do;
  const int Kr = (8 - K) & 7;
  const int Rr = mag - h - Df - Do;
  if (Df == 255 || Do == 255 || rc < -TAB_RMAX || rc > TAB_RMAX || Rr < -TAB_RMAX || Rr > TAB_RMAX) {
    *err = 1;
    return 0.0;
  }
  const double nc = tabs[TAB_POW + rc], ni = tabs[TAB_POW + Rr];
  const double cr = tabs[TAB_COS + kc] * nc, ci = tabs[TAB_SIN + kc] * nc;
  const double pr = tabs[TAB_COS + Kr] * ni, pi = tabs[TAB_SIN + Kr] * ni;
  const double tr = cr * pr - ci * pi;
  return (f == o) ? tr : tr + tr;
}

// Block = 256 consecutive pairs p0 + 256 c .. of one node u.  The nonzero
// terms are packed in pair order into t[u * ld + 256 c ...] and their count
// goes to cnt[u * nch + c] (nch = chunks per node): the sum then reads only
// the nonzero terms.  A zero term is an identity of the sum (it starts at +0
// and is never -0), so dropping it keeps the reference's rounding.
__global__ __launch_bounds__(256) void k_fold_terms(const uint64_t* __restrict__ rec, int64_t chi, int64_t p0,
                                                    int64_t np, int64_t ld, int64_t nch, int d,
                                                    const uint8_t* __restrict__ D, const int32_t* __restrict__ cs,
                                                    const double* __restrict__ tabs, double* __restrict__ t,
                                                    int32_t* __restrict__ cnt, int* __restrict__ err) {
  const int u = (int)(blockIdx.x / nch);
  const int64_t c = blockIdx.x - (int64_t)u * nch;
  const int64_t q = 256 * c + threadIdx.x;
  int j, k;
  node_of(u, &j, &k);
  const double val = q < np ? fold_term(rec, chi, p0 + q, u, leaves_of(d, j, k), D, cs, tabs, err) : 0.0;
  __shared__ int wc[4];
  const uint64_t m = __ballot(val != 0.0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) wc[wv] = __popcll(m);
  __syncthreads();
  int off = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) off += w < wv ? wc[w] : 0;
  const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
  if (val != 0.0) t[(size_t)u * ld + 256 * c + off + below] = val;
  if (threadIdx.x == 0) cnt[(size_t)u * nch + c] = wc[0] + wc[1] + wc[2] + wc[3];
}

// One wave per node: the node's packed terms, chunk by chunk in pair order,
// added one by one from LDS (every lane reads the same address).  A chunk's
// 256 slots are loaded whole (the adds read only its count of them), the next
// chunk's slots and count while the current chunk's dependent adds run, and
// the LDS reads of each group of 16 terms (the last one included) are issued
// before the adds of the group before it.
__global__ __launch_bounds__(64) void k_fold_sum(const double* __restrict__ t, const int32_t* __restrict__ cnt,
                                                 int64_t ld, int64_t nch, double* __restrict__ sums) {
  __shared__ double buf[256 + 16];  // 16 slots of padding: a last group may read past the chunk
  const int u = (int)blockIdx.x;
  const int lane = threadIdx.x;
  const double* col = t + (size_t)u * ld;
  const int32_t* cu = cnt + (size_t)u * nch;
  double s = sums[u];
  double v[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = col[64 * i + lane];
  int next = cu[0];
  for (int64_t c = 0; c < nch; ++c) {
    const int cur = __builtin_amdgcn_readfirstlane(next);
#pragma unroll
    for (int i = 0; i < 4; ++i) buf[64 * i + lane] = v[i];
    if (c + 1 < nch) {
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = col[256 * (c + 1) + 64 * i + lane];
      next = cu[c + 1];
    }
    __syncthreads();
    double a[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = buf[i];
    int k = 0;
    for (; k + 16 < cur; k += 16) {
      double b[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) b[i] = buf[k + 16 + i];
#pragma unroll
      for (int i = 0; i < 16; ++i) s = s + a[i];
#pragma unroll
      for (int i = 0; i < 16; ++i) a[i] = b[i];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (k + i < cur) s = s + a[i];
    __syncthreads();
  }
  if (lane == 0) sums[u] = s;
}

hipError_t launch_fold_diag(const uint64_t* rec, int64_t chi, int d, uint8_t* D, int* err, hipStream_t st) {
  if (chi <= 0) return hipSuccess;
  if (d < 1 || d > 6) return hipErrorInvalidValue;
  const int64_t nt = chi * FOLD_NODES;
  k_fold_diag<<<dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st>>>(rec, chi, d, D, err);
  return hipGetLastError();
}

hipError_t launch_fold_terms(const uint64_t* rec, int64_t chi, int64_t p0, int64_t np, int64_t ld, int d,
                             const uint8_t* D, const int32_t* cs, const double* tabs, double* t, int32_t* cnt,
                             int* err, hipStream_t st) {
  if (np <= 0) return hipSuccess;
  if (d < 1 || d > 6 || p0 < 0 || p0 + np > chi * (chi + 1) / 2 || np > ld) return hipErrorInvalidValue;
  const int nn = (2 << d) - 2;
  const int64_t nch = (np + 255) / 256;
  if (256 * nch > ld) return hipErrorInvalidValue;
  k_fold_terms<<<dim3((unsigned)(nch * nn)), dim3(256), 0, st>>>(rec, chi, p0, np, ld, nch, d, D, cs, tabs, t, cnt,
                                                                err);
  return hipGetLastError();
}

hipError_t launch_fold_sum(const double* t, const int32_t* cnt, int64_t np, int64_t ld, int nnodes, double* sums,
                           hipStream_t st) {
  if (np <= 0) return hipSuccess;
  const int64_t nch = (np + 255) / 256;
  if (nnodes < 1 || nnodes > FOLD_NODES || 256 * nch > ld) return hipErrorInvalidValue;
  k_fold_sum<<<dim3((unsigned)nnodes), dim3(64), 0, st>>>(t, cnt, ld, nch, sums);
  return hipGetLastError();
}

}  // namespace nx
