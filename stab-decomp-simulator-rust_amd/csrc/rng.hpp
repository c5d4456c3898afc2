// rng.hpp -- host RNG used by measure/sample.
//
// The reference draws from rand 0.8.5 StdRng::from_seed([u8; 32])
// (measurement.rs:12-15, sampling.rs:22-25).  StdRng is rand_chacha 0.3's
// ChaCha12Rng: a 64-bit block counter, stream 0, and a rand_core 0.6
// BlockRng buffer of four 16-word blocks.  Restated from the published
// algorithms: ChaCha (Bernstein 2008), rand's Standard f64 (53 high bits),
// UniformFloat<f64> ([1,2) mantissa trick), and rand_distr 0.4.3 Binomial
// (BINV below n*p = 10, BTPE of Kachitvichyanukul & Schmeiser 1988 above).
// Draws stay on the host in DFS order; only the norms run on the device.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace nx {

class ChaCha12Rng {
 public:
  explicit ChaCha12Rng(const uint8_t* seed) {
    for (int i = 0; i < 8; ++i)
      key_[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) | ((uint32_t)seed[4 * i + 2] << 16) |
                ((uint32_t)seed[4 * i + 3] << 24);
  }
  uint32_t next_u32() {
    if (index_ >= 64) refill();
    return buf_[index_++];
  }
  uint64_t next_u64() {
    uint64_t lo, hi;
    if (index_ < 63) {
      lo = buf_[index_];
      hi = buf_[index_ + 1];
      index_ += 2;
    } else if (index_ >= 64) {
      refill();
      lo = buf_[0];
      hi = buf_[1];
      index_ = 2;
    } else {
      lo = buf_[63];
      refill();
      hi = buf_[0];
      index_ = 1;
    }
    return (hi << 32) | lo;
  }
  // rand::distributions::Standard for f64
  double gen_f64() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

  static void block(const uint32_t key[8], uint64_t counter, int rounds, uint32_t out[16]) {
    uint32_t st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2], key[3],
                       key[4],      key[5],      key[6],      key[7],      (uint32_t)counter,
                       (uint32_t)(counter >> 32), 0u, 0u};
    uint32_t x[16];
    std::memcpy(x, st, sizeof st);
    auto rotl = [](uint32_t v, int k) { return (v << k) | (v >> (32 - k)); };
    auto qr = [&](int a, int b, int c, int d) {
      x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 16);
      x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 12);
      x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 8);
      x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 7);
    };
    for (int i = 0; i < rounds; i += 2) {
      qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
      qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + st[i];
  }

 private:
  void refill() {
    for (int b = 0; b < 4; ++b) block(key_, counter_ + (uint64_t)b, 12, buf_ + 16 * b);
    counter_ += 4;
    index_ = 0;
  }
  uint32_t key_[8];
  uint64_t counter_ = 0;
  uint32_t buf_[64];
  int index_ = 64;
};

namespace rngdetail {
inline double from_bits(uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; }
inline uint64_t to_bits(double d) { uint64_t b; std::memcpy(&b, &d, 8); return b; }
struct Unif {
  double low, scale;
  Unif(double lo, double hi) : low(lo) {
    const double max_rand = from_bits((uint64_t(1023) << 52) | (~uint64_t(0) >> 12)) - 1.0;
    scale = hi - lo;
    while (scale * max_rand + lo >= hi) scale = from_bits(to_bits(scale) - 1);
  }
  double sample(ChaCha12Rng& r) const {
    const double v12 = from_bits((uint64_t(1023) << 52) | (r.next_u64() >> 12));
    return (v12 - 1.0) * scale + low;
  }
};
inline int64_t f2i(double x) {
  if (std::isnan(x)) return 0;
  if (x >= 9.2233720368547758e18) return INT64_MAX;
  if (x <= -9.2233720368547758e18) return INT64_MIN;
  return (int64_t)x;
}
inline double powi(double a, int b) {  // f64::powi lowering (repeated squaring)
  const bool recip = b < 0;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}
inline double stirling(double a) {
  const double a2 = a * a;
  return (13860. - (462. - (132. - (99. - 140. / a2) / a2) / a2) / a2) / a / 166320.;
}
}  // namespace rngdetail

// rand_distr 0.4.3 Binomial::sample; caller checks 0 <= p <= 1
inline uint64_t binomial(ChaCha12Rng& rng, uint64_t nn, double pp) {
  using namespace rngdetail;
  if (pp == 0.0) return 0;
  if (pp == 1.0) return nn;
  const double p = pp <= 0.5 ? pp : 1.0 - pp;
  const double q = 1. - p;
  uint64_t result;
  if ((double)nn * p < 10. && nn <= (uint64_t)INT32_MAX) {
    const double s = p / q;
    const double a = ((double)(nn + 1)) * s;
    double r = powi(q, (int)nn);
    double u = rng.gen_f64();
    uint64_t x = 0;
    while (u > r) {
      u -= r;
      x += 1;
      r *= a / (double)x - s;
    }
    result = x;
  } else {
    const double n = (double)nn, np = n * p, npq = np * q, f_m = np + p;
    const int64_t m = f2i(f_m);
    const double p1 = std::floor(2.195 * std::sqrt(npq) - 4.6 * q) + 0.5;
    const double x_m = (double)m + 0.5, x_l = x_m - p1, x_r = x_m + p1;
    const double c = 0.134 + 20.5 / (15.3 + (double)m);
    const double p2 = p1 * (1. + 2. * c);
    auto lam = [](double a) { return a * (1. + 0.5 * a); };
    const double lambda_l = lam((f_m - x_l) / (f_m - x_l * p));
    const double lambda_r = lam((x_r - f_m) / (x_r * q));
    const double p3 = p2 + c / lambda_l, p4 = p3 + c / lambda_r;
    const Unif gen_u(0., p4), gen_v(0., 1.);
    int64_t y;
    for (;;) {
      const double u = gen_u.sample(rng);
      double v = gen_v.sample(rng);
      if (!(u > p1)) {
        y = f2i(x_m - p1 * v + u);
        break;
      }
      if (!(u > p2)) {
        const double x = x_l + (u - p1) / c;
        v = v * c + 1.0 - std::fabs(x - x_m) / p1;
        if (v > 1.) continue;
        y = f2i(x);
      } else if (!(u > p3)) {
        y = f2i(x_l + std::log(v) / lambda_l);
        if (y < 0) continue;
        v *= (u - p2) * lambda_l;
      } else {
        y = f2i(x_r - std::log(v) / lambda_r);
        if (y > 0 && (uint64_t)y > nn) continue;
        v *= (u - p3) * lambda_r;
      }
      const int64_t kk = y > m ? y - m : m - y;
      if (!(kk > 20 && (double)kk < 0.5 * npq - 1.)) {
        const double s = p / q, a = s * (n + 1.);
        double f = 1.0;
        if (m < y) {
          for (int64_t i = m;;) {
            i += 1;
            f *= a / (double)i - s;
            if (i == y) break;
          }
        } else if (m > y) {
          for (int64_t i = y;;) {
            i += 1;
            f /= a / (double)i - s;
            if (i == m) break;
          }
        }
        if (v > f) continue;
        break;
      }
      const double k = (double)kk;
      const double rho = (k / npq) * ((k * (k / 3. + 0.625) + 1. / 6.) / npq + 0.5);
      const double t = -0.5 * k * k / npq;
      const double alpha = std::log(v);
      if (alpha < t - rho) break;
      if (alpha > t + rho) continue;
      const double x1 = (double)(y + 1), f1 = (double)(m + 1);
      const double z = (double)(f2i(n) + 1 - m), w = (double)(f2i(n) - y + 1);
      if (alpha > x_m * std::log(f1 / x1) + (n - (double)m + 0.5) * std::log(z / w) +
                      (double)(y - m) * std::log(w * p / (x1 * q)) + stirling(f1) + stirling(z) - stirling(x1) -
                      stirling(w))
        continue;
      break;
    }
    result = (uint64_t)y;
  }
  return (p != pp) ? nn - result : result;
}

}  // namespace nx
