// psh_decimal.h -- exact decimal -> double conversion for the ASCII ingest
// (host and device).
//
// The reference reads coordinates with sscanf("%lf") (io/read_ascii.c:164-
// 233), i.e. glibc strtod: the correctly rounded (round-to-nearest-even)
// double of the decimal string.  This conversion returns the same double:
//  * a token [+-]digits[.digits][(e|E)[+-]digits] is split into the first
//    38 significant digits M (128 bits), a decimal exponent E and a flag for
//    nonzero digits dropped beyond the 38th;
//  * fast path (Clinger 1990): M <= 2^53 and |E| <= 22 with nothing dropped:
//    (double)M and 10^|E| are exact, so one IEEE multiply or divide rounds
//    the exact value once;
//  * otherwise the exact value M 10^E (E >= 0) or floor(M 2^s / 10^-E) with
//    a remainder flag (E < 0) is formed in 32-bit limbs and rounded to 53
//    bits, half to even, by hand;
//  * with dropped digits the value lies in (M 10^E, (M+1) 10^E); the result
//    is accepted when both ends round to the same double (otherwise the
//    token is rejected: no silent misrounding).
// Supported: results that are normal doubles (no overflow, no subnormals);
// inf, nan and hexadecimal forms are rejected, and so is a
// token with more than 19 significant digits whose value range straddles a
// rounding boundary (strtod would need the digits beyond the 38th).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define PSH_HD __host__ __device__
#else
#define PSH_HD
#endif

namespace pshdec {

constexpr int kLimbs = 48;   // 1536 bits: M 10^310, or M 2^s for 10^-360

PSH_HD inline int msb64(uint64_t x) { return 63 - __builtin_clzll(x); }

PSH_HD inline uint32_t pow10_u32(int e) {
  uint32_t r = 1;
  for (int i = 0; i < e; i++) r *= 10u;
  return r;
}

PSH_HD inline double pow10_exact(int e) {   // 10^e exactly, 0 <= e <= 22
  double r = 1.0;
  for (int i = 0; i < e; i++) r *= 10.0;   // every partial product is exact
  return r;
}

PSH_HD inline void mul_small(uint32_t *L, int &nl, uint32_t m) {
  uint64_t carry = 0;
  for (int i = 0; i < nl; i++) {
    const uint64_t t = (uint64_t)L[i] * m + carry;
    L[i] = (uint32_t)t;
    carry = t >> 32;
  }
  if (carry) L[nl++] = (uint32_t)carry;
}

PSH_HD inline void shl(uint32_t *L, int &nl, int s) {
  const int ls = s >> 5, bs = s & 31;
  if (bs) {
    uint32_t carry = 0;
    for (int i = 0; i < nl; i++) {
      const uint32_t v = L[i];
      L[i] = (v << bs) | carry;
      carry = v >> (32 - bs);
    }
    if (carry) L[nl++] = carry;
  }
  if (ls) {
    for (int i = nl - 1; i >= 0; i--) L[i + ls] = L[i];
    for (int i = 0; i < ls; i++) L[i] = 0;
    nl += ls;
  }
}

// L /= d; returns whether the remainder is nonzero.
PSH_HD inline bool div_small(uint32_t *L, int &nl, uint32_t d) {
  uint64_t rem = 0;
  for (int i = nl - 1; i >= 0; i--) {
    const uint64_t cur = (rem << 32) | L[i];
    L[i] = (uint32_t)(cur / d);
    rem = cur % d;
  }
  while (nl > 1 && L[nl - 1] == 0) nl--;
  return rem != 0;
}

PSH_HD inline int bit_at(const uint32_t *L, int b) { return (L[b >> 5] >> (b & 31)) & 1; }

PSH_HD inline bool any_below(const uint32_t *L, int b) {   // any bit < b set
  for (int i = 0; i < (b >> 5); i++)
    if (L[i]) return true;
  const int r = b & 31;
  return r && (L[b >> 5] & ((1u << r) - 1u));
}

// 128-bit significand (hi, lo) helpers.
PSH_HD inline void mul10_add(uint64_t &hi, uint64_t &lo, uint32_t d) {
  const uint64_t a0 = lo & 0xffffffffull, a1 = lo >> 32;
  const uint64_t t0 = a0 * 10 + d;
  const uint64_t t1 = a1 * 10 + (t0 >> 32);
  lo = (t1 << 32) | (t0 & 0xffffffffull);
  hi = hi * 10 + (t1 >> 32);
}

PSH_HD inline int msb128(uint64_t hi, uint64_t lo) { return hi ? 64 + msb64(hi) : msb64(lo); }

// Correctly rounded double of M 10^E (M = hi:lo >= 1), plus an
// infinitesimal when `above`.  Returns false outside the supported range.
PSH_HD inline bool round_exact(uint64_t Mhi, uint64_t Mlo, int E, bool above, double *out) {
  uint32_t L[kLimbs + 2];
  for (int i = 0; i < kLimbs + 2; i++) L[i] = 0;
  L[0] = (uint32_t)Mlo;
  L[1] = (uint32_t)(Mlo >> 32);
  L[2] = (uint32_t)Mhi;
  L[3] = (uint32_t)(Mhi >> 32);
  int nl = 4;
  while (nl > 1 && L[nl - 1] == 0) nl--;
  int s = 0;
  bool sticky = above;
  if (E >= 0) {
    if (E > 310) return false;
    for (int e = E; e > 0;) {
      const int step = e >= 9 ? 9 : e;
      mul_small(L, nl, pow10_u32(step));
      e -= step;
    }
  } else {
    const int k = -E;
    if (k > 360) return false;
    // s makes the quotient at least 2^55: 2^s >= 2^56 10^k / 2^msb(M)
    const int c = (int)(((int64_t)k * 3321928095LL + 999999999LL) / 1000000000LL);
    s = 56 + c - msb128(Mhi, Mlo);
    if (s < 0) s = 0;
    shl(L, nl, s);
    for (int e = k; e > 0;) {
      const int step = e >= 9 ? 9 : e;
      sticky |= div_small(L, nl, pow10_u32(step));
      e -= step;
    }
  }
  while (nl > 1 && L[nl - 1] == 0) nl--;
  const int B = 32 * (nl - 1) + (31 - __builtin_clz(L[nl - 1]));
  int sh = B - 52;
  uint64_t mant;
  if (sh <= 0) {
    // the value has at most 53 bits: exact (an infinitesimal above an
    // exactly representable value does not change its rounding)
    const uint64_t low = (uint64_t)L[0] | ((uint64_t)L[1] << 32);
    mant = low << (-sh);
  } else {
    mant = 0;
    for (int b = 52; b >= 0; b--) mant = (mant << 1) | (uint64_t)bit_at(L, sh + b);
    const bool rbit = bit_at(L, sh - 1);
    sticky |= any_below(L, sh - 1);
    if (rbit && (sticky || (mant & 1ull))) {
      mant++;
      if (mant == (1ull << 53)) {
        mant >>= 1;
        sh++;
      }
    }
  }
  int ex = sh - s;
  // normal doubles only: 2^52 <= mant < 2^53, so the value lies in
  // [2^(52 + ex), 2^(53 + ex))
  if (52 + ex < -1022 || 53 + ex > 1024) return false;
  double r = (double)mant;   // exact: mant < 2^53
  // scale by 2^ex in steps that keep every intermediate exact
  while (ex > 0) {
    const int st = ex > 60 ? 60 : ex;
    r *= (double)(1ull << st);
    ex -= st;
  }
  while (ex < 0) {
    const int st = -ex > 60 ? 60 : -ex;
    r /= (double)(1ull << st);
    ex += st;
  }
  *out = r;
  return true;
}

PSH_HD inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

// Parses the whole token [p, e); returns false when it is not a decimal
// number in the supported range (or too close to a rounding boundary after
// the 19th significant digit).
PSH_HD inline bool parse_double(const char *p, const char *e, double *out) {
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) {
    neg = *p == '-';
    p++;
  }
  uint64_t Mhi = 0, Mlo = 0;
  int nd = 0, dexp = 0;
  bool dropped = false, any = false;
  while (p < e && is_digit(*p)) {
    const int d = *p - '0';
    any = true;
    if (nd == 0 && d == 0) {
      p++;
      continue;
    }
    if (nd < 38) {
      mul10_add(Mhi, Mlo, (uint32_t)d);
      nd++;
    } else {
      dexp++;
      if (d) dropped = true;
    }
    p++;
  }
  if (p < e && *p == '.') {
    p++;
    while (p < e && is_digit(*p)) {
      const int d = *p - '0';
      any = true;
      if (nd == 0 && d == 0) {
        dexp--;
      } else if (nd < 38) {
        mul10_add(Mhi, Mlo, (uint32_t)d);
        nd++;
        dexp--;
      } else if (d) {
        dropped = true;
      }
      p++;
    }
  }
  if (!any) return false;
  if (p < e && (*p == 'e' || *p == 'E')) {
    p++;
    bool en = false;
    if (p < e && (*p == '+' || *p == '-')) {
      en = *p == '-';
      p++;
    }
    if (p >= e || !is_digit(*p)) return false;
    int ev = 0;
    while (p < e && is_digit(*p)) {
      if (ev < 100000) ev = ev * 10 + (*p - '0');
      p++;
    }
    dexp += en ? -ev : ev;
  }
  if (p != e) return false;
  double r;
  if (nd == 0) {
    r = 0.0;
  } else if (!dropped && Mhi == 0 && Mlo <= (1ull << 53) && dexp >= -22 && dexp <= 22) {
    r = (double)Mlo;
    if (dexp >= 0) r *= pow10_exact(dexp);
    else r /= pow10_exact(-dexp);
  } else {
    if (!round_exact(Mhi, Mlo, dexp, dropped, &r)) return false;
    if (dropped) {   // M + 1 < 10^38 < 2^127: no overflow
      const uint64_t lo1 = Mlo + 1, hi1 = Mhi + (lo1 == 0 ? 1 : 0);
      double up;
      if (!round_exact(hi1, lo1, dexp, false, &up) || up != r) return false;
    }
  }
  *out = neg ? -r : r;
  return true;
}

}  // namespace pshdec
