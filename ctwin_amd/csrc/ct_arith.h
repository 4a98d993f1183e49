// Exact modular and multi-word integer arithmetic shared by the gfx950 kernels
// and the host code of the squaring engine.
//
// Residues live modulo three primes p < 2^30 with 3 * 2^22 | p - 1, so that the
// transform modulo X^n - X^(n/2) + 1 exists for n <= 2^22 and the butterflies
// can work lazily on representatives in [0, 4p) (4p < 2^32; D. Harvey, "Faster
// arithmetic for number-theoretic transforms", J. Symbolic Comput. 60 (2014)
// 113-119).  A product by a table constant w uses Shoup's precomputed quotient
// w' = floor(w 2^32 / p) (V. Shoup, NTL): q = mulhi(a, w'), a w - q p lies in
// [0, 2p) for every 32-bit a -- three or four VALU instructions.  Products of
// two variables (the pointwise square) use Montgomery reduction with R = 2^32
// (P. L. Montgomery, Math. Comp. 44 (1985) 519-521); the CRT keeps the
// reduced [0, p) forms.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define CT_HD __host__ __device__ __forceinline__
#else
#define CT_HD inline
#endif

namespace ct {

// ---- prime set ------------------------------------------------------------

constexpr int kPrimes = 3;
constexpr uint32_t kP[kPrimes] = {943718401u, 918552577u, 880803841u};   // (75, 73, 70) * 3 * 2^22 + 1
constexpr uint32_t kG[kPrimes] = {7u, 5u, 26u};                           // primitive roots
constexpr int kMaxLog2N = 22;                                              // 3*2^22 | p-1 for all three

constexpr uint32_t inv32(uint32_t p)   // p^-1 mod 2^32 by Newton iteration
{
	uint32_t x = p;                      // correct to 3 bits for odd p
	for (int i = 0; i < 5; ++i) x *= 2u - p * x;
	return x;
}

constexpr uint32_t kQ[kPrimes] = {inv32(kP[0]), inv32(kP[1]), inv32(kP[2])};
static_assert(kQ[0] * kP[0] == 1u && kQ[1] * kP[1] == 1u && kQ[2] * kP[2] == 1u, "Montgomery inverse");
static_assert(kP[0] < 0x40000000u && kP[1] < 0x40000000u && kP[2] < 0x40000000u, "lazy [0, 4p) needs p < 2^30");
static_assert(kP[0] > kP[1] && kP[1] > kP[2], "the CRT assumes p1 > p2 > p3");

// ---- mod p, reduced forms [0, p) (CRT) ----------------------------------------

CT_HD uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }

CT_HD uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) { const uint32_t s = a + b; return umin32(s, s - p); }
CT_HD uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) { const uint32_t d = a - b; return umin32(d, d + p); }

#if defined(__HIP_DEVICE_COMPILE__)
CT_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return __umulhi(a, b); }
#else
CT_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
#endif

// Montgomery product a * b * 2^-32 mod p, a, b in [0, p).
CT_HD uint32_t mont_mul(uint32_t a, uint32_t b, uint32_t p, uint32_t q)
{
	const uint64_t t = (uint64_t)a * b;
	const uint32_t m = (uint32_t)t * q;
	const uint32_t r = (uint32_t)(t >> 32) - mulhi32(m, p);
	return umin32(r, r + p);
}

// Montgomery product by a constant w with wq = w * q mod 2^32 precomputed.
CT_HD uint32_t mont_mulc(uint32_t a, uint32_t w, uint32_t wq, uint32_t p)
{
	const uint32_t m = a * wq;
	const uint32_t r = mulhi32(a, w) - mulhi32(m, p);
	return umin32(r, r + p);
}

// ---- lazy forms (transform) -----------------------------------------------------

// [0, 4p) -> [0, 2p) and [0, 2p) -> [0, p): one subtract and one unsigned min
// (the subtraction wraps to a large value when no correction is due).
CT_HD uint32_t red2p(uint32_t a, uint32_t p) { return umin32(a, a - 2 * p); }
CT_HD uint32_t red1p(uint32_t a, uint32_t p) { return umin32(a, a - p); }

// Shoup product: a w mod p in [0, 2p) for any 32-bit a, w in [0, p), wp = floor(w 2^32 / p).
CT_HD uint32_t shoup(uint32_t a, uint32_t w, uint32_t wp, uint32_t p) { return a * w - mulhi32(a, wp) * p; }

// Negated Shoup product: with wn = 2^32 - w, returns 2^32 - t (mod 2^32) where
// t = shoup(a, w, wp, p); so x - result = x + t and x + result + 2p = x - t + 2p.
// (a wn + q p is one mul_lo and one mad_u64_u32 on gfx950.)
CT_HD uint32_t shoup_neg(uint32_t a, uint32_t wn, uint32_t wp, uint32_t p) { return a * wn + mulhi32(a, wp) * p; }

// Montgomery product of a, b in [0, 2p): a b 2^-32 mod p in (0, 2p).
CT_HD uint32_t mont_lazy(uint32_t a, uint32_t b, uint32_t p, uint32_t q)
{
	const uint64_t t = (uint64_t)a * b;                 // < 4 p^2
	const uint32_t m = (uint32_t)t * q;
	return (uint32_t)(t >> 32) - mulhi32(m, p) + p;     // (t - m p) / 2^32 in (-p, p)
}

// Signed 32-bit digit -> representative in [0, 3p): d + 3p for d < 0 (d >= -2^31
// and 3p > 2^31 keep it non-negative; 3p < 2^32), d itself otherwise (d < 2^31 < 3p).
static_assert(3ull * kP[2] > 0x80000000ull && 3ull * kP[0] < 0x100000000ull, "digit range");
CT_HD uint32_t from_signed(int32_t d, uint32_t p) { return (d < 0) ? (uint32_t)d + 3 * p : (uint32_t)d; }

// Two-prime mode only: there every digit is below 2^28 in magnitude
// (3 n D^2 < p1 p2 / 2 with n >= 32 gives D < 2^26.01, and a pending carry
// adds less than another D), far inside (-2p, 2p), so d + 2p is a
// representative in [0, 4p) -- the input range of the first forward level --
// in one instruction.
template <int NP>
CT_HD uint32_t from_signed_np(int32_t d, uint32_t p)
{
	if constexpr (NP == 2) return (uint32_t)d + 2 * p;
	else return from_signed(d, p);
}

// ---- 32-bit primality -----------------------------------------------------------

// Strong probable-prime test of odd p > 41 to the bases 2, 7 and 61
// (deterministic below 4,759,123,141, so exact for every 32-bit p; G. Jaeschke,
// Math. Comp. 61 (1993) 915-926), in 32-bit Montgomery form.  No 64-bit
// remainder: R = 2^32 mod p is 2^32 - p when p > 2^31 and one 32-bit remainder
// otherwise, and the bases' Montgomery forms are sums of modular doublings of
// R (7 = 8 - 1, 61 = 64 - 2 - 1).  spsp_base tests one base (0: 2, 1: 7,
// 2: 61), so that a caller can test base 2 on every candidate and the others
// on the few that pass it.
CT_HD bool spsp_base(uint32_t p, int which)
{
	uint32_t q = p;
	for (int i = 0; i < 5; ++i) q *= 2u - p * q;                    // p^-1 mod 2^32
	auto mul = [&](uint32_t a, uint32_t b) {
		const uint64_t t = (uint64_t)a * b;
		const uint32_t m = (uint32_t)t * q, th = (uint32_t)(t >> 32), mp = mulhi32(m, p);
		return (th < mp) ? th - mp + p : th - mp;
	};
	auto add = [&](uint32_t a, uint32_t b) { const uint32_t t = p - b; return (a >= t) ? a - t : a + b; };
	auto sub = [&](uint32_t a, uint32_t b) { return (a >= b) ? a - b : a - b + p; };
	const uint32_t one = (p & 0x80000000u) ? 0u - p : (0u - p) % p;
	const uint32_t m1 = p - one;                                     // -1
	const uint32_t b2 = add(one, one);
	uint32_t a = b2;
	if (which > 0)
	{
		const uint32_t b8 = add(add(b2, b2), add(b2, b2));
		if (which == 1) a = sub(b8, one);
		else
		{
			uint32_t x = add(b8, b8); x = add(x, x); x = add(x, x);   // 64
			a = sub(x, add(b2, one));
		}
	}
	if (a == 0) return true;                                         // p divides the base (p = 61)
	uint32_t d = p - 1;
	int s = 0;
	while ((d & 1) == 0) { d >>= 1; ++s; }
	uint32_t x = one;
	for (uint32_t e = d; e; e >>= 1)
	{
		if (e & 1) x = mul(x, a);
		a = mul(a, a);
	}
	if (x == one || x == m1) return true;
	for (int r = 1; r < s; ++r)
	{
		x = mul(x, x);
		if (x == m1) return true;
	}
	return false;
}

CT_HD bool spsp_2_7_61(uint32_t p) { return spsp_base(p, 0) && spsp_base(p, 1) && spsp_base(p, 2); }

// ---- host-side (or constexpr) plain modular helpers -------------------------

CT_HD uint32_t mulmod_plain(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)(((uint64_t)a * b) % p); }

CT_HD uint32_t powmod_plain(uint32_t a, uint64_t e, uint32_t p)
{
	uint64_t r = 1, y = a % p;
	while (e) { if (e & 1) r = (r * y) % p; y = (y * y) % p; e >>= 1; }
	return (uint32_t)r;
}

CT_HD uint32_t to_mont(uint32_t a, uint32_t p) { return (uint32_t)((((uint64_t)a) << 32) % p); }

// ---- signed 96-bit accumulator ----------------------------------------------

struct i96 { uint64_t lo; int32_t hi; };

CT_HD i96 i96_add(i96 x, i96 y)
{
	i96 r; r.lo = x.lo + y.lo;
	r.hi = x.hi + y.hi + ((r.lo < y.lo) ? 1 : 0);
	return r;
}

CT_HD i96 i96_from_i64(int64_t v) { i96 r; r.lo = (uint64_t)v; r.hi = (v < 0) ? -1 : 0; return r; }

CT_HD bool i96_neg(i96 x) { return x.hi < 0; }

CT_HD i96 i96_negate(i96 x)
{
	i96 r; r.lo = ~x.lo + 1;
	r.hi = ~x.hi + ((r.lo == 0) ? 1 : 0);
	return r;
}

// ---- exact division by the base b ---------------------------------------
//
// Division of a two-limb value (x1 < d) by a normalised divisor
// d in [2^31, 2^32) with the precomputed reciprocal v = floor((2^64-1)/d) - 2^32
// (N. Moller and T. Granlund, "Improved division by invariant integers",
// IEEE Trans. Computers 60 (2011), algorithm 4).  Quotient < 2^32.

struct divisor
{
	uint32_t d; uint32_t v; uint32_t sh; uint32_t b;   // normalised divisor, its 2-by-1 reciprocal, shift, b
	uint32_t m_lo, m_hi;                               // m = floor((2^64 - 1) / b), the 64-bit reciprocal
	uint32_t k_lo, k_hi, c_lo, c_hi;                   // K = c b, c = floor(2^62 / b) + 1 (divrem_i62)
};

CT_HD divisor make_divisor(uint32_t b)
{
	divisor r; r.b = b;
	uint32_t sh = 0; while (((b << sh) & 0x80000000u) == 0) ++sh;
	r.sh = sh; r.d = b << sh;
	r.v = (uint32_t)((~(uint64_t)0) / r.d - ((uint64_t)1 << 32));
	const uint64_t m = (~(uint64_t)0) / b;
	r.m_lo = (uint32_t)m; r.m_hi = (uint32_t)(m >> 32);
	const uint64_t c = ((uint64_t)1 << 62) / b + 1, k = c * b;
	r.c_lo = (uint32_t)c; r.c_hi = (uint32_t)(c >> 32);
	r.k_lo = (uint32_t)k; r.k_hi = (uint32_t)(k >> 32);
	return r;
}

CT_HD uint32_t div_2by1(uint32_t x1, uint32_t x0, const divisor &dv, uint32_t *rem)
{
	const uint64_t qq = (uint64_t)dv.v * x1 + (((uint64_t)x1 << 32) | x0);
	uint32_t q1 = (uint32_t)(qq >> 32) + 1;
	const uint32_t q0 = (uint32_t)qq;
	uint32_t r = x0 - q1 * dv.d;
	if (r > q0) { q1 -= 1; r += dv.d; }
	if (r >= dv.d) { q1 += 1; r -= dv.d; }
	*rem = r;
	return q1;
}

// |t| / b with t unsigned 96-bit (t2, t1, t0) and t < 2^64 * b:
// returns the 64-bit quotient and the remainder.
CT_HD uint64_t udiv96(uint32_t t2, uint32_t t1, uint32_t t0, const divisor &dv, uint32_t *rem)
{
	const uint32_t sh = dv.sh;
	// normalise: shift the dividend left by sh (t2 < b ensures no overflow)
	const uint32_t n2 = sh ? ((t2 << sh) | (t1 >> (32 - sh))) : t2;
	const uint32_t n1 = sh ? ((t1 << sh) | (t0 >> (32 - sh))) : t1;
	const uint32_t n0 = t0 << sh;
	uint32_t r;
	const uint32_t qh = div_2by1(n2, n1, dv, &r);
	const uint32_t ql = div_2by1(r, n0, dv, &r);
	*rem = r >> sh;
	return ((uint64_t)qh << 32) | ql;
}

// Signed division with truncation toward zero (C semantics), as the carry
// steps require: a = q * b + r with |r| < b and r of the sign of a.
CT_HD int32_t divrem_i96(i96 &a, const divisor &dv)
{
	const bool s = i96_neg(a);
	const i96 t = s ? i96_negate(a) : a;
	uint32_t r;
	const uint64_t q = udiv96((uint32_t)t.hi, (uint32_t)(t.lo >> 32), (uint32_t)t.lo, dv, &r);
	a = i96_from_i64(s ? -(int64_t)q : (int64_t)q);
	return s ? -(int32_t)r : (int32_t)r;
}

// |a| < 2^63 divided by b with the 64-bit reciprocal m = floor((2^64-1)/b):
// q' = floor(|a| m / 2^64) is q or q - 1 (|a| (2^64 - 1)/b - |a| - 1 < |a| m
// <= |a| 2^64 / b and |a| < 2^63), so one correction step gives the exact
// quotient; the remainder < 2b < 2^32 comes from the low words alone.
CT_HD uint64_t udiv64_recip(uint64_t u, const divisor &dv, uint32_t *rem)
{
	const uint32_t u0 = (uint32_t)u, u1 = (uint32_t)(u >> 32);
	const uint64_t s1 = (uint64_t)u0 * dv.m_hi + mulhi32(u0, dv.m_lo);             // < 2^64
	const uint64_t s2 = (uint64_t)u1 * dv.m_lo + (uint32_t)s1;                       // < 2^64
	uint64_t q = (uint64_t)u1 * dv.m_hi + ((s1 >> 32) + (s2 >> 32));                 // floor(u m / 2^64)
	uint32_t r = u0 - (uint32_t)q * dv.b;
	if (r >= dv.b) { r -= dv.b; q += 1; }
	*rem = r;
	return q;
}

// The same truncating division for |a| < 2^62 - b, through a nonnegative
// dividend: A = a + K (K = c b in (2^62, 2^62 + b]) lies in [0, 2^63); its
// quotient is floor(a / b) + c and its remainder a mod b in [0, b).  A
// negative a with a nonzero remainder then takes the truncated quotient (+1)
// and remainder (-b).  Fewer instructions than dividing |a| and restoring
// both signs.
CT_HD int32_t divrem_i62(int64_t &a, const divisor &dv)
{
	const uint64_t K = ((uint64_t)dv.k_hi << 32) | dv.k_lo, c = ((uint64_t)dv.c_hi << 32) | dv.c_lo;
	uint32_t r;
	const uint64_t q = udiv64_recip((uint64_t)a + K, dv, &r) - c;
	const bool fix = (a < 0) && (r != 0);
	a = (int64_t)q + (fix ? 1 : 0);
	return fix ? (int32_t)r - (int32_t)dv.b : (int32_t)r;
}

CT_HD int32_t divrem_i64(int64_t &a, const divisor &dv)
{
	const bool s = (a < 0);
	const uint64_t t = s ? (uint64_t)(-a) : (uint64_t)a;
	uint32_t r;
	const uint64_t q = udiv64_recip(t, dv, &r);
	a = s ? -(int64_t)q : (int64_t)q;
	return s ? -(int32_t)r : (int32_t)r;
}

// ---- Garner CRT constants for the engine primes --------------------------
//
// x = r3 + p3 * (u23 + p2 * u123) with
// u23 = (r2 - r3) / p3 mod p2, u13 = (r1 - r3) / p3 mod p1,
// u123 = (u13 - u23) / p2 mod p1; centred in (-M/2, M/2].

struct garner_consts
{
	uint32_t inv3_1, inv3_1q;   // Montgomery form of p3^-1 mod p1, and its q-product
	uint32_t inv3_2, inv3_2q;   // p3^-1 mod p2
	uint32_t inv2_1, inv2_1q;   // p2^-1 mod p1
	uint32_t p3_mod_p1, p3_mod_p2;
	uint32_t m_lo, m_mid, m_hi;        // M = p1 p2 p3 (96-bit)
	uint32_t h_lo, h_mid, h_hi;        // floor(M / 2)
	uint64_t m2, h2;                   // two-prime mode: M2 = p1 p2, floor(M2 / 2)
};

}   // namespace ct
