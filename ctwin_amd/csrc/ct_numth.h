// Number theory shared by the sieve kernels and their host fallbacks:
// Montgomery arithmetic for odd p < 2^32 (device) and p < 2^63 (device and
// host), Miller-Rabin, the 2-Sylow generator and 2^d-th roots mod p.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CT_HD2 __host__ __device__ __forceinline__
#else
#define CT_HD2 inline
#endif

#include <cstdint>

#include "ct_arith.h"

namespace ct {

// ---------------------------------------------------------------------------
// modular arithmetic: 32-bit Montgomery (odd p < 2^32, device and host),
// plain 128-bit arithmetic on the host (odd p < 2^63); same interface.

struct m32
{
	uint32_t p, q, one, r2;

	CT_HD2 explicit m32(uint32_t p_) : p(p_)
	{
		uint32_t x = p;
		for (int i = 0; i < 5; ++i) x *= 2u - p * x;
		q = x;
		one = (0u - p) % p;                                  // 2^32 mod p
		// 2^64 mod p without 64-bit remainders: 2^36 by four doublings, then
		// three Montgomery squarings (x^2 / 2^32): 2^40, 2^48, 2^64
		uint32_t t = add(one, one);
		t = add(t, t); t = add(t, t); t = add(t, t);
		t = mul(t, t); t = mul(t, t);
		r2 = mul(t, t);
	}
	CT_HD2 uint32_t mul(uint32_t a, uint32_t b) const
	{
		const uint64_t t = (uint64_t)a * b;
		const uint32_t m = (uint32_t)t * q, th = (uint32_t)(t >> 32), mp = mulhi32(m, p);
		const uint32_t r = th - mp;
		return (th < mp) ? r + p : r;
	}
	CT_HD2 uint32_t add(uint32_t a, uint32_t b) const { const uint32_t s = a + b; return (s < a || s >= p) ? s - p : s; }
	CT_HD2 uint32_t sub(uint32_t a, uint32_t b) const { return (a >= b) ? a - b : a - b + p; }
	CT_HD2 uint32_t to_m(uint32_t x) const { return mul(x % p, r2); }
	CT_HD2 uint32_t from_m(uint32_t x) const { return mul(x, 1u); }
	CT_HD2 uint32_t neg_one() const { return p - one; }
	// Square-and-multiply without branches: each step multiplies by a or by
	// the Montgomery one.  Exponents are below 2^32 (every caller's is below p).
	CT_HD2 uint32_t pow(uint32_t a, uint32_t e) const
	{
		uint32_t r = one;
		for (; e; e >>= 1)
		{
			r = mul(r, (e & 1) ? a : one);
			a = mul(a, a);
		}
		return r;
	}
};

struct m64h
{
	uint64_t p, one;
	explicit m64h(uint64_t p_) : p(p_), one(1) {}
	uint64_t mul(uint64_t a, uint64_t b) const { return (uint64_t)(((unsigned __int128)a * b) % p); }
	uint64_t add(uint64_t a, uint64_t b) const { const unsigned __int128 s = (unsigned __int128)a + b; return (uint64_t)(s >= p ? s - p : s); }
	uint64_t sub(uint64_t a, uint64_t b) const { return (a >= b) ? a - b : a + (p - b); }
	uint64_t to_m(uint64_t x) const { return x % p; }
	uint64_t from_m(uint64_t x) const { return x; }
	uint64_t neg_one() const { return p - 1; }
	uint64_t pow(uint64_t a, uint64_t e) const
	{
		uint64_t r = 1;
		while (e) { if (e & 1) r = mul(r, a); a = mul(a, a); e >>= 1; }
		return r;
	}
};


// Miller-Rabin with the given bases.
template <class M, typename U>
CT_HD2 bool mr_prime(const M &A, U p, const uint32_t *bases, int nb)
{
	U d = p - 1; int s = 0;
	while ((d & 1) == 0) { d >>= 1; ++s; }
	for (int i = 0; i < nb; ++i)
	{
		const U a = (U)bases[i] % p;
		if (a == 0) continue;
		U x = A.pow(A.to_m(a), d);
		if (x == A.one || x == A.neg_one()) continue;
		bool ok = false;
		for (int r = 1; r < s; ++r) { x = A.mul(x, x); if (x == A.neg_one()) { ok = true; break; } }
		if (!ok) return false;
	}
	return true;
}

// Quadratic-residue masks modulo the odd primes below 64: bit r is set when r
// is a nonzero square modulo c.
constexpr uint64_t qr_mask(uint32_t c)
{
	uint64_t m = 0;
	for (uint32_t x = 1; x < c; ++x) m |= 1ull << (x * x % c);
	return m;
}
constexpr uint32_t kOddPrimes64[17] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};
constexpr uint64_t kQrMask64[17] = {qr_mask(3), qr_mask(5), qr_mask(7), qr_mask(11), qr_mask(13), qr_mask(17),
                                    qr_mask(19), qr_mask(23), qr_mask(29), qr_mask(31), qr_mask(37), qr_mask(41),
                                    qr_mask(43), qr_mask(47), qr_mask(53), qr_mask(59), qr_mask(61)};

// The least quadratic non-residue c of an odd prime p.  c is prime (a
// product of residues is a residue).  Below 64 its character comes from p
// mod c: (2/p) = -1 iff p = 3, 5 mod 8, and for odd c != p quadratic
// reciprocity gives (c/p) = (p/c) (-1)^((c-1)/2 (p-1)/2).  Past 61 the Euler
// criterion decides.
template <class M, typename U>
CT_HD2 uint32_t least_nonresidue(const M &A, U p)
{
	const uint32_t p8 = (uint32_t)(p & 7u);
	if (p8 == 3 || p8 == 5) return 2;
	const bool p3 = (p & 3u) == 3;
#pragma unroll
	for (int i = 0; i < 17; ++i)
	{
		const uint32_t c = kOddPrimes64[i];
		const uint32_t r = (uint32_t)(p % c);
		if (r == 0) continue;                              // c = p
		const bool res = (kQrMask64[i] >> r) & 1u;          // (p/c) = 1
		const bool flip = ((c & 3u) == 3) && p3;
		if (res == flip) return c;
	}
	for (uint32_t c = 67;; ++c)
		if (A.pow(A.to_m((U)c), (p - 1) / 2) == A.neg_one()) return c;
}

// Generator of the 2-Sylow subgroup (order 2^v) of an odd prime p: w = c^q,
// p - 1 = 2^v q, c the least quadratic non-residue.
template <class M, typename U>
CT_HD2 U two_sylow_gen(const M &A, U p, U q)
{
	return A.pow(A.to_m((U)least_nonresidue(A, p)), q);
}

// a^e1 and a^e2 over one squaring chain (multiplications by a or by one,
// without branches).
template <class M, typename U>
CT_HD2 void pow_pair(const M &A, U a, uint64_t e1, uint64_t e2, U &r1, U &r2)
{
	U x = A.one, y = A.one;
	while (e1 | e2)
	{
		x = A.mul(x, (e1 & 1) ? a : A.one);
		y = A.mul(y, (e2 & 1) ? a : A.one);
		e1 >>= 1; e2 >>= 1;
		a = A.mul(a, a);
	}
	r1 = x; r2 = y;
}

// a^ea and b^eb in one loop: two independent chains, for latency.  E is the
// exponent type (32 bits suffice for p < 2^32).
template <class M, typename U, typename E = uint64_t>
CT_HD2 void pow_dual(const M &A, U a, E ea, U b, E eb, U &ra, U &rb)
{
	U x = A.one, y = A.one;
	while (ea | eb)
	{
		x = A.mul(x, (ea & 1) ? a : A.one);
		y = A.mul(y, (eb & 1) ? b : A.one);
		ea >>= 1; eb >>= 1;
		a = A.mul(a, a); b = A.mul(b, b);
	}
	ra = x; rb = y;
}

// a^e1, a^e2 and b^e2 in one loop (a's chain shared by the first two).
template <class M, typename U, typename E = uint64_t>
CT_HD2 void pow_tri(const M &A, U a, U b, E e1, E e2, U &r1, U &r2, U &r3)
{
	U x = A.one, y = A.one, z = A.one;
	while (e1 | e2)
	{
		x = A.mul(x, (e1 & 1) ? a : A.one);
		y = A.mul(y, (e2 & 1) ? a : A.one);
		z = A.mul(z, (e2 & 1) ? b : A.one);
		e1 >>= 1; e2 >>= 1;
		a = A.mul(a, a); b = A.mul(b, b);
	}
	r1 = x; r2 = y; r3 = z;
}

// alpha = 2^-d mod q (q odd) by d halvings modulo q: a/2 for even a,
// (a + q)/2 = floor(a/2) + floor(q/2) + 1 for odd a, which adds 2^i to c, so
// that alpha 2^d = 1 + c q -- no remainders.
CT_HD2 void inv_pow2_mod(uint64_t q, int d, uint64_t &alpha, int64_t &c)
{
	alpha = (q == 1) ? 0 : 1;                 // 1 mod q
	c = (q == 1) ? -1 : 0;
	for (int i = 0; i < d; ++i)
	{
		if (alpha & 1) { alpha = (alpha >> 1) + (q >> 1) + 1; c += (int64_t)1 << i; }
		else alpha >>= 1;
	}
}

// z^-1 for z in the 2-Sylow subgroup (z^(2^v) = 1): z^(2^v - 1) by x <- x^2 z.
template <class M, typename U>
CT_HD2 U inv_2sylow(const M &A, U z, int v)
{
	U x = z;
	for (int i = 1; i < v; ++i) x = A.mul(A.mul(x, x), z);
	return x;
}

// Whether y is a 2^d-th power, from z = y^q: y^((p-1)/2^e) = z^(2^(v-e)) = 1,
// e = min(v, d).
template <class M, typename U>
CT_HD2 bool is_pow2d(const M &A, U z, int v, int d)
{
	const int e = (v < d) ? v : d;
	U t = z;
	for (int i = 0; i < v - e; ++i) t = A.mul(t, t);
	return t == A.one;
}

// A 2^d-th root of y (a 2^d-th power, Montgomery form) from x1 = y^alpha and
// z = y^q (alpha 2^d = 1 + c q).  p - 1 = 2^v q, q odd, w generator of the
// 2-Sylow subgroup.  x1 is a 2^d-th root of y up to the 2-Sylow factor
// eta = y x1^(-2^d) = y^(1 - alpha 2^d) = z^(-c); eta = W^k with W = w^(2^d)
// of order 2^L, L = v - d, and the root is x1 w^k.
template <class M, typename U>
CT_HD2 U root_2d_finish(const M &A, int v, U w, U x1, U z, int64_t c, int d)
{
	if (d >= v) return x1;          // y lies in the odd part: x1^(2^d) = y
	const uint64_t vmask = ((uint64_t)1 << v) - 1;
	const U eta = A.pow(z, (uint64_t)(-c) & vmask);
	const int L = v - d;
	U W = w;
	for (int i = 0; i < d; ++i) W = A.mul(W, W);
	U Wi = inv_2sylow(A, W, L);     // W^-1, squared along
	U mu = eta;
	uint64_t k = 0;
	for (int j = 0; j < L; ++j)
	{
		U s = mu;
		for (int i = 0; i < L - 1 - j; ++i) s = A.mul(s, s);
		if (s != A.one) { k |= (uint64_t)1 << j; mu = A.mul(mu, Wi); }
		Wi = A.mul(Wi, Wi);
	}
	return A.mul(x1, A.pow(w, k));
}

// One solution of x^(2^d) = y (Montgomery form), or 0 when there is none.
// z = y^q and x1 = y^alpha share one squaring chain; for d = 1, alpha =
// (q + 1)/2 and both come from t = y^((q-1)/2): x1 = t y, z = t x1.
template <class M, typename U>
CT_HD2 U root_2d(const M &A, U p, U q, int v, U w, U y, int d)
{
	(void)p;
	uint64_t alpha;
	int64_t c;
	inv_pow2_mod((uint64_t)q, d, alpha, c);
	U z, x1;
	if (d == 1 && q > 1)
	{
		const U t = A.pow(y, ((uint64_t)q - 1) / 2);
		x1 = A.mul(t, y);
		z = A.mul(t, x1);
	}
	else pow_pair(A, y, (uint64_t)q, alpha, z, x1);
	if (!is_pow2d(A, z, v, d)) return 0;
	return root_2d_finish(A, v, w, x1, z, c, d);
}


#if defined(__HIPCC__)
// 64-bit Montgomery (R = 2^64) for odd p < 2^63, device side.
struct m64d
{
	uint64_t p, q, one, r2;

	__device__ explicit m64d(uint64_t p_) : p(p_)
	{
		uint64_t x = p;
		for (int i = 0; i < 6; ++i) x *= 2 - p * x;
		q = x;
		one = (0 - p) % p;
		uint64_t t = add(one, one); t = add(t, t); t = add(t, t); t = add(t, t);   // 16
		for (int i = 0; i < 4; ++i) t = mul(t, t);                              // 16^16 = 2^64
		r2 = t;
	}
	__device__ uint64_t mul(uint64_t a, uint64_t b) const
	{
		const uint64_t lo = a * b, hi = __umul64hi(a, b);
		const uint64_t mp = __umul64hi(lo * q, p);
		const uint64_t r = hi - mp;
		return (hi < mp) ? r + p : r;
	}
	__device__ uint64_t add(uint64_t a, uint64_t b) const { const uint64_t c = (a >= p - b) ? p : 0; return a + b - c; }
	__device__ uint64_t sub(uint64_t a, uint64_t b) const { const uint64_t c = (a < b) ? p : 0; return a - b + c; }
	__device__ uint64_t to_m(uint64_t x) const { return mul(x % p, r2); }
	__device__ uint64_t from_m(uint64_t x) const { return mul(x, 1); }
	__device__ uint64_t neg_one() const { return p - one; }
	__device__ uint64_t pow(uint64_t a, uint64_t e) const
	{
		uint64_t r = one;
		while (e) { if (e & 1) r = mul(r, a); a = mul(a, a); e >>= 1; }
		return r;
	}
};

#endif

}   // namespace ct
