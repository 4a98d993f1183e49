// Number theory shared by the sieve kernels and their host fallbacks:
// Montgomery arithmetic for odd p < 2^32 (device) and p < 2^63 (device and
// host), Miller-Rabin, the 2-Sylow generator and 2^d-th roots mod p.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ct {

// ---------------------------------------------------------------------------
// modular arithmetic: 32-bit Montgomery on the device (odd p < 2^32),
// plain 128-bit arithmetic on the host (odd p < 2^63); same interface.

struct m32
{
	uint32_t p, q, one, r2;

	__device__ explicit m32(uint32_t p_) : p(p_)
	{
		uint32_t x = p;
		for (int i = 0; i < 5; ++i) x *= 2u - p * x;
		q = x;
		one = (uint32_t)((0x100000000ull) % p);
		r2 = (uint32_t)(((uint64_t)one * one) % p);
	}
	__device__ uint32_t mul(uint32_t a, uint32_t b) const
	{
		const uint64_t t = (uint64_t)a * b;
		const uint32_t m = (uint32_t)t * q, th = (uint32_t)(t >> 32), mp = __umulhi(m, p);
		const uint32_t r = th - mp;
		return (th < mp) ? r + p : r;
	}
	__device__ uint32_t add(uint32_t a, uint32_t b) const { const uint32_t s = a + b; return (s < a || s >= p) ? s - p : s; }
	__device__ uint32_t sub(uint32_t a, uint32_t b) const { return (a >= b) ? a - b : a - b + p; }
	__device__ uint32_t to_m(uint32_t x) const { return mul(x % p, r2); }
	__device__ uint32_t from_m(uint32_t x) const { return mul(x, 1u); }
	__device__ uint32_t neg_one() const { return p - one; }
	__device__ uint32_t pow(uint32_t a, uint64_t e) const
	{
		uint32_t r = one;
		while (e) { if (e & 1) r = mul(r, a); a = mul(a, a); e >>= 1; }
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


#define CT_HD2 __host__ __device__ __forceinline__

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

// Generator of the 2-Sylow subgroup (order 2^v), from the first quadratic
// non-residue c: w = c^q, p - 1 = 2^v q.
template <class M, typename U>
CT_HD2 U two_sylow_gen(const M &A, U p, U q)
{
	for (uint32_t c = 2;; ++c)
	{
		const U cm = A.to_m((U)c);
		if (A.pow(cm, (p - 1) / 2) == A.neg_one()) return A.pow(cm, q);
	}
}

// One solution of x^(2^d) = y (Montgomery form), or 0 when there is none.
// p - 1 = 2^v q, q odd, w generator of the 2-Sylow subgroup.
template <class M, typename U>
CT_HD2 U root_2d(const M &A, U p, U q, int v, U w, U y, int d)
{
	const int e = (v < d) ? v : d;
	if (A.pow(y, (p - 1) >> e) != A.one) return 0;
	// alpha = 2^-d mod q (q odd) by d halvings modulo q: a/2 for even a,
	// (a + q)/2 = floor(a/2) + floor(q/2) + 1 for odd a -- no remainders
	uint64_t alpha = 1 % (uint64_t)q;
	for (int i = 0; i < d; ++i) alpha = (alpha & 1) ? (alpha >> 1) + ((uint64_t)q >> 1) + 1 : alpha >> 1;
	const U x1 = A.pow(y, alpha);
	if (d >= v) return x1;          // y lies in the odd part: x1^(2^d) = y
	// eta = y * x1^(-2^d) lies in the 2-Sylow subgroup and is a 2^d-th power there
	U x1d = x1;
	for (int i = 0; i < d; ++i) x1d = A.mul(x1d, x1d);
	const U eta = A.mul(y, A.pow(x1d, p - 2));
	// eta = W^k, W = w^(2^d) of order 2^L, L = v - d: bits of k from the bottom
	const int L = v - d;
	U W = w;
	for (int i = 0; i < d; ++i) W = A.mul(W, W);
	U Wi = A.pow(W, p - 2);         // W^-1, squared along
	U mu = eta;
	uint64_t k = 0;
	for (int j = 0; j < L; ++j)
	{
		U t = mu;
		for (int i = 0; i < L - 1 - j; ++i) t = A.mul(t, t);
		if (t != A.one) { k |= (uint64_t)1 << j; mu = A.mul(mu, Wi); }
		Wi = A.mul(Wi, Wi);
	}
	return A.mul(x1, A.pow(w, k));
}


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

}   // namespace ct
