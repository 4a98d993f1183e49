// Host check of the sieve's root construction in ctwin_amd/csrc/ct_numth.h (the
// same header, compiled by g++ with its 128-bit host arithmetic):
//  - two_sylow_gen picks c^q for the least quadratic non-residue c, found by a
//    scan of every integer c = 2, 3, ... with the Euler criterion;
//  - root_2d returns the same root as the direct construction below (Euler
//    check by y^((p-1)/2^e), inversions by Fermat's little theorem), and
//    the root it returns satisfies x^(2^d) = y; 0 exactly when y is no
//    2^d-th power;
//  - the two sides of the sieve's N- roots: for y1 = -1/y0, y1^q equals
//    -(y0^q)^-1 by inv_2sylow, is_pow2d on it agrees with the Euler check of
//    y1, and root_2d_finish from it gives the direct construction's root.
//  - the 32-bit Montgomery arithmetic m32 of the sieve kernels (constructor
//    constants, products, branch-free powers, the two- and three-chain power
//    loops) against 128-bit arithmetic, and
//    the generator and roots computed with it against the 128-bit ones.
// Over every prime p = 10k +- 1 below 2^18, the primes p = 1 mod 2^12 below
// 2^26, the Fermat primes, and random primes below 2^32 and 2^62.
#include <cstdio>
#include <random>
#include <vector>

#include "../../ctwin_amd/csrc/ct_numth.h"

using namespace ct;

static uint64_t gen_scan(const m64h &A, uint64_t p, uint64_t q)
{
	for (uint64_t c = 2;; ++c)
		if (A.pow(c % p, (p - 1) / 2) == p - 1) return A.pow(c % p, q);
}

static uint64_t root_direct(const m64h &A, uint64_t p, uint64_t q, int v, uint64_t w, uint64_t y, int d)
{
	const int e = (v < d) ? v : d;
	if (A.pow(y, (p - 1) >> e) != 1) return 0;
	uint64_t alpha = 1 % q;
	for (int i = 0; i < d; ++i) alpha = (alpha & 1) ? (alpha + q) / 2 : alpha / 2;
	const uint64_t x1 = A.pow(y, alpha);
	if (d >= v) return x1;
	uint64_t x1d = x1;
	for (int i = 0; i < d; ++i) x1d = A.mul(x1d, x1d);
	const uint64_t eta = A.mul(y, A.pow(x1d, p - 2));
	const int L = v - d;
	uint64_t W = w;
	for (int i = 0; i < d; ++i) W = A.mul(W, W);
	uint64_t Wi = A.pow(W, p - 2), mu = eta, k = 0;
	for (int j = 0; j < L; ++j)
	{
		uint64_t t = mu;
		for (int i = 0; i < L - 1 - j; ++i) t = A.mul(t, t);
		if (t != 1) { k |= 1ull << j; mu = A.mul(mu, Wi); }
		Wi = A.mul(Wi, Wi);
	}
	return A.mul(x1, A.pow(w, k));
}

static bool is_prime(uint64_t n)
{
	if (n < 2) return false;
	static const uint32_t b[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
	for (uint32_t s : b) { if (n == s) return true; if (n % s == 0) return false; }
	return mr_prime<m64h, uint64_t>(m64h(n), n, b, 12);
}

static long bad = 0, checked = 0, roots = 0, m32_checked = 0;

// m32 against m64h for an odd prime p < 2^32
static void check_m32(uint32_t p, std::mt19937_64 &rng)
{
	const m32 A(p);
	const m64h H(p);
	bool ok = A.one == (uint32_t)((1ull << 32) % p);
	ok = ok && A.r2 == (uint32_t)(((unsigned __int128)1 << 64) % p);
	for (int t = 0; t < 20 && ok; ++t)
	{
		const uint32_t a = (uint32_t)rng(), b = (uint32_t)rng(), e = (t < 2) ? (uint32_t)t : (uint32_t)rng() >> (t % 32);
		const uint32_t am = A.to_m(a), bm = A.to_m(b);
		ok = ok && A.from_m(am) == a % p;
		ok = ok && A.from_m(A.mul(am, bm)) == H.mul(a % p, b % p);
		ok = ok && A.from_m(A.add(am, bm)) == H.add(a % p, b % p);
		ok = ok && A.from_m(A.sub(am, bm)) == H.sub(a % p, b % p);
		ok = ok && A.from_m(A.pow(am, e)) == H.pow(a % p, e);
		uint32_t x1, x2;
		pow_pair<m32, uint32_t>(A, am, e, (uint32_t)rng(), x1, x2);
		ok = ok && A.from_m(x1) == H.pow(a % p, e);
		const uint32_t e2 = (uint32_t)rng() >> (t % 32);
		uint32_t d1, d2, t1, t2, t3;
		pow_dual<m32, uint32_t>(A, am, e, bm, e2, d1, d2);
		ok = ok && A.from_m(d1) == H.pow(a % p, e) && A.from_m(d2) == H.pow(b % p, e2);
		pow_tri<m32, uint32_t>(A, am, bm, e, e2, t1, t2, t3);
		ok = ok && A.from_m(t1) == H.pow(a % p, e) && A.from_m(t2) == H.pow(a % p, e2) && A.from_m(t3) == H.pow(b % p, e2);
	}
	if (p > 5)
	{
		uint32_t q = p - 1; int v = 0;
		while ((q & 1) == 0) { q >>= 1; ++v; }
		const uint32_t w = two_sylow_gen<m32, uint32_t>(A, p, q);
		const uint64_t wh = two_sylow_gen<m64h, uint64_t>(H, p, q);
		ok = ok && A.from_m(w) == wh;
		for (int d : {1, 11})
		{
			const uint32_t y = (uint32_t)(rng() % (p - 1)) + 1;
			const uint32_t x = root_2d<m32, uint32_t>(A, p, q, v, w, A.to_m(y), d);
			const uint64_t xh = root_2d<m64h, uint64_t>(H, p, q, v, wh, y, d);
			ok = ok && A.from_m(x) == xh;
		}
	}
	++m32_checked;
	if (!ok && bad++ < 10) std::printf("m32 p=%u\n", p);
}

static void check_prime(uint64_t p, std::mt19937_64 &rng, int ny)
{
	const m64h A(p);
	uint64_t q = p - 1; int v = 0;
	while ((q & 1) == 0) { q >>= 1; ++v; }
	const uint64_t w = two_sylow_gen<m64h, uint64_t>(A, p, q);
	if (w != gen_scan(A, p, q)) { if (bad++ < 10) std::printf("gen p=%llu\n", (unsigned long long)p); }
	static const int ds[] = {1, 2, 5, 11, 12, 20};
	for (int d : ds)
	{
		for (int t = 0; t < ny; ++t)
		{
			uint64_t y;
			if (t == 0) y = 5 % p;
			else if (t == 1) y = 1;
			else if (t == 2) y = p - 1;
			else if (t == 3) { y = rng() % p; for (int i = 0; i < d; ++i) y = A.mul(y, y); }   // a 2^d-th power
			else y = rng() % p;
			if (y == 0) continue;
			++checked;
			const uint64_t x = root_2d<m64h, uint64_t>(A, p, q, v, w, y, d);
			const uint64_t want = root_direct(A, p, q, v, w, y, d);
			bool ok = (x == want);
			if (x != 0)
			{
				uint64_t s = x;
				for (int i = 0; i < d; ++i) s = A.mul(s, s);
				ok = ok && (s == y);
				++roots;
			}
			else
			{
				const int e = (v < d) ? v : d;
				ok = ok && (A.pow(y, (p - 1) >> e) != 1);
			}
			if (!ok && bad++ < 10)
				std::printf("root p=%llu d=%d y=%llu got=%llu want=%llu\n", (unsigned long long)p, d,
				            (unsigned long long)y, (unsigned long long)x, (unsigned long long)want);
			// the other side y1 = -1/y, from z = y^q alone
			const uint64_t y1 = A.sub(0, A.pow(y, p - 2));
			const uint64_t z0 = A.pow(y, q);
			const uint64_t z1 = A.sub(0, inv_2sylow<m64h, uint64_t>(A, z0, v));
			const int e = (v < d) ? v : d;
			bool ok1 = (z1 == A.pow(y1, q));
			const bool pw1 = is_pow2d<m64h, uint64_t>(A, z1, v, d);
			ok1 = ok1 && (pw1 == (A.pow(y1, (p - 1) >> e) == 1));
			ok1 = ok1 && (is_pow2d<m64h, uint64_t>(A, z0, v, d) == (A.pow(y, (p - 1) >> e) == 1));
			if (pw1)
			{
				uint64_t alpha;
				int64_t c;
				inv_pow2_mod(q, d, alpha, c);
				ok1 = ok1 && (root_2d_finish<m64h, uint64_t>(A, v, w, A.pow(y1, alpha), z1, c, d) ==
				              root_direct(A, p, q, v, w, y1, d));
			}
			if (!ok1 && bad++ < 10) std::printf("side p=%llu d=%d y=%llu\n", (unsigned long long)p, d, (unsigned long long)y);
		}
	}
}

int main()
{
	std::mt19937_64 rng(20241);
	const uint32_t L = 1u << 18;
	std::vector<uint8_t> comp(L, 0);
	for (uint32_t i = 2; (uint64_t)i * i < L; ++i)
		if (!comp[i])
			for (uint32_t j = i * i; j < L; j += i) comp[j] = 1;
	for (uint32_t p = 11; p < L; p += 2)
		if (!comp[p] && (p % 10 == 1 || p % 10 == 9)) check_prime(p, rng, 5);
	for (uint64_t p = (1u << 12) + 1; p < (1u << 26); p += 1u << 12)
		if (is_prime(p)) check_prime(p, rng, 6);
	for (uint64_t p : {3ull, 5ull, 17ull, 257ull, 65537ull}) check_prime(p, rng, 6);
	for (int i = 0; i < 3000;)
	{
		const uint64_t p = (rng() >> 32) | 1;
		if (p > 61 && is_prime(p)) { check_prime(p, rng, 6); ++i; }
	}
	for (int i = 0; i < 300;)
	{
		const uint64_t p = (rng() >> 2) | 1;
		if (is_prime(p)) { check_prime(p, rng, 6); ++i; }
	}
	for (uint32_t p = 3; p < L; p += 2)
		if (!comp[p]) check_m32(p, rng);
	for (uint64_t p : {2147483647ull, 2147483659ull, 4294967291ull, 4294967279ull, 3221225473ull}) check_m32((uint32_t)p, rng);
	for (int i = 0; i < 20000;)
	{
		const uint64_t p = (rng() >> 32) | 1;
		if (p > 2 && is_prime(p)) { check_m32((uint32_t)p, rng); ++i; }
	}
	std::printf("%ld mismatches in %ld roots (%ld found), %ld m32 primes\n", bad, checked, roots, m32_checked);
	return bad ? 1 : 0;
}
