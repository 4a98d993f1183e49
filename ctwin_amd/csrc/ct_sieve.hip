// Window sieve for b^(2^m) - b^(2^(m-1)) +/- 1 on the GPU (the work of
// Sieve::check_pos / check_neg, sieve/CPU/ctsieve.cpp:678-787).
//
// A b in [b_min, b_max] is struck when some prime p of the side's form in the
// processed range divides N+(b) (p = 3k 2^m + 1) or N-(b) (p = 10k +/- 1):
//   + : b mod p is a primitive 3*2^m-th root of unity,
//   - : y = b^(2^(m-1)) mod p is a root of y^2 - y - 1, i.e. y = (1 +/- sqrt 5)/2.
// The reference filters p with a base-2 Euler test and verifies every strike;
// a composite p that survives it can only strike b that share a smaller prime
// factor with it, so the survivor set is the same as with exact primes.  Here
// primality is exact (+: Miller-Rabin to bases 2, 7, 61, deterministic below
// 4,759,123,141; -: a sieve of Eratosthenes) and the roots are computed
// directly:
//   + : c = a^k, k = (p-1)/(3 2^m), for the first a with c^(3 2^(m-1)) = -1
//       and c^(2^(m-1)) != -1 (then c has order 3 2^m); roots c g^t and c^5 g^t,
//       g = c^6, t < 2^(m-1).
//   - : x^(2^d) = y (d = m-1) solved once (y^alpha with alpha 2^d = 1 mod the
//       odd part of p-1, then a 2-group discrete logarithm), all 2^e roots by
//       a primitive 2^e-th root of unity (e = min(v2(p-1), d)).
// Pipeline per block of k: a candidate kernel (+: trial division and
// Miller-Rabin with wave-aggregated compaction; -: a segmented sieve of
// Eratosthenes over k), root kernels that write 1 into
// struck[b - b_min] (idempotent plain byte stores; no atomics needed).  Primes
// below 2^16 on the - side strike too many b per root for one thread: their
// (p, root) pairs are queued and struck by one workgroup each (+ primes below
// 2^20 too).  A - prime with
// more than 32 roots (2^e, e = min(v2(p-1), m-1)) queues its run of roots for a
// wavefront of its own, so that no lane walks up to 2^(m-1) roots while its
// wavefront waits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/ctwin.h"
#include "ct_engine_impl.h"
#include "ct_numth.h"

namespace ct {

// ---------------------------------------------------------------------------
// striking

__device__ __forceinline__ void strike_range(uint8_t *struck, uint64_t x, uint64_t p, uint64_t b_min, uint64_t b_max)
{
	uint64_t s = x;
	if (s < b_min) s += ((b_min - s + p - 1) / p) * p;
	for (; s <= b_max; s += p) struck[s - b_min] = 1;
}

__device__ __forceinline__ void strike_one(uint8_t *struck, uint32_t x, uint32_t p, uint64_t b_min, uint64_t b_max)
{
	if ((uint64_t)p > b_max) { if (x >= b_min && x <= b_max) struck[x - b_min] = 1; }
	else strike_range(struck, x, p, b_min, b_max);
}

__device__ __forceinline__ bool small_factor(uint32_t p)
{
	return (p % 3 == 0) || (p % 5 == 0) || (p % 7 == 0) || (p % 11 == 0) || (p % 13 == 0) || (p % 17 == 0)
	    || (p % 19 == 0) || (p % 23 == 0) || (p % 29 == 0) || (p % 31 == 0) || (p % 37 == 0) || (p % 41 == 0);
}

__device__ __forceinline__ bool is_prime32(uint32_t p)
{
	if (p < 2) return false;
	if (p < 43) return p == 2 || p == 3 || p == 5 || p == 7 || p == 11 || p == 13 || p == 17 || p == 19
	                || p == 23 || p == 29 || p == 31 || p == 37 || p == 41;
	if ((p & 1) == 0 || small_factor(p)) return false;
	return spsp_2_7_61(p);
}

// wave-aggregated append of p to list (one atomic per wave)
__device__ __forceinline__ void append(bool keep, uint32_t p, uint32_t *list, uint32_t *count)
{
	const uint64_t mask = __ballot(keep);
	if (mask == 0) return;
	const int lane = __lane_id();
	const int leader = __ffsll((long long)mask) - 1;
	uint32_t base = 0;
	if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
	base = __shfl(base, leader);
	if (keep) list[base + __popcll(mask & ((1ull << lane) - 1))] = p;
}

// ---------------------------------------------------------------------------
// + side

__global__ void k_primes_pos(uint64_t k0, uint64_t kcount, int m, uint32_t *list, uint32_t *count)
{
	const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	bool keep = false;
	uint32_t p = 0;
	if (i < kcount)
	{
		const uint64_t pp = 3 * ((k0 + i) << m) + 1;
		p = (uint32_t)pp;
		keep = (pp < 0x100000000ull) && is_prime32(p);
	}
	append(keep, p, list, count);
}

// Primes below these strike too many b per root for one thread (about
// 2^24 / p in a 2^24 window): their (p, root) pairs are queued for
// k_strike_small, one workgroup per pair.  On the + side every prime has 2^m
// roots, and one workgroup walking them for p = 3 2^12 + 1 would keep its CU
// busy for milliseconds (each byte store touches 64 cache lines).
constexpr uint32_t kSmallP = 1u << 16;
constexpr uint32_t kSmallPosP = 1u << 20;

// one workgroup per prime: find c, then 256 threads walk g = c^6
__global__ void k_roots_pos(const uint32_t *list, const uint32_t *count, int m, uint64_t b_min, uint64_t b_max, uint8_t *struck,
                            uint2 *small, uint32_t *small_count, uint32_t small_cap)
{
	const uint32_t idx = blockIdx.x;
	if (idx >= *count) return;
	const uint32_t p = list[idx];
	const m32 A(p);
	const uint64_t k = (uint64_t)(p - 1) / (3ull << m);
	__shared__ uint32_t sc;
	if (threadIdx.x == 0)
	{
		uint32_t c = 0;
		for (uint32_t a = 2;; ++a)
		{
			c = A.pow(A.to_m(a), k);
			uint32_t c2 = c;
			for (int i = 0; i < m - 1; ++i) c2 = A.mul(c2, c2);           // c^(2^(m-1))
			const uint32_t c6 = A.mul(A.mul(c2, c2), c2);                  // c^(3 2^(m-1))
			if (c6 == A.neg_one() && c2 != A.neg_one()) break;
		}
		sc = c;
	}
	__syncthreads();
	const uint32_t c = sc;
	const uint32_t c2 = A.mul(c, c), c4 = A.mul(c2, c2), c5 = A.mul(c4, c), g = A.mul(c4, c2);
	const uint32_t half = 1u << (m - 1);                 // t < 2^(m-1)
	const uint32_t nt = blockDim.x;
	const uint32_t per = (half + nt - 1) / nt;
	const uint32_t t0 = threadIdx.x * per;
	if (t0 >= half) return;
	const uint32_t t1 = min(half, t0 + per);
	uint32_t cur = A.pow(g, t0);
	for (uint32_t t = t0; t < t1; ++t)
	{
		const uint32_t b1 = A.from_m(A.mul(c, cur)), b5 = A.from_m(A.mul(c5, cur));
		if (p < kSmallPosP)
		{
			const uint32_t slot = atomicAdd(small_count, 2u);
			if (slot + 1 < small_cap) { small[slot] = make_uint2(p, b1); small[slot + 1] = make_uint2(p, b5); }
		}
		else
		{
			strike_one(struck, b1, p, b_min, b_max);
			strike_one(struck, b5, p, b_min, b_max);
		}
		cur = A.mul(cur, g);
	}
}

// ---------------------------------------------------------------------------
// - side

// Primes p = 10k +- 1 (5 a square mod p) by a sieve of Eratosthenes over k,
// one workgroup per segment of kSegK values of k, with a bitmap per side in
// LDS (bit k: p = 10k - 1 for side 0, 10k + 1 for side 1).  The sieving
// primes are every prime 3 <= q < 2^16 but 5 (q = 2 and q = 5 divide no
// 10k +- 1; every composite p < 2^32 has a prime factor below 2^16).  q
// divides 10k - 1 exactly for k = r0 mod q, r0 = 10^-1 mod q, and 10k + 1 for
// k = r1 = q - r0, so each q strikes one residue class of k per side; the
// strike of p = q itself is skipped.  Sieving primes below kCoopQ stride their
// many multiples with the whole workgroup; the others take one lane each.  The
// unstruck k become primes, gathered in an LDS list and appended to the global
// list with one atomic per workgroup (a list that outgrows LDS, possible only
// in the first segments where primes are dense, appends its excess directly).
// The list order is immaterial: every prime strikes its own b.
struct sieve_prime { uint32_t q, r; };    // r = r0 | r1 << 16

constexpr uint32_t kSegK = 1u << 15;      // k values per workgroup
constexpr uint32_t kSegCap = 12288;       // LDS prime list (2 kSegK candidates hold about 7,600 primes near 2^32)
constexpr uint32_t kCoopQ = 256;
constexpr uint32_t kCoopMax = 64;         // sieving primes below kCoopQ: 52

__device__ __forceinline__ void strike_k(uint32_t *bm, uint32_t k) { atomicOr(&bm[k >> 5], 1u << (k & 31)); }

__global__ __launch_bounds__(256) void k_sieve_neg(uint64_t k0, uint64_t kcount, const sieve_prime *sp, uint32_t nsp,
                                                  uint32_t ncoop, uint32_t *list, uint32_t *count)
{
	__shared__ uint32_t bm[2][kSegK / 32];
	__shared__ uint32_t off[2][kCoopMax];
	__shared__ uint32_t plist[kSegCap];
	__shared__ uint32_t npl, gbase;
	const uint32_t t = threadIdx.x, nt = blockDim.x;
	const uint64_t kseg = k0 + (uint64_t)blockIdx.x * kSegK;       // below 2^29: p < 2^32
	const uint64_t kend = min(k0 + kcount, kseg + kSegK);
	const uint32_t ks = (uint32_t)kseg;
	for (uint32_t i = t; i < 2 * (kSegK / 32); i += nt) (&bm[0][0])[i] = 0;
	if (t == 0) npl = 0;
	for (uint32_t i = t; i < ncoop; i += nt)
	{
		const uint32_t q = sp[i].q, r = sp[i].r, kr = ks % q;
		for (int s = 0; s < 2; ++s)
		{
			uint32_t o = (s ? (r >> 16) : (r & 0xffffu)) + q - kr;
			off[s][i] = (o >= q) ? o - q : o;
		}
	}
	__syncthreads();
	for (uint32_t i = 0; i < ncoop; ++i)
	{
		const uint32_t q = sp[i].q;
		for (int s = 0; s < 2; ++s)
		{
			const uint32_t o = off[s][i];
			for (uint32_t k = o + t * q; k < kSegK; k += nt * q)
			{
				if (k == o && 10 * (kseg + k) + 2 * s - 1 == q) continue;     // p = q
				strike_k(bm[s], k);
			}
		}
	}
	for (uint32_t i = ncoop + t; i < nsp; i += nt)
	{
		const uint32_t q = sp[i].q, r = sp[i].r, kr = ks % q;
		for (int s = 0; s < 2; ++s)
		{
			uint32_t o = (s ? (r >> 16) : (r & 0xffffu)) + q - kr;
			if (o >= q) o -= q;
			uint32_t k = o;
			if (k < kSegK && 10 * (kseg + k) + 2 * s - 1 == q) k += q;      // p = q
			for (; k < kSegK; k += q) strike_k(bm[s], k);
		}
	}
	__syncthreads();
	const int lane = __lane_id();
	for (uint32_t wi = t; wi < 2 * (kSegK / 32); wi += nt)        // uniform trip count: 2 kSegK / 32 is a multiple of nt
	{
		const int s = (int)(wi / (kSegK / 32));
		const uint32_t w = wi % (kSegK / 32);
		uint32_t u = ~bm[s][w];
		while (__ballot(u != 0))
		{
			bool keep = false;
			uint32_t p = 0;
			if (u)
			{
				const uint32_t b = __ffs(u) - 1;
				u &= u - 1;
				const uint64_t k = kseg + w * 32 + b;
				const uint64_t pp = 10 * k + 2 * s - 1;
				keep = k < kend && pp >= 11 && pp < 0x100000000ull;
				p = (uint32_t)pp;
			}
			const uint64_t mask = __ballot(keep);
			if (mask == 0) continue;
			const int leader = __ffsll((long long)mask) - 1;
			uint32_t base = 0;
			if (lane == leader) base = atomicAdd(&npl, (uint32_t)__popcll(mask));
			base = __shfl(base, leader);
			if (keep)
			{
				const uint32_t slot = base + __popcll(mask & ((1ull << lane) - 1));
				if (slot < kSegCap) plist[slot] = p;
				else list[atomicAdd(count, 1u)] = p;
			}
		}
	}
	__syncthreads();
	const uint32_t n = min(npl, kSegCap);
	if (t == 0) gbase = atomicAdd(count, n);
	__syncthreads();
	for (uint32_t j = t; j < n; j += nt) list[gbase + j] = plist[j];
}

// The sieving primes, built once on the host (a function-local static:
// initialised once even when several host threads call the sieve).
static std::vector<sieve_prime> build_sieve_primes_neg()
{
	std::vector<sieve_prime> v;
	const uint32_t L = 1u << 16;
	std::vector<uint8_t> comp(L, 0);
	for (uint32_t i = 2; i * i < L; ++i)
		if (!comp[i])
			for (uint32_t j = i * i; j < L; j += i) comp[j] = 1;
	for (uint32_t q = 3; q < L; ++q)
	{
		if (comp[q] || q == 5) continue;
		uint32_t r0 = 0;
		for (uint32_t m = 0; m < 10; ++m)
			if ((q * m + 1) % 10 == 0) { r0 = (q * m + 1) / 10; break; }   // 10 r0 = 1 mod q
		v.push_back(sieve_prime{q, r0 | ((q - r0) << 16)});
	}
	return v;
}

static const std::vector<sieve_prime> &sieve_primes_neg(uint32_t *ncoop)
{
	static const std::vector<sieve_prime> v = build_sieve_primes_neg();
	uint32_t nc = 0;
	while (nc < v.size() && v[nc].q < kCoopQ) ++nc;       // ascending q
	*ncoop = nc;
	return v;
}


// A run of roots x u^j (j < 2^e) of one prime, in Montgomery form.
struct root_run { uint32_t p, x, u, e; };

constexpr int kRunInline = 5;              // runs of up to 32 roots stay in their lane (8: 0.6 ms slower in all)

// Strikes (or queues, for p below kSmallP) the cnt roots r, r u, r u^2, ...
__device__ __forceinline__ void strike_roots(const m32 &A, uint32_t p, uint32_t r, uint32_t u, uint32_t cnt, uint8_t *struck,
                                             uint64_t b_min, uint64_t b_max, uint2 *small, uint32_t *small_count,
                                             uint32_t small_cap)
{
	for (uint32_t j = 0; j < cnt; ++j)
	{
		const uint32_t b = A.from_m(r);
		if (p < kSmallP)
		{
			const uint32_t slot = atomicAdd(small_count, 1u);
			if (slot < small_cap) small[slot] = make_uint2(p, b);
		}
		else strike_one(struck, b, p, b_min, b_max);
		r = A.mul(r, u);
	}
}

// One prime: sqrt(5), y = (1 +/- sqrt 5)/2, all 2^(m-1)-th roots of y.
__device__ __forceinline__ void roots_neg_one(uint32_t p, int m, uint64_t b_min, uint64_t b_max, uint8_t *struck,
                                              uint2 *small, uint32_t *small_count, uint32_t small_cap,
                                              root_run *big, uint32_t *big_count, uint32_t big_cap)
{
	const m32 A(p);
	uint32_t q = p - 1; int v = 0;
	while ((q & 1) == 0) { q >>= 1; ++v; }
	// The exponentiations run two or three chains per loop: the kernel is
	// bound by the latency of dependent products, not by their number.
	// w = c^q and t = 5^((q-1)/2) together; sqrt 5 = 5^alpha (2^d-th root for
	// d = 1, alpha = (q + 1)/2, c = 1) from x1 = 5 t, z = t x1 = 5^q.
	const uint32_t five = A.to_m(5u);
	uint32_t w, t5;
	pow_dual<m32, uint32_t, uint32_t>(A, A.to_m(least_nonresidue(A, p)), q, five, (q - 1) / 2, w, t5);
	if (q == 1) return;                                    // p = 2^v + 1: not of the form 10k +- 1
	const uint32_t x15 = A.mul(t5, five), z5 = A.mul(t5, x15);
	if (!is_pow2d<m32, uint32_t>(A, z5, v, 1)) return;
	const uint32_t s5 = root_2d_finish<m32, uint32_t>(A, v, w, x15, z5, 1, 1);
	const uint32_t inv2 = A.to_m((p + 1) / 2);
	const int d = m - 1, e = (v < d) ? v : d;
	uint32_t u = w;                                        // primitive 2^e-th root: w^(2^(v-e))
	for (int i = 0; i < v - e; ++i) u = A.mul(u, u);
	// y0 = (1 - sqrt 5)/2, y1 = (1 + sqrt 5)/2 = -1/y0 (y0 y1 = -1), so
	// y1^q = -(y0^q)^-1 (q odd).  For v <= d at most one of them is a 2^d-th
	// power (z0 = 1 or z0 = -1), for v > d both or neither.  z0 = y0^q,
	// y0^alpha and y1^alpha come from one loop.
	const uint32_t y0 = A.mul(A.sub(A.one, s5), inv2), y1 = A.mul(A.add(A.one, s5), inv2);
	uint64_t alpha;
	int64_t c;
	inv_pow2_mod((uint64_t)q, d, alpha, c);
	uint32_t z0, xa0, xa1;
	pow_tri<m32, uint32_t, uint32_t>(A, y0, y1, q, (uint32_t)alpha, z0, xa0, xa1);      // alpha < q
	const uint32_t z1 = A.sub(0u, inv_2sylow<m32, uint32_t>(A, z0, v));
	const bool ok0 = is_pow2d<m32, uint32_t>(A, z0, v, d), ok1 = is_pow2d<m32, uint32_t>(A, z1, v, d);
	for (int round = 0; round < 2; ++round)
	{
		const bool act = round ? (ok0 && ok1) : (ok0 || ok1);
		if (!act) continue;
		const bool take0 = round == 0 && ok0;
		const uint32_t x = root_2d_finish<m32, uint32_t>(A, v, w, take0 ? xa0 : xa1, take0 ? z0 : z1, c, d);
		if (e > kRunInline)
		{   // many roots: a wavefront of k_roots_run shares them, unless the queue is full
			const uint32_t slot = atomicAdd(big_count, 1u);
			if (slot < big_cap) { big[slot] = root_run{p, x, u, (uint32_t)e}; continue; }
		}
		strike_roots(A, p, x, u, 1u << e, struck, b_min, b_max, small, small_count, small_cap);
	}
}

// One thread per prime of list[0 .. *count), in a grid-stride loop: the grid
// is sized from the list's capacity, so the host launches it without reading
// the count back.
__global__ void k_roots_neg(const uint32_t *list, const uint32_t *count, int m, uint64_t b_min, uint64_t b_max,
                            uint8_t *struck, uint2 *small, uint32_t *small_count, uint32_t small_cap,
                            root_run *big, uint32_t *big_count, uint32_t big_cap)
{
	const uint32_t n = *count;
	for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x)
		roots_neg_one(list[idx], m, b_min, b_max, struck, small, small_count, small_cap, big, big_count, big_cap);
}

// The queued runs of 2^e > 2^kRunInline roots, one wavefront each: lane l
// takes the roots x u^j for j in [l 2^e / 64, (l + 1) 2^e / 64), or the single
// root j = l of a run shorter than 64.  A grid-stride loop covers every queued
// run.
__global__ __launch_bounds__(64) void k_roots_run(const root_run *big, const uint32_t *big_count, uint32_t big_cap,
                                                  uint64_t b_min, uint64_t b_max, uint8_t *struck, uint2 *small,
                                                  uint32_t *small_count, uint32_t small_cap)
{
	const uint32_t n = min(*big_count, big_cap);
	for (uint32_t i = blockIdx.x; i < n; i += gridDim.x)
	{
		const root_run R = big[i];
		const m32 A(R.p);
		const uint32_t cnt = 1u << R.e;
		const uint32_t per = (cnt >= 64) ? cnt / 64 : (threadIdx.x < cnt ? 1u : 0u);
		const uint32_t j0 = (cnt >= 64) ? threadIdx.x * per : threadIdx.x;
		const uint32_t r = A.mul(R.x, A.pow(R.u, j0));
		strike_roots(A, R.p, r, R.u, per, struck, b_min, b_max, small, small_count, small_cap);
	}
}

// one workgroup per (p, root) of a small prime: threads stride the window
__global__ void k_strike_small(const uint2 *small, const uint32_t *small_count, uint64_t b_min, uint64_t b_max, uint8_t *struck)
{
	if (blockIdx.x >= *small_count) return;
	const uint2 e = small[blockIdx.x];
	const uint64_t p = e.x;
	uint64_t s = e.y;
	if (s < b_min) s += ((b_min - s + p - 1) / p) * p;
	for (uint64_t t = s + threadIdx.x * p; t <= b_max; t += (uint64_t)blockDim.x * p) struck[t - b_min] = 1;
}

}   // namespace ct

using namespace ct;

#define HIPCHK(call) do { hipError_t _e = (call); if (_e != hipSuccess) { rc = hip_fail(_e, #call); goto done; } } while (0)

namespace {

constexpr uint64_t kWinBlockK = 1ull << 24;            // k values per block of the candidate kernels
constexpr uint32_t kSmallCap = 1u << 22;               // queued (p, root) pairs of small primes, all blocks
constexpr uint32_t kBigCap = 1u << 16;                 // queued runs of roots per block
constexpr uint32_t kRootsNegGrid = 1u << 16;           // workgroups of k_roots_neg (x 128 threads >= primes per block)
constexpr int kTimedBlocks = 32;                       // k_roots_neg launches timed per call (p < 2^32: at most 26)

// Device buffers, stream and events of the window sieve, kept per device for
// the life of the process (about 190 MiB plus the byte map, which grows to the
// largest window seen).  `mu` serialises calls on one device.
struct window_ws
{
	std::mutex mu;
	bool ready = false;
	hipStream_t st = nullptr;
	uint8_t *d_struck = nullptr;
	size_t struck_cap = 0;
	uint32_t *d_list = nullptr, *d_count = nullptr, *d_scount = nullptr;
	uint2 *d_small = nullptr;
	root_run *d_big = nullptr;
	sieve_prime *d_sp = nullptr;
	uint64_t *d_list64 = nullptr;
	uint32_t *d_count64 = nullptr;
	hipEvent_t ev[2 * kTimedBlocks + 4] = {};
	ct_sieve_timing last = {};
};

window_ws g_win[64];

int window_ws_init(window_ws &w, int device)
{
	if (w.ready) return CT_OK;
	hipError_t h;
	uint32_t ncoop = 0;
	const std::vector<sieve_prime> &sp = sieve_primes_neg(&ncoop);
	if ((h = hipSetDevice(device)) != hipSuccess) return hip_fail(h, "hipSetDevice");
	if ((h = hipStreamCreate(&w.st)) != hipSuccess) return hip_fail(h, "hipStreamCreate");
	for (hipEvent_t &e : w.ev)
		if ((h = hipEventCreate(&e)) != hipSuccess) return hip_fail(h, "hipEventCreate");
	if (hipMalloc((void **)&w.d_list, 2 * kWinBlockK * sizeof(uint32_t)) != hipSuccess ||
	    hipMalloc((void **)&w.d_count, sizeof(uint32_t)) != hipSuccess ||
	    hipMalloc((void **)&w.d_scount, 2 * sizeof(uint32_t)) != hipSuccess ||     // small pairs, queued runs
	    hipMalloc((void **)&w.d_small, kSmallCap * sizeof(uint2)) != hipSuccess ||
	    hipMalloc((void **)&w.d_big, kBigCap * sizeof(root_run)) != hipSuccess ||
	    hipMalloc((void **)&w.d_sp, sp.size() * sizeof(sieve_prime)) != hipSuccess ||
	    hipMalloc((void **)&w.d_list64, kWindow64List * sizeof(uint64_t)) != hipSuccess ||
	    hipMalloc((void **)&w.d_count64, sizeof(uint32_t)) != hipSuccess)
		return fail(CT_ERR_NOMEM, "hipMalloc window sieve buffers");
	if ((h = hipMemcpy(w.d_sp, sp.data(), sp.size() * sizeof(sieve_prime), hipMemcpyHostToDevice)) != hipSuccess)
		return hip_fail(h, "sieving primes");
	w.ready = true;
	return CT_OK;
}

}   // namespace

int ct_sieve_window(int device, int m, uint64_t b_min, uint64_t b_max, uint64_t p_min_pos, uint64_t p_max_pos,
                    uint64_t p_min_neg, uint64_t p_max_neg, int mode, uint8_t *struck)
{
	if (!struck) return fail(CT_ERR_ARG, "struck is null");
	if (device < 0 || device >= 64) return fail(CT_ERR_ARG, "device ordinal out of range");
	if (m < 8 || m > 24) return fail(CT_ERR_ARG, "m must be in [8, 24] (ctsieve.cpp:841)");
	if (b_max < b_min || b_max - b_min >= (1ull << 32)) return fail(CT_ERR_ARG, "bad b window");
	const uint64_t P32 = 1ull << 32;
	// k ranges with the reference's bound rule (ctsieve.cpp:680-684, 730-734)
	const uint64_t kp_min = (p_min_pos >> m) / 3;
	uint64_t kp_max = (p_max_pos >> m) / 3;
	while (3 * (kp_max << m) + 1 < p_max_pos) ++kp_max;
	const uint64_t kn_min = p_min_neg / 10;
	uint64_t kn_max = p_max_neg / 10;
	while (10 * kn_max + 1 < p_max_neg) ++kn_max;
	if ((mode >= 0 && p_max_pos >= (1ull << 62)) || (mode <= 0 && p_max_neg >= (1ull << 62)))
		return fail(CT_ERR_UNSUPPORTED, "p bounds must lie below 2^62");
	// primes from 2^32 up run through the 64-bit kernels: the k from which
	// every p of the side is >= 2^32
	const uint64_t kp64 = (P32 - 1 + (3ull << m) - 1) / (3ull << m);     // 3 k 2^m + 1 >= 2^32
	const uint64_t kn64 = (P32 + 10) / 10;                                // 10 k - 1 >= 2^32

	const size_t W = (size_t)(b_max - b_min + 1);
	uint32_t ncoop = 0;
	const std::vector<sieve_prime> &sp = sieve_primes_neg(&ncoop);
	if (ncoop > kCoopMax) return fail(CT_ERR_UNSUPPORTED, "sieving prime table larger than the kernel holds");
	window_ws &w = g_win[device];
	std::lock_guard<std::mutex> lock(w.mu);
	int rc = window_ws_init(w, device);
	if (rc) return rc;
	hipStream_t st = w.st;
	int timed = 0;
	w.last = ct_sieve_timing{};
	{
		HIPCHK(hipSetDevice(device));
		if (W > w.struck_cap)
		{
			HIPCHK(hipFree(w.d_struck));
			w.d_struck = nullptr; w.struck_cap = 0;
			HIPCHK(hipMalloc((void **)&w.d_struck, W));
			w.struck_cap = W;
		}
		HIPCHK(hipEventRecord(w.ev[2 * kTimedBlocks], st));
		HIPCHK(hipMemcpyAsync(w.d_struck, struck, W, hipMemcpyHostToDevice, st));
		HIPCHK(hipEventRecord(w.ev[2 * kTimedBlocks + 2], st));
		HIPCHK(hipMemsetAsync(w.d_scount, 0, 2 * sizeof(uint32_t), st));
		// + side: one block of k covers every p < 2^32 for m >= 9 (two for m = 8)
		if (mode >= 0 && kp_max >= kp_min)
		{
			for (uint64_t k0 = kp_min; k0 <= kp_max; k0 += kWinBlockK)
			{
				const uint64_t kc = std::min(kWinBlockK, kp_max - k0 + 1);
				if (3 * (k0 << m) + 1 >= P32) break;
				HIPCHK(hipMemsetAsync(w.d_count, 0, sizeof(uint32_t), st));
				hipLaunchKernelGGL(k_primes_pos, dim3((uint32_t)((kc + 255) / 256)), dim3(256), 0, st, k0, kc, m, w.d_list, w.d_count);
				HIPCHK(hipGetLastError());
				uint32_t cnt = 0;
				HIPCHK(hipMemcpyAsync(&cnt, w.d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
				HIPCHK(hipStreamSynchronize(st));
				if (cnt)
				{
					hipLaunchKernelGGL(k_roots_pos, dim3(cnt), dim3(256), 0, st, w.d_list, w.d_count, m, b_min, b_max, w.d_struck,
					                   w.d_small, w.d_scount, kSmallCap);
					HIPCHK(hipGetLastError());
				}
			}
		}
		// - side: the blocks run back to back with no host round trip; the
		// list, its count and the run queue are reused block after block in
		// stream order
		if (mode <= 0 && kn_max >= kn_min)
		{
			for (uint64_t k0 = kn_min; k0 <= kn_max; k0 += kWinBlockK)
			{
				const uint64_t kc = std::min(kWinBlockK, kn_max - k0 + 1);
				if (10 * k0 >= P32 + 1) break;
				HIPCHK(hipMemsetAsync(w.d_count, 0, sizeof(uint32_t), st));
				HIPCHK(hipMemsetAsync(w.d_scount + 1, 0, sizeof(uint32_t), st));
				hipLaunchKernelGGL(k_sieve_neg, dim3((uint32_t)((kc + kSegK - 1) / kSegK)), dim3(256), 0, st, k0, kc, w.d_sp,
				                   (uint32_t)sp.size(), ncoop, w.d_list, w.d_count);
				HIPCHK(hipGetLastError());
				const uint32_t grid = (uint32_t)std::min<uint64_t>(kRootsNegGrid, (2 * kc + 127) / 128);
				if (timed < kTimedBlocks) HIPCHK(hipEventRecord(w.ev[2 * timed], st));
				hipLaunchKernelGGL(k_roots_neg, dim3(grid), dim3(128), 0, st, w.d_list, w.d_count, m, b_min, b_max,
				                   w.d_struck, w.d_small, w.d_scount, kSmallCap, w.d_big, w.d_scount + 1, kBigCap);
				HIPCHK(hipGetLastError());
				if (timed < kTimedBlocks) { HIPCHK(hipEventRecord(w.ev[2 * timed + 1], st)); ++timed; }
				hipLaunchKernelGGL(k_roots_run, dim3(1024), dim3(64), 0, st, w.d_big, w.d_scount + 1, kBigCap, b_min, b_max,
				                   w.d_struck, w.d_small, w.d_scount, kSmallCap);
				HIPCHK(hipGetLastError());
			}
		}
		// the queued (p, root) pairs of small primes, both sides, in one launch
		{
			uint32_t scnt = 0;
			HIPCHK(hipMemcpyAsync(&scnt, w.d_scount, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
			HIPCHK(hipStreamSynchronize(st));
			if (scnt > kSmallCap) { rc = fail(CT_ERR_UNSUPPORTED, "too many small-prime roots"); goto done; }
			if (scnt)
			{
				hipLaunchKernelGGL(k_strike_small, dim3(scnt), dim3(256), 0, st, w.d_small, w.d_scount, b_min, b_max, w.d_struck);
				HIPCHK(hipGetLastError());
			}
		}
		// p >= 2^32 (up to the first p >= p_max, the reference's rule): the
		// 64-bit prime and root kernels strike the same byte map
		if (mode >= 0 && kp_max >= kp_min && kp_max >= kp64)
		{
			rc = strike_window64(st, m, +1, std::max(kp_min, kp64), kp_max, P32, 3 * (kp_max << m) + 1, b_min, b_max, w.d_struck,
			                     w.d_list64, w.d_count64);
			if (rc) goto done;
		}
		if (mode <= 0 && kn_max >= kn_min && kn_max >= kn64)
		{
			rc = strike_window64(st, m, -1, std::max(kn_min, kn64), kn_max, P32, 10 * kn_max + 1, b_min, b_max, w.d_struck,
			                     w.d_list64, w.d_count64);
			if (rc) goto done;
		}
		HIPCHK(hipEventRecord(w.ev[2 * kTimedBlocks + 3], st));
		HIPCHK(hipMemcpyAsync(struck, w.d_struck, W, hipMemcpyDeviceToHost, st));
		HIPCHK(hipEventRecord(w.ev[2 * kTimedBlocks + 1], st));
		HIPCHK(hipStreamSynchronize(st));
		for (int i = 0; i < timed; ++i)
		{
			float ms = 0;
			HIPCHK(hipEventElapsedTime(&ms, w.ev[2 * i], w.ev[2 * i + 1]));
			w.last.roots_neg_ms += ms;
		}
		w.last.roots_neg_launches = (uint32_t)timed;
		HIPCHK(hipEventElapsedTime(&w.last.device_ms, w.ev[2 * kTimedBlocks], w.ev[2 * kTimedBlocks + 1]));
		HIPCHK(hipEventElapsedTime(&w.last.kernels_ms, w.ev[2 * kTimedBlocks + 2], w.ev[2 * kTimedBlocks + 3]));
	}
done:
	if (rc) (void)hipStreamSynchronize(st);       // nothing of this call still runs when it returns
	return rc;
}

int ct_sieve_last_timing(int device, ct_sieve_timing *out)
{
	if (!out) return fail(CT_ERR_ARG, "out is null");
	if (device < 0 || device >= 64) return fail(CT_ERR_ARG, "device ordinal out of range");
	window_ws &w = g_win[device];
	std::lock_guard<std::mutex> lock(w.mu);
	*out = w.last;
	return CT_OK;
}
