// GPU factor emission (the path of sieve/GPU: kernel.cl:243-479 sequenced by
// sieve/GPU/src/engine.h:188-220 and gfsieve::check, ctsieve.h:326-431):
// every pair (p, b) with p a prime of the side's form in [p_min, p_max],
// 0 <= b < p a root of N+(b) = b^(2^m) - b^(2^(m-1)) + 1 (p = 3k 2^m + 1) or of
// N-(b) (p = 10k +/- 1) mod p, and b <= bMax(p) (5e9, or 1e10 above 1e16 for
// + and above 1e13 for -, kernel.cl:304 and :434).  Primality is exact
// (base-2 strong test, then Miller-Rabin to the first twelve prime bases,
// deterministic below 3.3e24); the roots come from the same constructions as
// the window sieve, with 64-bit Montgomery arithmetic.  The host re-checks
// every pair as saveFactors does (ctsieve.h:270-289).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <fstream>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ctwin.h"
#include "ct_engine_impl.h"
#include "ct_numth.h"

namespace ct {

__device__ __forceinline__ bool is_prime64(uint64_t p)
{
	if (p < 2) return false;
	const uint32_t small[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
	for (int i = 0; i < 12; ++i) { if (p == small[i]) return true; if (p % small[i] == 0) return false; }
	if (p < 1369) return true;
	const m64d A(p);
	if (!mr_prime<m64d, uint64_t>(A, p, small, 1)) return false;        // base 2 first: most composites stop here
	return mr_prime<m64d, uint64_t>(A, p, small + 1, 11);
}

__device__ __forceinline__ void fappend(bool keep, uint64_t p, uint64_t *list, uint32_t *count)
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

__device__ __forceinline__ void emit(uint64_t p, uint64_t b, uint64_t bmax, uint64_t *out, unsigned long long *ocount, uint64_t cap)
{
	if (b > bmax) return;
	const unsigned long long slot = atomicAdd(ocount, 1ull);
	if (slot < cap) { out[2 * slot] = p; out[2 * slot + 1] = b; }
}

// Where a root b of N+-(b) mod p goes: into the pair list (factor emission,
// b <= bMax(p) of kernel.cl:304 and :434, or b <= bmax_set when it is set), or
// onto a window of b (the window sieve for p of any size: struck[b' - b_min]
// for every b' = b mod p in [b_min, b_max]).
struct emit_sink
{
	uint64_t *out; unsigned long long *ocount; uint64_t cap; uint64_t bmax_set;
	__device__ __forceinline__ void operator()(uint64_t p, uint64_t b, uint64_t bmax) const
	{
		emit(p, b, bmax_set ? bmax_set : bmax, out, ocount, cap);
	}
};

struct strike_sink
{
	uint8_t *struck; uint64_t b_min, b_max;
	__device__ __forceinline__ void operator()(uint64_t p, uint64_t b, uint64_t) const
	{
		uint64_t s = b;
		if (s < b_min) s += ((b_min - s + p - 1) / p) * p;
		while (s <= b_max)
		{
			struck[s - b_min] = 1;
			if (s > ~0ull - p) break;                  // s + p would wrap
			s += p;
		}
	}
};

__global__ void k_fprimes(uint64_t k0, uint64_t kcount, int m, int mode, uint64_t p_min, uint64_t p_max,
                          uint64_t *list, uint32_t *count)
{
	const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	for (int s = 0; s < (mode > 0 ? 1 : 2); ++s)
	{
		bool keep = false;
		uint64_t p = 0;
		if (i < kcount)
		{
			const uint64_t k = k0 + i;
			p = (mode > 0) ? 3 * (k << m) + 1 : 10 * k + 2 * (uint64_t)s - 1;
			keep = (k > 0 || mode > 0) && p >= p_min && p <= p_max && p >= 11 && is_prime64(p);
		}
		fappend(keep, p, list, count);
	}
}

// + side: one workgroup per prime
template <typename Sink>
__global__ void k_froots_pos(const uint64_t *list, const uint32_t *count, int m, Sink sink)
{
	if (blockIdx.x >= *count) return;
	const uint64_t p = list[blockIdx.x];
	const m64d A(p);
	const uint64_t k = (p - 1) / (3ull << m);
	__shared__ uint64_t sc;
	if (threadIdx.x == 0)
	{
		uint64_t c = 0;
		for (uint64_t a = 2;; ++a)
		{
			c = A.pow(A.to_m(a), k);
			uint64_t c2 = c;
			for (int i = 0; i < m - 1; ++i) c2 = A.mul(c2, c2);
			if (A.mul(A.mul(c2, c2), c2) == A.neg_one() && c2 != A.neg_one()) break;
		}
		sc = c;
	}
	__syncthreads();
	const uint64_t c = sc;
	const uint64_t c2 = A.mul(c, c), c4 = A.mul(c2, c2), c5 = A.mul(c4, c), g = A.mul(c4, c2);
	const uint64_t bmax = (p < 10000000000000000ull) ? 5000000000ull : 10000000000ull;
	const uint32_t half = 1u << (m - 1), per = (half + blockDim.x - 1) / blockDim.x;
	const uint32_t t0 = threadIdx.x * per;
	if (t0 >= half) return;
	const uint32_t t1 = min(half, t0 + per);
	uint64_t cur = A.pow(g, t0);
	for (uint32_t t = t0; t < t1; ++t)
	{
		sink(p, A.from_m(A.mul(c, cur)), bmax);
		sink(p, A.from_m(A.mul(c5, cur)), bmax);
		cur = A.mul(cur, g);
	}
}

// - side: one thread per prime
template <typename Sink>
__global__ void k_froots_neg(const uint64_t *list, const uint32_t *count, int m, Sink sink)
{
	const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= *count) return;
	const uint64_t p = list[idx];
	const m64d A(p);
	uint64_t q = p - 1; int v = 0;
	while ((q & 1) == 0) { q >>= 1; ++v; }
	const uint64_t w = two_sylow_gen<m64d, uint64_t>(A, p, q);
	const uint64_t s5 = root_2d<m64d, uint64_t>(A, p, q, v, w, A.to_m(5), 1);
	if (s5 == 0) return;
	const uint64_t inv2 = A.to_m((p + 1) / 2);
	const int d = m - 1, e = (v < d) ? v : d;
	uint64_t u = w;
	for (int i = 0; i < v - e; ++i) u = A.mul(u, u);
	const uint64_t bmax = (p < 10000000000000ull) ? 5000000000ull : 10000000000ull;
	for (int side = 0; side < 2; ++side)
	{
		const uint64_t y = A.mul(side ? A.add(A.one, s5) : A.sub(A.one, s5), inv2);
		const uint64_t x = root_2d<m64d, uint64_t>(A, p, q, v, w, y, d);
		if (x == 0) continue;
		uint64_t r = x;
		for (uint64_t j = 0; j < (1ull << e); ++j) { sink(p, A.from_m(r), bmax); r = A.mul(r, u); }
	}
}

// On-device re-verification of emitted pairs (the check of saveFactors,
// ctsieve.h:270-289): p divides b^(2^m) - b^(2^(m-1)) + 1 (mode +1) or - 1.
__global__ void k_fverify(const uint64_t *pairs, const unsigned long long *ocount, uint64_t cap, int m, int mode,
                          uint32_t *bad)
{
	const uint64_t n = min((uint64_t)*ocount, cap);
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
	{
		const uint64_t p = pairs[2 * i], b = pairs[2 * i + 1];
		const m64d A(p);
		uint64_t x = A.to_m(b);
		for (int k = 0; k < m - 1; ++k) x = A.mul(x, x);
		const uint64_t r = A.from_m(A.sub(A.mul(x, x), x));
		if (r != ((mode > 0) ? p - 1 : 1)) atomicAdd(bad, 1u);
	}
}

// clears n words (a kernel, not hipMemsetAsync)
__global__ void k_fzero(uint32_t *w, uint32_t n)
{
	for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) w[i] = 0;
}

// host re-verification (saveFactors, ctsieve.h:270-289)
static bool host_divides(uint64_t p, uint64_t b, int m, int mode)
{
	const m64h A(p);
	uint64_t x = b % p;
	for (int i = 0; i < m - 1; ++i) x = A.mul(x, x);
	const uint64_t r = A.sub(A.mul(x, x), x);
	return (mode > 0) ? (r == p - 1) : (r == 1);
}

}   // namespace ct

using namespace ct;

#define HIPCHK(call) do { hipError_t _e = (call); if (_e != hipSuccess) { rc = hip_fail(_e, #call); goto done; } } while (0)

int ct_sieve_factors(int device, int m, uint64_t p_min, uint64_t p_max, int mode, uint64_t *pairs, uint64_t cap,
                     uint64_t *count)
{
	if (!count || (cap && !pairs)) return fail(CT_ERR_ARG, "bad argument");
	if (m < 8 || m > 24 || mode == 0) return fail(CT_ERR_ARG, "m in [8, 24], mode +1 or -1");
	if (p_max < p_min || p_max >= (1ull << 63)) return fail(CT_ERR_UNSUPPORTED, "p range must lie below 2^63");
	const uint64_t k_lo = (mode > 0) ? (p_min >> m) / 3 : p_min / 10;
	const uint64_t k_hi = (mode > 0) ? (p_max >> m) / 3 + 1 : p_max / 10 + 1;
	const uint64_t KB = 1ull << 22;
	const uint64_t dcap = std::max<uint64_t>(std::min<uint64_t>(cap, 1ull << 24), 1);
	int rc = CT_OK;
	uint64_t *d_list = nullptr, *d_out = nullptr;
	uint32_t *d_count = nullptr;
	unsigned long long *d_ocount = nullptr;
	hipStream_t st = nullptr;
	uint64_t total = 0;
	{
		HIPCHK(hipSetDevice(device));
		HIPCHK(hipStreamCreate(&st));
		HIPCHK(hipMalloc((void **)&d_list, 2 * KB * sizeof(uint64_t)));
		HIPCHK(hipMalloc((void **)&d_out, 2 * dcap * sizeof(uint64_t)));
		HIPCHK(hipMalloc((void **)&d_count, sizeof(uint32_t)));
		HIPCHK(hipMalloc((void **)&d_ocount, sizeof(unsigned long long)));
		HIPCHK(hipMemsetAsync(d_ocount, 0, sizeof(unsigned long long), st));
		for (uint64_t k0 = k_lo; k0 <= k_hi; k0 += KB)
		{
			const uint64_t kc = std::min(KB, k_hi - k0 + 1);
			HIPCHK(hipMemsetAsync(d_count, 0, sizeof(uint32_t), st));
			hipLaunchKernelGGL(k_fprimes, dim3((uint32_t)((kc + 255) / 256)), dim3(256), 0, st, k0, kc, m, mode, p_min, p_max,
			                   d_list, d_count);
			HIPCHK(hipGetLastError());
			uint32_t cnt = 0;
			HIPCHK(hipMemcpyAsync(&cnt, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
			HIPCHK(hipStreamSynchronize(st));
			if (!cnt) continue;
			const emit_sink sink{d_out, d_ocount, dcap, 0};
			if (mode > 0) hipLaunchKernelGGL(k_froots_pos<emit_sink>, dim3(cnt), dim3(256), 0, st, d_list, d_count, m, sink);
			else hipLaunchKernelGGL(k_froots_neg<emit_sink>, dim3((cnt + 127) / 128), dim3(128), 0, st, d_list, d_count, m, sink);
			HIPCHK(hipGetLastError());
		}
		unsigned long long oc = 0;
		HIPCHK(hipMemcpyAsync(&oc, d_ocount, sizeof(oc), hipMemcpyDeviceToHost, st));
		HIPCHK(hipStreamSynchronize(st));
		total = oc;
		*count = total;
		if (total > cap) { rc = fail(CT_ERR_NOMEM, "pairs buffer too small (count holds the size needed)"); goto done; }
		if (total > dcap) { rc = fail(CT_ERR_UNSUPPORTED, "more than 2^24 pairs in one call: split the p range"); goto done; }
		if (total) HIPCHK(hipMemcpy(pairs, d_out, 2 * total * sizeof(uint64_t), hipMemcpyDeviceToHost));
		for (uint64_t i = 0; i < total; ++i)
			if (!host_divides(pairs[2 * i], pairs[2 * i + 1], m, mode)) { rc = fail(CT_ERR_COMPUTE, "emitted pair does not divide"); goto done; }
	}
done:
	hipFree(d_list); hipFree(d_out); hipFree(d_count); hipFree(d_ocount);
	if (st) hipStreamDestroy(st);
	return rc;
}

// ---------------------------------------------------------------------------
// ct_sieve_factors without allocation: the same kernels with their scratch in
// a file-scope device array.  Layout of `scratch` (per device):
//   [0, 8 KB_NA)  the prime list of one block of k (at most 2 primes per k),
//   then four 32-bit words {prime count, mismatches, unused, unused} and the
//   64-bit emitted-pair counter.
// Every call on one device shares the array; calls that overlap on a device race.

namespace {

constexpr uint64_t KB_NA = 1ull << 20;                                  // k values per block
constexpr size_t kNaList = 2 * KB_NA * sizeof(uint64_t);                // 16 MiB
constexpr size_t kNaScratch = kNaList + 256;

__device__ __align__(256) uint8_t scratch[kNaScratch];

std::mutex g_na_mu;
void *g_na_ptr[64];                                                     // per device, from hipGetSymbolAddress

int na_scratch(int device, uint8_t **out)
{
	if (device < 0 || device >= 64) return fail(CT_ERR_ARG, "device ordinal out of range");
	std::lock_guard<std::mutex> g(g_na_mu);
	if (!g_na_ptr[device])
	{
		void *ptr = nullptr;
		const hipError_t h = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (h != hipSuccess) return hip_fail(h, "hipGetSymbolAddress(scratch)");
		g_na_ptr[device] = ptr;
	}
	*out = static_cast<uint8_t *>(g_na_ptr[device]);
	return CT_OK;
}

}   // namespace

#define NACHK(call) do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail(_e, #call); } while (0)

int ct_sieve_factors_noalloc(int device, int m, uint64_t p_min, uint64_t p_max, int mode, uint64_t *pairs, uint64_t cap,
                             uint64_t *count, void *stream)
{
	if (!count || (cap && !pairs)) return fail(CT_ERR_ARG, "bad argument");
	if (m < 8 || m > 24 || mode == 0) return fail(CT_ERR_ARG, "m in [8, 24], mode +1 or -1");
	if (p_max < p_min || p_max >= (1ull << 63)) return fail(CT_ERR_UNSUPPORTED, "p range must lie below 2^63");
	const uint64_t k_lo = (mode > 0) ? (p_min >> m) / 3 : p_min / 10;
	const uint64_t k_hi = (mode > 0) ? (p_max >> m) / 3 + 1 : p_max / 10 + 1;
	NACHK(hipSetDevice(device));
	uint8_t *base = nullptr;
	int rc = na_scratch(device, &base);
	if (rc) return rc;
	hipStream_t st = static_cast<hipStream_t>(stream);
	uint64_t *d_list = reinterpret_cast<uint64_t *>(base);
	uint32_t *d_words = reinterpret_cast<uint32_t *>(base + kNaList);      // {count, bad, -, -}
	uint32_t *d_count = d_words, *d_bad = d_words + 1;
	unsigned long long *d_ocount = reinterpret_cast<unsigned long long *>(base + kNaList + 16);
	hipLaunchKernelGGL(k_fzero, dim3(1), dim3(64), 0, st, d_words, 6u);     // the four words and the 64-bit counter
	NACHK(hipGetLastError());
	for (uint64_t k0 = k_lo; k0 <= k_hi; k0 += KB_NA)
	{
		const uint64_t kc = std::min(KB_NA, k_hi - k0 + 1);                 // at most 2 kc primes: the list holds them
		hipLaunchKernelGGL(k_fzero, dim3(1), dim3(64), 0, st, d_count, 1u);
		hipLaunchKernelGGL(k_fprimes, dim3((uint32_t)((kc + 255) / 256)), dim3(256), 0, st, k0, kc, m, mode, p_min, p_max,
		                   d_list, d_count);
		NACHK(hipGetLastError());
		uint32_t cnt = 0;
		NACHK(hipMemcpyAsync(&cnt, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
		NACHK(hipStreamSynchronize(st));
		if (!cnt) continue;
		const emit_sink sink{pairs, d_ocount, cap, 0};
		if (mode > 0) hipLaunchKernelGGL(k_froots_pos<emit_sink>, dim3(cnt), dim3(256), 0, st, d_list, d_count, m, sink);
		else hipLaunchKernelGGL(k_froots_neg<emit_sink>, dim3((cnt + 127) / 128), dim3(128), 0, st, d_list, d_count, m, sink);
		NACHK(hipGetLastError());
	}
	hipLaunchKernelGGL(k_fverify, dim3(1024), dim3(256), 0, st, pairs, d_ocount, cap, m, mode, d_bad);
	NACHK(hipGetLastError());
	unsigned long long oc = 0;
	uint32_t bad = 0;
	NACHK(hipMemcpyAsync(&oc, d_ocount, sizeof(oc), hipMemcpyDeviceToHost, st));
	NACHK(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st));
	NACHK(hipStreamSynchronize(st));
	*count = oc;
	if (bad) return fail(CT_ERR_COMPUTE, "emitted pair does not divide");
	if (oc > cap) return fail(CT_ERR_NOMEM, "pairs buffer too small (count holds the size needed)");
	return CT_OK;
}

// ---------------------------------------------------------------------------
// Block-wise emission: the blocks of gfsieve::check (sieve/GPU/src/ctsieve.h:
// 366-409, engine.h:188-220).  Block i holds k in [i 2^21, (i + 1) 2^21): the
// primes p = 3 k 2^m + 1 (mode +1) or p = 10 k -+ 1 (mode -1) of the block
// that lie in [p_min, p_max], their roots, and the pairs (p, b) with
// b <= bMax(p) (or every root).

namespace ct {

constexpr int kLog2Block = 21;                        // _log2GlobalWorkSize, ctsieve.h:174
constexpr uint64_t kBlockK = 1ull << kLog2Block;

// Runs the k range [k_lo, k_hi] (at most kBlockK values per launch of the
// prime kernel) into `sink`, with d_list (2 kBlockK words) and d_count scratch.
template <typename Sink>
static int run_k_range(hipStream_t st, int m, int mode, uint64_t k_lo, uint64_t k_hi, uint64_t p_min, uint64_t p_max,
                       uint64_t *d_list, uint32_t *d_count, const Sink &sink)
{
	for (uint64_t k0 = k_lo; k0 <= k_hi; k0 += kBlockK)
	{
		const uint64_t kc = std::min(kBlockK, k_hi - k0 + 1);
		hipLaunchKernelGGL(k_fzero, dim3(1), dim3(64), 0, st, d_count, 1u);
		hipLaunchKernelGGL(k_fprimes, dim3((uint32_t)((kc + 255) / 256)), dim3(256), 0, st, k0, kc, m, mode, p_min, p_max,
		                   d_list, d_count);
		hipError_t h = hipGetLastError();
		if (h != hipSuccess) return hip_fail(h, "prime kernel");
		uint32_t cnt = 0;
		if ((h = hipMemcpyAsync(&cnt, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, st)) != hipSuccess ||
		    (h = hipStreamSynchronize(st)) != hipSuccess)
			return hip_fail(h, "prime count");
		if (!cnt) continue;
		if (mode > 0) hipLaunchKernelGGL(k_froots_pos<Sink>, dim3(cnt), dim3(256), 0, st, d_list, d_count, m, sink);
		else hipLaunchKernelGGL(k_froots_neg<Sink>, dim3((cnt + 127) / 128), dim3(128), 0, st, d_list, d_count, m, sink);
		if ((h = hipGetLastError()) != hipSuccess) return hip_fail(h, "root kernel");
		if (k0 > k_hi - kc) break;                        // k0 + kBlockK would pass k_hi (or wrap)
	}
	return CT_OK;
}

// The window sieve for primes of any size below 2^63: strikes every b in
// [b_min, b_max] with p | N+-(b) for the primes p in [p_lo, p_hi] of the k
// range [k_lo, k_hi] (d_struck[b - b_min] = 1).
int strike_window64(hipStream_t st, int m, int mode, uint64_t k_lo, uint64_t k_hi, uint64_t p_lo, uint64_t p_hi,
                    uint64_t b_min, uint64_t b_max, uint8_t *d_struck, uint64_t *d_list, uint32_t *d_count)
{
	static_assert(kWindow64List == 2 * kBlockK, "list size of the 64-bit prime kernel");
	if (k_hi < k_lo || p_hi < p_lo) return CT_OK;
	const strike_sink sink{d_struck, b_min, b_max};
	return run_k_range(st, m, mode, k_lo, k_hi, p_lo, p_hi, d_list, d_count, sink);
}

}   // namespace ct

struct ct_fsieve
{
	int device = 0, m = 0, mode = 0;
	uint64_t p_min = 0, p_max = 0, i_min = 0, i_max = 0;
	uint64_t bmax = 0;                                // 0: the reference's bMax(p)
	hipStream_t st = nullptr;
	uint64_t *d_list = nullptr, *d_out = nullptr;
	uint32_t *d_words = nullptr;                     // {prime count, mismatches}
	unsigned long long *d_ocount = nullptr;
	uint64_t out_cap = 0;
};

int ct_fsieve_destroy(ct_fsieve *s)
{
	if (!s) return CT_OK;
	(void)hipSetDevice(s->device);
	if (s->st) (void)hipStreamSynchronize(s->st);
	hipFree(s->d_list); hipFree(s->d_out); hipFree(s->d_words); hipFree(s->d_ocount);
	if (s->st) (void)hipStreamDestroy(s->st);
	delete s;
	return CT_OK;
}

int ct_fsieve_create(int device, int m, uint64_t p_min, uint64_t p_max, int mode, ct_fsieve **out)
{
	if (!out) return fail(CT_ERR_ARG, "out is null");
	*out = nullptr;
	if (m < 8 || m > 24 || (mode != 1 && mode != -1)) return fail(CT_ERR_ARG, "m in [8, 24], mode +1 or -1");
	if (p_max < p_min || p_max >= (1ull << 63)) return fail(CT_ERR_UNSUPPORTED, "p range must lie below 2^63");
	ct_fsieve *s = new ct_fsieve();
	s->device = device; s->m = m; s->mode = mode; s->p_min = p_min; s->p_max = p_max;
	// the block range of gfsieve::check (ctsieve.h:375-377), in the same floating-point arithmetic
	const double f = (mode == 1) ? 1.0 / (double)(3ull << (kLog2Block + m)) : 1.0 / (double)(10u << kLog2Block);
	s->i_min = (uint64_t)std::floor((double)p_min * f);
	s->i_max = (uint64_t)std::ceil((double)p_max * f);
	hipError_t h = hipSetDevice(device);
	if (h == hipSuccess) h = hipStreamCreate(&s->st);
	if (h != hipSuccess) { ct_fsieve_destroy(s); return hip_fail(h, "stream create"); }
	if (hipMalloc((void **)&s->d_list, 2 * kBlockK * sizeof(uint64_t)) != hipSuccess ||
	    hipMalloc((void **)&s->d_words, 2 * sizeof(uint32_t)) != hipSuccess ||
	    hipMalloc((void **)&s->d_ocount, sizeof(unsigned long long)) != hipSuccess)
	{
		ct_fsieve_destroy(s);
		return fail(CT_ERR_NOMEM, "hipMalloc sieve buffers");
	}
	*out = s;
	return CT_OK;
}

int ct_fsieve_blocks(const ct_fsieve *s, uint64_t *i_min, uint64_t *i_max)
{
	if (!s) return fail(CT_ERR_ARG, "null sieve");
	if (i_min) *i_min = s->i_min;
	if (i_max) *i_max = s->i_max;
	return CT_OK;
}

int ct_fsieve_set_bmax(ct_fsieve *s, uint64_t bmax)
{
	if (!s) return fail(CT_ERR_ARG, "null sieve");
	s->bmax = bmax;
	return CT_OK;
}

int ct_fsieve_run(ct_fsieve *s, uint64_t i_begin, uint64_t nblocks, uint64_t *pairs, uint64_t cap, uint64_t *count)
{
	if (!s || !count || (cap && !pairs)) return fail(CT_ERR_ARG, "bad argument");
	*count = 0;
	if (nblocks == 0) return CT_OK;
	// every p of the blocks below 2^63: k < 2^63 / (3 2^m) (+) or 2^63 / 10 (-)
	const uint64_t k_end = (s->mode > 0) ? ((1ull << 63) / 3) >> s->m : (1ull << 63) / 10;
	const uint64_t i_end = k_end >> kLog2Block;
	if (i_begin >= i_end || nblocks > i_end - i_begin) return fail(CT_ERR_ARG, "blocks beyond p = 2^63");
	int rc = hipSetDevice(s->device) == hipSuccess ? CT_OK : fail(CT_ERR_HIP, "hipSetDevice");
	if (rc) return rc;
	const uint64_t dcap = std::max<uint64_t>(std::min<uint64_t>(cap, 1ull << 24), 1);
	if (dcap > s->out_cap)
	{
		hipFree(s->d_out); s->d_out = nullptr; s->out_cap = 0;
		if (hipMalloc((void **)&s->d_out, 2 * dcap * sizeof(uint64_t)) != hipSuccess) return fail(CT_ERR_NOMEM, "hipMalloc pairs");
		s->out_cap = dcap;
	}
	hipLaunchKernelGGL(k_fzero, dim3(1), dim3(64), 0, s->st, s->d_words, 2u);
	hipLaunchKernelGGL(k_fzero, dim3(1), dim3(64), 0, s->st, reinterpret_cast<uint32_t *>(s->d_ocount), 2u);
	const uint64_t k_lo = i_begin << kLog2Block, k_hi = ((i_begin + nblocks) << kLog2Block) - 1;
	const emit_sink sink{s->d_out, s->d_ocount, dcap, s->bmax};
	if ((rc = run_k_range(s->st, s->m, s->mode, k_lo, k_hi, s->p_min, s->p_max, s->d_list, s->d_words, sink))) return rc;
	hipLaunchKernelGGL(k_fverify, dim3(1024), dim3(256), 0, s->st, s->d_out, s->d_ocount, dcap, s->m, s->mode, s->d_words + 1);
	unsigned long long oc = 0;
	uint32_t bad = 0;
	hipError_t h = hipMemcpyAsync(&oc, s->d_ocount, sizeof(oc), hipMemcpyDeviceToHost, s->st);
	if (h == hipSuccess) h = hipMemcpyAsync(&bad, s->d_words + 1, sizeof(bad), hipMemcpyDeviceToHost, s->st);
	if (h == hipSuccess) h = hipStreamSynchronize(s->st);
	if (h != hipSuccess) return hip_fail(h, "pair count");
	*count = oc;
	if (bad) return fail(CT_ERR_COMPUTE, "emitted pair does not divide");
	if (oc > cap) return fail(CT_ERR_NOMEM, "pairs buffer too small (count holds the size needed)");
	if (oc > dcap) return fail(CT_ERR_UNSUPPORTED, "more than 2^24 pairs in one call: run fewer blocks");
	if (oc && (h = hipMemcpy(pairs, s->d_out, 2 * oc * sizeof(uint64_t), hipMemcpyDeviceToHost)) != hipSuccess)
		return hip_fail(h, "pairs read-out");
	return CT_OK;
}

// ---------------------------------------------------------------------------
// gfsieve::check (sieve/GPU/src/ctsieve.h:326-431) over the block cursor:
// the factor file fp<ext> / fn<ext> ("p b" per line), the checkpoint
// ctx<ext> / ctx_n<ext> (blocks done, elapsed seconds), ext =
// "<m>_<p_min / 10^12>_<p_max / 10^12>.txt", the adaptive cadence of reads
// (every j_sync blocks, aiming at record_s seconds and half the pair buffer),
// the quit flag.  The pairs of each read are written sorted by (p, b) (the
// reference writes them in the order its atomics produced them).

namespace {

std::atomic<int> g_fsieve_quit{0};

double fs_now()
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

std::string fs_path(const char *dir, const std::string &leaf)
{
	return (dir && *dir) ? std::string(dir) + "/" + leaf : leaf;
}

}   // namespace

void ct_fsieve_quit(void) { g_fsieve_quit.store(1); }

int ct_fsieve_check(int device, int m, uint64_t p_min, uint64_t p_max, int mode, const char *work_dir, double record_s,
                    uint64_t max_blocks, ct_fsieve_stats *stats)
{
	g_fsieve_quit.store(0);
	ct_fsieve *s = nullptr;
	int rc = ct_fsieve_create(device, m, p_min, p_max, mode, &s);
	if (rc) return rc;
	struct guard { ct_fsieve *s; ~guard() { ct_fsieve_destroy(s); } } g{s};
	const uint64_t tera = 1000000000000ull;
	const std::string ext = std::to_string(m) + "_" + std::to_string(p_min / tera) + "_" + std::to_string(p_max / tera) + ".txt";
	const std::string ctx = fs_path(work_dir, ((mode == 1) ? "ctx" : "ctx_n") + ext);
	const std::string res = fs_path(work_dir, ((mode == 1) ? "fp" : "fn") + ext);
	ct_fsieve_stats st{};
	uint64_t cnt = 0;
	double elapsed = 0;
	{
		std::ifstream in(ctx);
		if (in.is_open() && (in >> cnt >> elapsed)) st.resumed = 1;
		else { cnt = 0; elapsed = 0; }
	}
	const double t0 = fs_now() - elapsed;
	const uint64_t cap = 1ull << 24;                       // _factorSize, ctsieve.h:173
	std::vector<uint64_t> pairs;
	uint64_t j_sync = 10, done_here = 0;
	double last = fs_now();
	for (uint64_t i = s->i_min + cnt; i < s->i_max;)
	{
		if (g_fsieve_quit.load() || (max_blocks && done_here >= max_blocks)) { st.interrupted = 1; break; }
		uint64_t nb = std::min<uint64_t>(j_sync, s->i_max - i);
		if (max_blocks) nb = std::min<uint64_t>(nb, max_blocks - done_here);
		uint64_t count = 0;
		pairs.resize(2 * cap);
		rc = ct_fsieve_run(s, i, nb, pairs.data(), cap, &count);
		if (rc == CT_ERR_NOMEM || rc == CT_ERR_UNSUPPORTED)
		{   // more pairs than the buffer holds: one block at a time (the reference throws,
			// "factor count is too large", ctsieve.h:253)
			if (nb == 1) return fail(CT_ERR_UNSUPPORTED, "factor count of one block is too large");
			j_sync = 1;
			continue;
		}
		if (rc) return rc;
		std::vector<std::pair<uint64_t, uint64_t>> v(count);
		for (uint64_t q = 0; q < count; ++q) v[q] = {pairs[2 * q], pairs[2 * q + 1]};
		std::sort(v.begin(), v.end());
		{
			std::ofstream f(res, std::ios::app);
			if (!f.is_open()) return fail(CT_ERR_ARG, "cannot write the factor file");
			for (const auto &pb : v) f << pb.first << " " << pb.second << "\n";
		}
		i += nb; cnt += nb; done_here += nb; st.factors += count;
		const double now = fs_now(), dt = std::max(now - last, 1e-6);
		last = now;
		{
			std::ofstream c(ctx);
			if (!c.is_open()) return fail(CT_ERR_ARG, "cannot write the checkpoint");
			c << cnt << std::endl << (now - t0) << std::endl;
		}
		// next cadence: record_s seconds and half the pair buffer (ctsieve.h:389-398)
		const uint64_t by_time = (uint64_t)std::llround(nb * record_s / dt);
		const uint64_t by_size = count ? (uint64_t)std::llround(nb * (cap / 2.0) / count) : by_time;
		j_sync = std::max<uint64_t>(std::min(by_time, by_size), 2);
	}
	st.blocks_done = cnt;
	st.blocks_total = s->i_max - s->i_min;
	st.elapsed = fs_now() - t0;
	if (stats) *stats = st;
	return st.interrupted ? CT_INTERRUPTED : CT_OK;
}
