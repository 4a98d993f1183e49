// CPU comparator (test infrastructure and the bench's CPU baseline, never the
// product): the squaring engine's algorithm restated for host cores in C++
// with OpenMP -- the same transform modulo X^n - X^(n/2) + 1 over the
// engine's two NTT primes (root tree of cyclo/src/transform.h:336-366, lazy
// Harvey butterflies with Shoup products from ctwin_amd/csrc/ct_arith.h,
// root tables from ct_tables.h), the sqr2 leaf (kernel.cl:189-194), the
// two-prime Garner CRT and the 16-digit chunk carry of normalize1/normalize2
// (kernel.cl:1029-1081).  Its digits equal the engine's digit for digit.
//
// Parallel over candidates when there are several; over the butterflies of
// each level and over cache-sized blocks of the tree when there is one.  The
// levels whose blocks are larger than kBlock elements run as whole-array
// passes; below that each block runs its remaining forward levels, the
// pointwise square and the inverse levels up to its size while it is in cache.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../ctwin_amd/csrc/ct_arith.h"
#include "../../ctwin_amd/csrc/ct_tables.h"

using namespace ct;

namespace {

constexpr uint32_t kBlock = 1u << 13;        // elements per in-cache block (32 KiB per prime)
constexpr uint32_t kChunk = 16;              // carry chunk (csize)
constexpr uint32_t kTask = 2048;             // butterflies per task of a whole-array level

struct cpu_engine
{
	uint32_t ln = 0, n = 0, lanes = 0;
	std::vector<uint32_t> b;
	std::vector<divisor> dv;
	std::vector<tw_entry> tw;                // [prime][dir][n] as on the device
	garner_consts gc{};
	std::vector<uint32_t> res;               // [lanes][2][n] residues
	std::vector<int64_t> carry;              // [lanes][n / kChunk]
};

inline const tw_entry *fwd_tab(const cpu_engine &e, int P) { return &e.tw[(size_t)(P * 2) * e.n]; }
inline const tw_entry *inv_tab(const cpu_engine &e, int P) { return &e.tw[(size_t)(P * 2 + 1) * e.n]; }

// forward butterfly on block j of level s (block size m), elements [lo, hi) of its first half
inline void fwd_block(uint32_t *u, uint32_t base, uint32_t m, tw_entry w, uint32_t p, uint32_t lo, uint32_t hi)
{
	const uint32_t h = m / 2;
	for (uint32_t i = lo; i < hi; ++i)
	{
		const uint32_t x = red2p(u[base + i], p);
		const uint32_t s = shoup_neg(u[base + i + h], w.w, w.wp, p);
		u[base + i] = x - s;
		u[base + i + h] = x + s + 2 * p;
	}
}

inline void inv_block(uint32_t *u, uint32_t base, uint32_t m, tw_entry w, uint32_t p, uint32_t lo, uint32_t hi)
{
	const uint32_t h = m / 2;
	for (uint32_t i = lo; i < hi; ++i)
	{
		const uint32_t x = u[base + i], y = u[base + i + h];
		u[base + i] = red2p(x + y, p);
		u[base + i + h] = shoup(x + 2 * p - y, w.w, w.wp, p);
	}
}

// Levels below block size `top` of one block of `top` elements at level s0
// (s0 = n / top blocks, this is block j0): forward levels, the sqr2 leaf,
// inverse levels back to block size `top`.
void block_square(uint32_t *u, uint32_t n, uint32_t top, uint32_t j0, const tw_entry *fw, const tw_entry *iw, uint32_t p,
                  uint32_t q)
{
	const uint32_t base0 = j0 * top;
	for (uint32_t m = top, s = n / top; m >= 4; m /= 2, s *= 2)
	{
		const uint32_t per = top / m;        // blocks of this level inside the block
		for (uint32_t t = 0; t < per; ++t) fwd_block(u, base0 + t * m, m, fw[s + j0 * per + t], p, 0, m / 2);
	}
	const uint32_t s_leaf = n / 2;
	for (uint32_t z = 0; z < top; z += 2)
	{
		const tw_entry w = fw[s_leaf + (base0 + z) / 2];
		const uint32_t u0 = red2p(u[base0 + z], p), u1 = red2p(u[base0 + z + 1], p);
		const uint32_t u1w = 0u - shoup_neg(u[base0 + z + 1], w.w, w.wp, p);
		const uint32_t c = mont_lazy(u0, u1, p, q);
		u[base0 + z + 1] = red2p(c + c, p);
		u[base0 + z] = red2p(mont_lazy(u0, u0, p, q) + mont_lazy(u1w, u1w, p, q), p);
	}
	for (uint32_t m = 4, s = n / 4; m <= top; m *= 2, s /= 2)
	{
		const uint32_t per = top / m;
		for (uint32_t t = 0; t < per; ++t) inv_block(u, base0 + t * m, m, iw[s + j0 * per + t], p, 0, m / 2);
	}
}

// One prime's square of one residue vector u (digits already converted).
// par: run the whole-array levels and the blocks in parallel (one candidate).
void square_prime(uint32_t *u, const cpu_engine &e, int P, bool par)
{
	const uint32_t n = e.n, h = n / 2, p = kP[P], q = kQ[P];
	const tw_entry *fw = fwd_tab(e, P), *iw = inv_tab(e, P);
	const uint32_t top = std::min(kBlock, n / 2);
	// level 1: the Phi_6 split (x + r y, y + x - r y)
	{
		const tw_entry w = fw[1];
#pragma omp parallel for schedule(static) if (par)
		for (uint32_t i = 0; i < h; ++i)
		{
			const uint32_t x = red2p(u[i], p);
			const uint32_t s = shoup_neg(u[i + h], w.w, w.wp, p);
			const uint32_t yx = red2p(red2p(u[i + h], p) + x, p);
			u[i] = x - s;
			u[i + h] = yx + s + 2 * p;
		}
	}
	// whole-array levels with blocks larger than `top`
	for (uint32_t m = n / 2, s = 2; m > top; m /= 2, s *= 2)
	{
		// tasks of kTask contiguous butterflies (m / 2 >= kBlock / 2 >= kTask)
		const uint32_t per = (m / 2) / kTask;
#pragma omp parallel for schedule(static) if (par)
		for (uint32_t k = 0; k < (n / 2) / kTask; ++k)
		{
			const uint32_t j = k / per, i = (k % per) * kTask;
			fwd_block(u, j * m, m, fw[s + j], p, i, i + kTask);
		}
	}
#pragma omp parallel for schedule(static) if (par)
	for (uint32_t j0 = 0; j0 < n / top; ++j0) block_square(u, n, top, j0, fw, iw, p, q);
	for (uint32_t m = 2 * top, s = n / (2 * top); m <= n / 2; m *= 2, s /= 2)
	{
		const uint32_t per = (m / 2) / kTask;
#pragma omp parallel for schedule(static) if (par)
		for (uint32_t k = 0; k < (n / 2) / kTask; ++k)
		{
			const uint32_t j = k / per, i = (k % per) * kTask;
			inv_block(u, j * m, m, iw[s + j], p, i, i + kTask);
		}
	}
	// level-1 merge with the 2/n, 2^32 and (1 - 2r)^-1 scale in K
	{
		const tw_entry r = iw[0], K = iw[1];
#pragma omp parallel for schedule(static) if (par)
		for (uint32_t i = 0; i < h; ++i)
		{
			const uint32_t x = u[i], y = u[i + h];
			const uint32_t vh = shoup(x + y, r.w, r.wp, p);
			u[i + h] = red1p(shoup(y + 2 * p - x, K.w, K.wp, p), p);
			u[i] = red1p(shoup(x + 2 * p - vh, K.w, K.wp, p), p);
		}
	}
}

inline int64_t garner2_host(uint32_t r1, uint32_t r2, const garner_consts &g)
{
	constexpr uint32_t p1 = kP[0], p2 = kP[1];
	const uint32_t u = mont_mulc(sub_mod(r1, r2, p1), g.inv2_1, g.inv2_1q, p1);
	const uint64_t v = (uint64_t)p2 * u + r2;
	return (v > g.h2) ? (int64_t)(v - g.m2) : (int64_t)v;
}

// x <- x^2 * 2^dup for one lane (digits in place)
void square_lane(cpu_engine &e, uint32_t l, int32_t *x, bool dup, bool par)
{
	const uint32_t n = e.n, nc = n / kChunk;
	uint32_t *r0 = &e.res[(size_t)l * 2 * n], *r1 = r0 + n;
#pragma omp parallel for schedule(static) if (par)
	for (uint32_t k = 0; k < n; ++k)
	{
		r0[k] = (uint32_t)x[k] + 2 * kP[0];
		r1[k] = (uint32_t)x[k] + 2 * kP[1];
	}
	square_prime(r0, e, 0, par);
	square_prime(r1, e, 1, par);
	int64_t *f = &e.carry[(size_t)l * nc];
	const divisor &dv = e.dv[l];
	const int64_t mul = dup ? 2 : 1;
	// normalize1: per chunk, truncating division of the running sum
#pragma omp parallel for schedule(static) if (par)
	for (uint32_t g = 0; g < nc; ++g)
	{
		int64_t acc = 0;
		for (uint32_t c = 0; c < kChunk; ++c)
		{
			const uint32_t k = g * kChunk + c;
			acc += mul * garner2_host(r0[k], r1[k], e.gc);
			x[k] = divrem_i62(acc, dv);
		}
		f[g] = acc;
	}
	// normalize2: the previous chunk's carry (X^n = X^(n/2) - 1: the top
	// carry goes negated to chunk 0 and added to chunk nc/2)
#pragma omp parallel for schedule(static) if (par)
	for (uint32_t g = 0; g < nc; ++g)
	{
		int64_t c = f[(g + nc - 1) % nc];
		if (g == 0) c = -c;
		else if (g == nc / 2) c += f[nc - 1];
		int32_t *d = &x[g * kChunk];
		bool live = true;
		for (uint32_t k = 0; k < kChunk - 1 && live; ++k)
		{
			c += d[k];
			d[k] = divrem_i64(c, dv);
			live = (c != 0);
		}
		if (live) d[kChunk - 1] += (int32_t)c;
	}
}

}   // namespace

extern "C" {

// n = 2^ln digits, `lanes` candidates with b[lanes]; two-prime mode only
// (3 n b^2 < p1 p2 / 2, as the engine checks): returns null otherwise.
void *cp_new(uint32_t ln, uint32_t lanes, const uint32_t *b)
{
	if (ln < 5 || ln > (uint32_t)kMaxLog2N || lanes == 0) return nullptr;
	cpu_engine *e = new cpu_engine();
	e->ln = ln; e->n = 1u << ln; e->lanes = lanes;
	e->gc = make_garner();
	for (uint32_t l = 0; l < lanes; ++l)
	{
		const unsigned __int128 bound = (unsigned __int128)3 * e->n * b[l] * b[l];
		if (b[l] <= 32 || b[l] >= 0x80000000u || bound >= e->gc.h2) { delete e; return nullptr; }
		e->b.push_back(b[l]);
		e->dv.push_back(make_divisor(b[l]));
	}
	e->tw = build_tables(e->n);
	e->res.assign((size_t)lanes * 2 * e->n, 0);
	e->carry.assign((size_t)lanes * (e->n / kChunk), 0);
	return e;
}

void cp_free(void *ve) { delete static_cast<cpu_engine *>(ve); }

// `nsteps` PRP-loop squarings of every lane: step t multiplies by 2 when bit
// (i_hi - t) of the lane's exponent is set (bits[lane * words + i / 32]).
// threads: OpenMP threads (over lanes when lanes > 1, else inside the transform).
int cp_steps(void *ve, int32_t *digits, const uint32_t *bits, uint32_t words, int i_hi, int nsteps, int threads)
{
	cpu_engine &e = *static_cast<cpu_engine *>(ve);
	if (i_hi - nsteps + 1 < 0 || (uint64_t)i_hi >= (uint64_t)words * 32) return -1;
	omp_set_num_threads(std::max(1, threads));
	if (e.lanes > 1)
	{
#pragma omp parallel for schedule(dynamic, 1)
		for (uint32_t l = 0; l < e.lanes; ++l)
			for (int t = 0; t < nsteps; ++t)
			{
				const int i = i_hi - t;
				const bool dup = (bits[(size_t)l * words + i / 32] >> (i % 32)) & 1u;
				square_lane(e, l, &digits[(size_t)l * e.n], dup, false);
			}
		return 0;
	}
	for (int t = 0; t < nsteps; ++t)
	{
		const int i = i_hi - t;
		square_lane(e, 0, digits, (bits[i / 32] >> (i % 32)) & 1u, threads > 1);
	}
	return 0;
}

}
