// Device building blocks of the transform modulo X^n - X^(n/2) + 1 (tree of
// cyclo/src/transform.h:336-366, DIF forward / unnormalised inverse as in
// cyclo/ocl/kernel.cl:177-342), written for one "instance": a length-2^LNL
// sub-transform held by 2^LNL / 16 threads with 16 elements per thread.
//
// An instance starts at some level of the global tree.  `base` is s0 + j0 for
// its first level (s0 blocks at that level, j0 = the instance's block), so
// that its level-l twiddles are tw[((base * 16^L + j) << q) + k] in the
// radix-16 passes.  The whole transform is the instance with base = 1 and the
// Phi_6 split at its first level; a row of a two-dimensional schedule is an
// instance with base = R + row and no Phi_6 level.
//
// Layouts of an instance (tt = local thread, q = slot 0..15):
//   pass L < a : k = j*m + i + q*m/16, m = 2^LNL/16^L, i = tt mod m/16, j = tt div m/16
//   L == a     : k = 16 tt + q  (final pass / carry chunks)
// with a = floor((LNL-1)/4) radix-16 passes and f = (LNL-1) mod 4 levels in the
// final pass.  LDS holds one instance in 2^LNL + 2^LNL/32 words (one pad word
// per 32 against bank conflicts).
#pragma once

#include <hip/hip_runtime.h>

#include "ct_arith.h"
#include "ct_tables.h"

namespace ct {

constexpr int E = 16;

template <int LNL>
struct geo
{
	static constexpr uint32_t n = 1u << LNL;
	static constexpr uint32_t T = n / E;
	static constexpr int a = (LNL - 1) / 4;
	static constexpr int f = (LNL - 1) % 4;
	static constexpr uint32_t lds_words = n + (n >> 5);
};

// Padded LDS word address base(tt) + c(q), c a compile-time constant (one base
// VGPR + immediate offsets):
//   L < a : (j m + i + jm/32 + i/32) + (q mm + q mm/32)   (exact: m >= 32, and
//           q mm + i never crosses a multiple of 32 when mm <= 32, while both
//           are multiples of 32 when mm >= 32)
//   L = a : (16 tt + tt/2) + q
template <int LNL, int L>
__device__ __forceinline__ uint32_t lbase(uint32_t tt)
{
	if constexpr (L == geo<LNL>::a) return E * tt + (tt >> 1);
	else
	{
		constexpr uint32_t m = geo<LNL>::n >> (4 * L), mm = m / 16;
		const uint32_t i = tt % mm, jm = (tt / mm) * m;
		return jm + i + (jm >> 5) + (i >> 5);
	}
}

template <int LNL, int L>
__device__ __forceinline__ constexpr uint32_t loff(int q)
{
	if constexpr (L == geo<LNL>::a) return (uint32_t)q;
	else
	{
		constexpr uint32_t mm = (geo<LNL>::n >> (4 * L)) / 16;
		return q * mm + ((q * mm) >> 5);
	}
}

// Element index (unpadded) of slot q in layout L.
template <int LNL, int L>
__device__ __forceinline__ uint32_t elem(uint32_t tt, int q)
{
	if constexpr (L == geo<LNL>::a) return E * tt + q;
	else
	{
		constexpr uint32_t m = geo<LNL>::n >> (4 * L), mm = m / 16;
		return (tt / mm) * m + (tt % mm) + q * mm;
	}
}

template <int LNL, int FROM, int TO, typename V>
__device__ __forceinline__ void xchg(V (&u)[E], V *lds, uint32_t tt)
{
	V *wr = lds + lbase<LNL, FROM>(tt);
	V *rd = lds + lbase<LNL, TO>(tt);
	__syncthreads();
#pragma unroll
	for (int q = 0; q < E; ++q) wr[loff<LNL, FROM>(q)] = u[q];
	__syncthreads();
#pragma unroll
	for (int q = 0; q < E; ++q) u[q] = rd[loff<LNL, TO>(q)];
}

// ---- radix-16 DIF pass over the instance's levels 4L+1 .. 4L+4 -------------
//
// Value ranges (lazy butterflies, p < 2^30):
//   forward (x, y) -> (x + w y, x - w y): x, y in [0, 4p) -> [0, 4p)
//     (x is brought to [0, 2p), t = w y in [0, 2p) by Shoup);
//   inverse (x, y) -> (x + y, (x - y) w): x, y in [0, 2p) -> [0, 2p);
//   the level-1 (Phi_6) forward split takes inputs in [0, 3p) (digits), the
//   level-1 merge returns reduced values in [0, p).

template <int LNL, int P, int L, bool PHI6>
__device__ __forceinline__ void fwd_pass(uint32_t (&u)[E], const tw_entry *__restrict__ twf, uint32_t base, uint32_t tt)
{
	constexpr uint32_t p = kP[P];
	constexpr uint32_t mm = (geo<LNL>::n >> (4 * L)) / 16;
	const uint32_t j = (L == 0) ? 0u : tt / mm;
	const uint32_t sj = (base << (4 * L)) + j;
#pragma unroll
	for (int q = 0; q < 4; ++q)
	{
		const int h = 8 >> q;
		tw_entry w[8];
#pragma unroll
		for (int k = 0; k < (1 << q); ++k) w[k] = twf[(sj << q) + k];
#pragma unroll
		for (int v = 0; v < E; ++v)
		{
			if ((v & h) != 0) continue;
			const tw_entry wv = w[v >> (4 - q)];
			const uint32_t x = red2p(u[v], p);
			const uint32_t s = shoup_neg(u[v + h], wv.w, wv.wp, p);     // 2^32 - w y'
			if (PHI6 && L == 0 && q == 0)
			{   // level 1: mod X^(n/2) - r and X^(n/2) - r', r' = 1 - r:
				// (x + r y, y + x - r y)
				const uint32_t yx = red2p(red2p(u[v + h], p) + x, p);
				u[v] = x - s;
				u[v + h] = yx + s + 2 * p;
			}
			else
			{
				u[v] = x - s;
				u[v + h] = x + s + 2 * p;
			}
		}
	}
}

template <int LNL, int P, int L, bool PHI6>
__device__ __forceinline__ void inv_pass(uint32_t (&u)[E], const tw_entry *__restrict__ twi, uint32_t base, uint32_t tt)
{
	constexpr uint32_t p = kP[P];
	constexpr uint32_t mm = (geo<LNL>::n >> (4 * L)) / 16;
	const uint32_t j = (L == 0) ? 0u : tt / mm;
	const uint32_t sj = (base << (4 * L)) + j;
#pragma unroll
	for (int q = 3; q >= 0; --q)
	{
		const int h = 8 >> q;
		if (PHI6 && L == 0 && q == 0)
		{   // level-1 merge with the 2/n, 2^32 and (1-2r)^-1 scale folded into K
			const tw_entry r = twi[0], K = twi[1];
#pragma unroll
			for (int v = 0; v < h; ++v)
			{
				const uint32_t x = u[v], y = u[v + h];
				const uint32_t vh = shoup(x + y, r.w, r.wp, p);
				u[v + h] = red1p(shoup(y + 2 * p - x, K.w, K.wp, p), p);
				u[v] = red1p(shoup(x + 2 * p - vh, K.w, K.wp, p), p);
			}
			continue;
		}
		tw_entry w[8];
#pragma unroll
		for (int k = 0; k < (1 << q); ++k) w[k] = twi[(sj << q) + k];
#pragma unroll
		for (int v = 0; v < E; ++v)
		{
			if ((v & h) != 0) continue;
			const tw_entry wv = w[v >> (4 - q)];
			const uint32_t x = u[v], y = u[v + h];
			u[v] = red2p(x + y, p);
			u[v + h] = shoup(x + 2 * p - y, wv.w, wv.wp, p);
		}
	}
}

// ---- final pass: last f levels, pointwise square / product, f inverse levels

template <int LNL, int P, bool MUL>
__device__ __forceinline__ void final_pass(uint32_t (&u)[E], uint32_t (&vv)[E], const tw_entry *__restrict__ twf,
                                           const tw_entry *__restrict__ twi, uint32_t base, uint32_t tt)
{
	constexpr uint32_t p = kP[P], pq = kQ[P];
	constexpr int f = geo<LNL>::f, a = geo<LNL>::a;
	constexpr uint32_t n = geo<LNL>::n;
	const uint32_t k0 = E * tt;
#pragma unroll
	for (int q = 0; q < f; ++q)
	{
		const int h = 1 << (f - q);
		const uint32_t s = base << (4 * a + q);
		const tw_entry *tq = twf + s + (k0 >> (f + 1 - q));
#pragma unroll
		for (int z = 0; z < E; ++z)
		{
			if ((z & h) != 0) continue;
			const tw_entry w = tq[z >> (f + 1 - q)];
			{
				const uint32_t x = red2p(u[z], p), sn = shoup_neg(u[z + h], w.w, w.wp, p);
				u[z] = x - sn; u[z + h] = x + sn + 2 * p;
			}
			if (MUL)
			{
				const uint32_t x = red2p(vv[z], p), sn = shoup_neg(vv[z + h], w.w, w.wp, p);
				vv[z] = x - sn; vv[z + h] = x + sn + 2 * p;
			}
		}
	}
	// last level: (u0 + u1 X)(v0 + v1 X) mod X^2 - w^2; products leave 2^-32
	const tw_entry *tl = twf + base * (n / 2) + (k0 >> 1);
#pragma unroll
	for (int z = 0; z < E; z += 2)
	{
		const tw_entry w = tl[z >> 1];
		const uint32_t u0 = red2p(u[z], p), u1 = red2p(u[z + 1], p);
		const uint32_t u1w = 0u - shoup_neg(u[z + 1], w.w, w.wp, p);             // w u1 in [0, 2p)
		if (!MUL)
		{
			const uint32_t c = mont_lazy(u0, u1, p, pq);
			u[z + 1] = red2p(c + c, p);
			u[z] = red2p(mont_lazy(u0, u0, p, pq) + mont_lazy(u1w, u1w, p, pq), p);
		}
		else
		{
			const uint32_t v0 = red2p(vv[z], p), v1 = red2p(vv[z + 1], p);
			const uint32_t v1w = 0u - shoup_neg(vv[z + 1], w.w, w.wp, p);
			u[z + 1] = red2p(mont_lazy(u0, v1, p, pq) + mont_lazy(u1, v0, p, pq), p);
			u[z] = red2p(mont_lazy(u0, v0, p, pq) + mont_lazy(u1w, v1w, p, pq), p);
		}
	}
#pragma unroll
	for (int q = f - 1; q >= 0; --q)
	{
		const int h = 1 << (f - q);
		const uint32_t s = base << (4 * a + q);
		const tw_entry *tq = twi + s + (k0 >> (f + 1 - q));
#pragma unroll
		for (int z = 0; z < E; ++z)
		{
			if ((z & h) != 0) continue;
			const tw_entry w = tq[z >> (f + 1 - q)];
			const uint32_t x = u[z], y = u[z + h];
			u[z] = red2p(x + y, p);
			u[z + h] = shoup(x + 2 * p - y, w.w, w.wp, p);
		}
	}
}

// forward passes 0..a-1, each followed by the transpose to the next layout
// (the last one lands in layout a, the final-pass layout)
template <int LNL, int P, int L, bool PHI6>
__device__ __forceinline__ void fwd_chain(uint32_t (&u)[E], uint32_t *lds, const tw_entry *twf, uint32_t base, uint32_t tt)
{
	if constexpr (L < geo<LNL>::a)
	{
		fwd_pass<LNL, P, L, PHI6>(u, twf, base, tt);
		xchg<LNL, L, L + 1>(u, lds, tt);
		fwd_chain<LNL, P, L + 1, PHI6>(u, lds, twf, base, tt);
	}
}

// inverse passes L..0, each preceded by the transpose from layout L+1
template <int LNL, int P, int L, bool PHI6>
__device__ __forceinline__ void inv_chain(uint32_t (&u)[E], uint32_t *lds, const tw_entry *twi, uint32_t base, uint32_t tt)
{
	if constexpr (L >= 0)
	{
		xchg<LNL, L + 1, L>(u, lds, tt);
		inv_pass<LNL, P, L, PHI6>(u, twi, base, tt);
		inv_chain<LNL, P, L - 1, PHI6>(u, lds, twi, base, tt);
	}
}

// Garner CRT of the three engine residues to a signed 96-bit integer in
// (-M/2, M/2] (the role of garner3, kernel.cl:451-460).
__device__ __forceinline__ i96 garner(uint32_t r1, uint32_t r2, uint32_t r3, const garner_consts &g)
{
	constexpr uint32_t p1 = kP[0], p2 = kP[1], p3 = kP[2];
	const uint32_t u13 = mont_mulc(sub_mod(r1, r3, p1), g.inv3_1, g.inv3_1q, p1);   // r3 < p3 < p1
	const uint32_t u23 = mont_mulc(sub_mod(r2, r3, p2), g.inv3_2, g.inv3_2q, p2);   // r3 < p3 < p2
	const uint32_t u123 = mont_mulc(sub_mod(u13, u23, p1), g.inv2_1, g.inv2_1q, p1);// u23 < p2 < p1
	const uint64_t s = (uint64_t)p2 * u123 + u23;                                     // < 2^62
	const uint64_t lo = (uint64_t)p3 * (uint32_t)s + r3;
	const uint64_t hi = (uint64_t)p3 * (uint32_t)(s >> 32) + (lo >> 32);
	uint64_t v_lo = (hi << 32) | (uint32_t)lo;
	uint32_t v_hi = (uint32_t)(hi >> 32);
	const uint64_t h_lo = ((uint64_t)g.h_mid << 32) | g.h_lo;
	const bool big = (v_hi > g.h_hi) || ((v_hi == g.h_hi) && (v_lo > h_lo));
	if (big)
	{   // v - M < 0
		const uint64_t m_lo = ((uint64_t)g.m_mid << 32) | g.m_lo;
		const uint64_t d_lo = v_lo - m_lo;
		v_hi = v_hi - g.m_hi - ((v_lo < m_lo) ? 1u : 0u);
		v_lo = d_lo;
	}
	i96 r; r.lo = v_lo; r.hi = (int32_t)v_hi;
	return r;
}

// Two-prime CRT (used when every coefficient is known to lie below
// p1 p2 / 2 in magnitude): x = r2 + p2 ((r1 - r2) / p2 mod p1), centred.
__device__ __forceinline__ int64_t garner2(uint32_t r1, uint32_t r2, const garner_consts &g)
{
	constexpr uint32_t p1 = kP[0], p2 = kP[1];
	const uint32_t u = mont_mulc(sub_mod(r1, r2, p1), g.inv2_1, g.inv2_1q, p1);      // r2 < p2 < p1
	const uint64_t v = (uint64_t)p2 * u + r2;                                          // < p1 p2 < 2^62
	return (v > g.h2) ? (int64_t)(v - g.m2) : (int64_t)v;
}

// normalize1 with a 64-bit accumulator (two-prime mode: |l| < 2^61)
__device__ __forceinline__ int64_t carry_chunk64(const int64_t (&l)[E], int32_t (&dig)[E], bool dup, const divisor &dv)
{
	int64_t acc = 0;
#pragma unroll
	for (int q = 0; q < E; ++q)
	{
		acc += l[q];
		if (dup) acc += l[q];
		dig[q] = divrem_i64(acc, dv);
	}
	return acc;
}

// normalize1 on one 16-coefficient chunk (kernel.cl:1029-1051): returns the
// chunk's outgoing carry.
__device__ __forceinline__ int64_t carry_chunk(const i96 (&l)[E], int32_t (&dig)[E], bool dup, const divisor &dv)
{
	i96 acc; acc.lo = 0; acc.hi = 0;
#pragma unroll
	for (int q = 0; q < E; ++q)
	{
		acc = i96_add(acc, l[q]);
		if (dup) acc = i96_add(acc, l[q]);
		dig[q] = divrem_i96(acc, dv);
	}
	return (int64_t)acc.lo;
}

// normalize2 on one chunk (kernel.cl:1053-1081): the incoming carry c is
// propagated until it vanishes; the last digit absorbs what is left.
__device__ __forceinline__ void carry_in(int32_t (&dig)[E], int64_t c, const divisor &dv)
{
	bool live = true;
#pragma unroll
	for (int q = 0; q < E - 1; ++q)
	{
		if (live)
		{
			c += dig[q];
			dig[q] = divrem_i64(c, dv);
			live = (c != 0);
		}
	}
	if (live) dig[E - 1] += (int32_t)c;
}

}   // namespace ct
