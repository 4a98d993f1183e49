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

// Store of a residue that the next launch reads, possibly on another XCD: a
// system-scope relaxed store (gfx950 `sc0 sc1`) writes through the L2, so the
// plane drains to memory while the launch still computes instead of in the
// write-back of dirty L2 lines at its end (measured at n = 2^21: 1.4 us less
// per row-pass launch).  The value and the order of the stores are those of a
// plain store; the launch boundary orders them before the next launch's reads.
__device__ __forceinline__ void store_through(uint32_t *p, uint32_t v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

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

// Workgroup transpose from layout FROM to layout TO through lds.  In every
// layout the threads own disjoint sets of words, so the leading barrier (which
// orders the writes after the workgroup's earlier accesses of lds) can go
// (LEAD = false) when each thread's last access of lds since the previous
// workgroup barrier was a read of its own FROM-layout words -- the TO read of
// the exchange before it in a chain.
template <int LNL, int FROM, int TO, bool LEAD = true, typename V>
__device__ __forceinline__ void xchg(V (&u)[E], V *lds, uint32_t tt)
{
	V *wr = lds + lbase<LNL, FROM>(tt);
	V *rd = lds + lbase<LNL, TO>(tt);
	if constexpr (LEAD) __syncthreads();
#pragma unroll
	for (int q = 0; q < E; ++q) wr[loff<LNL, FROM>(q)] = u[q];
	__syncthreads();
#pragma unroll
	for (int q = 0; q < E; ++q) u[q] = rd[loff<LNL, TO>(q)];
}

// Layout L when tt is the workgroup's flat thread index: the 64 threads of a
// wavefront own 1024 consecutive elements -- a contiguous padded LDS range
// of their own -- in the final layout and whenever a block spans at most one
// wavefront (mm <= 64).
template <int LNL, int L>
__device__ __forceinline__ constexpr bool wave_local()
{
	if constexpr (L == geo<LNL>::a) return true;
	else return ((geo<LNL>::n >> (4 * L)) / 16) <= 64;
}

// Transpose for flat thread indices: when both layouts keep every wavefront's
// elements in its own LDS range, and the previous transpose already left each
// wavefront alone in its range, no other wavefront reads or writes it, so the
// workgroup barriers reduce to ordering within the wavefront (its LDS
// operations execute in order; wave_barrier keeps the compiler from moving
// them across).
template <int LNL, int FROM, int TO, bool LEAD = true, typename V>
__device__ __forceinline__ void xchg_flat(V (&u)[E], V *lds, uint32_t tt)
{
	if constexpr (wave_local<LNL, FROM>() && wave_local<LNL, TO>())
	{
		V *wr = lds + lbase<LNL, FROM>(tt);
		V *rd = lds + lbase<LNL, TO>(tt);
		__builtin_amdgcn_wave_barrier();
#pragma unroll
		for (int q = 0; q < E; ++q) wr[loff<LNL, FROM>(q)] = u[q];
		__builtin_amdgcn_wave_barrier();
#pragma unroll
		for (int q = 0; q < E; ++q) u[q] = rd[loff<LNL, TO>(q)];
		__builtin_amdgcn_wave_barrier();
	}
	else xchg<LNL, FROM, TO, LEAD>(u, lds, tt);
}

// ---- twiddles of one pass, loaded one pass ahead of their use -------------
//
// A radix-16 pass over levels 4L+1..4L+4 of an instance needs 1 + 2 + 4 + 8
// roots per thread: level q (0..3) uses tw[(sj << q) + k], k < 2^q, with
// sj = (base << 4L) + j.  They are loaded into registers while the previous
// pass computes, so the global-memory latency is not exposed between passes
// (the passes are separated by LDS transposes and workgroup barriers).

struct ptw                                // roots of one pass held in registers
{
	tw_entry w[15];                       // level q, root k at w[(1 << q) - 1 + k]
	__device__ __forceinline__ tw_entry get(int q, int k) const { return w[(1 << q) - 1 + k]; }
};
struct ptw_lazy                           // the same roots read where they are used
{
	const tw_entry *tw; uint32_t sj;
	__device__ __forceinline__ tw_entry get(int q, int k) const { return tw[(sj << q) + k]; }
};
struct mtw { tw_entry r, K; };           // level-1 merge constants (inverse table entries 0, 1)

// Block index of thread tt in pass L.  When a block spans a multiple of 64
// threads, every lane of a wavefront has the same block, and readfirstlane
// makes that visible to the compiler: the roots are then read with scalar
// loads into SGPRs.  (tt is the thread index in every caller that reaches a
// pass with mm >= 64: the resident kernels and the row pass.)
template <int LNL, int L>
__device__ __forceinline__ uint32_t pass_block(uint32_t tt)
{
	constexpr uint32_t mm = (geo<LNL>::n >> (4 * L)) / 16;
	if constexpr (L == 0) return 0u;
	else if constexpr (mm >= 64 && mm % 64 == 0) return __builtin_amdgcn_readfirstlane(tt / mm);
	else return tt / mm;
}

template <int LNL, int L>
__device__ __forceinline__ ptw_lazy lazy_ptw(const tw_entry *tw, uint32_t base, uint32_t tt)
{
	ptw_lazy T; T.tw = tw; T.sj = (base << (4 * L)) + pass_block<LNL, L>(tt);
	return T;
}

template <int LNL, int L>
__device__ __forceinline__ void load_ptw(ptw &T, const tw_entry *__restrict__ tw, uint32_t base, uint32_t tt)
{
	const uint32_t sj = (base << (4 * L)) + pass_block<LNL, L>(tt);
#pragma unroll
	for (int q = 0; q < 4; ++q)
	{
#pragma unroll
		for (int k = 0; k < (1 << q); ++k) T.w[(1 << q) - 1 + k] = tw[(sj << q) + k];
	}
}

__device__ __forceinline__ mtw load_mtw(const tw_entry *__restrict__ twi) { mtw m; m.r = twi[0]; m.K = twi[1]; return m; }

// Final-pass roots: f forward levels (level q holds 2^(3-f+q) distinct roots
// per thread), the leaf level (8), and the f inverse levels.
template <int LNL>
struct ftw
{
	static constexpr int f = geo<LNL>::f;
	static constexpr int nf = (f > 0) ? 8 - (8 >> f) : 1;
	tw_entry fq[nf];
	tw_entry last[8];
	tw_entry iq[nf];
	__device__ __forceinline__ tw_entry fwd(int q, int c) const { return fq[(8 >> f) * ((1 << q) - 1) + c]; }
	__device__ __forceinline__ tw_entry leaf(int c) const { return last[c]; }
	__device__ __forceinline__ tw_entry inv(int q, int c) const { return iq[(8 >> f) * ((1 << q) - 1) + c]; }
};

template <int LNL>
struct ftw_lazy
{
	static constexpr int f = geo<LNL>::f, a = geo<LNL>::a;
	const tw_entry *twf, *twi; uint32_t base, k0;
	__device__ __forceinline__ tw_entry fwd(int q, int c) const { return twf[(base << (4 * a + q)) + (k0 >> (f + 1 - q)) + c]; }
	__device__ __forceinline__ tw_entry leaf(int c) const { return twf[base * (geo<LNL>::n / 2) + (k0 >> 1) + c]; }
	__device__ __forceinline__ tw_entry inv(int q, int c) const { return twi[(base << (4 * a + q)) + (k0 >> (f + 1 - q)) + c]; }
};

template <int LNL>
__device__ __forceinline__ void load_ftw(ftw<LNL> &F, const tw_entry *__restrict__ twf, const tw_entry *__restrict__ twi,
                                         uint32_t base, uint32_t tt)
{
	constexpr int f = geo<LNL>::f, a = geo<LNL>::a;
	constexpr uint32_t n = geo<LNL>::n;
	const uint32_t k0 = E * tt;
#pragma unroll
	for (int q = 0; q < f; ++q)
	{
		const uint32_t s = base << (4 * a + q);
		const uint32_t off = (8 >> f) * ((1u << q) - 1), cnt = 16u >> (f + 1 - q);
#pragma unroll
		for (uint32_t c = 0; c < cnt; ++c)
		{
			F.fq[off + c] = twf[s + (k0 >> (f + 1 - q)) + c];
			F.iq[off + c] = twi[s + (k0 >> (f + 1 - q)) + c];
		}
	}
	const tw_entry *tl = twf + base * (n / 2) + (k0 >> 1);
#pragma unroll
	for (int z = 0; z < 8; ++z) F.last[z] = tl[z];
}

// ---- radix-16 DIF pass over the instance's levels 4L+1 .. 4L+4 -------------
//
// Value ranges (lazy butterflies, p < 2^30):
//   forward (x, y) -> (x + w y, x - w y): x, y in [0, 4p) -> [0, 4p)
//     (x is brought to [0, 2p), t = w y in [0, 2p) by Shoup);
//   inverse (x, y) -> (x + y, (x - y) w): x, y in [0, 2p) -> [0, 2p);
//   the level-1 (Phi_6) forward split takes inputs in [0, 3p) (digits), the
//   level-1 merge returns reduced values in [0, p).

template <int P, bool LEVEL1, typename TW>
__device__ __forceinline__ void fwd_pass(uint32_t (&u)[E], const TW &T)
{
	constexpr uint32_t p = kP[P];
#pragma unroll
	for (int q = 0; q < 4; ++q)
	{
		const int h = 8 >> q;
#pragma unroll
		for (int v = 0; v < E; ++v)
		{
			if ((v & h) != 0) continue;
			const tw_entry wv = T.get(q, v >> (4 - q));
			const uint32_t x = red2p(u[v], p);
			const uint32_t s = shoup_neg(u[v + h], wv.w, wv.wp, p);     // 2^32 - w y'
			if (LEVEL1 && q == 0)
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

template <int P, bool LEVEL1, typename TW>
__device__ __forceinline__ void inv_pass(uint32_t (&u)[E], const TW &T, const mtw &M)
{
	constexpr uint32_t p = kP[P];
#pragma unroll
	for (int q = 3; q >= 0; --q)
	{
		const int h = 8 >> q;
		if (LEVEL1 && q == 0)
		{   // level-1 merge with the 2/n, 2^32 and (1-2r)^-1 scale folded into K
#pragma unroll
			for (int v = 0; v < h; ++v)
			{
				const uint32_t x = u[v], y = u[v + h];
				const uint32_t vh = shoup(x + y, M.r.w, M.r.wp, p);
				u[v + h] = red1p(shoup(y + 2 * p - x, M.K.w, M.K.wp, p), p);
				u[v] = red1p(shoup(x + 2 * p - vh, M.K.w, M.K.wp, p), p);
			}
			continue;
		}
#pragma unroll
		for (int v = 0; v < E; ++v)
		{
			if ((v & h) != 0) continue;
			const tw_entry wv = T.get(q, v >> (4 - q));
			const uint32_t x = u[v], y = u[v + h];
			u[v] = red2p(x + y, p);
			u[v + h] = shoup(x + 2 * p - y, wv.w, wv.wp, p);
		}
	}
}

// ---- final pass: last f levels, pointwise square / product, f inverse levels

template <int LNL, int P, bool MUL, typename FT>
__device__ __forceinline__ void final_pass(uint32_t (&u)[E], uint32_t (&vv)[E], const FT &F)
{
	constexpr uint32_t p = kP[P], pq = kQ[P];
	constexpr int f = geo<LNL>::f;
#pragma unroll
	for (int q = 0; q < f; ++q)
	{
		const int h = 1 << (f - q);
#pragma unroll
		for (int z = 0; z < E; ++z)
		{
			if ((z & h) != 0) continue;
			const tw_entry w = F.fwd(q, z >> (f + 1 - q));
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
#pragma unroll
	for (int z = 0; z < E; z += 2)
	{
		const tw_entry w = F.leaf(z >> 1);
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
#pragma unroll
		for (int z = 0; z < E; ++z)
		{
			if ((z & h) != 0) continue;
			const tw_entry w = F.inv(q, z >> (f + 1 - q));
			const uint32_t x = u[z], y = u[z + h];
			u[z] = red2p(x + y, p);
			u[z + h] = shoup(x + 2 * p - y, w.w, w.wp, p);
		}
	}
}

// forward passes L..a-1, each followed by the transpose to the next layout
// (the last one lands in layout a, the final-pass layout).  PF: the roots of
// pass L+1 (and finally those of the final pass, into F) are loaded into
// registers before pass L computes, hiding their latency at the price of live
// VGPRs; otherwise each pass reads its roots where it uses them.  Every
// exchange after the first follows the previous one's read of its own words
// and needs no leading barrier; the first has one unless LEAD0 is false (since
// the workgroup's last barrier no thread has touched lds, or each has only
// read its own layout-0 words).
template <int LNL, int P, int L, bool PHI6, bool PF, bool LEAD0>
__device__ __forceinline__ void fwd_chain(uint32_t (&u)[E], uint32_t *lds, const tw_entry *twf, const tw_entry *twi,
                                          uint32_t base, uint32_t tt, const ptw &cur, ftw<LNL> &F)
{
	if constexpr (L < geo<LNL>::a)
	{
		constexpr bool last = (L + 1 == geo<LNL>::a);
		ptw nxt;
		if constexpr (PF)
		{
			if constexpr (!last) load_ptw<LNL, L + 1>(nxt, twf, base, tt);
			else load_ftw<LNL>(F, twf, twi, base, tt);
			fwd_pass<P, PHI6 && L == 0>(u, cur);
		}
		else fwd_pass<P, PHI6 && L == 0>(u, lazy_ptw<LNL, L>(twf, base, tt));
		xchg_flat<LNL, L, L + 1, (L == 0 && LEAD0)>(u, lds, tt);
		if constexpr (!last) fwd_chain<LNL, P, L + 1, PHI6, PF, LEAD0>(u, lds, twf, twi, base, tt, nxt, F);
	}
}

// Forward passes of one instance from layout 0.
template <int LNL, int P, bool PHI6, bool PF, bool LEAD0 = true>
__device__ __forceinline__ void fwd_all(uint32_t (&u)[E], uint32_t *lds, const tw_entry *twf, const tw_entry *twi,
                                        uint32_t base, uint32_t tt, ftw<LNL> &F)
{
	static_assert(geo<LNL>::a > 0, "instances have at least one radix-16 pass");
	ptw c0;
	if constexpr (PF) load_ptw<LNL, 0>(c0, twf, base, tt);
	fwd_chain<LNL, P, 0, PHI6, PF, LEAD0>(u, lds, twf, twi, base, tt, c0, F);
}

// inverse passes L..0, each preceded by the transpose from layout L+1 (the
// first follows the forward chain's read of layout a, each later one the
// previous exchange's read: no leading barriers)
template <int LNL, int P, int L, bool PHI6, bool PF>
__device__ __forceinline__ void inv_chain(uint32_t (&u)[E], uint32_t *lds, const tw_entry *twi, uint32_t base, uint32_t tt,
                                          const ptw &cur, const mtw &M)
{
	if constexpr (L >= 0)
	{
		ptw nxt;
		if constexpr (PF && L > 0) load_ptw<LNL, L - 1>(nxt, twi, base, tt);
		xchg_flat<LNL, L + 1, L, false>(u, lds, tt);
		if constexpr (PF) inv_pass<P, PHI6 && L == 0>(u, cur, M);
		else inv_pass<P, PHI6 && L == 0>(u, lazy_ptw<LNL, L>(twi, base, tt), M);
		if constexpr (L > 0) inv_chain<LNL, P, L - 1, PHI6, PF>(u, lds, twi, base, tt, nxt, M);
	}
}

// Final pass then the inverse passes down to layout 0 (F: final-pass roots
// loaded by fwd_all when PF).
template <int LNL, int P, bool PHI6, bool MUL, bool PF>
__device__ __forceinline__ void final_inv(uint32_t (&u)[E], uint32_t (&vv)[E], uint32_t *lds, const tw_entry *twf,
                                          const tw_entry *twi, uint32_t base, uint32_t tt, const ftw<LNL> &F)
{
	constexpr int a = geo<LNL>::a;
	mtw M;
	if constexpr (PHI6) M = load_mtw(twi);
	ptw c;
	if constexpr (PF)
	{
		load_ptw<LNL, a - 1>(c, twi, base, tt);
		final_pass<LNL, P, MUL>(u, vv, F);
	}
	else
	{
		ftw_lazy<LNL> FL; FL.twf = twf; FL.twi = twi; FL.base = base; FL.k0 = E * tt;
		final_pass<LNL, P, MUL>(u, vv, FL);
	}
	inv_chain<LNL, P, a - 1, PHI6, PF>(u, lds, twi, base, tt, c, M);
}

// ---- primes 0 and 1 of one instance, pass by pass ---------------------------
//
// The two-prime squaring runs both primes' passes in turn at every level
// instead of one whole transform after the other: prime 0 exchanges through
// l0 and prime 1 through l1, the two share each workgroup barrier, and the LDS
// traffic of one prime's exchange is in flight while the other prime's pass
// computes.  Every element sees the same butterflies as in the single-prime
// chains (roots read where they are used), so the residues are identical.
//
// A workgroup-wide exchange (the first pass's, when the instance spans more
// than one wavefront) is skewed between the primes: prime 0's plane is
// written, prime 1's pass runs, and after the barrier prime 0 is read back
// and prime 1's plane written while prime 0's next pass runs.  Every barrier
// then has a pass of work in front of it instead of both primes' LDS traffic
// (n = 2^14: 30.6 -> 29.5 us per squaring step of the transforms).

template <int LNL, int LAY>
__device__ __forceinline__ void write_plane(const uint32_t (&u)[E], uint32_t *l, uint32_t tt)
{
	uint32_t *w = l + lbase<LNL, LAY>(tt);
#pragma unroll
	for (int q = 0; q < E; ++q) w[loff<LNL, LAY>(q)] = u[q];
}

template <int LNL, int LAY>
__device__ __forceinline__ void read_plane(uint32_t (&u)[E], const uint32_t *l, uint32_t tt)
{
	const uint32_t *r = l + lbase<LNL, LAY>(tt);
#pragma unroll
	for (int q = 0; q < E; ++q) u[q] = r[loff<LNL, LAY>(q)];
}

template <int LNL, int FROM, int TO>
__device__ __forceinline__ void xchg2(uint32_t (&u0)[E], uint32_t (&u1)[E], uint32_t *l0, uint32_t *l1, uint32_t tt)
{
	uint32_t *w0 = l0 + lbase<LNL, FROM>(tt), *w1 = l1 + lbase<LNL, FROM>(tt);
	uint32_t *r0 = l0 + lbase<LNL, TO>(tt), *r1 = l1 + lbase<LNL, TO>(tt);
	constexpr bool local = wave_local<LNL, FROM>() && wave_local<LNL, TO>();
	if constexpr (local) __builtin_amdgcn_wave_barrier();
	else __syncthreads();
#pragma unroll
	for (int q = 0; q < E; ++q) w0[loff<LNL, FROM>(q)] = u0[q];
#pragma unroll
	for (int q = 0; q < E; ++q) w1[loff<LNL, FROM>(q)] = u1[q];
	if constexpr (local) __builtin_amdgcn_wave_barrier();
	else __syncthreads();
#pragma unroll
	for (int q = 0; q < E; ++q) u0[q] = r0[loff<LNL, TO>(q)];
#pragma unroll
	for (int q = 0; q < E; ++q) u1[q] = r1[loff<LNL, TO>(q)];
	if constexpr (local) __builtin_amdgcn_wave_barrier();
}

template <int LNL, int L, bool PHI6>
__device__ __forceinline__ void fwd_chain2(uint32_t (&u0)[E], uint32_t (&u1)[E], uint32_t *l0, uint32_t *l1,
                                           const tw_entry *twf0, const tw_entry *twf1, uint32_t base, uint32_t tt)
{
	if constexpr (L < geo<LNL>::a)
	{
		fwd_pass<0, PHI6 && L == 0>(u0, lazy_ptw<LNL, L>(twf0, base, tt));
		fwd_pass<1, PHI6 && L == 0>(u1, lazy_ptw<LNL, L>(twf1, base, tt));
		xchg2<LNL, L, L + 1>(u0, u1, l0, l1, tt);
		fwd_chain2<LNL, L + 1, PHI6>(u0, u1, l0, l1, twf0, twf1, base, tt);
	}
}

// inverse passes L down to STOP + 1, each preceded by the exchange from layout L+1
template <int LNL, int L, bool PHI6, int STOP = -1>
__device__ __forceinline__ void inv_chain2(uint32_t (&u0)[E], uint32_t (&u1)[E], uint32_t *l0, uint32_t *l1,
                                           const tw_entry *twi0, const tw_entry *twi1, uint32_t base, uint32_t tt,
                                           const mtw &M0, const mtw &M1)
{
	if constexpr (L > STOP)
	{
		xchg2<LNL, L + 1, L>(u0, u1, l0, l1, tt);
		inv_pass<0, PHI6 && L == 0>(u0, lazy_ptw<LNL, L>(twi0, base, tt), M0);
		inv_pass<1, PHI6 && L == 0>(u1, lazy_ptw<LNL, L>(twi1, base, tt), M1);
		inv_chain2<LNL, L - 1, PHI6, STOP>(u0, u1, l0, l1, twi0, twi1, base, tt, M0, M1);
	}
}

// The first exchange is workgroup-wide and every later one wavefront-local.
template <int LNL>
__device__ __forceinline__ constexpr bool skew_first()
{
	return geo<LNL>::a >= 2 && !wave_local<LNL, 0>() && wave_local<LNL, 1>();
}

// Forward passes, final pass (pointwise square) and inverse passes of primes 0
// and 1 (tw: the table of prime 0; prime 1's follows at tw + 2 n).
template <int LNL, bool PHI6>
__device__ __forceinline__ void square2(uint32_t (&u0)[E], uint32_t (&u1)[E], uint32_t *l0, uint32_t *l1,
                                        const tw_entry *tw, uint32_t base, uint32_t tt)
{
	constexpr uint32_t n = geo<LNL>::n;
	const tw_entry *twf0 = tw, *twi0 = tw + n, *twf1 = tw + 2 * n, *twi1 = tw + 3 * n;
	if constexpr (skew_first<LNL>())
	{   // layout 0 -> 1 skewed; the barrier ahead of prime 0's write orders it
		// after every earlier access of l0 by the workgroup
		fwd_pass<0, PHI6>(u0, lazy_ptw<LNL, 0>(twf0, base, tt));
		__syncthreads();
		write_plane<LNL, 0>(u0, l0, tt);
		fwd_pass<1, PHI6>(u1, lazy_ptw<LNL, 0>(twf1, base, tt));
		__syncthreads();
		read_plane<LNL, 1>(u0, l0, tt);
		write_plane<LNL, 0>(u1, l1, tt);
		fwd_pass<0, false>(u0, lazy_ptw<LNL, 1>(twf0, base, tt));
		__syncthreads();
		read_plane<LNL, 1>(u1, l1, tt);
		fwd_pass<1, false>(u1, lazy_ptw<LNL, 1>(twf1, base, tt));
		xchg2<LNL, 1, 2>(u0, u1, l0, l1, tt);
		fwd_chain2<LNL, 2, PHI6>(u0, u1, l0, l1, twf0, twf1, base, tt);
	}
	else fwd_chain2<LNL, 0, PHI6>(u0, u1, l0, l1, twf0, twf1, base, tt);
	mtw M0, M1;
	if constexpr (PHI6) { M0 = load_mtw(twi0); M1 = load_mtw(twi1); }
	uint32_t none[E];
#pragma unroll
	for (int q = 0; q < E; ++q) none[q] = 0;
	ftw_lazy<LNL> F0; F0.twf = twf0; F0.twi = twi0; F0.base = base; F0.k0 = E * tt;
	ftw_lazy<LNL> F1; F1.twf = twf1; F1.twi = twi1; F1.base = base; F1.k0 = E * tt;
	final_pass<LNL, 0, false>(u0, none, F0);
	final_pass<LNL, 1, false>(u1, none, F1);
	if constexpr (skew_first<LNL>())
	{   // layout 1 -> 0 skewed: the layout-1 writes land in each wavefront's
		// own range, which only that wavefront has touched since the last
		// workgroup barrier
		inv_chain2<LNL, geo<LNL>::a - 1, PHI6, 1>(u0, u1, l0, l1, twi0, twi1, base, tt, M0, M1);
		xchg2<LNL, 2, 1>(u0, u1, l0, l1, tt);
		inv_pass<0, false>(u0, lazy_ptw<LNL, 1>(twi0, base, tt), M0);
		write_plane<LNL, 1>(u0, l0, tt);
		inv_pass<1, false>(u1, lazy_ptw<LNL, 1>(twi1, base, tt), M1);
		__syncthreads();
		read_plane<LNL, 0>(u0, l0, tt);
		write_plane<LNL, 1>(u1, l1, tt);
		inv_pass<0, PHI6>(u0, lazy_ptw<LNL, 0>(twi0, base, tt), M0);
		__syncthreads();
		read_plane<LNL, 0>(u1, l1, tt);
		inv_pass<1, PHI6>(u1, lazy_ptw<LNL, 0>(twi1, base, tt), M1);
	}
	else inv_chain2<LNL, geo<LNL>::a - 1, PHI6>(u0, u1, l0, l1, twi0, twi1, base, tt, M0, M1);
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
// dup: multiply the chunk by the PRP base (2 or 3; base_m1 = base - 1) as it is
// carried, x^2 * base^bit (kernel.cl:1044 doubles; base 3 adds l twice).
__device__ __forceinline__ int64_t carry_chunk64(const int64_t (&l)[E], int32_t (&dig)[E], bool dup, int base_m1,
                                                 const divisor &dv)
{
	// the coefficient times m = 1, 2 or 3 (m is uniform over a candidate's
	// threads): m l = m lo(l) + 2^32 m hi(l) mod 2^64, one mad_u64_u32 and one
	// mul_lo instead of per-lane selects
	const uint32_t m = dup ? (uint32_t)(base_m1 + 1) : 1u;
	int64_t acc = 0;
#pragma unroll
	for (int q = 0; q < E; ++q)
	{
		const uint64_t lu = (uint64_t)l[q];
		const uint64_t ml = (uint64_t)(uint32_t)lu * m + ((uint64_t)((uint32_t)(lu >> 32) * m) << 32);
		acc += (int64_t)ml;
		dig[q] = divrem_i62(acc, dv);               // |acc| < 2^61 (the two-prime bound times 3)
	}
	return acc;
}

// normalize1 on one 16-coefficient chunk (kernel.cl:1029-1051): returns the
// chunk's outgoing carry.
__device__ __forceinline__ int64_t carry_chunk(const i96 (&l)[E], int32_t (&dig)[E], bool dup, int base_m1,
                                               const divisor &dv)
{
	i96 acc; acc.lo = 0; acc.hi = 0;
#pragma unroll
	for (int q = 0; q < E; ++q)
	{
		acc = i96_add(acc, l[q]);
		if (dup) { acc = i96_add(acc, l[q]); if (base_m1 == 2) acc = i96_add(acc, l[q]); }
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
