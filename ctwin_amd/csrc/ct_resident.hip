// Resident squaring kernels for n = 2^LN <= 2^14 digits: one workgroup owns one
// candidate for many consecutive squarings and keeps the whole state on chip.
//
// Each of the T = n/16 threads holds 16 elements per prime in VGPRs.  One
// squaring (the work of engine::squareDup, cyclo/src/engine.h:387-419) is, per
// prime: radix-16 register passes over the levels 1..4a of the DIF tree
// (a = floor((LN-1)/4)) with LDS transposes between them, a final register
// pass over the last (LN-1) mod 4 levels fused with the pointwise square (the
// sqr2 form of kernel.cl:189-194), and the mirrored inverse passes ending in
// the level-1 merge.  Then Garner CRT and the signed base-b carry in 16-digit
// chunks (normalize1/normalize2, kernel.cl:1029-1081) run on consecutive
// digits after one more LDS transpose, and the carried digits return to the
// transform layout.  Nothing but the exponent bits is read from HBM inside
// the step loop.
//
// Layouts (t = thread, q = slot 0..15):
//   pass L < a : k = j*m + i + q*m/16, m = n/16^L, i = t mod m/16, j = t div m/16
//   L == a     : k = 16 t + q  (final pass, and the carry chunks)
// Layout 0 (k = t + q*T) is also the layout of the digits in registers and in
// HBM between launches, so loads and stores are coalesced.
//
// MINUS (N- = b^n - b^(n/2) - 1, n = 2^(LN-1) digits): the transform has length
// 2n; the digits fill k < n (slots q < 8 of layout 0) and the rest stays zero,
// so the transform's product is the exact acyclic product c_0 .. c_{2n-2}.  It
// is folded by X^n = X^(n/2) + 1 in residue space inside each thread:
//   j < n/2:      d_j = c_j + c_{j+n} + c_{j+3n/2}
//   n/2 <= j < n: d_j = c_j + c_{j+n/2} + 2 c_{j+n}
// (k + n/2, k + n, k + 3n/2 are slots q + 4, q + 8, q + 12), then the CRT and
// the carry run on the n digits (threads t < T/2), the top carry wrapping to
// chunk 0 and chunk n/32 with a plus sign.
#include <hip/hip_runtime.h>

#include "ct_arith.h"
#include "ct_tables.h"
#include "ct_ntt.h"
#include "ct_kernels.h"

namespace ct {

// x <- x^2 (or x*y) mod (X^n - X^(n/2) + 1, p_P) -- or, MINUS, the folded
// product mod X^(n/2) - X^(n/4) - 1 in slots q < 8; the signed digits are read
// from LDS (layout 0, region D), the output residues (exact coefficients mod
// p, in [0, p)) are left in registers in layout 0.
template <int LN, int P, int NP, bool MUL, bool MINUS>
__device__ __forceinline__ void transform(const int32_t *D, const int32_t (&yd)[E], uint32_t (&out)[E],
                                          uint32_t *lds, const tw_entry *tw, uint32_t t)
{
	constexpr uint32_t p = kP[P], n = geo<LN>::n;
	const tw_entry *twf = tw + (size_t)(P * 2 + 0) * n, *twi = tw + (size_t)(P * 2 + 1) * n;
	uint32_t u[E], v[E];
	const int32_t *dr = D + lbase<LN, 0>(t);
#pragma unroll
	for (int k = 0; k < E; ++k) u[k] = from_signed_np<NP>(dr[loff<LN, 0>(k)], p);
	ftw<LN> F;
	// no leading barrier: the caller's last barrier follows every access of
	// lds by other threads, and an earlier prime's transform ends by reading
	// this thread's own layout-0 words
	fwd_all<LN, P, true, false, false>(u, lds, twf, twi, 1u, t, F);
	if (MUL)
	{
#pragma unroll
		for (int k = 0; k < E; ++k) v[k] = from_signed_np<NP>(yd[k], p);
		fwd_all<LN, P, true, false>(v, lds, twf, twi, 1u, t, F);
	}
	else
	{
#pragma unroll
		for (int k = 0; k < E; ++k) v[k] = 0;
	}
	final_inv<LN, P, true, MUL, false>(u, v, lds, twf, twi, 1u, t, F);
	if constexpr (MINUS)
	{
#pragma unroll
		for (int q = 0; q < 4; ++q) out[q] = red1p(red1p(u[q] + u[q + 8], p) + u[q + 12], p);
#pragma unroll
		for (int q = 4; q < 8; ++q) out[q] = red1p(red1p(u[q] + u[q + 4], p) + red1p(u[q + 8] + u[q + 8], p), p);
#pragma unroll
		for (int q = 8; q < E; ++q) out[q] = 0;
	}
	else
	{
#pragma unroll
		for (int k = 0; k < E; ++k) out[k] = u[k];
	}
}

// Two-prime squaring of the digits in D (layout 0): both primes' passes in
// turn at every level (square2), prime 1 exchanging through D itself, which is
// free once every thread has read its digits (each thread reads only its own
// layout-0 words, and the first exchange into D starts with a barrier, or is
// wave-local and reaches only the wave's own words).  Residues out as in
// transform (layout 0; MINUS: folded).
template <int LN, bool MINUS>
__device__ __forceinline__ void transform2(int32_t *D, uint32_t (&out0)[E], uint32_t (&out1)[E], uint32_t *lds,
                                           const tw_entry *tw, uint32_t t)
{
	constexpr uint32_t p0 = kP[0], p1 = kP[1];
	const int32_t *dr = D + lbase<LN, 0>(t);
#pragma unroll
	for (int k = 0; k < E; ++k)
	{
		const int32_t d = dr[loff<LN, 0>(k)];
		out0[k] = from_signed_np<2>(d, p0);
		out1[k] = from_signed_np<2>(d, p1);
	}
	square2<LN, true>(out0, out1, lds, reinterpret_cast<uint32_t *>(D), tw, 1u, t);
	if constexpr (MINUS)
	{
		uint32_t f0[E], f1[E];
#pragma unroll
		for (int q = 0; q < 4; ++q)
		{
			f0[q] = red1p(red1p(out0[q] + out0[q + 8], p0) + out0[q + 12], p0);
			f1[q] = red1p(red1p(out1[q] + out1[q + 8], p1) + out1[q + 12], p1);
		}
#pragma unroll
		for (int q = 4; q < 8; ++q)
		{
			f0[q] = red1p(red1p(out0[q] + out0[q + 4], p0) + red1p(out0[q + 8] + out0[q + 8], p0), p0);
			f1[q] = red1p(red1p(out1[q] + out1[q + 4], p1) + red1p(out1[q + 8] + out1[q + 8], p1), p1);
		}
#pragma unroll
		for (int q = 0; q < E; ++q)
		{
			out0[q] = (q < 8) ? f0[q] : 0u;
			out1[q] = (q < 8) ? f1[q] : 0u;
		}
	}
}

// One squaring (MUL = false, x <- x^2 * base^dup) or product (MUL = true,
// x <- x * y) with CRT and the chunked carry, over NP primes (NP = 2 when every
// coefficient is known to be below p1 p2 / 2, else 3).  x lives in LDS region
// D (layout 0) before and after; y (MUL only) in registers, layout 0.
template <int LN, bool MUL, int NP, bool MINUS>
__device__ __forceinline__ void step(int32_t *D, const int32_t (&yd)[E], bool dup, int base_m1, const divisor &dv,
                                     uint32_t *lds, int64_t *fbuf, const tw_entry *tw, const garner_consts &gc, uint32_t t)
{
	constexpr int a = geo<LN>::a;
	constexpr uint32_t T = geo<LN>::T;
	constexpr uint32_t TC = MINUS ? T / 2 : T;          // threads that own a chunk of digits
	uint32_t r0[E], r1[E], r2[E];
	if constexpr (NP == 2 && !MUL) transform2<LN, MINUS>(D, r0, r1, lds, tw, t);
	else
	{
		transform<LN, 0, NP, MUL, MINUS>(D, yd, r0, lds, tw, t);
		transform<LN, 1, NP, MUL, MINUS>(D, yd, r1, lds, tw, t);
		if (NP == 3) transform<LN, 2, NP, MUL, MINUS>(D, yd, r2, lds, tw, t);
	}
	// to the carry layout: primes 0 and 1 share one transpose, prime 1 through
	// D (every read of the digits came before the first barrier of prime 1's
	// transform; each thread later writes its carried digits to exactly the
	// words of D it reads here).  No barrier ahead of the writes: every
	// transform ends by reading the thread's own layout-0 words of lds and D,
	// the words it writes here, which no other thread has touched since the
	// transform's last workgroup barrier.
	{
		uint32_t *Du = reinterpret_cast<uint32_t *>(D);
#pragma unroll
		for (int q = 0; q < E; ++q)
		{
			lds[lbase<LN, 0>(t) + loff<LN, 0>(q)] = r0[q];
			Du[lbase<LN, 0>(t) + loff<LN, 0>(q)] = r1[q];
		}
		__syncthreads();
#pragma unroll
		for (int q = 0; q < E; ++q)
		{
			r0[q] = lds[lbase<LN, a>(t) + loff<LN, a>(q)];
			r1[q] = Du[lbase<LN, a>(t) + loff<LN, a>(q)];
		}
	}
	if (NP == 3) xchg<LN, 0, a>(r2, lds, t);

	// normalize1: CRT, optional multiplication by the base, signed digits of 16
	// consecutive coefficients
	int32_t dig[E];
	if (t < TC)
	{
		if (NP == 3)
		{
			i96 l[E];
#pragma unroll
			for (int q = 0; q < E; ++q) l[q] = garner(r0[q], r1[q], r2[q], gc);
			fbuf[t] = carry_chunk(l, dig, dup, base_m1, dv);
		}
		else
		{
			int64_t l[E];
#pragma unroll
			for (int q = 0; q < E; ++q) l[q] = garner2(r0[q], r1[q], gc);
			fbuf[t] = carry_chunk64(l, dig, dup, base_m1, dv);
		}
	}
	__syncthreads();
	// normalize2: carry in from the previous chunk; X^n = X^(n/2) - 1 (N+) wraps
	// the top carry to chunk 0 negated and adds it to chunk TC/2; for N-,
	// X^n = X^(n/2) + 1 adds it to both
	if (t < TC)
	{
		int64_t c = fbuf[(t + TC - 1) % TC];
		if (t == 0) c = MINUS ? c : -c;
		else if (t == TC / 2) c += fbuf[TC - 1];
		carry_in(dig, c, dv);
		// the carried chunk goes straight back to D (every reader of D in this
		// step has passed the barriers above)
		int32_t *dw = D + lbase<LN, a>(t);
#pragma unroll
		for (int q = 0; q < E; ++q) dw[loff<LN, a>(q)] = dig[q];
	}
	__syncthreads();
}

// Two-prime squaring with the primes in the two halves of a workgroup of 2T
// threads (thread tf: prime tf / T, role t = tf % T): each half runs its
// prime's transform with its own LDS region (both halves pass the same
// barriers, the halves being whole wavefronts), writes its residues to that
// region in layout 0, and the first half runs the CRT and the chunked carry
// on both primes' residues as step does.  Twice the wavefronts per candidate
// for n = 2^10 .. 2^13, the same digits.
template <int LN>
struct res_halves
{
	static constexpr bool ok = (geo<LN>::T % 64 == 0) && (2 * geo<LN>::T <= 1024);
};

template <int LN, bool MINUS>
__device__ __forceinline__ void step_h(int32_t *D, uint32_t P, bool dup, int base_m1, const divisor &dv,
                                       uint32_t (*lds)[geo<LN>::lds_words], int64_t *fbuf, const tw_entry *tw,
                                       const garner_consts &gc, uint32_t t)
{
	constexpr int a = geo<LN>::a;
	constexpr uint32_t T = geo<LN>::T;
	constexpr uint32_t TC = MINUS ? T / 2 : T;
	int32_t none[E];
#pragma unroll
	for (int q = 0; q < E; ++q) none[q] = 0;
	uint32_t r[E];
	if (P == 0) transform<LN, 0, 2, false, MINUS>(D, none, r, lds[0], tw, t);
	else transform<LN, 1, 2, false, MINUS>(D, none, r, lds[1], tw, t);
	// both primes' residues, layout 0, each in its own region (no barrier
	// ahead: the transform ended by reading this thread's own layout-0 words
	// of its region, the words written here)
	{
		uint32_t *w = lds[P] + lbase<LN, 0>(t);
#pragma unroll
		for (int q = 0; q < E; ++q) w[loff<LN, 0>(q)] = r[q];
	}
	__syncthreads();
	int32_t dig[E];
	if (P == 0 && t < TC)
	{
		uint32_t r0[E], r1[E];
		const uint32_t *s0 = lds[0] + lbase<LN, a>(t), *s1 = lds[1] + lbase<LN, a>(t);
#pragma unroll
		for (int q = 0; q < E; ++q) { r0[q] = s0[loff<LN, a>(q)]; r1[q] = s1[loff<LN, a>(q)]; }
		int64_t l[E];
#pragma unroll
		for (int q = 0; q < E; ++q) l[q] = garner2(r0[q], r1[q], gc);
		fbuf[t] = carry_chunk64(l, dig, dup, base_m1, dv);
	}
	__syncthreads();
	if (P == 0 && t < TC)
	{
		int64_t c = fbuf[(t + TC - 1) % TC];
		if (t == 0) c = MINUS ? c : -c;
		else if (t == TC / 2) c += fbuf[TC - 1];
		carry_in(dig, c, dv);
		int32_t *dw = D + lbase<LN, a>(t);
#pragma unroll
		for (int q = 0; q < E; ++q) dw[loff<LN, a>(q)] = dig[q];
	}
	__syncthreads();
}

// ---- kernels -----------------------------------------------------------------

template <int LN, bool MINUS>
__global__ __launch_bounds__(2 * geo<LN>::T) void k_res_steps_h(res_args A)
{
	static_assert(res_halves<LN>::ok, "halves of whole wavefronts, at most 1024 threads");
	constexpr uint32_t T = geo<LN>::T, n = MINUS ? geo<LN>::n / 2 : geo<LN>::n;
	constexpr int QD = MINUS ? E / 2 : E;
	__shared__ uint32_t lds[2][geo<LN>::lds_words];
	__shared__ int32_t D[geo<LN>::lds_words];
	__shared__ int64_t fbuf[T];
	const uint32_t c = blockIdx.x, tf = threadIdx.x, P = tf / T, t = tf % T;
	const int i_top = A.lane_i ? A.lane_i[c] : A.i_hi;
	if (A.lane_i && i_top < 0) return;          // idle lane (uniform per workgroup)
	int32_t *x = A.x + (size_t)c * n;
	if (P == 0)
	{
		int32_t *dw = D + lbase<LN, 0>(t);
#pragma unroll
		for (int q = 0; q < E; ++q) dw[loff<LN, 0>(q)] = (q < QD) ? x[t + q * T] : 0;
	}
	__syncthreads();
	const divisor dv = A.dv[c];
	const uint32_t *bits = A.bits ? A.bits + (size_t)c * A.bit_words : nullptr;
	const bool mdup = A.mask ? ((A.mask[c >> 5] >> (c & 31)) & 1u) : false;
	for (int it = 0; it < A.nsteps; ++it)
	{
		const int i = i_top - it;
		const bool dup = bits ? ((bits[i >> 5] >> (i & 31)) & 1u) : mdup;
		step_h<LN, MINUS>(D, P, dup, A.base_m1, dv, lds, fbuf, A.tw, A.gc, t);
	}
	if (P == 0)
	{
		const int32_t *dr = D + lbase<LN, 0>(t);
#pragma unroll
		for (int q = 0; q < QD; ++q) x[t + q * T] = dr[loff<LN, 0>(q)];
	}
}

template <int LN, int NP, bool MINUS>
__global__ __launch_bounds__(geo<LN>::T) void k_res_steps(res_args A)
{
	constexpr uint32_t T = geo<LN>::T, n = MINUS ? geo<LN>::n / 2 : geo<LN>::n;    // digits per candidate
	constexpr int QD = MINUS ? E / 2 : E;                                           // layout-0 slots holding digits
	__shared__ uint32_t lds[geo<LN>::lds_words];
	__shared__ int32_t D[geo<LN>::lds_words];
	__shared__ int64_t fbuf[T];
	const uint32_t c = blockIdx.x, t = threadIdx.x;
	const int i_top = A.lane_i ? A.lane_i[c] : A.i_hi;
	if (A.lane_i && i_top < 0) return;          // idle lane (uniform per workgroup)
	int32_t *x = A.x + (size_t)c * n;
	{
		int32_t *dw = D + lbase<LN, 0>(t);
#pragma unroll
		for (int q = 0; q < E; ++q) dw[loff<LN, 0>(q)] = (q < QD) ? x[t + q * T] : 0;
	}
	__syncthreads();
	const divisor dv = A.dv[c];
	const uint32_t *bits = A.bits ? A.bits + (size_t)c * A.bit_words : nullptr;
	const bool mdup = A.mask ? ((A.mask[c >> 5] >> (c & 31)) & 1u) : false;
	int32_t none[E];
#pragma unroll
	for (int q = 0; q < E; ++q) none[q] = 0;
	for (int it = 0; it < A.nsteps; ++it)
	{
		const int i = i_top - it;
		const bool dup = bits ? ((bits[i >> 5] >> (i & 31)) & 1u) : mdup;
		step<LN, false, NP, MINUS>(D, none, dup, A.base_m1, dv, lds, fbuf, A.tw, A.gc, t);
	}
	const int32_t *dr = D + lbase<LN, 0>(t);
#pragma unroll
	for (int q = 0; q < QD; ++q) x[t + q * T] = dr[loff<LN, 0>(q)];
}

template <int LN, int NP, bool MINUS>
__global__ __launch_bounds__(geo<LN>::T) void k_res_mul(mul_args A)
{
	constexpr uint32_t T = geo<LN>::T, n = MINUS ? geo<LN>::n / 2 : geo<LN>::n;
	constexpr int QD = MINUS ? E / 2 : E;
	__shared__ uint32_t lds[geo<LN>::lds_words];
	__shared__ int32_t D[geo<LN>::lds_words];
	__shared__ int64_t fbuf[T];
	const uint32_t c = blockIdx.x, t = threadIdx.x;
	const int32_t *xa = A.a + (size_t)c * n, *xb = A.b + (size_t)c * n;
	int32_t yd[E];
	{
		int32_t *dw = D + lbase<LN, 0>(t);
#pragma unroll
		for (int q = 0; q < E; ++q)
		{
			dw[loff<LN, 0>(q)] = (q < QD) ? xa[t + q * T] : 0;
			yd[q] = (q < QD) ? xb[t + q * T] : 0;
		}
	}
	__syncthreads();
	const divisor dv = A.dv[c];
	step<LN, true, NP, MINUS>(D, yd, false, 1, dv, lds, fbuf, A.tw, A.gc, t);
	int32_t *xo = A.out + (size_t)c * n;
	const int32_t *dr = D + lbase<LN, 0>(t);
#pragma unroll
	for (int q = 0; q < QD; ++q) xo[t + q * T] = dr[loff<LN, 0>(q)];
}

template <int LN>
static hipError_t launch_steps_ln(int np, bool minus, bool halves, const res_args &A, uint32_t lanes, hipStream_t st)
{
	const dim3 g(lanes), b(geo<LN>::T);
	if constexpr (res_halves<LN>::ok)
		if (halves && np == 2)
		{
			const dim3 bh(2 * geo<LN>::T);
			if (minus) hipLaunchKernelGGL((k_res_steps_h<LN, true>), g, bh, 0, st, A);
			else hipLaunchKernelGGL((k_res_steps_h<LN, false>), g, bh, 0, st, A);
			return hipGetLastError();
		}
	if (minus)
	{
		if (np == 2) hipLaunchKernelGGL((k_res_steps<LN, 2, true>), g, b, 0, st, A);
		else hipLaunchKernelGGL((k_res_steps<LN, 3, true>), g, b, 0, st, A);
	}
	else if (np == 2) hipLaunchKernelGGL((k_res_steps<LN, 2, false>), g, b, 0, st, A);
	else hipLaunchKernelGGL((k_res_steps<LN, 3, false>), g, b, 0, st, A);
	return hipGetLastError();
}

template <int LN>
static hipError_t launch_mul_ln(int np, bool minus, const mul_args &A, uint32_t lanes, hipStream_t st)
{
	const dim3 g(lanes), b(geo<LN>::T);
	if (minus)
	{
		if (np == 2) hipLaunchKernelGGL((k_res_mul<LN, 2, true>), g, b, 0, st, A);
		else hipLaunchKernelGGL((k_res_mul<LN, 3, true>), g, b, 0, st, A);
	}
	else if (np == 2) hipLaunchKernelGGL((k_res_mul<LN, 2, false>), g, b, 0, st, A);
	else hipLaunchKernelGGL((k_res_mul<LN, 3, false>), g, b, 0, st, A);
	return hipGetLastError();
}

// tln: log2 of the transform length (n digits for N+, 2n for N-); halves:
// two primes in the halves of one workgroup where the halves are whole
// wavefronts (resident_halves_ok), else the one-prime-after-the-other kernel
bool resident_halves_ok(int tln) { return tln >= 10 && tln <= 13; }

hipError_t resident_steps(int tln, int np, bool minus, bool halves, const res_args &A, uint32_t lanes, hipStream_t st)
{
	switch (tln)
	{
	case 5: return launch_steps_ln<5>(np, minus, halves, A, lanes, st);
	case 6: return launch_steps_ln<6>(np, minus, halves, A, lanes, st);
	case 7: return launch_steps_ln<7>(np, minus, halves, A, lanes, st);
	case 8: return launch_steps_ln<8>(np, minus, halves, A, lanes, st);
	case 9: return launch_steps_ln<9>(np, minus, halves, A, lanes, st);
	case 10: return launch_steps_ln<10>(np, minus, halves, A, lanes, st);
	case 11: return launch_steps_ln<11>(np, minus, halves, A, lanes, st);
	case 12: return launch_steps_ln<12>(np, minus, halves, A, lanes, st);
	case 13: return launch_steps_ln<13>(np, minus, halves, A, lanes, st);
	case 14: return launch_steps_ln<14>(np, minus, halves, A, lanes, st);
	default: return hipErrorInvalidValue;
	}
}

hipError_t resident_mul(int tln, int np, bool minus, const mul_args &A, uint32_t lanes, hipStream_t st)
{
	switch (tln)
	{
	case 5: return launch_mul_ln<5>(np, minus, A, lanes, st);
	case 6: return launch_mul_ln<6>(np, minus, A, lanes, st);
	case 7: return launch_mul_ln<7>(np, minus, A, lanes, st);
	case 8: return launch_mul_ln<8>(np, minus, A, lanes, st);
	case 9: return launch_mul_ln<9>(np, minus, A, lanes, st);
	case 10: return launch_mul_ln<10>(np, minus, A, lanes, st);
	case 11: return launch_mul_ln<11>(np, minus, A, lanes, st);
	case 12: return launch_mul_ln<12>(np, minus, A, lanes, st);
	case 13: return launch_mul_ln<13>(np, minus, A, lanes, st);
	case 14: return launch_mul_ln<14>(np, minus, A, lanes, st);
	default: return hipErrorInvalidValue;
	}
}

}   // namespace ct
