// Plain multipass squaring engine: one squaring is three launches over the
// whole GPU, for n = 2^LN digits with 10 <= LN <= 22 (one candidate per launch
// in the "plain" configuration; any number of lanes share each launch).
//
// n = R x C (R = 2^RL rows, C = 2^CL columns, k = r C + c).  The first RL
// levels of the DIF tree pair elements of one column; the remaining levels
// stay inside one row.  So a squaring is
//   A  column pass: load 16-digit chunks of W adjacent columns, apply the
//      pending carries of the previous squaring (normalize2, kernel.cl:1053),
//      then per prime the RL column levels (radix-16 passes, LDS transposes)
//      -> three residue planes in HBM;
//   B  row pass (one workgroup per row and prime): levels RL+1..LN-1, the
//      pointwise square (sqr2) and the inverse row levels, in place;
//   C  column pass: inverse column levels and the Phi_6 merge per prime,
//      Garner CRT and normalize1 on 16-digit chunks -> digits + chunk carries.
// HBM traffic per squaring and digit: A 4+0.5 read, 12 write; B 12 + 12;
// C 12 read, 4+0.5 write = 57 bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/ctwin.h"
#include "ct_engine_impl.h"
#include "ct_ntt.h"

namespace ct {

struct plain_args
{
	int32_t *x;                 // digits of the source / destination register
	const int32_t *xb;          // second operand digits (product only)
	uint32_t *res;              // [3][lanes][n] residue planes
	uint32_t *res2;             // second operand planes (product only)
	int64_t *carry;             // [lanes][n/16] chunk carries, chunk g at cslot(g)
	const int64_t *carry_in;    // pending carries to apply on load (or null)
	const tw_entry *tw;
	const divisor *dv;
	const uint32_t *bits;       // exponent bits or null
	const uint32_t *mask;       // dup bits when bits is null (or null)
	uint32_t bit_words;
	const int32_t *run;         // device {i0, iter}: top bit index of the run, squarings done by earlier graph replays
	const int32_t *lane_i;      // per-lane top bit index (or null, then i0); negative = idle lane
	int step;                   // squaring index within this launch sequence: bit = top - (iter + step)
	int base_m1;                // PRP base - 1 (a set bit multiplies by the base)
	int split_a;                // pass A: one workgroup per column group and prime
	int row_form;               // pass B: 0 per (row, prime), 1 per row, primes in sequence, 2 per row, primes in slices
	uint32_t lanes;
	garner_consts gc;
};

// Column-pass geometry: W_ adjacent columns per workgroup (the LDS tile;
// 64-byte row segments for W = 16, 128-byte ones for W = 32).  The engine's
// autotuner picks W per transform length.
template <int LN, int RL, int W_>
struct colgeo
{
	static constexpr uint32_t R = 1u << RL;
	static constexpr uint32_t Tc = R / 16;                 // threads per column
	static constexpr uint32_t W = W_;                      // columns per workgroup
	static constexpr uint32_t threads = W * Tc;            // = chunks per workgroup (R W / 16)
	// padded column stride in LDS: a column's rows r sit at r + r/32 (the
	// layout of the in-column transposes); the odd stride puts the W columns
	// that a wavefront touches at once in distinct banks
	static constexpr uint32_t RP = R + (R >> 5) + 1;
	static constexpr int passes = RL / 4;
};

// Workgroup b -> column group, XCD-aware.  Consecutive workgroups go to the
// 8 XCDs in turn; with a tile of 16 columns (64-byte row segments) two
// neighbouring groups share every 128-byte line, so neighbouring groups are
// given to workgroups of one XCD: b -> (b mod 8) total/8 + b div 8 (a
// bijection when 8 divides total) hands each XCD a contiguous range of
// columns, whose lines its L2 then fetches once.
__device__ __forceinline__ uint32_t xcd_group(uint32_t b, uint32_t total)
{
	return (total % 8u == 0u) ? (b % 8u) * (total / 8u) + b / 8u : b;
}

// column forward passes 0..passes-1 (transposes between them only)
// Slot of chunk g (16 digits; row g / cpr, chunk g % cpr of the row, cpr =
// chunks per row) in a lane's carry array: chunk-column major, so that the
// column passes, whose threads hold one chunk per row, touch consecutive slots.
__host__ __device__ __forceinline__ uint32_t cslot(uint32_t g, uint32_t cpr, uint32_t rows)
{
	return (g % cpr) * rows + g / cpr;
}

template <int RL, int P, int L>
__device__ __forceinline__ void col_fwd(uint32_t (&u)[E], uint32_t *X, const tw_entry *twf, uint32_t tt, const ptw &cur)
{
	ptw nxt;
	if constexpr (L + 1 < RL / 4) load_ptw<RL, L + 1>(nxt, twf, 1u, tt);
	fwd_pass<P, L == 0>(u, cur);
	if constexpr (L + 1 < RL / 4)
	{
		xchg<RL, L, L + 1>(u, X, tt);
		col_fwd<RL, P, L + 1>(u, X, twf, tt, nxt);
	}
}

template <int RL, int P, int L>
__device__ __forceinline__ void col_inv(uint32_t (&u)[E], uint32_t *X, const tw_entry *twi, uint32_t tt, const ptw &cur,
                                        const mtw &M)
{
	ptw nxt;
	if constexpr (L > 0) load_ptw<RL, L - 1>(nxt, twi, 1u, tt);
	inv_pass<P, L == 0>(u, cur, M);
	if constexpr (L > 0)
	{
		xchg<RL, L, L - 1>(u, X, tt);
		col_inv<RL, P, L - 1>(u, X, twi, tt, nxt, M);
	}
}

// only: the prime this workgroup transforms (-1: all of them)
template <int LN, int RL, int NP, int W>
__device__ __forceinline__ void colA_prime_all(const int32_t *D, uint32_t *X, uint32_t *res_lane, const tw_entry *tw,
                                               uint32_t c0, uint32_t cw, uint32_t tt, size_t plane, int only)
{
	using G = colgeo<LN, RL, W>;
	constexpr uint32_t n = 1u << LN, C = n >> RL;
	constexpr int La = G::passes - 1;
#pragma unroll
	for (int P = 0; P < NP; ++P)
	{
		if (only >= 0 && P != only) continue;
		uint32_t u[E];
		const int32_t *dr = D + cw * G::RP + lbase<RL, 0>(tt);
		const uint32_t p = kP[P];
#pragma unroll
		for (int q = 0; q < E; ++q) u[q] = from_signed_np<NP>(dr[loff<RL, 0>(q)], p);
		const tw_entry *twf = tw + (size_t)(P * 2) * n;
		uint32_t *Xc = X + cw * G::RP;
		ptw w0;
		load_ptw<RL, 0>(w0, twf, 1u, tt);
		if (P == 0) col_fwd<RL, 0, 0>(u, Xc, twf, tt, w0);
		else if (P == 1) col_fwd<RL, 1, 0>(u, Xc, twf, tt, w0);
		else col_fwd<RL, 2, 0>(u, Xc, twf, tt, w0);
		uint32_t *out = res_lane + P * plane + c0 + cw;
#pragma unroll
		for (int q = 0; q < E; ++q) store_through(&out[(size_t)elem<RL, La>(tt, q) * C], u[q]);
	}
}

// Forward column levels of every prime for 256-row columns around one
// transpose that all primes share (one barrier): the digits (layout 0 of D)
// are read once, prime P goes through pass 0, is transposed through its own
// free LDS plane bufs[P], then pass 1, and its residues are stored.  D_REUSED:
// one of the planes is D itself, and a barrier separates every thread's digit
// reads from the plane writes.
template <int LN, int RL, int NP, int W, bool D_REUSED>
__device__ __forceinline__ void colA_all_shared(const int32_t *D, uint32_t *const (&bufs)[NP], uint32_t *res_lane,
                                                const tw_entry *tw, uint32_t c0, uint32_t cw, uint32_t tt, size_t plane)
{
	using G = colgeo<LN, RL, W>;
	constexpr uint32_t n = 1u << LN, C = n >> RL;
	int32_t d[E];
	{
		const int32_t *dr = D + cw * G::RP + lbase<RL, 0>(tt);
#pragma unroll
		for (int q = 0; q < E; ++q) d[q] = dr[loff<RL, 0>(q)];
	}
	const tw_entry *twf[NP];
#pragma unroll
	for (int P = 0; P < NP; ++P) twf[P] = tw + (size_t)(P * 2) * n;
	uint32_t u[NP][E];
#pragma unroll
	for (int P = 0; P < NP; ++P)
#pragma unroll
		for (int q = 0; q < E; ++q) u[P][q] = from_signed_np<NP>(d[q], kP[P]);
	col_fwd_lo<RL, 0>(u[0], twf[0], tt);
	col_fwd_lo<RL, 1>(u[1], twf[1], tt);
	if constexpr (NP == 3) col_fwd_lo<RL, 2>(u[NP - 1], twf[NP - 1], tt);
	if constexpr (D_REUSED) __syncthreads();
#pragma unroll
	for (int P = 0; P < NP; ++P)
	{
		uint32_t *bw = bufs[P] + cw * G::RP + lbase<RL, 0>(tt);
#pragma unroll
		for (int q = 0; q < E; ++q) bw[loff<RL, 0>(q)] = u[P][q];
	}
	__syncthreads();
#pragma unroll
	for (int P = 0; P < NP; ++P)
	{
		const uint32_t *br = bufs[P] + cw * G::RP + lbase<RL, 1>(tt);
#pragma unroll
		for (int q = 0; q < E; ++q) u[P][q] = br[loff<RL, 1>(q)];
	}
	col_fwd_hi<RL, 0>(u[0], twf[0], tt);
	col_fwd_hi<RL, 1>(u[1], twf[1], tt);
	if constexpr (NP == 3) col_fwd_hi<RL, 2>(u[NP - 1], twf[NP - 1], tt);
#pragma unroll
	for (int P = 0; P < NP; ++P)
	{
		uint32_t *out = res_lane + P * plane + c0 + cw;
#pragma unroll
		for (int q = 0; q < E; ++q) store_through(&out[(size_t)elem<RL, 1>(tt, q) * C], u[P][q]);
	}
}

// A: chunk load (+ pending carry), column levels, residue planes.  MINUS: the
// transform has length 2n (n = nd digits in rows < R/2, zeros below), for the
// exact product modulo N- = b^nd - b^(nd/2) - 1 (folded in pass C).
// SPLIT: one workgroup per column group and prime (blockIdx.z), twice the
// workgroups of the combined form for the same work (both load the digits).
template <int LN, int RL, int NP, int W, bool MINUS, bool SPLIT>
__global__ __launch_bounds__((colgeo<LN, RL, W>::threads)) void k_colA(plain_args A, int operand)
{
	using G = colgeo<LN, RL, W>;
	constexpr uint32_t n = 1u << LN, C = n >> RL;
	constexpr uint32_t nd = MINUS ? n / 2 : n, nc = nd / 16, RD = MINUS ? G::R / 2 : G::R;   // digits, chunks, digit rows
	__shared__ int32_t D[G::W * G::RP];
	__shared__ uint32_t X[G::W * G::RP];
	__shared__ uint32_t Y[(NP == 3 && G::passes == 2 && !SPLIT) ? G::W * G::RP : 1];    // third transpose plane
	const uint32_t t = threadIdx.x, lane = blockIdx.y, c0 = xcd_group(blockIdx.x, gridDim.x) * G::W;
	if (A.lane_i && A.lane_i[lane] < 0) return;
	const int32_t *x = (operand ? A.xb : A.x) + (size_t)lane * nd;
	uint32_t *res = (operand ? A.res2 : A.res) + (size_t)lane * n;
	// chunk role: row r, chunk h of the W columns
	{
		constexpr uint32_t hpr = G::W / 16;
		const uint32_t r = t / hpr, h = t % hpr;
		const uint32_t k = r * C + c0 + 16 * h;
		int32_t dig[E];
#pragma unroll
		for (int q = 0; q < E; ++q) dig[q] = 0;
		if (r < RD)
		{
			const int4 *src = reinterpret_cast<const int4 *>(x + k);
#pragma unroll
			for (int v = 0; v < 4; ++v) { const int4 w = src[v]; dig[4 * v] = w.x; dig[4 * v + 1] = w.y; dig[4 * v + 2] = w.z; dig[4 * v + 3] = w.w; }
			if (A.carry_in && operand == 0)
			{
				// X^nd = X^(nd/2) -+ 1: the top carry goes to chunk 0 (negated for
				// N+) and is added to chunk nc/2
				const uint32_t g = k / 16;
				const int64_t *f = A.carry_in + (size_t)lane * nc;
				int64_t c = f[cslot((g + nc - 1) % nc, C / 16, RD)];
				if (g == 0) c = MINUS ? c : -c;
				else if (g == nc / 2) c += f[cslot(nc - 1, C / 16, RD)];
				carry_in(dig, c, A.dv[lane]);
			}
		}
#pragma unroll
		for (int q = 0; q < E; ++q) D[(16 * h + q) * G::RP + r + (r >> 5)] = dig[q];
	}
	__syncthreads();
	const uint32_t cw = t % G::W, tt = t / G::W;
	if constexpr (G::passes == 2 && !SPLIT)
	{
		uint32_t *bufs[NP];
		bufs[0] = X;
		if constexpr (NP == 3) bufs[1] = Y;
		bufs[NP - 1] = reinterpret_cast<uint32_t *>(D);
		colA_all_shared<LN, RL, NP, W, true>(D, bufs, res, A.tw, c0, cw, tt, (size_t)A.lanes * n);
	}
	else colA_prime_all<LN, RL, NP, W>(D, X, res, A.tw, c0, cw, tt, (size_t)A.lanes * n, SPLIT ? (int)blockIdx.z : -1);
}

// B: row levels with the pointwise square / product, in place
template <int LN, int RL, int P, bool MUL>
__device__ __forceinline__ void rowB_prime(plain_args &A, uint32_t *lds, uint32_t r, uint32_t lane, uint32_t t)
{
	constexpr int CL = LN - RL;
	constexpr uint32_t n = 1u << LN, C = 1u << CL, T = C / 16, R = 1u << RL;
	const size_t plane = (size_t)A.lanes * n;
	uint32_t *row = A.res + P * plane + (size_t)lane * n + (size_t)r * C;
	const tw_entry *twf = A.tw + (size_t)(P * 2) * n, *twi = A.tw + (size_t)(P * 2 + 1) * n;
	const uint32_t base = R + r;
	uint32_t u[E], v[E];
#pragma unroll
	for (int q = 0; q < E; ++q) u[q] = row[t + q * T];
	ftw<CL> F;
	fwd_all<CL, P, false, true, false>(u, lds, twf, twi, base, t, F);     // lds untouched so far
	if (MUL)
	{
		const uint32_t *row2 = A.res2 + P * plane + (size_t)lane * n + (size_t)r * C;
#pragma unroll
		for (int q = 0; q < E; ++q) v[q] = row2[t + q * T];
		fwd_all<CL, P, false, true>(v, lds, twf, twi, base, t, F);
	}
	else
	{
#pragma unroll
		for (int q = 0; q < E; ++q) v[q] = 0;
	}
	final_inv<CL, P, false, MUL, true>(u, v, lds, twf, twi, base, t, F);
#pragma unroll
	for (int q = 0; q < E; ++q) store_through(&row[t + q * T], u[q]);
}

template <int LN, int RL, bool MUL>
__global__ __launch_bounds__((1u << (LN - RL)) / 16) void k_rowB(plain_args A)
{
	constexpr int CL = LN - RL;
	__shared__ uint32_t lds[geo<CL>::lds_words];
	const uint32_t r = blockIdx.x, P = blockIdx.y, lane = blockIdx.z, t = threadIdx.x;
	if (A.lane_i && A.lane_i[lane] < 0) return;
	if (P == 0) rowB_prime<LN, RL, 0, MUL>(A, lds, r, lane, t);
	else if (P == 1) rowB_prime<LN, RL, 1, MUL>(A, lds, r, lane, t);
	else rowB_prime<LN, RL, 2, MUL>(A, lds, r, lane, t);
}

// B with the primes of a row in one workgroup of NP slices, slice P on prime
// P (every slice whole wavefronts, so the prime is uniform over a wavefront
// and every slice passes the same barriers); each slice exchanges through its
// own LDS region.  One workgroup per row instead of NP.
template <int LN, int RL, int NP>
struct rowh
{
	static constexpr uint32_t T = (1u << (LN - RL)) / 16;
	static constexpr bool ok = (T % 64 == 0) && (NP * T <= 1024);
};

template <int LN, int RL, int NP>
__global__ __launch_bounds__((NP * rowh<LN, RL, NP>::T)) void k_rowH(plain_args A)
{
	constexpr int CL = LN - RL;
	constexpr uint32_t T = rowh<LN, RL, NP>::T;
	static_assert(rowh<LN, RL, NP>::ok, "slices of whole wavefronts, at most 1024 threads");
	__shared__ uint32_t lds[NP][geo<CL>::lds_words];
	const uint32_t r = blockIdx.x, lane = blockIdx.y, P = threadIdx.x / T, t = threadIdx.x % T;
	if (A.lane_i && A.lane_i[lane] < 0) return;
	if (P == 0) rowB_prime<LN, RL, 0, false>(A, lds[0], r, lane, t);
	else if (P == 1) rowB_prime<LN, RL, 1, false>(A, lds[1], r, lane, t);
	else rowB_prime<LN, RL, 2, false>(A, lds[NP - 1], r, lane, t);
}

// B with the primes of a row in one workgroup, one after the other: the next
// prime's row is requested before the current one is transformed (its loads
// are in flight during the transform), and the current one is stored before
// the next is transformed.  Half the workgroups of k_rowB, each with half of
// its global-memory phases exposed.
template <int LN, int RL, int P, int NP>
__device__ __forceinline__ void rowS_prime(plain_args &A, uint32_t *lds, uint32_t r, uint32_t lane, uint32_t t,
                                           uint32_t (&u)[E])
{
	constexpr int CL = LN - RL;
	constexpr uint32_t n = 1u << LN, C = 1u << CL, T = C / 16, R = 1u << RL;
	const size_t plane = (size_t)A.lanes * n;
	uint32_t nx[E];
	if constexpr (P + 1 < NP)
	{
		const uint32_t *nrow = A.res + (P + 1) * plane + (size_t)lane * n + (size_t)r * C;
#pragma unroll
		for (int q = 0; q < E; ++q) nx[q] = nrow[t + q * T];
	}
	const tw_entry *twf = A.tw + (size_t)(P * 2) * n, *twi = A.tw + (size_t)(P * 2 + 1) * n;
	uint32_t dummy[E];
	ftw<CL> F;
	// no leading barrier: lds is untouched so far (P = 0), or the barrier
	// after the previous prime's transposes came last (P > 0)
	fwd_all<CL, P, false, true, false>(u, lds, twf, twi, R + r, t, F);
	final_inv<CL, P, false, false, true>(u, dummy, lds, twf, twi, R + r, t, F);
	uint32_t *row = A.res + P * plane + (size_t)lane * n + (size_t)r * C;
#pragma unroll
	for (int q = 0; q < E; ++q) store_through(&row[t + q * T], u[q]);
	if constexpr (P + 1 < NP)
	{
		__syncthreads();                 // every wave is done with this prime's LDS transposes
#pragma unroll
		for (int q = 0; q < E; ++q) u[q] = nx[q];
		rowS_prime<LN, RL, P + 1, NP>(A, lds, r, lane, t, u);
	}
}

template <int LN, int RL, int NP>
__global__ __launch_bounds__((1u << (LN - RL)) / 16) void k_rowS(plain_args A)
{
	constexpr int CL = LN - RL;
	constexpr uint32_t n = 1u << LN, C = 1u << CL, T = C / 16;
	__shared__ uint32_t lds[geo<CL>::lds_words];
	const uint32_t r = blockIdx.x, lane = blockIdx.y, t = threadIdx.x;
	if (A.lane_i && A.lane_i[lane] < 0) return;
	const uint32_t *row0 = A.res + (size_t)lane * n + (size_t)r * C;
	uint32_t u[E];
#pragma unroll
	for (int q = 0; q < E; ++q) u[q] = row0[t + q * T];
	rowS_prime<LN, RL, 0, NP>(A, lds, r, lane, t, u);
}

// C: inverse column levels, CRT, normalize1.  MINUS: the 2n coefficients of
// each column are folded by X^nd = X^(nd/2) + 1 in residue space first
// (rows r < R/4: c_r + c_{r+R/2} + c_{r+3R/4}; R/4 <= r < R/2: c_r + c_{r+R/4}
// + 2 c_{r+R/2}), leaving the nd = n/2 digit coefficients in rows < R/2.
template <int P>
__device__ __forceinline__ uint32_t fold_minus(const uint32_t *S, uint32_t col_base, uint32_t r, uint32_t R)
{
	constexpr uint32_t p = kP[P];
	auto at = [&](uint32_t row) { return S[col_base + row + (row >> 5)]; };
	if (r < R / 4) return red1p(red1p(at(r) + at(r + R / 2), p) + at(r + 3 * R / 4), p);
	const uint32_t c2 = at(r + R / 4 + R / 4);
	return red1p(red1p(at(r) + at(r + R / 4), p) + red1p(c2 + c2, p), p);
}

// Inverse column levels of prime P for 256-row columns, in two halves around
// one transpose that every prime shares: pass 1 in registers, then (after
// the caller's write of layout 1, a barrier and the read of layout 0) pass 0
// with the level-1 merge.
template <int RL, int P>
__device__ __forceinline__ void col_inv_hi(uint32_t (&u)[E], const tw_entry *twi, uint32_t tt)
{
	ptw c;
	load_ptw<RL, 1>(c, twi, 1u, tt);
	inv_pass<P, false>(u, c, load_mtw(twi));
}
template <int RL, int P>
__device__ __forceinline__ void col_inv_lo(uint32_t (&u)[E], const tw_entry *twi, uint32_t tt)
{
	ptw c;
	load_ptw<RL, 0>(c, twi, 1u, tt);
	inv_pass<P, true>(u, c, load_mtw(twi));
}
// Forward column levels of prime P for 256-row columns around one shared
// transpose: pass 0 (with the Phi6 split) from layout 0, then pass 1.
template <int RL, int P>
__device__ __forceinline__ void col_fwd_lo(uint32_t (&u)[E], const tw_entry *twf, uint32_t tt)
{
	ptw c;
	load_ptw<RL, 0>(c, twf, 1u, tt);
	fwd_pass<P, true>(u, c);
}
template <int RL, int P>
__device__ __forceinline__ void col_fwd_hi(uint32_t (&u)[E], const tw_entry *twf, uint32_t tt)
{
	ptw c;
	load_ptw<RL, 1>(c, twf, 1u, tt);
	fwd_pass<P, false>(u, c);
}

// The inverse column levels, the CRT and normalize1 of column pass C for the
// chunk role of thread t (row r, chunk h of the group's W columns): the
// chunk's digits and outgoing carry.  Returns false for threads whose row
// holds no digits (MINUS, r >= R/2).  Every thread passes the barriers.
template <int LN, int RL, int NP, int W, bool MINUS>
__device__ __forceinline__ bool colC_chunk(const plain_args &A, uint32_t (*S)[colgeo<LN, RL, W>::W * colgeo<LN, RL, W>::RP],
                                           uint32_t *X, uint32_t lane, uint32_t c0, uint32_t t, int no_dup,
                                           int32_t (&dig)[E], int64_t &f)
{
	using G = colgeo<LN, RL, W>;
	constexpr uint32_t n = 1u << LN, C = n >> RL;
	constexpr uint32_t RD = MINUS ? G::R / 2 : G::R;
	constexpr int La = G::passes - 1;
	const uint32_t cw = t % G::W, tt = t / G::W;
	const size_t plane = (size_t)A.lanes * n;
	const uint32_t *res = A.res + (size_t)lane * n;
	// every prime's residues are requested before any of them is transformed
	uint32_t uu[NP][E];
#pragma unroll
	for (int P = 0; P < NP; ++P)
	{
		const uint32_t *in = res + P * plane + c0 + cw;
#pragma unroll
		for (int q = 0; q < E; ++q) uu[P][q] = in[(size_t)elem<RL, La>(tt, q) * C];
	}
	if constexpr (G::passes == 2)
	{
		// 256-row columns: every prime's pass 1, one transpose through each
		// prime's own S[P] (a single barrier for all of them), pass 0 and the
		// merge, the result written back in place (each thread rewrites exactly
		// the words it read)
		const tw_entry *twi[NP];
#pragma unroll
		for (int P = 0; P < NP; ++P) twi[P] = A.tw + (size_t)(P * 2 + 1) * n;
		col_inv_hi<RL, 0>(uu[0], twi[0], tt);
		col_inv_hi<RL, 1>(uu[1], twi[1], tt);
		if constexpr (NP == 3) col_inv_hi<RL, 2>(uu[NP - 1], twi[NP - 1], tt);
#pragma unroll
		for (int P = 0; P < NP; ++P)
		{
			uint32_t *sw = S[P] + cw * G::RP + lbase<RL, 1>(tt);
#pragma unroll
			for (int q = 0; q < E; ++q) sw[loff<RL, 1>(q)] = uu[P][q];
		}
		__syncthreads();
#pragma unroll
		for (int P = 0; P < NP; ++P)
		{
			const uint32_t *sr = S[P] + cw * G::RP + lbase<RL, 0>(tt);
#pragma unroll
			for (int q = 0; q < E; ++q) uu[P][q] = sr[loff<RL, 0>(q)];
		}
		col_inv_lo<RL, 0>(uu[0], twi[0], tt);
		col_inv_lo<RL, 1>(uu[1], twi[1], tt);
		if constexpr (NP == 3) col_inv_lo<RL, 2>(uu[NP - 1], twi[NP - 1], tt);
#pragma unroll
		for (int P = 0; P < NP; ++P)
		{
			uint32_t *sw = S[P] + cw * G::RP + lbase<RL, 0>(tt);
#pragma unroll
			for (int q = 0; q < E; ++q) sw[loff<RL, 0>(q)] = uu[P][q];
		}
	}
	else
	{
#pragma unroll
		for (int P = 0; P < NP; ++P)
		{
			const tw_entry *twi = A.tw + (size_t)(P * 2 + 1) * n;
			uint32_t *Xc = X + cw * G::RP;
			ptw c;
			load_ptw<RL, La>(c, twi, 1u, tt);
			const mtw M = load_mtw(twi);
			if (P == 0) col_inv<RL, 0, La>(uu[0], Xc, twi, tt, c, M);
			else if (P == 1) col_inv<RL, 1, La>(uu[1], Xc, twi, tt, c, M);
			else col_inv<RL, 2, La>(uu[NP - 1], Xc, twi, tt, c, M);
			uint32_t *sw = S[P] + cw * G::RP + lbase<RL, 0>(tt);
#pragma unroll
			for (int q = 0; q < E; ++q) sw[loff<RL, 0>(q)] = uu[P][q];
		}
	}
	__syncthreads();
	// chunk role (threads whose row holds digits)
	constexpr uint32_t hpr = G::W / 16;
	const uint32_t r = t / hpr, h = t % hpr;
	if (r >= RD) return false;
	const uint32_t rr = r + (r >> 5);
	bool dup = false;
	if (!no_dup)
	{
		const int bi = (A.lane_i ? A.lane_i[lane] : A.run[0]) - (A.run[1] + A.step);
		if (A.bits) dup = (A.bits[(size_t)lane * A.bit_words + (bi >> 5)] >> (bi & 31)) & 1u;
		else if (A.mask) dup = (A.mask[lane >> 5] >> (lane & 31)) & 1u;
	}
	if (NP == 3)
	{
		i96 l[E];
#pragma unroll
		for (int q = 0; q < E; ++q)
		{
			const uint32_t cb = (16 * h + q) * G::RP, a = cb + rr;
			if constexpr (MINUS)
				l[q] = garner(fold_minus<0>(S[0], cb, r, G::R), fold_minus<1>(S[1], cb, r, G::R),
				              fold_minus<2>(S[NP - 1], cb, r, G::R), A.gc);
			else l[q] = garner(S[0][a], S[1][a], S[NP - 1][a], A.gc);
		}
		f = carry_chunk(l, dig, dup, A.base_m1, A.dv[lane]);
	}
	else
	{
		int64_t l[E];
#pragma unroll
		for (int q = 0; q < E; ++q)
		{
			const uint32_t cb = (16 * h + q) * G::RP, a = cb + rr;
			if constexpr (MINUS) l[q] = garner2(fold_minus<0>(S[0], cb, r, G::R), fold_minus<1>(S[1], cb, r, G::R), A.gc);
			else l[q] = garner2(S[0][a], S[1][a], A.gc);
		}
		f = carry_chunk64(l, dig, dup, A.base_m1, A.dv[lane]);
	}
	return true;
}

template <int LN, int RL, int NP, int W, bool MINUS>
__global__ __launch_bounds__((colgeo<LN, RL, W>::threads)) void k_colC(plain_args A, int no_dup)
{
	using G = colgeo<LN, RL, W>;
	constexpr uint32_t n = 1u << LN, C = n >> RL;
	constexpr uint32_t nd = MINUS ? n / 2 : n, nc = nd / 16, RD = MINUS ? G::R / 2 : G::R;
	__shared__ uint32_t S[NP][G::W * G::RP];
	__shared__ uint32_t X[G::W * G::RP];
	const uint32_t t = threadIdx.x, lane = blockIdx.y, c0 = xcd_group(blockIdx.x, gridDim.x) * G::W;
	if (A.lane_i && A.lane_i[lane] < 0) return;
	int32_t dig[E];
	int64_t f;
	if (!colC_chunk<LN, RL, NP, W, MINUS>(A, S, X, lane, c0, t, no_dup, dig, f)) return;
	constexpr uint32_t hpr = G::W / 16;
	const uint32_t r = t / hpr, h = t % hpr;
	const uint32_t k = r * C + c0 + 16 * h;
	int4 *dst = reinterpret_cast<int4 *>(A.x + (size_t)lane * nd + k);
#pragma unroll
	for (int v = 0; v < 4; ++v) dst[v] = make_int4(dig[4 * v], dig[4 * v + 1], dig[4 * v + 2], dig[4 * v + 3]);
	A.carry[(size_t)lane * nc + cslot(k / 16, C / 16, RD)] = f;
}

// Column pass C of one squaring fused with column pass A of the next (N+):
// the inverse column levels, CRT and normalize1 as in k_colC, then the chunk
// carries cross to the next column group (normalize2 needs the carry of the
// chunk before each chunk, kernel.cl:1053-1081), then the carried digits go
// through LDS straight into the forward column levels of the next squaring.
//
// Exchange.  Group j publishes the carries of its last chunk per row in xc,
// then raises flags[j] to this launch's epoch and waits until flags[j - 1]
// holds it (group 0 waits for the last group, whose rows feed the next row's
// first chunk and, with X^n = X^(n/2) - 1, chunks 0 and n/32).  Every group
// publishes before it waits, so the waits end once every group has started:
// when the grid fits the device at once its workgroups are named by their
// block index (through xcd_group);
// otherwise (use_ticket) each takes a ticket (ctl[1], one atomic add; the
// holder of the last ticket resets the counter) and the ticket names its
// (lane, column group), so a group waits only for groups that started before
// it, except group 0, which waits for the last ticket of its lane.  The epoch
// is ctl[2] + step + 1: ctl[2] advances by the length of each launch
// sequence, so no two launches share an epoch and flags left by an earlier
// launch never match.  A wait that exceeds its bound sets ctl[0] and the run
// is reported as failed.
template <int LN, int RL, int NP, int W>
__global__ __launch_bounds__((colgeo<LN, RL, W>::threads)) void k_colCA(plain_args A, int32_t *flags, int32_t *ctl,
                                                                       int64_t *xc, int use_ticket)
{
	using G = colgeo<LN, RL, W>;
	constexpr uint32_t n = 1u << LN, C = n >> RL, GW = C / G::W, hpr = G::W / 16, R = G::R;
	__shared__ uint32_t S[NP][G::W * G::RP];
	__shared__ uint32_t X[G::W * G::RP];
	__shared__ int64_t fl[G::threads];
	__shared__ uint32_t ticket;
	const uint32_t t = threadIdx.x;
	uint32_t grp = xcd_group(blockIdx.x, gridDim.x);
	if (use_ticket)
	{
		if (t == 0)
		{
			const uint32_t total = GW * A.lanes;
			const uint32_t tk = (uint32_t)__hip_atomic_fetch_add(&ctl[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if (tk == total - 1) __hip_atomic_store(&ctl[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			ticket = tk;
		}
		__syncthreads();
		grp = ticket;
	}
	const uint32_t lane = grp / GW, j = grp % GW, c0 = j * G::W;
	// ctl[2] changes only between launches (k_epoch_add), so a plain load (one
	// scalar load per wavefront) reads it; unsigned: the epoch wraps after 2^32
	// launches (flags are only compared for equality)
	const int32_t epoch = (int32_t)((uint32_t)ctl[2] + (uint32_t)A.step + 1u);
	// an idle lane (per-lane runs): all its groups leave; waits stay within a lane
	if (A.lane_i && A.lane_i[lane] < 0) return;
	int32_t dig[E];
	int64_t f;
	colC_chunk<LN, RL, NP, W, false>(A, S, X, lane, c0, t, 0, dig, f);      // every row holds digits (N+)
	const uint32_t r = t / hpr, h = t % hpr;
	int64_t *xl = xc + (size_t)lane * GW * R;                               // [group][row]: carry of the row's last chunk
	int32_t *fg = flags + (size_t)lane * GW;
	// Ordering without cache maintenance: the carries, the flags and their
	// reads are device-scope atomics (gfx950 `sc1` accesses, coherent across
	// the XCDs' L2s).  Each storing thread waits for its store to complete
	// before the barrier, the flag is stored after the barrier, and the
	// carries are read after the barrier that follows the wait.  (Release and
	// acquire orderings would write back and invalidate a whole L2 per access
	// on this target.)
	fl[t] = f;
	if (h == hpr - 1) __hip_atomic_store(&xl[(size_t)j * R + r], f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	__builtin_amdgcn_s_waitcnt(0);
	__syncthreads();
	const uint32_t jp = (j + GW - 1) % GW;
	if (t == 0)
	{
		__hip_atomic_store(&fg[j], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		uint32_t spins = 0;
		while (__hip_atomic_load(&fg[jp], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch)
		{
			if (++spins > (1u << 22)) { __hip_atomic_store(&ctl[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); break; }
			__builtin_amdgcn_s_sleep(2);
		}
	}
	__syncthreads();
	// incoming carry of chunk (r, h) = chunk g = r C/16 + j hpr + h
	int64_t c;
	if (h > 0) c = fl[t - 1];
	else if (j > 0) c = __hip_atomic_load(&xl[(size_t)jp * R + r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	else
	{
		const int64_t top = __hip_atomic_load(&xl[(size_t)(GW - 1) * R + (R - 1)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (r == 0) c = -top;
		else
		{
			c = __hip_atomic_load(&xl[(size_t)(GW - 1) * R + (r - 1)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if (r == R / 2) c += top;
		}
	}
	carry_in(dig, c, A.dv[lane]);
	// the digits of the next squaring in the column layout of pass A (they do
	// not go to HBM: only the run's last pass C writes the register); S[0] is
	// free (every read of S came before the barrier above)
	int32_t *D = reinterpret_cast<int32_t *>(S[0]);
#pragma unroll
	for (int q = 0; q < E; ++q) D[(16 * h + q) * G::RP + r + (r >> 5)] = dig[q];
	__syncthreads();
	// forward column levels of every prime around one shared transpose: prime
	// P transposes through bufs[P] (X, then the S[P] that the CRT has finished
	// reading; S[0] holds the digits)
	uint32_t *bufs[NP];
	bufs[0] = X;
#pragma unroll
	for (int P = 1; P < NP; ++P) bufs[P] = S[P];
	colA_all_shared<LN, RL, NP, W, false>(D, bufs, A.res + (size_t)lane * n, A.tw, c0, t % G::W, t / G::W,
	                                      (size_t)A.lanes * n);
}

// k_colCA in two-prime mode with the primes in separate halves of a
// workgroup twice as large (thread t: prime t / TH, column role t % TH): each
// half runs its prime's inverse column levels, the first half the CRT,
// normalize1, the carry exchange and normalize2 (the CRT reads both primes'
// planes in LDS), then each half its prime's forward column levels from the
// digits in LDS.  Twice the waves per column group, half the registers per
// thread.  The exchange and its ordering are those of k_colCA.
template <int LN, int RL, int W>
__global__ __launch_bounds__((2 * colgeo<LN, RL, W>::threads)) void k_colCA2(plain_args A, int32_t *flags, int32_t *ctl,
                                                                            int64_t *xc, int use_ticket)
{
	using G = colgeo<LN, RL, W>;
	constexpr int NP = 2;
	constexpr uint32_t n = 1u << LN, C = n >> RL, GW = C / G::W, hpr = G::W / 16, R = G::R, TH = G::threads;
	static_assert(G::passes == 2 && TH % 64 == 0, "256-row columns; each half made of whole wavefronts");
	__shared__ uint32_t S[NP][G::W * G::RP];
	__shared__ uint32_t X[G::W * G::RP];
	__shared__ int64_t fl[TH];
	__shared__ uint32_t ticket;
	const uint32_t tf = threadIdx.x;
	const uint32_t P = tf / TH, t = tf % TH;             // P is uniform over a wavefront
	uint32_t grp = xcd_group(blockIdx.x, gridDim.x);
	if (use_ticket)
	{
		if (tf == 0)
		{
			const uint32_t total = GW * A.lanes;
			const uint32_t tk = (uint32_t)__hip_atomic_fetch_add(&ctl[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if (tk == total - 1) __hip_atomic_store(&ctl[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			ticket = tk;
		}
		__syncthreads();
		grp = ticket;
	}
	const uint32_t lane = grp / GW, j = grp % GW, c0 = j * G::W;
	// ctl[2] changes only between launches (k_epoch_add), so a plain load (one
	// scalar load per wavefront) reads it; unsigned: the epoch wraps after 2^32
	// launches (flags are only compared for equality)
	const int32_t epoch = (int32_t)((uint32_t)ctl[2] + (uint32_t)A.step + 1u);
	if (A.lane_i && A.lane_i[lane] < 0) return;
	const uint32_t cw = t % G::W, tt = t / G::W;
	const size_t plane = (size_t)A.lanes * n;
	// 1. inverse column levels and the merge of prime P, into S[P] (layout 0)
	{
		uint32_t u[E];
		const uint32_t *in = A.res + P * plane + (size_t)lane * n + c0 + cw;
#pragma unroll
		for (int q = 0; q < E; ++q) u[q] = in[(size_t)elem<RL, 1>(tt, q) * C];
		const tw_entry *twi = A.tw + (size_t)(P * 2 + 1) * n;
		if (P == 0) col_inv_hi<RL, 0>(u, twi, tt);
		else col_inv_hi<RL, 1>(u, twi, tt);
		uint32_t *sw = S[P] + cw * G::RP + lbase<RL, 1>(tt);
#pragma unroll
		for (int q = 0; q < E; ++q) sw[loff<RL, 1>(q)] = u[q];
		__syncthreads();
		uint32_t *sr = S[P] + cw * G::RP + lbase<RL, 0>(tt);
#pragma unroll
		for (int q = 0; q < E; ++q) u[q] = sr[loff<RL, 0>(q)];
		if (P == 0) col_inv_lo<RL, 0>(u, twi, tt);
		else col_inv_lo<RL, 1>(u, twi, tt);
#pragma unroll
		for (int q = 0; q < E; ++q) sr[loff<RL, 0>(q)] = u[q];        // the words this thread read
	}
	__syncthreads();
	// 2. first half: CRT and normalize1 of chunk (r, h), publish the row's last carry
	const uint32_t r = t / hpr, h = t % hpr;
	int32_t dig[E];
	int64_t *xl = xc + (size_t)lane * GW * R;
	int32_t *fg = flags + (size_t)lane * GW;
	if (P == 0)
	{
		const uint32_t rr = r + (r >> 5);
		const int bi = (A.lane_i ? A.lane_i[lane] : A.run[0]) - (A.run[1] + A.step);
		bool dup = false;
		if (A.bits) dup = (A.bits[(size_t)lane * A.bit_words + (bi >> 5)] >> (bi & 31)) & 1u;
		else if (A.mask) dup = (A.mask[lane >> 5] >> (lane & 31)) & 1u;
		int64_t l[E];
#pragma unroll
		for (int q = 0; q < E; ++q)
		{
			const uint32_t a = (16 * h + q) * G::RP + rr;
			l[q] = garner2(S[0][a], S[1][a], A.gc);
		}
		const int64_t f = carry_chunk64(l, dig, dup, A.base_m1, A.dv[lane]);
		fl[t] = f;
		if (h == hpr - 1) __hip_atomic_store(&xl[(size_t)j * R + r], f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		__builtin_amdgcn_s_waitcnt(0);
	}
	__syncthreads();
	const uint32_t jp = (j + GW - 1) % GW;
	if (tf == 0)
	{
		__hip_atomic_store(&fg[j], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		uint32_t spins = 0;
		while (__hip_atomic_load(&fg[jp], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch)
		{
			if (++spins > (1u << 22)) { __hip_atomic_store(&ctl[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); break; }
			__builtin_amdgcn_s_sleep(2);
		}
	}
	__syncthreads();
	// normalize2 and the digits into D = S[0] (every read of S came before the barriers above)
	int32_t *D = reinterpret_cast<int32_t *>(S[0]);
	if (P == 0)
	{
		int64_t c;
		if (h > 0) c = fl[t - 1];
		else if (j > 0) c = __hip_atomic_load(&xl[(size_t)jp * R + r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		else
		{
			const int64_t top = __hip_atomic_load(&xl[(size_t)(GW - 1) * R + (R - 1)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if (r == 0) c = -top;
			else
			{
				c = __hip_atomic_load(&xl[(size_t)(GW - 1) * R + (r - 1)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				if (r == R / 2) c += top;
			}
		}
		carry_in(dig, c, A.dv[lane]);
#pragma unroll
		for (int q = 0; q < E; ++q) D[(16 * h + q) * G::RP + r + (r >> 5)] = dig[q];
	}
	__syncthreads();
	// 3. forward column levels of prime P from the digits; prime 0 transposes
	// through X, prime 1 through S[1] (free since the CRT)
	{
		uint32_t u[E];
		const int32_t *dr = D + cw * G::RP + lbase<RL, 0>(tt);
#pragma unroll
		for (int q = 0; q < E; ++q) u[q] = from_signed_np<NP>(dr[loff<RL, 0>(q)], kP[P]);
		const tw_entry *twf = A.tw + (size_t)(P * 2) * n;
		if (P == 0) col_fwd_lo<RL, 0>(u, twf, tt);
		else col_fwd_lo<RL, 1>(u, twf, tt);
		uint32_t *buf = (P == 0) ? X : S[1];
		uint32_t *bw = buf + cw * G::RP + lbase<RL, 0>(tt);
#pragma unroll
		for (int q = 0; q < E; ++q) bw[loff<RL, 0>(q)] = u[q];
		__syncthreads();
		const uint32_t *br = buf + cw * G::RP + lbase<RL, 1>(tt);
#pragma unroll
		for (int q = 0; q < E; ++q) u[q] = br[loff<RL, 1>(q)];
		if (P == 0) col_fwd_hi<RL, 0>(u, twf, tt);
		else col_fwd_hi<RL, 1>(u, twf, tt);
		uint32_t *out = A.res + P * plane + (size_t)lane * n + c0 + cw;
#pragma unroll
		for (int q = 0; q < E; ++q) store_through(&out[(size_t)elem<RL, 1>(tt, q) * C], u[q]);
	}
}

// normalize2 alone: materialises the pending carries into the digits
// n = digits per lane, rows = rows holding digits, minus: N- wrap
__global__ void k_carry(int32_t *x, const int64_t *carry, const divisor *dvs, uint32_t n, uint32_t lanes, uint32_t rows,
                        int minus)
{
	const uint32_t nc = n / 16, cpr = nc / rows;
	const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (id >= (size_t)nc * lanes) return;
	const uint32_t lane = (uint32_t)(id / nc), g = (uint32_t)(id % nc);
	const int64_t *f = carry + (size_t)lane * nc;
	int64_t c = f[cslot((g + nc - 1) % nc, cpr, rows)];
	if (g == 0) c = minus ? c : -c;
	else if (g == nc / 2) c += f[cslot(nc - 1, cpr, rows)];
	int4 *p = reinterpret_cast<int4 *>(x + (size_t)lane * n + 16 * (size_t)g);
	int32_t dig[E];
#pragma unroll
	for (int v = 0; v < 4; ++v) { const int4 w = p[v]; dig[4 * v] = w.x; dig[4 * v + 1] = w.y; dig[4 * v + 2] = w.z; dig[4 * v + 3] = w.w; }
	carry_in(dig, c, dvs[lane]);
#pragma unroll
	for (int v = 0; v < 4; ++v) p[v] = make_int4(dig[4 * v], dig[4 * v + 1], dig[4 * v + 2], dig[4 * v + 3]);
}

// run control: {i0, iter} (one thread; kernel nodes only, so that the
// sequence can be captured in a graph)
__global__ void k_run_set(int32_t *run, int32_t i0) { run[0] = i0; run[1] = 0; }
__global__ void k_run_add(int32_t *run, int32_t k) { run[1] += k; }
// fused pass: advance the epoch base ctl[2] past a launch sequence
__global__ void k_epoch_add(int32_t *ctl, int32_t k)
{
	const uint32_t e = (uint32_t)__hip_atomic_load(&ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + (uint32_t)k;
	__hip_atomic_store(&ctl[2], (int32_t)e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// host sequencing

static int rows_log2(int ln) { return (ln <= 12) ? 4 : 8; }

// Column tiles compiled for a transform length: W = 64 with 16-row columns
// (n <= 2^12); W = 16 or 32 with 256-row columns.
static bool tile_ok(int ln, uint32_t w) { return (ln <= 12) ? (w == 64) : (w == 16 || w == 32); }
uint32_t plain_default_tile(int ln) { return (ln <= 12) ? 64 : ((ln >= 19) ? 32 : 16); }

// parts of a squaring: 1 = column pass A, 2 = row pass B, 4 = column pass C
template <int LN, int RL, int NP, int W, bool MINUS>
static hipError_t square_ln(const plain_args &A, uint32_t lanes, hipStream_t st, int parts)
{
	using G = colgeo<LN, RL, W>;
	constexpr uint32_t C = 1u << (LN - RL), R = 1u << RL;
	const dim3 gcol(C / G::W, lanes);
	if (parts & 1)
	{
		if (A.split_a) hipLaunchKernelGGL((k_colA<LN, RL, NP, W, MINUS, true>), dim3(C / G::W, lanes, NP), dim3(G::threads), 0, st, A, 0);
		else hipLaunchKernelGGL((k_colA<LN, RL, NP, W, MINUS, false>), gcol, dim3(G::threads), 0, st, A, 0);
	}
	if (parts & 2)
	{
		// row form 2 where its slices fit one workgroup, else form 0
		bool done = false;
		if constexpr (rowh<LN, RL, NP>::ok)
			if (A.row_form == 2)
			{
				hipLaunchKernelGGL((k_rowH<LN, RL, NP>), dim3(R, lanes), dim3(NP * (C / 16)), 0, st, A);
				done = true;
			}
		if (!done)
		{
			if (A.row_form == 1) hipLaunchKernelGGL((k_rowS<LN, RL, NP>), dim3(R, lanes), dim3(C / 16), 0, st, A);
			else hipLaunchKernelGGL((k_rowB<LN, RL, false>), dim3(R, NP, lanes), dim3(C / 16), 0, st, A);
		}
	}
	if (parts & 4) hipLaunchKernelGGL((k_colC<LN, RL, NP, W, MINUS>), gcol, dim3(G::threads), 0, st, A, 0);
	return hipGetLastError();
}

// Fused pass C + next pass A (N+, 256-row layout): flags per (lane, column
// group), ctl = {error word, ticket counter, epoch base}, carries; whether the
// grid fits the device at once (fits != null) or the launch.
struct fused_ctl { int32_t *flags; int32_t *ctl; int64_t *xc; int use_ticket; };

template <int LN, int NP, int W>
static hipError_t fused_ln(const plain_args &A, uint32_t lanes, hipStream_t st, const fused_ctl &F, bool *fits)
{
	using G = colgeo<LN, 8, W>;
	constexpr uint32_t C = 1u << (LN - 8);
	const bool halves = (NP == 2) && A.split_a;           // the primes in separate halves (k_colCA2)
	const void *fn = halves ? reinterpret_cast<const void *>(&k_colCA2<LN, 8, W>)
	                        : reinterpret_cast<const void *>(&k_colCA<LN, 8, NP, W>);
	const uint32_t threads = halves ? 2 * G::threads : G::threads;
	if (fits)
	{
		int dev = 0, cus = 0, per = 0;
		hipError_t h = hipGetDevice(&dev);
		if (h == hipSuccess) h = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
		if (h == hipSuccess) h = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, (int)threads, 0);
		if (h != hipSuccess) return h;
		*fits = (uint64_t)per * (uint64_t)cus >= (uint64_t)(C / G::W) * lanes;
		return hipSuccess;
	}
	if (halves)
		hipLaunchKernelGGL((k_colCA2<LN, 8, W>), dim3((C / G::W) * lanes), dim3(threads), 0, st, A, F.flags, F.ctl, F.xc,
		                   F.use_ticket);
	else
		hipLaunchKernelGGL((k_colCA<LN, 8, NP, W>), dim3((C / G::W) * lanes), dim3(threads), 0, st, A, F.flags, F.ctl, F.xc,
		                   F.use_ticket);
	return hipGetLastError();
}

template <int LN, int RL, int NP, int W, bool MINUS>
static hipError_t mul_ln(const plain_args &In, const plain_args &Out, uint32_t lanes, hipStream_t st)
{
	using G = colgeo<LN, RL, W>;
	constexpr uint32_t C = 1u << (LN - RL), R = 1u << RL;
	const dim3 gcol(C / G::W, lanes);
	hipLaunchKernelGGL((k_colA<LN, RL, NP, W, MINUS, false>), gcol, dim3(G::threads), 0, st, In, 0);
	hipLaunchKernelGGL((k_colA<LN, RL, NP, W, MINUS, false>), gcol, dim3(G::threads), 0, st, In, 1);
	hipLaunchKernelGGL((k_rowB<LN, RL, true>), dim3(R, NP, lanes), dim3(C / 16), 0, st, Out);
	hipLaunchKernelGGL((k_colC<LN, RL, NP, W, MINUS>), gcol, dim3(G::threads), 0, st, Out, 1);
	return hipGetLastError();
}

template <int LN, int RL, int W, bool MINUS>
static hipError_t square_np(int np, const plain_args &A, uint32_t lanes, hipStream_t st, int parts)
{
	return (np == 2) ? square_ln<LN, RL, 2, W, MINUS>(A, lanes, st, parts) : square_ln<LN, RL, 3, W, MINUS>(A, lanes, st, parts);
}

template <int LN, int RL, int W, bool MINUS>
static hipError_t mul_np(int np, const plain_args &In, const plain_args &Out, uint32_t lanes, hipStream_t st)
{
	return (np == 2) ? mul_ln<LN, RL, 2, W, MINUS>(In, Out, lanes, st) : mul_ln<LN, RL, 3, W, MINUS>(In, Out, lanes, st);
}

template <int LN, int RL, bool MINUS>
static hipError_t square_w(int np, uint32_t w, const plain_args &A, uint32_t lanes, hipStream_t st, int parts)
{
	if constexpr (RL == 4) return square_np<LN, RL, 64, MINUS>(np, A, lanes, st, parts);
	else return (w == 32) ? square_np<LN, RL, 32, MINUS>(np, A, lanes, st, parts)
	                      : square_np<LN, RL, 16, MINUS>(np, A, lanes, st, parts);
}

template <int LN, int RL, bool MINUS>
static hipError_t mul_w(int np, uint32_t w, const plain_args &In, const plain_args &Out, uint32_t lanes, hipStream_t st)
{
	if constexpr (RL == 4) return mul_np<LN, RL, 64, MINUS>(np, In, Out, lanes, st);
	else return (w == 32) ? mul_np<LN, RL, 32, MINUS>(np, In, Out, lanes, st) : mul_np<LN, RL, 16, MINUS>(np, In, Out, lanes, st);
}

template <int LN, int RL>
static hipError_t square_m(int np, uint32_t w, bool minus, const plain_args &A, uint32_t lanes, hipStream_t st, int parts)
{
	if constexpr (LN >= 11)
		if (minus) return square_w<LN, RL, true>(np, w, A, lanes, st, parts);
	return minus ? hipErrorInvalidValue : square_w<LN, RL, false>(np, w, A, lanes, st, parts);
}

template <int LN, int RL>
static hipError_t fused_m(int np, uint32_t w, const plain_args &A, uint32_t lanes, hipStream_t st, const fused_ctl &F,
                          bool *fits)
{
	if constexpr (RL != 8) return hipErrorInvalidValue;
	else
	{
		if (np == 2) return (w == 32) ? fused_ln<LN, 2, 32>(A, lanes, st, F, fits) : fused_ln<LN, 2, 16>(A, lanes, st, F, fits);
		return (w == 32) ? fused_ln<LN, 3, 32>(A, lanes, st, F, fits) : fused_ln<LN, 3, 16>(A, lanes, st, F, fits);
	}
}

template <int LN, int RL>
static hipError_t mul_m(int np, uint32_t w, bool minus, const plain_args &In, const plain_args &Out, uint32_t lanes,
                        hipStream_t st)
{
	if constexpr (LN >= 11)
		if (minus) return mul_w<LN, RL, true>(np, w, In, Out, lanes, st);
	return minus ? hipErrorInvalidValue : mul_w<LN, RL, false>(np, w, In, Out, lanes, st);
}

#define CT_PLAIN_SIZES(X) X(10, 4) X(11, 4) X(12, 4) X(13, 8) X(14, 8) X(15, 8) X(16, 8) X(17, 8) \
	X(18, 8) X(19, 8) X(20, 8) X(21, 8) X(22, 8)

// tln: log2 of the transform length (n digits for N+, 2n for N-)
static hipError_t square_any(int tln, int np, uint32_t w, bool minus, const plain_args &A, uint32_t lanes, hipStream_t st,
                             int parts = 7)
{
	switch (tln)
	{
#define CT_SQ_CASE(L_, R_) case L_: return square_m<L_, R_>(np, w, minus, A, lanes, st, parts);
	CT_PLAIN_SIZES(CT_SQ_CASE)
#undef CT_SQ_CASE
	default: return hipErrorInvalidValue;
	}
}

static hipError_t fused_any(int tln, int np, uint32_t w, const plain_args &A, uint32_t lanes, hipStream_t st,
                            const fused_ctl &F, bool *fits = nullptr)
{
	switch (tln)
	{
#define CT_FU_CASE(L_, R_) case L_: return fused_m<L_, R_>(np, w, A, lanes, st, F, fits);
	CT_PLAIN_SIZES(CT_FU_CASE)
#undef CT_FU_CASE
	default: return hipErrorInvalidValue;
	}
}

static hipError_t mul_any(int tln, int np, uint32_t w, bool minus, const plain_args &In, const plain_args &Out,
                          uint32_t lanes, hipStream_t st)
{
	switch (tln)
	{
#define CT_MUL_CASE(L_, R_) case L_: return mul_m<L_, R_>(np, w, minus, In, Out, lanes, st);
	CT_PLAIN_SIZES(CT_MUL_CASE)
#undef CT_MUL_CASE
	default: return hipErrorInvalidValue;
	}
}

#define HIPCHK(call) do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail(_e, #call); } while (0)

bool plain_tile_ok(int ln, uint32_t w) { return tile_ok(ln, w); }
int plain_rows_log2(int ln) { return rows_log2(ln); }

int plain_alloc(ct_engine *e)
{
	if (e->tln < 10 || e->tln > 22 || (e->sign < 0 && e->tln < 11))
		return fail(CT_ERR_UNSUPPORTED, "multipass engine needs a transform length 2^10 .. 2^22 (2^11 .. for N-)");
	e->rows_log2 = (uint32_t)rows_log2((int)e->tln);
	if (!tile_ok((int)e->tln, e->col_w)) e->col_w = plain_default_tile((int)e->tln);
	const size_t words = (size_t)3 * e->lanes * e->tn;
	if (hipMalloc((void **)&e->d_res, words * sizeof(uint32_t)) != hipSuccess) return fail(CT_ERR_NOMEM, "hipMalloc residues");
	if (hipMalloc((void **)&e->d_res2, words * sizeof(uint32_t)) != hipSuccess) return fail(CT_ERR_NOMEM, "hipMalloc residues");
	if (hipMalloc((void **)&e->d_carry, (size_t)e->lanes * (e->n / 16) * sizeof(int64_t)) != hipSuccess)
		return fail(CT_ERR_NOMEM, "hipMalloc carries");
	if (hipMalloc((void **)&e->d_run, 2 * sizeof(int32_t)) != hipSuccess) return fail(CT_ERR_NOMEM, "hipMalloc run");
	// fused column pass: flags and carries per (lane, column group) of the
	// narrowest tile, then {error word, ticket counter, epoch base}
	const size_t groups = (size_t)e->lanes * ((e->tn >> e->rows_log2) / 16);
	e->fused_bytes = (groups + 3) * sizeof(int32_t);
	if (hipMalloc((void **)&e->d_flags, e->fused_bytes) != hipSuccess ||
	    hipMalloc((void **)&e->d_xc, groups * (1u << e->rows_log2) * sizeof(int64_t)) != hipSuccess)
		return fail(CT_ERR_NOMEM, "hipMalloc fused-pass buffers");
	if (hipMemset(e->d_flags, 0, e->fused_bytes) != hipSuccess) return fail(CT_ERR_HIP, "hipMemset flags");
	e->d_ctl = e->d_flags + groups;
	e->fused_w = 0;
	return CT_OK;
}

// Destroys the cached executables.  Replays of them may still be queued on
// the engine's stream (plain_steps launches a run as consecutive replays and
// may come here between two of them), so the stream drains first: an
// executable is destroyed only after its last replay has completed.
static void drop_graphs(ct_engine *e)
{
	if (e->graphs.empty()) return;
	if (e->st) (void)hipStreamSynchronize(e->st);
	for (auto &g : e->graphs) hipGraphExecDestroy(g.exec);
	e->graphs.clear();
}

void plain_free(ct_engine *e)
{
	drop_graphs(e);
	hipFree(e->d_res); hipFree(e->d_res2); hipFree(e->d_carry); hipFree(e->d_run); hipFree(e->d_flags); hipFree(e->d_xc);
	e->d_res = e->d_res2 = nullptr; e->d_carry = nullptr; e->d_run = nullptr; e->d_flags = e->d_ctl = nullptr; e->d_xc = nullptr;
}

static plain_args make_args(ct_engine *e)
{
	plain_args A;
	std::memset(&A, 0, sizeof(A));
	A.res = e->d_res; A.res2 = e->d_res2; A.carry = e->d_carry;
	A.tw = e->d_tw; A.dv = e->d_dv; A.lanes = e->lanes; A.gc = e->gc; A.run = e->d_run; A.base_m1 = (int)e->base - 1;
	A.split_a = e->split_a ? 1 : 0;
	A.row_form = e->row_form;
	return A;
}

// Applies the pending carries of register 0 (after squarings).
int plain_materialize(ct_engine *e)
{
	if (!e->pending) return CT_OK;
	const size_t total = (size_t)e->lanes * (e->n / 16);
	hipLaunchKernelGGL(k_carry, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, e->st,
	                   e->reg(0), (const int64_t *)e->d_carry, (const divisor *)e->d_dv, e->n, e->lanes, e->digit_rows(), e->sign < 0 ? 1 : 0);
	HIPCHK(hipGetLastError());
	e->pending = false;
	if (e->fused_check)
	{   // the fused runs since the last read: did every carry exchange complete?
		e->fused_check = false;
		int32_t err = 0;
		HIPCHK(hipMemcpyAsync(&err, e->d_ctl, sizeof(err), hipMemcpyDeviceToHost, e->st));
		HIPCHK(hipStreamSynchronize(e->st));
		if (err)
		{
			e->use_fused = false;
			e->fused_w = 0;
			return fail(CT_ERR_HIP, "fused column pass: a carry exchange between workgroups timed out");
		}
	}
	return CT_OK;
}

// Launches `len` squarings of register 0, step indices relative to the run
// counter: A B C per squaring, or with `fused` (len >= 2) A B, then (CA B) per
// further squaring, then C, then the epoch base advances.  Either way the last
// pass C leaves its carries pending.
static hipError_t launch_seq(ct_engine *e, plain_args A, int len, bool fused, bool first_carry)
{
	const int tln = (int)e->tln;
	hipError_t h = hipSuccess;
	if (!fused || len < 2)
	{
		for (int j = 0; j < len && h == hipSuccess; ++j)
		{
			A.step = j;
			A.carry_in = (j == 0 && !first_carry) ? nullptr : e->d_carry;
			h = square_any(tln, e->np, e->col_w, e->sign < 0, A, e->lanes, e->st);
		}
		return h;
	}
	const fused_ctl F{e->d_flags, e->d_ctl, e->d_xc, e->fused_ticket ? 1 : 0};
	A.step = 0;
	A.carry_in = first_carry ? e->d_carry : nullptr;
	h = square_any(tln, e->np, e->col_w, false, A, e->lanes, e->st, 3);
	for (int j = 1; j < len && h == hipSuccess; ++j)
	{
		A.step = j - 1;                             // pass C belongs to squaring j - 1
		h = fused_any(tln, e->np, e->col_w, A, e->lanes, e->st, F);
		A.step = j;
		if (h == hipSuccess) h = square_any(tln, e->np, e->col_w, false, A, e->lanes, e->st, 2);
	}
	A.step = len - 1;
	if (h == hipSuccess) h = square_any(tln, e->np, e->col_w, false, A, e->lanes, e->st, 4);
	if (h == hipSuccess)
	{
		hipLaunchKernelGGL(k_epoch_add, dim3(1), dim3(1), 0, e->st, e->d_ctl, (int32_t)len);
		h = hipGetLastError();
	}
	return h;
}

// A run of squarings is launched as replays of captured graphs of up to
// kGraphMax squarings (one graph per distinct length, form, first-carry state
// and argument set; the last node advances run[1]), or eagerly.
static constexpr int kGraphMax = 128;
static constexpr size_t kGraphCache = 16;

static int graph_for(ct_engine *e, plain_args A, int len, bool fused, hipGraphExec_t *out)
{
	const bool fc = e->pending;
	for (const auto &g : e->graphs)
		if (g.nsteps == len && g.np == e->np && g.col_w == e->col_w && g.base == e->base && g.split == e->split_a &&
		    g.row_form == e->row_form &&
		    g.fused == fused && g.ticket == (fused && e->fused_ticket) && g.first_carry == fc && g.bits == A.bits &&
		    g.mask == A.mask &&
		    g.lane_i == A.lane_i && g.words == A.bit_words) { *out = g.exec; return CT_OK; }
	if (e->graphs.size() >= kGraphCache) drop_graphs(e);
	hipGraph_t graph = nullptr;
	HIPCHK(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
	hipError_t h = launch_seq(e, A, len, fused, fc);
	if (h == hipSuccess)
	{
		hipLaunchKernelGGL(k_run_add, dim3(1), dim3(1), 0, e->st, e->d_run, (int32_t)len);
		h = hipGetLastError();
	}
	const hipError_t h2 = hipStreamEndCapture(e->st, &graph);
	if (h != hipSuccess || h2 != hipSuccess)
	{
		if (graph) hipGraphDestroy(graph);
		return hip_fail(h != hipSuccess ? h : h2, "graph capture of the squaring sequence");
	}
	hipGraphExec_t exec = nullptr;
	h = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
	hipGraphDestroy(graph);
	if (h != hipSuccess) return hip_fail(h, "hipGraphInstantiate");
	e->graphs.push_back({len, e->np, e->col_w, e->base, e->split_a, fused, fused && e->fused_ticket, fc, e->row_form, A.bits,
	                     A.mask, A.lane_i, A.bit_words, exec});
	*out = exec;
	return CT_OK;
}

int plain_steps(ct_engine *e, const uint32_t *bits, uint32_t bit_words, int i_hi, int nsteps, const uint32_t *mask_words,
                const int32_t *lane_i)
{
	plain_args A = make_args(e);
	A.x = e->reg(0); A.bits = bits; A.bit_words = bit_words; A.mask = mask_words; A.lane_i = lane_i;
	A.run = e->d_run;
	hipLaunchKernelGGL(k_run_set, dim3(1), dim3(1), 0, e->st, e->d_run, (int32_t)i_hi);
	HIPCHK(hipGetLastError());
	// the fused form: N+, 256-row layout
	const bool fused = e->use_fused && e->sign > 0 && e->rows_log2 == 8 && nsteps > 1;
	if (fused && (e->fused_w != e->col_w || e->fused_np != e->np || e->fused_split != e->split_a))
	{   // a new kernel or grid shape: clear the flags, the words and the epoch
		// base; tickets only when the grid does not fit the device at once
		bool fits = false;
		const fused_ctl F{e->d_flags, e->d_ctl, e->d_xc, 0};
		HIPCHK(fused_any((int)e->tln, e->np, e->col_w, A, e->lanes, e->st, F, &fits));
		e->fused_ticket = !fits || e->fused_force_ticket;
		HIPCHK(hipMemsetAsync(e->d_flags, 0, e->fused_bytes, e->st));
		e->fused_w = e->col_w;
		e->fused_np = e->np;
		e->fused_split = e->split_a;
	}
	if (e->use_graphs && nsteps > 1)
	{
		for (int s = 0; s < nsteps;)
		{
			const int len = std::min(kGraphMax, nsteps - s);
			hipGraphExec_t exec = nullptr;
			int rc = graph_for(e, A, len, fused && len > 1, &exec);
			if (rc) return rc;
			HIPCHK(hipGraphLaunch(exec, e->st));
			e->pending = true;
			s += len;
		}
	}
	else
	{
		HIPCHK(launch_seq(e, A, nsteps, fused, e->pending));
		e->pending = true;
	}
	if (fused) e->fused_check = true;
	return CT_OK;
}

int plain_mul(ct_engine *e, uint32_t dst, uint32_t a, uint32_t b)
{
	int rc = plain_materialize(e);
	if (rc) return rc;
	// In: the two operands (column pass A reads x and xb); Out: dst (row pass
	// B and column pass C).  dst may alias a or b: both are read into the
	// residue planes before C writes dst.
	plain_args In = make_args(e);
	In.x = e->reg(a); In.xb = e->reg(b);
	plain_args Out = make_args(e);
	Out.x = e->reg(dst);
	HIPCHK(mul_any((int)e->tln, e->np, e->col_w, e->sign < 0, In, Out, e->lanes, e->st));
	// materialise dst at once (only register 0 may carry pending carries)
	const size_t total = (size_t)e->lanes * (e->n / 16);
	hipLaunchKernelGGL(k_carry, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, e->st,
	                   e->reg(dst), (const int64_t *)e->d_carry, (const divisor *)e->d_dv, e->n, e->lanes, e->digit_rows(), e->sign < 0 ? 1 : 0);
	HIPCHK(hipGetLastError());
	return CT_OK;
}

}   // namespace ct
