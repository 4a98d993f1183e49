// Launch interfaces shared by the engine host code and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "ct_arith.h"
#include "ct_tables.h"

namespace ct {

// k_res_steps: x <- x^2 * base^bit for nsteps steps, bit i = i_hi, i_hi - 1, ...
struct res_args
{
	int32_t *x;                 // [lanes][n] signed digits (register 0)
	const tw_entry *tw;         // [3][2][n] root tables
	const divisor *dv;          // [lanes] base b as a normalised divisor
	const uint32_t *bits;       // [lanes][bit_words] exponent bits, or null
	const uint32_t *mask;       // [lanes/32] dup bits when bits is null (or null = no dup)
	uint32_t bit_words;
	int i_hi;
	int nsteps;
	const int32_t *lane_i;      // per-lane top bit index (continuous batching), or null: i_hi;
	                            // a negative entry marks an idle lane (its workgroup exits)
	int base_m1;                // PRP base - 1 (a set bit multiplies by the base)
	garner_consts gc;
};

// k_res_mul: out <- a * b (no doubling)
struct mul_args
{
	const int32_t *a, *b;
	int32_t *out;
	const tw_entry *tw;
	const divisor *dv;
	garner_consts gc;
};

// halves: the two primes in the halves of one workgroup (two primes and
// resident_halves_ok(tln) only; otherwise ignored)
hipError_t resident_steps(int tln, int np, bool minus, bool halves, const res_args &A, uint32_t lanes, hipStream_t st);
bool resident_halves_ok(int tln);
hipError_t resident_mul(int tln, int np, bool minus, const mul_args &A, uint32_t lanes, hipStream_t st);

}   // namespace ct
