// Internal engine state shared by the host translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/ctwin.h"
#include "ct_arith.h"
#include "ct_tables.h"

struct ct_engine
{
	int device = 0;
	uint32_t ln = 0, n = 0, lanes = 0, csize = 16;   // n = 2^ln digits per candidate
	int sign = 1;                     // N = b^n - b^(n/2) + sign
	uint32_t tln = 0, tn = 0;         // transform length 2^tln: n (N+) or 2n (N-, exact product then fold)
	int kind = 0;
	hipStream_t st = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	int32_t *d_x = nullptr;           // 3 registers x [lanes][n] signed digits
	ct::tw_entry *d_tw = nullptr;     // [3][2][n]
	ct::divisor *d_dv = nullptr;      // [lanes]
	uint32_t *d_mask = nullptr;       // [lanes/32] dup bits of one squareDup
	uint32_t *d_bits = nullptr;       // [lanes][bit_words] exponent bits
	uint32_t bit_words = 0;
	bool b_set = false;
	std::vector<uint32_t> b;
	ct::garner_consts gc{};
	// multipass (plain) engine scratch
	uint32_t *d_res = nullptr;        // [3][lanes][n] residues between passes
	uint32_t *d_res2 = nullptr;       // second operand of a product
	int64_t *d_carry = nullptr;       // [lanes][n/16] chunk carries
	uint32_t rows_log2 = 0;           // column length R = 2^rows_log2
	uint32_t col_w = 0;               // column-pass tile width (0 = default for the length)
	bool split_a = false;             // column pass A: one workgroup per column group and prime
	bool res_halves = false;          // resident kernels: two primes in the halves of one workgroup
	int row_form = 0;                 // row pass B: 0 a workgroup per (row, prime), 1 per row with its primes
	                                  // in sequence, 2 per row with its primes in workgroup slices
	bool use_fused = false;           // fuse column pass C with the next pass A
	uint32_t fused_w = 0;             // tile of the grid that last used the flags (0: clear them)
	int fused_np = 0;                 // and its prime count
	bool fused_split = false;         // and its form (primes in separate workgroup halves)
	bool fused_ticket = false;        // the grid does not fit the device at once: order workgroups by ticket
	bool fused_force_ticket = false;  // order by ticket whether or not it fits
	bool fused_check = false;         // fused runs since the last read: check the error word
	size_t fused_bytes = 0;           // flags + {error word, ticket counter, epoch base}
	int32_t *d_flags = nullptr, *d_ctl = nullptr;   // fused pass: flags per (lane, column group); d_ctl = {error, ticket, epoch}
	int64_t *d_xc = nullptr;          // fused pass: last-chunk carries per (lane, column group, row)
	// autotuner timings (ms per squaring): resident, multipass W=16, W=32, and
	// the same two tiles with column pass A split by prime
	float tune_ms[CT_TUNE_TIMINGS] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
	bool tuned = false;               // created with csize 0 (autotuned)
	int tuned_np = 0;                 // prime count the multipass parameters were tuned for
	bool pending = false;             // register 0 still owes the chunk carries (plain engine)
	int np = 3;                       // primes per squaring: 2 when 3 n D^2 < p1 p2 / 2
	uint64_t b_max = 0;               // largest b over the lanes
	uint64_t state_bound = 1;         // bound on |digit| set or written into the registers
	bool force3 = false;              // always use three primes
	uint32_t base = 2;                // PRP base: a set exponent bit multiplies by it (2 or 3)
	// captured squaring sequences of the multipass engine (hipGraph), keyed
	// by what the captured kernel arguments depend on
	struct graph_entry { int nsteps, np; uint32_t col_w, base; bool split, fused, ticket, first_carry; int row_form; const void *bits, *mask, *lane_i; uint32_t words; hipGraphExec_t exec; };
	std::vector<graph_entry> graphs;
	bool use_graphs = true;
	int32_t *d_run = nullptr;         // {top bit index of the run, squarings done by replayed graphs}

	int32_t *reg(uint32_t r) const { return d_x + (size_t)r * lanes * n; }
	// multipass: rows of the R x C transform layout that hold digits
	uint32_t digit_rows() const { return (1u << rows_log2) >> (sign < 0 ? 1 : 0); }
	size_t reg_bytes() const { return (size_t)lanes * n * sizeof(int32_t); }
};

namespace ct {

int fail(int code, const char *what);
int hip_fail(hipError_t e, const char *where);
int canonical(const int32_t *d, uint32_t n, uint32_t b, int sign, uint32_t *x);
int steps(ct_engine *e, const uint32_t *bits, uint32_t bit_words, int i_hi, int nsteps, const uint32_t *mask_words,
          const int32_t *lane_i = nullptr);
int mul_regs(ct_engine *e, uint32_t dst, uint32_t a, uint32_t b);

hipError_t launch_set(int32_t *x, uint32_t n, uint32_t lanes, int32_t a, hipStream_t st);

int plain_alloc(ct_engine *e);
void plain_free(ct_engine *e);
int plain_steps(ct_engine *e, const uint32_t *bits, uint32_t bit_words, int i_hi, int nsteps, const uint32_t *mask_words,
                const int32_t *lane_i = nullptr);
int plain_mul(ct_engine *e, uint32_t dst, uint32_t a, uint32_t b);
int plain_materialize(ct_engine *e);
void choose_primes(ct_engine *e);
bool plain_tile_ok(int ln, uint32_t w);
int plain_rows_log2(int ln);          // log2 of the column length of the multipass layout
// The window sieve's primes from 2^32 up: d_list holds kWindow64List words
// and d_count one word, both owned by the caller.
constexpr size_t kWindow64List = 2ull << 21;
int strike_window64(hipStream_t st, int m, int mode, uint64_t k_lo, uint64_t k_hi, uint64_t p_lo, uint64_t p_hi,
                    uint64_t b_min, uint64_t b_max, uint8_t *d_struck, uint64_t *d_list, uint32_t *d_count);
uint32_t plain_default_tile(int ln);

}   // namespace ct
