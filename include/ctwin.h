/*
 * ctwin_amd -- C ABI of the MI355X (gfx950) engines for galloty/ctwin's two
 * data-parallel paths: the exact Fermat squaring engine for
 * N = b^n - b^(n/2) + 1 (n = 2^m digits in base b) and the b-window sieve.
 *
 * Plain C types only; every call returns a ct_status (0 = success) and never
 * exits the process (the reference's engine calls exit() on errors,
 * cyclo/src/ocl.h:106-113).  One engine per host thread; calls on one engine
 * are stream-ordered and only the read-out calls block.
 */
#ifndef CTWIN_H
#define CTWIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
	CT_OK = 0,
	CT_ERR_ARG = -1,          /* invalid argument */
	CT_ERR_HIP = -2,          /* HIP runtime error (no device, launch failure) */
	CT_ERR_NOMEM = -3,        /* device allocation failed */
	CT_ERR_UNSUPPORTED = -4,  /* size outside the engine's range */
	CT_ERR_GETINT = -5,       /* canonical read-out failed ("getInt failed", transform.h:589) */
	CT_ERR_COMPUTE = -6,      /* residue 0 ("computation failed", transform.h:611 / cyclo.h:369) */
	CT_ERR_GERBICZ = -7,      /* Gerbicz-Li mismatch ("Gerbicz failed", cyclo.h:377) */
	CT_INTERRUPTED = 1        /* not an error: a batch run stopped on request, checkpoint saved */
} ct_status;

typedef struct ct_engine ct_engine;

/* Engine kinds. */
#define CT_KIND_AUTO      0   /* resident kernels for n <= 2^14, else multipass */
#define CT_KIND_RESIDENT  1   /* one workgroup per candidate, state on chip (n <= 2^14) */
#define CT_KIND_PLAIN     2   /* per-squaring launch sequence over the whole GPU */

/* Replaces engine(platform, d, verbose) + transform(size, engine, isBoinc,
 * csize) (cyclo/src/engine.h:40, cyclo/src/transform.h:380-433): device
 * ordinal, log2 n (5..22; resident kernels to 14, multipass from 10), number of candidates ("lanes"; the reference
 * fixes VSIZE = 64, engine.h:19), carry chunk csize, engine kind.  The
 * carry chunk is 16 digits (the carried digits depend on it, the residues do
 * not).  csize = 16 builds the engine as asked (AUTO: resident kernels for
 * n <= 2^14, else multipass).  csize = 0 runs the autotuner first, as the
 * reference does for csize 0 (transform.h:383-419, which times CSIZE 8..64):
 * it times squarings under each compiled configuration -- resident kernels
 * vs the multipass engine when kind is AUTO and n <= 2^14; for the resident
 * kernels, two primes one after the other or in workgroup halves; for the
 * multipass engine the column tile, the split of column pass A, the fused
 * column pass and the row-pass form -- keeps the fastest, and caches the choice per (device, n,
 * lanes, kind, sign, prime count) for the process.  The engine is tuned for
 * three primes at creation; when ct_engine_set_b or ct_engine_set_primes
 * moves a tuned multipass engine to another prime count, its multipass
 * parameters are tuned again for that count (the engine kind stays).
 * Allocates 3 digit registers per lane. */
int ct_engine_create(int device, uint32_t log2_n, uint32_t lanes, uint32_t csize, int kind, ct_engine **out);

/* ct_engine_create for either twin member: sign = +1 gives the engine above
 * (N+ = b^n - b^(n/2) + 1), sign = -1 an engine modulo N- = b^n - b^(n/2) - 1
 * (no PRP engine in the reference; its sieve defines N-, ctsieve.cpp:728-787).
 * The N- engine computes the exact product with the same transform at length
 * 2n (digits zero-padded) and folds it by X^n = X^(n/2) + 1; log2_n <= 21, resident
 * kernels for n <= 2^13.  Every call below works on both kinds of engine; the
 * context file and the batch driver are N+ only. */
int ct_engine_create_ex(int device, uint32_t log2_n, uint32_t lanes, uint32_t csize, int kind, int sign, ct_engine **out);

/* Multipass engine: columns per workgroup of the column passes (16 or 32 for
 * n >= 2^13; 64 below).  Results do not depend on it.  No counterpart in the
 * reference (its work-group geometry is fixed per kernel, kernel.cl). */
int ct_engine_set_column_tile(ct_engine *e, uint32_t w);

/* Multipass engine: column pass A as one workgroup per column group (on = 0)
 * or one per column group and prime (on = 1: twice the workgroups, both load
 * the digits); with the fused column pass and two primes, on = 1 puts the
 * primes in separate halves of a workgroup twice as large.  Results do not
 * depend on it.  No counterpart in the reference. */
int ct_engine_set_column_split(ct_engine *e, int on);

/* Multipass engine (N+, n >= 2^13): fuse column pass C of each squaring with
 * column pass A of the next (on = 1), the column workgroups exchanging their
 * chunk carries through device memory within one launch, or keep three
 * launches per squaring (on = 0).  on = 2 is on = 1 with the workgroups
 * always ordered by ticket (otherwise only when the grid does not fit the
 * device at once).  Results do not depend on it.  A failed exchange is
 * reported by the next read of register 0 and switches the fused form off.
 * The autotuner (csize 0) chooses it when it is faster.  No counterpart in the
 * reference. */
int ct_engine_set_fused(ct_engine *e, int on);

/* Multipass engine: the workgroup form of row pass B.  form 0: one workgroup
 * per row and prime; 1: one workgroup per row that transforms its primes one
 * after the other, loading the next prime's row while it transforms the
 * current one; 2: one workgroup per row with one slice of threads per prime
 * (used where the slices are whole wavefronts and fit 1024 threads, else
 * form 0).  Results do not depend on it.  No counterpart in the reference. */
int ct_engine_set_row_form(ct_engine *e, int form);

/* Resident engine: run the two primes of a two-prime squaring in the two
 * halves of one workgroup twice the size (on = 1), for transform lengths
 * 2^10 .. 2^13, where each half is whole wavefronts; the first half then runs
 * the CRT and the carry.  Ignored for other lengths and with three primes.
 * Results do not depend on it.  No counterpart in the reference. */
int ct_engine_set_prime_halves(ct_engine *e, int on);

/* What the autotuner chose and measured: the column tile (0 for resident
 * kernels), the split of the column passes, the fused column pass, the row
 * form (0, 1 or 2, ct_engine_set_row_form), and ms per squaring (-1 where not
 * timed or when the engine was not tuned) in ms[CT_TUNE_TIMINGS]: {resident,
 * multipass W=16, W=32, W=16 split, W=32 split, W=16 fused, W=32 fused, W=16
 * fused split, W=32 fused split, the fastest of these with row form 1, with
 * row form 2, resident with two primes one after the other, resident with
 * two primes in workgroup halves}.  For resident engines, split reports the
 * primes in workgroup halves (ct_engine_set_prime_halves).  Any output may be
 * null. */
#define CT_TUNE_TIMINGS 13
int ct_engine_tuning(const ct_engine *e, uint32_t *col_w, int *split, int *fused, int *row_form, float *ms);
int ct_engine_destroy(ct_engine *e);

/* transform::set_b (transform.h:472-486): b[lanes], 32 < b < 2^31. */
int ct_engine_set_b(ct_engine *e, const uint32_t *b);

/* engine::set (engine.h:231-237, kernel.cl:514-520): register 0 := a. */
int ct_engine_set(ct_engine *e, uint32_t a);

/* engine::copy (engine.h:240-247): register dst := register src (0..2). */
int ct_engine_copy(ct_engine *e, uint32_t dst, uint32_t src);

/* PRP base (2, the default and the reference's, or 3): a set bit of the
 * squaring steps below multiplies by it, so ct_prp_check computes
 * base^(N-1) mod N.  No counterpart in the reference (base 2 is fixed in its
 * normalize1, kernel.cl:1044). */
int ct_engine_set_base(ct_engine *e, uint32_t base);

/* engine::squareDup (engine.h:387-419): register 0 := x^2 * base^bit(lane),
 * bit of lane l = (dup[l / 64] >> (l % 64)) & 1; dup may be NULL (no doubling). */
int ct_engine_square_dup(ct_engine *e, const uint64_t *dup);

/* engine::mul (engine.h:421-433): register 0 := register 0 * register 1.
 * Unlike the reference, register 1 is preserved. */
int ct_engine_mul(ct_engine *e);

/* Same product into any register: dst := a * b. */
int ct_engine_mul_regs(ct_engine *e, uint32_t dst, uint32_t a, uint32_t b);

/* transform::getInt (transform.h:501-597): canonical digits in [0, b) of
 * register reg (the reference always reads device register 0) as
 * digits[lane * n + j]. */
int ct_engine_get_int(ct_engine *e, uint32_t reg, uint32_t *digits);

/* transform::isPrime (transform.h:600-636) over canonical digits:
 * prm[lane] = (value == 1), res64[lane] = value mod 2^64; returns
 * CT_ERR_COMPUTE when some value is 0.  Host-only. */
int ct_engine_is_prime(const ct_engine *e, const uint32_t *digits, int *prm, uint64_t *res64);

/* Raw signed-digit state of a register (readMemory_x123 / writeMemory_x123,
 * engine.h:204-225): digits[lane * n + j], the same non-canonical digits as
 * the reference's device registers for the same chunk size. */
int ct_engine_read_state(ct_engine *e, uint32_t reg, int32_t *digits);
int ct_engine_write_state(ct_engine *e, uint32_t reg, const int32_t *digits);

/* Device-resident exponent bits: bits[lane * words + w], bit i of lane l is
 * (bits[l * words + i / 32] >> (i % 32)) & 1.  Replaces the per-iteration
 * host mpz_tstbit loop of cyclo.h:244-249. */
int ct_engine_set_exponents(ct_engine *e, const uint32_t *bits, uint32_t words);

/* The PRP loop of cyclo.h:277-308 on the device: for i = i_hi down to i_lo,
 * register 0 := x^2 * base^(bit i); when gl_block > 0, after the step for bit i
 * with i % gl_block == 0 and i / gl_block != 0, register 1 := register 0 *
 * register 1 (Gerbicz-Li d(t)). */
int ct_engine_run(ct_engine *e, int i_hi, int i_lo, int gl_block);

/* cyclo::check for one batch (cyclo.h:252-411): Fermat PRP (the engine's base,
 * 2 by default) of b^n - b^(n/2) + 1 (or - 1 for an N- engine) for every lane with
 * Gerbicz-Li verification.  Outputs
 * per lane: prm (probable prime), res64 and gl_ok; the canonical residues
 * (optional, may be NULL) as residue[lane * n + j].  Returns CT_ERR_GERBICZ if
 * any lane fails verification, CT_ERR_COMPUTE if any residue is 0. */
int ct_prp_check(ct_engine *e, const uint32_t *b, int *prm, uint64_t *res64, int *gl_ok, uint32_t *residue);

/* ct_prp_check split at its loop, as cyclo::check is split by its
 * checkpoints (cyclo.h:256-276 before the loop, :310-377 after it):
 * ct_prp_begin sets b and uploads the exponent bits of E = N - 1 (GMP,
 * cyclo.h:258-271), returning the top bit index i0 and the Gerbicz-Li block B
 * (cyclo.h:272).  The caller then sets register 0 := 1, copies it to
 * register 1 (d(t) = 1) and runs ct_engine_run(e, i0, 0, B), in as many
 * pieces as it likes.  ct_prp_finish is the tail of cyclo::check (the last
 * d(t) product, the Gerbicz-Li verification run, RES64 and the verdict) with
 * the outputs and status of ct_prp_check; a job is finished once
 * (CT_ERR_ARG the second time) and freed with ct_prp_job_free. */
typedef struct ct_prp_job ct_prp_job;
int ct_prp_begin(ct_engine *e, const uint32_t *b, ct_prp_job **job, int *i0, int *gl_block);
int ct_prp_finish(ct_engine *e, ct_prp_job *job, int *prm, uint64_t *res64, int *gl_ok, uint32_t *residue);
int ct_prp_job_free(ct_prp_job *job);

/* Wall time of the device work issued since the last call (ms), measured with
 * HIP events on the engine's stream. */
int ct_engine_timer_start(ct_engine *e);
int ct_engine_timer_stop(ct_engine *e, float *ms);

/* Number of NTT primes per squaring.  The engine uses two primes when every
 * coefficient provably fits (3 n D^2 < p1 p2 / 2, D = max(b, |digits set or
 * written|)), else three (the reference always uses three, transform.h:26-29);
 * the carried digits are identical either way.  np = 3 forces three primes,
 * np = 0 restores the automatic choice. */
int ct_engine_set_primes(ct_engine *e, int np);
int ct_engine_primes(const ct_engine *e, int *np);

/* Multipass engine: launch runs of squarings as replays of captured HIP
 * graphs (on = 1, the default) or as individual kernel launches (on = 0);
 * the results are identical.  No counterpart in the reference, which enqueues
 * every kernel (engine.h:387-419).  No effect on the resident kernels, which
 * run a whole run in one launch. */
int ct_engine_set_graphs(ct_engine *e, int on);

/* ct_engine_set_graphs for every engine created after the call, the
 * autotuner's timing engines included (default on).  Off keeps HIP graphs out
 * of a process altogether, e.g. one traced by a profiler. */
int ct_set_graphs_default(int on);

/* Engine facts: n, lanes, kind actually used. */
int ct_engine_info(const ct_engine *e, uint32_t *n, uint32_t *lanes, int *kind);

/* The engine's NTT primes p[0] > p[1] > p[2] (host only, no device needed).
 * The reference fixes its primes as #defines (transform.h:26-29); this build
 * uses primes below 2^30 so that its butterflies can stay lazy.  Two primes
 * serve when 3 n D^2 < p[0] p[1] / 2 (see ct_engine_set_primes). */
int ct_engine_moduli(uint32_t *p);

/* ---------------- serving loop ---------------- */

/* Continuous batching of PRP-loop requests of mixed sizes.  Replaces the
 * batch loop of cyclo::check(filename) (cyclo.h:414-485), which groups the
 * candidates of one n into fixed 64-lane batches (parseFile, cyclo.h:216-241)
 * and runs one batch at a time.  A request (id, 2^log2_n digits, b, steps)
 * yields x = 2^(top `steps` bits of N - 1) mod N, N = b^n - b^(n/2) + 1.
 * Requests are admitted into free lanes of per-size engines as lanes free up;
 * a new size class is created only within hbm_budget bytes of device memory.
 * slots = lanes per size class; chunk = squarings per launch of a class whose
 * transform has up to 2^14 points (larger transforms, of 2^tl points, launch
 * chunk / 2^(tl - 14), at least 4; the transform of an N- class is twice its
 * digit count); max_steps bounds `steps`.  10 <= log2_n <= 22. */
typedef struct ct_server ct_server;
int ct_server_create(int device, uint64_t hbm_budget, uint32_t slots, uint32_t chunk, uint32_t max_steps, ct_server **out);
int ct_server_destroy(ct_server *s);
int ct_server_submit(ct_server *s, uint64_t id, uint32_t log2_n, uint32_t b, uint32_t steps);
/* ct_server_submit for either twin member: sign = +1 is the request above,
 * sign = -1 asks for x = 2^(top `steps` bits of N - 1) mod N with
 * N = b^n - b^(n/2) - 1 (10 <= log2_n <= 21), in size classes of their own
 * (engines of ct_engine_create_ex).  No counterpart in the reference: its
 * batch loop tests N+ only, and its sieve defines N- (ctsieve.cpp:728-787). */
int ct_server_submit_ex(ct_server *s, uint64_t id, uint32_t log2_n, uint32_t b, uint32_t steps, int sign);
/* One step: read out the size classes whose chunk has completed, admit
 * queued requests into free lanes of classes that are not running a chunk,
 * launch the next chunk on every idle class with work (their streams
 * overlap), then wait until at least one class's chunk completes -- not all
 * of them, so a small class is never paced by a large one. */
int ct_server_step(ct_server *s);
/* Finished requests since the last poll: id, RES64 of the canonical residue,
 * a 64-bit FNV-1a hash of its digits, latency from submit (s). */
int ct_server_poll(ct_server *s, uint64_t *ids, uint64_t *res64, uint64_t *hash, double *latency_s, uint32_t max,
                   uint32_t *count);
int ct_server_stats(const ct_server *s, uint32_t *queued, uint32_t *active, uint64_t *hbm_used, uint64_t *squarings);
/* ct_engine_set_graphs for every size class of the server (default on). */
int ct_server_set_graphs(ct_server *s, int on);

/* ---------------- window sieve ---------------- */

/* Sieve::check_pos / check_neg (sieve/CPU/ctsieve.cpp:678-787) on the GPU:
 * strike every b in [b_min, b_max] with a prime factor p of
 * b^(2^m) - b^(2^(m-1)) + 1 (p = 3k 2^m + 1 for k from (p_min_pos / 2^m) / 3
 * up to the first p >= p_max_pos; mode >= 0) or of b^(2^m) - b^(2^(m-1)) - 1
 * (p = 10k +/- 1 for k from p_min_neg / 10, same bounds rule with p_max_neg;
 * mode <= 0).  struck[b - b_min] is read and updated (1 = struck); p bounds
 * below 2^62 (the reference's ini bounds, ctsieve.cpp:797-801, and beyond),
 * b_max - b_min < 2^32.  Primes below 2^32 run through 32-bit kernels, the
 * others through the 64-bit ones of the factor emission.  The survivor list
 * is the set of unstruck b in ascending order (cand_*.txt of
 * ctsieve.cpp:621-626).  The device buffers (about 190 MiB, plus the byte map
 * of the largest window seen) and the stream are created by the first call on
 * a device and kept until the process ends; calls on one device run one at a
 * time. */
int ct_sieve_window(int device, int m, uint64_t b_min, uint64_t b_max, uint64_t p_min_pos, uint64_t p_max_pos,
                    uint64_t p_min_neg, uint64_t p_max_neg, int mode, uint8_t *struck);

/* Device time of the last ct_sieve_window call on a device, from HIP events
 * recorded on its stream (a measurement aid; no counterpart in the
 * reference): roots_neg_ms sums the launches of the N- root kernel
 * (roots_neg_launches of them, at most 32 timed), device_ms spans the call's
 * device work from the upload of the byte map to the end of its download,
 * kernels_ms the part between the two copies. */
typedef struct
{
	float roots_neg_ms;
	uint32_t roots_neg_launches;
	float device_ms;
	float kernels_ms;
} ct_sieve_timing;
int ct_sieve_last_timing(int device, ct_sieve_timing *out);

/* GPU factor emission: the work of `ctsieveg <n> <p_min> <p_max> <+|->`
 * (sieve/GPU/src/main.cpp:113-169, gfsieve::check, ctsieve.h:326-431, kernels
 * sieve/GPU/ocl/kernel.cl:243-479).  Every pair (p, b) with p prime in
 * [p_min, p_max] of the form 3k 2^m + 1 (mode +1) or 10k +/- 1 (mode -1),
 * 0 <= b < p a root of b^(2^m) - b^(2^(m-1)) +/- 1 mod p and b <= bMax(p) (5e9;
 * 1e10 above 1e16 (+) or 1e13 (-)), every pair re-verified on the host as in
 * saveFactors (ctsieve.h:270-289).  pairs[2i] = p, pairs[2i+1] = b, unordered
 * (the reference appends with atomics too); *count = number of pairs;
 * CT_ERR_NOMEM when it exceeds cap (count then holds the size needed).
 * p_max < 2^63. */
int ct_sieve_factors(int device, int m, uint64_t p_min, uint64_t p_max, int mode, uint64_t *pairs, uint64_t cap,
                     uint64_t *count);

/* ct_sieve_factors that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (serving loops that must not
 * allocate, harnesses that forbid hipMalloc); the call itself reads counters
 * back and synchronises its stream, so it is not for use inside a graph
 * capture.  pairs is a
 * DEVICE buffer of 2 * cap uint64 words owned by the caller; the same pairs
 * as ct_sieve_factors (unordered), re-verified on the device instead of on
 * the host; *count = number of pairs, CT_ERR_NOMEM when it exceeds cap (the
 * first cap pairs are written).  stream: a hipStream_t (NULL: the default
 * stream); the call returns after the stream has finished its work.  Scratch:
 * a fixed 16 MiB + 256 B device array per device (the prime list of one block
 * of 2^20 k values and the counters), kept until the process ends; the p range
 * is processed block by block, so every range below 2^63 fits it.  Every call
 * on a device shares that array: calls that overlap on one device race. */
int ct_sieve_factors_noalloc(int device, int m, uint64_t p_min, uint64_t p_max, int mode, uint64_t *pairs,
                             uint64_t cap, uint64_t *count, void *stream);

/* Block cursor over the GPU factor emission, the per-block driving of
 * gfsieve::check (sieve/GPU/src/ctsieve.h:366-409) and engine.h:188-220
 * (generatePrimes / initFactors / generateFactors / readFactors per block):
 * block i holds k in [i 2^21, (i + 1) 2^21), p = 3 k 2^m + 1 (mode +1) or
 * p = 10 k -+ 1 (mode -1); the block range [i_min, i_max) of a p range is
 * the reference's (ctsieve.h:375-377).  ct_fsieve_run emits the pairs of
 * nblocks blocks from i_begin with p in [p_min, p_max] and b <= bMax(p) (or
 * b <= bmax after ct_fsieve_set_bmax(s, bmax) with bmax != 0; 2^64 - 1 emits
 * every root), re-verified on the device
 * (the check of saveFactors, ctsieve.h:270-289), unordered; CT_ERR_NOMEM when
 * they exceed cap (count = the size needed), at most 2^24 pairs per call. */
typedef struct ct_fsieve ct_fsieve;
int ct_fsieve_create(int device, int m, uint64_t p_min, uint64_t p_max, int mode, ct_fsieve **out);
int ct_fsieve_blocks(const ct_fsieve *s, uint64_t *i_min, uint64_t *i_max);
int ct_fsieve_set_bmax(ct_fsieve *s, uint64_t bmax);
int ct_fsieve_run(ct_fsieve *s, uint64_t i_begin, uint64_t nblocks, uint64_t *pairs, uint64_t cap, uint64_t *count);
int ct_fsieve_destroy(ct_fsieve *s);

typedef struct
{
	uint64_t blocks_done;   /* blocks of [i_min, i_max) done, this call and the checkpoint's */
	uint64_t blocks_total;  /* i_max - i_min */
	uint64_t factors;       /* pairs written in this call */
	double elapsed;         /* seconds, including the time restored from the checkpoint */
	int resumed;            /* 1 if the call resumed from its checkpoint */
	int interrupted;        /* 1 if it stopped on ct_fsieve_quit() or max_blocks */
} ct_fsieve_stats;

/* gfsieve::check (ctsieve.h:326-431), `ctsieveg <m> <p_min> <p_max> <+|->`:
 * appends "p b" lines (sorted by p, then b, within each read) to
 * <work_dir>/fp<ext> (mode +1) or fn<ext> (mode -1), ext =
 * "<m>_<p_min / 10^12>_<p_max / 10^12>.txt", and keeps the text checkpoint
 * ctx<ext> / ctx_n<ext> (blocks done, elapsed seconds) after every read,
 * reading every j_sync blocks with j_sync adapted to record_s seconds and
 * half of a 2^24-pair buffer (ctsieve.h:389-398).  Resumes from the
 * checkpoint.  Stops with CT_INTERRUPTED after ct_fsieve_quit() or max_blocks
 * blocks in this call (0 = no limit). */
int ct_fsieve_check(int device, int m, uint64_t p_min, uint64_t p_max, int mode, const char *work_dir, double record_s,
                    uint64_t max_blocks, ct_fsieve_stats *stats);
/* Asks a running ct_fsieve_check to stop after its current read (the role of
 * the reference's quit flag, ctsieve.h:166-170). */
void ct_fsieve_quit(void);

/* The ctools merge (tools/ctools.cpp:151-248, read_cand / read_factors /
 * write_cand): removes from a candidate list every b with a known factor.
 * - cand_path: one b per line; a b not above the previous one is skipped.
 * - factor_paths[i] (mode modes[i]: +1 for fp files, p | N+; -1 for fn
 *   files, p | N-): "p b" lines; a line repeating the previous pair is
 *   skipped.  A pair with b_min <= b <= b_max is verified on the device
 *   (x = b^(2^(m-1)) mod p, x^2 - x = -1 for +, +1 for -, p < 2^64) and, if
 *   it holds, removes b.  p outside [p_min_t[i], p_max_t[i]] x 10^12 is counted
 *   as out of range and still used; p_min_t / p_max_t may be NULL.
 * - out_path: the remaining candidates, one per line, ascending.
 * stats may be NULL.  A pair that fails the check is counted, not fatal, as
 * in the reference. */
typedef struct
{
	uint64_t candidates;    /* candidates read */
	uint64_t pairs;         /* pairs read, repeated lines dropped */
	uint64_t in_window;     /* pairs with b_min <= b <= b_max */
	uint64_t factors;       /* pairs that hold */
	uint64_t bad;           /* pairs in the window that do not hold */
	uint64_t out_of_range;  /* pairs with p outside their file's p range */
	uint64_t remaining;     /* candidates written */
} ct_merge_stats;
int ct_merge_factors(int device, int m, uint64_t b_min, uint64_t b_max, const char *cand_path, int nfiles,
                     const char *const *factor_paths, const int *modes, const uint64_t *p_min_t,
                     const uint64_t *p_max_t, const char *out_path, ct_merge_stats *stats);

/* ---------------- files and the batch driver ---------------- */

/* CRC-32 (IEEE 802.3, reflected) continued over len more bytes; the files
 * below end with ~crc ^ 0xa23777ac of all preceding bytes (file.h:80-121,
 * ctsieve.cpp:497-529). */
uint32_t ct_crc32(uint32_t crc, const void *buf, size_t len);

/* Sieve::write / Sieve::read (sieve/CPU/ctsieve.cpp:551-629): the state file
 * <dir>/sv_<m>_<b_min>_<b_max>.dat = [p_min_pos u64][p_min_neg u64][count u64]
 * [index u32 of every unstruck b - b_min][crc], the previous one rotated to
 * .old; with write_cand, also cand_<m>_<b_min>_<b_max>.txt (b per line,
 * ascending).  struck[b - b_min] as in ct_sieve_window; dir may be NULL. */
int ct_sieve_save(const char *dir, int m, uint64_t b_min, uint64_t b_max, uint64_t p_min_pos, uint64_t p_min_neg,
                  const uint8_t *struck, int write_cand);
int ct_sieve_load(const char *dir, int m, uint64_t b_min, uint64_t b_max, uint64_t *p_min_pos, uint64_t *p_min_neg,
                  uint8_t *struck);

/* cyclo::saveContext / readContext (cyclo.h:107-176, transform.h:445-468):
 * [int32 version 1][int32 k][int32 i][double elapsed][registers 0 and 1 as the
 * reference's device buffers: x12 (residues of each carried digit modulo the
 * reference primes P1, P2; index (reg n + j) lanes + lane) then x3 (modulo
 * P3)][crc]; written via path.new with the previous file rotated to
 * path.old; load falls back to path.old.  With 64 lanes the file is the
 * reference's own cyclo.ctx (k = batch, i = next exponent bit). */
int ct_context_save(ct_engine *e, const char *path, int k, int i, double elapsed);
int ct_context_load(ct_engine *e, const char *path, int *k, int *i, double *elapsed);

typedef struct
{
	uint64_t candidates;   /* distinct b tested in this call */
	uint64_t prps;         /* probable primes found */
	uint64_t squarings;    /* PRP-loop squarings run in this call (per lane) */
	double elapsed;        /* seconds, including the time restored from a checkpoint */
	int resumed;           /* 1 if the call resumed from cyclo.ctx */
} ct_cyclo_stats;

/* cyclo::check(filename) (cyclo.h:414-485): base-2 PRP with Gerbicz-Li of
 * b^(2^log2_n) - b^(2^(log2_n - 1)) + 1 for every b of cand_file (one per line),
 * lanes at a time (the reference: 64), appending
 * "<b>^<n> - <b>^<n/2> + 1 is a probable prime." or "... is composite. RES64:
 * <16 hex>" to <work_dir>/result_c.txt and probable primes to
 * <work_dir>/prp_<log2_n>.txt.  Checkpoints <work_dir>/cyclo.ctx every
 * checkpoint_s seconds at Gerbicz-Li points and resumes from it.  Stops with
 * the checkpoint saved and returns CT_INTERRUPTED after ct_cyclo_quit() or
 * once max_steps squarings ran in this call (0 = no limit).  On completion the
 * checkpoint files are removed.  stats may be NULL. */
int ct_cyclo_check_file(int device, uint32_t log2_n, const char *cand_file, const char *work_dir, uint32_t lanes,
                        double checkpoint_s, uint64_t max_steps, ct_cyclo_stats *stats);
/* Asks a running ct_cyclo_check_file to stop at its next segment boundary
 * (async-signal-safe; the role of the reference's SIGINT handler). */
void ct_cyclo_quit(void);

/* Last error text of this thread (static storage). */
const char *ct_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
