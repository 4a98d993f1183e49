// Host side of the squaring engine: device buffers, kernel sequencing and the
// C ABI of include/ctwin.h (the drop-in for cyclo/src/engine.h and the
// engine-facing part of cyclo/src/transform.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ctwin.h"
#include "ct_arith.h"
#include "ct_tables.h"
#include "ct_kernels.h"
#include "ct_engine_impl.h"

using namespace ct;

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_err;

const char *ct_last_error(void) { return g_err.c_str(); }

int ct::fail(int code, const char *what)
{
	g_err = what;
	return code;
}

int ct::hip_fail(hipError_t e, const char *where)
{
	g_err = std::string(where) + ": " + hipGetErrorString(e);
	return CT_ERR_HIP;
}

#define HIPCHK(call) do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail(_e, #call); } while (0)

// ---------------------------------------------------------------------------
// engine lifetime

// Graph replay for engines created from now on (the autotuner's timing
// engines included); ct_engine_set_graphs changes one engine.
static std::atomic<int> g_graphs_default{1};

int ct_set_graphs_default(int on)
{
	g_graphs_default.store(on != 0 ? 1 : 0);
	return CT_OK;
}

static int create_impl(int device, uint32_t log2_n, uint32_t lanes, int kind, uint32_t col_w, bool split, int sign,
                       ct_engine **out)
{
	const uint32_t tln = log2_n + (sign < 0 ? 1 : 0);
	if (kind == CT_KIND_AUTO) kind = (tln <= 14) ? CT_KIND_RESIDENT : CT_KIND_PLAIN;
	if (kind == CT_KIND_RESIDENT && tln > 14) return fail(CT_ERR_UNSUPPORTED, "resident kernels need a transform length <= 2^14");
	if (kind == CT_KIND_PLAIN && (tln < 10 || (sign < 0 && tln < 11)))
		return fail(CT_ERR_UNSUPPORTED, "multipass engine needs log2_n >= 10");
	if (kind != CT_KIND_RESIDENT && kind != CT_KIND_PLAIN) return fail(CT_ERR_ARG, "unknown engine kind");
	int ndev = 0;
	HIPCHK(hipGetDeviceCount(&ndev));
	if (device < 0 || device >= ndev) return fail(CT_ERR_HIP, "no such HIP device");
	HIPCHK(hipSetDevice(device));

	ct_engine *e = new ct_engine();
	e->device = device; e->ln = log2_n; e->n = 1u << log2_n; e->lanes = lanes; e->csize = 16; e->kind = kind;
	e->sign = (sign < 0) ? -1 : 1; e->tln = tln; e->tn = 1u << tln;
	e->col_w = col_w; e->split_a = split;
	e->use_graphs = g_graphs_default.load() != 0;
	e->gc = make_garner();
	e->b.assign(lanes, 0);
	const size_t reg_words = (size_t)lanes * e->n;
	int rc = CT_OK;
	auto alloc = [&](void **p, size_t bytes) {
		if (rc != CT_OK) return;
		hipError_t h = hipMalloc(p, bytes);
		if (h != hipSuccess) rc = fail(CT_ERR_NOMEM, "hipMalloc failed");
	};
	alloc((void **)&e->d_x, 3 * reg_words * sizeof(int32_t));
	alloc((void **)&e->d_tw, 6 * (size_t)e->tn * sizeof(tw_entry));
	alloc((void **)&e->d_dv, lanes * sizeof(divisor));
	alloc((void **)&e->d_mask, ((lanes + 31) / 32) * sizeof(uint32_t));
	if (rc == CT_OK && kind == CT_KIND_PLAIN) rc = plain_alloc(e);
	if (rc != CT_OK) { ct_engine_destroy(e); return rc; }
	hipError_t h = hipStreamCreate(&e->st);
	if (h == hipSuccess) h = hipEventCreate(&e->ev0);
	if (h == hipSuccess) h = hipEventCreate(&e->ev1);
	if (h != hipSuccess) { ct_engine_destroy(e); return hip_fail(h, "stream/event create"); }
	const std::vector<tw_entry> tw = build_tables(e->tn);
	h = hipMemcpy(e->d_tw, tw.data(), tw.size() * sizeof(tw_entry), hipMemcpyHostToDevice);
	if (h == hipSuccess) h = hipMemset(e->d_x, 0, 3 * reg_words * sizeof(int32_t));
	if (h != hipSuccess) { ct_engine_destroy(e); return hip_fail(h, "table upload"); }
	*out = e;
	return CT_OK;
}

// ---------------------------------------------------------------------------
// autotuner (the role of the CSIZE autotuner, transform.h:383-419): for a
// transform length and lane count, times squarings of a synthetic state under
// every compiled configuration -- resident kernels vs the multipass engine
// (n <= 2^14; one resident workgroup per candidate loses to the whole-GPU
// multipass launch when there are few candidates), the multipass column tile
// (16 or 32 columns per workgroup), column pass A combined or split by prime
// and, for N+ with 256-row columns, pass C fused with the next pass A -- and
// keeps the fastest.  Results are cached per (device, log2 n, lanes, kind,
// sign, prime count) for the life of the process.  An engine is tuned at
// creation for three primes (b is not known yet); when set_b or set_primes
// later moves a tuned multipass engine to two primes (or back), its column
// tile, split and fused form are tuned again for that prime count.

namespace {

struct tune_key { int device; uint32_t ln, lanes; int kind, sign, np;
	bool operator<(const tune_key &o) const
	{ return std::tie(device, ln, lanes, kind, sign, np) < std::tie(o.device, o.ln, o.lanes, o.kind, o.sign, o.np); } };
struct tune_val { int kind; uint32_t col_w; bool split, fused; int row_form; bool res_halves; float ms[CT_TUNE_TIMINGS]; };
std::mutex g_tune_mu;
std::map<tune_key, tune_val> g_tune;

constexpr int kTuneSteps = 32, kTuneReps = 4;

// A timed multipass configuration and its slot in tune_val::ms.
struct plain_cand { uint32_t w; bool split, fused; int form, slot; float ms; };

// ms per squaring of engine e with np primes, or a negative value on failure:
// the fastest of kTuneReps timed runs of kTuneSteps squarings after a warm-up
// run of the same length (which also captures the graph of that length).
// Three primes: b = 2^30 + 9699690 i as in transform.h:391 (but below 2^31);
// two primes: b = 65537 + 2 i, for which 4 n b^2 < p1 p2 / 2 up to n = 2^22.
float time_squarings(ct_engine *e, int np)
{
	std::vector<uint32_t> b(e->lanes);
	for (uint32_t i = 0; i < e->lanes; ++i) b[i] = (np == 3) ? (1u << 30) + 9699690u * (i % 100) : 65537u + 2u * (i % 100);
	if (ct_engine_set_b(e, b.data()) || ct_engine_set(e, 3)) return -1;
	if (e->np != np) return -1;
	if (ct::steps(e, nullptr, 0, 0, kTuneSteps, nullptr)) return -1;
	float best = 1e30f;
	for (int r = 0; r < kTuneReps; ++r)
	{
		float ms = 0;
		if (ct_engine_timer_start(e) || ct::steps(e, nullptr, 0, 0, kTuneSteps, nullptr) || ct_engine_timer_stop(e, &ms))
			return -1;
		best = std::min(best, ms);
	}
	return best / kTuneSteps;
}

}   // namespace

static int tune(int device, uint32_t ln, uint32_t lanes, int kind, int sign, int np, tune_val *out)
{
	const tune_key key{device, ln, lanes, kind, sign, np};
	const uint32_t tln = ln + (sign < 0 ? 1 : 0);
	{
		std::lock_guard<std::mutex> g(g_tune_mu);
		auto it = g_tune.find(key);
		if (it != g_tune.end()) { *out = it->second; return CT_OK; }
	}
	tune_val v{CT_KIND_AUTO, 0, false, false, 0, false, {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}};
	float best = 1e30f;
	const bool try_res = (kind == CT_KIND_RESIDENT) || (kind == CT_KIND_AUTO && tln <= 14);
	const bool try_plain = (kind == CT_KIND_PLAIN) || (kind == CT_KIND_AUTO && tln >= (sign < 0 ? 11u : 10u));
	if (try_res)
	{
		ct_engine *e = nullptr;
		int rc = create_impl(device, ln, lanes, CT_KIND_RESIDENT, 0, false, sign, &e);
		if (rc) return rc;
		v.ms[0] = time_squarings(e, np);
		if (v.ms[0] >= 0 && resident_halves_ok((int)tln))
		{   // two primes (the configuration whose primes can run in workgroup
			// halves): one prime after the other, and in halves
			v.ms[11] = time_squarings(e, 2);
			e->res_halves = true;
			v.ms[12] = time_squarings(e, 2);
			e->res_halves = false;
			if (v.ms[11] >= 0 && v.ms[12] >= 0) v.res_halves = v.ms[12] < v.ms[11];
			else
			{   // not timed (the size takes three primes): the one-after-the-other form
				(void)hipGetLastError();
				v.ms[11] = v.ms[12] = -1;
			}
		}
		ct_engine_destroy(e);
		if (v.ms[0] < 0) return fail(CT_ERR_HIP, "autotuner: timing the resident kernels failed");
		if (v.ms[0] < best) { best = v.ms[0]; v.kind = CT_KIND_RESIDENT; v.col_w = 0; }
	}
	if (try_plain)
	{
		ct_engine *e = nullptr;
		int rc = create_impl(device, ln, lanes, CT_KIND_PLAIN, 0, false, sign, &e);
		if (rc) return rc;
		std::vector<plain_cand> cands;
		const uint32_t tiles[2] = {16, 32};
		for (int sp = 0; sp < 2; ++sp)
			for (int t = 0; t < 2; ++t)
			{
				if (!plain_tile_ok((int)tln, tiles[t])) continue;
				e->col_w = tiles[t]; e->split_a = (sp != 0);
				const float ms = time_squarings(e, np);
				if (ms < 0) { ct_engine_destroy(e); return fail(CT_ERR_HIP, "autotuner: timing the multipass engine failed"); }
				v.ms[1 + t + 2 * sp] = ms;
				cands.push_back({tiles[t], sp != 0, false, 0, 1 + t + 2 * sp, ms});
				if (ms < best) { best = ms; v.kind = CT_KIND_PLAIN; v.col_w = tiles[t]; v.split = (sp != 0); v.fused = false; }
			}
		if (sign > 0 && tln >= 13)
			for (int sp = 0; sp < (np == 2 ? 2 : 1); ++sp)       // sp: primes in separate workgroup halves
				for (int t = 0; t < 2; ++t)
				{
					if (!plain_tile_ok((int)tln, tiles[t])) continue;
					e->col_w = tiles[t]; e->split_a = (sp != 0); e->use_fused = true;
					float ms = time_squarings(e, np);
					// a failed exchange surfaces at the next read of register 0
					if (ms >= 0 && plain_materialize(e) != CT_OK) ms = -1;
					e->use_fused = false;
					if (ms < 0) continue;              // not a candidate
					v.ms[5 + t + 2 * sp] = ms;
					cands.push_back({tiles[t], sp != 0, true, 0, 5 + t + 2 * sp, ms});
					if (ms < best) { best = ms; v.kind = CT_KIND_PLAIN; v.col_w = tiles[t]; v.split = (sp != 0); v.fused = true; }
				}
		if (!plain_tile_ok((int)tln, 16) && !plain_tile_ok((int)tln, 32))
		{   // small n: one tile only
			const float ms = time_squarings(e, np);
			if (ms < 0) { ct_engine_destroy(e); return fail(CT_ERR_HIP, "autotuner: timing the multipass engine failed"); }
			v.ms[1] = ms;
			if (ms < best) { best = ms; v.kind = CT_KIND_PLAIN; v.col_w = plain_default_tile((int)tln); }
		}
		if (v.kind == CT_KIND_PLAIN)
		{   // the best multipass configuration with the other row-pass workgroup
			// forms: one workgroup per row with its primes in sequence (1), or in
			// slices of one workgroup (2, where the slices fit)
			const uint32_t T = (1u << (tln - plain_rows_log2((int)tln))) / 16;
			const tune_val base = v;
			for (int form = 1; form <= 2; ++form)
			{
				if (form == 2 && (T % 64 != 0 || (uint32_t)np * T > 1024)) continue;
				e->col_w = base.col_w; e->split_a = base.split; e->use_fused = base.fused; e->row_form = form; e->fused_w = 0;
				float ms = time_squarings(e, np);
				if (ms >= 0 && base.fused && plain_materialize(e) != CT_OK) ms = -1;
				e->row_form = 0; e->use_fused = false;
				if (ms >= 0)
				{
					v.ms[8 + form] = ms;
					cands.push_back({base.col_w, base.split, base.fused, form, 8 + form, ms});
					if (ms < best) { best = ms; v.row_form = form; }
				}
			}
			// the two fastest, timed again in turn: configurations a few percent
			// apart are not told apart reliably by one series each (the clock
			// and the caches drift between series)
			std::sort(cands.begin(), cands.end(), [](const plain_cand &x, const plain_cand &y) { return x.ms < y.ms; });
			if (cands.size() >= 2 && cands[1].ms < 1.1f * cands[0].ms)
			{
				for (int round = 0; round < 3; ++round)
					for (int c = 0; c < 2; ++c)
					{
						plain_cand &k = cands[c];
						e->col_w = k.w; e->split_a = k.split; e->use_fused = k.fused; e->row_form = k.form; e->fused_w = 0;
						float ms = time_squarings(e, np);
						if (ms >= 0 && k.fused && plain_materialize(e) != CT_OK) ms = -1;
						e->row_form = 0; e->use_fused = false;
						if (ms >= 0) k.ms = std::min(k.ms, ms);
					}
				const plain_cand &win = (cands[1].ms < cands[0].ms) ? cands[1] : cands[0];
				v.ms[cands[0].slot] = cands[0].ms;
				v.ms[cands[1].slot] = cands[1].ms;
				v.col_w = win.w; v.split = win.split; v.fused = win.fused; v.row_form = win.form;
			}
		}
		ct_engine_destroy(e);
	}
	if (v.kind == CT_KIND_AUTO) return fail(CT_ERR_UNSUPPORTED, "autotuner: no configuration for this size");
	{
		std::lock_guard<std::mutex> g(g_tune_mu);
		g_tune[key] = v;
	}
	*out = v;
	return CT_OK;
}

int ct_engine_create_ex(int device, uint32_t log2_n, uint32_t lanes, uint32_t csize, int kind, int sign, ct_engine **out)
{
	if (!out) return fail(CT_ERR_ARG, "out is null");
	*out = nullptr;
	if (sign != 1 && sign != -1) return fail(CT_ERR_ARG, "sign must be +1 or -1");
	const uint32_t tln = log2_n + (sign < 0 ? 1 : 0);
	if (log2_n < 5 || tln > (uint32_t)kMaxLog2N)
		return fail(CT_ERR_UNSUPPORTED, "log2_n must be in [5, 22] ([5, 21] for N-)");
	if (lanes == 0) return fail(CT_ERR_ARG, "lanes must be > 0");
	if (csize != 0 && csize != 16) return fail(CT_ERR_UNSUPPORTED, "csize must be 0 or 16");
	if (kind != CT_KIND_AUTO && kind != CT_KIND_RESIDENT && kind != CT_KIND_PLAIN) return fail(CT_ERR_ARG, "unknown engine kind");
	if (kind == CT_KIND_RESIDENT && tln > 14) return fail(CT_ERR_UNSUPPORTED, "resident kernels need a transform length <= 2^14");
	if (kind == CT_KIND_PLAIN && (tln < 10 || (sign < 0 && tln < 11)))
		return fail(CT_ERR_UNSUPPORTED, "multipass engine needs log2_n >= 10");
	if (csize == 16) return create_impl(device, log2_n, lanes, kind, 0, false, sign, out);
	tune_val v;
	int rc = tune(device, log2_n, lanes, kind, sign, 3, &v);
	if (rc) return rc;
	rc = create_impl(device, log2_n, lanes, v.kind, v.col_w, v.split, sign, out);
	if (rc == CT_OK)
	{
		std::memcpy((*out)->tune_ms, v.ms, sizeof(v.ms));
		(*out)->use_fused = v.fused;
		(*out)->row_form = v.row_form;
		(*out)->res_halves = v.res_halves;
		(*out)->tuned = true;
		(*out)->tuned_np = 3;
	}
	return rc;
}

int ct_engine_create(int device, uint32_t log2_n, uint32_t lanes, uint32_t csize, int kind, ct_engine **out)
{
	return ct_engine_create_ex(device, log2_n, lanes, csize, kind, 1, out);
}

int ct_engine_set_column_tile(ct_engine *e, uint32_t w)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	if (e->kind != CT_KIND_PLAIN) return fail(CT_ERR_UNSUPPORTED, "column tiles belong to the multipass engine");
	if (!plain_tile_ok((int)e->tln, w)) return fail(CT_ERR_UNSUPPORTED, "tile width not compiled for this length");
	e->col_w = w;
	return CT_OK;
}

int ct_engine_set_fused(ct_engine *e, int on)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	if (e->kind != CT_KIND_PLAIN) return fail(CT_ERR_UNSUPPORTED, "column passes belong to the multipass engine");
	if (on < 0 || on > 2) return fail(CT_ERR_ARG, "on must be 0, 1 or 2");
	e->use_fused = (on != 0);
	e->fused_force_ticket = (on == 2);
	e->fused_w = 0;                    // start from cleared flags
	return CT_OK;
}

int ct_engine_set_row_form(ct_engine *e, int form)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	if (e->kind != CT_KIND_PLAIN) return fail(CT_ERR_UNSUPPORTED, "the row pass belongs to the multipass engine");
	if (form < 0 || form > 2) return fail(CT_ERR_ARG, "row form must be 0, 1 or 2");
	e->row_form = form;
	return CT_OK;
}

int ct_engine_set_prime_halves(ct_engine *e, int on)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	if (e->kind != CT_KIND_RESIDENT) return fail(CT_ERR_UNSUPPORTED, "prime halves belong to the resident kernels");
	e->res_halves = (on != 0);
	return CT_OK;
}

int ct_engine_set_column_split(ct_engine *e, int on)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	if (e->kind != CT_KIND_PLAIN) return fail(CT_ERR_UNSUPPORTED, "column passes belong to the multipass engine");
	e->split_a = (on != 0);
	return CT_OK;
}

int ct_engine_tuning(const ct_engine *e, uint32_t *col_w, int *split, int *fused, int *row_form, float *ms)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	if (col_w) *col_w = (e->kind == CT_KIND_PLAIN) ? e->col_w : 0;
	if (split) *split = ((e->kind == CT_KIND_PLAIN && e->split_a) || (e->kind == CT_KIND_RESIDENT && e->res_halves)) ? 1 : 0;
	if (fused) *fused = (e->kind == CT_KIND_PLAIN && e->use_fused) ? 1 : 0;
	if (row_form) *row_form = (e->kind == CT_KIND_PLAIN) ? e->row_form : 0;
	if (ms) for (int i = 0; i < CT_TUNE_TIMINGS; ++i) ms[i] = e->tune_ms[i];
	return CT_OK;
}

int ct_engine_destroy(ct_engine *e)
{
	if (!e) return CT_OK;
	hipSetDevice(e->device);
	if (e->st) hipStreamSynchronize(e->st);
	hipFree(e->d_x); hipFree(e->d_tw); hipFree(e->d_dv); hipFree(e->d_mask); hipFree(e->d_bits);
	plain_free(e);
	if (e->ev0) hipEventDestroy(e->ev0);
	if (e->ev1) hipEventDestroy(e->ev1);
	if (e->st) hipStreamDestroy(e->st);
	delete e;
	return CT_OK;
}

int ct_engine_info(const ct_engine *e, uint32_t *n, uint32_t *lanes, int *kind)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	if (n) *n = e->n;
	if (lanes) *lanes = e->lanes;
	if (kind) *kind = e->kind;
	return CT_OK;
}

int ct_engine_set_base(ct_engine *e, uint32_t base)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	if (base != 2 && base != 3) return fail(CT_ERR_ARG, "the PRP base must be 2 or 3");
	e->base = base;
	return CT_OK;
}

int ct_engine_set_graphs(ct_engine *e, int on)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	e->use_graphs = (on != 0);
	return CT_OK;
}

int ct_engine_moduli(uint32_t *p)
{
	if (!p) return fail(CT_ERR_ARG, "null pointer");
	for (int i = 0; i < kPrimes; ++i) p[i] = kP[i];
	return CT_OK;
}

// ---------------------------------------------------------------------------
// parameters and registers

// A tuned multipass engine whose prime count changed: tune its column tile,
// pass-A split and fused form for the new count (the engine kind stays).
// Tuning is an optimisation: if it cannot run (e.g. no device memory for
// its timing engines), the engine keeps its configuration and the call
// succeeds; the count is not tried again.
static int retune_for_primes(ct_engine *e)
{
	if (!e->tuned || e->kind != CT_KIND_PLAIN || e->np == e->tuned_np) return CT_OK;
	e->tuned_np = e->np;
	tune_val v;
	const int rc = tune(e->device, e->ln, e->lanes, CT_KIND_PLAIN, e->sign, e->np, &v);
	HIPCHK(hipSetDevice(e->device));
	if (rc != CT_OK)
	{
		(void)hipGetLastError();      // a failed allocation must not surface at the next launch check
		return CT_OK;
	}
	e->col_w = v.col_w; e->split_a = v.split; e->use_fused = v.fused; e->row_form = v.row_form; e->fused_w = 0;
	for (int i = 1; i < CT_TUNE_TIMINGS; ++i) e->tune_ms[i] = v.ms[i];
	return CT_OK;
}

int ct_engine_set_b(ct_engine *e, const uint32_t *b)
{
	if (!e || !b) return fail(CT_ERR_ARG, "null argument");
	std::vector<divisor> dv(e->lanes);
	for (uint32_t l = 0; l < e->lanes; ++l)
	{
		if (b[l] <= 32 || b[l] >= 0x80000000u) return fail(CT_ERR_ARG, "b must satisfy 32 < b < 2^31");
		e->b[l] = b[l];
		dv[l] = make_divisor(b[l]);
	}
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->st));
	HIPCHK(hipMemcpy(e->d_dv, dv.data(), dv.size() * sizeof(divisor), hipMemcpyHostToDevice));
	e->b_set = true;
	e->b_max = *std::max_element(e->b.begin(), e->b.end());
	e->state_bound = 1;     // registers must be (re)initialised for the new modulus
	choose_primes(e);
	return retune_for_primes(e);
}

int ct_engine_set(ct_engine *e, uint32_t a)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(launch_set(e->d_x, e->n, e->lanes, (int32_t)a, e->st));
	e->pending = false;
	e->state_bound = std::max<uint64_t>(e->state_bound, a);
	choose_primes(e);
	return CT_OK;
}

// Every coefficient of a square or product is bounded by 3 n D^2 (N+) or
// 4 n D^2 (N-), D >= |digit|.  N+: each product a_i a_k lands on at most two
// positions of X^n - X^(n/2) + 1 and at most 1.5 n pairs meet at one position.
// N-: the acyclic product has at most n D^2 per coefficient and the fold by
// X^n = X^(n/2) + 1 adds at most four of them (2 c_{j+n} counts twice).  The
// multiplication by the base happens after the CRT, on exact integers.  The
// carry leaves |digit| <= b, so D = max(b, digits written or set).  Two primes
// reconstruct exactly when that bound is below p1 p2 / 2.
void ct::choose_primes(ct_engine *e)
{
	const uint64_t D = std::max<uint64_t>(e->b_max, e->state_bound);
	const unsigned __int128 bound = (unsigned __int128)(e->sign < 0 ? 4 : 3) * e->n * D * D;
	e->np = (e->force3 || bound >= e->gc.h2) ? 3 : 2;
}

int ct_engine_set_primes(ct_engine *e, int np)
{
	if (!e || (np != 0 && np != 3)) return fail(CT_ERR_ARG, "np must be 0 (automatic) or 3");
	e->force3 = (np == 3);
	choose_primes(e);
	return retune_for_primes(e);
}

int ct_engine_primes(const ct_engine *e, int *np)
{
	if (!e || !np) return fail(CT_ERR_ARG, "bad argument");
	*np = e->np;
	return CT_OK;
}

// Brings register 0 to its carried form before any operation other than a
// squaring reads it (the plain engine defers the last carry step).
static int settle(ct_engine *e)
{
	if (e->kind == CT_KIND_PLAIN && e->pending) return plain_materialize(e);
	return CT_OK;
}

int ct_engine_copy(ct_engine *e, uint32_t dst, uint32_t src)
{
	if (!e || dst > 2 || src > 2) return fail(CT_ERR_ARG, "register must be 0, 1 or 2");
	if (dst == src) return CT_OK;
	HIPCHK(hipSetDevice(e->device));
	int rc = settle(e);
	if (rc) return rc;
	HIPCHK(hipMemcpyAsync(e->reg(dst), e->reg(src), e->reg_bytes(), hipMemcpyDeviceToDevice, e->st));
	return CT_OK;
}

int ct_engine_read_state(ct_engine *e, uint32_t reg, int32_t *digits)
{
	if (!e || !digits || reg > 2) return fail(CT_ERR_ARG, "bad argument");
	HIPCHK(hipSetDevice(e->device));
	int rc = settle(e);
	if (rc) return rc;
	HIPCHK(hipStreamSynchronize(e->st));
	HIPCHK(hipMemcpy(digits, e->reg(reg), e->reg_bytes(), hipMemcpyDeviceToHost));
	return CT_OK;
}

int ct_engine_write_state(ct_engine *e, uint32_t reg, const int32_t *digits)
{
	if (!e || !digits || reg > 2) return fail(CT_ERR_ARG, "bad argument");
	uint64_t dmax = 0;
	for (uint32_t l = 0; l < e->lanes; ++l)
		for (uint32_t j = 0; j < e->n; ++j)
		{
			const int64_t d = digits[(size_t)l * e->n + j];
			const uint64_t ad = (uint64_t)(d < 0 ? -d : d);
			if (e->b_set && ad > 2 * (uint64_t)e->b[l]) return fail(CT_ERR_ARG, "digits must satisfy |d| <= 2b");
			dmax = std::max(dmax, ad);
		}
	HIPCHK(hipSetDevice(e->device));
	int rc = settle(e);
	if (rc) return rc;
	e->state_bound = std::max(e->state_bound, dmax);
	choose_primes(e);
	HIPCHK(hipStreamSynchronize(e->st));
	HIPCHK(hipMemcpy(e->reg(reg), digits, e->reg_bytes(), hipMemcpyHostToDevice));
	return CT_OK;
}

int ct_engine_set_exponents(ct_engine *e, const uint32_t *bits, uint32_t words)
{
	if (!e || !bits || words == 0) return fail(CT_ERR_ARG, "bad argument");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->st));
	const size_t bytes = (size_t)e->lanes * words * sizeof(uint32_t);
	if (words != e->bit_words)
	{
		hipFree(e->d_bits); e->d_bits = nullptr; e->bit_words = 0;
		if (hipMalloc((void **)&e->d_bits, bytes) != hipSuccess) return fail(CT_ERR_NOMEM, "hipMalloc exponent bits");
		e->bit_words = words;
	}
	HIPCHK(hipMemcpy(e->d_bits, bits, bytes, hipMemcpyHostToDevice));
	return CT_OK;
}

// ---------------------------------------------------------------------------
// squarings and products

static int check_ready(const ct_engine *e)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	if (!e->b_set) return fail(CT_ERR_ARG, "set_b has not been called");
	return CT_OK;
}

int ct::steps(ct_engine *e, const uint32_t *bits, uint32_t bit_words, int i_hi, int nsteps, const uint32_t *mask_words,
              const int32_t *lane_i)
{
	if (nsteps <= 0) return CT_OK;
	if (e->kind == CT_KIND_PLAIN) return plain_steps(e, bits, bit_words, i_hi, nsteps, mask_words, lane_i);
	res_args A;
	A.x = e->reg(0); A.tw = e->d_tw; A.dv = e->d_dv;
	A.bits = bits; A.bit_words = bit_words; A.mask = mask_words; A.lane_i = lane_i;
	A.i_hi = i_hi; A.nsteps = nsteps; A.gc = e->gc; A.base_m1 = (int)e->base - 1;
	HIPCHK(resident_steps((int)e->tln, e->np, e->sign < 0, e->res_halves, A, e->lanes, e->st));
	return CT_OK;
}

int ct::mul_regs(ct_engine *e, uint32_t dst, uint32_t a, uint32_t b)
{
	if (e->kind == CT_KIND_PLAIN) return plain_mul(e, dst, a, b);
	mul_args A;
	A.a = e->reg(a); A.b = e->reg(b); A.out = e->reg(dst); A.tw = e->d_tw; A.dv = e->d_dv; A.gc = e->gc;
	HIPCHK(resident_mul((int)e->tln, e->np, e->sign < 0, A, e->lanes, e->st));
	return CT_OK;
}

int ct_engine_square_dup(ct_engine *e, const uint64_t *dup)
{
	int rc = check_ready(e);
	if (rc) return rc;
	HIPCHK(hipSetDevice(e->device));
	const uint32_t words = (e->lanes + 31) / 32;
	const uint32_t *mask = nullptr;
	if (dup)
	{
		std::vector<uint32_t> m(words, 0);
		for (uint32_t l = 0; l < e->lanes; ++l)
			if ((dup[l / 64] >> (l % 64)) & 1u) m[l / 32] |= 1u << (l % 32);
		HIPCHK(hipStreamSynchronize(e->st));
		HIPCHK(hipMemcpy(e->d_mask, m.data(), words * sizeof(uint32_t), hipMemcpyHostToDevice));
		mask = e->d_mask;
	}
	return steps(e, nullptr, 0, 0, 1, mask);
}

int ct_engine_mul(ct_engine *e) { return ct_engine_mul_regs(e, 0, 0, 1); }

int ct_engine_mul_regs(ct_engine *e, uint32_t dst, uint32_t a, uint32_t b)
{
	int rc = check_ready(e);
	if (rc) return rc;
	if (dst > 2 || a > 2 || b > 2) return fail(CT_ERR_ARG, "register must be 0, 1 or 2");
	HIPCHK(hipSetDevice(e->device));
	return mul_regs(e, dst, a, b);
}

int ct_engine_run(ct_engine *e, int i_hi, int i_lo, int gl_block)
{
	int rc = check_ready(e);
	if (rc) return rc;
	if (!e->d_bits) return fail(CT_ERR_ARG, "set_exponents has not been called");
	if (i_lo < 0 || i_hi < i_lo) return CT_OK;
	if ((uint64_t)i_hi >= (uint64_t)e->bit_words * 32) return fail(CT_ERR_ARG, "i_hi beyond the exponent bits");
	HIPCHK(hipSetDevice(e->device));
	int i = i_hi;
	while (i >= i_lo)
	{
		// next Gerbicz-Li point g <= i (i % B == 0, i / B != 0), cyclo.h:301
		int stop = i_lo;
		if (gl_block > 0)
		{
			const int g = (i / gl_block) * gl_block;
			if (g >= i_lo && g / gl_block != 0) stop = g;
		}
		rc = steps(e, e->d_bits, e->bit_words, i, i - stop + 1, nullptr);
		if (rc) return rc;
		if (gl_block > 0 && stop % gl_block == 0 && stop / gl_block != 0)
		{
			rc = mul_regs(e, 1, 0, 1);
			if (rc) return rc;
		}
		i = stop - 1;
	}
	return CT_OK;
}

// ---------------------------------------------------------------------------
// read-out

// Canonical digits of one lane, the algorithm of transform.h:501-597:
// floor-carry in base b, wrap X^n = X^(n/2) - 1 until no carry is left, then
// subtract N once when the value is >= N.
// N- = b^n - b^(n/2) - 1 (no counterpart in the reference): carry with
// floor division, wrap the top carry f as f (b^(n/2) + 1), then subtract N-
// once if the value reached it (digits of N-: b - 1 everywhere but b - 2 at n/2).
static int canonical_minus(const int32_t *d, uint32_t n, uint32_t b, uint32_t *x)
{
	auto fl = [b](int64_t &f) -> uint32_t {
		int64_t q = f / (int64_t)b, r = f % (int64_t)b;
		if (r < 0) { r += b; --q; }
		f = q;
		return (uint32_t)r;
	};
	int64_t f = 0;
	for (uint32_t j = 0; j < n; ++j) { int64_t v = f + d[j]; x[j] = fl(v); f = v; }
	int guard = 0;
	while (f != 0)
	{
		if (++guard > 64) return CT_ERR_GETINT;
		int64_t v = f;
		for (uint32_t j = 0; j < n / 2; ++j) { v += (int64_t)x[j]; x[j] = fl(v); }
		v += f;
		for (uint32_t j = n / 2; j < n; ++j) { v += (int64_t)x[j]; x[j] = fl(v); }
		f = v;
	}
	auto nd = [&](uint32_t j) -> uint32_t { return (j == n / 2) ? b - 2 : b - 1; };
	int cmp = 0;
	for (uint32_t j = n; j-- > 0;) if (x[j] != nd(j)) { cmp = (x[j] > nd(j)) ? 1 : -1; break; }
	if (cmp >= 0)
	{   // value - N-
		int64_t borrow = 0;
		for (uint32_t j = 0; j < n; ++j)
		{
			int64_t v = (int64_t)x[j] - nd(j) - borrow;
			borrow = (v < 0) ? 1 : 0;
			x[j] = (uint32_t)(v + (borrow ? (int64_t)b : 0));
		}
		if (borrow) return CT_ERR_GETINT;
	}
	return CT_OK;
}

int ct::canonical(const int32_t *d, uint32_t n, uint32_t b, int sign, uint32_t *x)
{
	if (sign < 0) return canonical_minus(d, n, b, x);
	auto fl = [b](int64_t &f) -> uint32_t {
		int64_t q = f / (int64_t)b, r = f % (int64_t)b;
		if (r < 0) { r += b; --q; }
		f = q;
		return (uint32_t)r;
	};
	int64_t f = 0;
	for (uint32_t j = 0; j < n; ++j) { int64_t v = f + d[j]; x[j] = fl(v); f = v; }
	int guard = 0;
	while (f != 0)
	{
		if (++guard > 64) return CT_ERR_GETINT;
		int64_t v = -f;
		for (uint32_t j = 0; j < n / 2; ++j) { v += (int64_t)x[j]; x[j] = fl(v); }
		v += f;
		for (uint32_t j = n / 2; j < n; ++j) { v += (int64_t)x[j]; x[j] = fl(v); if (v == 0) break; }
		f = v;
	}
	bool over = true;
	for (uint32_t j = n / 2; j < n; ++j) if (x[j] != b - 1) { over = false; break; }
	if (over)
	{
		bool low = false;
		for (uint32_t j = 0; j < n / 2; ++j) if (x[j] != 0) { low = true; break; }
		over = low;
	}
	if (over)
	{   // value - N = (low half) - 1, top half cleared
		int32_t borrow = 1;
		for (uint32_t j = 0; j < n / 2 && borrow; ++j)
		{
			if (x[j] >= (uint32_t)borrow) { x[j] -= borrow; borrow = 0; }
			else { x[j] = b - 1; }
		}
		if (borrow) return CT_ERR_GETINT;
		for (uint32_t j = n / 2; j < n; ++j) x[j] = 0;
	}
	return CT_OK;
}

int ct_engine_get_int(ct_engine *e, uint32_t reg, uint32_t *digits)
{
	if (!e || !digits || reg > 2) return fail(CT_ERR_ARG, "bad argument");
	if (!e->b_set) return fail(CT_ERR_ARG, "set_b has not been called");
	std::vector<int32_t> raw((size_t)e->lanes * e->n);
	int rc = ct_engine_read_state(e, reg, raw.data());
	if (rc) return rc;
	for (uint32_t l = 0; l < e->lanes; ++l)
	{
		rc = canonical(&raw[(size_t)l * e->n], e->n, e->b[l], e->sign, &digits[(size_t)l * e->n]);
		if (rc) return fail(CT_ERR_GETINT, "getInt failed");
	}
	return CT_OK;
}

int ct_engine_is_prime(const ct_engine *e, const uint32_t *digits, int *prm, uint64_t *res64)
{
	if (!e || !digits || !prm || !res64) return fail(CT_ERR_ARG, "bad argument");
	bool err = false;
	for (uint32_t l = 0; l < e->lanes; ++l)
	{
		const uint32_t *x = &digits[(size_t)l * e->n];
		bool one = (x[0] == 1), zero = (x[0] == 0);
		uint64_t r = x[0], bi = e->b[l];
		for (uint32_t j = 1; j < e->n; ++j)
		{
			one &= (x[j] == 0); zero &= (x[j] == 0);
			r += (uint64_t)x[j] * bi;
			bi *= e->b[l];
		}
		prm[l] = one ? 1 : 0; res64[l] = r;
		err |= zero;
	}
	return err ? fail(CT_ERR_COMPUTE, "computation failed") : CT_OK;
}

// ---------------------------------------------------------------------------
// timing

int ct_engine_timer_start(ct_engine *e)
{
	if (!e) return fail(CT_ERR_ARG, "null engine");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipEventRecord(e->ev0, e->st));
	return CT_OK;
}

int ct_engine_timer_stop(ct_engine *e, float *ms)
{
	if (!e || !ms) return fail(CT_ERR_ARG, "bad argument");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipEventRecord(e->ev1, e->st));
	HIPCHK(hipEventSynchronize(e->ev1));
	HIPCHK(hipEventElapsedTime(ms, e->ev0, e->ev1));
	return CT_OK;
}
