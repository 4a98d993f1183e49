// Serving loop: continuous batching of PRP-loop requests of mixed sizes.
//
// A request (id, n = 2^ln, b, t, sign) asks for x_t = 2^(top t bits of E) mod
// N, N = b^n - b^(n/2) + 1 (sign +1) or b^n - b^(n/2) - 1 (sign -1), E = N - 1:
// t squarings of the PRP loop of cyclo.h:277-308 from x = 1.  Requests are
// grouped by (n, sign) into size classes; each class owns one engine with
// `slots` lanes (resident kernels for transforms of up to 2^14 points -- N-
// runs its transform at length 2n -- the multipass engine above) and its own
// HIP stream.  A request is
// admitted into a free lane of its class as soon as one frees up (its digits,
// b, exponent window and bit offset are written for that lane only; the other
// lanes keep running), and a new class engine is created only while the sum
// of the engines' device allocations stays within the HBM budget; idle classes
// are released when a queued request needs the memory.  Every class runs on
// its own cadence: a chunk of min(chunk / 2^(tl - 14) (at least 4), shortest
// remaining) squarings, 2^tl the transform length, is launched on the class
// stream with an event behind it, and the class is read out, re-admitted and
// relaunched as soon as that event has completed, while the chunks of the
// other classes are still running (their kernels overlap on the device).  A
// small class is thus never held back by a large one.
#include <hip/hip_runtime.h>
#include <gmp.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/ctwin.h"
#include "ct_engine_impl.h"

using namespace ct;

namespace {

double now_s()
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct request { uint64_t id; uint32_t ln, b, steps; int sign; double t_submit; };

struct lane_state { bool busy = false; request req{}; int remaining = 0; };

struct size_class
{
	ct_engine *e = nullptr;
	uint32_t ln = 0, words = 0;
	int sign = 1;
	int32_t *d_lane_i = nullptr;
	std::vector<int32_t> lane_i;
	std::vector<lane_state> lanes;
	std::vector<uint32_t> stage_bits;
	std::vector<int32_t> stage_digits;
	std::vector<divisor> stage_dv;
	std::vector<int32_t> stage_lane_i;      // lane offsets of the chunk in flight (read by its upload)
	uint64_t bytes = 0;
	hipEvent_t ev = nullptr;                // recorded behind the chunk in flight
	bool in_flight = false;
	int k = 0;                              // squarings of the chunk in flight
};

struct result { uint64_t id, res64, hash; double latency; };

// Top k bits of E = b^n - b^(n/2) (sign +1) or b^n - b^(n/2) - 2 (sign -1):
// w = E >> s (certified truncated squaring of b; exact arithmetic when E is
// short).  Returns false when not certified.
bool exponent_window(uint32_t b, uint32_t n, uint32_t k, int sign, mpz_t w)
{
	const uint64_t ebits = (uint64_t)n * 32;
	if (ebits < 4 * (uint64_t)k + 4096)
	{
		mpz_t h, E; mpz_inits(h, E, NULL);
		mpz_ui_pow_ui(h, b, n / 2);
		mpz_mul(E, h, h); mpz_sub(E, E, h);
		if (sign < 0) mpz_sub_ui(E, E, 2);
		const size_t len = mpz_sizeinbase(E, 2);
		if (len > k) mpz_fdiv_q_2exp(w, E, len - k); else mpz_set(w, E);
		mpz_clears(h, E, NULL);
		return true;
	}
	const size_t P = k + 96;
	mpz_t lo, hi, t; mpz_inits(lo, hi, t, NULL);
	mpz_set_ui(lo, b); mpz_set_ui(hi, b);
	uint64_t e = 0;
	for (uint32_t m = n; m > 1; m >>= 1)
	{
		mpz_mul(lo, lo, lo); mpz_mul(hi, hi, hi); e *= 2;
		const size_t hb = mpz_sizeinbase(hi, 2);
		if (hb > P)
		{
			const size_t sh = hb - P;
			mpz_fdiv_q_2exp(lo, lo, sh); mpz_fdiv_q_2exp(hi, hi, sh); mpz_add_ui(hi, hi, 1);
			e += sh;
		}
	}
	bool ok = mpz_sizeinbase(lo, 2) == mpz_sizeinbase(hi, 2);
	const size_t sl = mpz_sizeinbase(hi, 2) - k;
	mpz_fdiv_q_2exp(w, lo, sl); mpz_fdiv_q_2exp(t, hi, sl);
	ok = ok && mpz_cmp(w, t) == 0;
	// b^n = w 2^(e + sl) + r with r >= t 2^e, t = lo mod 2^sl; subtracting
	// b^(n/2) (+ 2) < 2^(h + 1), h = (n/2) bits(b), leaves the window alone
	// when t 2^e >= 2^(h + 1)
	const uint64_t bl = 64 - __builtin_clzll((uint64_t)b);
	mpz_fdiv_r_2exp(t, lo, sl);
	ok = ok && mpz_sgn(t) != 0 && e + mpz_sizeinbase(t, 2) - 1 >= (uint64_t)(n / 2) * bl + 1;
	mpz_clears(lo, hi, t, NULL);
	return ok;
}

// Device bytes of a class: digit registers (n per lane), tables and residues
// (transform length tn), exponent words, carries.
uint64_t class_bytes(uint32_t ln, int sign, uint32_t slots, uint32_t words, bool multipass)
{
	const uint64_t n = 1ull << ln, tn = sign < 0 ? 2 * n : n;
	uint64_t b = 3 * (uint64_t)slots * n * 4 + 48 * tn + (uint64_t)slots * (words * 4 + 64);
	if (multipass) b += 2 * 3 * (uint64_t)slots * tn * 4 + (uint64_t)slots * (n / 16) * 8;
	return b;
}

// Size classes are keyed by (n, sign).
uint32_t class_key(uint32_t ln, int sign) { return 2 * ln + (sign < 0 ? 1 : 0); }

}   // namespace

struct ct_server
{
	int device = 0;
	uint64_t budget = 0, used = 0;
	uint32_t slots = 0, chunk = 0, max_steps = 0;
	std::deque<request> queue;
	std::map<uint32_t, std::unique_ptr<size_class>> classes;
	std::vector<result> done;
	uint64_t squarings = 0;
	bool graphs = true;          // multipass classes replay captured graphs
};

static void destroy_class(ct_server *s, uint32_t key)
{
	auto it = s->classes.find(key);
	if (it == s->classes.end()) return;
	size_class &c = *it->second;
	if (c.e && c.e->st) (void)hipStreamSynchronize(c.e->st);
	if (c.ev) (void)hipEventDestroy(c.ev);
	hipFree(c.d_lane_i);
	ct_engine_destroy(c.e);
	s->used -= c.bytes;
	s->classes.erase(it);
}

int ct_server_create(int device, uint64_t hbm_budget, uint32_t slots, uint32_t chunk, uint32_t max_steps, ct_server **out)
{
	if (!out || slots == 0 || chunk == 0 || max_steps == 0) return fail(CT_ERR_ARG, "bad argument");
	ct_server *s = new ct_server();
	s->device = device; s->budget = hbm_budget; s->slots = slots; s->chunk = chunk; s->max_steps = max_steps;
	*out = s;
	return CT_OK;
}

int ct_server_destroy(ct_server *s)
{
	if (!s) return CT_OK;
	while (!s->classes.empty()) destroy_class(s, s->classes.begin()->first);
	delete s;
	return CT_OK;
}

int ct_server_submit_ex(ct_server *s, uint64_t id, uint32_t log2_n, uint32_t b, uint32_t steps, int sign)
{
	if (!s) return fail(CT_ERR_ARG, "null server");
	if (sign != 1 && sign != -1) return fail(CT_ERR_ARG, "sign must be +1 or -1");
	if (log2_n < 10 || log2_n > (sign > 0 ? 22u : 21u))
		return fail(CT_ERR_UNSUPPORTED, "served sizes are 2^10 .. 2^22 (N+), 2^10 .. 2^21 (N-)");
	if (b <= 32 || b >= 0x80000000u) return fail(CT_ERR_ARG, "b must satisfy 32 < b < 2^31");
	if (steps == 0 || steps > s->max_steps) return fail(CT_ERR_ARG, "steps out of range");
	s->queue.push_back(request{id, log2_n, b, steps, sign, now_s()});
	return CT_OK;
}

int ct_server_submit(ct_server *s, uint64_t id, uint32_t log2_n, uint32_t b, uint32_t steps)
{
	return ct_server_submit_ex(s, id, log2_n, b, steps, 1);
}

static int make_class(ct_server *s, uint32_t ln, int sign, size_class **out)
{
	const uint32_t words = (s->max_steps + 31) / 32 + 1;
	// admission is checked against the larger (multipass) footprint; the
	// autotuned engine kind then fixes what is charged
	if (s->used + class_bytes(ln, sign, s->slots, words, true) > s->budget) return CT_ERR_NOMEM;
	std::unique_ptr<size_class> c(new size_class());
	int rc = ct_engine_create_ex(s->device, ln, s->slots, 0, CT_KIND_AUTO, sign, &c->e);
	if (rc) return rc;
	const uint64_t bytes = class_bytes(ln, sign, s->slots, words, c->e->kind == CT_KIND_PLAIN);
	c->e->use_graphs = c->e->use_graphs && s->graphs;
	c->ln = ln; c->sign = sign; c->words = words; c->bytes = bytes;
	c->lanes.resize(s->slots);
	c->lane_i.assign(s->slots, -1);
	// every lane gets a valid b before the first launch (idle lanes never run)
	std::vector<uint32_t> b0(s->slots, 33);
	if ((rc = ct_engine_set_b(c->e, b0.data()))) { ct_engine_destroy(c->e); return rc; }
	c->e->b_max = 0; c->e->state_bound = 1;
	std::vector<uint32_t> zeros((size_t)s->slots * words, 0);
	if ((rc = ct_engine_set_exponents(c->e, zeros.data(), words))) { ct_engine_destroy(c->e); return rc; }
	if (hipMalloc((void **)&c->d_lane_i, s->slots * sizeof(int32_t)) != hipSuccess)
	{ ct_engine_destroy(c->e); return fail(CT_ERR_NOMEM, "hipMalloc lane offsets"); }
	if (hipEventCreateWithFlags(&c->ev, hipEventDisableTiming) != hipSuccess)
	{ hipFree(c->d_lane_i); ct_engine_destroy(c->e); return fail(CT_ERR_HIP, "event create"); }
	c->stage_bits.assign((size_t)s->slots * words, 0);
	c->stage_digits.assign((size_t)s->slots * c->e->n, 0);
	c->stage_dv.assign(s->slots, divisor{});
	c->stage_lane_i.assign(s->slots, -1);
	s->used += bytes;
	*out = c.get();
	s->classes[class_key(ln, sign)] = std::move(c);
	return CT_OK;
}

// Writes request r into lane l of class c (stream-ordered on the class stream).
static int admit(size_class &c, uint32_t l, const request &r)
{
	ct_engine *e = c.e;
	const uint32_t n = e->n;
	mpz_t w; mpz_init(w);
	if (!exponent_window(r.b, n, r.steps, r.sign, w)) { mpz_clear(w); return fail(CT_ERR_COMPUTE, "exponent window not certified"); }
	uint32_t *bits = &c.stage_bits[(size_t)l * c.words];
	std::memset(bits, 0, c.words * sizeof(uint32_t));
	size_t cnt = 0;
	mpz_export(bits, &cnt, -1, sizeof(uint32_t), 0, 0, w);
	mpz_clear(w);
	int32_t *dig = &c.stage_digits[(size_t)l * n];
	std::memset(dig, 0, n * sizeof(int32_t)); dig[0] = 1;
	e->b[l] = r.b;
	e->b_max = std::max<uint64_t>(e->b_max, r.b);
	choose_primes(e);
	c.stage_dv[l] = make_divisor(r.b);
	// stream-ordered uploads from per-class staging that is not touched again
	// before the class's next chunk has completed
	hipError_t h = hipMemcpyAsync(e->d_dv + l, &c.stage_dv[l], sizeof(divisor), hipMemcpyHostToDevice, e->st);
	if (h == hipSuccess)
		h = hipMemcpyAsync(e->d_bits + (size_t)l * c.words, bits, c.words * sizeof(uint32_t), hipMemcpyHostToDevice, e->st);
	if (h == hipSuccess) h = hipMemcpyAsync(e->reg(0) + (size_t)l * n, dig, n * sizeof(int32_t), hipMemcpyHostToDevice, e->st);
	if (h == hipSuccess && e->kind == CT_KIND_PLAIN)
		h = hipMemsetAsync(e->d_carry + (size_t)l * (n / 16), 0, (n / 16) * sizeof(int64_t), e->st);
	if (h != hipSuccess) return hip_fail(h, "admission upload");
	c.lanes[l].busy = true; c.lanes[l].req = r; c.lanes[l].remaining = (int)r.steps;
	c.lane_i[l] = (int32_t)r.steps - 1;
	return CT_OK;
}

// The chunk in flight of class c has completed: advance its lanes and read
// out the finished ones.
static int complete_class(ct_server *s, size_class &c)
{
	c.in_flight = false;
	ct_engine *e = c.e;
	bool finished = false;
	for (uint32_t l = 0; l < c.lanes.size(); ++l)
		if (c.lanes[l].busy)
		{
			c.lanes[l].remaining -= c.k; c.lane_i[l] -= c.k; s->squarings += (uint64_t)c.k;
			finished |= (c.lanes[l].remaining == 0);
		}
	if (!finished) return CT_OK;
	int rc;
	if (e->kind == CT_KIND_PLAIN && (rc = plain_materialize(e))) return rc;
	if (hipStreamSynchronize(e->st) != hipSuccess) return fail(CT_ERR_HIP, "class stream failed");
	const uint32_t n = e->n;
	std::vector<int32_t> raw(n);
	std::vector<uint32_t> canon(n);
	for (uint32_t l = 0; l < c.lanes.size(); ++l)
	{
		lane_state &ls = c.lanes[l];
		if (!(ls.busy && ls.remaining == 0)) continue;
		if (hipMemcpy(raw.data(), e->reg(0) + (size_t)l * n, n * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
			return fail(CT_ERR_HIP, "read-out");
		if (canonical(raw.data(), n, ls.req.b, c.sign, canon.data())) return fail(CT_ERR_GETINT, "getInt failed");
		uint64_t r64 = canon[0], bi = ls.req.b, h = 1469598103934665603ull;
		for (uint32_t j = 1; j < n; ++j) { r64 += (uint64_t)canon[j] * bi; bi *= ls.req.b; }
		for (uint32_t j = 0; j < n; ++j) { h ^= canon[j]; h *= 1099511628211ull; }
		s->done.push_back(result{ls.req.id, r64, h, now_s() - ls.req.t_submit});
		ls.busy = false;
		c.lane_i[l] = -1;
	}
	return CT_OK;
}

// Classes whose chunk has completed; with `wait`, until at least one has
// (or none is in flight).  Returns the number completed, or a negative status.
static int poll_classes(ct_server *s, bool wait)
{
	int done = 0;
	for (;;)
	{
		bool any = false;
		for (auto &kv : s->classes)
		{
			size_class &c = *kv.second;
			if (!c.in_flight) continue;
			any = true;
			const hipError_t q = hipEventQuery(c.ev);
			if (q == hipErrorNotReady) continue;
			if (q != hipSuccess) return hip_fail(q, "class chunk");
			const int rc = complete_class(s, c);
			if (rc) return rc;
			++done;
		}
		if (done || !wait || !any) return done;
		std::this_thread::sleep_for(std::chrono::microseconds(20));
	}
}

// Launches the next chunk of an idle class that has busy lanes.
static int launch_class(ct_server *s, size_class &c)
{
	if (c.in_flight) return CT_OK;
	// chunks sized by cost: a class whose transform has 2^tl points launches
	// chunk / 2^(tl - 14) squarings (at least 4), so that a large class's
	// launch sequences stay short and the kernels of small classes that share
	// its hardware queue are not held behind them
	const uint32_t tl = c.e->tln;
	int k = std::max(4, (int)(s->chunk >> (tl > 14 ? tl - 14 : 0)));
	bool any = false;
	for (auto &ls : c.lanes) if (ls.busy) { any = true; k = std::min(k, ls.remaining); }
	if (!any) return CT_OK;
	ct_engine *e = c.e;
	c.stage_lane_i = c.lane_i;            // read by the upload below; untouched until the chunk completes
	hipError_t h = hipMemcpyAsync(c.d_lane_i, c.stage_lane_i.data(), c.stage_lane_i.size() * sizeof(int32_t),
	                              hipMemcpyHostToDevice, e->st);
	if (h != hipSuccess) return hip_fail(h, "lane offsets upload");
	const int rc = steps(e, e->d_bits, c.words, 0, k, nullptr, c.d_lane_i);
	if (rc) return rc;
	if ((h = hipEventRecord(c.ev, e->st)) != hipSuccess) return hip_fail(h, "event record");
	c.in_flight = true;
	c.k = k;
	return CT_OK;
}

int ct_server_step(ct_server *s)
{
	if (!s) return fail(CT_ERR_ARG, "null server");
	int rc = hipSetDevice(s->device) == hipSuccess ? CT_OK : fail(CT_ERR_HIP, "hipSetDevice");
	if (rc) return rc;
	// 1. classes whose chunk has completed: read out, free their lanes
	if ((rc = poll_classes(s, false)) < 0) return rc;
	// 2. admission: FIFO into free lanes of classes that are not running a
	//    chunk; a request may overtake one whose class is full, running, or
	//    cannot be allocated within the budget
	for (auto it = s->queue.begin(); it != s->queue.end();)
	{
		const request r = *it;
		auto ci = s->classes.find(class_key(r.ln, r.sign));
		size_class *c = (ci != s->classes.end()) ? ci->second.get() : nullptr;
		if (!c)
		{
			rc = make_class(s, r.ln, r.sign, &c);
			if (rc == CT_ERR_NOMEM)
			{   // release idle classes and retry once
				std::vector<uint32_t> idle;
				for (auto &kv : s->classes)
				{
					bool any = kv.second->in_flight;
					for (auto &ls : kv.second->lanes) any |= ls.busy;
					if (!any) idle.push_back(kv.first);
				}
				for (uint32_t key : idle) destroy_class(s, key);
				rc = make_class(s, r.ln, r.sign, &c);
			}
			if (rc == CT_ERR_NOMEM)
			{
				if (s->classes.empty()) return fail(CT_ERR_NOMEM, "HBM budget below one size class");
				++it; continue;
			}
			if (rc) return rc;
		}
		if (c->in_flight) { ++it; continue; }
		uint32_t l = 0;
		while (l < c->lanes.size() && c->lanes[l].busy) ++l;
		if (l == c->lanes.size()) { ++it; continue; }
		if ((rc = admit(*c, l, r))) return rc;
		it = s->queue.erase(it);
	}
	// 3. launch every idle class that has work (streams overlap)
	for (auto &kv : s->classes)
		if ((rc = launch_class(s, *kv.second))) return rc;
	// 4. wait until some class completes its chunk (not all of them)
	if ((rc = poll_classes(s, true)) < 0) return rc;
	return CT_OK;
}

int ct_server_poll(ct_server *s, uint64_t *ids, uint64_t *res64, uint64_t *hash, double *latency_s, uint32_t max,
                   uint32_t *count)
{
	if (!s || !count) return fail(CT_ERR_ARG, "bad argument");
	const uint32_t k = (uint32_t)std::min<size_t>(max, s->done.size());
	for (uint32_t i = 0; i < k; ++i)
	{
		const result &r = s->done[i];
		if (ids) ids[i] = r.id;
		if (res64) res64[i] = r.res64;
		if (hash) hash[i] = r.hash;
		if (latency_s) latency_s[i] = r.latency;
	}
	s->done.erase(s->done.begin(), s->done.begin() + k);
	*count = k;
	return CT_OK;
}

int ct_server_set_graphs(ct_server *s, int on)
{
	if (!s) return fail(CT_ERR_ARG, "null server");
	s->graphs = (on != 0);
	for (auto &kv : s->classes) kv.second->e->use_graphs = s->graphs;
	return CT_OK;
}

int ct_server_stats(const ct_server *s, uint32_t *queued, uint32_t *active, uint64_t *hbm_used, uint64_t *squarings)
{
	if (!s) return fail(CT_ERR_ARG, "null server");
	uint32_t a = 0;
	for (auto &kv : s->classes) for (auto &ls : kv.second->lanes) a += ls.busy ? 1 : 0;
	if (queued) *queued = (uint32_t)s->queue.size();
	if (active) *active = a;
	if (hbm_used) *hbm_used = s->used;
	if (squarings) *squarings = s->squarings;
	return CT_OK;
}
