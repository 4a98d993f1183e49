// Batch PRP driver over a candidate file: cyclo::check(filename) of
// cyclo/src/cyclo.h:414-485 with its checkpointing (cyclo.h:252-300) and result
// files (cyclo.h:379-408).  Candidates (one b per line, as written by the
// sieve's cand_*.txt) are taken `lanes` at a time, the last batch padded with
// its last b (parseFile, cyclo.h:216-241); each batch runs the base-2 PRP loop
// with Gerbicz-Li on the device in segments that end at the Gerbicz-Li points,
// where the checkpoint cyclo.ctx (the reference's format) is written every
// `checkpoint_s` seconds, and where a quit request or the step budget stops the
// run with the checkpoint saved.  A later call resumes from cyclo.ctx (or
// cyclo.ctx.old).  Results are appended to result_c.txt and probable primes
// to prp_<n>.txt in the reference's wording.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/ctwin.h"
#include "ct_engine_impl.h"
#include "ct_prp.h"

using namespace ct;

namespace {

std::atomic<int> g_quit{0};

std::string path_in(const char *dir, const std::string &leaf)
{
	return (dir && *dir) ? std::string(dir) + "/" + leaf : leaf;
}

double now_s()
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// 16 upper-case hex digits (res64String, cyclo.h:51-57)
std::string res64_string(uint64_t r)
{
	char buf[17];
	std::snprintf(buf, sizeof(buf), "%016llX", (unsigned long long)r);
	return buf;
}

}   // namespace

void ct_cyclo_quit(void) { g_quit.store(1); }

int ct_cyclo_check_file(int device, uint32_t log2_n, const char *cand_file, const char *work_dir, uint32_t lanes,
                        double checkpoint_s, uint64_t max_steps, ct_cyclo_stats *stats)
{
	if (!cand_file || lanes == 0) return fail(CT_ERR_ARG, "bad argument");
	if (log2_n < 5 || log2_n > 22) return fail(CT_ERR_UNSUPPORTED, "log2_n must be in [5, 22]");
	g_quit.store(0);
	ct_cyclo_stats st{};

	// parseFile: batches of `lanes`, the last one padded with its last b
	std::vector<std::vector<uint32_t>> batches;
	{
		std::ifstream in(cand_file);
		if (!in.is_open()) return fail(CT_ERR_ARG, "cannot open input file");
		std::vector<uint32_t> cur;
		std::string line;
		while (std::getline(in, line))
		{
			if (line.empty()) continue;
			const unsigned long long v = std::stoull(line);
			if (v <= 32 || v >= 0x80000000ull) return fail(CT_ERR_ARG, "candidate b must satisfy 32 < b < 2^31");
			cur.push_back((uint32_t)v);
			if (cur.size() == lanes) { batches.push_back(cur); cur.clear(); }
		}
		if (!cur.empty())
		{
			while (cur.size() < lanes) cur.push_back(cur.back());
			batches.push_back(cur);
		}
	}

	ct_engine *e = nullptr;
	int rc = ct_engine_create(device, log2_n, lanes, 0, CT_KIND_AUTO, &e);
	if (rc) return rc;
	struct guard { ct_engine *e; ~guard() { ct_engine_destroy(e); } } g{e};

	const std::string ctx = path_in(work_dir, "cyclo.ctx");
	int rk = 0, ri = 0;
	double restored = 0;
	if (ct_context_load(e, ctx.c_str(), &rk, &ri, &restored) == CT_OK) st.resumed = 1;
	else { rk = ri = 0; restored = 0; }
	const double t0 = now_s() - restored;
	double last_save = now_s();
	constexpr uint64_t kAhead = 8;      // Gerbicz-Li segments in flight at most
	uint64_t segments = 0;
	const uint32_t n = 1u << log2_n;

	for (size_t k = (size_t)rk; k < batches.size(); ++k)
	{
		const std::vector<uint32_t> &b = batches[k];
		prp_job J;
		if ((rc = prp_begin(e, b.data(), J))) return rc;
		if (!st.resumed || k != (size_t)rk)
		{
			ri = 0;
			if ((rc = ct_engine_set(e, 1)) || (rc = ct_engine_copy(e, 1, 0))) return rc;   // x = 1, d(t) = 1
		}
		const int B = J.B;
		int i = (ri != 0) ? ri : J.i0;
		while (i >= 0)
		{
			// (no checkpoint at i = 0: a saved i of 0 reads back as "start of the
			// batch", cyclo.h:274)
			if (i > 0 && i % B == 0 && now_s() - last_save > checkpoint_s)
			{
				if ((rc = ct_context_save(e, ctx.c_str(), (int)k, i, now_s() - t0))) return rc;
				last_save = now_s();
			}
			if (i > 0 && (g_quit.load() || (max_steps != 0 && st.squarings >= max_steps)))
			{
				if ((rc = ct_context_save(e, ctx.c_str(), (int)k, i, now_s() - t0))) return rc;
				st.elapsed = now_s() - t0;
				if (stats) *stats = st;
				return CT_INTERRUPTED;
			}
			// bits i .. lo, lo just above the next Gerbicz-Li point below i
			const int next = (i % B == 0) ? i - B : (i / B) * B;
			int lo = std::max(next + 1, 0);
			if (max_steps != 0 && st.squarings < max_steps)
				lo = (int)std::max<int64_t>(lo, (int64_t)i - (int64_t)(max_steps - st.squarings) + 1);
			if ((rc = ct_engine_run(e, i, lo, B))) return rc;
			st.squarings += (uint64_t)(i - lo + 1);
			i = lo - 1;
			// bounded run-ahead: the host checks the quit flag and the
			// checkpoint clock as it enqueues segments, so it waits for the
			// device every kAhead segments (the reference forces a queue finish
			// every 1024 launches, ocl.h:541-545)
			if (++segments % kAhead == 0 && hipStreamSynchronize(e->st) != hipSuccess)
				return fail(CT_ERR_HIP, "engine stream failed");
		}
		std::vector<int> prm(lanes), glok(lanes);
		std::vector<uint64_t> r64(lanes);
		if ((rc = prp_finish(e, J, prm.data(), r64.data(), glok.data(), nullptr))) return rc;

		std::string text;
		std::vector<uint32_t> prps;
		for (uint32_t l = 0; l < lanes; ++l)
		{
			if (l > 0 && b[l] == b[l - 1]) continue;
			text += std::to_string(b[l]) + "^" + std::to_string(n) + " - " + std::to_string(b[l]) + "^" +
			        std::to_string(n / 2) + " + 1 is ";
			if (prm[l]) { text += "a probable prime.\n"; prps.push_back(b[l]); }
			else text += "composite. RES64: " + res64_string(r64[l]) + "\n";
			++st.candidates;
		}
		{
			std::ofstream res(path_in(work_dir, "result_c.txt"), std::ios::app);
			if (!res.is_open()) return fail(CT_ERR_ARG, "cannot write result_c.txt");
			res << text;
		}
		if (!prps.empty())
		{
			std::ofstream pf(path_in(work_dir, "prp_" + std::to_string(log2_n) + ".txt"), std::ios::app);
			if (!pf.is_open()) return fail(CT_ERR_ARG, "cannot write the prp file");
			for (uint32_t x : prps) pf << x << "\n";
			st.prps += prps.size();
		}
	}
	// done: the checkpoint would only restart the last batch
	std::remove(ctx.c_str());
	std::remove((ctx + ".old").c_str());
	st.elapsed = now_s() - t0;
	if (stats) *stats = st;
	return CT_OK;
}
