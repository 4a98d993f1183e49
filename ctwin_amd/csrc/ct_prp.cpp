// Fermat PRP driver with Gerbicz-Li verification: cyclo::check of
// cyclo/src/cyclo.h:252-411 over the C ABI of the engine (base 2 as in the
// reference, or the engine's base 3; N+ as in the reference, or N- for an N-
// engine).  The exponent E = N - 1 is built with GMP as in the reference (cyclo.h:258-271),
// but its bits are uploaded once and read by the kernels instead of being
// gathered on the host every iteration (cyclo.h:244-249).
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/ctwin.h"
#include "ct_engine_impl.h"
#include "ct_prp.h"

using namespace ct;

namespace {

int ilog2_32(uint32_t v) { return 31 - __builtin_clz(v); }

// Pack mpz values into [lanes][words] little-endian 32-bit words.
std::vector<uint32_t> pack_bits(std::vector<__mpz_struct> &v, uint32_t words)
{
	std::vector<uint32_t> out((size_t)v.size() * words, 0);
	for (size_t l = 0; l < v.size(); ++l)
	{
		size_t count = 0;
		const size_t need = (mpz_sizeinbase(&v[l], 2) + 31) / 32;
		std::vector<uint32_t> tmp(std::max<size_t>(need, 1), 0);
		mpz_export(tmp.data(), &count, -1, sizeof(uint32_t), 0, 0, &v[l]);
		std::memcpy(&out[l * words], tmp.data(), std::min<size_t>(count, words) * sizeof(uint32_t));
	}
	return out;
}

}   // namespace

ct::prp_job::~prp_job() { for (auto &z : ex) mpz_clear(&z); }

int ct::prp_begin(ct_engine *e, const uint32_t *b, prp_job &J)
{
	int rc = ct_engine_set_b(e, b);
	if (rc) return rc;
	const uint32_t lanes = e->lanes, n = e->n;
	for (auto &z : J.ex) mpz_clear(&z);
	J.ex.assign(lanes, __mpz_struct());
	// E = N - 1 = (b^(n/2))^2 - b^(n/2) (N+; minus 2 more for N-); i0 = top bit
	// over all lanes (cyclo.h:256-272)
	J.i0 = 0;
	mpz_t h; mpz_init(h);
	for (uint32_t l = 0; l < lanes; ++l)
	{
		mpz_init(&J.ex[l]);
		mpz_ui_pow_ui(h, b[l], n / 2);
		mpz_mul(&J.ex[l], h, h);
		mpz_sub(&J.ex[l], &J.ex[l], h);
		if (e->sign < 0) mpz_sub_ui(&J.ex[l], &J.ex[l], 2);     // N- - 1 = b^n - b^(n/2) - 2
		J.i0 = std::max(J.i0, (int)mpz_sizeinbase(&J.ex[l], 2) - 1);
	}
	mpz_clear(h);
	const int L = 2 << (ilog2_32((uint32_t)J.i0) / 2);
	J.B = (J.i0 - 1) / L + 1;
	const uint32_t words = (uint32_t)(J.i0 / 32 + 1);
	const std::vector<uint32_t> bits = pack_bits(J.ex, words);
	return ct_engine_set_exponents(e, bits.data(), words);
}

int ct::prp_finish(ct_engine *e, prp_job &J, int *prm, uint64_t *res64, int *gl_ok, uint32_t *residue)
{
	const uint32_t lanes = e->lanes, n = e->n;
	const int B = J.B;
	int rc;
	std::vector<uint32_t> canon((size_t)lanes * n), y((size_t)lanes * n), d1((size_t)lanes * n);
	if ((rc = ct_engine_get_int(e, 0, canon.data()))) return rc;

	// d(t+1) = d(t) * result into register 2 (cyclo.h:317-320)
	if ((rc = ct_engine_mul_regs(e, 2, 0, 1))) return rc;

	// res = sum of the B-bit chunks of E (cyclo.h:322-338)
	int i1 = 0;
	{
		mpz_t res, tmp; mpz_init(res); mpz_init(tmp);
		for (uint32_t l = 0; l < lanes; ++l)
		{
			mpz_set_ui(res, 0);
			while (mpz_sgn(&J.ex[l]) != 0)
			{
				mpz_fdiv_r_2exp(tmp, &J.ex[l], (mp_bitcnt_t)B);
				mpz_add(res, res, tmp);
				mpz_fdiv_q_2exp(&J.ex[l], &J.ex[l], (mp_bitcnt_t)B);
			}
			mpz_set(&J.ex[l], res);
			i1 = std::max(i1, (int)mpz_sizeinbase(res, 2) - 1);
		}
		mpz_clear(res); mpz_clear(tmp);
	}
	const uint32_t words = (uint32_t)(std::max(i1, B - 1) / 32 + 1);
	const std::vector<uint32_t> bits = pack_bits(J.ex, words);
	if ((rc = ct_engine_set_exponents(e, bits.data(), words))) return rc;

	// (2^(res >> B) d(t))^(2^B) 2^(res mod 2^B) (cyclo.h:340-363)
	if ((rc = ct_engine_set(e, 1))) return rc;
	if (i1 >= B && (rc = ct_engine_run(e, i1, B, 0))) return rc;
	if ((rc = ct_engine_mul_regs(e, 0, 0, 1))) return rc;
	if ((rc = ct_engine_run(e, B - 1, 0, 0))) return rc;

	if ((rc = ct_engine_get_int(e, 0, y.data()))) return rc;
	if ((rc = ct_engine_get_int(e, 2, d1.data()))) return rc;

	bool all_ok = true;
	for (uint32_t l = 0; l < lanes; ++l)
	{
		gl_ok[l] = std::memcmp(&y[(size_t)l * n], &d1[(size_t)l * n], n * sizeof(uint32_t)) == 0 ? 1 : 0;
		all_ok &= (gl_ok[l] != 0);
	}
	if (residue) std::memcpy(residue, canon.data(), canon.size() * sizeof(uint32_t));
	rc = ct_engine_is_prime(e, canon.data(), prm, res64);
	if (rc) return rc;
	return all_ok ? CT_OK : fail(CT_ERR_GERBICZ, "Gerbicz failed");
}

int ct_prp_check(ct_engine *e, const uint32_t *b, int *prm, uint64_t *res64, int *gl_ok, uint32_t *residue)
{
	if (!e || !b || !prm || !res64 || !gl_ok) return fail(CT_ERR_ARG, "bad argument");
	prp_job J;
	int rc = prp_begin(e, b, J);
	if (rc) return rc;
	// x = 1, d = 1; PRP loop with d(t) <- d(t) x every B iterations
	if ((rc = ct_engine_set(e, 1)) || (rc = ct_engine_copy(e, 1, 0)) || (rc = ct_engine_run(e, J.i0, 0, J.B))) return rc;
	return prp_finish(e, J, prm, res64, gl_ok, residue);
}

// ct_prp_check split at its loop, for callers that drive the loop themselves
// (checkpointing, interruption, or a test that alters the state mid-run).
struct ct_prp_job
{
	prp_job J;
	bool finished = false;
};

int ct_prp_begin(ct_engine *e, const uint32_t *b, ct_prp_job **out, int *i0, int *gl_block)
{
	if (!e || !b || !out) return fail(CT_ERR_ARG, "bad argument");
	*out = nullptr;
	ct_prp_job *job = new ct_prp_job();
	const int rc = prp_begin(e, b, job->J);
	if (rc) { delete job; return rc; }
	if (i0) *i0 = job->J.i0;
	if (gl_block) *gl_block = job->J.B;
	*out = job;
	return CT_OK;
}

int ct_prp_finish(ct_engine *e, ct_prp_job *job, int *prm, uint64_t *res64, int *gl_ok, uint32_t *residue)
{
	if (!e || !job || !prm || !res64 || !gl_ok) return fail(CT_ERR_ARG, "bad argument");
	if (job->finished) return fail(CT_ERR_ARG, "the PRP job has already been finished");
	job->finished = true;          // prp_finish consumes the exponents (their B-bit chunk sums replace them)
	return prp_finish(e, job->J, prm, res64, gl_ok, residue);
}

int ct_prp_job_free(ct_prp_job *job)
{
	delete job;
	return CT_OK;
}
