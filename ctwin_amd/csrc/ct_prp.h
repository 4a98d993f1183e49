// Pieces of the PRP test shared by ct_prp_check and the batch driver.
#pragma once

#include <gmp.h>

#include <vector>

struct ct_engine;

namespace ct {

struct prp_job
{
	std::vector<__mpz_struct> ex;   // E = b^n - b^(n/2) per lane (then the Gerbicz-Li residue sum)
	int i0 = 0;                     // top bit index over the lanes
	int B = 0;                      // Gerbicz-Li block (cyclo.h:272)
	prp_job() = default;
	prp_job(const prp_job &) = delete;
	prp_job &operator=(const prp_job &) = delete;
	~prp_job();
};

// set_b, exponents built and uploaded; registers untouched.
int prp_begin(ct_engine *e, const uint32_t *b, prp_job &J);
// After the PRP loop (registers 0 = x, 1 = d(t)): RES64, primality, Gerbicz-Li.
int prp_finish(ct_engine *e, prp_job &J, int *prm, uint64_t *res64, int *gl_ok, uint32_t *residue);

}   // namespace ct
