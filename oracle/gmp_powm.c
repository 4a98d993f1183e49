/* Test infrastructure: GMP's mpz_powm as an independent big-integer oracle for
 * the PRP residues base^(N-1) mod N at small sizes.  Integers cross the
 * boundary as little-endian byte strings. */
#include <gmp.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

int or_powm(const uint8_t *base, size_t nb, const uint8_t *ex, size_t ne, const uint8_t *mod, size_t nm,
            uint8_t *out, size_t nout)
{
	mpz_t a, e, m, r;
	mpz_inits(a, e, m, r, NULL);
	mpz_import(a, nb, -1, 1, 0, 0, base);
	mpz_import(e, ne, -1, 1, 0, 0, ex);
	mpz_import(m, nm, -1, 1, 0, 0, mod);
	int rc = 0;
	if (mpz_sgn(m) <= 0) rc = -1;
	else
	{
		mpz_powm(r, a, e, m);
		size_t count = 0;
		if ((mpz_sizeinbase(r, 2) + 7) / 8 > nout) rc = -2;
		else
		{
			memset(out, 0, nout);
			mpz_export(out, &count, -1, 1, 0, 0, r);
		}
	}
	mpz_clears(a, e, m, r, NULL);
	return rc;
}
