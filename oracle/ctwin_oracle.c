/*
 * ctwin oracle -- TEST INFRASTRUCTURE ONLY (see ctwin_oracle.h).
 *
 * Plain-C restatement of galloty/ctwin: every function cites the reference
 * lines it follows.  Plain modular arithmetic replaces the reference's
 * Montgomery forms; every operation is fully reduced mod p, so the residues
 * are the same as the reference's device residues up to the Montgomery factor
 * and the carried digits are identical.
 */
#include "ctwin_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef __int128 i128;

/* ------------------------------------------------------------------ */
/* mod-p helpers (transform.h:32-78, Zp)                               */
/* ------------------------------------------------------------------ */

static inline uint32_t addm(uint32_t a, uint32_t b, uint32_t p) { uint64_t s = (uint64_t)a + b; return (uint32_t)(s >= p ? s - p : s); }
static inline uint32_t subm(uint32_t a, uint32_t b, uint32_t p) { return (a >= b) ? a - b : (uint32_t)((uint64_t)a + p - b); }
static inline uint32_t mulm(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)(((uint64_t)a * b) % p); }
static inline uint32_t negm(uint32_t a, uint32_t p) { return a ? p - a : 0; }
/* algo.cpp:43 Zp::half */
static inline uint32_t halfm(uint32_t a, uint32_t p) { return (a % 2 == 0) ? a / 2 : (uint32_t)(((uint64_t)a + p) / 2); }

uint32_t or_powmod32(uint32_t a, uint64_t e, uint32_t p)
{
	uint64_t r = 1 % p, y = a % p;
	while (e) { if (e & 1) r = (r * y) % p; y = (y * y) % p; e >>= 1; }
	return (uint32_t)r;
}

static inline uint32_t invm(uint32_t a, uint32_t p) { return or_powmod32(a, p - 2, p); }
/* Zp(int32) of transform.h:43 */
static inline uint32_t seti(int64_t i, uint32_t p) { int64_t r = i % (int64_t)p; return (uint32_t)(r < 0 ? r + p : r); }

/* transform.h:245-249 */
static uint32_t bitrev(uint32_t i, uint32_t n)
{
	uint32_t r = 0;
	for (uint32_t k = n, j = i; k > 1; k /= 2, j /= 2) r = (2 * r) | (j % 2);
	return r;
}

/* ------------------------------------------------------------------ */
/* root tables: transform.h:336-366, algo.cpp:74-85                    */
/* ------------------------------------------------------------------ */

void or_transform_tables(uint32_t p, uint32_t g, uint32_t n, uint32_t *w, uint32_t *wi)
{
	memset(w, 0, sizeof(uint32_t) * n); memset(wi, 0, sizeof(uint32_t) * n);
	/* norm(n/2) = -((p-1)/(n/2)) = 2/n mod p, transform.h:75 */
	const uint32_t nn_2 = negm((p - 1) / (n / 2), p);
	for (uint32_t s = 1; s < n; s *= 2)
	{
		const uint32_t r_s = or_powmod32(g, (p - 1) / (6 * s), p);
		for (uint32_t j = 0; j < s; ++j)
		{
			const uint32_t j_s = 3 * bitrev(j, s) + 1 + (2 * j) / s;
			const uint32_t wrsj = or_powmod32(r_s, j_s, p);
			wi[s + s - j - 1] = negm(wrsj, p);
			w[s + j] = wrsj;
		}
		if (s == 1)
		{
			wi[0] = w[1];
			const uint32_t dsr_inv = invm(subm(1, addm(r_s, r_s, p), p), p);
			wi[1] = mulm(addm(dsr_inv, dsr_inv, p), nn_2, p);
		}
	}
}

/* ------------------------------------------------------------------ */
/* squaring transform: kernel.cl frwd41_0 :244, frwd2 :177, sqr2 :189, */
/* bkwd2 :183, bkwd41_0 :262, sequenced as engine.h:387-419            */
/* ------------------------------------------------------------------ */

static void forward_levels(uint32_t p, uint32_t n, const uint32_t *w, uint32_t *x, int last_level)
{
	const uint32_t h = n / 2;
	/* level 1: X^n - X^(n/2) + 1 = (X^(n/2) - r)(X^(n/2) - r'), r + r' = 1 */
	for (uint32_t i = 0; i < h; ++i)
	{
		const uint32_t u0 = x[i], u1 = x[i + h], t = mulm(u1, w[1], p);
		x[i] = addm(u0, t, p);
		x[i + h] = addm(u1, subm(u0, t, p), p);
	}
	/* levels s = 2 ... (m = block size); stop before m = 2 unless last_level */
	const uint32_t m_stop = last_level ? 2 : 4;
	for (uint32_t s = 2, m = n / 2; m >= m_stop; s *= 2, m /= 2)
	{
		for (uint32_t j = 0; j < s; ++j)
		{
			const uint32_t ws = w[s + j];
			for (uint32_t i = 0; i < m / 2; ++i)
			{
				const uint32_t k = j * m + i;
				const uint32_t u0 = x[k], t = mulm(x[k + m / 2], ws, p);
				x[k] = addm(u0, t, p); x[k + m / 2] = subm(u0, t, p);
			}
		}
	}
}

static void inverse_levels(uint32_t p, uint32_t n, const uint32_t *wi, uint32_t *x, int first_level)
{
	/* inverse levels from block size 4 (or 2 when first_level) up to n/2 */
	for (uint32_t s = first_level ? n / 2 : n / 4, m = n / s; s >= 2; s /= 2, m *= 2)
	{
		for (uint32_t j = 0; j < s; ++j)
		{
			const uint32_t wis = wi[s + j];
			for (uint32_t i = 0; i < m / 2; ++i)
			{
				const uint32_t k = j * m + i;
				const uint32_t u0 = x[k], u1 = x[k + m / 2];
				x[k] = addm(u0, u1, p); x[k + m / 2] = mulm(subm(u0, u1, p), wis, p);
			}
		}
	}
	/* level 1 merge with the 2/n scale (bkwd41_0, kernel.cl:262-269) */
	const uint32_t h = n / 2;
	for (uint32_t i = 0; i < h; ++i)
	{
		const uint32_t a = x[i], b = x[i + h];
		const uint32_t vh = mulm(addm(a, b, p), wi[0], p);
		x[i + h] = mulm(subm(b, a, p), wi[1], p);
		x[i] = mulm(subm(a, vh, p), wi[1], p);
	}
}

void or_square_mod(uint32_t p, uint32_t n, const uint32_t *w, const uint32_t *wi, uint32_t *x)
{
	forward_levels(p, n, w, x, 0);
	/* innermost level fused with the square: (u0 + u1 X)^2 mod X^2 - w^2 */
	for (uint32_t j = 0; j < n / 2; ++j)
	{
		const uint32_t k = 2 * j, u0 = x[k], u1 = x[k + 1];
		const uint32_t u1w = mulm(u1, w[n / 2 + j], p);
		x[k + 1] = mulm(addm(u0, u0, p), u1, p);
		x[k] = addm(mulm(u0, u0, p), mulm(u1w, u1w, p), p);
	}
	inverse_levels(p, n, wi, x, 0);
}

void or_mul_mod(uint32_t p, uint32_t n, const uint32_t *w, const uint32_t *wi, uint32_t *x, const uint32_t *y)
{
	uint32_t *v = (uint32_t *)malloc(sizeof(uint32_t) * n);
	memcpy(v, y, sizeof(uint32_t) * n);
	forward_levels(p, n, w, x, 1);
	forward_levels(p, n, w, v, 1);
	/* mul2/mul4: pointwise product, halved (kernel.cl:706) */
	for (uint32_t k = 0; k < n; ++k) x[k] = halfm(mulm(x[k], v[k], p), p);
	inverse_levels(p, n, wi, x, 1);
	free(v);
}

/* ------------------------------------------------------------------ */
/* algo.cpp:87-141, the single-prime CPU transform (P1, generator 19)  */
/* ------------------------------------------------------------------ */

void or_poly_sqr_algo(uint32_t n, uint32_t *P)
{
	const uint32_t p = 4095u * (1u << 20) + 1, g = 19;
	uint32_t *w = (uint32_t *)calloc(n, sizeof(uint32_t));
	for (uint32_t s = 1; s < n; s *= 2)
	{
		const uint32_t r_s = or_powmod32(g, (p - 1) / (6 * s), p);
		for (uint32_t j = 0; j < s; ++j) w[s + j] = or_powmod32(r_s, 3 * bitrev(j, s) + 1 + (2 * j) / s, p);
	}
	const uint32_t w_1 = w[1];
	const uint32_t d_inv = invm(subm(1, addm(w_1, w_1, p), p), p), d2_inv = addm(d_inv, d_inv, p);
	const uint32_t n_2 = n / 2;
	for (uint32_t i = 0; i < n_2; ++i)
	{
		const uint32_t u0 = P[i], u1 = P[i + n_2], u1w = mulm(u1, w_1, p);
		P[i] = addm(u0, u1w, p);
		P[i + n_2] = addm(P[i + n_2], subm(u0, u1w, p), p);
	}
	for (uint32_t s = 2, m = n_2; m >= 2; s *= 2, m /= 2)
		for (uint32_t j = 0; j < s; ++j)
			for (uint32_t i = 0; i < m / 2; ++i)
			{
				const uint32_t k = j * m + i, u0 = P[k], t = mulm(P[k + m / 2], w[s + j], p);
				P[k] = addm(u0, t, p); P[k + m / 2] = subm(u0, t, p);
			}
	for (uint32_t k = 0; k < n; ++k) P[k] = mulm(P[k], P[k], p);
	for (uint32_t s = n_2, m = 2; m <= n_2; s /= 2, m *= 2)
		for (uint32_t j = 0; j < s; ++j)
		{
			const uint32_t winv = invm(w[s + j], p);
			for (uint32_t i = 0; i < m / 2; ++i)
			{
				const uint32_t k = j * m + i, u0 = P[k], u1 = P[k + m / 2];
				P[k] = halfm(addm(u0, u1, p), p);
				P[k + m / 2] = mulm(halfm(subm(u0, u1, p), p), winv, p);
			}
		}
	for (uint32_t i = 0; i < n_2; ++i)
	{
		const uint32_t u0 = P[i], u1 = P[i + n_2];
		P[i] = mulm(d2_inv, halfm(subm(u0, mulm(w_1, addm(u0, u1, p), p), p), p), p);
		P[i + n_2] = mulm(d2_inv, halfm(subm(u1, u0, p), p), p);
	}
	free(w);
}

void or_poly_mul_exact(uint32_t n, const int64_t *a, const int64_t *b, i128 *c)
{
	/* X^(n + k) = X^(n/2 + k) - X^k, applied from the top degree down */
	i128 *t = (i128 *)calloc(2 * n, sizeof(i128));
	for (uint32_t i = 0; i < n; ++i)
		for (uint32_t j = 0; j < n; ++j) t[i + j] += (i128)a[i] * b[j];
	for (uint32_t d = 2 * n - 1; d >= n; --d)
	{
		const i128 v = t[d]; t[d] = 0;
		t[d - n / 2] += v; t[d - n] -= v;
	}
	for (uint32_t i = 0; i < n; ++i) c[i] = t[i];
	free(t);
}

/* ------------------------------------------------------------------ */
/* Garner CRT: kernel.cl:451-460                                       */
/* ------------------------------------------------------------------ */

i128 or_garner3(uint32_t r1, uint32_t r2, uint32_t r3, uint32_t p1, uint32_t p2, uint32_t p3)
{
	const uint32_t u13 = mulm(subm(r1, r3 % p1, p1), invm(p3 % p1, p1), p1);
	const uint32_t u23 = mulm(subm(r2, r3 % p2, p2), invm(p3 % p2, p2), p2);
	const uint32_t u123 = mulm(subm(u13, u23 % p1, p1), invm(p2 % p1, p1), p1);
	const u128 M = (u128)p1 * p2 * p3, M_2 = M / 2;
	const u128 v = (u128)p2 * p3 * u123 + (u128)u23 * p3 + r3;
	return (v > M_2) ? (i128)v - (i128)M : (i128)v;
}

void or_garner3_words(uint32_t r1, uint32_t r2, uint32_t r3, uint32_t p1, uint32_t p2, uint32_t p3,
                      uint64_t *lo, int64_t *hi)
{
	const i128 v = or_garner3(r1, r2, r3, p1, p2, p3);
	*lo = (uint64_t)v; *hi = (int64_t)(v >> 64);
}

/* ------------------------------------------------------------------ */
/* carry: normalize1 kernel.cl:1029-1051, normalize2 :1053-1081        */
/* reduce96/reduce64 return the remainder with the sign of the        */
/* dividend and the truncated quotient (kernel.cl:480-510): C's / %.  */
/* ------------------------------------------------------------------ */

void or_normalize(uint32_t n, uint32_t csize, uint32_t b, int dup, const i128 *coef, int32_t *digits)
{
	const uint32_t nc = n / csize;
	int64_t *f = (int64_t *)malloc(sizeof(int64_t) * nc);
	for (uint32_t j = 0; j < nc; ++j)
	{
		i128 a = 0;
		for (uint32_t c = 0; c < csize; ++c)
		{
			const uint32_t k = j * csize + c;
			a += coef[k]; if (dup) a += coef[k];
			const i128 r = a % (i128)b;
			a /= (i128)b;
			digits[k] = (int32_t)r;
		}
		f[j] = (int64_t)a;
	}
	for (uint32_t j = 0; j < nc; ++j)
	{
		const uint32_t src = (j + nc - 1) % nc;
		int64_t a = f[src];
		if (j == 0) a = -a;
		else if (j == nc / 2) a += f[nc - 1];
		uint32_t c;
		int done = 0;
		for (c = 0; c < csize - 1; ++c)
		{
			const uint32_t k = j * csize + c;
			a += digits[k];
			const int64_t r = a % (int64_t)b;
			a /= (int64_t)b;
			digits[k] = (int32_t)r;
			if (a == 0) { done = 1; break; }
		}
		if (!done) digits[j * csize + c] += (int32_t)a;
	}
	free(f);
}

/* ------------------------------------------------------------------ */
/* engine context                                                      */
/* ------------------------------------------------------------------ */

typedef struct
{
	uint32_t n, csize, b;
	uint32_t p[3], g[3];
	uint32_t *w[3], *wi[3];
	uint32_t *x[3], *y[3];
	i128 *coef;
} or_engine;

void *or_engine_new(const uint32_t *p, const uint32_t *g, uint32_t n, uint32_t csize, uint32_t b)
{
	or_engine *e = (or_engine *)calloc(1, sizeof(or_engine));
	e->n = n; e->csize = csize; e->b = b;
	for (int i = 0; i < 3; ++i)
	{
		e->p[i] = p[i]; e->g[i] = g[i];
		e->w[i] = (uint32_t *)malloc(sizeof(uint32_t) * n);
		e->wi[i] = (uint32_t *)malloc(sizeof(uint32_t) * n);
		e->x[i] = (uint32_t *)malloc(sizeof(uint32_t) * n);
		e->y[i] = (uint32_t *)malloc(sizeof(uint32_t) * n);
		or_transform_tables(p[i], g[i], n, e->w[i], e->wi[i]);
	}
	e->coef = (i128 *)malloc(sizeof(i128) * n);
	return e;
}

void or_engine_free(void *ve)
{
	or_engine *e = (or_engine *)ve;
	if (!e) return;
	for (int i = 0; i < 3; ++i) { free(e->w[i]); free(e->wi[i]); free(e->x[i]); free(e->y[i]); }
	free(e->coef); free(e);
}

static void square_coefs(or_engine *e, const int32_t *digits)
{
	const uint32_t n = e->n;
	for (int i = 0; i < 3; ++i)
	{
		for (uint32_t k = 0; k < n; ++k) e->x[i][k] = seti(digits[k], e->p[i]);
		or_square_mod(e->p[i], n, e->w[i], e->wi[i], e->x[i]);
	}
	for (uint32_t k = 0; k < n; ++k)
		e->coef[k] = or_garner3(e->x[0][k], e->x[1][k], e->x[2][k], e->p[0], e->p[1], e->p[2]);
}

void or_engine_square_dup(void *ve, int dup, int32_t *digits)
{
	or_engine *e = (or_engine *)ve;
	square_coefs(e, digits);
	or_normalize(e->n, e->csize, e->b, dup, e->coef, digits);
}

void or_engine_square_coefs(void *ve, const int32_t *digits, uint64_t *lo, int64_t *hi)
{
	or_engine *e = (or_engine *)ve;
	square_coefs(e, digits);
	for (uint32_t k = 0; k < e->n; ++k) { lo[k] = (uint64_t)e->coef[k]; hi[k] = (int64_t)(e->coef[k] >> 64); }
}

void or_engine_mul(void *ve, int32_t *xd, const int32_t *yd)
{
	or_engine *e = (or_engine *)ve;
	const uint32_t n = e->n;
	for (int i = 0; i < 3; ++i)
	{
		for (uint32_t k = 0; k < n; ++k) { e->x[i][k] = seti(xd[k], e->p[i]); e->y[i][k] = seti(yd[k], e->p[i]); }
		or_mul_mod(e->p[i], n, e->w[i], e->wi[i], e->x[i], e->y[i]);
	}
	for (uint32_t k = 0; k < n; ++k)
		e->coef[k] = or_garner3(e->x[0][k], e->x[1][k], e->x[2][k], e->p[0], e->p[1], e->p[2]);
	or_normalize(n, e->csize, e->b, 0, e->coef, xd);
}

/* ------------------------------------------------------------------ */
/* canonical read-out: transform.h:488-597; RES64: :600-636            */
/* ------------------------------------------------------------------ */

/* transform.h:488-497: f = f' b + r with 0 <= r < b */
static uint32_t reduce_floor(int64_t *f, uint32_t b)
{
	int64_t q = *f / (int64_t)b, r = *f % (int64_t)b;
	if (r < 0) { r += b; --q; }
	*f = q;
	return (uint32_t)r;
}

int or_get_int(uint32_t n, uint32_t b, const int32_t *digits, uint32_t *x)
{
	int64_t f = 0;
	for (uint32_t j = 0; j < n; ++j)
	{
		int64_t e = f + digits[j];
		x[j] = reduce_floor(&e, b);
		f = e;
	}
	while (f != 0)
	{
		int64_t e = -f;
		for (uint32_t j = 0; j < n / 2; ++j) { e += (int32_t)x[j]; x[j] = reduce_floor(&e, b); }
		e += f;
		for (uint32_t j = n / 2; j < n; ++j) { e += (int32_t)x[j]; x[j] = reduce_floor(&e, b); if (e == 0) break; }
		f = e;
	}
	int overflow = 1;
	for (uint32_t j = n - 1; j >= n / 2; --j) if (x[j] != b - 1) { overflow = 0; break; }
	if (overflow)
	{
		overflow = 0;
		for (uint32_t j = n / 2 - 1; j > 0; --j) if (x[j] != 0) { overflow = 1; break; }
		if (!overflow) overflow = (x[0] != 0);
	}
	if (overflow)
	{
		int32_t carry = 0;
		for (uint32_t j = 0; j < n / 2; ++j)
		{
			const int32_t y = (j == 0) ? 1 : 0;
			const int32_t s = (int32_t)x[j] - y + carry;
			if (s < 0) { x[j] = (uint32_t)(s + (int32_t)b); carry = -1; } else { x[j] = (uint32_t)s; carry = 0; break; }
		}
		if (carry != 0) return -1;
		for (uint32_t j = n / 2; j < n; ++j) x[j] = 0;
	}
	return 0;
}

int or_is_prime(uint32_t n, uint32_t b, const uint32_t *x, int *prm, uint64_t *res64)
{
	int p = (x[0] == 1), err = (x[0] == 0);
	uint64_t r = x[0], bi = b;
	for (uint32_t j = 1; j < n; ++j)
	{
		p &= (x[j] == 0); err &= (x[j] == 0);
		r += (uint64_t)x[j] * bi;
		bi *= b;
	}
	*prm = p; *res64 = r;
	return err;
}

/* ------------------------------------------------------------------ */
/* exponent E = b^n - b^(n/2) (cyclo.h:258-271, GMP there)             */
/* ------------------------------------------------------------------ */

size_t or_exponent(uint32_t n, uint32_t b, uint32_t *w)
{
	/* h = b^(n/2) by repeated multiplication by b */
	const size_t cap = n + 2;
	uint32_t *h = (uint32_t *)calloc(cap, sizeof(uint32_t));
	size_t hs = 1; h[0] = 1;
	for (uint32_t i = 0; i < n / 2; ++i)
	{
		uint64_t c = 0;
		for (size_t k = 0; k < hs; ++k) { const uint64_t t = (uint64_t)h[k] * b + c; h[k] = (uint32_t)t; c = t >> 32; }
		if (c) h[hs++] = (uint32_t)c;
	}
	/* w = h^2 - h */
	memset(w, 0, sizeof(uint32_t) * cap);
	for (size_t i = 0; i < hs; ++i)
	{
		uint64_t c = 0;
		for (size_t j = 0; j < hs; ++j)
		{
			const uint64_t t = (uint64_t)h[i] * h[j] + w[i + j] + c;
			w[i + j] = (uint32_t)t; c = t >> 32;
		}
		size_t k = i + hs;
		while (c) { const uint64_t t = (uint64_t)w[k] + c; w[k] = (uint32_t)t; c = t >> 32; ++k; }
	}
	int64_t br = 0;
	for (size_t k = 0; k < 2 * hs; ++k)
	{
		int64_t t = (int64_t)w[k] - (k < hs ? h[k] : 0) + br;
		br = 0; if (t < 0) { t += ((int64_t)1 << 32); br = -1; }
		w[k] = (uint32_t)t;
	}
	size_t ws = 2 * hs;
	while (ws > 1 && w[ws - 1] == 0) --ws;
	free(h);
	return ws;
}

static inline int bit_of(const uint32_t *w, size_t ws, long i)
{
	if (i < 0 || (size_t)(i / 32) >= ws) return 0;
	return (w[i / 32] >> (i % 32)) & 1;
}

static int top_bit(const uint32_t *w, size_t ws)
{
	for (long k = (long)ws - 1; k >= 0; --k)
		if (w[k]) return (int)(k * 32 + 31 - __builtin_clz(w[k]));
	return -1;
}

static inline int ilog2_32(uint32_t v) { return 31 - __builtin_clz(v); }

void or_prp_steps(void *ve, int i0, int steps, int32_t *digits)
{
	or_engine *e = (or_engine *)ve;
	const uint32_t n = e->n;
	uint32_t *E = (uint32_t *)malloc(sizeof(uint32_t) * (n + 2));
	const size_t es = or_exponent(n, e->b, E);
	if (i0 <= 0) i0 = top_bit(E, es);
	memset(digits, 0, sizeof(int32_t) * n); digits[0] = 1;
	for (int t = 0, i = i0; t < steps && i >= 0; ++t, --i) or_engine_square_dup(e, bit_of(E, es, i), digits);
	free(E);
}

void or_prp_steps_window(void *ve, const uint32_t *win, int i_lo_win, int i0, int steps, int32_t *digits)
{
	or_engine *e = (or_engine *)ve;
	memset(digits, 0, sizeof(int32_t) * e->n); digits[0] = 1;
	for (int t = 0, i = i0; t < steps && i >= 0; ++t, --i)
	{
		const int o = i - i_lo_win;
		const int bit = (o >= 0) ? (int)((win[o / 32] >> (o % 32)) & 1u) : 0;
		or_engine_square_dup(e, bit, digits);
	}
}

int or_prp(void *ve, int i0, int flip_at, uint32_t *canon, uint64_t *res64, int *prm, int *gl_ok)
{
	or_engine *e = (or_engine *)ve;
	const uint32_t n = e->n, b = e->b;
	uint32_t *E = (uint32_t *)malloc(sizeof(uint32_t) * (n + 2));
	const size_t es = or_exponent(n, b, E);
	if (i0 <= 0) i0 = top_bit(E, es);
	/* cyclo.h:272 */
	const int L = 2 << (ilog2_32((uint32_t)i0) / 2), B_GL = (i0 - 1) / L + 1;

	int32_t *x = (int32_t *)calloc(n, sizeof(int32_t)), *d = (int32_t *)calloc(n, sizeof(int32_t));
	int32_t *t = (int32_t *)calloc(n, sizeof(int32_t));
	uint32_t *c1 = (uint32_t *)malloc(sizeof(uint32_t) * n), *c2 = (uint32_t *)malloc(sizeof(uint32_t) * n);
	x[0] = 1; d[0] = 1;   /* set(1); copy(1, 0) : cyclo.h:447-448 */

	/* cyclo.h:277-308 */
	for (int i = i0; i >= 0; --i)
	{
		int bit = bit_of(E, es, i);
		if (i == flip_at) bit ^= 1;
		or_engine_square_dup(e, bit, x);
		if ((i % B_GL == 0) && (i / B_GL != 0))
		{
			memcpy(t, x, sizeof(int32_t) * n);       /* copy(2, 0) */
			or_engine_mul(e, t, d);                   /* mul: x * d */
			memcpy(d, t, sizeof(int32_t) * n);       /* copy(1, 0); copy(0, 2) */
		}
	}

	int rc = or_get_int(n, b, x, canon);   /* getInt(0) */

	/* d(t+1) = d(t) * result : cyclo.h:317-320 */
	int32_t *dt1 = (int32_t *)malloc(sizeof(int32_t) * n);
	memcpy(dt1, x, sizeof(int32_t) * n);
	or_engine_mul(e, dt1, d);

	/* res = sum of the B-bit chunks of E : cyclo.h:322-338 */
	const int cw = (B_GL + 31) / 32;
	uint32_t *R = (uint32_t *)calloc(es + cw + 2, sizeof(uint32_t));
	uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * cw);
	size_t rs = 1;
	{
		const long nb = (long)es * 32;
		for (long base = 0; base < nb; base += B_GL)
		{
			/* chunk value as a multiword number, added into R */
			uint64_t carry = 0;
			size_t k = 0;
			memset(tmp, 0, sizeof(uint32_t) * cw);
			for (int q = 0; q < B_GL; ++q) if (bit_of(E, es, base + q)) tmp[q / 32] |= 1u << (q % 32);
			for (k = 0; k < (size_t)cw || carry; ++k)
			{
				const uint64_t s = (uint64_t)R[k] + (k < (size_t)cw ? tmp[k] : 0) + carry;
				R[k] = (uint32_t)s; carry = s >> 32;
			}
			if (k > rs) rs = k;
		}
	}
	const int i1 = top_bit(R, rs);

	/* 2^res * d(t)^(2^B) : cyclo.h:341-363 */
	int32_t *y = (int32_t *)calloc(n, sizeof(int32_t)); y[0] = 1;
	for (int i = i1; i >= B_GL; --i) or_engine_square_dup(e, bit_of(R, rs, i), y);
	or_engine_mul(e, y, d);
	for (int i = B_GL - 1; i >= 0; --i) or_engine_square_dup(e, (i <= i1) ? bit_of(R, rs, i) : 0, y);

	const int err = or_is_prime(n, b, canon, prm, res64);
	if (rc == 0 && or_get_int(n, b, y, c1) != 0) rc = -1;
	if (rc == 0 && or_get_int(n, b, dt1, c2) != 0) rc = -1;
	*gl_ok = (rc == 0) && (memcmp(c1, c2, sizeof(uint32_t) * n) == 0);
	if (rc == 0 && err) rc = 1;

	free(E); free(x); free(d); free(t); free(c1); free(c2); free(dt1); free(R); free(tmp); free(y);
	return rc;
}

/* ------------------------------------------------------------------ */
/* path B: ctsieve.cpp                                                  */
/* ------------------------------------------------------------------ */

/* ctsieve.cpp:25-65 */
int or_jacobi(uint64_t x, uint64_t y)
{
	static const int jac2[4] = { 1, -1, -1, 1 };
	static const int jac3[6] = { 1, 0, -1, -1, 0, 1 };
	static const int jac5[5] = { 1, -1, 0, -1, 1 };
	static const int jac6[12] = { 1, 0, 1, -1, 0, -1, -1, 0, -1, 1, 0, 1 };
	static const int jac7[14] = { 1, 1, -1, 0, 1, -1, -1, -1, -1, 1, 0, -1, 1, 1 };
	static const int jac8[4] = { 1, -1, -1, 1 };
	static const int jac9[9] = { 1, 0, 1, 1, 0, 1, 1, 0, 1 };
	const uint64_t y_2 = y / 2;
	if (x == 2) return jac2[y_2 % 4];
	if (x == 3) return jac3[y_2 % 6];
	if (x == 4) return 1;
	if (x == 5) return jac5[y_2 % 5];
	if (x == 6) return jac6[y_2 % 12];
	if (x == 7) return jac7[y_2 % 14];
	if (x == 8) return jac8[y_2 % 4];
	if (x == 9) return jac9[y_2 % 9];
	uint64_t m = x, n = y;
	int k = 1;
	while (m != 0)
	{
		int odd = 0;
		while (m % 2 == 0) { m /= 2; odd = !odd; }
		if (odd && (n % 8 != 1) && (n % 8 != 7)) k = -k;
		if (m == 1) return k;
		if ((m % 4 == 3) && (n % 4 == 3)) k = -k;
		const uint64_t t = n; n = m; m = t;
		m %= n;
	}
	return 0;
}

/* Mod class, ctsieve.cpp:68-172 */
/* Work counters for the sieve (a measurement aid, not part of the algorithm):
   Montgomery products, and the products of the strike verification. */
static uint64_t g_mp_products, g_md_products;
static inline uint64_t md_mul(uint64_t a, uint64_t b, uint64_t n) { ++g_md_products; return (uint64_t)(((u128)a * b) % n); }
static inline uint64_t md_sub(uint64_t a, uint64_t b, uint64_t n) { const uint64_t c = (a < b) ? n : 0; return a - b + c; }
static inline int ilog2_64(uint64_t x) { return 63 - __builtin_clzll(x); }
static uint64_t md_pow(uint64_t a, uint64_t e, uint64_t n)
{
	uint64_t r = a;
	for (int b = ilog2_64(e) - 1; b >= 0; --b)
	{
		r = md_mul(r, r, n);
		if ((e & ((uint64_t)1 << b)) != 0) r = md_mul(r, a, n);
	}
	return r;
}
static int md_spsp(uint64_t n, uint64_t p)
{
	uint64_t r = n - 1; int k = 0;
	for (; r % 2 == 0; r /= 2) ++k;
	uint64_t x = md_pow(p, r, n);
	if (x == 1) return 1;
	for (; k > 0; --k) { if (x == n - 1) return 1; x = md_mul(x, x, n); }
	return 0;
}
int or_isprime64(uint64_t n)
{
	if (n < 2) return 0;
	if (n % 2 == 0) return (n == 2);
	if (n < 9) return 1;
	if (!md_spsp(n, 2)) return 0;
	if (n < 2047ull) return 1;
	if (!md_spsp(n, 3)) return 0;
	if (n < 1373653ull) return 1;
	if (!md_spsp(n, 5)) return 0;
	if (n < 25326001ull) return 1;
	if (!md_spsp(n, 7)) return 0;
	if (n < 3215031751ull) return 1;
	if (!md_spsp(n, 11)) return 0;
	if (n < 2152302898747ull) return 1;
	if (!md_spsp(n, 13)) return 0;
	if (n < 3474749660383ull) return 1;
	if (!md_spsp(n, 17)) return 0;
	if (n < 341550071728321ull) return 1;
	if (!md_spsp(n, 19)) return 0;
	if (!md_spsp(n, 23)) return 0;
	if (n < 3825123056546413051ull) return 1;
	if (!md_spsp(n, 29)) return 0;
	if (!md_spsp(n, 31)) return 0;
	if (!md_spsp(n, 37)) return 0;
	return 1;
}

/* MpArith, ctsieve.cpp:175-395: 64-bit Montgomery */
typedef struct { uint64_t p, q, one, r2; } mpa;

static inline uint64_t mp_mulhi(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) >> 64); }
static inline uint64_t mp_redc(const mpa *m, u128 t)
{
	const uint64_t mp = mp_mulhi((uint64_t)t * m->q, m->p), t_hi = (uint64_t)(t >> 64), r = t_hi - mp;
	return (t_hi < mp) ? r + m->p : r;
}
static inline uint64_t mp_add(const mpa *m, uint64_t a, uint64_t b) { const uint64_t c = (a >= m->p - b) ? m->p : 0; return a + b - c; }
static inline uint64_t mp_sub(const mpa *m, uint64_t a, uint64_t b) { const uint64_t c = (a < b) ? m->p : 0; return a - b + c; }
static inline uint64_t mp_mul(const mpa *m, uint64_t a, uint64_t b) { ++g_mp_products; return mp_redc(m, (u128)a * b); }
static inline uint64_t mp_half(const mpa *m, uint64_t a) { return (a % 2 == 0) ? a / 2 : a / 2 + m->p / 2 + 1; }
static inline uint64_t mp_toint(const mpa *m, uint64_t r) { const uint64_t mp = mp_mulhi(r * m->q, m->p); return (mp != 0) ? m->p - mp : 0; }
static uint64_t mp_pow(const mpa *m, uint64_t a, uint64_t e)
{
	uint64_t r = a;
	for (int b = ilog2_64(e) - 1; b >= 0; --b)
	{
		r = mp_mul(m, r, r);
		if ((e & ((uint64_t)1 << b)) != 0) r = mp_mul(m, r, a);
	}
	return r;
}
static void mp_init(mpa *m, uint64_t p)
{
	uint64_t p_inv = 1, prev = 0;
	while (p_inv != prev) { prev = p_inv; p_inv *= 2 - p * p_inv; }
	m->p = p; m->q = p_inv; m->one = (-p) % p;
	uint64_t t = mp_add(m, m->one, m->one); t = mp_add(m, t, t);
	t = mp_add(m, t, t); t = mp_add(m, t, t);
	for (int i = 0; i < 4; ++i) t = mp_mul(m, t, t);
	m->r2 = t;
}
static inline uint64_t mp_tomp(const mpa *m, uint64_t n) { return mp_mul(m, n, m->r2); }

static int mp_prp(const mpa *m)
{
	const uint64_t e = (m->p - 1) / 2;
	int b = ilog2_64(e) - 1;
	uint64_t r = mp_add(m, m->one, m->one); r = mp_add(m, r, r);
	if ((e & ((uint64_t)1 << b)) != 0) r = mp_add(m, r, r);
	for (--b; b >= 0; --b)
	{
		r = mp_mul(m, r, r);
		if ((e & ((uint64_t)1 << b)) != 0) r = mp_add(m, r, r);
	}
	return (r == m->one) || (r == m->p - m->one);
}

/* ctsieve.cpp:264-309 */
static uint64_t mp_sqrt(const mpa *m, uint64_t a)
{
	const uint64_t p = m->p;
	if ((a == 0) || (a == m->one)) return a;
	if (p % 4 == 3) return mp_pow(m, a, (p + 1) / 4);
	if (p % 8 == 5)
	{
		const uint64_t b = mp_add(m, a, a);
		const uint64_t v = mp_pow(m, b, (p - 5) / 8);
		const uint64_t i = mp_mul(m, b, mp_mul(m, v, v));
		return mp_mul(m, mp_mul(m, a, v), mp_sub(m, i, m->one));
	}
	uint64_t q = p - 1; unsigned int e = 0; while (q % 2 == 0) { q /= 2; ++e; }
	uint64_t z = 3; while (or_jacobi(z, p) != -1) ++z;
	z = mp_pow(m, mp_tomp(m, z), q);
	uint64_t y = z;
	unsigned int r = e;
	uint64_t x = mp_pow(m, a, (q - 1) / 2);
	uint64_t b = mp_mul(m, a, mp_mul(m, x, x));
	x = mp_mul(m, a, x);
	for (int j = 0; j < 64; ++j)
	{
		if (b == m->one) return x;
		unsigned int mm = 1;
		uint64_t t = b; while (mm < r) { t = mp_mul(m, t, t); if (t == m->one) break; ++mm; }
		if (mm == r) break;
		t = y; for (unsigned int i = 0; i < r - 1 - mm; ++i) t = mp_mul(m, t, t);
		y = mp_mul(m, t, t);
		r = mm;
		x = mp_mul(m, x, t);
		b = mp_mul(m, b, y);
	}
	return 0;
}

/* ctsieve.cpp:311-324; returns -1 for "Calculation error" */
static int mp_sqrt_checked(const mpa *m, uint64_t a, uint64_t *out)
{
	const uint64_t s = mp_sqrt(m, a);
	if (mp_mul(m, s, s) != a)
	{
		if (or_isprime64(m->p)) return -1;
		*out = 0; return 0;
	}
	*out = s; return 0;
}

typedef struct { uint64_t *v; size_t n, cap; } u64vec;
static void vpush(u64vec *L, uint64_t x)
{
	if (L->n == L->cap) { L->cap = L->cap ? 2 * L->cap : 16; L->v = (uint64_t *)realloc(L->v, L->cap * sizeof(uint64_t)); }
	L->v[L->n++] = x;
}

/* ctsieve.cpp:326-379 */
static int mp_sqrtn(const mpa *m, uint64_t a, int n, u64vec *L)
{
	const uint64_t p = m->p;
	int e = 1; for (uint64_t q = (p - 1) / 2; q % 2 == 0; q /= 2) e++;
	if (n < e) e = n;
	if ((e > 1) && mp_pow(m, a, (p - 1) >> e) != m->one) return 0;
	uint64_t r = a, q = (p - 1) / 2, mm = 1;
	for (int i = 1; i <= n; ++i)
	{
		if (q % 2 == 0)
		{
			if (mp_sqrt_checked(m, r, &r) < 0) return -1;
			q /= 2;
			if (r == 0) break;
		}
		else
		{
			if (mm % 2 != 0) mm += q;
			mm /= 2;
		}
	}
	if (mm > 1) r = mp_pow(m, r, mm);
	if (r == 0) return or_isprime64(p) ? -1 : 0;
	uint64_t u = 2;
	while (1)
	{
		if (or_jacobi(u, p) == -1) break;
		++u; if ((u == 4) || (u == 9)) ++u;
	}
	u = mp_pow(m, mp_tomp(m, u), (p - 1) >> e);
	const size_t s = (size_t)1 << e;
	for (size_t i = 0; i < s; i += 2)
	{
		const uint64_t b = mp_toint(m, r);
		vpush(L, b); vpush(L, p - b);
		r = mp_mul(m, r, u);
	}
	return 0;
}

/* ctsieve.cpp:653-676: 1 = continue, 0 = stop (composite p), -1 = error */
static int check_root(uint8_t *struck, uint64_t b, uint64_t b_min, uint64_t b_max, uint64_t p, int n, int is_pos)
{
	const uint64_t res = is_pos ? (p - 1) : 1;
	for (uint64_t s = b; s <= b_max; s += p)
	{
		if (s >= b_min && !struck[s - b_min])
		{
			const uint64_t x = md_pow(s, (uint64_t)1 << (n - 1), p), r = md_sub(md_mul(x, x, p), x, p);
			if (r == res) struck[s - b_min] = 1;
			else if (or_isprime64(p)) return -1;
			else return 0;
		}
	}
	return 1;
}

/* ctsieve.cpp:678-726 */
static int check_pos(int n, uint64_t b_min, uint64_t b_max, uint64_t p_min, uint64_t p_max, uint8_t *struck)
{
	const uint64_t k_min = (p_min >> n) / 3;
	uint64_t k_max = (p_max >> n) / 3;
	if (p_max != (uint64_t)-1) while (3 * (k_max << n) + 1 < p_max) ++k_max;
	for (uint64_t k = k_min; k <= k_max; ++k)
	{
		const uint64_t p = 3 * (k << n) + 1;
		if ((p % 5 == 0) || (p % 7 == 0) || (p % 11 == 0) || (p % 13 == 0) || (p % 17 == 0)
		 || (p % 19 == 0) || (p % 23 == 0) || (p % 29 == 0) || (p % 31 == 0)) continue;
		mpa m; mp_init(&m, p);
		if (!mp_prp(&m)) continue;
		uint64_t a = 5, ma = mp_add(&m, mp_add(&m, mp_add(&m, m.one, m.one), mp_add(&m, m.one, m.one)), m.one);
		while (1)
		{
			if ((or_jacobi(a, p) == -1) && (mp_pow(&m, ma, k << (n - 1)) != p - m.one)) break;
			++a; ma = mp_add(&m, ma, m.one);
			if (a % 3 == 0) { ++a; ma = mp_add(&m, ma, m.one); }
		}
		const uint64_t c = mp_pow(&m, ma, k), b2 = mp_mul(&m, c, c), b4 = mp_mul(&m, b2, b2);
		uint64_t b = mp_toint(&m, c);
		for (uint64_t i = 1; i < ((uint64_t)3 << (n - 1)); i += 6)
		{
			int rc;
			if ((rc = check_root(struck, b, b_min, b_max, p, n, 1)) <= 0) { if (rc < 0) return -1; break; }
			if ((rc = check_root(struck, p - b, b_min, b_max, p, n, 1)) <= 0) { if (rc < 0) return -1; break; }
			b = mp_mul(&m, b, b4);
			if ((rc = check_root(struck, b, b_min, b_max, p, n, 1)) <= 0) { if (rc < 0) return -1; break; }
			if ((rc = check_root(struck, p - b, b_min, b_max, p, n, 1)) <= 0) { if (rc < 0) return -1; break; }
			b = mp_mul(&m, b, b2);
		}
	}
	return 0;
}

/* ctsieve.cpp:728-787 */
static int check_neg(int n, uint64_t b_min, uint64_t b_max, uint64_t p_min, uint64_t p_max, uint8_t *struck)
{
	const uint64_t k_min = p_min / 10;
	uint64_t k_max = p_max / 10;
	if (p_max != (uint64_t)-1) while (10 * k_max + 1 < p_max) ++k_max;
	u64vec L = { 0, 0, 0 }, L1 = { 0, 0, 0 }, L2 = { 0, 0, 0 };
	int ret = 0;
	for (uint64_t k = k_min; k <= k_max && ret == 0; ++k)
	{
		for (uint64_t i = 0; i <= 1; ++i)
		{
			const uint64_t p = 10 * k + 2 * i - 1;
			if ((p % 3 == 0) || (p % 7 == 0)) continue;
			if (p > 31)
			{
				if ((p % 11 == 0) || (p % 13 == 0) || (p % 17 == 0) || (p % 19 == 0)
				 || (p % 23 == 0) || (p % 29 == 0) || (p % 31 == 0)) continue;
			}
			mpa m; mp_init(&m, p);
			if (!(mp_prp(&m) && (p != 1093ull * 1093) && (p != 3511ull * 3511))) continue;
			L.n = 0;
			uint64_t s5;
			if (mp_sqrt_checked(&m, mp_tomp(&m, 5), &s5) < 0) { ret = -1; break; }
			if (s5 != 0)
			{
				const uint64_t r = mp_half(&m, mp_sub(&m, m.one, s5));
				if (p % 4 == 3)
				{
					const uint64_t rs = (or_jacobi(mp_toint(&m, r), p) != 1) ? mp_sub(&m, m.one, r) : r;
					if (mp_sqrtn(&m, rs, n - 1, &L) < 0) { ret = -1; break; }
				}
				else if (or_jacobi(mp_toint(&m, r), p) == 1)
				{
					L1.n = 0; L2.n = 0;
					if (mp_sqrtn(&m, r, n - 1, &L1) < 0) { ret = -1; break; }
					if (mp_sqrtn(&m, mp_sub(&m, m.one, r), n - 1, &L2) < 0) { ret = -1; break; }
					for (size_t t = 0; t < L1.n; ++t) vpush(&L, L1.v[t]);
					for (size_t t = 0; t < L2.n; ++t) vpush(&L, L2.v[t]);
				}
			}
			for (size_t t = 0; t < L.n; ++t)
			{
				const int rc = check_root(struck, L.v[t], b_min, b_max, p, n, 0);
				if (rc < 0) { ret = -1; break; }
				if (rc == 0) break;
			}
			if (ret) break;
		}
	}
	free(L.v); free(L1.v); free(L2.v);
	return ret;
}

int or_sieve_window(int m, uint64_t b_min, uint64_t b_max, uint64_t p_min_pos, uint64_t p_max_pos,
                    uint64_t p_min_neg, uint64_t p_max_neg, int mode, uint8_t *struck)
{
	if (mode >= 0 && check_pos(m, b_min, b_max, p_min_pos, p_max_pos, struck) < 0) return -1;
	if (mode <= 0 && check_neg(m, b_min, b_max, p_min_neg, p_max_neg, struck) < 0) return -1;
	return 0;
}

/* ------------------------------------------------------------------ */
/* GPU factor emission: sieve/GPU/ocl/kernel.cl, sieve/GPU/src/ctsieve.h */
/* ------------------------------------------------------------------ */

/* kernel.cl:88-100: base-2 Euler test in Montgomery form */
static int fe_prp(const mpa *m) { return mp_prp(m); }

/* kernel.cl:103-128 */
static int fe_jacobi(uint64_t x, uint64_t y) { return or_jacobi(x, y); }

/* kernel.cl:131-140 (p = 3 mod 4), :142-166 (p = 5 mod 8 Atkin), Cipolla :142-166 */
static uint64_t fe_sqrt_3_4(const mpa *m, uint64_t a) { return mp_pow(m, a, (m->p + 1) / 4); }
static uint64_t fe_sqrt_5_8(const mpa *m, uint64_t a)
{
	const uint64_t b = mp_add(m, a, a);
	const uint64_t v = mp_pow(m, b, (m->p - 5) / 8);
	const uint64_t i = mp_mul(m, b, mp_mul(m, v, v));
	return mp_mul(m, mp_mul(m, a, v), mp_sub(m, i, m->one));
}
static uint64_t fe_sqrt_cipolla(const mpa *m, uint64_t a)
{
	const uint64_t p = m->p;
	if ((a == 0) || (a == m->one)) return a;
	uint32_t d = 2; uint64_t md = mp_add(m, m->one, m->one), ai = p - mp_toint(m, a);
	while (d < 256)
	{
		if (fe_jacobi((uint64_t)d * d + ai, p) == -1) break;
		++d; md = mp_add(m, md, m->one);
	}
	if (d >= 256) return 0;
	const uint64_t w2 = mp_sub(m, mp_mul(m, md, md), a), e = (p + 1) / 2;
	uint64_t x = md, y = m->one;
	for (int b = ilog2_64(e) - 1; b >= 0; --b)
	{
		const uint64_t t = mp_add(m, mp_mul(m, x, x), mp_mul(m, mp_mul(m, y, y), w2));
		y = mp_mul(m, mp_add(m, x, x), y); x = t;
		if ((e & ((uint64_t)1 << b)) != 0)
		{
			const uint64_t t2 = mp_add(m, mp_mul(m, x, md), mp_mul(m, y, w2));
			y = mp_add(m, x, mp_mul(m, y, md)); x = t2;
		}
	}
	return (mp_mul(m, x, x) == a) ? x : 0;
}

/* kernel.cl:168-231 sqrtn variants: sel = 0 (3 mod 4), 1 (5 mod 8), 2 (Cipolla) */
static uint64_t fe_sqrtn(const mpa *m, uint64_t a, uint32_t n, int sel)
{
	const uint64_t p = m->p;
	uint64_t r = a, k = (p - 1) / 2, mm = 1;
	if (sel == 0)
	{
		for (uint32_t i = 1; i <= n; ++i) { if (mm % 2 != 0) mm += k; mm /= 2; }
		if (mm > 1) r = mp_pow(m, r, mm);
		return r;
	}
	uint32_t e = 1; for (uint64_t q = (p - 1) / 2; q % 2 == 0; q /= 2) ++e;
	if (n < e) e = n;
	if ((e > 1) && mp_pow(m, a, (p - 1) >> e) != m->one) return 0;
	for (uint32_t i = 1; i <= n; ++i)
	{
		if (k % 2 == 0)
		{
			r = (sel == 1) ? fe_sqrt_5_8(m, r) : fe_sqrt_cipolla(m, r);
			k /= 2;
			if (r == 0) break;
		}
		else { if (mm % 2 != 0) mm += k; mm /= 2; }
	}
	if (mm > 1) r = mp_pow(m, r, mm);
	return r;
}

typedef struct { uint64_t *pairs; uint64_t cap, count; int m; int mode; int err; } fe_out;

/* kernel.cl:233-241 push_factor + ctsieve.h:270-289 host verification */
static void fe_push(fe_out *o, uint64_t p, uint64_t b, uint64_t bmax)
{
	if (b > bmax) return;
	const uint64_t x = md_pow(b, (uint64_t)1 << (o->m - 1), p), r = md_sub(md_mul(x, x, p), x, p);
	if (!(((o->mode > 0) && (r == p - 1)) || ((o->mode < 0) && (r == 1))))
	{
		if (or_isprime64(p)) o->err = 1;
		return;
	}
	if (o->count < o->cap) { o->pairs[2 * o->count] = p; o->pairs[2 * o->count + 1] = b; }
	++o->count;
}

int or_sieve_factors(int m, uint64_t p_min, uint64_t p_max, int mode, uint64_t *pairs, uint64_t cap, uint64_t *count)
{
	fe_out o = { pairs, cap, 0, m, mode, 0 };
	if (mode > 0)
	{
		for (uint64_t k = (p_min >> m) / 3; ; ++k)
		{
			const uint64_t p = 3 * (k << m) + 1;
			if (p > p_max) break;
			if (p < p_min) continue;
			mpa M; mp_init(&M, p);
			if (!fe_prp(&M)) continue;
			/* init_factors_pos, kernel.cl:258-288: odd a >= 5, a != 0 mod 3 */
			const uint64_t one = M.one, two = mp_add(&M, one, one), four = mp_add(&M, two, two);
			const uint64_t pm1_6 = (p - 1) / 6;
			uint32_t a = 5; uint64_t ma = mp_add(&M, four, one);
			while (a < 256)
			{
				if ((fe_jacobi(p % a, a) == -1) && (mp_pow(&M, ma, pm1_6) != p - one)) break;
				a += 2; ma = mp_add(&M, ma, two);
				if ((fe_jacobi(p % a, a) == -1) && (mp_pow(&M, ma, pm1_6) != p - one)) break;
				a += 4; ma = mp_add(&M, ma, four);
			}
			if (a >= 256) continue;
			uint64_t b = mp_pow(&M, ma, pm1_6 >> (m - 1));
			/* generate_factors_pos, kernel.cl:290-320 */
			const uint64_t b2 = mp_mul(&M, b, b), b4 = mp_mul(&M, b2, b2);
			b = mp_toint(&M, b);
			const uint64_t bmax = (p < 10000000000000000ull) ? 5000000000ull : 10000000000ull;
			for (uint32_t j = 1; j < (3u << (m - 1)); j += 6)
			{
				fe_push(&o, p, b, bmax); fe_push(&o, p, p - b, bmax);
				b = mp_mul(&M, b, b4);
				fe_push(&o, p, b, bmax); fe_push(&o, p, p - b, bmax);
				b = mp_mul(&M, b, b2);
			}
		}
	}
	else
	{
		for (uint64_t k = p_min / 10; ; ++k)
		{
			if (10 * k - 1 > p_max && k > 0) break;
			for (int i = 0; i <= 1; ++i)
			{
				if (k == 0 && i == 0) continue;
				const uint64_t p = 10 * k + 2 * (uint64_t)i - 1;
				if (p < p_min || p > p_max || p < 11) continue;
				mpa M; mp_init(&M, p);
				if (!fe_prp(&M)) continue;
				const uint64_t one = M.one, two = mp_add(&M, one, one), five = mp_add(&M, mp_add(&M, two, two), one);
				uint64_t r1 = 0, r2 = 0;
				if (p % 4 == 3)
				{   /* init_3_4_factors_neg, kernel.cl:363-381 */
					const uint64_t s5 = fe_sqrt_3_4(&M, five);
					if (s5 == 0) continue;
					const uint64_t r = mp_half(&M, mp_sub(&M, one, s5));
					const uint64_t rs = (fe_jacobi(mp_toint(&M, r), p) == 1) ? r : mp_sub(&M, one, r);
					r1 = fe_sqrtn(&M, rs, m - 1, 0);
				}
				else
				{   /* init_5_8_factors_neg :383-401, init_factors_neg :403-421 */
					const int sel = (p % 8 == 5) ? 1 : 2;
					const uint64_t s5 = (sel == 1) ? fe_sqrt_5_8(&M, five) : fe_sqrt_cipolla(&M, five);
					if (s5 == 0) continue;
					const uint64_t r = mp_half(&M, mp_sub(&M, one, s5));
					if (fe_jacobi(mp_toint(&M, r), p) != 1) continue;
					r1 = fe_sqrtn(&M, r, m - 1, sel);
					r2 = fe_sqrtn(&M, mp_sub(&M, one, r), m - 1, sel);
				}
				/* generate_factors_neg, kernel.cl:423-471 */
				const uint64_t bmax = (p < 10000000000000ull) ? 5000000000ull : 10000000000ull;
				uint32_t e = 1; for (uint64_t q = (p - 1) / 2; q % 2 == 0; q /= 2) ++e;
				if (e > (uint32_t)(m - 1)) e = (uint32_t)(m - 1);
				uint32_t d = 2; uint64_t md = two;
				while (d < 256)
				{
					if (fe_jacobi(d, p) == -1) break;
					++d; md = mp_add(&M, md, one); if ((d == 4) || (d == 9)) { ++d; md = mp_add(&M, md, one); }
				}
				if (d >= 256) continue;
				const uint64_t u = mp_pow(&M, md, (p - 1) >> e);
				for (int w = 0; w < 2; ++w)
				{
					const uint64_t rr = w ? r2 : r1;
					if (rr == 0) continue;
					uint64_t b = mp_toint(&M, rr);
					for (uint32_t t = 0, s = 1u << e; t < s; t += 2)
					{
						fe_push(&o, p, b, bmax); fe_push(&o, p, p - b, bmax);
						b = mp_mul(&M, b, u);
					}
				}
			}
		}
	}
	*count = o.count;
	return o.err ? -1 : 0;
}

/* Reads (and with reset != 0 clears) the sieve's work counters. */
void or_sieve_products(int reset, uint64_t *mont, uint64_t *verify)
{
	*mont = g_mp_products; *verify = g_md_products;
	if (reset) { g_mp_products = 0; g_md_products = 0; }
}
