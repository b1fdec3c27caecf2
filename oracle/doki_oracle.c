/*
 * doki_oracle.c — CPU restatement of Doki's state-vector core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is what the tests compare the HIP
 * product path against; it is never linked into, loaded by, or called
 * from the product library or the `doki` extension.  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * It restates, in single-threaded C99, the algorithm of the reference
 * (Mowstyl/Doki 1.6.1).  Every function cites the reference lines it follows.
 * The arithmetic is written with the same C99 `double _Complex` operations in
 * the same order as the reference, so that, built with gcc -O3 without
 * fast-math (as the reference is, meson.build:7-14), its results equal the
 * reference's bit for bit when the reference runs with one OpenMP thread.
 * That equality is what the golden fixtures under tests/golden/ pin.
 *
 * Storage convention (qstate.h:44-59, :84-86): a state holds raw amplitudes
 * and a normalisation constant; every read returns raw / norm.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

typedef struct orc_state {
	unsigned nq;      /* number of qubits; 0 = empty (fully measured) */
	int64_t len;      /* 2^nq amplitudes */
	cplx *amp;        /* raw amplitudes, index bit q = qubit q */
	double norm;      /* reads divide by this (qstate.h:86) */
	int phase_cached; /* get_global_phase cache (qstate.h:55-58) */
	double phase;
} orc_state;

#define ORC_MAX_QUBITS 62 /* log2_64(INTMAX_MAX), platform.h:270 */

/* state_get, qstate.h:86: complex divided by the real norm constant. */
static inline cplx rd(const orc_state *s, int64_t i)
{
	return s->amp[i] / s->norm;
}

/* state_init, qstate.c:23-88 (one chunk; norm 1; phase cache unset). */
static orc_state *orc_alloc(unsigned nq, int zero_state)
{
	orc_state *s;
	if (nq > ORC_MAX_QUBITS)
		return NULL;
	s = (orc_state *)calloc(1, sizeof(*s));
	if (!s)
		return NULL;
	s->nq = nq;
	s->len = (int64_t)1 << nq;
	s->norm = 1.0;
	s->phase_cached = 0;
	s->phase = -10.0;
	s->amp = zero_state ? (cplx *)calloc((size_t)s->len, sizeof(cplx))
			    : (cplx *)malloc((size_t)s->len * sizeof(cplx));
	if (!s->amp) {
		free(s);
		return NULL;
	}
	if (zero_state)
		s->amp[0] = 1.0;
	return s;
}

orc_state *orc_new(unsigned nq, int zero_state)
{
	return orc_alloc(nq, zero_state);
}

/* registry_new_data, doki.c:625-720: values stored as given, norm 1 (Q3). */
orc_state *orc_from_data(unsigned nq, const double *interleaved)
{
	orc_state *s = orc_alloc(nq, 0);
	int64_t i;
	if (!s)
		return NULL;
	for (i = 0; i < s->len; i++)
		s->amp[i] = interleaved[2 * i] + I * interleaved[2 * i + 1];
	return s;
}

/* state_clone, qstate.c:90-105: destination holds the normalised values. */
orc_state *orc_clone(const orc_state *src)
{
	orc_state *d = orc_alloc(src->nq, 0);
	int64_t i;
	if (!d)
		return NULL;
	for (i = 0; i < src->len; i++)
		d->amp[i] = rd(src, i);
	return d;
}

void orc_free(orc_state *s)
{
	if (!s)
		return;
	free(s->amp);
	free(s);
}

unsigned orc_num_qubits(const orc_state *s)
{
	return s->nq;
}

double orc_norm(const orc_state *s)
{
	return s->norm;
}

/* Normalised read-out of every amplitude (state_get over all i). */
void orc_read(const orc_state *s, double *out)
{
	int64_t i;
	for (i = 0; i < s->len; i++) {
		cplx v = rd(s, i);
		out[2 * i] = creal(v);
		out[2 * i + 1] = cimag(v);
	}
}

/* Raw stored amplitudes (no division), for storage-level checks. */
void orc_read_raw(const orc_state *s, double *out)
{
	int64_t i;
	for (i = 0; i < s->len; i++) {
		out[2 * i] = creal(s->amp[i]);
		out[2 * i + 1] = cimag(s->amp[i]);
	}
}

/*
 * apply_gate, qops.c:176-261.  M is the 2^k x 2^k gate, row-major,
 * interleaved re/im; row = output basis state (Q2), targets[0] is the gate's
 * least significant bit (Q1).  Amplitudes whose controls are not all 1 or
 * whose anticontrols are not all 0 are copied (normalised, Q4).  The output
 * norm constant is sqrt(sum |out_i|^2) (Q3).
 */
orc_state *orc_apply(const orc_state *s, unsigned k, const double *M,
		     const unsigned *targets, const unsigned *ctrl,
		     unsigned nctrl, const unsigned *actrl, unsigned nactrl)
{
	orc_state *o;
	int64_t cmask = 0, amask = 0, i, idx;
	unsigned gsz = 1u << k, r, c, b, row;
	cplx acc, *G;
	double nsum = 0;

	o = orc_alloc(s->nq, 0);
	if (!o)
		return NULL;
	G = (cplx *)malloc((size_t)gsz * gsz * sizeof(cplx));
	if (!G) {
		orc_free(o);
		return NULL;
	}
	for (r = 0; r < gsz * gsz; r++)
		G[r] = M[2 * r] + I * M[2 * r + 1];
	for (b = 0; b < nctrl; b++)
		cmask |= (int64_t)1 << ctrl[b];
	for (b = 0; b < nactrl; b++)
		amask |= (int64_t)1 << actrl[b];

	for (i = 0; i < s->len; i++) {
		if ((i & cmask) == cmask && (i & amask) == 0) {
			row = 0;
			for (b = 0; b < k; b++)
				row |= (unsigned)((i >> targets[b]) & 1) << b;
			acc = 0;
			idx = i;
			for (c = 0; c < gsz; c++) {
				for (b = 0; b < k; b++) {
					if ((c >> b) & 1)
						idx |= (int64_t)1 << targets[b];
					else
						idx &= ~((int64_t)1 << targets[b]);
				}
				acc = acc + rd(s, idx) * G[(size_t)row * gsz + c];
			}
		} else {
			acc = rd(s, i);
		}
		o->amp[i] = acc;
		nsum += pow(creal(acc), 2) + pow(cimag(acc), 2);
	}
	o->norm = sqrt(nsum);
	free(G);
	return o;
}

/*
 * probability, qops.c:72-95: P(qubit q = 1), summed over the 2^(n-1)
 * indices with bit q inserted as 1, in increasing order.
 */
double orc_prob(const orc_state *s, unsigned q)
{
	int64_t half = s->len >> 1, bit = (int64_t)1 << q, lo = bit - 1, j;
	double p = 0;
	for (j = 0; j < half; j++) {
		cplx v = rd(s, ((j & ~lo) << 1) + bit + (j & lo));
		p += creal(v) * creal(v) + cimag(v) * cimag(v);
	}
	return p;
}

/*
 * collapse, qops.c:138-174: keep the half with bit q == value, drop qubit q
 * (higher qubits shift down), store the parent's normalised values, and set
 * norm = sqrt(P) with P = p1 or 1 - p1 (Q6).  A 1-qubit state collapses to
 * the empty state (nq = 0).
 */
orc_state *orc_collapse(const orc_state *s, unsigned q, int value, double p1)
{
	orc_state *o;
	int64_t bit = (int64_t)1 << q, lo = bit - 1, off, j;
	if (s->nq == 1) {
		o = (orc_state *)calloc(1, sizeof(*o));
		return o; /* nq = 0, no amplitudes */
	}
	o = orc_alloc(s->nq - 1, 0);
	if (!o)
		return NULL;
	off = value ? bit : 0;
	if (!value)
		p1 = 1 - p1;
	for (j = 0; j < o->len; j++)
		o->amp[j] = rd(s, ((j & ~lo) << 1) + off + (j & lo));
	o->norm = sqrt(p1);
	return o;
}

/*
 * The measurement loop of registry_measure, doki.c:1023-1192, with measure,
 * qops.c:124-136.  The state is cloned first (doki.c:1076), qubits are
 * visited from the most significant down (doki.c:1097-1098), a roll is taken
 * only for masked qubits (doki.c:1110-1121) and the outcome is P1 > roll
 * (qops.c:132).  res[i] describes qubit nq-1-i: -1 not measured, 0 or 1.
 * Returns 0 on success; on success *out is the remaining state, or NULL when
 * every qubit was measured (doki.c:1180-1190).
 * Errors: 1 roll missing or outside [0, 1) (doki.c:1112-1120),
 *         2 zero normalisation after collapse (doki.c:1139-1149),
 *         3 allocation failure.
 */
int orc_measure(const orc_state *s, uint64_t mask, const double *rolls,
		int nrolls, signed char *res, orc_state **out)
{
	orc_state *cur, *nxt;
	unsigned n = s->nq, i, q, measured = 0;
	int ri = 0;
	*out = NULL;
	cur = orc_clone(s);
	if (!cur)
		return 3;
	for (i = 0; i < n; i++) {
		q = n - i - 1;
		res[i] = -1;
		if (!((mask >> q) & 1))
			continue;
		if (ri >= nrolls || rolls[ri] < 0 || rolls[ri] >= 1) {
			orc_free(cur);
			return 1;
		}
		{
			double roll = rolls[ri++];
			double p1 = orc_prob(cur, q);
			int v = p1 > roll;
			nxt = orc_collapse(cur, q, v, p1);
			if (!nxt) {
				orc_free(cur);
				return 3;
			}
			if (nxt->nq > 0 && nxt->norm == 0.0) {
				orc_free(nxt);
				orc_free(cur);
				return 2;
			}
			res[i] = (signed char)v;
		}
		measured++;
		orc_free(cur);
		cur = nxt;
	}
	if (n - measured > 0) {
		*out = cur;
	} else {
		orc_free(cur);
		*out = NULL;
	}
	return 0;
}

/*
 * join, qops.c:97-122: r[i * 2^n2 + j] = a_i * b_j with the first argument
 * in the high qubits (Q11); normalised values, norm 1.
 */
orc_state *orc_join(const orc_state *most, const orc_state *least)
{
	orc_state *o = orc_alloc(most->nq + least->nq, 0);
	int64_t i, j;
	if (!o)
		return NULL;
	for (i = 0; i < most->len; i++) {
		cplx a = rd(most, i);
		for (j = 0; j < least->len; j++)
			o->amp[i * least->len + j] = a * rd(least, j);
	}
	return o;
}

/*
 * get_global_phase, qops.c:46-70: argument of the first nonzero normalised
 * amplitude, but only when its imaginary part is nonzero; 0 otherwise (Q8).
 * Cached on the state.
 */
double orc_phase(orc_state *s)
{
	int64_t i;
	double ph = 0.0;
	if (s->phase_cached)
		return s->phase;
	for (i = 0; i < s->len; i++) {
		cplx v = rd(s, i);
		if (creal(v) != 0. || cimag(v) != 0.) {
			if (cimag(v) != 0.)
				ph = carg(v);
			break;
		}
	}
	s->phase = ph;
	s->phase_cached = 1;
	return ph;
}

/*
 * registry_get, doki.c:546-595: the normalised amplitude, multiplied by
 * exp(-i phase) when canonical.
 */
void orc_get(orc_state *s, int64_t id, int canonical, double *out)
{
	cplx v = rd(s, id);
	if (canonical) {
		double ph = orc_phase(s);
		cplx rot = cos(ph) + I * -sin(ph);
		v = v * rot;
	}
	out[0] = creal(v);
	out[1] = cimag(v);
}

/*
 * Density-matrix element of a chain of partial traces, following
 * density_matrix / _densityFun (qops.c:266-293, 580-601) and
 * partial_trace / _PartialTFunct / _GetElemIndex (funmatrix.c:905-944,
 * 999-1037), evaluated through getitem's rules (funmatrix.c:238-340):
 * each level is checked against its shape and scaled by its scalar (1).
 *
 * chain[0..depth-1] are the traced positions in order of application: the
 * first partial trace acts on the full density matrix.  Returns 0 on
 * success, 7 for an element outside a level's shape and 8 for a NaN result
 * (getitem's codes, funmatrix.h:80-93).
 */
static int64_t ins_bit(int value, int64_t pos, int bit)
{
	int64_t step = (int64_t)1 << bit;
	return pos % step + (pos / step) * (step << 1) + value * step;
}

static int dens_get(const orc_state *s, const int *chain, int depth,
		    int64_t i, int64_t j, cplx *sol)
{
	int64_t dim = s->len >> depth;
	const cplx one = 1.0;
	*sol = NAN;
	if (!(i < dim && j < dim))
		return 7;
	if (depth == 0) {
		*sol = rd(s, i) * conj(rd(s, j));
	} else {
		cplx a = 0, b = 0;
		int e = chain[depth - 1];
		cplx val = 0;
		if (dens_get(s, chain, depth - 1, ins_bit(0, i, e),
			     ins_bit(0, j, e), &a) == 0 &&
		    dens_get(s, chain, depth - 1, ins_bit(1, i, e),
			     ins_bit(1, j, e), &b) == 0)
			val = a + b;
		else
			val = a; /* first term only, NaN when it failed */
		*sol = val;
	}
	if (isnan(creal(*sol)) || isnan(cimag(*sol)))
		return 8;
	*sol = *sol * one;
	return 0;
}

int orc_density_get(const orc_state *s, const int *chain, int depth,
		    int64_t i, int64_t j, double *out)
{
	cplx v = 0;
	int rc = dens_get(s, chain, depth, i, j, &v);
	out[0] = creal(v);
	out[1] = cimag(v);
	return rc;
}

/* Whole reduced matrix (dim = 2^(nq-depth)) in row-major interleaved form. */
int orc_density_matrix(const orc_state *s, const int *chain, int depth,
		       double *out)
{
	int64_t dim = s->len >> depth, i, j;
	for (i = 0; i < dim; i++)
		for (j = 0; j < dim; j++) {
			int rc = orc_density_get(s, chain, depth, i, j,
						 out + 2 * (i * dim + j));
			if (rc)
				return rc;
		}
	return 0;
}

/* funmatrix_trace, doki.c:2131-2215: diagonal summed in order. */
int orc_density_trace(const orc_state *s, const int *chain, int depth,
		      double *out)
{
	int64_t dim = s->len >> depth, i;
	cplx acc = 0, v;
	for (i = 0; i < dim; i++) {
		int rc = dens_get(s, chain, depth, i, i, &v);
		if (rc)
			return rc;
		acc = acc + v;
	}
	out[0] = creal(acc);
	out[1] = cimag(acc);
	return 0;
}

/* state_mem_size, qstate.c:123-135 on LP64 with one chunk (Q17). */
uint64_t orc_mem_size(unsigned nq)
{
	return 64u + 16u * ((uint64_t)1 << nq);
}
