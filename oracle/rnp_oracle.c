/*
 * rnp_oracle.c -- CPU restatement of RNP2's LA::QR path (TEST INFRASTRUCTURE,
 * not the product).  See rnp_oracle.h for the list of restated reference
 * functions with their file:line.  Built by oracle/Makefile into
 * oracle/liboracle.so; loaded only by tests/, smoke() and bench.py's
 * cpu_baseline leg.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include <complex.h>
#include "rnp_oracle.h"

/* RNPAssert, RNP/Debug.hpp:9-17: print and continue. */
static void orc_assert_fail(const char *cond, int line){
	fprintf(stderr, "Assertion failed in rnp_oracle_gen.h:%d: %s\n", line, cond);
}

/* Traits<real>::hypot3, RNP/Types.hpp:73-92. */
double orc_hypot3(double x, double y, double z){
	double xabs = fabs(x), yabs = fabs(y), zabs = fabs(z);
	double w = xabs;
	if(yabs > w){ w = yabs; }
	if(zabs > w){ w = zabs; }
	if(0.0 == w){
		return xabs + yabs + zabs;
	}
	xabs /= w; yabs /= w; zabs /= w;
	return w * sqrt(xabs*xabs + yabs*yabs + zabs*zabs);
}

/* ---------------------------------------------------------------------
 * Random::UniformRealVector, RNP/Random.hpp:45-183.
 * 48-bit multiplicative congruential generator, multiplier
 * 33952834046453, modulus 2^48.  The reference keeps a table of the first
 * 128 powers of the multiplier split into 12-bit limbs (lines 58-123); here
 * the powers are computed on first use.  Value j of a call is
 * seed * a^((j mod 128)+1) mod 2^48, the seed advancing by a^128 every 128
 * values and to the last value at the end of the call (lines 171-182), so
 * successive calls continue one stream.  The conversion
 * 1 - v/2^48 is exact in double (lines 157-169).
 * ------------------------------------------------------------------- */
#define ORC_MCG_A 33952834046453ULL
#define ORC_MASK48 ((1ULL<<48)-1)
static unsigned long long orc_mcg_pow[128];
static int orc_mcg_ready = 0;
static int orc_internal_iseed[4] = {2398, 691, 782, 721};

static void orc_mcg_init(void){
	if(orc_mcg_ready){ return; }
	unsigned __int128 p = 1;
	for(int i = 0; i < 128; ++i){
		p = (p * ORC_MCG_A) & ORC_MASK48;
		orc_mcg_pow[i] = (unsigned long long)p;
	}
	orc_mcg_ready = 1;
}

void orc_reset_internal_seed(void){
	orc_internal_iseed[0] = 2398; orc_internal_iseed[1] = 691;
	orc_internal_iseed[2] = 782;  orc_internal_iseed[3] = 721;
}

void orc_uniform_real_vector(size_t n, double *x, int *iseed){
	if(0 == n){ return; }
	orc_mcg_init();
	int *s = iseed ? iseed : orc_internal_iseed;
	unsigned long long v = 0;
	for(size_t j = 0; j < n; ++j){
		unsigned long long seed =
			((unsigned long long)(s[0] & 4095) << 36) | ((unsigned long long)(s[1] & 4095) << 24) |
			((unsigned long long)(s[2] & 4095) << 12) | (unsigned long long)((s[3] | 1) & 4095);
		const size_t i = j % 128;
		v = (unsigned long long)(((unsigned __int128)seed * orc_mcg_pow[i]) & ORC_MASK48);
		x[j] = 1.0 - (double)v * 0x1p-48;
		if(i+1 == 128){
			s[0] = (int)((v >> 36) & 4095); s[1] = (int)((v >> 24) & 4095);
			s[2] = (int)((v >> 12) & 4095); s[3] = (int)(v & 4095);
		}
	}
	s[0] = (int)((v >> 36) & 4095); s[1] = (int)((v >> 24) & 4095);
	s[2] = (int)((v >> 12) & 4095); s[3] = (int)(v & 4095);
}

/* 2*pi as written at RNP/Types.hpp:93. */
static const double orc_twopi = 6.28318530717958647692528676656;

/* Random::GenerateVector for real T, RNP/Random.hpp:262-320 with
 * Utility<T> (lines 185-213). */
void orc_d_generate_vector(int dist, size_t n, double *x, int *iseed){
	double pool[128];
	switch(dist){
	case ORC_UNIFORM01:
		orc_uniform_real_vector(n, x, iseed);
		break;
	case ORC_UNIFORM_11:
		orc_uniform_real_vector(n, x, iseed);
		for(size_t i = 0; i < n; ++i){ x[i] = 2.0*x[i] - 1.0; }
		break;
	case ORC_NORMAL01:
		for(size_t i0 = 0; i0 < n; i0 += 64){
			size_t ilim = 64; if(n < i0+ilim){ ilim = n-i0; }
			orc_uniform_real_vector(2*ilim, pool, iseed);
			for(size_t i = 0; i < ilim; ++i){
				double a = pool[2*i+0], b = pool[2*i+1];
				x[i0+i] = sqrt(-2.0 * log1p(-a)) * cos(orc_twopi * b);
			}
		}
		break;
	case ORC_UNITDISC:
		for(size_t i0 = 0; i0 < n; i0 += 64){
			size_t ilim = 64; if(n < i0+ilim){ ilim = n-i0; }
			orc_uniform_real_vector(128, pool, iseed);
			for(size_t i = 0; i < ilim; ++i){ x[i0+i] = 2.0*pool[2*i+0] - 1.0; }
		}
		break;
	case ORC_UNITCIRCLE:
		for(size_t i0 = 0; i0 < n; i0 += 64){
			size_t ilim = 64; if(n < i0+ilim){ ilim = n-i0; }
			orc_uniform_real_vector(128, pool, iseed);
			for(size_t i = 0; i < ilim; ++i){ x[i0+i] = (2.0*pool[2*i+0] < 1.0) ? -1.0 : 1.0; }
		}
		break;
	default: break;
	}
}

/* Random::GenerateVector for complex T (Utility<std::complex<T>>,
 * RNP/Random.hpp:198-213): Uniform01 draws 2n reals into the interleaved
 * storage; Normal01 is sqrt(-2 log1p(-a)) * (cos 2 pi b, sin 2 pi b). */
void orc_z_generate_vector(int dist, size_t n, orc_z *x, int *iseed){
	double pool[128];
	double *xr = (double*)x;
	switch(dist){
	case ORC_UNIFORM01:
		orc_uniform_real_vector(2*n, xr, iseed);
		break;
	case ORC_UNIFORM_11:
		orc_uniform_real_vector(2*n, xr, iseed);
		for(size_t i = 0; i < n; ++i){ xr[2*i] = 2.0*xr[2*i] - 1.0; xr[2*i+1] = 2.0*xr[2*i+1] - 1.0; }
		break;
	case ORC_NORMAL01:
		for(size_t i0 = 0; i0 < n; i0 += 64){
			size_t ilim = 64; if(n < i0+ilim){ ilim = n-i0; }
			orc_uniform_real_vector(2*ilim, pool, iseed);
			for(size_t i = 0; i < ilim; ++i){
				double a = pool[2*i+0], b = pool[2*i+1];
				double r = sqrt(-2.0 * log1p(-a));
				double ang = orc_twopi * b;
				xr[2*(i0+i)] = r * cos(ang);
				xr[2*(i0+i)+1] = r * sin(ang);
			}
		}
		break;
	case ORC_UNITDISC:
		for(size_t i0 = 0; i0 < n; i0 += 64){
			size_t ilim = 64; if(n < i0+ilim){ ilim = n-i0; }
			orc_uniform_real_vector(128, pool, iseed);
			for(size_t i = 0; i < ilim; ++i){
				double r = sqrt(pool[2*i+0]);
				double ang = orc_twopi * pool[2*i+1];
				xr[2*(i0+i)] = r * cos(ang);
				xr[2*(i0+i)+1] = r * sin(ang);
			}
		}
		break;
	case ORC_UNITCIRCLE:
		for(size_t i0 = 0; i0 < n; i0 += 64){
			size_t ilim = 64; if(n < i0+ilim){ ilim = n-i0; }
			orc_uniform_real_vector(128, pool, iseed);
			for(size_t i = 0; i < ilim; ++i){
				double ang = orc_twopi * pool[2*i+0];
				xr[2*(i0+i)] = cos(ang);
				xr[2*(i0+i)+1] = sin(ang);
			}
		}
		break;
	default: break;
	}
}

/* MatrixGen::GenerateDiagonal, RNP/MatrixGen.hpp:31-109 (real T). */
void orc_generate_diagonal(int dist, int mode, int reverse, int randsign, double cond,
                           double dscale, size_t n, double *d, int *iseed){
	if(0 == n){ return; }
	const double rcond = 1.0/cond;
	switch(mode){
	case ORC_DIAG_FIRST1_REMAINING_COND:
		d[0] = 1.0;
		for(size_t i = 1; i < n; ++i){ d[i] = rcond; }
		break;
	case ORC_DIAG_LASTCOND_REMAINING1:
		for(size_t i = 0; i+1 < n; ++i){ d[i] = 1.0; }
		d[n-1] = rcond;
		break;
	case ORC_DIAG_GRADED_EXPONENTIALLY:
		for(size_t i = 0; i < n; ++i){
			double t = (double)i / (double)(n-1);
			d[i] = pow(rcond, t);
		}
		break;
	case ORC_DIAG_GRADED_LINEARLY:
		for(size_t i = 0; i < n; ++i){
			double t = (double)i / (double)(n-1);
			double t1 = 1.0 - t;
			d[i] = t1 + t*rcond;
		}
		break;
	case ORC_DIAG_RANDOM_EXPONENTIAL:
		orc_d_generate_vector(ORC_UNIFORM01, n, d, iseed);
		for(size_t i = 0; i < n; ++i){ d[i] = pow(rcond, d[i]); }
		break;
	case ORC_DIAG_RANDOM_DIST:
		orc_d_generate_vector(dist, n, d, iseed);
		break;
	default: break;
	}
	if(reverse){
		/* the reference swaps d[i] and d[n-1-i] for every i (lines 79-85),
		 * which swaps each pair twice and leaves d unchanged. */
		for(size_t i = 0; i < n; ++i){
			double t = d[i]; d[i] = d[n-1-i]; d[n-1-i] = t;
		}
	}
	if(randsign){
		for(size_t i = 0; i < n; ++i){
			double r;
			orc_uniform_real_vector(1, &r, iseed);
			if(2.0*r > 1.0){ d[i] = -d[i]; }
		}
	}
	if(ORC_DIAG_GIVEN != mode && ORC_DIAG_RANDOM_DIST != mode){
		double maxabs = fabs(d[0]);
		for(size_t i = 1; i < n; ++i){
			double c = fabs(d[i]);
			if(c > maxabs){ maxabs = c; }
		}
		const double s = dscale / maxabs;
		for(size_t i = 0; i < n; ++i){ d[i] *= s; }
	}
}

/* ---- real instantiation ---- */
#define T double
#define FN(name) orc_d_##name
#define IS_CPLX 0
#define RE(x) (x)
#define IM(x) (0.0)
#define CONJ(x) (x)
#define MKT(re, im) (re)
#define TABS(x) fabs(x)
#include "rnp_oracle_gen.h"
#undef T
#undef FN
#undef IS_CPLX
#undef RE
#undef IM
#undef CONJ
#undef MKT
#undef TABS

/* ---- complex instantiation ---- */
#define T orc_z
#define FN(name) orc_z_##name
#define IS_CPLX 1
#define RE(x) creal(x)
#define IM(x) cimag(x)
#define CONJ(x) conj(x)
#define MKT(re, im) CMPLX((re), (im))
#define TABS(x) cabs(x)
#include "rnp_oracle_gen.h"
#undef T
#undef FN
#undef IS_CPLX
#undef RE
#undef IM
#undef CONJ
#undef MKT
#undef TABS

/* The CPU-baseline sample: QR::Factor's blocked loop (QR.hpp:173-190) run
 * for the first `npanels` panels only, with block size nb.  work must hold
 * n*nb doubles. */
void orc_d_factor_panels(size_t m, size_t n, double *a, size_t lda, double *tau,
                         double *work, size_t nb, size_t npanels){
	const size_t k = m < n ? m : n;
	size_t p = 0;
	for(size_t i = 0; i < k && p < npanels; i += nb, ++p){
		const size_t ib = (k-i < nb ? k-i : nb);
		orc_d_factor_unblocked(m-i, ib, &a[i+i*lda], lda, &tau[i], work);
		if(i+ib < n){
			orc_d_generate_block_tr(m-i, ib, &a[i+i*lda], lda, &tau[i], work, n);
			orc_d_apply_block("L", "C", m-i, n-i-ib, ib, &a[i+i*lda], lda,
			                  work, n, &a[i+(i+ib)*lda], lda, &work[ib], n);
		}
	}
}
