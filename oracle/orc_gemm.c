/*
 * orc_gemm.c -- the oracle's xgemm (TEST INFRASTRUCTURE, not the product).
 *
 * RNP's MultMM forwards to the external Fortran dgemm_/zgemm_
 * (src/BLAS_mix.cpp:498-517, RNP/BLAS/FBLAS.hpp:384-390, 804).  This file
 * restates the BLAS xgemm contract, C = alpha*op(A)*op(B) + beta*C with
 * op in {N, T, C} and beta == 0 clearing C, as a cache-blocked loop nest
 * (packed panels, 8x4 register block) parallelised with OpenMP over column
 * blocks of C.  It is what makes the oracle's blocked QR fast enough to be
 * the timed CPU baseline.  Summation order differs from any particular
 * BLAS; results agree to rounding.
 */
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#include <omp.h>
#include "rnp_oracle.h"

#define MC 128
#define KC 256
#define NCB 64

static int orc_nthreads = 0;

void orc_set_threads(int nthreads){ orc_nthreads = nthreads; }
int orc_get_threads(void){ return orc_nthreads > 0 ? orc_nthreads : omp_get_max_threads(); }

static inline double opd(char t, const double *a, size_t lda, size_t r, size_t c){
	return ('N' == t) ? a[r + c*lda] : a[c + r*lda];
}

void orc_d_gemm(const char *transa, const char *transb, size_t m, size_t n, size_t k,
                double alpha, const double *a, size_t lda, const double *b, size_t ldb,
                double beta, double *c, size_t ldc){
	if(0 == m || 0 == n){ return; }
	if(1.0 != beta){
		for(size_t j = 0; j < n; ++j){
			for(size_t i = 0; i < m; ++i){
				c[i+j*ldc] = (0.0 == beta) ? 0.0 : beta*c[i+j*ldc];
			}
		}
	}
	if(0.0 == alpha || 0 == k){ return; }
	const char ta = transa[0], tb = transb[0];
	const size_t nblk = (n + NCB - 1) / NCB;
	const double work = (double)m * (double)n * (double)k;
	const int nth = (work > 2.0e6) ? orc_get_threads() : 1;
	#pragma omp parallel for schedule(dynamic) num_threads(nth)
	for(size_t jb = 0; jb < nblk; ++jb){
		const size_t j0 = jb*NCB;
		const size_t ncb = (n - j0 < NCB) ? n - j0 : NCB;
		double *Ap = (double*)aligned_alloc(64, sizeof(double)*MC*KC);
		double *Bp = (double*)aligned_alloc(64, sizeof(double)*KC*NCB);
		for(size_t k0 = 0; k0 < k; k0 += KC){
			const size_t kc = (k - k0 < KC) ? k - k0 : KC;
			for(size_t j = 0; j < NCB; ++j){
				for(size_t l = 0; l < kc; ++l){
					Bp[l + j*KC] = (j < ncb) ? alpha*opd(tb, b, ldb, k0+l, j0+j) : 0.0;
				}
			}
			for(size_t i0 = 0; i0 < m; i0 += MC){
				const size_t mc = (m - i0 < MC) ? m - i0 : MC;
				const size_t mcp = (mc + 7) & ~(size_t)7;
				for(size_t l = 0; l < kc; ++l){
					for(size_t i = 0; i < mcp; ++i){
						Ap[i + l*MC] = (i < mc) ? opd(ta, a, lda, i0+i, k0+l) : 0.0;
					}
				}
				for(size_t jj = 0; jj < ncb; jj += 4){
					for(size_t ii = 0; ii < mc; ii += 8){
						double acc[4][8];
						memset(acc, 0, sizeof(acc));
						for(size_t l = 0; l < kc; ++l){
							const double *ap = Ap + ii + l*MC;
							const double *bp = Bp + l + jj*KC;
							const double b0 = bp[0], b1 = bp[KC], b2 = bp[2*KC], b3 = bp[3*KC];
							for(int r = 0; r < 8; ++r){
								const double av = ap[r];
								acc[0][r] += av*b0;
								acc[1][r] += av*b1;
								acc[2][r] += av*b2;
								acc[3][r] += av*b3;
							}
						}
						for(int q = 0; q < 4; ++q){
							if(jj+q >= ncb){ break; }
							double *cc = c + i0 + ii + (j0+jj+q)*ldc;
							for(int r = 0; r < 8; ++r){
								if(ii + r < mc){ cc[r] += acc[q][r]; }
							}
						}
					}
				}
			}
		}
		free(Ap);
		free(Bp);
	}
}

static inline orc_z opz(char t, const orc_z *a, size_t lda, size_t r, size_t c){
	if('N' == t){ return a[r + c*lda]; }
	if('T' == t){ return a[c + r*lda]; }
	return conj(a[c + r*lda]);
}

void orc_z_gemm(const char *transa, const char *transb, size_t m, size_t n, size_t k,
                orc_z alpha, const orc_z *a, size_t lda, const orc_z *b, size_t ldb,
                orc_z beta, orc_z *c, size_t ldc){
	if(0 == m || 0 == n){ return; }
	if(1.0 != beta){
		for(size_t j = 0; j < n; ++j){
			for(size_t i = 0; i < m; ++i){
				c[i+j*ldc] = (0.0 == beta) ? 0.0 : beta*c[i+j*ldc];
			}
		}
	}
	if(0.0 == alpha || 0 == k){ return; }
	const char ta = transa[0], tb = transb[0];
	const size_t nblk = (n + NCB - 1) / NCB;
	const double work = 4.0 * (double)m * (double)n * (double)k;
	const int nth = (work > 2.0e6) ? orc_get_threads() : 1;
	#pragma omp parallel for schedule(dynamic) num_threads(nth)
	for(size_t jb = 0; jb < nblk; ++jb){
		const size_t j0 = jb*NCB;
		const size_t ncb = (n - j0 < NCB) ? n - j0 : NCB;
		double *Ar = (double*)aligned_alloc(64, sizeof(double)*MC*KC);
		double *Ai = (double*)aligned_alloc(64, sizeof(double)*MC*KC);
		double *Br = (double*)aligned_alloc(64, sizeof(double)*KC*NCB);
		double *Bi = (double*)aligned_alloc(64, sizeof(double)*KC*NCB);
		for(size_t k0 = 0; k0 < k; k0 += KC){
			const size_t kc = (k - k0 < KC) ? k - k0 : KC;
			for(size_t j = 0; j < NCB; ++j){
				for(size_t l = 0; l < kc; ++l){
					orc_z v = (j < ncb) ? alpha*opz(tb, b, ldb, k0+l, j0+j) : 0.0;
					Br[l + j*KC] = creal(v);
					Bi[l + j*KC] = cimag(v);
				}
			}
			for(size_t i0 = 0; i0 < m; i0 += MC){
				const size_t mc = (m - i0 < MC) ? m - i0 : MC;
				const size_t mcp = (mc + 7) & ~(size_t)7;
				for(size_t l = 0; l < kc; ++l){
					for(size_t i = 0; i < mcp; ++i){
						orc_z v = (i < mc) ? opz(ta, a, lda, i0+i, k0+l) : 0.0;
						Ar[i + l*MC] = creal(v);
						Ai[i + l*MC] = cimag(v);
					}
				}
				for(size_t jj = 0; jj < ncb; jj += 4){
					for(size_t ii = 0; ii < mc; ii += 8){
						double accr[4][8], acci[4][8];
						memset(accr, 0, sizeof(accr));
						memset(acci, 0, sizeof(acci));
						for(size_t l = 0; l < kc; ++l){
							const double *apr = Ar + ii + l*MC;
							const double *api = Ai + ii + l*MC;
							for(int q = 0; q < 4; ++q){
								const double br = Br[l + (jj+q)*KC], bi = Bi[l + (jj+q)*KC];
								for(int r = 0; r < 8; ++r){
									accr[q][r] += apr[r]*br - api[r]*bi;
									acci[q][r] += apr[r]*bi + api[r]*br;
								}
							}
						}
						for(int q = 0; q < 4; ++q){
							if(jj+q >= ncb){ break; }
							orc_z *cc = c + i0 + ii + (j0+jj+q)*ldc;
							for(int r = 0; r < 8; ++r){
								if(ii + r < mc){ cc[r] += CMPLX(accr[q][r], acci[q][r]); }
							}
						}
					}
				}
			}
		}
		free(Ar); free(Ai); free(Br); free(Bi);
	}
}
