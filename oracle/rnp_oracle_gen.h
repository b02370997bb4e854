/*
 * rnp_oracle_gen.h -- type-generic body of the oracle (test infrastructure).
 * Included twice by rnp_oracle.c: once with T = double, once with
 * T = double complex.  Every function cites the reference file:line it
 * restates.  Macros supplied by the includer:
 *   T, FN(name), IS_CPLX, RE(x), IM(x), CONJ(x), MKT(re,im), TABS(x)
 */

/* BLAS::Norm2, RNP/BLAS/BLAS_mix.hpp:725-758: one-pass scaled sum of squares,
 * real and imaginary parts fed separately, in index order. */
double FN(norm2)(size_t n, const T *x, size_t incx){
	double ssq = 1.0, scale = 0.0;
	while(n-- > 0){
		double parts[2];
		int np = 1;
		parts[0] = RE(*x);
#if IS_CPLX
		parts[1] = IM(*x);
		np = 2;
#endif
		for(int p = 0; p < np; ++p){
			if(0.0 != parts[p]){
				double temp = fabs(parts[p]);
				if(scale < temp){
					double r = scale/temp;
					ssq = ssq*r*r + 1.0;
					scale = temp;
				}else{
					double r = temp/scale;
					ssq += r*r;
				}
			}
		}
		x += incx;
	}
	return scale*sqrt(ssq);
}

/* BLAS::Scale, RNP/BLAS/BLAS_mix.hpp:580-586 (scalar of type T). */
static void FN(scal)(size_t n, T alpha, T *x, size_t incx){
	while(n-- > 0){ *x *= alpha; x += incx; }
}
/* BLAS::Scale with a real scale factor (TS = real_type). */
static void FN(scal_real)(size_t n, double alpha, T *x, size_t incx){
	while(n-- > 0){ *x *= alpha; x += incx; }
}

/* Reflector::LastNonzeroColumnLimit, Reflector.hpp:36-48.  For m == 0 the
 * reference reads a[m-1+...] with m-1 wrapped around; here that case
 * returns 0 (an empty row range has no nonzero column). */
size_t FN(last_nonzero_column_limit)(size_t m, size_t n, const T *a, size_t lda){
	if(0 == n){ return 0; }
	if(0 == m){ return 0; }
	if(0.0 != a[0+(n-1)*lda] || 0.0 != a[m-1+(n-1)*lda]){ return n; }
	size_t j = n;
	while(j-- > 0){
		for(size_t i = 0; i < m; ++i){
			if(0.0 != a[i+j*lda]){ return j+1; }
		}
	}
	return 0;
}

/* Reflector::LastNonzeroRowLimit, Reflector.hpp:63-75 (n == 0 returns 0). */
size_t FN(last_nonzero_row_limit)(size_t m, size_t n, const T *a, size_t lda){
	if(0 == m){ return 0; }
	if(0 == n){ return 0; }
	if(0.0 != a[m-1+0*lda] || 0.0 != a[m-1+(n-1)*lda]){ return m; }
	size_t i = m;
	while(i-- > 0){
		for(size_t j = 0; j < n; ++j){
			if(0.0 != a[i+j*lda]){ return i+1; }
		}
	}
	return 0;
}

/* ---------------------------------------------------------------------
 * Level-2 BLAS restatement (what src/BLAS_mix.cpp forwards to Fortran).
 * ------------------------------------------------------------------- */

/* MultMV -> xgemv (src/BLAS_mix.cpp:104-123): y = alpha*op(A)*x + beta*y. */
static void FN(gemv)(const char *trans, size_t m, size_t n, T alpha, const T *a, size_t lda,
                     const T *x, size_t incx, T beta, T *y, size_t incy){
	const int notr = ('N' == trans[0]);
	const int conjt = ('C' == trans[0]);
	const size_t leny = notr ? m : n;
	const size_t lenx = notr ? n : m;
	if(0 == m || 0 == n){ /* y = beta*y still applies when the other dim is 0 */ }
	if(1.0 != beta){
		for(size_t i = 0; i < leny; ++i){
			y[i*incy] = (0.0 == beta) ? (T)0 : beta*y[i*incy];
		}
	}
	if(0.0 == alpha || 0 == lenx){ return; }
	if(notr){
		for(size_t j = 0; j < n; ++j){
			T temp = alpha*x[j*incx];
			for(size_t i = 0; i < m; ++i){ y[i*incy] += temp*a[i+j*lda]; }
		}
	}else{
		for(size_t j = 0; j < n; ++j){
			T temp = 0;
			if(conjt){
				for(size_t i = 0; i < m; ++i){ temp += CONJ(a[i+j*lda])*x[i*incx]; }
			}else{
				for(size_t i = 0; i < m; ++i){ temp += a[i+j*lda]*x[i*incx]; }
			}
			y[j*incy] += alpha*temp;
		}
	}
}

/* ConjugateRank1Update -> xger / xgerc (src/BLAS_mix.cpp:366-385):
 * A += alpha * x * y^H. */
static void FN(gerc)(size_t m, size_t n, T alpha, const T *x, size_t incx,
                     const T *y, size_t incy, T *a, size_t lda){
	if(0 == m || 0 == n || 0.0 == alpha){ return; }
	for(size_t j = 0; j < n; ++j){
		T yj = y[j*incy];
		if(0.0 != yj){
			T temp = alpha*CONJ(yj);
			for(size_t i = 0; i < m; ++i){ a[i+j*lda] += x[i*incx]*temp; }
		}
	}
}

/* MultTrV -> xtrmv (src/BLAS_mix.cpp:242-257), uplo "U", trans "N",
 * diag "N" -- the only form GenerateBlockTr uses (Reflector.hpp:618). */
static void FN(trmv_unn)(size_t n, const T *a, size_t lda, T *x){
	for(size_t j = 0; j < n; ++j){
		if(0.0 != x[j]){
			T temp = x[j];
			for(size_t i = 0; i < j; ++i){ x[i] += temp*a[i+j*lda]; }
			x[j] *= a[j+j*lda];
		}
	}
}

/* MultTrM -> xtrmm (src/BLAS_mix.cpp:653-672), side "R" only:
 * B = alpha * B * op(A), A k-by-k triangular, B m-by-k. */
static void FN(trmm_r)(const char *uplo, const char *transa, const char *diag,
                       size_t m, size_t k, T alpha, const T *a, size_t lda, T *b, size_t ldb){
	if(0 == m || 0 == k){ return; }
	const int upper = ('U' == uplo[0]);
	const int nounit = ('N' == diag[0]);
	const char tr = transa[0];
	/* Work column by column into a scratch copy of one row block. */
	T *tmp = (T*)malloc(sizeof(T)*k);
	for(size_t i = 0; i < m; ++i){
		for(size_t j = 0; j < k; ++j){ tmp[j] = b[i+j*ldb]; }
		for(size_t j = 0; j < k; ++j){
			/* (B*op(A))(i,j) = sum_l B(i,l) op(A)(l,j) */
			T s = 0;
			for(size_t l = 0; l < k; ++l){
				T opa;
				size_t r, c;
				if('N' == tr){ r = l; c = j; }else{ r = j; c = l; }
				int inside = upper ? (r <= c) : (r >= c);
				if(!inside){ continue; }
				if(r == c && !nounit){ opa = 1; }
				else{ opa = a[r+c*lda]; if('C' == tr){ opa = CONJ(opa); } }
				s += tmp[l]*opa;
			}
			b[i+j*ldb] = alpha*s;
		}
	}
	free(tmp);
}

/* SolveTrM -> xtrsm (src/BLAS_mix.cpp:674-693), "L","U","N","N":
 * B = inv(A) * B with A upper triangular, non-unit. */
static void FN(trsm_lunn)(size_t m, size_t n, const T *a, size_t lda, T *b, size_t ldb){
	for(size_t j = 0; j < n; ++j){
		size_t k = m;
		while(k-- > 0){
			if(0.0 != b[k+j*ldb]){
				b[k+j*ldb] /= a[k+k*lda];
				T bk = b[k+j*ldb];
				for(size_t i = 0; i < k; ++i){ b[i+j*ldb] -= bk*a[i+k*lda]; }
			}
		}
	}
}

/* ---------------------------------------------------------------------
 * Reflectors
 * ------------------------------------------------------------------- */

/* Reflector::Generate, Reflector.hpp:117-158.
 * Sign rule (line 131-132): beta = +hypot3(...) unless real(alpha) > 0, so
 * beta > 0 when real(alpha) == 0 (LAPACK's xlarfg makes it negative there).
 * safmin = DBL_MIN/(2*DBL_EPSILON) = 2^-971 (line 133).
 * Rescale loop (lines 139-144): its condition calls the C library's
 * int abs(int) on the double beta; the reading that terminates (and that an
 * optimizing build of the reference executes) is "truncate beta to an int
 * and test it against 0", i.e. the loop repeats while |beta| < 1.  For any
 * finite nonzero beta below 2^-971 that is exactly two scalings by 2^971. */
void FN(generate)(size_t n, T *alpha, T *x, size_t incx, T *tau){
	if(0 == n){ *tau = 0; return; }
	double xnorm = FN(norm2)(n-1, x, incx);
	if(0.0 == xnorm && 0.0 == IM(*alpha)){
		*tau = 0;
		return;
	}
	double beta = orc_hypot3(RE(*alpha), IM(*alpha), xnorm);
	if(RE(*alpha) > 0){ beta = -beta; }
	const double safmin = DBL_MIN / (2*DBL_EPSILON);
	const double rsafmn = 1.0 / safmin;
	size_t knt = 0;
	if(fabs(beta) < safmin){
		do{
			++knt;
			FN(scal_real)(n-1, rsafmn, x, incx);
			beta *= rsafmn;
			*alpha *= rsafmn;
		}while(fabs(beta) < 1.0);
		xnorm = FN(norm2)(n-1, x, incx);
		beta = orc_hypot3(RE(*alpha), IM(*alpha), xnorm);
		if(RE(*alpha) > 0){ beta = -beta; }
	}
	/* tau = (beta - alpha)/beta : complex over real, componentwise */
	*tau = MKT((beta - RE(*alpha))/beta, (-IM(*alpha))/beta);
	/* alpha = 1/(alpha - beta) : real(1) over T (std::complex division) */
	{
		T den = MKT(RE(*alpha) - beta, IM(*alpha));
		T one = MKT(1.0, 0.0);
		*alpha = one / den;
	}
	FN(scal)(n-1, *alpha, x, incx);
	while(knt-- > 0){ beta *= safmin; }
	*alpha = MKT(beta, 0.0);
}

/* Reflector::Apply, Reflector.hpp:191-360, for vconj == false and
 * vone in {0, 1} -- the branches QR::Factor_unblocked (QR.hpp:95-98) and
 * QR::MultQ_unblocked (QR.hpp:275, 295) reach.  H = I - tau v v^H. */
void FN(apply)(const char *side, int vone, size_t m, size_t n, const T *v, size_t incv,
               T tau, T *c, size_t ldc, T *work){
	size_t lenv = 0, lenc = 0;
	const int left = ('L' == side[0]);
	if(0.0 != tau){
		lenv = left ? m : n;
		if(0 == lenv){ return; }
		size_t i = (lenv-1)*incv;
		while(lenv > 0 && 0.0 == v[i]){
			--lenv;
			i -= incv;
		}
		if(left){
			lenc = FN(last_nonzero_column_limit)(lenv, n, c, ldc);
		}else{
			lenc = FN(last_nonzero_row_limit)(m, lenv, c, ldc);
		}
		if(0 == lenc){ return; }
	}
	const T one = 1;
	if(0 == vone){
		if(left){
			FN(gemv)("C", lenv, lenc, one, c, ldc, v, incv, 0, work, 1);
			FN(gerc)(lenv, lenc, -tau, v, incv, work, 1, c, ldc);
		}else{
			FN(gemv)("N", lenc, lenv, one, c, ldc, v, incv, 0, work, 1);
			FN(gerc)(lenc, lenv, -tau, work, 1, v, incv, c, ldc);
		}
	}else{
		if(left){
			if(lenv > 1){
				for(size_t j = 0; j < lenc; ++j){ work[j] = CONJ(c[0+j*ldc]); }
				FN(gemv)("C", lenv-1, lenc, one, &c[1], ldc, &v[incv], incv, one, work, 1);
			}else{
				for(size_t j = 0; j < lenc; ++j){ work[j] = CONJ(c[0+j*ldc]); }
			}
			FN(gerc)(1, lenc, -tau, &one, incv, work, 1, c, ldc);
			if(lenv > 1){
				FN(gerc)(lenv-1, lenc, -tau, &v[incv], incv, work, 1, &c[1], ldc);
			}
		}else{
			for(size_t i = 0; i < lenc; ++i){ work[i] = c[i]; }
			if(lenv > 1){
				FN(gemv)("N", lenc, lenv-1, one, &c[0+1*ldc], ldc, &v[incv], incv, one, work, 1);
			}
			/* Axpy(lenc, -tau, work, 1, c, 1) */
			for(size_t i = 0; i < lenc; ++i){ c[i] += -tau*work[i]; }
			if(lenv > 1){
				FN(gerc)(lenc, lenv-1, -tau, work, 1, &v[incv], incv, &c[0+1*ldc], ldc);
			}
		}
	}
}

/* Reflector::GenerateBlockTr, Reflector.hpp:573-626, dir "F", storev "C":
 * T upper k-by-k with H[0] H[1] ... H[k-1] = I - V T V^H. */
void FN(generate_block_tr)(size_t n, size_t k, const T *v, size_t ldv, const T *tau,
                           T *t, size_t ldt){
	if(0 == n){ return; }
	size_t prevlastv = n-1;
	for(size_t i = 0; i < k; ++i){
		if(i > prevlastv){ prevlastv = i; }
		if(0.0 == tau[i]){
			for(size_t j = 0; j <= i; ++j){ t[j+i*ldt] = 0; }
		}else{
			size_t lastv;
			for(lastv = n-1; lastv > i; --lastv){
				if(0.0 != v[lastv+i*ldv]){ break; }
			}
			for(size_t j = 0; j < i; ++j){
				t[j+i*ldt] = -tau[i] * CONJ(v[i+j*ldv]);
			}
			size_t jlim = (lastv < prevlastv ? lastv : prevlastv);
			FN(gemv)("C", jlim-i, i, -tau[i], &v[i+1+0*ldv], ldv,
			         &v[i+1+i*ldv], 1, 1, &t[0+i*ldt], 1);
			FN(trmv_unn)(i, t, ldt, &t[0+i*ldt]);
			t[i+i*ldt] = tau[i];
			if(i > 0){
				if(lastv > prevlastv){ prevlastv = lastv; }
			}else{
				prevlastv = lastv;
			}
		}
	}
}

/* Reflector::ApplyBlock, Reflector.hpp:725-815, dir "F", storev "C":
 * side "L": C = H C or H^H C;  side "R": C = C H or C H^H,
 * H = I - V T V^H, work is ldwork-by-k. */
void FN(apply_block)(const char *side, const char *trans, size_t m, size_t n, size_t k,
                     const T *v, size_t ldv, const T *t, size_t ldt,
                     T *c, size_t ldc, T *work, size_t ldwork){
	if(0 == m || 0 == n){ return; }
	const char *transt = ('N' == trans[0] ? "C" : "N");
	const T one = 1, mone = -1;
	if('L' == side[0]){
		size_t lastv = FN(last_nonzero_row_limit)(m, k, v, ldv);
		if(k > lastv){ lastv = k; }
		const size_t lastc = FN(last_nonzero_column_limit)(lastv, n, c, ldc);
		for(size_t j = 0; j < k; ++j){
			for(size_t i = 0; i < lastc; ++i){ work[i+j*ldwork] = CONJ(c[j+i*ldc]); }
		}
		FN(trmm_r)("L", "N", "U", lastc, k, one, v, ldv, work, ldwork);
		if(lastv > k){
			FN(gemm)("C", "N", lastc, k, lastv-k, one, &c[k], ldc, &v[k], ldv, one, work, ldwork);
		}
		FN(trmm_r)("U", transt, "N", lastc, k, one, t, ldt, work, ldwork);
		if(m > k){
			FN(gemm)("N", "C", lastv-k, lastc, k, mone, &v[k], ldv, work, ldwork, one, &c[k], ldc);
		}
		FN(trmm_r)("L", "C", "U", lastc, k, one, v, ldv, work, ldwork);
		for(size_t j = 0; j < k; ++j){
			for(size_t i = 0; i < lastc; ++i){ c[j+i*ldc] -= CONJ(work[i+j*ldwork]); }
		}
	}else{
		size_t lastv = FN(last_nonzero_row_limit)(n, k, v, ldv);
		if(k > lastv){ lastv = k; }
		const size_t lastc = FN(last_nonzero_row_limit)(m, lastv, c, ldc);
		for(size_t j = 0; j < k; ++j){
			for(size_t i = 0; i < lastc; ++i){ work[i+j*ldwork] = c[i+j*ldc]; }
		}
		FN(trmm_r)("L", "N", "U", lastc, k, one, v, ldv, work, ldwork);
		if(lastv > k){
			FN(gemm)("N", "N", lastc, k, lastv-k, one, &c[k*ldc], ldc, &v[k], ldv, one, work, ldwork);
		}
		FN(trmm_r)("U", trans, "N", lastc, k, one, t, ldt, work, ldwork);
		if(lastv > k){
			FN(gemm)("N", "C", lastc, lastv-k, k, mone, work, ldwork, &v[k], ldv, one, &c[k*ldc], ldc);
		}
		FN(trmm_r)("L", "C", "U", lastc, k, one, v, ldv, work, ldwork);
		for(size_t j = 0; j < k; ++j){
			for(size_t i = 0; i < lastc; ++i){ c[i+j*ldc] -= work[i+j*ldwork]; }
		}
	}
}

/* ---------------------------------------------------------------------
 * QR
 * ------------------------------------------------------------------- */

/* QR::Factor_unblocked, QR.hpp:78-102. */
void FN(factor_unblocked)(size_t m, size_t n, T *a, size_t lda, T *tau, T *work){
	size_t k = m < n ? m : n;
	for(size_t i = 0; i < k; ++i){
		size_t row = m-1; if(i+1 < row){ row = i+1; }
		FN(generate)(m-i, &a[i+i*lda], &a[row+i*lda], 1, &tau[i]);
		if(i < n-1){
			T alpha = a[i+i*lda];
			a[i+i*lda] = 1;
			FN(apply)("L", 0, m-i, n-i-1, &a[i+i*lda], 1, CONJ(tau[i]), &a[i+(i+1)*lda], lda, work);
			a[i+i*lda] = alpha;
		}
	}
}

/* QR::Factor, QR.hpp:129-197 (workspace query and short-lwork degrade
 * included; the Tuning hook is the optional `tune` argument). */
void FN(factor)(size_t m, size_t n, T *a, size_t lda, T *tau, size_t *lwork, T *work,
                const orc_tuning *tune){
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(NULL == lwork){ orc_assert_fail("NULL != lwork", __LINE__); return; }
	if(!(0 == *lwork || *lwork >= m)){ orc_assert_fail("0 == *lwork || *lwork >= m", __LINE__); }
	const size_t k = (m < n ? m : n);
	if(0 == k){ return; }
	size_t nb = tune ? tune->nb : 32;
	if(NULL == work || 0 == *lwork){
		*lwork = nb*n;
		return;
	}
	size_t nbmin = 2, nx = 0, iws = n, ldwork = n;
	if(nb > 1 && nb < k){
		nx = tune ? tune->nx : 128;
		if(nx < k){
			iws = ldwork*nb;
			if(*lwork < iws){
				nb = *lwork / ldwork;
				nbmin = tune ? tune->nbmin : 2;
				if(2 > nbmin){ nbmin = 2; }
			}
		}
	}
	size_t i;
	if(nb >= nbmin && nb < k && nx < k){
		for(i = 0; i+nx < k; i += nb){
			const size_t ib = (k-i < nb ? k-i : nb);
			FN(factor_unblocked)(m-i, ib, &a[i+i*lda], lda, &tau[i], work);
			if(i+ib < n){
				FN(generate_block_tr)(m-i, ib, &a[i+i*lda], lda, &tau[i], work, ldwork);
				FN(apply_block)("L", "C", m-i, n-i-ib, ib, &a[i+i*lda], lda,
				                work, ldwork, &a[i+(i+ib)*lda], lda, &work[ib], ldwork);
			}
		}
	}else{
		i = 0;
	}
	if(i < k){
		FN(factor_unblocked)(m-i, n-i, &a[i+i*lda], lda, &tau[i], work);
	}
}

/* QR::MultQ_unblocked, QR.hpp:234-298. */
void FN(multq_unblocked)(const char *side, const char *trans, size_t m, size_t n, size_t k,
                         const T *a, size_t lda, const T *tau, T *c, size_t ldc, T *work){
	if(!(ldc >= m)){ orc_assert_fail("ldc >= m", __LINE__); }
	const int left = ('L' == side[0]);
	const int notran = ('N' == trans[0]);
	if(!((left && k <= m) || (!left && k <= n))){ orc_assert_fail("k range", __LINE__); }
	if(!(lda >= k)){ orc_assert_fail("lda >= k", __LINE__); }
	if(m < 1 || n < 1 || k < 1){ return; }
	size_t mi = m, ni = n, ic = 0, jc = 0;
	const int forward = ((left && !notran) || (!left && notran));
	for(size_t s = 0; s < k; ++s){
		size_t i = forward ? s : k-1-s;
		if(left){ mi = m-i; ic = i; }else{ ni = n-i; jc = i; }
		T taui = notran ? tau[i] : CONJ(tau[i]);
		FN(apply)(side, 1, mi, ni, &a[i+i*lda], 1, taui, &c[ic+jc*ldc], ldc, work);
	}
}

/* QR::MultQ, QR.hpp:337-429 (T recomputed for every block, as there). */
void FN(multq)(const char *side, const char *trans, size_t m, size_t n, size_t k,
               const T *a, size_t lda, const T *tau, T *c, size_t ldc,
               size_t *lwork, T *work, const orc_tuning *tune){
	if(!(ldc >= m)){ orc_assert_fail("ldc >= m", __LINE__); }
	if(0 == m || 0 == n || 0 == k){ return; }
	const int left = ('L' == side[0]);
	const int notran = ('N' == trans[0]);
	if(!((left && k <= m) || (!left && k <= n))){ orc_assert_fail("k range", __LINE__); }
	if(!(lda >= k)){ orc_assert_fail("lda >= k", __LINE__); }
	const size_t nq = left ? m : n;
	const size_t nw = left ? n : m;
	size_t nb = tune ? tune->nb : 32;
	if(0 == *lwork || NULL == work){
		*lwork = nb*nw + nb*nb;
		return;
	}
	size_t nbmin = 2, ldwork = nw;
	T *t = NULL;
	size_t ldt = nb;
	if(nb > 1 && nb < k){
		size_t iws = nw*nb + nb*nb;
		t = work + nb*ldwork;
		if(*lwork < iws){
			nb = *lwork / (2*ldwork);
			t = work + nb*ldwork;
			ldt = nb;
			nbmin = tune ? tune->nbmin : 2;
		}
	}
	if(nb < nbmin || nb >= k){
		FN(multq_unblocked)(side, trans, m, n, k, a, lda, tau, c, ldc, work);
		return;
	}
	size_t mi = m, ni = n, ic = 0, jc = 0;
	if((left && !notran) || (!left && notran)){
		for(size_t i = 0; i < k; i += nb){
			size_t ib = (k < nb+i ? k-i : nb);
			FN(generate_block_tr)(nq-i, ib, &a[i+i*lda], lda, &tau[i], t, ldt);
			if(left){ mi = m-i; ic = i; }else{ ni = n-i; jc = i; }
			FN(apply_block)(side, trans, mi, ni, ib, &a[i+i*lda], lda, t, ldt,
			                &c[ic+jc*ldc], ldc, work, ldwork);
		}
	}else{
		size_t i = ((k-1)/nb)*nb + nb;
		do{
			i -= nb;
			size_t ib = (k < nb+i ? k-i : nb);
			FN(generate_block_tr)(nq-i, ib, &a[i+i*lda], lda, &tau[i], t, ldt);
			if(left){ mi = m-i; ic = i; }else{ ni = n-i; jc = i; }
			FN(apply_block)(side, trans, mi, ni, ib, &a[i+i*lda], lda, t, ldt,
			                &c[ic+jc*ldc], ldc, work, ldwork);
		}while(i > 0);
	}
}

/* QR::GenerateQ_unblocked, QR.hpp:451-485. */
void FN(generateq_unblocked)(size_t m, size_t n, size_t k, T *a, size_t lda, const T *tau,
                             T *work){
	if(!(m >= n)){ orc_assert_fail("m >= n", __LINE__); }
	if(!(k <= n)){ orc_assert_fail("k <= n", __LINE__); }
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(n < 1){ return; }
	for(size_t j = k; j < n; ++j){
		for(size_t i = 0; i < m; ++i){ a[i+j*lda] = 0; }
		a[j+j*lda] = 1;
	}
	size_t i = k;
	while(i-- > 0){
		if(i+1 < n){
			a[i+i*lda] = 1;
			FN(apply)("L", 0, m-i, n-i-1, &a[i+i*lda], 1, tau[i], &a[i+(i+1)*lda], lda, work);
		}
		if(i+1 < m){
			FN(scal)(m-i-1, -tau[i], &a[i+1+i*lda], 1);
		}
		a[i+i*lda] = 1.0 - tau[i];
		for(size_t l = 0; l < i; ++l){ a[l+i*lda] = 0; }
	}
}

/* QR::GenerateQ, QR.hpp:509-570. */
void FN(generateq)(size_t m, size_t n, size_t k, T *a, size_t lda, const T *tau,
                   size_t *lwork, T *work, const orc_tuning *tune){
	if(!(k <= n)){ orc_assert_fail("k <= n", __LINE__); }
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(0 == n){ return; }
	size_t nb = tune ? tune->nb : 32;
	if(0 == *lwork || NULL == work){
		*lwork = n*nb;
		return;
	}
	size_t nbmin = 2, nx = 0, ldwork = n;
	if(nb > 1 && nb < k){
		nx = tune ? tune->nx : 128;
		if(nx < k){
			size_t iws = ldwork*nb;
			if(*lwork < iws){
				nb = *lwork / ldwork;
				nbmin = tune ? tune->nbmin : 2;
			}
		}
	}
	size_t ki = 0, kk = 0;
	if(nb >= nbmin && nb < k && nx < k){
		ki = ((k-nx-1)/nb) * nb;
		kk = (k < ki+nb ? k : ki+nb);
		for(size_t j = kk; j < n; ++j){
			for(size_t i = 0; i < kk; ++i){ a[i+j*lda] = 0; }
		}
	}
	if(kk < n){
		FN(generateq_unblocked)(m-kk, n-kk, k-kk, &a[kk+kk*lda], lda, &tau[kk], work);
	}
	if(kk > 0){
		size_t i = ki + nb;
		do{
			i -= nb;
			const size_t ib = (nb+i < k ? nb : k-i);
			if(i+ib < n){
				FN(generate_block_tr)(m-i, ib, &a[i+i*lda], lda, &tau[i], work, ldwork);
				FN(apply_block)("L", "N", m-i, n-i-ib, ib, &a[i+i*lda], lda, work, ldwork,
				                &a[i+(i+ib)*lda], lda, &work[ib], ldwork);
			}
			FN(generateq_unblocked)(m-i, ib, ib, &a[i+i*lda], lda, &tau[i], work);
			for(size_t j = i; j < i+ib; ++j){
				for(size_t l = 0; l < i; ++l){ a[l+j*lda] = 0; }
			}
		}while(i > 0);
	}
}

/* ---------------------------------------------------------------------
 * LQ (RNP/LA/LQ.hpp): reflectors stored in rows, conjugated.
 * ------------------------------------------------------------------- */

/* BLAS::Conjugate, RNP/BLAS/BLAS_mix.hpp:233-239. */
static void FN(conjugate)(size_t n, T *x, size_t incx){
	for(size_t i = 0; i < n; ++i){ x[i*incx] = CONJ(x[i*incx]); }
}

/* Reflector::Apply, Reflector.hpp:238-291, for vone == 1 and vconj == true:
 * v = (1, conj(stored[1..])), H = I - tau v v^H.  LQ::MultQ_unblocked
 * (LQ.hpp:255, 268) reaches these branches.  Side "L" follows the
 * operations at Reflector.hpp:240-264: w = conj(C(0,:) + C(1:,:)^T v_s),
 * C(0,:) -= tau conj(w), C(i,j) -= tau conj(v_s[i]) conj(w_j).  The
 * reference's loop there writes c[i+0*ldc] for every j (Reflector.hpp:261),
 * i.e. only column 0; this restatement uses column j, the update the lines
 * compute (LQ.hpp:195-196 defines the result as Q C). */
static void FN(apply_vconj)(const char *side, size_t m, size_t n, const T *v, size_t incv, T tau, T *c,
                            size_t ldc, T *work){
	size_t lenv = 0, lenc = 0;
	const int left = ('L' == side[0]);
	if(0.0 != tau){
		lenv = left ? m : n;
		if(0 == lenv){ return; }
		size_t i = (lenv-1)*incv;
		while(lenv > 0 && 0.0 == v[i]){
			--lenv;
			i -= incv;
		}
		if(left){
			lenc = FN(last_nonzero_column_limit)(lenv, n, c, ldc);
		}else{
			lenc = FN(last_nonzero_row_limit)(m, lenv, c, ldc);
		}
		if(0 == lenc){ return; }
	}
	if(left){
		for(size_t j = 0; j < lenc; ++j){
			T w = 0;
			for(size_t i = 1; i < lenv; ++i){ w += c[i+j*ldc] * v[i*incv]; }
			w += c[0+j*ldc];
			work[j] = CONJ(w);
		}
		for(size_t j = 0; j < lenc; ++j){ c[0+j*ldc] += -tau * CONJ(work[j]); }
		for(size_t j = 0; j < lenc; ++j){
			for(size_t i = 1; i < lenv; ++i){
				c[i+j*ldc] -= tau * CONJ(v[i*incv]) * CONJ(work[j]);
			}
		}
	}else{
		for(size_t i = 0; i < lenc; ++i){ work[i] = c[i]; }
		for(size_t j = 1; j < lenv; ++j){
			const T cvj = CONJ(v[j*incv]);
			for(size_t i = 0; i < lenc; ++i){ work[i] += c[i+j*ldc] * cvj; }
		}
		for(size_t i = 0; i < lenc; ++i){ c[i] += -tau * work[i]; }
		/* Rank1Update (geru): C(:,1:) += -tau w v_s^T */
		for(size_t j = 1; j < lenv; ++j){
			for(size_t i = 0; i < lenc; ++i){ c[i+j*ldc] += -tau * work[i] * v[j*incv]; }
		}
	}
}

/* LQ::Factor_unblocked, LQ.hpp:78-102. */
void FN(lq_factor_unblocked)(size_t m, size_t n, T *a, size_t lda, T *tau, T *work){
	const size_t k = m < n ? m : n;
	for(size_t i = 0; i < k; ++i){
		FN(conjugate)(n-i, &a[i+i*lda], lda);
		T alpha = a[i+i*lda];
		T *x = (i+1 < n) ? &a[i+(i+1)*lda] : &a[i+i*lda];  /* n-i-1 == 0: not read */
		FN(generate)(n-i, &alpha, x, lda, &tau[i]);
		if(i+1 < m){
			a[i+i*lda] = 1;
			FN(apply)("R", 0, m-i-1, n-i, &a[i+i*lda], lda, tau[i], &a[i+1+i*lda], lda, work);
		}
		a[i+i*lda] = alpha;
		FN(conjugate)(n-i, &a[i+i*lda], lda);
	}
}

/* LQ::Factor, LQ.hpp:126-190: argument checks and the workspace query
 * (nb n, nb = LQ::Tuning default 32) as there.  The blocked loop
 * (LQ.hpp:163-183: Factor_unblocked on row blocks, GenerateBlockTr("F","R"),
 * ApplyBlock("R","N","F","R")) computes the same factors as the unblocked
 * loop in exact arithmetic; the restatement runs the unblocked loop for all
 * sizes, so its rounding is that of the reference's unblocked path. */
void FN(lq_factor)(size_t m, size_t n, T *a, size_t lda, T *tau, size_t *lwork, T *work){
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(NULL == lwork){ orc_assert_fail("NULL != lwork", __LINE__); return; }
	if(!(0 == *lwork || *lwork >= m)){ orc_assert_fail("0 == *lwork || *lwork >= m", __LINE__); }
	const size_t k = m < n ? m : n;
	if(0 == k){ return; }
	if(NULL == work || 0 == *lwork){
		*lwork = 32*n;
		return;
	}
	FN(lq_factor_unblocked)(m, n, a, lda, tau, work);
}

/* LQ::MultQ_unblocked, LQ.hpp:220-270: forward loop for (L,N) and (R,C),
 * backward otherwise; tau conjugated for "N". */
void FN(lq_multq_unblocked)(const char *side, const char *trans, size_t m, size_t n, size_t k,
                            const T *a, size_t lda, const T *tau, T *c, size_t ldc, T *work){
	if(!(ldc >= m)){ orc_assert_fail("ldc >= m", __LINE__); }
	if(0 == m || 0 == n || 0 == k){ return; }
	const int left = ('L' == side[0]);
	const int notran = ('N' == trans[0]);
	if(!((left && k <= m) || (!left && k <= n))){ orc_assert_fail("k range", __LINE__); }
	if(!(lda >= k)){ orc_assert_fail("lda >= k", __LINE__); }
	size_t mi = m, ni = n, ic = 0, jc = 0;
	const int forward = ((left && notran) || (!left && !notran));
	for(size_t s = 0; s < k; ++s){
		const size_t i = forward ? s : k-1-s;
		if(left){ mi = m-i; ic = i; }else{ ni = n-i; jc = i; }
		const T taui = notran ? CONJ(tau[i]) : tau[i];
		FN(apply_vconj)(side, mi, ni, &a[i+i*lda], lda, taui, &c[ic+jc*ldc], ldc, work);
	}
}

/* LQ::MultQ, LQ.hpp:309-405: checks and the workspace query (nb nw + nb^2);
 * the product itself by the unblocked loop (the blocked one applies the
 * same reflectors in the same order). */
void FN(lq_multq)(const char *side, const char *trans, size_t m, size_t n, size_t k,
                  const T *a, size_t lda, const T *tau, T *c, size_t ldc, size_t *lwork, T *work){
	if(!(ldc >= m)){ orc_assert_fail("ldc >= m", __LINE__); }
	if(0 == m || 0 == n || 0 == k){ return; }
	const int left = ('L' == side[0]);
	const size_t nw = left ? n : m;
	if(0 == *lwork || NULL == work){
		*lwork = 32*nw + 32*32;
		return;
	}
	FN(lq_multq_unblocked)(side, trans, m, n, k, a, lda, tau, c, ldc, work);
}

/* LQ::GenerateQ_unblocked, LQ.hpp:429-469. */
void FN(lq_generateq_unblocked)(size_t m, size_t n, size_t k, T *a, size_t lda, const T *tau, T *work){
	if(!(n >= m)){ orc_assert_fail("n >= m", __LINE__); }
	if(!(k <= m)){ orc_assert_fail("k <= m", __LINE__); }
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(0 == m){ return; }
	if(k < m){
		for(size_t j = 0; j < n; ++j){
			for(size_t i = k; i < m; ++i){ a[i+j*lda] = 0; }
			if(j >= k && j < m){ a[j+j*lda] = 1; }
		}
	}
	size_t i = k;
	while(i-- > 0){
		if(i+1 < n){
			FN(conjugate)(n-i-1, &a[i+(i+1)*lda], lda);
			if(i+1 < m){
				a[i+i*lda] = 1;
				FN(apply)("R", 0, m-i-1, n-i, &a[i+i*lda], lda, CONJ(tau[i]), &a[i+1+i*lda], lda, work);
			}
			for(size_t j = i+1; j < n; ++j){ a[i+j*lda] *= -tau[i]; }
			FN(conjugate)(n-i-1, &a[i+(i+1)*lda], lda);
		}
		a[i+i*lda] = 1.0 - CONJ(tau[i]);
		for(size_t l = 0; l < i; ++l){ a[i+l*lda] = 0; }
	}
}

/* LQ::GenerateQ, LQ.hpp:491-559: checks and the workspace query (n nb); Q by
 * the unblocked loop. */
void FN(lq_generateq)(size_t m, size_t n, size_t k, T *a, size_t lda, const T *tau, size_t *lwork, T *work){
	if(!(k <= m)){ orc_assert_fail("k <= m", __LINE__); }
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(0 == m){ return; }
	if(0 == *lwork || NULL == work){
		*lwork = n*32;
		return;
	}
	FN(lq_generateq_unblocked)(m, n, k, a, lda, tau, work);
}

/* ---------------------------------------------------------------------
 * QL (RNP/LA/QL.hpp) and RQ (RNP/LA/RQ.hpp): reflectors generated from the
 * last column (row) backwards, the implied 1 at the bottom (right) end of
 * each vector.
 * ------------------------------------------------------------------- */

/* Reflector::Apply with vone == -1 (Reflector.hpp:297-357): v's last entry
 * is an implied 1 (the stored value there is only read by the trailing-zero
 * scan).  vconj == 0: H = I - tau v v^H with v as stored (QL::MultQ_unblocked,
 * QL.hpp:272, 287).  vconj == 1: v = conj(stored) (RQ::MultQ_unblocked,
 * RQ.hpp:279, 295). */
static void FN(apply_bottom)(const char *side, int vconj, size_t m, size_t n, const T *v, size_t incv, T tau,
                             T *c, size_t ldc, T *work){
	size_t lenv = 0, lenc = 0;
	const int left = ('L' == side[0]);
	if(0.0 != tau){
		lenv = left ? m : n;
		if(0 == lenv){ return; }
		size_t i = (lenv-1)*incv;
		while(lenv > 0 && 0.0 == v[i]){
			--lenv;
			i -= incv;
		}
		if(left){
			lenc = FN(last_nonzero_column_limit)(lenv, n, c, ldc);
		}else{
			lenc = FN(last_nonzero_row_limit)(m, lenv, c, ldc);
		}
		if(0 == lenc){ return; }
	}
	if(left){
		const size_t r = lenv-1;  /* the implied-1 row */
		for(size_t j = 0; j < lenc; ++j){
			T w = 0;
			if(vconj){
				for(size_t i = 0; i < r; ++i){ w += c[i+j*ldc] * v[i*incv]; }
				w = CONJ(w + c[r+j*ldc]);
			}else{
				w = CONJ(c[r+j*ldc]);
				for(size_t i = 0; i < r; ++i){ w += CONJ(c[i+j*ldc]) * v[i*incv]; }
			}
			work[j] = w;
		}
		for(size_t j = 0; j < lenc; ++j){ c[r+j*ldc] += -tau * CONJ(work[j]); }
		for(size_t j = 0; j < lenc; ++j){
			for(size_t i = 0; i < r; ++i){
				const T vi = vconj ? CONJ(v[i*incv]) : v[i*incv];
				c[i+j*ldc] -= tau * vi * CONJ(work[j]);
			}
		}
	}else{
		const size_t r = lenv-1;  /* the implied-1 column */
		for(size_t i = 0; i < lenc; ++i){ work[i] = c[i+r*ldc]; }
		for(size_t j = 0; j < r; ++j){
			const T vj = vconj ? CONJ(v[j*incv]) : v[j*incv];
			for(size_t i = 0; i < lenc; ++i){ work[i] += c[i+j*ldc] * vj; }
		}
		for(size_t i = 0; i < lenc; ++i){ c[i+r*ldc] += -tau * work[i]; }
		for(size_t j = 0; j < r; ++j){
			const T vj = vconj ? v[j*incv] : CONJ(v[j*incv]);  /* v^H entries */
			for(size_t i = 0; i < lenc; ++i){ c[i+j*ldc] += -tau * work[i] * vj; }
		}
	}
}

/* QL::Factor_unblocked, QL.hpp:79-100. */
void FN(ql_factor_unblocked)(size_t m, size_t n, T *a, size_t lda, T *tau, T *work){
	const size_t k = m < n ? m : n;
	size_t i = k;
	while(i-- > 0){
		const size_t r = m-k+i, cc = n-k+i;
		T alpha = a[r+cc*lda];
		FN(generate)(r+1, &alpha, &a[0+cc*lda], 1, &tau[i]);
		a[r+cc*lda] = 1;
		FN(apply)("L", 0, r+1, cc, &a[0+cc*lda], 1, CONJ(tau[i]), a, lda, work);
		a[r+cc*lda] = alpha;
	}
}

/* QL::Factor, QL.hpp:126-205: the checks (lda >= m, lwork given) and the
 * workspace query (nb n, nb = QL::Tuning default 32) as there; the factors
 * by the unblocked loop for all sizes (the blocked loop, QL.hpp:175-196,
 * computes the same factors in exact arithmetic). */
void FN(ql_factor)(size_t m, size_t n, T *a, size_t lda, T *tau, size_t *lwork, T *work){
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(NULL == lwork){ orc_assert_fail("NULL != lwork", __LINE__); return; }
	const size_t k = m < n ? m : n;
	if(0 == k){ return; }
	if(NULL == work || 0 == *lwork){
		*lwork = 32*n;
		return;
	}
	FN(ql_factor_unblocked)(m, n, a, lda, tau, work);
}

/* QL::MultQ_unblocked, QL.hpp:245-300: forward loop for (L,N) and (R,C),
 * backward otherwise; tau conjugated for "C".  Reflector i acts on the
 * leading nq-k+i+1 rows (columns) of C. */
void FN(ql_multq_unblocked)(const char *side, const char *trans, size_t m, size_t n, size_t k,
                            const T *a, size_t lda, const T *tau, T *c, size_t ldc, T *work){
	if(!(ldc >= m)){ orc_assert_fail("ldc >= m", __LINE__); }
	const int left = ('L' == side[0]);
	const int notran = ('N' == trans[0]);
	if(!((left && k <= m) || (!left && k <= n))){ orc_assert_fail("k range", __LINE__); }
	if(!(lda >= k)){ orc_assert_fail("lda >= k", __LINE__); }
	if(m < 1 || n < 1 || k < 1){ return; }
	const int forward = ((left && notran) || (!left && !notran));
	for(size_t s = 0; s < k; ++s){
		const size_t i = forward ? s : k-1-s;
		const size_t mi = left ? m-k+i+1 : m, ni = left ? n : n-k+i+1;
		const T taui = notran ? tau[i] : CONJ(tau[i]);
		FN(apply_bottom)(side, 0, mi, ni, &a[0+i*lda], 1, taui, c, ldc, work);
	}
}

/* QL::MultQ, QL.hpp:335-430: checks and the workspace query (nb nw + nb^2);
 * the product by the unblocked loop. */
void FN(ql_multq)(const char *side, const char *trans, size_t m, size_t n, size_t k,
                  const T *a, size_t lda, const T *tau, T *c, size_t ldc, size_t *lwork, T *work){
	if(!(ldc >= m)){ orc_assert_fail("ldc >= m", __LINE__); }
	if(0 == m || 0 == n || 0 == k){ return; }
	const int left = ('L' == side[0]);
	const size_t nw = left ? n : m;
	if(0 == *lwork || NULL == work){
		*lwork = 32*nw + 32*32;
		return;
	}
	FN(ql_multq_unblocked)(side, trans, m, n, k, a, lda, tau, c, ldc, work);
}

/* QL::GenerateQ_unblocked, QL.hpp:445-471: the last n columns of
 * H[k-1] ... H[0]. */
void FN(ql_generateq_unblocked)(size_t m, size_t n, size_t k, T *a, size_t lda, const T *tau, T *work){
	if(!(m >= n)){ orc_assert_fail("m >= n", __LINE__); }
	if(!(k <= n)){ orc_assert_fail("k <= n", __LINE__); }
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(n < 1){ return; }
	for(size_t j = 0; j+k < n; ++j){
		for(size_t i = 0; i < m; ++i){ a[i+j*lda] = 0; }
		a[m-n+j+j*lda] = 1;
	}
	for(size_t i = 0; i < k; ++i){
		const size_t ii = n-k+i, r = m-n+ii;
		a[r+ii*lda] = 1;
		FN(apply)("L", 0, r+1, ii, &a[0+ii*lda], 1, tau[i], a, lda, work);
		for(size_t l = 0; l < r; ++l){ a[l+ii*lda] *= -tau[i]; }
		a[r+ii*lda] = 1.0 - tau[i];
		for(size_t l = r+1; l < m; ++l){ a[l+ii*lda] = 0; }
	}
}

/* QL::GenerateQ, QL.hpp:501-565: checks and the workspace query (n nb, nb
 * from QR::Tuning as there); Q by the unblocked loop. */
void FN(ql_generateq)(size_t m, size_t n, size_t k, T *a, size_t lda, const T *tau, size_t *lwork, T *work){
	if(!(k <= n)){ orc_assert_fail("k <= n", __LINE__); }
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(0 == n){ return; }
	if(0 == *lwork || NULL == work){
		*lwork = n*32;
		return;
	}
	FN(ql_generateq_unblocked)(m, n, k, a, lda, tau, work);
}

/* RQ::Factor_unblocked, RQ.hpp:80-102: row m-k+i conjugated, its reflector
 * generated with the 1 at column n-k+i, applied from the right to the rows
 * above, the stored part conjugated back. */
void FN(rq_factor_unblocked)(size_t m, size_t n, T *a, size_t lda, T *tau, T *work){
	const size_t k = m < n ? m : n;
	size_t i = k;
	while(i-- > 0){
		const size_t r = m-k+i, cc = n-k+i;
		FN(conjugate)(cc+1, &a[r+0*lda], lda);
		T alpha = a[r+cc*lda];
		FN(generate)(cc+1, &alpha, &a[r+0*lda], lda, &tau[i]);
		a[r+cc*lda] = 1;
		FN(apply)("R", 0, r, cc+1, &a[r+0*lda], lda, tau[i], a, lda, work);
		a[r+cc*lda] = alpha;
		FN(conjugate)(cc, &a[r+0*lda], lda);
	}
}

/* RQ::Factor, RQ.hpp:131-205: checks (lda >= m, lwork given, 0 or >= m) and
 * the workspace query (nb m, nb = RQ::Tuning default 32); the unblocked
 * loop for all sizes. */
void FN(rq_factor)(size_t m, size_t n, T *a, size_t lda, T *tau, size_t *lwork, T *work){
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(NULL == lwork){ orc_assert_fail("NULL != lwork", __LINE__); return; }
	if(!(0 == *lwork || *lwork >= m)){ orc_assert_fail("0 == *lwork || *lwork >= m", __LINE__); }
	const size_t k = m < n ? m : n;
	if(0 == k){ return; }
	if(NULL == work || 0 == *lwork){
		*lwork = 32*m;
		return;
	}
	FN(rq_factor_unblocked)(m, n, a, lda, tau, work);
}

/* RQ::MultQ_unblocked, RQ.hpp:245-303: forward loop for (L,C) and (R,N),
 * backward otherwise; tau conjugated for "N"; the stored rows conjugated
 * (vconj). */
void FN(rq_multq_unblocked)(const char *side, const char *trans, size_t m, size_t n, size_t k,
                            const T *a, size_t lda, const T *tau, T *c, size_t ldc, T *work){
	if(!(ldc >= m)){ orc_assert_fail("ldc >= m", __LINE__); }
	const int left = ('L' == side[0]);
	const int notran = ('N' == trans[0]);
	if(!((left && k <= m) || (!left && k <= n))){ orc_assert_fail("k range", __LINE__); }
	if(!(lda >= k)){ orc_assert_fail("lda >= k", __LINE__); }
	if(m < 1 || n < 1 || k < 1){ return; }
	const int forward = ((left && !notran) || (!left && notran));
	for(size_t s = 0; s < k; ++s){
		const size_t i = forward ? s : k-1-s;
		const size_t mi = left ? m-k+i+1 : m, ni = left ? n : n-k+i+1;
		const T taui = notran ? CONJ(tau[i]) : tau[i];
		FN(apply_bottom)(side, 1, mi, ni, &a[i+0*lda], lda, taui, c, ldc, work);
	}
}

/* RQ::MultQ, RQ.hpp:344-440: checks and the workspace query (nb nw + nb^2);
 * the product by the unblocked loop. */
void FN(rq_multq)(const char *side, const char *trans, size_t m, size_t n, size_t k,
                  const T *a, size_t lda, const T *tau, T *c, size_t ldc, size_t *lwork, T *work){
	if(!(ldc >= m)){ orc_assert_fail("ldc >= m", __LINE__); }
	if(0 == m || 0 == n || 0 == k){ return; }
	const int left = ('L' == side[0]);
	const size_t nw = left ? n : m;
	if(0 == *lwork || NULL == work){
		*lwork = 32*nw + 32*32;
		return;
	}
	FN(rq_multq_unblocked)(side, trans, m, n, k, a, lda, tau, c, ldc, work);
}

/* RQ::GenerateQ_unblocked, RQ.hpp:457-491: the last m rows of
 * H[0]^H ... H[k-1]^H. */
void FN(rq_generateq_unblocked)(size_t m, size_t n, size_t k, T *a, size_t lda, const T *tau, T *work){
	if(!(n >= m)){ orc_assert_fail("n >= m", __LINE__); }
	if(!(k <= m)){ orc_assert_fail("k <= m", __LINE__); }
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(0 == m){ return; }
	if(k < m){
		for(size_t j = 0; j < n; ++j){
			for(size_t i = 0; i+k < m; ++i){ a[i+j*lda] = 0; }
			if(j >= n-m && j < n-k){ a[m-n+j+j*lda] = 1; }
		}
	}
	for(size_t i = 0; i < k; ++i){
		const size_t ii = m-k+i, cc = n-m+ii;
		FN(conjugate)(cc, &a[ii+0*lda], lda);
		a[ii+cc*lda] = 1;
		FN(apply)("R", 0, ii, cc+1, &a[ii+0*lda], lda, CONJ(tau[i]), a, lda, work);
		for(size_t l = 0; l < cc; ++l){ a[ii+l*lda] *= -tau[i]; }
		FN(conjugate)(cc, &a[ii+0*lda], lda);
		a[ii+cc*lda] = 1.0 - CONJ(tau[i]);
		for(size_t l = cc+1; l < n; ++l){ a[ii+l*lda] = 0; }
	}
}

/* RQ::GenerateQ, RQ.hpp:519-580: checks and the workspace query (m nb, nb
 * from QR::Tuning as there); Q by the unblocked loop. */
void FN(rq_generateq)(size_t m, size_t n, size_t k, T *a, size_t lda, const T *tau, size_t *lwork, T *work){
	if(!(n >= m)){ orc_assert_fail("n >= m", __LINE__); }
	if(!(k <= m)){ orc_assert_fail("k <= m", __LINE__); }
	if(!(lda >= m)){ orc_assert_fail("lda >= m", __LINE__); }
	if(0 == m){ return; }
	if(0 == *lwork || NULL == work){
		*lwork = m*32;
		return;
	}
	FN(rq_generateq_unblocked)(m, n, k, a, lda, tau, work);
}

/* ---------------------------------------------------------------------
 * Hessenberg (RNP/LA/Hessenberg.hpp): the Q of a reduction, through QR.
 * ------------------------------------------------------------------- */

/* Hessenberg::MultQ, Hessenberg.hpp:478-514: QR::MultQ on the reflector
 * block one row below the diagonal (nh = ihi-ilo-1 reflectors from column
 * ilo), applied to rows (side "L") or columns (side "R") ilo+1.. of C. */
void FN(hess_multq)(const char *side, const char *trans, size_t m, size_t n, size_t ilo, size_t ihi,
                    const T *a, size_t lda, const T *tau, T *c, size_t ldc, size_t *lwork, T *work){
	const int left = ('L' == side[0]);
	const size_t nh = ihi-ilo-1;
	const size_t nq = left ? m : n, nw = left ? n : m;
	if(!('L' == side[0] || 'R' == side[0])){ orc_assert_fail("side", __LINE__); }
	if(!('N' == trans[0] || 'T' == trans[0] || 'C' == trans[0])){ orc_assert_fail("trans", __LINE__); }
	if(!(ilo <= nq)){ orc_assert_fail("ilo <= nq", __LINE__); }
	if(!(ilo < ihi)){ orc_assert_fail("ilo < ihi", __LINE__); }
	if(!(ihi <= nq)){ orc_assert_fail("ihi <= nq", __LINE__); }
	if(NULL == lwork){ orc_assert_fail("NULL != lwork", __LINE__); return; }
	if(!(0 == *lwork || *lwork >= nw)){ orc_assert_fail("0 == *lwork || *lwork >= nw", __LINE__); }
	const size_t mi = left ? nh : m, ni = left ? n : nh;
	const size_t i1 = left ? ilo+1 : 0, i2 = left ? 0 : ilo+1;
	FN(multq)(side, trans, mi, ni, nh, &a[ilo+1+ilo*lda], lda, &tau[ilo], &c[i1+i2*ldc], ldc, lwork, work, NULL);
}

/* Hessenberg::GenerateQ, Hessenberg.hpp:540-588: the reflectors shifted one
 * column right, identity outside rows/columns [ilo+1, ihi), then
 * QR::GenerateQ on the nh x nh block. */
void FN(hess_generateq)(size_t n, size_t ilo, size_t ihi, T *a, size_t lda, const T *tau, size_t *lwork, T *work){
	const size_t nh = ihi-ilo-1;
	if(!(ilo <= n)){ orc_assert_fail("ilo <= n", __LINE__); }
	if(!(ilo < ihi)){ orc_assert_fail("ilo < ihi", __LINE__); }
	if(!(ihi <= n)){ orc_assert_fail("ihi <= n", __LINE__); }
	if(!(lda >= n)){ orc_assert_fail("lda >= n", __LINE__); }
	if(NULL == lwork){ orc_assert_fail("NULL != lwork", __LINE__); return; }
	if(!(0 == *lwork || *lwork >= nh)){ orc_assert_fail("0 == *lwork || *lwork >= nh", __LINE__); }
	if(0 == *lwork || NULL == work){
		*lwork = nh*32;
		return;
	}
	if(0 == n){ return; }
	size_t j = ihi;
	while(j-- > ilo+1){
		for(size_t i = 0; i < j; ++i){ a[i+j*lda] = 0; }
		for(size_t i = j+1; i < ihi; ++i){ a[i+j*lda] = a[i+(j-1)*lda]; }
		for(size_t i = ihi; i < n; ++i){ a[i+j*lda] = 0; }
	}
	for(j = 0; j <= ilo; ++j){
		for(size_t i = 0; i < n; ++i){ a[i+j*lda] = 0; }
		a[j+j*lda] = 1;
	}
	for(j = ihi; j < n; ++j){
		for(size_t i = 0; i < n; ++i){ a[i+j*lda] = 0; }
		a[j+j*lda] = 1;
	}
	if(nh > 0){
		size_t lw = *lwork;
		FN(generateq)(nh, nh, nh, &a[ilo+1+(ilo+1)*lda], lda, &tau[ilo], &lw, work, NULL);
	}
}

/* ---------------------------------------------------------------------
 * LU (RNP/LA/LU.hpp): partial pivoting, pivots are 0-based absolute rows.
 * ------------------------------------------------------------------- */

/* Traits<T>::norm1: |re| + |im|. */
static double FN(norm1s)(T x){ return fabs(RE(x)) + fabs(IM(x)); }

/* BLAS::MaximumIndex, BLAS_mix.hpp:798-810: the first index of the largest
 * |re| + |im| (a NaN is never preferred to the first entry). */
static size_t FN(max_index)(size_t n, const T *x, size_t incx){
	if(n < 1){ return 0; }
	double mv = FN(norm1s)(x[0]);
	size_t mi = 0;
	for(size_t i = 1; i < n; ++i){
		const double cv = FN(norm1s)(x[i*incx]);
		if(cv > mv){ mi = i; mv = cv; }
	}
	return mi;
}

static void FN(swap_rows)(size_t ncols, T *a, size_t lda, size_t r1, size_t r2){
	for(size_t c = 0; c < ncols; ++c){
		const T t = a[r1+c*lda];
		a[r1+c*lda] = a[r2+c*lda];
		a[r2+c*lda] = t;
	}
}

/* LU::Factor_unblocked, LU.hpp:52-75, with its return value as written
 * there: the last column j whose pivot is exactly zero, returned as j+1,
 * which reads as success (0) when that column is column 0. */
int FN(lu_factor_unblocked)(size_t m, size_t n, T *a, size_t lda, size_t *piv){
	int info = 0;
	const size_t k = m < n ? m : n;
	for(size_t j = 0; j < k; ++j){
		const size_t jp = j + FN(max_index)(m-j, &a[j+j*lda], 1);
		piv[j] = jp;
		if(0.0 != a[jp+j*lda]){
			if(jp != j){ FN(swap_rows)(n, a, lda, j, jp); }
			const T r = 1.0 / a[j+j*lda];
			for(size_t i = j+1; i < m; ++i){ a[i+j*lda] *= r; }
		}else{
			info = (int)j;
		}
		/* Rank1Update(m-j-1, n-j-1, -1, x = A(j+1:, j), y = A(j, j+1:)) */
		for(size_t c = j+1; c < n; ++c){
			const T t = -a[j+c*lda];
			for(size_t i = j+1; i < m; ++i){ a[i+c*lda] += a[i+j*lda] * t; }
		}
	}
	return info ? info+1 : 0;
}

/* SolveTrM("L","L","N","U"): B <- L^-1 B, L unit lower (m x m). */
static void FN(trsm_llnu)(size_t m, size_t n, const T *a, size_t lda, T *b, size_t ldb){
	for(size_t c = 0; c < n; ++c){
		for(size_t k = 0; k < m; ++k){
			const T bk = b[k+c*ldb];
			if(0.0 == bk){ continue; }
			for(size_t i = k+1; i < m; ++i){ b[i+c*ldb] -= bk * a[i+k*lda]; }
		}
	}
}

/* LU::Factor, LU.hpp:77-115: blocks of nb = 64 columns (LU::Tuning): the
 * unblocked panel, pivots made absolute, the panel's swaps applied to the
 * columns left and right of it (ApplyPermutations "L","F", Permutation.hpp:
 * 33-52), U12 = L11^-1 A12, A22 -= L21 U12; unblocked throughout when
 * min(m,n) <= 64.  The return value is the reference's arithmetic: the
 * first panel whose Factor_unblocked reported iinfo > 0 gives info =
 * iinfo + j, returned as info + 1. */
int FN(lu_factor)(size_t m, size_t n, T *a, size_t lda, size_t *piv){
	const size_t nb = 64;
	const size_t k = m < n ? m : n;
	if(k <= nb){ return FN(lu_factor_unblocked)(m, n, a, lda, piv); }
	int info = 0;
	const T one = 1, mone = -1;
	for(size_t j = 0; j < k; j += nb){
		const size_t jb = (k < nb+j ? k-j : nb);
		const int iinfo = FN(lu_factor_unblocked)(m-j, jb, &a[j+j*lda], lda, &piv[j]);
		if(0 == info && iinfo > 0){ info = iinfo + (int)j; }
		const size_t ilim = (m < j+jb ? m : j+jb);
		for(size_t i = j; i < ilim; ++i){ piv[i] += j; }
		for(size_t i = j; i < j+jb; ++i){
			if(piv[i] != i){ FN(swap_rows)(j, a, lda, i, piv[i]); }
		}
		if(j+jb < n){
			for(size_t i = j; i < j+jb; ++i){
				if(piv[i] != i){ FN(swap_rows)(n-j-jb, &a[(j+jb)*lda], lda, i, piv[i]); }
			}
			FN(trsm_llnu)(jb, n-j-jb, &a[j+j*lda], lda, &a[j+(j+jb)*lda], lda);
			if(j+jb < m){
				FN(gemm)("N", "N", m-j-jb, n-j-jb, jb, mone, &a[j+jb+j*lda], lda, &a[j+(j+jb)*lda], lda, one,
				         &a[j+jb+(j+jb)*lda], lda);
			}
		}
	}
	return info ? info+1 : 0;
}

/* LU::Invert, RNP/LA/LU.hpp:146-214: inv(A) in place from its LU factors.
 *  - Triangular::Invert("U","N") (RNP/LA/Triangular.hpp:207-253): returns
 *    -(i+1) for the first zero U_ii and leaves a untouched; otherwise inv(U)
 *    in the upper triangle.  Restated in its unblocked form
 *    (Triangular.hpp:150-172: U_jj <- 1/U_jj, then column j above the
 *    diagonal <- -U_jj * triu(U(0:j,0:j)) U(0:j,j), a triangular
 *    matrix-vector product); the blocked form (Triangular.hpp:227-236) is the
 *    same product reassociated.
 *  - inv(A) L = inv(U) for inv(A), columns from the right (LU.hpp:172-185):
 *    save and zero L's column j, then a(:,j) -= a(:,j+1:n) l(j+1:n).
 *  - the column interchanges backward (LU.hpp:210, ApplyPermutations "R" "B",
 *    RNP/LA/Permutation.hpp:60-66).
 * The reference's blocked branch (LU.hpp:186-207, taken for n > 64 with a
 * full workspace) saves columns jj instead of j+jj of L and so is wrong for
 * every block but the first; this restatement computes the inverse both
 * branches intend (LAPACK dgetri's), which the unblocked branch delivers.
 * work holds n elements. */
int FN(lu_invert)(size_t n, T *a, size_t lda, const size_t *ipiv, T *work){
	if(0 == n){ return 0; }
	for(size_t i = 0; i < n; ++i){
		if((T)0 == a[i+i*lda]){ return -(int)(i+1); }
	}
	for(size_t j = 0; j < n; ++j){
		a[j+j*lda] = (T)1 / a[j+j*lda];
		const T ajj = -a[j+j*lda];
		/* x = triu(U(0:j,0:j)) x in place, x = a(0:j, j): row i uses x_k, k >= i */
		for(size_t i = 0; i < j; ++i){
			T acc = 0;
			for(size_t k = i; k < j; ++k){ acc += a[i+k*lda] * a[k+j*lda]; }
			a[i+j*lda] = acc;
		}
		for(size_t i = 0; i < j; ++i){ a[i+j*lda] *= ajj; }
	}
	size_t j = n;
	while(j-- > 0){
		for(size_t i = j+1; i < n; ++i){
			work[i] = a[i+j*lda];
			a[i+j*lda] = 0;
		}
		if(j+1 < n){
			for(size_t i = 0; i < n; ++i){
				T acc = 0;
				for(size_t k = j+1; k < n; ++k){ acc += a[i+k*lda] * work[k]; }
				a[i+j*lda] -= acc;
			}
		}
	}
	j = n;
	while(j-- > 0){
		const size_t jp = ipiv[j];
		if(jp == j){ continue; }
		for(size_t i = 0; i < n; ++i){
			const T t = a[i+j*lda];
			a[i+j*lda] = a[i+jp*lda];
			a[i+jp*lda] = t;
		}
	}
	return 0;
}

/* One triangular solve of LU::Solve (LU.hpp:117-144) on every column of b:
 * op(M) x = b with M the unit lower (lower) or upper part of a. */
static void FN(lu_trsv)(int lower, int unit, char op, size_t n, size_t nrhs, const T *a, size_t lda, T *b,
                        size_t ldb){
	for(size_t c = 0; c < nrhs; ++c){
		T *x = &b[c*ldb];
		/* op(M) is lower for (lower, N) and (upper, T/C): forward */
		const int fwd = (lower && 'N' == op) || (!lower && 'N' != op);
		for(size_t s = 0; s < n; ++s){
			const size_t i = fwd ? s : n-1-s;
			T acc = x[i];
			for(size_t t = 0; t < s; ++t){
				const size_t q = fwd ? t : n-1-t;  /* already solved */
				T mij = ('N' == op) ? a[i+q*lda] : a[q+i*lda];
				if('C' == op){ mij = CONJ(mij); }
				acc -= mij * x[q];
			}
			if(!unit){
				T d = a[i+i*lda];
				if('C' == op){ d = CONJ(d); }
				acc = acc / d;
			}
			x[i] = acc;
		}
	}
}

/* LU::Solve, LU.hpp:117-144: "N": the swaps forward, L (unit) then U;
 * "T"/"C": op(U) then op(L), the swaps backward. */
void FN(lu_solve)(const char *trans, size_t n, size_t nrhs, const T *a, size_t lda, const size_t *ipiv, T *b,
                  size_t ldb){
	if(0 == n || 0 == nrhs){ return; }
	if('N' == trans[0]){
		for(size_t i = 0; i < n; ++i){
			if(ipiv[i] != i){ FN(swap_rows)(nrhs, b, ldb, i, ipiv[i]); }
		}
		FN(lu_trsv)(1, 1, 'N', n, nrhs, a, lda, b, ldb);
		FN(lu_trsv)(0, 0, 'N', n, nrhs, a, lda, b, ldb);
	}else{
		const char op = ('T' == trans[0]) ? 'T' : 'C';
		FN(lu_trsv)(0, 0, op, n, nrhs, a, lda, b, ldb);
		FN(lu_trsv)(1, 1, op, n, nrhs, a, lda, b, ldb);
		size_t i = n;
		while(i-- > 0){
			if(ipiv[i] != i){ FN(swap_rows)(nrhs, b, ldb, i, ipiv[i]); }
		}
	}
}

/* MatrixNorm1Estimate_unblocked, MatrixNormEstimates.hpp:17-107: Hager's
 * 1-norm estimate with Higham's refinements (zlacn2), for an operator
 * given as aop(trans, n, x, data) that overwrites x with op(B) x. */
static double FN(norm1_estimate)(size_t n, void (*aop)(char, size_t, T *, void *), void *data, T *work){
	const double safmin = 2.2250738585072014e-308;
	T *v = work, *x = work + n;
	for(size_t i = 0; i < n; ++i){ x[i] = 1.0 / (double)n; }
	aop('N', n, x, data);
	if(1 == n){ return TABS(x[0]); }
	double est = 0.0;
	for(size_t i = 0; i < n; ++i){ est += TABS(x[i]); }
	for(size_t i = 0; i < n; ++i){
		const double ax = TABS(x[i]);
		if(ax > safmin){ x[i] /= ax; }else{ x[i] = 1.0; }
	}
	aop('C', n, x, data);
	size_t imax = 0;
	{
		double armax = 0.0;
		for(size_t i = 0; i < n; ++i){
			const double ar = fabs(RE(x[i]));
			if(ar > armax){ armax = ar; imax = i; }
		}
	}
	size_t iter = 1, jlast = 0;
	do{
		++iter;
		for(size_t i = 0; i < n; ++i){ x[i] = 0.0; }
		x[imax] = 1.0;
		aop('N', n, x, data);
		for(size_t i = 0; i < n; ++i){ v[i] = x[i]; }
		const double estold = est;
		est = 0.0;
		for(size_t i = 0; i < n; ++i){ est += TABS(v[i]); }
		if(est <= estold){ break; }
		for(size_t i = 0; i < n; ++i){
			const double ax = TABS(x[i]);
			if(ax > safmin){ x[i] /= ax; }else{ x[i] = 1.0; }
		}
		aop('C', n, x, data);
		jlast = imax;
		imax = 0;
		double armax = 0.0;
		for(size_t i = 0; i < n; ++i){
			const double ar = fabs(RE(x[i]));
			if(ar > armax){ armax = ar; imax = i; }
		}
	}while(TABS(x[jlast]) != TABS(x[imax]) && iter < 5);
	double altsgn = 1.0;
	for(size_t i = 0; i < n; ++i){
		x[i] = altsgn * (1.0 + (double)i / (double)(n-1));
		altsgn = -altsgn;
	}
	aop('N', n, x, data);
	double t = 0.0;
	for(size_t i = 0; i < n; ++i){ t += TABS(x[i]); }
	t = 2.0 * t / (double)(3*n);
	if(t >= est){ est = t; }
	return est;
}

/* Util::Norm1Estimator, LU.hpp:34-50 (with the LU solve it intends: the
 * reference calls an undeclared PLUSolve and reads data->ldav): "N" solves
 * with op_t then scales by w; "C" scales then solves with op_n. */
typedef struct { const T *af; size_t ldaf; const size_t *ipiv; char tn, tt; const double *w; } FN(refine_est);
static void FN(refine_aop)(char trans, size_t n, T *x, void *data_){
	const FN(refine_est) *d = (const FN(refine_est) *)data_;
	if('N' == trans){
		const char t[2] = {d->tt, 0};
		FN(lu_solve)(t, n, 1, d->af, d->ldaf, d->ipiv, x, n);
		for(size_t i = 0; i < n; ++i){ x[i] *= d->w[i]; }
	}else{
		for(size_t i = 0; i < n; ++i){ x[i] *= d->w[i]; }
		const char t[2] = {d->tn, 0};
		FN(lu_solve)(t, n, 1, d->af, d->ldaf, d->ipiv, x, n);
	}
}

/* LU::RefineSolution, LU.hpp:216-366: per right-hand side, residual
 * r = b - op(A) x, componentwise backward error
 * berr = max_i |r_i| / (|op(A)| |x| + |b|)_i (safe1 added where the
 * denominator is below safe2), a correction by LU::Solve while berr > eps,
 * berr at least halves and fewer than refine_max_iters = 5 tries; then
 * ferr = ||inv(op(A)) diag(w)||_1 estimate / max |x_i| with
 * w = |r| + nz eps (|op(A)||x| + |b|) (+ safe1 where the latter is below
 * safe2, as the comment at LU.hpp:322-323 and LAPACK's gerfs state; the
 * line at LU.hpp:329 multiplies by safe1 instead). */
void FN(lu_refine)(const char *trans, size_t n, size_t nrhs, const T *a, size_t lda, const T *af, size_t ldaf,
                   const size_t *ipiv, const T *b, size_t ldb, T *x, size_t ldx, double *ferr, double *berr){
	if(0 == n || 0 == nrhs){
		for(size_t j = 0; j < nrhs; ++j){ if(ferr){ ferr[j] = 0; } if(berr){ berr[j] = 0; } }
		return;
	}
	const double eps = 2.220446049250313e-16;  /* Traits<double>::eps(): numeric_limits epsilon, 2^-52 */
	const double dmin = 2.2250738585072014e-308;
	const size_t nz = n + 1;
	const double safe1 = (double)nz * dmin, safe2 = safe1 / eps;
	const char tr = trans[0];
	T *work = (T *)malloc(sizeof(T) * 2 * n);
	T *est_work = (T *)malloc(sizeof(T) * 2 * n);
	double *rw = (double *)malloc(sizeof(double) * n);
	for(size_t j = 0; j < nrhs; ++j){
		size_t count = 1;
		double lstres = 3.0, s = 0.0;
		for(;;){
			for(size_t i = 0; i < n; ++i){ work[i] = b[i+j*ldb]; }
			for(size_t k = 0; k < n; ++k){
				for(size_t i = 0; i < n; ++i){
					if('N' == tr){ work[i] -= a[i+k*lda] * x[k+j*ldx]; }
					else{
						const T aik = ('C' == tr) ? CONJ(a[k+i*lda]) : a[k+i*lda];
						work[i] -= aik * x[k+j*ldx];
					}
				}
			}
			for(size_t i = 0; i < n; ++i){ rw[i] = FN(norm1s)(b[i+j*ldb]); }
			if('N' == tr){
				for(size_t k = 0; k < n; ++k){
					const double xk = FN(norm1s)(x[k+j*ldx]);
					for(size_t i = 0; i < n; ++i){ rw[i] += FN(norm1s)(a[i+k*lda]) * xk; }
				}
			}else{
				for(size_t k = 0; k < n; ++k){
					double t = 0.0;
					for(size_t i = 0; i < n; ++i){ t += FN(norm1s)(a[i+k*lda]) * FN(norm1s)(x[i+j*ldx]); }
					rw[k] += t;
				}
			}
			s = 0.0;
			for(size_t i = 0; i < n; ++i){
				const double r = (rw[i] > safe2) ? FN(norm1s)(work[i]) / rw[i]
				                                 : (FN(norm1s)(work[i]) + safe1) / (rw[i] + safe1);
				if(r > s){ s = r; }
			}
			if(berr){ berr[j] = s; }
			if(s > eps && 2.0 * s <= lstres && count < 5){
				FN(lu_solve)(trans, n, 1, af, ldaf, ipiv, work, n);
				for(size_t i = 0; i < n; ++i){ x[i+j*ldx] += work[i]; }
				lstres = s;
				++count;
				continue;
			}
			break;
		}
		for(size_t i = 0; i < n; ++i){
			if(rw[i] > safe2){ rw[i] = FN(norm1s)(work[i]) + (double)nz * eps * rw[i]; }
			else{ rw[i] = FN(norm1s)(work[i]) + (double)nz * eps * rw[i] + safe1; }
		}
		FN(refine_est) ed = {af, ldaf, ipiv, ('N' == tr) ? 'N' : 'C', ('N' == tr) ? 'C' : 'N', rw};
		const double fe = FN(norm1_estimate)(n, FN(refine_aop), &ed, est_work);
		double xm = 0.0;
		for(size_t i = 0; i < n; ++i){
			const double v = FN(norm1s)(x[i+j*ldx]);
			if(v > xm){ xm = v; }
		}
		if(ferr){ ferr[j] = (0.0 != xm) ? fe / xm : fe; }
	}
	free(work);
	free(est_work);
	free(rw);
}

/* Least squares by QR: Factor (QR.hpp:129), MultQ("L","C") on B
 * (QR.hpp:337), then R x = (Q^H B)(0:n,:) by back substitution
 * (SolveTrM "L","U","N","N", BLAS_mix.hpp:2096). */
void FN(ls_solve)(size_t m, size_t n, size_t nrhs, T *a, size_t lda, T *b, size_t ldb){
	size_t lw = 0;
	FN(factor)(m, n, a, lda, NULL, &lw, NULL, NULL);
	size_t k = m < n ? m : n;
	T *tau = (T*)calloc(k ? k : 1, sizeof(T));
	T *work = (T*)calloc(lw ? lw : 1, sizeof(T));
	FN(factor)(m, n, a, lda, tau, &lw, work, NULL);
	free(work);
	lw = 0;
	FN(multq)("L", "C", m, nrhs, n, a, lda, tau, b, ldb, &lw, NULL, NULL);
	work = (T*)calloc(lw ? lw : 1, sizeof(T));
	FN(multq)("L", "C", m, nrhs, n, a, lda, tau, b, ldb, &lw, work, NULL);
	free(work);
	FN(trsm_lunn)(n, nrhs, a, lda, b, ldb);
	free(tau);
}

#if !IS_CPLX
/* RNP::MatrixGen::RandomMatrix, MatrixGen.hpp:120-198, Symmetry::Nonsymmetric:
 * diag(d) pre- and post-multiplied by random Householder reflectors drawn
 * from Normal01 (mindim of them on each side, last index first). */
void FN(random_matrix)(int dist, int mode, int reverse, double cond, double dmax,
                       size_t m, size_t n, T *a, size_t lda, int *iseed){
	const size_t mindim = (m < n ? m : n);
	if(0 == mindim){ return; }
	double *d = (double*)malloc(sizeof(double)*mindim);
	T *work = (T*)malloc(sizeof(T)*(m+n));
	orc_generate_diagonal(dist, mode, reverse, 0, cond, dmax, mindim, d, iseed);
	for(size_t j = 0; j < n; ++j){
		for(size_t i = 0; i < m; ++i){ a[i+j*lda] = (i == j) ? (T)d[i] : (T)0; }
	}
	free(d);
	size_t i = mindim;
	while(i-- > 0){
		for(int pass = 0; pass < 2; ++pass){
			const int leftside = (0 == pass);
			const size_t len = leftside ? m-i : n-i;
			if(!(i+1 < (leftside ? m : n))){ continue; }
			FN(generate_vector)(ORC_NORMAL01, len, work, iseed);
			double wn = FN(norm2)(len, work, 1);
			T wa = (wn / TABS(work[0])) * work[0];
			T tau = 0;
			if(0.0 != wn){
				T wb = work[0] + wa;
				{
					T inv = MKT(1.0, 0.0) / wb;
					for(size_t q = 1; q < len; ++q){ work[q] *= inv; }
				}
				work[0] = 1;
				tau = RE(wb / wa);
			}
			if(leftside){
				FN(gemv)("C", m-i, n-i, 1, &a[i+i*lda], lda, work, 1, 0, &work[m], 1);
				FN(gerc)(m-i, n-i, -tau, work, 1, &work[m], 1, &a[i+i*lda], lda);
			}else{
				FN(gemv)("N", m-i, n-i, 1, &a[i+i*lda], lda, work, 1, 0, &work[n], 1);
				FN(gerc)(m-i, n-i, -tau, &work[n], 1, work, 1, &a[i+i*lda], lda);
			}
		}
	}
	free(work);
}
#endif /* !IS_CPLX: the reference's RandomMatrix takes Traits<real>::abs of a T
        (MatrixGen.hpp:172), which only compiles for real T */
