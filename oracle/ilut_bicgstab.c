/*
 * CPU restatement of the reference's default linear solve, for the CPU baseline and as a
 * second FP64 solver in the tests (test infrastructure only; nothing in the GPU product
 * links it).
 *
 * The reference's solve_linear (src/solver.cpp:53-80, defaults include/edgefem/solver.hpp:16-36)
 * calls Eigen 3.4's BiCGSTAB with its IncompleteLUT preconditioner (fill factor 10, drop
 * tolerance 1e-4), relative tolerance 1e-10, at most 10000 iterations, x0 = 0.  Eigen is not
 * vendored and is absent from this machine, so both routines are restated from their
 * published algorithms:
 *
 *  - ILUT: Y. Saad, "ILUT: a dual threshold incomplete LU factorization", Numer. Linear
 *    Algebra Appl. 1 (1994) 387-402, row-wise IKJ elimination, with the parameter meanings
 *    Eigen documents for IncompleteLUT: per row, keep at most fill/2 entries in L and in U
 *    where fill = nnz(A) * fillfactor / n + 1; a multiplier l_ik with |l_ik| <= droptol is
 *    dropped; an entry of U with |u_ij| <= droptol * ||a_i||_2 is dropped; a zero pivot is
 *    replaced by sqrt(droptol) * ||a_i||_2.  The factorization runs on the symmetrically
 *    permuted matrix P A P^T; Eigen orders with AMD on the pattern of A + A^T, which is not
 *    restated here: the caller passes the permutation (reverse Cuthill-McKee in this
 *    build).
 *  - BiCGSTAB: H. A. van der Vorst, SIAM J. Sci. Stat. Comput. 13 (1992) 631-644, right
 *    preconditioned, with the Hermitian inner products Eigen's dot() uses and its restart
 *    rule (when |r0^H r| < eps^2 ||r0||^2 the shadow residual r0 is reset to the current
 *    true residual; the first restart also resets the iteration count).
 *
 * Build: oracle/Makefile (gcc -O3 -march=x86-64-v3 -shared -fPIC).  C99 complex doubles.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef double complex cplx;

typedef struct {
  int64_t n;
  int64_t *lp, *up;  /* row pointers of L (strict lower, unit diagonal implied) and U (diagonal first) */
  int32_t *lc, *uc;
  cplx *lv, *uv;
  int32_t *perm;     /* new -> old (NULL: identity) */
} ilut_t;

void efo_ilut_free(void *h);

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* binary min-heap of column indices */
static void heap_push(int32_t *h, int *len, int32_t v) {
  int i = (*len)++;
  while (i > 0) {
    int p = (i - 1) >> 1;
    if (h[p] <= v)
      break;
    h[i] = h[p];
    i = p;
  }
  h[i] = v;
}
static int32_t heap_pop(int32_t *h, int *len) {
  int32_t top = h[0];
  int32_t v = h[--(*len)];
  int i = 0;
  for (;;) {
    int c = 2 * i + 1;
    if (c >= *len)
      break;
    if (c + 1 < *len && h[c + 1] < h[c])
      c++;
    if (v <= h[c])
      break;
    h[i] = h[c];
    i = c;
  }
  if (*len > 0)
    h[i] = v;
  return top;
}

/* Keep the `keep` entries of largest magnitude among idx[0..m) (partial selection, order of
 * the kept entries unspecified).  mag[] holds |w| of each listed column. */
static void select_largest(int32_t *idx, double *mag, int m, int keep) {
  if (keep >= m)
    return;
  int lo = 0, hi = m - 1;
  while (lo < hi) {
    double piv = mag[(lo + hi) >> 1];
    int i = lo, j = hi;
    while (i <= j) {
      while (mag[i] > piv) i++;
      while (mag[j] < piv) j--;
      if (i <= j) {
        double tm = mag[i]; mag[i] = mag[j]; mag[j] = tm;
        int32_t ti = idx[i]; idx[i] = idx[j]; idx[j] = ti;
        i++; j--;
      }
    }
    if (keep - 1 <= j)
      hi = j;
    else if (keep - 1 >= i)
      lo = i;
    else
      break;
  }
}

/* Returns 0 on success, 1 on a zero row, 2 on allocation failure, 3 when time_limit > 0
 * seconds passed before the last row (then *out is NULL and *rows_done says how far the
 * factorization got: a rate for a bounded sample). */
int efo_ilut_factor(int64_t n, const int64_t *rp, const int32_t *ci, const cplx *v, const int32_t *perm,
                    double droptol, int fillfactor, double time_limit, void **out, int64_t *rows_done,
                    double *seconds) {
  const double t0 = now_s();
  *out = NULL;
  ilut_t *F = (ilut_t *)calloc(1, sizeof(ilut_t));
  if (!F)
    return 2;
  F->n = n;
  int32_t *inv = NULL;
  if (perm) {
    F->perm = (int32_t *)malloc(n * sizeof(int32_t));
    inv = (int32_t *)malloc(n * sizeof(int32_t));
    if (!F->perm || !inv)
      return 2;
    memcpy(F->perm, perm, n * sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i)
      inv[perm[i]] = (int32_t)i;
  }
  const int64_t nnz = rp[n];
  int64_t fill = nnz * (int64_t)fillfactor / n + 1;
  if (fill > n)
    fill = n;
  const int nnzL = (int)(fill / 2), nnzU = (int)(fill / 2);
  int64_t capL = (int64_t)n * nnzL + 16, capU = (int64_t)n * (nnzU + 1) + 16;
  F->lp = (int64_t *)malloc((n + 1) * sizeof(int64_t));
  F->up = (int64_t *)malloc((n + 1) * sizeof(int64_t));
  F->lc = (int32_t *)malloc(capL * sizeof(int32_t));
  F->lv = (cplx *)malloc(capL * sizeof(cplx));
  F->uc = (int32_t *)malloc(capU * sizeof(int32_t));
  F->uv = (cplx *)malloc(capU * sizeof(cplx));
  cplx *w = (cplx *)calloc(n, sizeof(cplx));
  int32_t *mark = (int32_t *)malloc(n * sizeof(int32_t));  /* 1 = in the row's pattern */
  int32_t *ul = (int32_t *)malloc(n * sizeof(int32_t));    /* U-part columns of the row */
  int32_t *ll = (int32_t *)malloc(n * sizeof(int32_t));    /* kept multipliers' columns */
  int32_t *heap = (int32_t *)malloc(n * sizeof(int32_t));
  double *mag = (double *)malloc(n * sizeof(double));
  if (!F->lp || !F->up || !F->lc || !F->lv || !F->uc || !F->uv || !w || !mark || !ul || !ll || !heap || !mag)
    return 2;
  for (int64_t i = 0; i < n; ++i)
    mark[i] = 0;
  int64_t nl = 0, nu = 0;
  F->lp[0] = F->up[0] = 0;
  int timed_out = 0;
  int64_t i;
  for (i = 0; i < n; ++i) {
    if (time_limit > 0 && (i & 1023) == 0 && now_s() - t0 > time_limit) {
      timed_out = 1;
      break;
    }
    const int64_t ro = perm ? perm[i] : i;  /* row i of P A P^T is row perm[i] of A */
    int hl = 0, nul = 0, nll = 0;
    double rownorm = 0.0;
    int diag_in = 0;
    for (int64_t k = rp[ro]; k < rp[ro + 1]; ++k) {
      const int32_t j = perm ? inv[ci[k]] : ci[k];
      if (!mark[j]) {
        mark[j] = 1;
        w[j] = 0;
        if (j < i)
          heap_push(heap, &hl, j);
        else if (j > i)
          ul[nul++] = j;
        else
          diag_in = 1;
      }
      w[j] += v[k];
      rownorm += creal(v[k]) * creal(v[k]) + cimag(v[k]) * cimag(v[k]);
    }
    if (!diag_in) {
      mark[i] = 1;
      w[i] = 0;
    }
    rownorm = sqrt(rownorm);
    if (rownorm == 0.0)
      return 1;
    /* eliminate with the previous rows, smallest column first */
    while (hl > 0) {
      const int32_t k = heap_pop(heap, &hl);
      const cplx fact = w[k] / F->uv[F->up[k]];
      if (cabs(fact) <= droptol) {  /* dropped multiplier */
        mark[k] = 0;
        w[k] = 0;
        continue;
      }
      for (int64_t q = F->up[k] + 1; q < F->up[k + 1]; ++q) {
        const int32_t j = F->uc[q];
        if (!mark[j]) {
          mark[j] = 1;
          w[j] = 0;
          if (j < i)
            heap_push(heap, &hl, j);
          else if (j > i)
            ul[nul++] = j;
        }
        w[j] -= fact * F->uv[q];
      }
      w[k] = fact;
      ll[nll++] = k;
    }
    /* L: the nnzL largest multipliers */
    for (int q = 0; q < nll; ++q)
      mag[q] = cabs(w[ll[q]]);
    select_largest(ll, mag, nll, nnzL);
    const int keepL = nll < nnzL ? nll : nnzL;
    for (int q = 0; q < keepL; ++q) {
      F->lc[nl] = ll[q];
      F->lv[nl++] = w[ll[q]];
    }
    for (int q = 0; q < nll; ++q) {
      mark[ll[q]] = 0;
      w[ll[q]] = 0;
    }
    /* diagonal, with the zero-pivot shift */
    cplx d = w[i];
    if (d == 0)
      d = sqrt(droptol) * rownorm;
    F->uc[nu] = (int32_t)i;
    F->uv[nu++] = d;
    mark[i] = 0;
    w[i] = 0;
    /* U: drop small entries, then the nnzU - 1 largest off the diagonal (the candidates go
     * to uk, the now empty heap array, so that ul still lists every touched column) */
    int32_t *uk = heap;
    int m = 0;
    for (int q = 0; q < nul; ++q) {
      const int32_t j = ul[q];
      const double a = cabs(w[j]);
      if (a > droptol * rownorm) {
        uk[m] = j;
        mag[m++] = a;
      }
    }
    select_largest(uk, mag, m, nnzU - 1);
    const int keepU = m < nnzU - 1 ? m : nnzU - 1;
    for (int q = 0; q < keepU; ++q) {
      F->uc[nu] = uk[q];
      F->uv[nu++] = w[uk[q]];
    }
    for (int q = 0; q < nul; ++q) {
      mark[ul[q]] = 0;
      w[ul[q]] = 0;
    }
    F->lp[i + 1] = nl;
    F->up[i + 1] = nu;
  }
  free(w); free(mark); free(ul); free(ll); free(heap); free(mag); free(inv);
  if (rows_done)
    *rows_done = i;
  if (seconds)
    *seconds = now_s() - t0;
  if (timed_out) {
    efo_ilut_free(F);
    return 3;
  }
  *out = F;
  return 0;
}

int64_t efo_ilut_nnz(void *h) {
  ilut_t *F = (ilut_t *)h;
  return F->lp[F->n] + F->up[F->n];
}

void efo_ilut_free(void *h) {
  ilut_t *F = (ilut_t *)h;
  if (!F)
    return;
  free(F->lp); free(F->up); free(F->lc); free(F->uc); free(F->lv); free(F->uv); free(F->perm);
  free(F);
}

/* y = P^T U^-1 L^-1 P x  (t: work vector of length n) */
static void ilut_apply(const ilut_t *F, const cplx *x, cplx *y, cplx *t) {
  const int64_t n = F->n;
  for (int64_t i = 0; i < n; ++i) {
    cplx s = F->perm ? x[F->perm[i]] : x[i];
    for (int64_t q = F->lp[i]; q < F->lp[i + 1]; ++q)
      s -= F->lv[q] * t[F->lc[q]];
    t[i] = s;
  }
  for (int64_t i = n - 1; i >= 0; --i) {
    cplx s = t[i];
    for (int64_t q = F->up[i] + 1; q < F->up[i + 1]; ++q)
      s -= F->uv[q] * t[F->uc[q]];
    t[i] = s / F->uv[F->up[i]];
  }
  for (int64_t i = 0; i < n; ++i) {
    if (F->perm)
      y[F->perm[i]] = t[i];
    else
      y[i] = t[i];
  }
}

/* y = P^T U^-1 L^-1 P x, the preconditioner's action (exported for the tests). */
int efo_ilut_apply(void *h, const cplx *x, cplx *y) {
  const ilut_t *F = (const ilut_t *)h;
  cplx *t = (cplx *)malloc(F->n * sizeof(cplx));
  if (!t)
    return 2;
  ilut_apply(F, x, y, t);
  free(t);
  return 0;
}

static void spmv(int64_t n, const int64_t *rp, const int32_t *ci, const cplx *v, const cplx *x, cplx *y) {
  for (int64_t i = 0; i < n; ++i) {
    cplx s = 0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
      s += v[k] * x[ci[k]];
    y[i] = s;
  }
}

static cplx dotc(int64_t n, const cplx *a, const cplx *b) {  /* a^H b */
  cplx s = 0;
  for (int64_t i = 0; i < n; ++i)
    s += conj(a[i]) * b[i];
  return s;
}

static double nrm2sq(int64_t n, const cplx *a) {
  double s = 0;
  for (int64_t i = 0; i < n; ++i)
    s += creal(a[i]) * creal(a[i]) + cimag(a[i]) * cimag(a[i]);
  return s;
}

/* BiCGSTAB on A x = b with the ILUT preconditioner h (NULL: none), x0 = 0.
 * Stops when ||r|| <= tol ||b||, after max_iter iterations, or when time_limit > 0
 * seconds have passed (status 2).  Returns 0 converged, 1 not converged, 2 time limit,
 * 3 allocation failure.  *iters, *resid (recursive relative residual), *seconds. */
int efo_bicgstab(int64_t n, const int64_t *rp, const int32_t *ci, const cplx *v, void *h, const cplx *b,
                 cplx *x, double tol, int max_iter, double time_limit, int *iters, double *resid,
                 double *seconds) {
  const double t0 = now_s();
  const ilut_t *F = (const ilut_t *)h;
  cplx *r = (cplx *)malloc(n * sizeof(cplx)), *r0 = (cplx *)malloc(n * sizeof(cplx));
  cplx *p = (cplx *)calloc(n, sizeof(cplx)), *vv = (cplx *)calloc(n, sizeof(cplx));
  cplx *y = (cplx *)malloc(n * sizeof(cplx)), *z = (cplx *)malloc(n * sizeof(cplx));
  cplx *s = (cplx *)malloc(n * sizeof(cplx)), *t = (cplx *)malloc(n * sizeof(cplx));
  cplx *work = (cplx *)malloc(n * sizeof(cplx));
  if (!r || !r0 || !p || !vv || !y || !z || !s || !t || !work)
    return 3;
  for (int64_t i = 0; i < n; ++i) {
    x[i] = 0;
    r[i] = b[i];
    r0[i] = b[i];
  }
  double r0sq = nrm2sq(n, r0);
  const double bsq = nrm2sq(n, b);
  int status = 1, it = 0, restarts = 0;
  if (bsq == 0.0) {
    *iters = 0;
    *resid = 0.0;
    status = 0;
    goto done;
  }
  const double tol2 = tol * tol * bsq, eps2 = DBL_EPSILON * DBL_EPSILON;
  cplx rho = 1, alpha = 1, om = 1;
  while (nrm2sq(n, r) > tol2 && it < max_iter) {
    if (time_limit > 0 && now_s() - t0 > time_limit) {
      status = 2;
      break;
    }
    const cplx rho_old = rho;
    rho = dotc(n, r0, r);
    if (cabs(rho) < eps2 * r0sq) {  /* restart with the true residual as shadow */
      spmv(n, rp, ci, v, x, t);
      for (int64_t i = 0; i < n; ++i) {
        r[i] = b[i] - t[i];
        r0[i] = r[i];
      }
      rho = r0sq = nrm2sq(n, r);
      if (restarts++ == 0)
        it = 0;
    }
    const cplx beta = (rho / rho_old) * (alpha / om);
    for (int64_t i = 0; i < n; ++i)
      p[i] = r[i] + beta * (p[i] - om * vv[i]);
    if (F) ilut_apply(F, p, y, work); else memcpy(y, p, n * sizeof(cplx));
    spmv(n, rp, ci, v, y, vv);
    alpha = rho / dotc(n, r0, vv);
    for (int64_t i = 0; i < n; ++i)
      s[i] = r[i] - alpha * vv[i];
    if (F) ilut_apply(F, s, z, work); else memcpy(z, s, n * sizeof(cplx));
    spmv(n, rp, ci, v, z, t);
    const double tt = nrm2sq(n, t);
    om = tt > 0.0 ? dotc(n, t, s) / tt : 0.0;
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * y[i] + om * z[i];
      r[i] = s[i] - om * t[i];
    }
    ++it;
  }
  if (status != 2)
    status = nrm2sq(n, r) <= tol2 ? 0 : 1;
  *iters = it;
  *resid = sqrt(nrm2sq(n, r) / bsq);
done:
  free(r); free(r0); free(p); free(vv); free(y); free(z); free(s); free(t); free(work);
  if (seconds)
    *seconds = now_s() - t0;
  return status;
}
