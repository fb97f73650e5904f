/* FP64 CPU restatement of the FIRTEZ-dz per-pixel Levenberg-Marquardt step:
 * node basis, equivalent response functions, chi^2 / damping update, Jacobian
 * normalisation, the damped pseudo-inverse solve and the perturbation
 * application with its limits.
 *
 * TEST INFRASTRUCTURE (the oracle).  Only tests/ and the smoke check call it,
 * as the checker.  Each routine cites what it restates.  Arithmetic keeps the
 * reference's REAL(4)/REAL(8) split: the model is float, everything else
 * double.  The pseudo-inverse uses a cyclic Jacobi eigen-decomposition of the
 * symmetric damped Hessian; for a symmetric matrix the singular values are the
 * absolute eigenvalues, so the truncated pseudo-inverse is the one the
 * reference builds from LAPACK DGESVD (lm.f90:61-133), to rounding.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fzcpu.h"

/* get_dmodel.f90:31 */
static const double PARAM_NORMS[8] = {6000.0, 1.0e5, 1.0e-5, 1.0e3, 1.0e3, 1.0e3, 1.0e6, 6.0e2};
/* model column (0-based, MODEL(NZ,13)) of each physical parameter:
 * get_dmodel.f90:1887-1894 (ATMPOS) and the pointer selects of :1937-1962 */
static const int ATMCOL[8] = {0, 1, 2, 5, 6, 7, 10, 1};
/* hard bounds of REARANGE_DA_MODEL (get_dmodel.f90:1903-1919) */
static const double LOWLIM[8] = {2500.0, 1.0e-2, 1.0e-10, -5.0e4, -5.0e4, -5.0e4, -1.0e7, 1.0e-8};
static const double UPPLIM[8] = {1.8e4, 1.0e12, 1.0e10, 5.0e4, 5.0e4, 5.0e4, 1.0e7, 1.0e12};

/* ------------------------------------------------ node basis */

/* GET_POSITIONS + GET_INT_COEFS (get_dmodel.f90:463-477, 563-644):
 * hat functions centred on nj equidistant positions in height-index space;
 * coefs[j*nz + k] = COEFS(k+1, j+1). */
int fzc_int_coefs(int nz, int nj, double* coefs) {
  if (nj < 1 || nj > nz) return 5;
  memset(coefs, 0, sizeof(double) * (size_t)nz * nj);
  if (nj == 1) {
    for (int k = 0; k < nz; ++k) coefs[k] = 1.0;
    return 0;
  }
  if (nj == nz) {
    for (int k = 0; k < nz; ++k) coefs[(size_t)k * nz + k] = 1.0;
    return 0;
  }
  double* pos = (double*)malloc(sizeof(double) * nj);
  double step = (double)(nz - 1) / (double)(nj - 1);
  for (int j = 0; j < nj; ++j) pos[j] = ((double)(j + 1) - 1.0) * step + 1.0;
  for (int j = 0; j < nj; ++j) {
    double ji = pos[j > 0 ? j - 1 : 0];
    double jf = pos[j + 1 < nj ? j + 1 : nj - 1];
    double jm = pos[j];
    for (int i = 1; i <= nz; ++i) {
      double x = (double)i;
      if (x < ji || x > jf) continue;
      double c;
      if (x == jm)
        c = 1.0;
      else if (x < jm)
        c = (1.0 - 0.0) / (jm - ji) * (x - ji);
      else
        c = (1.0 - 0.0) / (jf - jm) * (jf - x);
      coefs[(size_t)j * nz + (i - 1)] = c;
    }
  }
  free(pos);
  return 0;
}

/* GET_NJEVALS (get_dmodel.f90:375-407): nodes requested for parameter apar
 * (1-based) in cycle curic. */
int fzc_njevals(int apar, int curic, int maxiter, int cycpow, int toffset, int nz) {
  if (apar == 1) {
    if (curic == 0 || curic == maxiter) return nz;
    double v = pow((double)cycpow, (double)(curic + toffset - 1));
    int iv = (int)v;
    return iv < nz ? iv : nz;
  }
  if (curic == 0) return nz;
  if (curic == 1) return 1;
  if (curic == maxiter) return nz;
  double v = pow((double)cycpow, (double)(curic - 1));
  int iv = (int)v;
  return iv < nz ? iv : nz;
}

/* EXPAND_VECTOR / CONTRACT_VECTOR (get_dmodel.f90:839-880).  The contraction
 * takes a REAL(SP) height profile and divides by the column sums of COEFS. */
static void expand(int nz, int nj, const double* iv, const double* coefs, double* ov) {
  for (int k = 0; k < nz; ++k) {
    double s = 0.0;
    for (int j = 0; j < nj; ++j) s = s + iv[j] * coefs[(size_t)j * nz + k];
    ov[k] = s;
  }
}
static void contract(int nz, int nj, const float* iv, const double* coefs, double* ov) {
  for (int j = 0; j < nj; ++j) {
    double s = 0.0, w = 0.0;
    for (int k = 0; k < nz; ++k) s = s + (double)iv[k] * coefs[(size_t)j * nz + k];
    for (int k = 0; k < nz; ++k) w += coefs[(size_t)j * nz + k];
    ov[j] = s / w;
  }
}

/* ------------------------------------------------ equivalent response functions */

/* CALCULATE_EQUIVALENT_DSYN (forward.f90:552-681, no LSF): for each inverted
 * parameter and inverted Stokes parameter, EQVDSYN = DSYN1D(s,p,:,:) x COEFS
 * times GAIN1D, stored REAL(SP) as EQVDSYN(NFREQ, IFREEP) column-major.
 * dsyn is the oracle's propagated response function, already multiplied by
 * GAIN1D: dsyn[((k*numw + l)*8 + p)*4 + s]. */
int fzc_eqv_dsyn(const fzc_lmcfg* c, const double* coefs, const double* dsyn, float* eqv) {
  int nz = c->nz, numw = c->numw;
  int nstk = 0;
  for (int s = 0; s < 4; ++s) nstk += c->invstk[s] != 0;
  int nfreq = nstk * numw;
  int ifreep = 0;
  for (int p = 0; p < 8; ++p)
    if (c->atmpar[p]) ifreep += c->nodes[p];
  memset(eqv, 0, sizeof(float) * (size_t)nfreq * ifreep);
  int jcnt = 0;
  for (int p = 0; p < 8; ++p) {
    if (!c->atmpar[p]) continue;
    int nj = c->nodes[p];
    int is = 0;
    for (int s = 0; s < 4; ++s) {
      if (!c->invstk[s]) continue;
      for (int j = 0; j < nj; ++j) {
        const double* cf = coefs + (size_t)(jcnt + j) * nz;
        for (int l = 0; l < numw; ++l) {
          double acc = 0.0;
          for (int k = 0; k < nz; ++k) acc += dsyn[(((size_t)k * numw + l) * 8 + p) * 4 + s] * cf[k];
          eqv[(size_t)(jcnt + j) * nfreq + is * numw + l] = (float)acc;
        }
      }
      ++is;
    }
    jcnt += nj;
  }
  /* columns whose absolute sum overflows REAL(SP) are zeroed (forward.f90:672-678) */
  for (int j = 0; j < ifreep; ++j) {
    float t = 0.0f;
    for (int f = 0; f < nfreq; ++f) t += fabsf(eqv[(size_t)j * nfreq + f]);
    if (t > 3.40282347e38f)
      for (int f = 0; f < nfreq; ++f) eqv[(size_t)j * nfreq + f] = 0.0f;
  }
  return 0;
}

/* ------------------------------------------------ damped solve (lm.f90) */

/* Symmetric eigen-decomposition by the cyclic Jacobi method (Golub & Van
 * Loan, "Matrix Computations", sec. 8.5): a[n*n] row-major is destroyed, its
 * eigenvalues are left in d, eigenvectors in the columns of v (row-major). */
static void jacobi_eig(int n, double* a, double* d, double* v) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) v[i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        tot += a[i * n + j] * a[i * n + j];
        if (i != j) off += a[i * n + j] * a[i * n + j];
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = a[p * n + q];
        if (apq == 0.0) continue;
        double app = a[p * n + p], aqq = a[q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
        double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
        for (int k = 0; k < n; ++k) { /* A <- A J */
          double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = cs * akp - sn * akq;
          a[k * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) { /* A <- J^T A */
          double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = cs * apk - sn * aqk;
          a[q * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = cs * vkp - sn * vkq;
          v[k * n + q] = sn * vkp + cs * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) d[i] = a[i * n + i];
}

/* NEW_CALCULATE_DELTA_UBICGSTAB (lm.f90:375-453): a conjugate-gradient
 * iteration on the damped system, stopped on a relative residual test.
 * alp[n*n] column-major ALP(NPAR,NPAR). */
static void cg_solve(int n, double* delt, const double* alp, const double* bet) {
  double* xv = (double*)calloc(n, sizeof(double));
  double* pv = (double*)malloc(sizeof(double) * n);
  double* rv = (double*)malloc(sizeof(double) * n);
  double* qv = (double*)malloc(sizeof(double) * n);
  double* ck = (double*)malloc(sizeof(double) * n);
#define MATVEC(out, x)                                                        \
  for (int i_ = 0; i_ < n; ++i_) {                                            \
    double s_ = 0.0;                                                          \
    for (int j_ = 0; j_ < n; ++j_) s_ += alp[(size_t)j_ * n + i_] * (x)[j_]; \
    (out)[i_] = s_;                                                           \
  }
  for (int i = 0; i < n; ++i) delt[i] = 0.0;
  for (int i = 0; i < n; ++i) pv[i] = bet[i];
  MATVEC(rv, xv);
  for (int i = 0; i < n; ++i) rv[i] = bet[i] - rv[i];
  double crho = 0.0;
  for (int i = 0; i < n; ++i) crho += rv[i] * rv[i];
  MATVEC(qv, pv);
  double calpha = 0.0;
  for (int i = 0; i < n; ++i) calpha += pv[i] * qv[i];
  calpha = crho / calpha;
  for (int i = 0; i < n; ++i) xv[i] = xv[i] + calpha * pv[i];
  for (int i = 0; i < n; ++i) rv[i] = rv[i] - calpha * qv[i];
  double pcondit = 1.0e29, dcondit = -1.0e29;
  int converged = 0;
  for (int it = 1; it <= 100000; ++it) {
    double crho0 = crho;
    crho = 0.0;
    for (int i = 0; i < n; ++i) crho += rv[i] * rv[i];
    double cbeta = crho / crho0;
    for (int i = 0; i < n; ++i) pv[i] = rv[i] + cbeta * pv[i];
    MATVEC(qv, pv);
    calpha = 0.0;
    for (int i = 0; i < n; ++i) calpha += pv[i] * qv[i];
    calpha = crho / calpha;
    for (int i = 0; i < n; ++i) xv[i] = xv[i] + calpha * pv[i];
    for (int i = 0; i < n; ++i) rv[i] = rv[i] - calpha * qv[i];
    MATVEC(ck, xv);
    double condit = 0.0;
    for (int i = 0; i < n; ++i) {
      double e = (ck[i] - bet[i]) / (fabs(bet[i]) + 1.0e-5);
      condit += e * e;
    }
    condit = condit / (double)n;
    if (condit < 1.0e-3 || dcondit > 1.0e30) {
      for (int i = 0; i < n; ++i) delt[i] = xv[i];
      converged = 1;
      break;
    }
    dcondit = condit - pcondit;
    pcondit = condit;
  }
  /* The reference leaves its FAILED flag unset when the loop runs out; the
   * restatement treats that case as failed (delta = x/10). */
  if (!converged)
    for (int i = 0; i < n; ++i) delt[i] = xv[i] / 10.0;
#undef MATVEC
  free(xv);
  free(pv);
  free(rv);
  free(qv);
  free(ck);
}

/* GET_PERTURBATION (lm.f90:137-204) with HESSIAN (:30-39), GET_TERM1
 * (:43-57) and GET_INVERSE (:61-133):
 *   H = J^T J + REG_HSS, TERM1 = H + lambda diag(H),
 *   TERM2 = J^T DIFF - REG_PEN,  DELTA = TERM1^+ TERM2
 * with the pseudo-inverse dropping singular values below svdtol * max.
 * jac[(size_t)p*nf + f] = IJACOBIAN(f+1,p+1); reg_hss column-major. */
int fzc_get_perturbation(int np, int nf, const double* jac, const double* diff, double lambda,
                         double* delta, const double* reg_pen, const double* reg_hss, int dtype,
                         double svdtol) {
  double* h = (double*)malloc(sizeof(double) * np * np);
  double* t2 = (double*)malloc(sizeof(double) * np);
  for (int i = 0; i < np; ++i) {
    for (int j = 0; j < np; ++j) {
      double s = 0.0;
      for (int f = 0; f < nf; ++f) s += jac[(size_t)i * nf + f] * jac[(size_t)j * nf + f];
      h[i * np + j] = s;
    }
    double s = 0.0;
    for (int f = 0; f < nf; ++f) s += jac[(size_t)i * nf + f] * diff[f];
    t2[i] = s - (reg_pen ? reg_pen[i] : 0.0);
  }
  if (reg_hss)
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) h[i * np + j] += reg_hss[(size_t)j * np + i];
  for (int i = 0; i < np; ++i) h[i * np + i] = h[i * np + i] + h[i * np + i] * lambda;
  int rc = 0;
  if (dtype == 1) {
    /* symmetrise (H is symmetric up to rounding) and invert */
    double* a = (double*)malloc(sizeof(double) * np * np);
    double* v = (double*)malloc(sizeof(double) * np * np);
    double* d = (double*)malloc(sizeof(double) * np);
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) a[i * np + j] = 0.5 * (h[i * np + j] + h[j * np + i]);
    jacobi_eig(np, a, d, v);
    double smax = 0.0;
    for (int i = 0; i < np; ++i) smax = fabs(d[i]) > smax ? fabs(d[i]) : smax;
    double thr = smax * (svdtol > 0 ? svdtol : 10.0e-4);
    for (int i = 0; i < np; ++i) delta[i] = 0.0;
    for (int e = 0; e < np; ++e) {
      if (fabs(d[e]) < thr) continue;
      double proj = 0.0;
      for (int i = 0; i < np; ++i) proj += v[i * np + e] * t2[i];
      proj /= d[e];
      for (int i = 0; i < np; ++i) delta[i] += v[i * np + e] * proj;
    }
    free(a);
    free(v);
    free(d);
  } else if (dtype == 2) {
    double* cm = (double*)malloc(sizeof(double) * np * np);
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) cm[(size_t)j * np + i] = h[i * np + j];
    cg_solve(np, delta, cm, t2);
    free(cm);
  } else {
    rc = 8;
  }
  free(h);
  free(t2);
  return rc;
}

/* ------------------------------------------------ the per-pixel step */

static int ifreep_of(const fzc_lmcfg* c) {
  int n = 0;
  for (int p = 0; p < 8; ++p)
    if (c->atmpar[p]) n += c->nodes[p];
  return n;
}

/* GET_IT_PENALTIES (get_dmodel.f90:1041-1197) on the current model:
 * per inverted parameter, node-sampled model / PARAM_NORM, a Tikhonov
 * operator (type 0: identity/NJ, type 1: first differences) scaled by
 * sqrt(PEN_ALP/IFREEP); PEN_RES = L^T L m, PEN_HSS = L^T L (block diagonal). */
static void penalties(const fzc_lmcfg* c, const double* coefs, const float* model_cur, double* pen_res,
                      double* pen_hss) {
  int nz = c->nz, ifreep = ifreep_of(c);
  memset(pen_res, 0, sizeof(double) * ifreep);
  memset(pen_hss, 0, sizeof(double) * (size_t)ifreep * ifreep);
  int ji = 0;
  for (int p = 0; p < 8; ++p) {
    if (!c->atmpar[p]) continue;
    int jd = c->nodes[p];
    const double* cf = coefs + (size_t)ji * nz;
    double alp = sqrt(c->pen_alp[p] / (double)ifreep);
    double* sm = (double*)malloc(sizeof(double) * jd);
    double* L = (double*)calloc((size_t)jd * jd, sizeof(double)); /* row-major L[k][j] */
    double* pen = (double*)malloc(sizeof(double) * jd);
    contract(nz, jd, model_cur + (size_t)ATMCOL[p] * nz, cf, sm);
    for (int j = 0; j < jd; ++j) sm[j] = sm[j] / PARAM_NORMS[p];
    if (c->pen_typ[p] == 0) {
      for (int k = 0; k < jd; ++k) L[k * jd + k] = 1.0 * (1.0 / (double)jd);
    } else if (c->pen_typ[p] == 1) {
      for (int k = 1; k < jd; ++k) {
        L[k * jd + k] = 1.0 * 1.0;
        L[k * jd + k - 1] = -1.0 * 1.0;
      }
    }
    for (int i = 0; i < jd * jd; ++i) L[i] = L[i] * alp;
    for (int k = 0; k < jd; ++k) {
      double s = 0.0;
      for (int j = 0; j < jd; ++j) s += L[k * jd + j] * sm[j];
      pen[k] = s;
    }
    for (int j = 0; j < jd; ++j) {
      double s = 0.0;
      for (int k = 0; k < jd; ++k) s += L[k * jd + j] * pen[k];
      pen_res[ji + j] = s;
    }
    for (int i = 0; i < jd; ++i)
      for (int j = 0; j < jd; ++j) {
        double s = 0.0;
        for (int k = 0; k < jd; ++k) s += L[k * jd + i] * L[k * jd + j];
        pen_hss[(size_t)(ji + j) * ifreep + (ji + i)] = s;
      }
    free(sm);
    free(L);
    free(pen);
    ji += jd;
  }
}

/* UPDATE_LAMBDA (get_dmodel.f90:917-956) */
static void update_lambda(double* chicur, double* chipre, double* lambda, double* steps, double* precyc,
                          double factor) {
  const double lmin = 1.0e-2;
  if ((*chipre - *chicur) > 0) {
    if ((*chipre - *chicur) > factor * *chipre) {
      double v = *steps - 1.0;
      *steps = v > 0.0 ? v : 0.0;
    } else {
      *steps = *steps + 1.0;
    }
    *chipre = *chicur;
    *lambda = *lambda / sqrt(10.0);
  } else {
    if (fabs(*precyc - 1.0) < 1.0e-2) {
      *precyc = 0.0;
      *lambda = *lambda * 3.0;
    } else {
      *lambda = *lambda * 10.0;
    }
    *steps = *steps + 1.0;
  }
  *lambda = *lambda > lmin ? *lambda : lmin;
}

/* SORT_DELTA_TO_ATMOSPHERE (get_dmodel.f90:1715-1836): scale the node
 * perturbation, limit it to max(absolute, mpert*|node mean|), expand it to
 * heights and apply the hard bounds.  On return ipar holds the height
 * perturbation (float); delta[jofs..] the limited node perturbation. */
static void sort_delta(const fzc_lmcfg* c, double* delta, int jofs, int nj, const double* cf, double factor,
                       double inf, double sup, double absolute, float* ipar) {
  int nz = c->nz;
  double* itd = (double*)malloc(sizeof(double) * nj);
  double* avg = (double*)malloc(sizeof(double) * nj);
  double* lim = (double*)malloc(sizeof(double) * nj);
  double* nzp = (double*)malloc(sizeof(double) * nz);
  float* opar = (float*)malloc(sizeof(float) * nz);
  float* tmp = (float*)malloc(sizeof(float) * nz);
  for (int j = 0; j < nj; ++j) itd[j] = delta[jofs + j] * factor;
  contract(nz, nj, ipar, cf, avg);
  for (int j = 0; j < nj; ++j) {
    if (absolute > 0) {
      double a = fabs(c->mpert * avg[j]);
      lim[j] = absolute > a ? absolute : a;
    } else {
      lim[j] = fabs(c->mpert * avg[j]);
    }
    if (avg[j] == 0) lim[j] = absolute;
  }
  for (int j = 0; j < nj; ++j)
    if (fabs(itd[j]) > lim[j]) itd[j] = itd[j] / fabs(itd[j]) * lim[j];
  for (int j = 0; j < nj; ++j) delta[jofs + j] = itd[j];
  double* sumv = (double*)malloc(sizeof(double) * nj);
  for (int j = 0; j < nj; ++j) sumv[j] = avg[j] + itd[j];
  expand(nz, nj, sumv, cf, nzp);
  if (c->node_expansion == 2) {
    for (int k = 0; k < nz; ++k) {
      opar[k] = (float)nzp[k];
      if ((double)opar[k] < inf) opar[k] = (float)inf;
      if ((double)opar[k] > sup) opar[k] = (float)sup;
    }
    for (int k = 0; k < nz; ++k) ipar[k] = opar[k] - ipar[k];
  } else if (c->node_expansion == 1) {
    for (int k = 0; k < nz; ++k) nzp[k] = nzp[k] - (double)ipar[k];
    for (int k = 0; k < nz; ++k) tmp[k] = (float)nzp[k];
    contract(nz, nj, tmp, cf, itd);
    expand(nz, nj, itd, cf, nzp);
    for (int j = 0; j < nj; ++j) delta[jofs + j] = itd[j];
    for (int k = 0; k < nz; ++k) opar[k] = ipar[k] + (float)nzp[k];
    for (int k = 0; k < nz; ++k) {
      if ((double)opar[k] < inf) nzp[k] = inf - (double)ipar[k];
      if ((double)opar[k] > sup) nzp[k] = sup - (double)ipar[k];
    }
    for (int k = 0; k < nz; ++k) ipar[k] = (float)nzp[k];
  }
  free(sumv);
  free(itd);
  free(avg);
  free(lim);
  free(nzp);
  free(opar);
  free(tmp);
}

/* REARANGE_DA_MODEL (get_dmodel.f90:1843-2016, ERRORS absent): apply the
 * limited perturbation to the current model, parameter by parameter, with
 * the temperature tail rule in log tau5 (:1970-1984). */
static void rearange(const fzc_lmcfg* c, const double* coefs, double* delta, float* model_cur) {
  int nz = c->nz;
  int jpar = 0;
  float* bumod = (float*)malloc(sizeof(float) * nz);
  const float* tau = model_cur + (size_t)11 * nz;
  for (int p = 0; p < 8; ++p) {
    if (!c->atmpar[p]) continue;
    int nj = c->nodes[p];
    float* col = model_cur + (size_t)ATMCOL[p] * nz;
    memcpy(bumod, col, sizeof(float) * nz);
    sort_delta(c, delta, jpar, nj, coefs + (size_t)jpar * nz, PARAM_NORMS[p], LOWLIM[p], UPPLIM[p],
               PARAM_NORMS[p] * c->mpert, col);
    if (p == 0) {
      const float tllim = -11.5f, tulim = 0.5f;
      for (int k = nz - 1; k >= 0; --k) {
        if (tau[k] < tllim) {
          float x = fabsf(tau[k] - tllim) / 0.5f;
          col[k] = expf(-(x * x)) * col[k];
        } else if (tau[k] > tulim) {
          /* the reference reads one past the top for k = NZ; left unchanged here */
          if (k + 1 < nz) col[k] = col[k + 1];
        }
      }
    }
    for (int k = 0; k < nz; ++k) col[k] = col[k] + bumod[k];
    jpar += nj;
  }
  free(bumod);
}

/* GET_PERT_1D (get_dmodel.f90:1201-1244) for one pixel:
 *   GET_IT_PENALTIES -> NEW_CALCULATE_CHI2 (:960-1032, with UPDATE_LAMBDA)
 *   -> NORMALIZE_JACOB (:725-770) -> NEW_SOLVE_PERTURBATION (:780-829)
 *   -> LIMIT_PERTURBATION (:884-907).
 * model[(w*13 + col)*nz + k] = MODEL2D_RCV(k+1, col+1, w+1) (w=0 current,
 * w=1 best), updated; sla[9] = INV_SLA(1:9); jac_cur / jac_best are
 * EQVDSYN(NFREQ, IFREEP) column-major; best_syn, jac_best updated; delta
 * (ifreep, optional) receives the applied node perturbation. */
int fzc_lm_step(const fzc_lmcfg* c, const double* coefs, const double* obs, const double* syn,
                double* best_syn, double* sla, const float* jac_cur, float* jac_best, const double* isigmap,
                float* model, double* delta_out) {
  int nz = c->nz, numw = c->numw;
  int nstk = 0;
  for (int s = 0; s < 4; ++s) nstk += c->invstk[s] != 0;
  int nf = nstk * numw, np = ifreep_of(c);
  if (np <= 0 || nf <= 0) return 5;
  float* cur = model;
  float* best = model + (size_t)13 * nz;
  double* jac = (double*)malloc(sizeof(double) * (size_t)nf * np);
  double* r = (double*)malloc(sizeof(double) * nf);
  double* pen_res = (double*)calloc(np, sizeof(double));
  double* pen_hss = (double*)calloc((size_t)np * np, sizeof(double));
  double* delta = (double*)calloc(np, sizeof(double));
  for (size_t i = 0; i < (size_t)nf * np; ++i) jac[i] = (double)jac_cur[i];

  if (c->regul) penalties(c, coefs, cur, pen_res, pen_hss);

  /* NEW_CALCULATE_CHI2 */
  double chi = 0.0;
  for (int f = 0; f < nf; ++f) {
    double d = obs[f] - syn[f];
    chi += d * d * isigmap[f];
  }
  sla[1] = chi / c->inu;
  if (c->regul) {
    double s = 0.0;
    for (int i = 0; i < np; ++i) s += pen_res[i] * 1.0 * 1.0 * pen_res[i];
    sla[1] = sla[1] + s;
  }
  if (sla[1] < sla[2]) {
    memcpy(best, cur, sizeof(float) * 13 * nz);
    for (size_t i = 0; i < (size_t)nf * np; ++i) jac_best[i] = (float)jac[i];
    memcpy(best_syn, syn, sizeof(double) * nf);
  } else {
    for (size_t i = 0; i < (size_t)nf * np; ++i) jac[i] = (double)jac_best[i];
    memcpy(cur, best, sizeof(float) * 13 * nz);
    if (c->regul) penalties(c, coefs, cur, pen_res, pen_hss);
  }
  update_lambda(&sla[1], &sla[2], &sla[8], &sla[0], &sla[3], 0.01);
  for (int f = 0; f < nf; ++f) r[f] = (obs[f] - best_syn[f]) * sqrt(isigmap[f]) / sqrt(c->inu);

  /* NORMALIZE_JACOB */
  for (int j = 0; j < np; ++j)
    for (int f = 0; f < nf; ++f) jac[(size_t)j * nf + f] = jac[(size_t)j * nf + f] * sqrt(isigmap[f]) / sqrt(c->inu);
  int jinit = 0;
  for (int p = 0; p < 8; ++p) {
    if (!c->atmpar[p]) continue;
    for (int j = jinit; j < jinit + c->nodes[p]; ++j)
      for (int f = 0; f < nf; ++f) jac[(size_t)j * nf + f] = jac[(size_t)j * nf + f] * (1.0 * PARAM_NORMS[p]);
    jinit += c->nodes[p];
  }

  /* NEW_SOLVE_PERTURBATION */
  if (!c->regul) {
    memset(pen_res, 0, sizeof(double) * np);
    memset(pen_hss, 0, sizeof(double) * (size_t)np * np);
  }
  int rc = fzc_get_perturbation(np, nf, jac, r, sla[8], delta, pen_res, pen_hss, c->solver, c->svdtol);
  double sum = 0.0;
  for (int i = 0; i < np; ++i) sum += delta[i];
  if (sum != sum)
    for (int i = 0; i < np; ++i)
      if (delta[i] != delta[i]) delta[i] = 0.0;

  /* LIMIT_PERTURBATION -> REARANGE_DA_MODEL on the current model */
  if (rc == 0) rearange(c, coefs, delta, cur);
  if (delta_out) memcpy(delta_out, delta, sizeof(double) * np);
  free(jac);
  free(r);
  free(pen_res);
  free(pen_hss);
  free(delta);
  return rc;
}
