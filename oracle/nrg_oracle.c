/*
 * nrg_oracle.c -- CPU restatement of NRG's chemistry step (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library; the product never does.  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off, glibc libm: the same arithmetic the reference's
 * x86-64 Fortran build performs, with no FMA contraction).
 *
 * Each routine cites the reference lines it restates.  Fortran evaluates
 * a op b op c left to right and never reassociates; every expression below
 * is parenthesised to that order.  Fortran MAX/MIN/SIGN/ABS are written as
 * fmax-free comparisons / copysign / fabs.
 */
#include "nrg_oracle.h"

#include <float.h>
#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define R_GAS_J 8.31446261815324            /* package_library/src/global_data.f90:36 */
#define T_REF 298.15                        /* :38 */
#define P_ATM 101325.0                      /* :39 */
#define T_CEILING 6000.0                    /* :44 (default-real literal, exact) */

static inline double dmax(double a, double b) { return (a >= b) ? a : b; }
static inline double dmin(double a, double b) { return (a <= b) ? a : b; }

/* ======================================================================
 * BLAS-1 / LINPACK
 * ====================================================================== */

/* DNRM2, Lawson's four-phase scaled 2-norm (sl/dnrm2.f90:84-160). */
double or_dnrm2(int n, const double *dx)
{
  const double cutlo = 8.232e-11, cuthi = 1.304e19;
  double sum = 0.0, xmax = 0.0, hitest, r;
  int i, j, next;
  if (n <= 0) return 0.0;
  next = 30;
  i = 0;
L20:
  switch (next) {
    case 30: goto L30;
    case 50: goto L50;
    case 70: goto L70;
    default: goto L110;
  }
L30:
  if (fabs(dx[i]) > cutlo) goto L85;
  next = 50;
  xmax = 0.0;
L50:
  if (dx[i] == 0.0) goto L200;
  if (fabs(dx[i]) > cutlo) goto L85;
  next = 70;
  goto L105;
L100:
  i = j;
  next = 110;
  sum = (sum / dx[i]) / dx[i];
L105:
  xmax = fabs(dx[i]);
  goto L115;
L70:
  if (fabs(dx[i]) > cutlo) goto L75;
L110:
  if (fabs(dx[i]) <= xmax) goto L115;
  r = xmax / dx[i];
  sum = 1.0 + sum * (r * r);
  xmax = fabs(dx[i]);
  goto L200;
L115:
  r = dx[i] / xmax;
  sum = sum + r * r;
  goto L200;
L75:
  sum = (sum * xmax) * xmax;
L85:
  hitest = cuthi / (double)n;
  for (j = i; j < n; j++) {
    if (fabs(dx[j]) >= hitest) goto L100;
    sum = sum + dx[j] * dx[j];
  }
  return sqrt(sum);
L200:
  i = i + 1;
  if (i < n) goto L20;
  return xmax * sqrt(sum);
}

/* DAXPY with unit strides (sl/daxpy.f90:50-83): no-op when da == 0. */
static void daxpy1(int n, double da, const double *dx, double *dy)
{
  if (n <= 0 || da == 0.0) return;
  for (int i = 0; i < n; i++) dy[i] = dy[i] + da * dx[i];
}

/* IDAMAX (sl/idamax.f90:48-81): first index of the largest |x|. */
static int idamax1(int n, const double *dx)
{
  if (n <= 0) return -1;
  int imax = 0;
  double dmx = fabs(dx[0]);
  for (int i = 1; i < n; i++) {
    double xm = fabs(dx[i]);
    if (xm > dmx) { imax = i; dmx = xm; }
  }
  return imax;
}

/* DGEFA (sl/dgefa.f90:70-116). a(i,j) = a[i + lda*j]; ipvt 0-based; returns INFO (1-based). */
int or_dgefa(double *a, int lda, int n, int *ipvt)
{
  int info = 0;
  for (int k = 0; k < n - 1; k++) {
    int l = idamax1(n - k, &a[k + lda * k]) + k;
    ipvt[k] = l;
    if (a[l + lda * k] == 0.0) { info = k + 1; continue; }
    if (l != k) {
      double t = a[l + lda * k];
      a[l + lda * k] = a[k + lda * k];
      a[k + lda * k] = t;
    }
    double t = -1.0 / a[k + lda * k];
    for (int i = k + 1; i < n; i++) a[i + lda * k] = t * a[i + lda * k];   /* DSCAL */
    for (int j = k + 1; j < n; j++) {
      double tj = a[l + lda * j];
      if (l != k) {
        a[l + lda * j] = a[k + lda * j];
        a[k + lda * j] = tj;
      }
      daxpy1(n - k - 1, tj, &a[k + 1 + lda * k], &a[k + 1 + lda * j]);
    }
  }
  ipvt[n - 1] = n - 1;
  if (a[(n - 1) + lda * (n - 1)] == 0.0) info = n;
  return info;
}

/* DGESL, JOB = 0 (sl/dgesl.f90:78-103). */
void or_dgesl(const double *a, int lda, int n, const int *ipvt, double *b)
{
  for (int k = 0; k < n - 1; k++) {
    int l = ipvt[k];
    double t = b[l];
    if (l != k) { b[l] = b[k]; b[k] = t; }
    daxpy1(n - k - 1, t, &a[k + 1 + lda * k], &b[k + 1]);
  }
  for (int kb = 0; kb < n; kb++) {
    int k = n - 1 - kb;
    b[k] = b[k] / a[k + lda * k];
    double t = -b[k];
    daxpy1(k, t, &a[lda * k], &b[0]);
  }
}

/* DDCST, MINT = 2 (sl/ddcst.f90:39-86).  el: EL(13,12) column-major, tq: TQ(3,12). */
void or_ddcst(int maxord, double *el, double *tq)
{
#define EL(i, j) el[((i) - 1) + 13 * ((j) - 1)]
#define TQ(i, j) tq[((i) - 1) + 3 * ((j) - 1)]
  double factrl[13];
  factrl[1] = 1.0;
  for (int i = 2; i <= maxord; i++) factrl[i] = (double)i * factrl[i - 1];
  EL(1, 1) = 1.0;
  EL(2, 1) = 1.0;
  for (int j = 2; j <= maxord; j++) {
    EL(1, j) = factrl[j];
    for (int i = 2; i <= j; i++) EL(i, j) = (double)j * EL(i, j - 1) + EL(i - 1, j - 1);
    EL(j + 1, j) = 1.0;
  }
  double sum = 1.0;
  for (int j = 2; j <= maxord; j++) {
    sum = sum + 1.0 / (double)j;
    for (int i = 1; i <= j + 1; i++) EL(i, j) = EL(i, j) / (factrl[j] * sum);
  }
  for (int j = 1; j <= maxord; j++) {
    if (j > 1) TQ(1, j) = 1.0 / factrl[j - 1];
    TQ(2, j) = (double)(j + 1) / EL(1, j);
    TQ(3, j) = (double)(j + 2) / EL(1, j);
  }
#undef EL
#undef TQ
}

/* ======================================================================
 * SDRIVE Gear-BDF integrator, MITER = 2, IMPL = 0, IERROR = 3
 * ====================================================================== */

enum { MXTRY = 50, MXFAIL = 3, MXITER = 3, NDJSTP = 10 };
static const double BIAS1 = 1.3, BIAS2 = 1.2, BIAS3 = 1.4, RCTEST = 0.3, RMFAIL = 2.0,
                    RMNORM = 10.0, TRSHLD = 1.0, RMINIT = 10000.0, FACMAX = 0.5, BU = 0.5;

typedef struct bdf_state {
  int n, maxord;
  double eps, ewt, hmax, uround, sqrtn;
  double t, h, hold, rc, rmax, trend, hused, avgh, avgord;
  int nq, nwait, jtask, jstepl, nstep, nfe, nje, ncorr, natt, nqused, convrg, ier;
  double yh[6][OR_NMAX];            /* yh[j][i] = YH(i, j+1) */
  double y[OR_NMAX], ywt[OR_NMAX], save1[OR_NMAX], save2[OR_NMAX], fac[OR_NMAX];
  double dfdy[OR_NMAX * OR_NMAX];   /* column-major, leading dimension n */
  int ipvt[OR_NMAX];
  double el[13 * 12], tq[3 * 12];
  double br, bl, bp, facmin;        /* DDPST constants, u**.875 etc. */
  or_rhs_fn f;
  void *ctx;
} bdf_state;

#define ELv(s, i, j) ((s)->el[((i) - 1) + 13 * ((j) - 1)])
#define TQv(s, i, j) ((s)->tq[((i) - 1) + 3 * ((j) - 1)])

static void call_f(bdf_state *s, const double *y, double *out)
{
  s->f(s->ctx, s->n, s->t, y, out);
}

/* DDPSC (sl/ddpsc.f90:26-38): Pascal-triangle prediction (+1) or retraction (-1). */
static void ddpsc(bdf_state *s, int ksgn)
{
  int n = s->n, nq = s->nq;
  for (int j1 = 1; j1 <= nq; j1++)
    for (int j2 = j1; j2 <= nq; j2++) {
      int j = nq - j2 + j1;          /* 1-based column */
      double *a = s->yh[j - 1], *b = s->yh[j];
      if (ksgn > 0)
        for (int i = 0; i < n; i++) a[i] = a[i] + b[i];
      else
        for (int i = 0; i < n; i++) a[i] = a[i] - b[i];
    }
}

/* DDSCL (sl/ddscl.f90:24-35): rescale the Nordsieck history by RH. */
static void ddscl(bdf_state *s, double rh)
{
  double h = s->h, ah = fabs(h);
  if (h < 1.0)
    rh = dmin(dmin(ah * rh, ah * s->rmax), s->hmax) / ah;
  else
    rh = dmin(dmin(rh, s->rmax), s->hmax / ah);
  double r1 = 1.0;
  for (int j = 1; j <= s->nq; j++) {
    r1 = r1 * rh;
    for (int i = 0; i < s->n; i++) s->yh[j][i] = s->yh[j][i] * r1;
  }
  s->h = s->h * rh;
  s->rc = s->rc * rh;
}

/* DDNTL for JTASK >= 0, IMPL = 0, MITER = 2 (sl/ddntl.f90:52-63,152-161). */
static void ddntl(bdf_state *s)
{
  int n = s->n;
  s->ier = 0;
  if (s->jtask == 0) {
    or_ddcst(s->maxord, s->el, s->tq);
    s->rmax = RMINIT;
  }
  s->rc = 0.0;
  s->convrg = 0;
  s->trend = 1.0;
  s->nq = 1;
  s->nwait = 3;
  call_f(s, s->y, s->save2);
  s->nfe++;
  for (int i = 0; i < n; i++) s->save1[i] = s->save2[i] / dmax(1.0, s->ywt[i]);
  double sum = or_dnrm2(n, s->save1) / s->sqrtn;
  if (sum > s->eps / fabs(s->h)) s->h = copysign(s->eps / sum, s->h);
  for (int i = 0; i < n; i++) s->yh[1][i] = s->h * s->save2[i];
  for (int i = 0; i < n; i++) s->fac[i] = sqrt(s->uround);
}

/* DDPST, MITER = 2, IMPL = 0 (sl/ddpst.f90:39-40,53-114,142-143):
 * finite-difference Jacobian with adaptive increments, P = I - h*l0*J, LU. */
static void ddpst(bdf_state *s)
{
  int n = s->n;
  double *save1 = s->save1, *save2 = s->save2, *y = s->y, *fac = s->fac;
  s->nje++;
  s->ier = 0;
  for (int j = 0; j < n; j++) {
    double ys = dmax(fabs(s->ywt[j]), fabs(y[j]));
    double dy;
    for (;;) {
      dy = fac[j] * ys;
      if (dy == 0.0) {
        if (fac[j] < FACMAX) {
          fac[j] = dmin(100.0 * fac[j], FACMAX);
          continue;
        }
        dy = ys;
      }
      break;
    }
    if (s->nq == 1)
      dy = copysign(dy, save2[j]);
    else
      dy = copysign(dy, s->yh[2][j]);
    dy = (y[j] + dy) - y[j];
    double yj = y[j];
    y[j] = y[j] + dy;
    call_f(s, y, save1);
    y[j] = yj;
    double factor = (-ELv(s, 1, s->nq) * s->h) / dy;
    for (int i = 0; i < n; i++) s->dfdy[i + n * j] = (save1[i] - save2[i]) * factor;
    double diff = fabs(save2[0] - save1[0]);
    int imax = 0;
    for (int i = 1; i < n; i++) {
      if (fabs(save2[i] - save1[i]) > diff) {
        imax = i;
        diff = fabs(save2[i] - save1[i]);
      }
    }
    if (dmin(fabs(save2[imax]), fabs(save1[imax])) > 0.0) {
      double scale = dmax(fabs(save2[imax]), fabs(save1[imax]));
      if (diff > BU * scale)
        fac[j] = dmax(s->facmin, fac[j] * 0.5);
      else if (s->br * scale <= diff && diff <= s->bl * scale)
        fac[j] = dmin(fac[j] * 2.0, FACMAX);
      else if (diff < s->br * scale)
        fac[j] = dmin(s->bp * fac[j], FACMAX);
    }
  }
  s->nfe += n;
  for (int i = 0; i < n; i++) s->dfdy[i + n * i] = s->dfdy[i + n * i] + 1.0;
  if (or_dgefa(s->dfdy, n, n, s->ipvt) != 0) s->ier = 1;
}

/* DDCOR, MITER = 2, IMPL = 0, IERROR = 3 (sl/ddcor.f90:51-54,98-108). */
static double ddcor(bdf_state *s)
{
  int n = s->n;
  for (int i = 0; i < n; i++)
    s->save2[i] = ((s->h * s->save2[i]) - s->yh[1][i]) - s->save1[i];
  or_dgesl(s->dfdy, n, n, s->ipvt, s->save2);
  for (int i = 0; i < n; i++) {
    s->save1[i] = s->save1[i] + s->save2[i];
    s->save2[i] = s->save2[i] / dmax(fabs(s->y[i]), s->ywt[i]);
  }
  s->ncorr++;
  return or_dnrm2(n, s->save2) / s->sqrtn;
}

/* DDSTP, MINT = 2, ISWFLG = 2 (no method switching), IERROR = 3
 * (sl/ddstp.f90:91-437).  Returns JSTATE. */
static int ddstp(bdf_state *s)
{
  int n = s->n, ntry = 0, nfail = 0, iter = 0, evaljc;
  double told = s->t, d = 0.0, d1 = 0.0, ctest = 0.0, etest, rh;

  if (s->jtask <= 0) {
    ddntl(s);
    if (s->h == 0.0) goto L400;
    if (s->ier) goto L420;
  }
L100:
  ntry++;
  if (ntry > MXTRY) goto L410;
  s->natt++;
  s->t = s->t + s->h;
  ddpsc(s, 1);
  evaljc = (fabs(s->rc - 1.0) > RCTEST) || (s->nstep >= s->jstepl + NDJSTP);
L110:
  iter = 0;
  for (int i = 0; i < n; i++) s->y[i] = s->yh[0][i];
  call_f(s, s->y, s->save2);
  s->nfe++;
  if (evaljc || s->ier) {
    ddpst(s);
    if (s->ier) goto L160;
    s->convrg = 0;
    s->rc = 1.0;
    s->jstepl = s->nstep;
  }
  for (int i = 0; i < n; i++) s->save1[i] = 0.0;
L130:
  d = ddcor(s);
  if (iter > 0) s->trend = dmax(0.9 * s->trend, d / d1);
  d1 = d;
  ctest = dmin(2.0 * s->trend, 1.0) * d;
  if (ctest <= s->eps) goto L170;
  iter++;
  if (iter < MXITER) {
    double l0 = ELv(s, 1, s->nq);
    for (int i = 0; i < n; i++) s->y[i] = s->yh[0][i] + l0 * s->save1[i];
    call_f(s, s->y, s->save2);
    s->nfe++;
    goto L130;
  }
  if (s->convrg) {
    evaljc = 1;
    goto L110;
  }
L160:
  s->t = told;
  ddpsc(s, -1);
  s->nwait = s->nq + 2;
  if (s->jtask != 0 && s->jtask != 2) s->rmax = RMFAIL;
  if (iter == 0)
    rh = 0.3;
  else
    rh = 0.9 * pow(s->eps / ctest, 0.2);
  if (rh * s->h == 0.0) goto L400;
  ddscl(s, rh);
  goto L100;

L170:
  s->convrg = 1;
  for (int i = 0; i < n; i++) s->save2[i] = s->save1[i] / dmax(fabs(s->y[i]), s->ywt[i]);
  etest = or_dnrm2(n, s->save2) / (TQv(s, 2, s->nq) * s->sqrtn);
  if (etest > s->eps) {
    s->t = told;
    ddpsc(s, -1);
    nfail++;
    if (nfail < MXFAIL || s->nq == 1) {
      if (s->jtask != 0 && s->jtask != 2) s->rmax = RMFAIL;
      double rh2 = 1.0 / (BIAS2 * pow(etest / s->eps, 1.0 / (double)(s->nq + 1)));
      if (s->nq > 1) {
        for (int i = 0; i < n; i++)
          s->save2[i] = s->yh[s->nq][i] / dmax(fabs(s->y[i]), s->ywt[i]);
        double erdn = or_dnrm2(n, s->save2) / (TQv(s, 1, s->nq) * s->sqrtn);
        double rh1 = 1.0 / dmax(1.0, BIAS1 * pow(erdn / s->eps, 1.0 / (double)s->nq));
        if (rh2 < rh1) {
          s->nq = s->nq - 1;
          s->rc = (s->rc * ELv(s, 1, s->nq)) / ELv(s, 1, s->nq + 1);
          rh = rh1;
        } else {
          rh = rh2;
        }
      } else {
        rh = rh2;
      }
      s->nwait = s->nq + 2;
      if (rh * s->h == 0.0) goto L400;
      ddscl(s, rh);
      goto L100;
    }
    nfail = 0;
    s->jtask = 2;
    for (int i = 0; i < n; i++) s->y[i] = s->yh[0][i];
    ddntl(s);
    s->rmax = RMNORM;
    if (s->h == 0.0) goto L400;
    if (s->ier) goto L420;
    goto L100;
  }

  /* accepted step (sl/ddstp.f90:291-300) */
  s->nstep++;
  s->hused = s->h;
  s->nqused = s->nq;
  s->avgh = ((double)(s->nstep - 1) * s->avgh + s->h) / (double)s->nstep;
  s->avgord = ((double)(s->nstep - 1) * s->avgord + (double)s->nq) / (double)s->nstep;
  for (int j = 1; j <= s->nq + 1; j++) {
    double e = ELv(s, j, s->nq);
    for (int i = 0; i < n; i++) s->yh[j - 1][i] = s->yh[j - 1][i] + e * s->save1[i];
  }
  for (int i = 0; i < n; i++) s->y[i] = s->yh[0][i];

  /* step size / order selection (sl/ddstp.f90:360-432) */
  if (s->jtask == 0 || s->jtask == 2) {
    rh = 1.0 / dmax(s->uround, BIAS2 * pow(etest / s->eps, 1.0 / (double)(s->nq + 1)));
    if (rh > TRSHLD) ddscl(s, rh);
  } else if (s->nwait > 1) {
    s->nwait = s->nwait - 1;
    if (s->nwait == 1 && s->nq < s->maxord)
      for (int i = 0; i < n; i++) s->yh[s->maxord][i] = s->save1[i];
  } else {
    double rh1, rh2, rh3;
    if (s->nq == 1) {
      rh1 = 0.0;
    } else {
      for (int i = 0; i < n; i++)
        s->save2[i] = s->yh[s->nq][i] / dmax(fabs(s->y[i]), s->ywt[i]);
      double erdn = or_dnrm2(n, s->save2) / (TQv(s, 1, s->nq) * s->sqrtn);
      rh1 = 1.0 / dmax(s->uround, BIAS1 * pow(erdn / s->eps, 1.0 / (double)s->nq));
    }
    rh2 = 1.0 / dmax(s->uround, BIAS2 * pow(etest / s->eps, 1.0 / (double)(s->nq + 1)));
    if (s->nq == s->maxord) {
      rh3 = 0.0;
    } else {
      for (int i = 0; i < n; i++)
        s->save2[i] = (s->save1[i] - s->yh[s->maxord][i]) / dmax(fabs(s->y[i]), s->ywt[i]);
      double erup = or_dnrm2(n, s->save2) / (TQv(s, 3, s->nq) * s->sqrtn);
      rh3 = 1.0 / dmax(s->uround, BIAS3 * pow(erup / s->eps, 1.0 / (double)(s->nq + 2)));
    }
    if (rh1 > rh2 && rh1 >= rh3) {
      rh = rh1;
      if (rh <= TRSHLD) goto L380;
      s->nq = s->nq - 1;
      s->rc = (s->rc * ELv(s, 1, s->nq)) / ELv(s, 1, s->nq + 1);
    } else if (rh2 >= rh1 && rh2 >= rh3) {
      rh = rh2;
      if (rh <= TRSHLD) goto L380;
    } else {
      rh = rh3;
      if (rh <= TRSHLD) goto L380;
      double e = ELv(s, s->nq + 1, s->nq);
      for (int i = 0; i < n; i++) s->yh[s->nq + 1][i] = (s->save1[i] * e) / (double)(s->nq + 1);
      s->nq = s->nq + 1;
      s->rc = (s->rc * ELv(s, 1, s->nq)) / ELv(s, 1, s->nq - 1);
    }
    ddscl(s, rh);
    s->rmax = RMNORM;
  L380:
    s->nwait = s->nq + 2;
  }
  s->hold = s->h;
  return 1;

L400:
  s->hold = s->h;
  for (int i = 0; i < n; i++) s->y[i] = s->yh[0][i];
  return 2;
L410:
  s->hold = s->h;
  return 3;
L420:
  s->hold = s->h;
  return 4;
}

/* DDNTP with K = 0 (sl/ddntp.f90:25-32): interpolate the Nordsieck history at tout. */
static void ddntp(const bdf_state *s, double tout, double *y)
{
  int n = s->n, nq = s->nq;
  for (int i = 0; i < n; i++) y[i] = s->yh[nq][i];
  double r = (tout - s->t) / s->h;
  for (int jj = 1; jj <= nq; jj++) {
    int j = nq + 1 - jj;                           /* 1-based column */
    for (int i = 0; i < n; i++) y[i] = s->yh[j - 1][i] + r * y[i];
  }
}

/* DDRIV3 (sl/ddriv3.f90): NSTATE = 1 entry (:930-982), the NTASK = 1 loop
 * 260 -> 410 -> DDSTP -> 470 (:1292-1449) and the return paths (:560-580). */
void or_ddriv3(int n, double *y, double tout, const or_bdf_opts *o,
               or_rhs_fn f, void *ctx, or_bdf_out *out)
{
  bdf_state s_local;
  bdf_state *s = &s_local;
  memset(s, 0, sizeof(*s));
  s->n = n;
  s->f = f;
  s->ctx = ctx;
  s->maxord = o->mxord < 5 ? o->mxord : 5;
  s->eps = o->eps;
  s->ewt = o->ewt;
  s->hmax = o->hmax;
  s->uround = 2.220446049250313080847e-16;      /* D1MACH(4) = epsilon(1d0) (sl/d1mach.f90:75-77) */
  s->sqrtn = sqrt((double)n);
  s->br = pow(s->uround, 0.875);
  s->bl = pow(s->uround, 0.75);
  s->bp = pow(s->uround, -0.15);
  s->facmin = pow(s->uround, 0.78);
  s->t = 0.0;
  double h = (tout - s->t) * (1.0 - 4.0 * s->uround);
  h = copysign(dmin(fabs(h), s->hmax), h);
  s->h = h;
  double hsign = copysign(1.0, h);
  s->jtask = 0;
  for (int i = 0; i < n; i++) s->yh[0][i] = y[i];
  out->ierflg = 0;
  out->nstate = 2;
  if (s->t == tout) goto done;
  int nstepl = s->nstep;
  for (int i = 0; i < n; i++) y[i] = s->yh[0][i];
  memcpy(s->y, y, sizeof(double) * n);
  for (;;) {
    /* 260: IERROR = 3 weights */
    for (int i = 0; i < n; i++) s->ywt[i] = dmax(s->ewt, fabs(s->y[i]));
    /* 410 */
    for (int i = 0; i < n; i++) s->save2[i] = s->y[i] / s->ywt[i];
    double sum = or_dnrm2(n, s->save2) / s->sqrtn;
    sum = dmax(1.0, sum);
    if (s->eps < sum * s->uround) {
      s->eps = sum * s->uround * (1.0 + 10.0 * s->uround);
      out->ierflg = 4;
      out->nstate = 4;
      for (int i = 0; i < n; i++) y[i] = s->yh[0][i];
      goto done;
    }
    if (s->nstep - nstepl == o->mxstep) {
      out->ierflg = 3;
      out->nstate = 3;
      for (int i = 0; i < n; i++) y[i] = s->yh[0][i];
      goto done;
    }
    int jstate = ddstp(s);
    if (jstate != 1) {
      out->ierflg = (jstate == 2) ? 41 : (jstate == 3) ? 42 : 43;
      out->nstate = 12;
      for (int i = 0; i < n; i++) y[i] = s->y[i];
      goto done;
    }
    s->jtask = 1;
    if (s->t * hsign < tout * hsign) continue;      /* GO TO 260 */
    ddntp(s, tout, y);
    out->nstate = 2;
    goto done;
  }
done:
  out->nstep = s->nstep;
  out->nfe = s->nfe;
  out->nje = s->nje;
  out->ncorr = s->ncorr;
  out->natt = s->natt;
  out->nqused = s->nqused;
  out->hused = s->hused;
  out->t = s->t;
}

/* ======================================================================
 * Thermochemistry (package_library/src/thermophysical_properties_class.f90)
 * ====================================================================== */

static inline const double *acoef(const nrg_chem_mech *m, int s, double temp)
{
  int set = (temp >= 1000.0) ? 1 : 0;            /* :236, :255, :277 */
  return m->a_coeffs + 14 * s + 7 * set;
}

static inline double clip_t(double T) { return dmin(T, T_CEILING); }   /* :876-881 */

double or_cp_molar(const nrg_chem_mech *m, double T, int s)
{
  double temp = clip_t(T);
  const double *a = acoef(m, s, temp);
  double cp = a[4];
  for (int i = 4; i >= 1; i--) cp = cp * temp + a[i - 1];
  return cp;
}

double or_h_molar(const nrg_chem_mech *m, double T, int s)
{
  double temp = clip_t(T);
  const double *a = acoef(m, s, temp);
  double hs = a[4] / 5.0;
  for (int i = 4; i >= 1; i--) hs = hs * temp + a[i - 1] / (double)i;
  hs = hs + a[5] / temp;
  hs = hs * temp;
  return hs;
}

double or_s_molar(const nrg_chem_mech *m, double T, int s)
{
  double temp = clip_t(T);
  const double *a = acoef(m, s, temp);
  double sm = a[4] / 4.0;
  for (int i = 4; i >= 2; i--) sm = sm * temp + a[i - 1] / (double)(i - 1);
  sm = (sm * temp + a[6]) + a[0] * log(temp);
  return sm;
}

double or_e_molar(const nrg_chem_mech *m, double T, int s)
{
  return or_h_molar(m, T, s) - R_GAS_J * T;       /* unclipped T (:298) */
}

double or_mixture_cp_molar(const nrg_chem_mech *m, double T, const double *X)
{
  double cp = 0.0;
  for (int s = 0; s < m->species_number; s++) cp = cp + or_cp_molar(m, T, s) * X[s];
  return cp;
}

double or_mixture_h_molar(const nrg_chem_mech *m, double T, const double *X)
{
  double h = 0.0;
  for (int s = 0; s < m->species_number; s++) h = h + or_h_molar(m, T, s) * X[s];
  return h;
}

double or_mixture_e_molar(const nrg_chem_mech *m, double T, const double *X)
{
  double e = 0.0;
  for (int s = 0; s < m->species_number; s++) e = e + or_e_molar(m, T, s) * X[s];
  return e;
}

/* calculate_temperature: frozen-cp Newton on e (:409-452).  eps = 1e-08 is a
 * default-real literal: (double)1e-8f. */
double or_calculate_temperature(const nrg_chem_mech *m, double T, double e_i, const double *X)
{
  const double eps = (double)1e-08f;
  double temp = clip_t(T);
  double h_const = or_mixture_h_molar(m, T_REF, X);
  double cp = or_mixture_cp_molar(m, temp, X);
  for (int iter = 1; iter <= 1000; iter++) {
    double e_int = or_mixture_e_molar(m, temp, X) - h_const;
    double residual = (e_i - e_int) / cp;
    temp = temp + residual;
    if (fabs(residual) < eps) break;
  }
  return temp;
}

/* ======================================================================
 * Kinetics (computing_module/src/current_build/chemical_kinetics_solver.f90)
 * ====================================================================== */

#define CC(m, c, r, side) ((m)->chemical_coeffs[(c) + 4 * (r) + 4 * (m)->reactions_number * (side)])
#define TROE(m, r, k) ((m)->Troe_coeffs[(r) + (m)->reactions_number * (k)])
#define EFF(m, r, s) ((m)->enhanced_efficiencies[(r) + (m)->reactions_number * (s)])

void or_mixture_concentration(const nrg_chem_mech *m, const double *Y, double rho, double *Mc)
{
  int S = m->species_number, R = m->reactions_number;
  for (int r = 0; r < R; r++) {
    Mc[r] = 0.0;
    for (int s = 0; s < S; s++)
      if (m->molar_masses[s] != 0.0)
        Mc[r] = Mc[r] + ((Y[s] * rho) / m->molar_masses[s]) * EFF(m, r, s);
  }
}

void or_rate_constants(const nrg_chem_mech *m, double Tin, const double *Mc, double *kf, double *kr)
{
  int S = m->species_number, R = m->reactions_number;
  double sv[OR_NMAX], hv[OR_NMAX];
  double T = dmin(Tin, 10000.0);                                   /* :382 */
  for (int c = 0; c < S; c++) {
    sv[c] = or_s_molar(m, T, c);
    hv[c] = or_h_molar(m, T, c);
  }
  for (int r = 0; r < R; r++) {
    double d_s = 0.0, d_h = 0.0, d_nu;
    int nl = CC(m, 0, r, 0), nr = CC(m, 0, r, 1);
    for (int c = 1; c <= nr; c++) {
      int sp = CC(m, c, r, 1);
      if (sp == S + 1) continue;
      d_s = d_s + sv[sp - 1];
      d_h = d_h + hv[sp - 1];
    }
    for (int c = 1; c <= nl; c++) {
      int sp = CC(m, c, r, 0);
      if (sp == S + 1) continue;
      d_s = d_s - sv[sp - 1];
      d_h = d_h - hv[sp - 1];
    }
    d_nu = (double)(nr - nl);                                        /* :410 */
    double Kp, Kc, k0, kinf, pr;
    switch (m->reactions_type[r]) {
      case 0:
      case 1:
        kf[r] = (m->A[r] * pow(T, m->beta[r])) * exp(((-m->E_act[r]) / R_GAS_J) / T);
        Kp = exp(d_s / R_GAS_J - (d_h / R_GAS_J) / T);
        Kc = Kp * pow((P_ATM / R_GAS_J) / T, d_nu);
        kr[r] = kf[r] / Kc;
        break;
      case 2:
        k0 = (m->A_low[r] * pow(T, m->beta_low[r])) * exp(((-m->E_act_low[r]) / R_GAS_J) / T);
        kinf = (m->A[r] * pow(T, m->beta[r])) * exp(((-m->E_act[r]) / R_GAS_J) / T);
        pr = (k0 * Mc[r]) / kinf;
        kf[r] = kinf * (pr / (1.0 + pr));
        Kp = exp(d_s / R_GAS_J - (d_h / R_GAS_J) / T);
        Kc = Kp * pow((P_ATM / R_GAS_J) / T, d_nu);
        kr[r] = kf[r] / Kc;
        break;
      case 3: {
        k0 = exp(((-m->E_act_low[r]) / R_GAS_J) / T);
        kinf = exp(((-m->E_act[r]) / R_GAS_J) / T);
        k0 = (m->A_low[r] * pow(T, m->beta_low[r])) * k0;
        kinf = (m->A[r] * pow(T, m->beta[r])) * kinf;
        pr = (k0 * Mc[r]) / kinf;
        double a = TROE(m, r, 0);
        double fcent = ((1.0 - a) * exp(-(T / TROE(m, r, 1))) + a * exp(-(T / TROE(m, r, 2))))
                       + exp(-(TROE(m, r, 3) / T));
        double c = -0.4 - 0.67 * log10(fcent);
        double nn = 0.75 - 1.27 * log10(fcent);
        double d = 0.14;
        double blend;
        if (pr != 0.0) {
          double x = (log10(pr) + c) / (nn - d * (log10(pr) + c));
          blend = pow(fcent, 1.0 / (1.0 + x * x));
        } else {
          blend = pow(fcent, -1.0 / d);
        }
        kf[r] = (kinf * (pr / (1.0 + pr))) * blend;
        Kp = exp(d_s / R_GAS_J - (d_h / R_GAS_J) / T);
        Kc = Kp * pow((P_ATM / R_GAS_J) / T, d_nu);
        kr[r] = kf[r] / Kc;
        break;
      }
      case 5:
      case 6:
        kf[r] = (m->A[r] * pow(T, m->beta[r])) * exp(((-m->E_act[r]) / R_GAS_J) / T);
        kr[r] = 0.0;
        break;
      case 7:
        k0 = (m->A_low[r] * pow(T, m->beta_low[r])) * exp(((-m->E_act_low[r]) / R_GAS_J) / T);
        kinf = (m->A[r] * pow(T, m->beta[r])) * exp(((-m->E_act[r]) / R_GAS_J) / T);
        pr = (k0 * Mc[r]) / kinf;
        kf[r] = kinf * (pr / (1.0 + pr));
        kr[r] = 0.0;
        break;
      default:                                  /* type 4 never produced by the parser */
        break;
    }
  }
}

void or_rhs(const nrg_chem_mech *m, const double *kf, const double *kr, const double *Mc,
            const double *Z, double *Zdot)
{
  int S = m->species_number, R = m->reactions_number;
  for (int s = 0; s < S; s++) Zdot[s] = 0.0;
  for (int r = 0; r < R; r++) {
    int nl = CC(m, 0, r, 0), nr = CC(m, 0, r, 1);
    double fwd = kf[r], rev = kr[r];
    for (int c = 1; c <= nl; c++) {
      int sp = CC(m, c, r, 0);
      fwd = fwd * ((sp == S + 1) ? Mc[r] : Z[sp - 1]);
    }
    for (int c = 1; c <= nr; c++) {
      int sp = CC(m, c, r, 1);
      rev = rev * ((sp == S + 1) ? Mc[r] : Z[sp - 1]);
    }
    for (int s = 1; s <= S; s++) {
      for (int c = 1; c <= nl; c++)
        if (CC(m, c, r, 0) == s) Zdot[s - 1] = (Zdot[s - 1] - fwd) + rev;
      for (int c = 1; c <= nr; c++)
        if (CC(m, c, r, 1) == s) Zdot[s - 1] = (Zdot[s - 1] - rev) + fwd;
    }
  }
}

typedef struct chem_ctx {
  const nrg_chem_mech *m;
  double kf[OR_NMAX * 8], kr[OR_NMAX * 8], Mc[OR_NMAX * 8];
} chem_ctx;

static void chem_fun(void *ctx, int n, double t, const double *y, double *ydot)
{
  (void)n;
  (void)t;
  chem_ctx *c = (chem_ctx *)ctx;
  or_rhs(c->m, c->kf, c->kr, c->Mc, y, ydot);
}

size_t or_chem_ctx_size(void) { return sizeof(chem_ctx); }

/* Per-cell setup of solve_chemical_kinetics (:235-265) into an opaque context
 * usable with or_chem_fun as the DDRIV3 right-hand side; cin receives C_in. */
void or_chem_ctx_init(void *ctx, const nrg_chem_mech *m, double T, double rho, const double *Y,
                      double *cin)
{
  chem_ctx *c = (chem_ctx *)ctx;
  c->m = m;
  or_mixture_concentration(m, Y, rho, c->Mc);
  or_rate_constants(m, T, c->Mc, c->kf, c->kr);
  for (int s = 0; s < m->species_number; s++)
    cin[s] = (m->molar_masses[s] != 0.0) ? (Y[s] * rho) / m->molar_masses[s] : Y[s];
}

void or_chem_fun(void *ctx, int n, double t, const double *y, double *ydot)
{
  chem_fun(ctx, n, t, y, ydot);
}

void or_default_opts(nrg_chem_opts *o)
{
  memset(o, 0, sizeof(*o));
  o->eps = 1.0e-9;
  o->ewt = 1.0e-4;
  o->hmax = (double)1.0e-7f;
  o->t_gate = 305.0;
  o->mxstep = 10000;
  o->mxord = 5;
  o->jacobian = NRG_JAC_FD_SLATEC;
}

int or_chem_cell(const nrg_chem_mech *m, const nrg_chem_opts *o, double dt,
                 double T, double rho, const double *Y, int8_t bc,
                 double *Y_prod, double *Y_prod2, double *E_f_prod, or_cell_out *out)
{
  int S = m->species_number;
  memset(out, 0, sizeof(*out));
  if (!(T >= o->t_gate && bc == 0)) {                              /* :228-231, :341-346 */
    for (int s = 0; s < S; s++) Y_prod[s] = 0.0;
    *E_f_prod = 0.0;
    out->status = NRG_CELL_GATED;
    return NRG_CELL_GATED;
  }
  chem_ctx ctx;
  ctx.m = m;
  double ef = 0.0;
  or_mixture_concentration(m, Y, rho, ctx.Mc);
  or_rate_constants(m, T, ctx.Mc, ctx.kf, ctx.kr);
  double cin[OR_NMAX], cout[OR_NMAX];
  for (int s = 0; s < S; s++) {
    if (m->molar_masses[s] != 0.0)
      cin[s] = (Y[s] * rho) / m->molar_masses[s];
    else
      cin[s] = Y[s];
    cout[s] = cin[s];
  }
  or_bdf_opts bo = {o->eps, o->ewt, o->hmax, o->mxstep, o->mxord};
  or_ddriv3(S, cout, dt, &bo, chem_fun, &ctx, &out->bdf);
  for (int s = 0; s < S; s++) {
    if (m->molar_masses[s] != 0.0) {
      Y_prod[s] = ((cout[s] - cin[s]) * m->molar_masses[s]) / dt;
      Y_prod2[s] = cout[s] - cin[s];
    } else {
      Y_prod[s] = cout[s] - cin[s];
    }
  }
  for (int s = 0; s < S; s++) {
    double hs = or_h_molar(m, T_REF, s);
    ef = ef - (hs * (cout[s] - cin[s])) / dt;
  }
  *E_f_prod = ef;
  if (out->bdf.nstate == 2)
    out->status = NRG_CELL_OK;
  else if (out->bdf.nstate == 3)
    out->status = NRG_CELL_MXSTEP;
  else if (out->bdf.ierflg == 41)
    out->status = NRG_CELL_STEP_ZERO;
  else if (out->bdf.ierflg == 42)
    out->status = NRG_CELL_REPEATED_FAIL;
  else
    out->status = NRG_CELL_SINGULAR;
  return out->status;
}

static int status_to_rc(int st)
{
  if (st == NRG_CELL_STEP_ZERO) return NRG_CHEM_E_STEP_ZERO;
  if (st == NRG_CELL_REPEATED_FAIL || st == NRG_CELL_SINGULAR) return NRG_CHEM_E_REPEATED_FAIL;
  return NRG_CHEM_OK;
}

int or_chem_step(const nrg_chem_mech *m, const nrg_chem_opts *o, const nrg_grid *g, double dt,
                 const double *T, const double *rho, const double *const *Y, const int8_t *bc,
                 double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                 int32_t *cell_status, nrg_chem_stats *stats, int nthreads)
{
  int S = m->species_number;
  int64_t nx = g->hi[0] - g->lo[0] + 1, ny = g->hi[1] - g->lo[1] + 1;
  int64_t ix = g->inner_hi[0] - g->inner_lo[0] + 1, iy = g->inner_hi[1] - g->inner_lo[1] + 1,
          iz = g->inner_hi[2] - g->inner_lo[2] + 1;
  int64_t ninner = ix * iy * iz;
  int rc = NRG_CHEM_OK;
  int64_t integ = 0, nstep_sum = 0, nfe_sum = 0, nje_sum = 0, ncorr_sum = 0, natt_sum = 0;
  int nstep_max = 0, nfe_max = 0;
  int64_t cnt[8] = {0};
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) \
    reduction(+ : integ, nstep_sum, nfe_sum, nje_sum, ncorr_sum, natt_sum) reduction(max : nstep_max, nfe_max, rc)
#endif
  for (int64_t q = 0; q < ninner; q++) {
    int64_t i = q % ix + g->inner_lo[0], j = (q / ix) % iy + g->inner_lo[1], k = q / (ix * iy) + g->inner_lo[2];
    int64_t off = (i - g->lo[0]) + nx * ((j - g->lo[1]) + ny * (k - g->lo[2]));
    double yv[OR_NMAX], yp[OR_NMAX], yp2[OR_NMAX], ef;
    for (int s = 0; s < S; s++) {
      yv[s] = Y[s][off];
      yp2[s] = Y_prod2[s][off];
    }
    or_cell_out co;
    int st = or_chem_cell(m, o, dt, T[off], rho[off], yv, bc ? bc[off] : 0, yp, yp2, &ef, &co);
    for (int s = 0; s < S; s++) {
      Y_prod[s][off] = yp[s];
      Y_prod2[s][off] = yp2[s];
    }
    E_f_prod[off] = ef;
    if (cell_status) cell_status[off] = st;
    if (st != NRG_CELL_GATED) {
      integ++;
      nstep_sum += co.bdf.nstep;
      nfe_sum += co.bdf.nfe;
      nje_sum += co.bdf.nje;
      ncorr_sum += co.bdf.ncorr;
      natt_sum += co.bdf.natt;
      if (co.bdf.nstep > nstep_max) nstep_max = co.bdf.nstep;
      if (co.bdf.nfe > nfe_max) nfe_max = co.bdf.nfe;
    }
    int r2 = status_to_rc(st);
    if (r2 > rc) rc = r2;
#ifdef _OPENMP
#pragma omp atomic
#endif
    cnt[st]++;
  }
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->cells_inner = ninner;
    stats->cells_integrated = integ;
    stats->nstep_sum = nstep_sum;
    stats->nfe_sum = nfe_sum;
    stats->nje_sum = nje_sum;
    stats->ncorr_sum = ncorr_sum;
    stats->natt_sum = natt_sum;
    stats->nstep_max = nstep_max;
    stats->nfe_max = nfe_max;
    for (int s = 0; s < 8; s++) stats->status_count[s] = cnt[s];
  }
  return rc;
}

int or_chem_cells(const nrg_chem_mech *m, const nrg_chem_opts *o, int64_t ncells, double dt,
                  const double *T, const double *rho, const double *const *Y, const int8_t *bc,
                  double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                  int32_t *cell_status, int32_t *nstep, int32_t *nfe, int32_t *nje, int nthreads)
{
  int S = m->species_number;
  int rc = NRG_CHEM_OK;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads) reduction(max : rc)
#endif
  for (int64_t q = 0; q < ncells; q++) {
    double yv[OR_NMAX], yp[OR_NMAX], yp2[OR_NMAX], ef;
    for (int s = 0; s < S; s++) {
      yv[s] = Y[s][q];
      yp2[s] = Y_prod2[s][q];
    }
    or_cell_out co;
    int st = or_chem_cell(m, o, dt, T[q], rho[q], yv, bc ? bc[q] : 0, yp, yp2, &ef, &co);
    for (int s = 0; s < S; s++) {
      Y_prod[s][q] = yp[s];
      Y_prod2[s][q] = yp2[s];
    }
    E_f_prod[q] = ef;
    if (cell_status) cell_status[q] = st;
    if (nstep) nstep[q] = co.bdf.nstep;
    if (nfe) nfe[q] = co.bdf.nfe;
    if (nje) nje[q] = co.bdf.nje;
    int r2 = status_to_rc(st);
    if (r2 > rc) rc = r2;
  }
  return rc;
}

/* interpolate_chemical_kinetics (cm/chemical_kinetics_solver.f90:650-685): every
 * bracketing interval overwrites c (the last one wins), then the two range
 * rules.  Fortran evaluates a*b/c left to right. */
static void or_table_interpolate(int S, int32_t rows, const double *T_t, const double *dC_t,
                                 double T, double *c)
{
  for (int32_t i = 0; i + 1 < rows; i++) {
    if (T >= T_t[i] && T <= T_t[i + 1]) {
      for (int s = 0; s < S; s++)
        c[s] = (dC_t[s * rows + i + 1] * (T - T_t[i])) / (T_t[i + 1] - T_t[i]) +
               (dC_t[s * rows + i] * (T - T_t[i + 1])) / (T_t[i] - T_t[i + 1]);
    }
  }
  if (T < T_t[0])
    for (int s = 0; s < S; s++) c[s] = 0.0;
  if (T > T_t[rows - 1])
    for (int s = 0; s < S; s++) c[s] = dC_t[s * rows + rows - 1];
}

void or_table_cells(const nrg_chem_mech *m, int64_t ncells, double dt, double t_gate,
                    int32_t rows, const double *T_t, const double *dC_t,
                    const double *T, const double *const *Y, const int8_t *bc,
                    double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                    int32_t *cell_status)
{
  const int S = m->species_number;
  double href[OR_NMAX];
  for (int s = 0; s < S; s++) href[s] = or_h_molar(m, T_REF, s);     /* :331 */
  for (int64_t q = 0; q < ncells; q++) {
    if (!(T[q] >= t_gate && (bc ? bc[q] : 0) == 0)) {                  /* :228-231, :341-346 */
      for (int s = 0; s < S; s++) Y_prod[s][q] = 0.0;
      E_f_prod[q] = 0.0;
      if (cell_status) cell_status[q] = NRG_CELL_GATED;
      continue;
    }
    double cin[OR_NMAX], c[OR_NMAX];
    for (int s = 0; s < S; s++) c[s] = cin[s] = Y[s][q];              /* :255-262 */
    or_table_interpolate(S, rows, T_t, dC_t, T[q], c);                 /* :314-315 */
    double ef = 0.0;                                                   /* :235 */
    for (int s = 0; s < S; s++) {
      const double W = m->molar_masses[s];
      Y_prod2[s][q] = c[s] - cin[s];                                   /* :321 */
      Y_prod[s][q] = (c[s] * W) / dt;                                  /* :323-325 */
      ef = ef - (href[s] * c[s]) / dt;                                 /* :334-335 */
    }
    E_f_prod[q] = ef;
    if (cell_status) cell_status[q] = NRG_CELL_OK;
  }
}

/* ======================================================================
 * Low-Mach (FDS) coupling: the per-cell consumers of the chemistry outputs in
 * computing_module/src/current_build/fds_low_mach_solver.f90 and the low-Mach
 * equation of state (table_approximated_real_gas.f90:570-677)
 * ====================================================================== */

void or_low_mach_eos_cells(const nrg_chem_mech *m, int64_t ncells, double dt, int predictor,
                           const double *rho, const double *const *Y, const double *const *v,
                           int ndim, const double *dp_stat_dt, const int8_t *bc, double *p_stat,
                           double *p_stat_old, double *T, double *p, double *mix_mol_mass,
                           double *mixture_cp, double *gamma, double *h_s, double *v_s,
                           double *e_i, double *E_f)
{
  const int S = m->species_number;
  for (int64_t q = 0; q < ncells; q++) {
    if (bc && bc[q] != 0) continue;                                    /* :622 */
    double avg = 0.0;                                                  /* :626-631 */
    for (int s = 0; s < S; s++) avg = avg + Y[s][q] / m->molar_masses[s];
    avg = 1.0 / avg;
    const double W = avg;
    mix_mol_mass[q] = W;
    double X[OR_NMAX];
    for (int s = 0; s < S; s++) X[s] = (Y[s][q] * W) / m->molar_masses[s];   /* :633-635 */
    double ps;
    if (predictor) {                                                   /* :637-642 */
      p_stat_old[q] = p_stat[q];
      ps = p_stat[q] + dp_stat_dt[q] * dt;
    } else {
      ps = 0.5 * ((p_stat[q] + p_stat_old[q]) + dp_stat_dt[q] * dt);
    }
    p_stat[q] = ps;
    const double Tq = ((ps / rho[q]) / R_GAS_J) * W;                   /* :644 */
    T[q] = Tq;
    const double cp = or_mixture_cp_molar(m, Tq, X);                   /* :646-649 */
    const double cv = cp - R_GAS_J;
    const double g = cp / cv;
    if (gamma) gamma[q] = g;
    if (h_s) h_s[q] = (or_mixture_h_molar(m, Tq, X) - or_mixture_h_molar(m, T_REF, X)) / W;   /* :651 */
    if (mixture_cp) mixture_cp[q] = cp;
    if (p) p[q] = ps;                                                  /* :655 */
    if (v_s) v_s[q] = sqrt((g * ps) / rho[q]);                         /* :658 */
    const double ei = (ps / rho[q]) / (g - 1.0);                       /* :659 */
    if (e_i) e_i[q] = ei;
    if (E_f)
      for (int d = 0; d < ndim; d++) E_f[q] = ei + 0.5 * (v[d][q] * v[d][q]);   /* :661-663 */
  }
}

void or_low_mach_species_cells(const nrg_chem_mech *m, int64_t ncells, double dt, int corrector,
                               const double *rho_a, const double *const *Y_a,
                               const double *rho_b, const double *const *Y_b,
                               const double *const *rhs, const double *const *Y_prod_diff,
                               const double *const *Y_prod_chem, const int8_t *bc,
                               double *rho_out, double *const *Y_out)
{
  const int S = m->species_number;
  for (int64_t q = 0; q < ncells; q++) {
    if (bc && bc[q] != 0) continue;                                    /* :848, :3044 */
    double r[OR_NMAX];
    for (int s = 0; s < S; s++) {
      const double f = rhs ? rhs[s][q] : 0.0;                          /* rhs_vec = 0 (:865) + fluxes */
      if (!corrector) {                                                /* :906-909 */
        r[s] = rho_a[q] * Y_a[s][q] + f * dt;
        if (Y_prod_diff) r[s] = r[s] + Y_prod_diff[s][q] * dt;
        if (Y_prod_chem) r[s] = r[s] + Y_prod_chem[s][q] * dt;
      } else {                                                         /* :3103-3107 */
        r[s] = 0.5 * (rho_a[q] * Y_a[s][q] + rho_b[q] * Y_b[s][q]) + f * (0.5 * dt);
        if (Y_prod_diff) r[s] = r[s] + (0.5 * Y_prod_diff[s][q]) * dt;
        if (Y_prod_chem) r[s] = r[s] + (0.5 * Y_prod_chem[s][q]) * dt;
      }
    }
    double sum = 0.0;                                                  /* :918-925, :3116-3123 */
    for (int s = 0; s < S; s++) {
      r[s] = dmax(r[s], 0.0);
      sum = sum + r[s];
    }
    if (sum > DBL_MIN) {                                               /* tiny(1.0_dp) */
      rho_out[q] = sum;
      for (int s = 0; s < S; s++) Y_out[s][q] = r[s] / sum;
    }
  }
}

void or_low_mach_chem_divergence_cells(const nrg_chem_mech *m, int64_t ncells, const double *rho,
                                       const double *T, const double *const *Y,
                                       const double *mix_mol_mass, const double *E_f_prod_chem,
                                       const double *const *Y_prod_chem, const int8_t *bc,
                                       double *div_v)
{
  const int S = m->species_number;
  for (int64_t q = 0; q < ncells; q++) {
    if (bc && bc[q] != 0) continue;                                    /* :1034 */
    const double W = mix_mol_mass[q];
    double d = 0.0;                                                    /* :1057 */
    d = d + E_f_prod_chem[q];                                          /* :1063 */
    double X[OR_NMAX];
    for (int s = 0; s < S; s++) X[s] = (Y[s][q] * W) / m->molar_masses[s];   /* :1075-1078 */
    const double cp = or_mixture_cp_molar(m, T[q], X);                 /* :1117 */
    d = d / (((rho[q] * cp) * T[q]) / W);                              /* :1119 */
    for (int s = 0; s < S; s++) {
      const double hs = (or_h_molar(m, T[q], s) - or_h_molar(m, T_REF, s)) / m->molar_masses[s];   /* :1123 */
      d = d + (((W / m->molar_masses[s]) / rho[q]) - hs / (((rho[q] * cp) * T[q]) / W)) *
                  Y_prod_chem[s][q];                                   /* :1161-1162 */
    }
    div_v[q] = d;
  }
}

void or_closure_cells(const nrg_chem_mech *m, int64_t ncells, double dt,
                      const double *rho, const double *E_f_prod, const double *const *Y_prod,
                      const int8_t *bc, double *E_f, double *const *Y, double *T, double *p,
                      int nthreads)
{
  int S = m->species_number;
  (void)nthreads;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
  for (int64_t q = 0; q < ncells; q++) {
    if (bc && bc[q] != 0) continue;
    /* cabaret_solver.f90:3316-3318 (E_f_prod, Y_prod start at zero), 3386-3416 */
    double efp = 0.0 + E_f_prod[q] * dt;
    E_f[q] = E_f[q] + efp / rho[q];
    double yv[OR_NMAX], sum = 0.0;
    for (int s = 0; s < S; s++) {
      double ypr = 0.0 + Y_prod[s][q] * dt;
      yv[s] = Y[s][q] + ypr / rho[q];
      sum = sum + yv[s];
    }
    for (int s = 0; s < S; s++) yv[s] = dmax(yv[s], 0.0) / sum;
    /* table_approximated_real_gas.f90:412-440 (velocity zero: e_i - 0.5*(0*0)) */
    double e_i = E_f[q];
    double avg = 0.0;
    for (int s = 0; s < S; s++)
      if (m->molar_masses[s] != 0.0) avg = avg + yv[s] / m->molar_masses[s];
    double wmix = 1.0 / avg;
    double X[OR_NMAX];
    for (int s = 0; s < S; s++)
      X[s] = (m->molar_masses[s] != 0.0) ? (yv[s] / m->molar_masses[s]) * wmix : 0.0;
    e_i = e_i * wmix;
    double Tn = or_calculate_temperature(m, T[q], e_i, X);
    T[q] = Tn;
    if (p) p[q] = ((Tn * rho[q]) * R_GAS_J) / wmix;
    for (int s = 0; s < S; s++) Y[s][q] = yv[s];
  }
}

void or_energy_from_temperature_cells(const nrg_chem_mech *m, int64_t ncells, const double *T,
                                      const double *const *Y, double *E_f)
{
  int S = m->species_number;
  for (int64_t q = 0; q < ncells; q++) {
    double avg = 0.0, X[OR_NMAX];
    for (int s = 0; s < S; s++)
      if (m->molar_masses[s] != 0.0) avg = avg + Y[s][q] / m->molar_masses[s];
    double wmix = 1.0 / avg;
    for (int s = 0; s < S; s++)
      X[s] = (m->molar_masses[s] != 0.0) ? (Y[s][q] / m->molar_masses[s]) * wmix : 0.0;
    double e = or_mixture_e_molar(m, T[q], X) - or_mixture_h_molar(m, T_REF, X);
    E_f[q] = e / wmix;
  }
}
