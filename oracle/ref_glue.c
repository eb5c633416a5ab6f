/*
 * ref_glue.c -- C entry to the reference's own SLATEC DDRIV3 (TEST INFRASTRUCTURE).
 *
 * oracle/Makefile compiles the reference's vendored SLATEC sources
 * (every .f90 file in computing_module/src/third_party/slatec, unmodified, from
 * where they lie) with flang and links them with this file into oracle/_ref/libslatec_ref.so.
 * Tests use it to check the C restatement in nrg_oracle.c bit for bit.
 *
 * ref_ddriv3 makes the call chemical_kinetics_solver.f90:311-313 makes, with
 * the same settings (:270-295), but with the right-hand side supplied from C.
 */
#include <stdint.h>
#include <string.h>

typedef void (*ref_rhs_fn)(void *ctx, int n, double t, const double *y, double *ydot);

typedef void (*f77_sub)(int *, double *, double *, double *);
typedef double (*f77_g)(int *, double *, double *, int *);

extern void ddriv3_(int *n, double *t, double *y, f77_sub f, int *nstate, double *tout, int *ntask,
                    int *nroot, double *eps, double *ewt, int *ierror, int *mint, int *miter,
                    int *impl, int *ml, int *mu, int *mxord, double *hmax, double *work, int *lenw,
                    int *iwork, int *leniw, f77_sub jacobn, f77_sub fa, int *nde, int *mxstep,
                    f77_g g, f77_sub users, int *ierflg);

static ref_rhs_fn g_rhs;
static void *g_ctx;

static void f_bridge(int *n, double *t, double *y, double *ydot)
{
  g_rhs(g_ctx, *n, *t, y, ydot);
}

static double g_dummy(int *n, double *t, double *y, int *iroot)
{
  (void)n; (void)y; (void)iroot;
  return *t;
}

/* Integrate y' = f(y) from 0 to tout with DDRIV3.  iwork_out (>= 59 ints) receives
 * IWORK (NQUSED = iwork[1], NSTEP = iwork[2], NFE = iwork[3], NJE = iwork[4]).
 * Returns NSTATE; *ierflg gets IERFLG.  Not thread-safe (as DDRIV3 itself is not). */
int ref_ddriv3(int n, double *y, double tout, double eps, double ewt, double hmax,
               int mxstep, int mxord, ref_rhs_fn rhs, void *ctx, int *ierflg, int *iwork_out,
               double *work_out)
{
  double t = 0.0, ewtv[1] = {ewt};
  int nstate = 1, ntask = 1, nroot = 0, ierror = 3, mint = 2, miter = 2, impl = 0, ml = 0, mu = 0;
  int lenw = n * n + (mxord + 5) * n + 2 * nroot + 250, leniw = n + 50, nde = n, ierr = 0;
  double work[64 * 64 + 20 * 64 + 250];
  int iwork[64 + 50];
  if (n > 64) return -1;
  memset(work, 0, sizeof(work));
  memset(iwork, 0, sizeof(iwork));
  g_rhs = rhs;
  g_ctx = ctx;
  ddriv3_(&n, &t, y, f_bridge, &nstate, &tout, &ntask, &nroot, &eps, ewtv, &ierror, &mint, &miter,
          &impl, &ml, &mu, &mxord, &hmax, work, &lenw, iwork, &leniw, f_bridge, f_bridge, &nde,
          &mxstep, g_dummy, f_bridge, &ierr);
  if (ierflg) *ierflg = ierr;
  if (iwork_out) memcpy(iwork_out, iwork, sizeof(int) * (size_t)leniw);
  if (work_out) memcpy(work_out, work, sizeof(double) * (size_t)lenw);
  return nstate;
}
