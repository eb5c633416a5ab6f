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
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include "nrg_oracle.h"

typedef void (*ref_rhs_fn)(void *ctx, int n, double t, const double *y, double *ydot);

typedef void (*f77_sub)(int *, double *, double *, double *);
typedef double (*f77_g)(int *, double *, double *, int *);

extern void ddriv3_(int *n, double *t, double *y, f77_sub f, int *nstate, double *tout, int *ntask,
                    int *nroot, double *eps, double *ewt, int *ierror, int *mint, int *miter,
                    int *impl, int *ml, int *mu, int *mxord, double *hmax, double *work, int *lenw,
                    int *iwork, int *leniw, f77_sub jacobn, f77_sub fa, int *nde, int *mxstep,
                    f77_g g, f77_sub users, int *ierflg);

/* per thread: ref_chem_cells calls DDRIV3 from several threads at once */
static __thread ref_rhs_fn g_rhs;
static __thread void *g_ctx;

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

/* ---- the chemistry step over a list of cells, integrated by the reference's DDRIV3 ----
 *
 * One cell of solve_chemical_kinetics (chemical_kinetics_solver.f90:228-346):
 * gate, [M] and rate constants (:235-252), C_in (:254-265), DDRIV3 with the
 * reference's settings (:270-313) -- here the reference's own SLATEC build --
 * and the output assembly (:318-346).  [M], the rate constants and the
 * right-hand side `fun` come from the oracle (or_chem_ctx_init, or_chem_fun),
 * because the reference's chemistry module does not compile unmodified.
 * Used as the CPU baseline's integrator and checked bit for bit against
 * or_chem_cells in tests/test_oracle_slatec.py. */
#define T_REF_K 298.15

typedef struct ref_job {
  const nrg_chem_mech *m;
  const nrg_chem_opts *o;
  int64_t n;
  double dt;
  const double *T, *rho;
  const double *const *Y;
  double *const *Y_prod;
  double *const *Y_prod2;
  double *E_f_prod;
  int32_t *status, *nstep, *nfe, *nje;
  int64_t next;           /* next unclaimed cell (atomic) */
} ref_job;

static void ref_cell(const ref_job *j, int64_t q)
{
  const nrg_chem_mech *m = j->m;
  const int S = m->species_number;
  const double T = j->T[q];
  if (!(T >= j->o->t_gate)) {                                      /* :228-231, :341-346 */
    for (int s = 0; s < S; s++) j->Y_prod[s][q] = 0.0;
    j->E_f_prod[q] = 0.0;
    if (j->status) j->status[q] = NRG_CELL_GATED;
    if (j->nstep) j->nstep[q] = j->nfe[q] = j->nje[q] = 0;
    return;
  }
  unsigned char ctx[16384] __attribute__((aligned(16)));
  double y[OR_NMAX] = {0.0}, cin[OR_NMAX], cout[OR_NMAX];
  for (int s = 0; s < S; s++) y[s] = j->Y[s][q];
  or_chem_ctx_init(ctx, m, T, j->rho[q], y, cin);
  for (int s = 0; s < S; s++) cout[s] = cin[s];
  int ierflg = 0, iwork[64 + 50];
  const int nstate = ref_ddriv3(S, cout, j->dt, j->o->eps, j->o->ewt, j->o->hmax, j->o->mxstep,
                                j->o->mxord, or_chem_fun, ctx, &ierflg, iwork, NULL);
  double ef = 0.0;
  for (int s = 0; s < S; s++) {
    const double dc = cout[s] - cin[s];
    j->Y_prod[s][q] = (dc * m->molar_masses[s]) / j->dt;
    j->Y_prod2[s][q] = dc;
  }
  for (int s = 0; s < S; s++) ef = ef - (or_h_molar(m, T_REF_K, s) * (cout[s] - cin[s])) / j->dt;
  j->E_f_prod[q] = ef;
  if (j->status)
    j->status[q] = nstate == 2 ? NRG_CELL_OK
                 : nstate == 3 ? NRG_CELL_MXSTEP
                 : ierflg == 41 ? NRG_CELL_STEP_ZERO
                 : ierflg == 42 ? NRG_CELL_REPEATED_FAIL : NRG_CELL_SINGULAR;
  if (j->nstep) {
    j->nstep[q] = iwork[2];
    j->nfe[q] = iwork[3];
    j->nje[q] = iwork[4];
  }
}

static void *ref_worker(void *arg)
{
  ref_job *j = (ref_job *)arg;
  for (;;) {                   /* dynamic schedule in chunks of 16 cells, as the oracle's loop */
    const int64_t q0 = __atomic_fetch_add(&j->next, 16, __ATOMIC_RELAXED);
    if (q0 >= j->n) break;
    const int64_t q1 = q0 + 16 < j->n ? q0 + 16 : j->n;
    for (int64_t q = q0; q < q1; q++) ref_cell(j, q);
  }
  return NULL;
}

/* Flat list of cells (SoA, stride 1), no markers: the arguments of
 * or_chem_cells.  nthreads host threads (pthreads).  Returns 0, or -1 when the
 * context does not fit or a thread cannot start. */
int ref_chem_cells(const nrg_chem_mech *m, const nrg_chem_opts *o, int64_t ncells, double dt,
                   const double *T, const double *rho, const double *const *Y,
                   double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                   int32_t *cell_status, int32_t *nstep, int32_t *nfe, int32_t *nje, int nthreads)
{
  if (or_chem_ctx_size() > 16384 || m->species_number > 64) return -1;
  ref_job j = {m, o, ncells, dt, T, rho, Y, Y_prod, Y_prod2, E_f_prod,
               cell_status, nstep, nfe, nje, 0};
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t tid[256];
  int started = 0, rc = 0;
  for (int i = 1; i < nthreads; i++) {
    if (pthread_create(&tid[started], NULL, ref_worker, &j) != 0) {
      rc = -1;
      break;
    }
    started++;
  }
  ref_worker(&j);
  for (int i = 0; i < started; i++) pthread_join(tid[i], NULL);
  return rc;
}
