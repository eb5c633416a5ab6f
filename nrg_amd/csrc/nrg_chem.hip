// nrg_chem.hip -- kernels and entry points of the gfx950 chemistry step for
// one reaction topology.
//
// Entry points are declared in include/nrg_chem.h; each replaces a piece of
// NRG's chemical_kinetics_solver (computing_module/src/current_build/
// chemical_kinetics_solver.f90) or of the split-step closure the flow solvers
// apply after it.  All device work runs on the handle's stream.
//
// The kernels unroll the mechanism's reaction lists at compile time, so this
// file is built once per topology (the Makefile sets NRG_TOPOLOGY_HEADER,
// NRG_TOPO_NS, NRG_TOPO_OPS_FN, NRG_TOPO_NAME and NRG_TOPO_PRIMARY).  Each build
// lives in its own namespace and hands its entry points to the C-ABI
// (nrg_api.cpp) as a table (topo_api.h); nrg_chem_create picks the table whose
// topology the mechanism has.
//
// NRG_PART splits a build into two objects: 1 = everything but the
// integrator kernel's body, 2 = the integrator kernel alone (so that it can be
// compiled with scheduling options of its own); 0 (default) = both in one.
// With NRG_PART 2, NRG_BDF_JAC selects the kernel's instantiations: 1 = those
// with the finite-difference Jacobian, 2 = those with the analytic one, unset
// = all four.
#ifndef NRG_PART
#define NRG_PART 0
#endif
#ifndef NRG_TOPO_NS
#define NRG_TOPO_NS nrg_topo_keromnes
#define NRG_TOPO_OPS_FN nrg_topo_ops_keromnes
#define NRG_TOPO_NAME "KEROMNES"
#define NRG_TOPO_PRIMARY 1
#endif
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/nrg_chem.h"
#include "../../include/nrg_low_mach.h"
#include "../../include/nrg_surrogate.h"
#include "chem_device.h"
#include "mlp_internal.h"
#include "nrg_common.h"
#include "topo_api.h"

using namespace nrg;

// nothing of this build is exported: the C-ABI reaches it through its table
#pragma GCC visibility push(hidden)
namespace NRG_TOPO_NS {

// This build's handle; the C-ABI sees it only as an opaque pointer.
struct nrg_chem_handle;
typedef nrg_chem_handle nrg_chem_t;
void nrg_chem_destroy(nrg_chem_t *h);
int nrg_chem_table_step_dev(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T,
                            const double *p, const double *rho, const double *const *Y,
                            const int8_t *bc, double *const *Y_prod, double *const *Y_prod2,
                            double *E_f_prod, int32_t *cell_status, nrg_chem_stats *stats);

namespace kern {

// Field addressing of one call (Fortran column-major with ghost layers).
struct GridIdx {
  int lo0, lo1, lo2;
  int nx, ny;                 // allocated extents in dims 1, 2
  int ilo0, ilo1, ilo2;       // inner lower bounds
  int ix, iy, iz;             // inner extents
  int idx32;                  // 1: inner indices and allocation offsets below 2^31 (32-bit division)
  long long ninner;
  long long pad0, pad1;       // allocation offsets [pad0, pad1) holding every inner cell; pad0 % 64 == 0
};

// Is allocation offset p an inner cell (not a ghost)?
NRG_DI bool inner_at(const GridIdx &g, long long p)
{
  long long i, j, k;
  if (g.idx32) {
    const unsigned pu = (unsigned)p, nx = (unsigned)g.nx, ny = (unsigned)g.ny;
    const unsigned r = pu / nx;
    i = pu - r * nx;
    j = r % ny;
    k = r / ny;
  } else {
    const long long r = p / g.nx;
    i = p - r * g.nx;
    j = r % g.ny;
    k = r / g.ny;
  }
  return (unsigned long long)(i - (g.ilo0 - g.lo0)) < (unsigned long long)g.ix &&
         (unsigned long long)(j - (g.ilo1 - g.lo1)) < (unsigned long long)g.iy &&
         (unsigned long long)(k - (g.ilo2 - g.lo2)) < (unsigned long long)g.iz;
}

// The aligned walk: thread t of the launch takes allocation offset pad0 + t,
// from the 64-cell slice of the first inner cell to the last inner cell, and
// skips ghost cells.  With ghost layers a row of inner cells starts off a cache
// line; walking inner cells in order would make every wave touch one more line
// per field.  One offset per thread: a grid-stride loop here let the compiler
// hoist the LDS-staged mechanism into registers (closure 72 -> 169 VGPRs) and
// was slower for every kernel that uses it.
template <class F>
NRG_DI void at_inner_cell(const GridIdx &g, F &&f)
{
  const long long off = g.pad0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (off < g.pad1 && inner_at(g, off)) f(off);
}

NRG_DI long long cell_offset(const GridIdx &g, long long q)
{
  long long i, j, kk;
  if (g.idx32) {
    const unsigned qu = (unsigned)q, ix = (unsigned)g.ix, iy = (unsigned)g.iy;
    const unsigned r = qu / ix;
    i = (long long)(qu - r * ix) + g.ilo0;
    j = (long long)(r % iy) + g.ilo1;
    kk = (long long)(r / iy) + g.ilo2;
  } else {
    i = q % g.ix + g.ilo0;
    j = (q / g.ix) % g.iy + g.ilo1;
    kk = q / ((long long)g.ix * g.iy) + g.ilo2;
  }
  return (i - g.lo0) + (long long)g.nx * ((j - g.lo1) + (long long)g.ny * (kk - g.lo2));
}

struct StepArgs {
  GridIdx g;
  double dt;
  BdfConst k;
  const MechCoef *mech;
  const double *T, *rho;
  const double *Y[S];
  const int8_t *bc;
  double *Yp[S], *Yp2[S];
  double *Efp;
  int32_t *status;
  int32_t *nstep, *nfe, *nje;          // optional per-cell work counters
  unsigned long long *acc;              // [16] work/status accumulators
  unsigned long long *queue;            // next unclaimed cell of the chunk (persistent form)
  int jac_pct;                          // run a Jacobian batch when >= this % of pending lanes want one
  const double *rates;                  // cell records of the chunk's positions (REC_PAIRS pairs each)
  double *rates_out;                    // same buffer, written by rates_kernel
  long long base, nchunk;               // this launch covers queue positions [base, base+nchunk)
  const int32_t *list;                  // NULL: position t is inner cell t; else inner cell list[t]
  int32_t *hist;                        // optional: NSTEP of each inner cell (0 if gated), for ordering
};

// Inner cell (inner-loop order) at position base + t of the launch's queue.
// The 64-bit form alone: the integrator kernel's one call (at a refill) keeps
// its register allocation without the second path.
NRG_DI long long cell_offset_wide(const GridIdx &g, long long q)
{
  const long long i = q % g.ix + g.ilo0;
  const long long j = (q / g.ix) % g.iy + g.ilo1;
  const long long kk = q / ((long long)g.ix * g.iy) + g.ilo2;
  return (i - g.lo0) + (long long)g.nx * ((j - g.lo1) + (long long)g.ny * (kk - g.lo2));
}

NRG_DI long long cell_of(const StepArgs &a, long long t)
{
  return a.list ? (long long)a.list[a.base + t] : a.base + t;
}

// Block-wide staging of the mechanism coefficients into LDS: each thread
// issues its loads (up to 4) before the first store, so a block waits for one
// round trip to memory instead of one per element it copies.
NRG_DI void stage_mech(MechCoef &dst, const MechCoef *src)
{
  const double *s = reinterpret_cast<const double *>(src);
  double *d = reinterpret_cast<double *>(&dst);
  constexpr int n = sizeof(MechCoef) / sizeof(double);
  constexpr int K = 4;
  const int stride = blockDim.x;
  for (int i0 = threadIdx.x; i0 < n; i0 += K * stride) {
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = (i0 + k * stride < n) ? s[i0 + k * stride] : 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (i0 + k * stride < n) d[i0 + k * stride] = v[k];
  }
  __syncthreads();
}

NRG_DI void stage_bdf(BdfCoef &dst, const MechCoef *src)
{
  for (int i = threadIdx.x; i < MAXORD * (MAXORD + 1); i += blockDim.x)
    (&dst.el[0][0])[i] = (&src->el[0][0])[i];
  for (int i = threadIdx.x; i < MAXORD * 3; i += blockDim.x) (&dst.tq[0][0])[i] = (&src->tq[0][0])[i];
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    dst.W[i] = src->W[i];
    dst.h_ref[i] = src->h_ref[i];
  }
  __syncthreads();
}

NRG_DI unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

NRG_DI int wave_max_i32(int v)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

// accumulator slots
enum { ACC_INTEG = 0, ACC_NSTEP, ACC_NFE, ACC_NJE, ACC_NCORR, ACC_NATT, ACC_NSTEP_MAX, ACC_NFE_MAX,
       ACC_STATUS0,     // ACC_STATUS0 + NRG_CELL_* (8 slots)
       ACC_QUEUE0 = 16, // one queue head per chunk
       MAX_CHUNKS = 48,
       ACC_SLOTS = ACC_QUEUE0 + MAX_CHUNKS };

// Zero the work accumulators and the queue head (a kernel, so the launch
// sequence can be captured in a graph without memset nodes).
#if NRG_PART != 2
__global__ void reset_kernel(unsigned long long *acc, int n)
{
  if ((int)threadIdx.x < n) acc[threadIdx.x] = 0ull;
}
#endif

typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(3))) double lds_double;

// The integrator keeps the whole cell record in LDS when four single-wave
// blocks still fit a CU's 160 KiB (every mechanism up to ~37 record pairs);
// a larger mechanism keeps the rate constants and [M] there and reads the
// initial concentrations and the gate flag from HBM.
constexpr size_t LDS_PER_CU = 160 * 1024;
constexpr bool REC_C_IN_LDS = 4 * ((size_t)REC_PAIRS * 16 * WAVE + sizeof(BdfCoef)) <= LDS_PER_CU;
constexpr int REC_LDS_PAIRS = REC_C_IN_LDS ? REC_PAIRS : REC_C;

// Load one cell and set it up (chemical_kinetics_solver.f90:228-265).  The
// cell record of position t (rates_kernel) goes from HBM straight into the
// lane's LDS slots (recl: the block's record array, kl: this lane's view of
// it): one round trip, no registers in between.  Returns false for a gated
// cell, whose outputs are written here (:341-346).
NRG_DI bool start_cell(const StepArgs &a, long long t, long long q, long long &off, Cell &c,
                       lds_double *recl, const double *kl)
{
  const double *src = a.rates + 2 * t;
#pragma unroll
  for (int p = 0; p < REC_LDS_PAIRS; ++p)
    __builtin_amdgcn_global_load_lds((const void *)(src + 2 * (long long)p * a.nchunk),
                                     (lds_void *)(recl + 2 * p * WAVE), 16, 0, 0);   // M0: LDS address
  double cs[S + 1];
  if (!REC_C_IN_LDS) {
#pragma unroll
    for (int s = 0; s <= S; ++s) cs[s] = src[rec_slot(REC_C, s, a.nchunk)];
  }
  off = cell_offset_wide(a.g, q);
  // a direct load's data is in LDS once vmcnt has counted it
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  if (REC_C_IN_LDS) {
#pragma unroll
    for (int s = 0; s <= S; ++s) cs[s] = rec_c(kl, s);
  }
  if (cs[REC_FLAG] == 0.0) {
#pragma unroll
    for (int s = 0; s < S; ++s) a.Yp[s][off] = 0.0;
    a.Efp[off] = 0.0;
    if (a.status) a.status[off] = NRG_CELL_GATED;
    if (a.nstep) {
      a.nstep[q] = 0;
      a.nfe[q] = 0;
      a.nje[q] = 0;
    }
    if (a.hist) a.hist[q] = 0;
    return false;
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (kSp.act[s] >= 0)
      c.y[kSp.act[s]] = cs[s];
    else
      c.yi[kSp.islot[s]] = cs[s];
  }
  bdf_init(c, a.dt, a.k);
  return true;
}

#if NRG_PART != 2
// Cell records of the chunk's positions: the gate (:224-227), and for every
// integrated cell its concentrations (:228-233), rate constants and [M]
// (:235-252, 360-493), kept apart from the integrator so its ~80
// transcendentals run coherently.  A block covers
// RATES_CELLS cells RATES_GROUPS times: the waves of group g compute
// reactions [g R / G, (g+1) R / G) (a wave-uniform split, so each wave's code
// and registers cover a part of the mechanism: occupancy 4 instead of 1-2).
#ifndef RATES_GROUPS
#define RATES_GROUPS 2
#endif
constexpr int RATES_CELLS = 256 / RATES_GROUPS;

__global__ void __launch_bounds__(256) rates_kernel(StepArgs a)
{
  __shared__ MechCoef mc;
  stage_mech(mc, a.mech);
  const int grp = __builtin_amdgcn_readfirstlane(threadIdx.x / RATES_CELLS);   // wave-uniform
  const long long t = (long long)blockIdx.x * RATES_CELLS + (threadIdx.x % RATES_CELLS);
  if (t >= a.nchunk) return;
  const long long q = cell_of(a, t);
  const long long off = cell_offset(a.g, q);
  const double T = a.T[off];
  const int8_t bcv = a.bc ? a.bc[off] : (int8_t)0;
  double *out = a.rates_out + 2 * t;
  if (!(T >= a.k.t_gate && bcv == 0)) {
    if (grp == 0)   // the gate flag's pair
      *reinterpret_cast<double2 *>(out + rec_slot(REC_C, REC_FLAG & ~1, a.nchunk)) = make_double2(0.0, 0.0);
    return;
  }
  const double rho = a.rho[off];
  double C[S];
#pragma unroll
  for (int s = 0; s < S; ++s) C[s] = (a.Y[s][off] * rho) / mc.W[s];
  constexpr int G = RATES_GROUPS;
  static_assert(G == 1 || G == 2 || G == 4, "RATES_GROUPS must be 1, 2 or 4");
  if (G == 1 || grp == 0)
    rate_constants<InlineMath, 0, R / G>(mc, T, C, out, a.nchunk);
  else if (G == 2 || grp == 1)
    rate_constants<InlineMath, R / G, 2 * R / G>(mc, T, C, out, a.nchunk);
  else if (grp == 2)
    rate_constants<InlineMath, 2 * R / G, 3 * R / G>(mc, T, C, out, a.nchunk);
  else
    rate_constants<InlineMath, 3 * R / G, R>(mc, T, C, out, a.nchunk);
  // the concentrations and the gate flag, written last (stored ahead of the
  // rate constants, they doubled the kernel's registers: occupancy 4 -> 2)
  if (grp == 0) {
#pragma unroll
    for (int j = 0; j <= REC_FLAG; j += 2) {
      const double lo = j < S ? C[j] : 1.0;
      const double hi = j + 1 < S ? C[j + 1] : (j + 1 == REC_FLAG ? 1.0 : 0.0);
      *reinterpret_cast<double2 *>(out + rec_slot(REC_C, j, a.nchunk)) = make_double2(lo, hi);
    }
  }
}

#endif

// Output assembly (:318-340) once the cell's DDRIV3 has returned; returns the
// cell's NRG_CELL_* status.
NRG_DI int finish_cell(const StepArgs &a, const BdfCoef &mc, long long q, long long off,
                       const Cell &c, const double *kl)
{
  const double rho = REC_C_IN_LDS ? 0.0 : a.rho[off];
  double ef = 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    // the initial concentration as rates_kernel recorded it (the same expression)
    const double cin = REC_C_IN_LDS ? rec_c(kl, s) : (a.Y[s][off] * rho) / mc.W[s];
    const double ys = (kSp.act[s] >= 0) ? c.y[kSp.act[s]] : c.yi[kSp.islot[s]];
    const double dc = ys - cin;
    a.Yp[s][off] = (dc * mc.W[s]) / a.dt;
    a.Yp2[s][off] = dc;
    ef = ef - (mc.h_ref[s] * dc) / a.dt;
  }
  a.Efp[off] = ef;
  const int status = (c.outcome == OUT_OK) ? NRG_CELL_OK
                   : (c.outcome == OUT_MXSTEP || c.outcome == OUT_EPS) ? NRG_CELL_MXSTEP
                   : (c.outcome == OUT_STEP_ZERO) ? NRG_CELL_STEP_ZERO
                   : (c.outcome == OUT_REPEATED_FAIL) ? NRG_CELL_REPEATED_FAIL
                   : NRG_CELL_SINGULAR;
  if (a.status) a.status[off] = status;
  if (a.nstep) {
    a.nstep[q] = c.nstep;
    a.nfe[q] = c.nfe;
    a.nje[q] = c.nje;
  }
  if (a.hist) a.hist[q] = c.nstep;
  return status;
}

// Per-lane work and status counts, reduced once per wave at the end.
struct Tally {
  unsigned long long cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  unsigned long long work[6] = {0, 0, 0, 0, 0, 0};
  int mx_step = 0, mx_fe = 0;

  NRG_DI void integrated(const Cell &c, int status)
  {
#pragma unroll
    for (int st = 0; st < 8; ++st) cnt[st] += (status == st) ? 1ull : 0ull;
    work[0] += 1;
    work[1] += c.nstep;
    work[2] += c.nfe;
    work[3] += c.nje;
    work[4] += c.ncorr;
    work[5] += c.natt;
    mx_step = max(mx_step, c.nstep);
    mx_fe = max(mx_fe, c.nfe);
  }

  NRG_DI void flush(unsigned long long *acc, int lane)
  {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const unsigned long long w = wave_sum_u64(work[i]);
      if (lane == 0 && w) atomicAdd(&acc[ACC_INTEG + i], w);
    }
    const int ms = wave_max_i32(mx_step);
    const int mf = wave_max_i32(mx_fe);
    if (lane == 0) {
      atomicMax(&acc[ACC_NSTEP_MAX], (unsigned long long)ms);
      atomicMax(&acc[ACC_NFE_MAX], (unsigned long long)mf);
    }
#pragma unroll
    for (int st = 0; st < 6; ++st) {
      const unsigned long long w = wave_sum_u64(cnt[st]);
      if (lane == 0 && w) atomicAdd(&acc[ACC_STATUS0 + st], w);
    }
  }
};

// One chemistry step (solve_chemical_kinetics, :224-346), one cell per lane.
// PERSISTENT = false: lane i of block b owns inner cell 64*b + i (the plain
//   launch).  PERSISTENT = true: a resident grid; once k.refill_idle lanes
//   of a wave have finished (or all have), they take the next unclaimed cells
//   from a.queue (one atomic per wave per refill).
// Each iteration either evaluates f for the lanes that asked, or -- when at
// least jac_pct % of the pending lanes wait for a Jacobian, or nothing else is
// pending -- runs one batched FD Jacobian + LU for the waiting lanes.
constexpr int REFILL_MIN_IDLE_DEFAULT = 63;   // measured: 48 -> 41.9, 56 -> 40.5, 63 -> 40.0 ms per 256^3 step

#ifndef NRG_BDF_WAVES_PER_EU
#define NRG_BDF_WAVES_PER_EU 1   // resident waves per SIMD the register budget is sized for
#endif
template <bool PERSISTENT, bool ANALYTIC>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(NRG_BDF_WAVES_PER_EU)))
chem_bdf_kernel(StepArgs a)
#if NRG_PART == 1
;   // defined in the integrator part of this build
extern template __global__ void chem_bdf_kernel<false, false>(StepArgs);
extern template __global__ void chem_bdf_kernel<false, true>(StepArgs);
extern template __global__ void chem_bdf_kernel<true, false>(StepArgs);
extern template __global__ void chem_bdf_kernel<true, true>(StepArgs);
#else
{
  __shared__ BdfCoef mc;
  __shared__ __align__(16) double recl[2 * REC_LDS_PAIRS * WAVE];   // the lanes' cell records
  stage_bdf(mc, a.mech);
  const int lane = threadIdx.x;
  const double *kl = recl + 2 * lane;
  lds_double *recl_lds = (lds_double *)recl;   // the array's LDS address (direct loads' M0)
  Cell c;
  c.req = REQ_DONE;
  long long q = -1, off = 0;
  bool have = false, exhausted = false;
  // The persistent form tallies each finished cell as it goes; a lane of the
  // plain form owns one cell, whose counts are read from its state at the end
  // (keeping the tally out of the registers live across the loop).
  Tally tl;
  bool gated = false;
  int status = 0;
  if (!PERSISTENT) {
    const long long t = (long long)blockIdx.x * WAVE + lane;
    if (t < a.nchunk) {
      q = cell_of(a, t);
      have = start_cell(a, t, q, off, c, recl_lds, kl);
      gated = !have;
    }
  }
  for (;;) {
    if (PERSISTENT) {
      // refill in batches: new cells start together, so the wave's lanes stay
      // near the same stage (refilling every pass measured 30 % slower)
      bool need = !have && !exhausted;
      if (wave_any(have) && __popcll(__ballot(need ? 1 : 0)) < a.k.refill_idle) need = false;
      while (wave_any(need)) {
        const unsigned long long mask = __ballot(need ? 1 : 0);
        const int leader = __ffsll((long long)mask) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.queue, (unsigned long long)__popcll(mask));
        base = __shfl(base, leader, WAVE);
        if (need) {
          const long long t =
              (long long)(base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull)));
          if (t >= a.nchunk) {
            exhausted = true;
          } else {
            q = cell_of(a, t);
            have = start_cell(a, t, q, off, c, recl_lds, kl);
            if (!have) tl.cnt[NRG_CELL_GATED]++;
          }
        }
        need = !have && !exhausted;
      }
    }
    if (!wave_any(have)) break;
    const bool wantj = have && c.req == REQ_JAC;
    const bool wantf = have && c.req == REQ_F;
    const int nj = __popcll(__ballot(wantj ? 1 : 0));
    const int nf = __popcll(__ballot(wantf ? 1 : 0));
    const bool run_jac = nj > 0 && (nf == 0 || nj * 100 >= a.jac_pct * (nj + nf));
    if (run_jac) {
      if (ANALYTIC)
        jac_analytic(c, wantj, mc, kl);
      else
        jac_batch(c, wantj, a.k, mc, kl);
    }
    else if (wantf)
      rhs(kl, c.y, c.save2);
    if (have && (c.req == REQ_NONE || (run_jac ? wantj : wantf))) {
      c.req = REQ_NONE;
      advance(c, a.dt, a.k, mc);
    }
    if (have && c.req == REQ_DONE) {
      status = finish_cell(a, mc, q, off, c, kl);
      if (PERSISTENT) tl.integrated(c, status);
      have = false;
    }
    if (!PERSISTENT && !wave_any(have)) break;
  }
  if (!PERSISTENT) {
    if (gated)
      tl.cnt[NRG_CELL_GATED]++;
    else if (q >= 0)
      tl.integrated(c, status);
  }
  tl.flush(a.acc, lane);   // one atomic per wave and slot
}
#endif

#if NRG_PART == 2
// The integrator part ends here: its kernels, compiled with their own
// scheduling options (the Makefile), and nothing else.
#if !defined(NRG_BDF_JAC) || NRG_BDF_JAC == 1
template __global__ void chem_bdf_kernel<false, false>(StepArgs);
template __global__ void chem_bdf_kernel<true, false>(StepArgs);
#endif
#if !defined(NRG_BDF_JAC) || NRG_BDF_JAC == 2
template __global__ void chem_bdf_kernel<false, true>(StepArgs);
template __global__ void chem_bdf_kernel<true, true>(StepArgs);
#endif
}  // namespace kern
}  // namespace NRG_TOPO_NS
#pragma GCC visibility pop
#else

// ---------------------------------------------------------------------------
// Split-step closure and EOS (cabaret_solver.f90:3316-3318,3386-3416;
// table_approximated_real_gas.f90:410-440; thermophysical_properties_class.f90:409-452)
// ---------------------------------------------------------------------------
struct ClosureArgs {
  GridIdx g;
  double dt;
  int apply_sources;                    // 0: EOS only
  int ndim;
  const MechCoef *mech;
  const double *rho;
  const double *v[3];
  const double *Efp;
  const double *Yp[S];
  const int8_t *bc;
  double *E_f;
  double *Y[S];
  double *T, *p;
};

NRG_DI double calculate_temperature(const MechCoef &mc, double T, double e_i, const double (&X)[S])
{
  const double eps = (double)1e-08f;
  double temp = clip_t(T);
  double h_const = 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) h_const = h_const + mc.h_ref[s] * X[s];
  double cp = 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) cp = cp + cp_molar(mc, temp, s) * X[s];
  for (int iter = 1; iter <= 1000; ++iter) {
    double e = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) e = e + e_molar(mc, temp, s) * X[s];
    const double e_int = e - h_const;
    const double residual = (e_i - e_int) / cp;
    temp = temp + residual;
    if (fabs(residual) < eps) break;
  }
  return temp;
}

__global__ void __launch_bounds__(256) closure_kernel(ClosureArgs a)
{
  __shared__ MechCoef mc;
  stage_mech(mc, a.mech);
  at_inner_cell(a.g, [&](long long off) {
    if (a.bc && a.bc[off] != 0) return;
    const double rho = a.rho[off];
    double Ef = a.E_f[off];
    double y[S];
    if (a.apply_sources) {
      const double efp = 0.0 + a.Efp[off] * a.dt;
      Ef = Ef + efp / rho;
      double sum = 0.0;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const double ypr = 0.0 + a.Yp[s][off] * a.dt;
        y[s] = a.Y[s][off] + ypr / rho;
        sum = sum + y[s];
      }
#pragma unroll
      for (int s = 0; s < S; ++s) y[s] = dmax(y[s], 0.0) / sum;
      a.E_f[off] = Ef;
#pragma unroll
      for (int s = 0; s < S; ++s) a.Y[s][off] = y[s];
    } else {
#pragma unroll
      for (int s = 0; s < S; ++s) y[s] = a.Y[s][off];
    }
    double e_i = Ef;
    for (int d = 0; d < a.ndim; ++d) {
      const double vd = a.v[d][off];
      e_i = e_i - 0.5 * (vd * vd);
    }
    double avg = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) avg = avg + y[s] / mc.W[s];
    const double wmix = 1.0 / avg;
    double X[S];
#pragma unroll
    for (int s = 0; s < S; ++s) X[s] = (y[s] / mc.W[s]) * wmix;
    e_i = e_i * wmix;
    const double Tn = calculate_temperature(mc, a.T[off], e_i, X);
    a.T[off] = Tn;
    if (a.p) a.p[off] = ((Tn * rho) * R_GAS_J) / wmix;
  });
}

struct EnergyArgs {
  GridIdx g;
  int ndim;
  const MechCoef *mech;
  const double *T;
  const double *Y[S];
  const double *v[3];
  double *E_f;
};

__global__ void __launch_bounds__(256) energy_kernel(EnergyArgs a)
{
  __shared__ MechCoef mc;
  stage_mech(mc, a.mech);
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.g.ninner) return;
  const long long off = cell_offset(a.g, q);
  double y[S], X[S], avg = 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    y[s] = a.Y[s][off];
    avg = avg + y[s] / mc.W[s];
  }
  const double wmix = 1.0 / avg;
#pragma unroll
  for (int s = 0; s < S; ++s) X[s] = (y[s] / mc.W[s]) * wmix;
  const double T = a.T[off];
  double e = 0.0, hc = 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) e = e + e_molar(mc, T, s) * X[s];
#pragma unroll
  for (int s = 0; s < S; ++s) hc = hc + mc.h_ref[s] * X[s];
  double Ef = (e - hc) / wmix;
  for (int d = 0; d < a.ndim; ++d) {
    const double vd = a.v[d][off];
    Ef = Ef + 0.5 * (vd * vd);
  }
  a.E_f[off] = Ef;
}

// ---------------------------------------------------------------------------
// Table-approximated chemistry (chemical_kinetics_solver.f90:250-346 with
// ODE_solver = 'table_approximated', interpolate_chemical_kinetics :650-685):
// the "concentrations" are the mass fractions themselves, replaced by the
// table's dC(T) interpolated linearly in T (the last bracketing interval
// wins; below the table 0, above it the last row); then Y_prod = C_out*W/dt,
// Y_prod2 = C_out - Y, E_f_prod = -sum_s h_s(T_ref)*C_out/dt.
// ---------------------------------------------------------------------------
struct TableArgs {
  GridIdx g;
  double dt, t_gate;
  const MechCoef *mech;
  const double *T;
  const double *Y[S];
  const int8_t *bc;
  double *Yp[S], *Yp2[S];
  double *Efp;
  int32_t *status;
  const double *table;      // T_t[rows], dC_t[S][rows]
  int rows;
  int ascending;            // T_t strictly increasing: bisection finds the same interval
  unsigned long long *acc;  // [0] integrated cells
};

// The interval the reference's scan ends on (the last i with
// T_t[i] <= T <= T_t[i+1]), or -1 if none.  For a strictly increasing table
// that is the largest i with T_t[i] <= T, capped at rows-2 (T on a node k
// sits in intervals k-1 and k; k wins), so bisection gives it in log2(rows)
// probes.  Any other table is scanned in full.
__device__ __forceinline__ int table_interval(const double *Tt, int rows, int ascending, double T)
{
  if (ascending) {
    if (!(T >= Tt[0] && T <= Tt[rows - 1])) return -1;
    int lo = 0, hi = rows - 1;      // Tt[lo] <= T; the answer lies in [lo, hi]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (Tt[mid] <= T) lo = mid;
      else hi = mid - 1;
    }
    return lo < rows - 1 ? lo : rows - 2;
  }
  int last = -1;
  for (int i = 0; i + 1 < rows; ++i)
    if (T >= Tt[i] && T <= Tt[i + 1]) last = i;
  return last;
}

constexpr int TABLE_THREADS = 256;
constexpr int TABLE_LDS_BYTES = 32768;        // a table up to this size is staged in LDS

// A grid of as many blocks as fit on the CUs strides over the cells, so the
// table is staged into LDS once per block (not once per 256 cells) and the
// integrated-cell count takes one global atomic per block.
template <bool TAB_LDS>
__global__ void __launch_bounds__(TABLE_THREADS) table_kernel(TableArgs a)
{
  extern __shared__ double tab_lds[];         // T_t[rows], dC_t[S][rows] when TAB_LDS
  __shared__ double W[S], href[S];
  __shared__ unsigned long long n_integ;
  if (threadIdx.x < S) {
    W[threadIdx.x] = a.mech->W[threadIdx.x];
    href[threadIdx.x] = a.mech->h_ref[threadIdx.x];
  }
  if (threadIdx.x == 0) n_integ = 0;
  if (TAB_LDS)
    for (int k = threadIdx.x; k < (S + 1) * a.rows; k += TABLE_THREADS) tab_lds[k] = a.table[k];
  __syncthreads();
  const double *Tt = TAB_LDS ? tab_lds : a.table;
  const double *dC = Tt + a.rows;
  unsigned long long cnt = 0;
  const long long stride = (long long)gridDim.x * TABLE_THREADS;
  // Waves walk the allocation in 64-cell slices aligned with it (from the
  // first inner cell's slice to the last inner cell) and skip ghost cells:
  // with ghost layers a row of inner cells starts off a cache line, and
  // walking inner cells would make every wave touch one line more per field.
  for (long long off = a.g.pad0 + (long long)blockIdx.x * TABLE_THREADS + threadIdx.x;
       off < a.g.pad1; off += stride) {
    if (!inner_at(a.g, off)) continue;
    const double T = a.T[off];
    const int8_t bcv = a.bc ? a.bc[off] : (int8_t)0;
    if (!(T >= a.t_gate && bcv == 0)) {
#pragma unroll
      for (int s = 0; s < S; ++s) a.Yp[s][off] = 0.0;
      a.Efp[off] = 0.0;
      if (a.status) a.status[off] = NRG_CELL_GATED;
      continue;
    }
    ++cnt;
    double cin[S], c[S];
#pragma unroll
    for (int s = 0; s < S; ++s) c[s] = cin[s] = a.Y[s][off];
    const int i = table_interval(Tt, a.rows, a.ascending, T);
    if (i >= 0) {
      const double t0 = Tt[i], t1 = Tt[i + 1];
#pragma unroll
      for (int s = 0; s < S; ++s)
        c[s] = (dC[s * a.rows + i + 1] * (T - t0)) / (t1 - t0) +
               (dC[s * a.rows + i] * (T - t1)) / (t0 - t1);
    }
    if (T < Tt[0]) {
#pragma unroll
      for (int s = 0; s < S; ++s) c[s] = 0.0;
    }
    if (T > Tt[a.rows - 1]) {
#pragma unroll
      for (int s = 0; s < S; ++s) c[s] = dC[s * a.rows + a.rows - 1];
    }
    double ef = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      a.Yp[s][off] = (c[s] * W[s]) / a.dt;
      a.Yp2[s][off] = c[s] - cin[s];
      ef = ef - (href[s] * c[s]) / a.dt;
    }
    a.Efp[off] = ef;
    if (a.status) a.status[off] = NRG_CELL_OK;
  }
  if (cnt) atomicAdd(&n_integ, cnt);
  __syncthreads();
  if (threadIdx.x == 0 && n_integ) atomicAdd(a.acc, n_integ);
}

// ---------------------------------------------------------------------------
// Low-Mach (FDS) coupling (fds_low_mach_solver.f90, table_approximated_real_gas.
// f90:570-677).  Per-cell, HBM-bound; one thread per offset of the aligned
// walk, the mechanism staged in LDS per block.
// ---------------------------------------------------------------------------
constexpr int LM_THREADS = 256;

struct LowMachEosArgs {
  GridIdx g;
  double dt;
  int predictor, ndim;
  const MechCoef *mech;
  const double *rho, *Y[S], *v[3], *dp;
  const int8_t *bc;
  double *p_stat, *p_stat_old, *T, *mmm;
  double *p, *cp, *gamma, *h_s, *v_s, *e_i, *E_f;   // optional
};

__global__ void __launch_bounds__(LM_THREADS) low_mach_eos_kernel(LowMachEosArgs a)
{
  __shared__ MechCoef mc;
  stage_mech(mc, a.mech);
  at_inner_cell(a.g, [&](long long off) {
    if (a.bc && a.bc[off] != 0) return;
    double X[S], avg = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      X[s] = a.Y[s][off];
      avg = avg + X[s] / mc.W[s];
    }
    const double W = 1.0 / avg;
    a.mmm[off] = W;
#pragma unroll
    for (int s = 0; s < S; ++s) X[s] = (X[s] * W) / mc.W[s];
    double ps;
    if (a.predictor) {
      const double p0 = a.p_stat[off];
      a.p_stat_old[off] = p0;
      ps = p0 + a.dp[off] * a.dt;
    } else {
      ps = 0.5 * ((a.p_stat[off] + a.p_stat_old[off]) + a.dp[off] * a.dt);
    }
    a.p_stat[off] = ps;
    const double rho = a.rho[off];
    const double T = ((ps / rho) / R_GAS_J) * W;
    a.T[off] = T;
    double cp = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) cp = cp + cp_molar(mc, T, s) * X[s];
    const double g = cp / (cp - R_GAS_J);
    if (a.gamma) a.gamma[off] = g;
    if (a.h_s) {
      double hT = 0.0, h0 = 0.0;
#pragma unroll
      for (int s = 0; s < S; ++s) hT = hT + h_molar(mc, T, s) * X[s];
#pragma unroll
      for (int s = 0; s < S; ++s) h0 = h0 + mc.h_ref[s] * X[s];
      a.h_s[off] = (hT - h0) / W;
    }
    if (a.cp) a.cp[off] = cp;
    if (a.p) a.p[off] = ps;
    if (a.v_s) a.v_s[off] = sqrt((g * ps) / rho);
    const double ei = (ps / rho) / (g - 1.0);
    if (a.e_i) a.e_i[off] = ei;
    if (a.E_f && a.ndim > 0) {
      const double vd = a.v[a.ndim - 1][off];
      a.E_f[off] = ei + 0.5 * (vd * vd);
    }
  });
}

struct LowMachSpeciesArgs {
  GridIdx g;
  double dt;
  int corrector;
  const double *rho_a, *Y_a[S], *rho_b, *Y_b[S];
  const double *rhs[S], *Yd[S], *Yc[S];   // rhs[0], Yd[0], Yc[0] NULL: term absent
  const int8_t *bc;
  double *rho_out, *Y_out[S];
};

__global__ void __launch_bounds__(LM_THREADS) low_mach_species_kernel(LowMachSpeciesArgs a)
{
  at_inner_cell(a.g, [&](long long off) {
    if (a.bc && a.bc[off] != 0) return;
    double r[S];
    const double ra = a.rho_a[off];
    const double rb = a.corrector ? a.rho_b[off] : 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double f = a.rhs[0] ? a.rhs[s][off] : 0.0;
      if (!a.corrector) {
        r[s] = ra * a.Y_a[s][off] + f * a.dt;
        if (a.Yd[0]) r[s] = r[s] + a.Yd[s][off] * a.dt;
        if (a.Yc[0]) r[s] = r[s] + a.Yc[s][off] * a.dt;
      } else {
        r[s] = 0.5 * (ra * a.Y_a[s][off] + rb * a.Y_b[s][off]) + f * (0.5 * a.dt);
        if (a.Yd[0]) r[s] = r[s] + (0.5 * a.Yd[s][off]) * a.dt;
        if (a.Yc[0]) r[s] = r[s] + (0.5 * a.Yc[s][off]) * a.dt;
      }
    }
    double sum = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      r[s] = dmax(r[s], 0.0);
      sum = sum + r[s];
    }
    if (sum > DBL_MIN) {
      a.rho_out[off] = sum;
#pragma unroll
      for (int s = 0; s < S; ++s) a.Y_out[s][off] = r[s] / sum;
    }
  });
}

struct LowMachDivArgs {
  GridIdx g;
  const MechCoef *mech;
  const double *rho, *T, *Y[S], *mmm, *Efc, *Yc[S];
  const int8_t *bc;
  double *div;
};

__global__ void __launch_bounds__(LM_THREADS) low_mach_div_kernel(LowMachDivArgs a)
{
  __shared__ MechCoef mc;
  stage_mech(mc, a.mech);
  at_inner_cell(a.g, [&](long long off) {
    if (a.bc && a.bc[off] != 0) return;
    const double W = a.mmm[off], rho = a.rho[off], T = a.T[off];
    double d = 0.0;
    d = d + a.Efc[off];
    double cp = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) cp = cp + cp_molar(mc, T, s) * ((a.Y[s][off] * W) / mc.W[s]);
    d = d / (((rho * cp) * T) / W);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double hs = (h_molar(mc, T, s) - mc.h_ref[s]) / mc.W[s];
      d = d + (((W / mc.W[s]) / rho) - hs / (((rho * cp) * T) / W)) * a.Yc[s][off];
    }
    a.div[off] = d;
  });
}

// ---------------------------------------------------------------------------
// Stiffness-ordered queue: bin the inner cells by expected cost, most
// expensive bin first.  A block bins a contiguous run of BIN_RUN cells (one
// global atomic per bin and block); within a bin, the cells of a run come in
// roughly increasing index order.
// ---------------------------------------------------------------------------
constexpr int NBIN = 64;
constexpr int BIN_THREADS = 256;
constexpr int BIN_RUN = 32 * BIN_THREADS;

struct BinArgs {
  GridIdx g;
  const double *T;
  const int8_t *bc;
  double t_gate;
  const int32_t *hist;              // NSTEP of the previous call, or NULL (use T)
  unsigned long long *count;        // [NBIN]
  unsigned long long *cursor;       // [NBIN] first queue position of each bin
  const int32_t *in_list;           // cells to order (NULL: all inner cells 0..npos-1)
  long long npos;                   // number of cells to order
  int32_t *list;                    // [npos] the ordered cells
};

// 0 = gated (no integration); otherwise grows with the expected cost.
NRG_DI int cost_bin(const BinArgs &a, long long q)
{
  const long long off = cell_offset(a.g, q);
  const int8_t b = a.bc ? a.bc[off] : (int8_t)0;
  const double T = a.T[off];
  if (!(T >= a.t_gate && b == 0)) return 0;
  if (a.hist) {
    const int n = max(a.hist[q], 0);
    const int k = n <= 31 ? n : 31 + (n - 28) / 4;     // unit bins to 31 steps, then 4-step bins
    return 1 + min(NBIN - 2, k);
  }
  const double x = (T - a.t_gate) * (1.0 / 50.0);                     // first call: 50 K bins
  return 1 + (x < (double)(NBIN - 2) ? (int)x : NBIN - 2);
}

// Histogram of the block's run into LDS.
NRG_DI void bin_histogram(const BinArgs &a, unsigned int *h)
{
  if (threadIdx.x < NBIN) h[threadIdx.x] = 0u;
  __syncthreads();
  const long long p0 = (long long)blockIdx.x * BIN_RUN + threadIdx.x;
  for (int k = 0; k < BIN_RUN / BIN_THREADS; ++k) {
    const long long p = p0 + (long long)k * BIN_THREADS;
    if (p < a.npos) atomicAdd(&h[cost_bin(a, a.in_list ? (long long)a.in_list[p] : p)], 1u);
  }
  __syncthreads();
}

__global__ void __launch_bounds__(BIN_THREADS) bin_count_kernel(BinArgs a)
{
  __shared__ unsigned int h[NBIN];
  bin_histogram(a, h);
  if (threadIdx.x < NBIN && h[threadIdx.x]) atomicAdd(&a.count[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// cursor[b] = number of cells in the bins more expensive than b (one wave).
__global__ void bin_offsets_kernel(BinArgs a)
{
  if (threadIdx.x == 0) {
    unsigned long long pos = 0;
    for (int k = NBIN - 1; k >= 0; --k) {
      a.cursor[k] = pos;
      pos += a.count[k];
    }
  }
}

// Each block claims its bins' slots with one atomic per bin, then places its
// run's cells (same run as bin_count_kernel, binned again).
__global__ void __launch_bounds__(BIN_THREADS) bin_scatter_kernel(BinArgs a)
{
  __shared__ unsigned int h[NBIN];
  __shared__ unsigned long long base[NBIN];
  bin_histogram(a, h);
  if (threadIdx.x < NBIN) {
    base[threadIdx.x] = h[threadIdx.x] ? atomicAdd(&a.cursor[threadIdx.x], (unsigned long long)h[threadIdx.x]) : 0ull;
    h[threadIdx.x] = 0u;
  }
  __syncthreads();
  const long long p0 = (long long)blockIdx.x * BIN_RUN + threadIdx.x;
  for (int k = 0; k < BIN_RUN / BIN_THREADS; ++k) {
    const long long p = p0 + (long long)k * BIN_THREADS;
    if (p < a.npos) {
      const long long q = a.in_list ? (long long)a.in_list[p] : p;
      const int b = cost_bin(a, q);
      const unsigned int rank = atomicAdd(&h[b], 1u);
      a.list[base[b] + rank] = (int32_t)q;
    }
  }
}

// ---------------------------------------------------------------------------
// Learned step: features -> network -> gate -> FP64 fallback list
// ---------------------------------------------------------------------------
constexpr int NIN = S + 3, NOUT = S + 1;

enum { SUR_FALLBACK = 0, SUR_ACCEPT, SUR_GATED, SUR_R_DOMAIN, SUR_R_ELEM, SUR_R_NEG, SUR_R_ENERGY,
       SUR_R_CONF, SUR_R_RATE, SUR_FROZEN, SUR_CAND, SUR_RECHECK, SUR_SLOTS = 12 };
constexpr int SUR_NREASONS = SUR_R_RATE - SUR_R_DOMAIN + 1;

struct SurArgs {
  GridIdx g;
  double dt, t_gate;
  const MechCoef *mech;
  const double *T, *p, *rho;
  const double *Y[S];
  const int8_t *bc;
  float in_shift[NIN], in_scale[NIN], out_shift[NOUT], out_scale[NOUT];
  float y_floor;
  double t_min, t_max, p_min, p_max;
  int dt_ok;
  double elem_tol, elem_floor, y_neg_tol, energy_tol, dT_max, dT_freeze, dY_freeze, rate_tol;
  int ne;
  double E[NRG_SUR_MAX_ELEMENTS][S];
  const int32_t *cand;        // NULL: position p is inner cell p; else inner cell cand[p]
  long long npos;             // positions the network, rate check and gate visit
  int32_t *screen_out;        // the screen: where the cells it does not freeze go
  int screen_slot;            // ... and the counter that numbers them (SUR_CAND, SUR_FALLBACK)
  int screen_fp64;            // 1: the FP64 estimate decides every live cell
  int32_t *recheck;           // cells screen_kernel leaves to rate_screen_kernel ...
  const unsigned long long *recheck_n;   // ... and their number (acc[SUR_RECHECK])
  bf16 *X;                    // npos x MLP_IN_PAD
  const float *Ypred;         // npos x NOUT
  double *Yp[S], *Yp2[S];
  double *Efp;
  int32_t *status;
  int32_t *list;              // fallback cells (inner index)
  int8_t *rate_bad;           // per inner cell: 1 = the rate-consistency test failed
  unsigned long long *acc;    // SUR_* counters
};

NRG_DI bool gated_cell(const SurArgs &a, long long off)
{
  const int8_t bcv = a.bc ? a.bc[off] : (int8_t)0;
  return !(a.T[off] >= a.t_gate && bcv == 0);
}

__global__ void __launch_bounds__(256) featurize_kernel(SurArgs a)
{
  const long long pos = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= a.npos) return;
  const long long q = a.cand ? (long long)a.cand[pos] : pos;
  const long long off = cell_offset(a.g, q);
  bf16 *x = a.X + pos * MLP_IN_PAD;
  float f[MLP_IN_PAD];
#pragma unroll
  for (int i = 0; i < MLP_IN_PAD; ++i) f[i] = 0.0f;
  if (gated_cell(a, off)) {
    // the FP64 step's handling of a skipped cell (chemical_kinetics_solver.f90:341-346)
#pragma unroll
    for (int s = 0; s < S; ++s) a.Yp[s][off] = 0.0;
    a.Efp[off] = 0.0;
    if (a.status) a.status[off] = NRG_CELL_GATED;
  } else {
    f[0] = (float)a.T[off];
    f[1] = (float)log(a.p[off]);
#pragma unroll
    for (int s = 0; s < S; ++s) f[2 + s] = logf((float)a.Y[s][off] + a.y_floor);
    f[S + 2] = (float)log(a.dt);
#pragma unroll
    for (int i = 0; i < NIN; ++i) f[i] = (f[i] - a.in_shift[i]) * a.in_scale[i];
  }
#pragma unroll
  for (int i = 0; i < MLP_IN_PAD; ++i) x[i] = (bf16)f[i];
}

// Solve the n x n system M z = b in place (n <= 4, partial pivoting).  Loops
// run to the compile-time bound with n as a predicate, so M and b stay in
// registers (row swaps are selects).
NRG_DI bool small_solve(double (&M)[NRG_SUR_MAX_ELEMENTS][NRG_SUR_MAX_ELEMENTS],
                        double (&b)[NRG_SUR_MAX_ELEMENTS], int n)
{
  constexpr int NE = NRG_SUR_MAX_ELEMENTS;
  bool ok = true;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    if (k >= n) break;
    int piv = k;
    double best = fabs(M[k][k]);
#pragma unroll
    for (int i = k + 1; i < NE; ++i)
      if (i < n && fabs(M[i][k]) > best) {
        piv = i;
        best = fabs(M[i][k]);
      }
    if (best == 0.0) ok = false;
#pragma unroll
    for (int i = k + 1; i < NE; ++i) {
      const bool sw = (piv == i);
#pragma unroll
      for (int j = 0; j < NE; ++j) {
        const double t = M[k][j];
        M[k][j] = sw ? M[i][j] : M[k][j];
        M[i][j] = sw ? t : M[i][j];
      }
      const double t = b[k];
      b[k] = sw ? b[i] : b[k];
      b[i] = sw ? t : b[i];
    }
#pragma unroll
    for (int i = k + 1; i < NE; ++i) {
      if (i >= n) break;
      const double f = M[i][k] / M[k][k];
#pragma unroll
      for (int j = k; j < NE; ++j) M[i][j] -= f * M[k][j];
      b[i] -= f * b[k];
    }
  }
#pragma unroll
  for (int k = NE - 1; k >= 0; --k) {
    if (k >= n) continue;
    double s = b[k];
#pragma unroll
    for (int j = k + 1; j < NE; ++j)
      if (j < n) s -= M[k][j] * b[j];
    b[k] = s / M[k][k];
  }
  return ok;
}

// The per-reason counts are summed per block in LDS and added to the global
// counters once per block.
// The rate-consistency test of the gate, in a kernel of its own (the gate
// kernel's registers leave no room for the rate constants).  For every live
// cell the answer the gate would give before its conservation checks (the
// network's increment, or zero for a cell it freezes) is compared with the
// explicit estimate of the step, C_s(t+dt) - C_s = omega_s(C, T) dt, from the
// FP64 rate constants (EstimateMath transcendentals) and mass action at the
// cell's own state;
// rate_bad[q] = 1 where some active species differs by more than rate_tol in
// mass fraction.  As in rates_kernel, waves 0-1 and 2-3 of a block take the
// two halves of the reactions for the same 128 cells (one half per thread
// fits in half the registers) and add their parts of omega in LDS.
constexpr int RCHECK_CELLS = 128;

__global__ void __launch_bounds__(256) rate_check_kernel(SurArgs a)
{
  __shared__ MechCoef mc;
  __shared__ double part[RCHECK_CELLS][SA];
  stage_mech(mc, a.mech);
  const int group = __builtin_amdgcn_readfirstlane(threadIdx.x / RCHECK_CELLS);
  const int local = threadIdx.x % RCHECK_CELLS;
  const long long pos = (long long)blockIdx.x * RCHECK_CELLS + local;
  const long long q = pos < a.npos ? (a.cand ? (long long)a.cand[pos] : pos) : 0;
  const long long off = pos < a.npos ? cell_offset(a.g, q) : 0;
  const bool live = pos < a.npos && !gated_cell(a, off);
  double y[S], C[S], Zd[SA];
  double rho = 1.0;
  if (live) {
    rho = a.rho[off];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      y[s] = a.Y[s][off];
      C[s] = (y[s] * rho) / mc.W[s];
    }
    if (group == 0) explicit_rates<EstimateMath, 0, R / 2>(mc, a.T[off], C, Zd);
    else explicit_rates<EstimateMath, R / 2, R>(mc, a.T[off], C, Zd);
    if (group == 1) {
#pragma unroll
      for (int i = 0; i < SA; ++i) part[local][i] = Zd[i];
    }
  }
  __syncthreads();
  if (group != 0 || pos >= a.npos) return;
  int8_t bad = 0;
  if (live) {
    const float *yp = a.Ypred + pos * NOUT;
    const double dT = (double)yp[0] * a.out_scale[0] + a.out_shift[0];
    double dy[S], dymax = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      dy[s] = (double)yp[1 + s] * a.out_scale[1 + s] + a.out_shift[1 + s];
      dymax = fmax(dymax, fabs(dy[s]));
    }
    const bool frozen = a.dT_freeze > 0.0 && fabs(dT) <= a.dT_freeze && dymax <= a.dY_freeze;
#pragma unroll
    for (int i = 0; i < SA; ++i) {
      const int s = kSp.full[i];
      const double est = (((Zd[i] + part[local][i]) * a.dt) * mc.W[s]) / rho;
      if (!(fabs(est - (frozen ? 0.0 : dy[s])) <= a.rate_tol)) bad = 1;
    }
  }
  a.rate_bad[pos] = bad;
}

// The explicit-rate screen: the learned step's freeze test without the
// network.  For every live cell it forms the explicit estimate of the step,
// dY_s = omega_s(C, T) dt W_s / rho (rate constants and mass action as in
// rate_check_kernel), and the temperature change that estimate implies at the
// cell's frozen cv, dT = -(sum_s e_s(T) dY_s / W_s) / cv with
// cv = sum_s (Y_s / W_s)(cp_s(T) - R).  A cell with every |dY_s| <= dY_freeze
// and |dT| <= dT_freeze is taken as inert and gets the zero increment (the FP64
// step's outputs for a cell whose concentrations do not change).  Every other
// live cell is appended to screen_out: the network's candidates, or, for the
// method without the network, the FP64 fallback list.  Gated cells get the
// FP64 step's outputs for a skipped cell (chemical_kinetics_solver.f90:341-346).
//
// Two kernels share the work, and the decision is always the one the FP64
// estimate gives.  screen_kernel visits every inner cell and evaluates the
// estimate with FP64 exponents and FP32 exponentials (the transcendentals
// that make up most of the FP64 estimate's cost), carrying a bound on its
// distance from the FP64 estimate.  Where the bound settles the decision it
// decides; the other live cells go to a list, which rate_screen_kernel then
// decides with the FP64 estimate itself.

NRG_DI void freeze_outputs(const SurArgs &a, long long off)
{
#pragma unroll
  for (int s = 0; s < S; ++s) {
    a.Yp[s][off] = 0.0;
    a.Yp2[s][off] = 0.0;
  }
  a.Efp[off] = 0.0;
  if (a.status) a.status[off] = NRG_CELL_SURROGATE;
}

// Appends the wave's lanes with `take` to out[] at the counter *ctr.
NRG_DI void wave_append(bool take, int lane, long long q, int32_t *out, unsigned long long *ctr)
{
  const unsigned long long mask = __ballot(take ? 1 : 0);
  if (!mask) return;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(ctr, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader, WAVE);
  if (take) out[base + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)q;
}

// The FP64 decision on the cells screen_kernel lists (recheck[0, *recheck_n)).
// Waves 0-1 and 2-3 of a block take the two halves of the reactions for the
// same 128 cells, as in rate_check_kernel; blocks stride over the list.
__global__ void __launch_bounds__(256) rate_screen_kernel(SurArgs a)
{
  const long long npos = (long long)*a.recheck_n;
  if ((long long)blockIdx.x * RCHECK_CELLS >= npos) return;   // the whole block
  __shared__ MechCoef mc;
  __shared__ double part[RCHECK_CELLS][SA];
  __shared__ double rW[S];   // 1 / W_s: the estimate multiplies where the step divides
  __shared__ unsigned long long blk[SUR_SLOTS];
  if (threadIdx.x < SUR_SLOTS) blk[threadIdx.x] = 0ull;
  if (threadIdx.x < S) rW[threadIdx.x] = 1.0 / a.mech->W[threadIdx.x];
  stage_mech(mc, a.mech);
  const int group = __builtin_amdgcn_readfirstlane(threadIdx.x / RCHECK_CELLS);
  const int lane = threadIdx.x & (WAVE - 1);
  const int local = threadIdx.x % RCHECK_CELLS;
  for (long long p0 = (long long)blockIdx.x * RCHECK_CELLS; p0 < npos;
       p0 += (long long)gridDim.x * RCHECK_CELLS) {
    // the coefficients in LDS are read where they are used, not held in
    // registers across the loop
    asm volatile("" ::: "memory");
    const long long pos = p0 + local;
    const bool live = pos < npos;   // listed cells are live
    const long long q = live ? (long long)a.recheck[pos] : 0;
    const long long off = live ? cell_offset(a.g, q) : 0;
    double C[S], Zd[SA];
    double rho = 1.0, T = 0.0;
    if (live) {
      rho = a.rho[off];
      T = a.T[off];
#pragma unroll
      for (int s = 0; s < S; ++s) C[s] = (a.Y[s][off] * rho) * rW[s];
      if (group == 0) explicit_rates<EstimateMath, 0, R / 2>(mc, T, C, Zd);
      else explicit_rates<EstimateMath, R / 2, R>(mc, T, C, Zd);
      if (group == 1) {
#pragma unroll
        for (int i = 0; i < SA; ++i) part[local][i] = Zd[i];
      }
    }
    __syncthreads();
    if (group == 0) {
      bool frozen = false;
      if (live) {
        double dymax = 0.0, de = 0.0, cv = 0.0;
        const double rrho = 1.0 / rho;
#pragma unroll
        for (int i = 0; i < SA; ++i) {
          const int s = kSp.full[i];
          const double est = (((Zd[i] + part[local][i]) * a.dt) * mc.W[s]) * rrho;
          dymax = fmax(dymax, fabs(est));
          de += e_molar(mc, T, s) * (est * rW[s]);
        }
#pragma unroll
        for (int s = 0; s < S; ++s) cv += (C[s] * rrho) * (cp_molar(mc, T, s) - R_GAS_J);
        const double dT = -de / cv;
        frozen = a.dT_freeze > 0.0 && dymax <= a.dY_freeze && fabs(dT) <= a.dT_freeze;
        if (frozen) freeze_outputs(a, off);
      }
      wave_append(live && !frozen, lane, q, a.screen_out, &a.acc[a.screen_slot]);
      const unsigned long long c_frz = __popcll(__ballot(live && frozen ? 1 : 0));
      if (lane == 0 && c_frz) {
        atomicAdd(&blk[SUR_ACCEPT], c_frz);
        atomicAdd(&blk[SUR_FROZEN], c_frz);
      }
    }
    __syncthreads();   // part[] is rewritten by the next slice
  }
  if (threadIdx.x < SUR_SLOTS && (int)threadIdx.x != a.screen_slot && blk[threadIdx.x])
    atomicAdd(&a.acc[threadIdx.x], blk[threadIdx.x]);
}

// ln x for x > 0 (normal or subnormal): frexp's exponent plus an FP32 log2 of
// the mantissa m in [0.5, 1).  Rounding m to FP32 moves log2 m by at most
// 2^-24 / ln 2 and log2f is within 2 ulp (2^-23 on [-1, 0]), so
// |ln_m32(x) - ln x| < 2^-22, bounded below by SCR_LN_ERR.
constexpr double SCR_LN_ERR = 0x1p-20;
constexpr double SCR_LN2 = 0.69314718055994530942;
constexpr double SCR_LOG2E = 1.4426950408889634074;
constexpr double SCR_LN10 = 2.30258509299404568402;

NRG_DI double ln_m32(double x)
{
  int e;
  const double m = frexp(x, &e);
  return ((double)e + (double)log2f((float)m)) * SCR_LN2;
}

// v = e^x as FP32 exp2 of an FP64 exponent x that is within ex of the exact
// one, and err, a bound on |v - e^x_exact|: rounding x log2(e) to FP32 moves
// the value by at most |x| 2^-24 relative and exp2f is within 2 ulp (2^-22),
// here taken as 2^-23 |x| and 2^-21; e^ex - 1 <= 1.01 ex for ex <= 0.01; and
// FP32 results below its normal range add at most 2^-125 (exponents below
// -100 give the value 0 with that bound).  Returns false where the value
// overflows FP32, is NaN, or the bound exceeds 1 %.
NRG_DI bool exp_term(double x, double ex, double &v, double &err)
{
  if (x == -INFINITY) {   // a zero rate constant or third-body concentration
    v = 0.0;
    err = 0.0;
    return true;
  }
  if (x < -100.0) {       // below FP32's range: e^(x + ex) < e^-99 < 2^-125 for ex <= 1
    v = 0.0;
    err = 0x1p-125;
    return ex <= 1.0;
  }
  const double rel = ex + fabs(x) * 0x1p-23 + 0x1p-21;
  const float y = exp2f((float)(x * SCR_LOG2E));
  v = (double)y;
  err = rel * 1.01 * v + 0x1p-125;
  return y <= 3.0e38f && rel <= 0.01;
}

// Stoichiometry of species s in reaction r: right minus left count, and the
// number of times it appears (the bound adds a term's error once per
// appearance, as the right-hand side adds the term).
constexpr int nu_net(int r, int s)
{
  int n = 0;
  for (int c = 0; c < 3; ++c) {
    if (c < topo::nr[r] && topo::right[r][c] == s) ++n;
    if (c < topo::nl[r] && topo::left[r][c] == s) --n;
  }
  return n;
}
constexpr int nu_occ(int r, int s)
{
  int n = 0;
  for (int c = 0; c < 3; ++c) {
    if (c < topo::nr[r] && topo::right[r][c] == s) ++n;
    if (c < topo::nl[r] && topo::left[r][c] == s) ++n;
  }
  return n;
}

// Screen estimate of one cell: FP64 exponents (ln k_f, ln K_p, ln [M], ln C
// from ln_m32) and FP32 exponentials.  acc[i] = sum over the reactions of
// nu (F - B) for active species i, bnd[i] a bound on its distance from the
// FP64 estimate's sum.  The reactions' arithmetic follows reaction_k (rate
// constants, chemical_kinetics_solver.f90:394-485) and rhs_reaction (mass
// action, :507-536) in logarithms:
//   ln k_f = ln A + b ln T - E/(R T)
//   fall-off: ln k_f = ln k_inf - ln(1 + 1/P_r) [+ ln F for Troe, with
//             ln F = ln F_cent / (1 + x^2)], ln P_r = ln k_0 + ln [M] - ln k_inf
//   ln k_r = ln k_f - ln K_p + d_nu ln(R_u T / P_atm)
// Returns false (the FP64 path decides the cell) where a term leaves FP32's
// range, a bound exceeds 1 % or a quantity the logarithms need is not
// positive (a negative concentration or [M], a Troe denominator near zero).
NRG_DI bool screen_estimate(const MechCoef &mc, double Tin, const double (&C)[S], double (&acc)[SA],
                            double (&bnd)[SA], double (&H)[S])
{
  const double T = dmin(Tin, 10000.0);
  const double lnT = log(T), rT = 1.0 / T;
  const double ln_x = mc.ln_patm_r - lnT;   // ln(P_atm / (R_u T))
  bool ok = true;
  double lnC[S], G[S];                    // G_s = s_s/R - h_s/(R T): ln K_p = sum nu G
  unsigned absent = 0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    ok = ok && C[s] >= 0.0;
    absent |= (C[s] == 0.0) ? 1u << s : 0u;
    lnC[s] = C[s] > 0.0 ? ln_m32(C[s]) : 0.0;
    H[s] = h_molar<EstimateMath>(mc, T, s);
    G[s] = s_molar<EstimateMath>(mc, T, s) * (1.0 / R_GAS_J) - (H[s] * (1.0 / R_GAS_J)) * rT;
    // each species' values are formed here, in turn: left to the compiler,
    // every species' coefficients were loaded at once (twice the registers)
    asm volatile("" : "+v"(G[s]), "+v"(lnC[s]), "+v"(H[s]));
  }
#pragma unroll
  for (int i = 0; i < SA; ++i) {
    acc[i] = 0.0;
    bnd[i] = 0.0;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool fwd = (absent & side_mask(r, 0)) == 0;
    const bool rev = has_reverse(r) && (absent & side_mask(r, 1)) == 0;
    if (!__any((fwd || rev) ? 1 : 0)) continue;
    const int rt = topo::rtype[r];
    double m = 0.0, lm = 0.0;
    if (uses_m(r)) {
#pragma unroll
      for (int s = 0; s < S; ++s) m = m + C[s] * mc.eff[r][s];
      lm = m > 0.0 ? ln_m32(m) : -INFINITY;
    }
    // ln k_f and a bound on its error
    // (rok: this reaction's checks, which matter only where the cell has a
    // term of it)
    double lkf = mc.lnA[r] + mc.beta[r] * lnT + mc.mE_R[r] * rT, ekf = 0.0;
    bool rok = true;
    if (rt == 2 || rt == 3 || rt == 7) {
      const double lk0 = mc.lnA_low[r] + mc.beta_low[r] * lnT + mc.mElow_R[r] * rT;
      // P_r = 0 ([M] = 0: every efficiency of the cell's species zero, or
      // k_0 = 0): k_f = k_inf * 0 * F = 0 exactly in the FP64 estimate too
      const bool zero_pr = m == 0.0 || lk0 == -INFINITY;
      const double lpr = zero_pr ? 0.0 : lk0 + lm - lkf;
      rok = m >= 0.0 && lpr > -1e4 && lpr < 1e4;
      // ln(P_r / (1 + P_r)) = -softplus(-ln P_r) in FP32: the argument's
      // rounding, expf and log1pf (2 ulp each) and the sum stay within
      // (|ln P_r| + 1) 2^-23 + 2^-21; ln [M] adds SCR_LN_ERR
      const float z = (float)(-lpr);
      const float sp = z > 0.0f ? z + log1pf(expf(-z)) : log1pf(expf(z));
      lkf = lkf - (double)sp;
      ekf = SCR_LN_ERR + (fabs(lpr) + 1.0) * 0x1p-23 + 0x1p-21;
      if (rt == 3) {
        double e1, e2, e3, r1, r2, r3;
        rok = exp_term(-(T / mc.troe[1][r]), 0.0, e1, r1) && rok;
        rok = exp_term(-(T / mc.troe[2][r]), 0.0, e2, r2) && rok;
        rok = exp_term(-(mc.troe[3][r] * rT), 0.0, e3, r3) && rok;
        const double fcent = (mc.one_m_alpha[r] * e1 + mc.troe[0][r] * e2) + e3;
        const double efc = fabs(mc.one_m_alpha[r]) * r1 + fabs(mc.troe[0][r]) * r2 + r3;
        // (with P_r = 0 the FP64 estimate multiplies 0 by F_cent^(-1/0.14),
        // finite for F_cent above 1e-40)
        rok = rok && fcent > 0.0 && efc <= 0.001 * fcent && (!zero_pr || fcent > 1e-40);
        const double lnfc = ln_m32(fcent > 0.0 ? fcent : 1.0);
        const double elnfc = SCR_LN_ERR + 1.01 * efc / (fcent > 0.0 ? fcent : 1.0);
        const double lf = lnfc * (1.0 / SCR_LN10);
        const double cc = -0.4 - 0.67 * lf, nn = 0.75 - 1.27 * lf;
        const double nume = lpr * (1.0 / SCR_LN10) + cc;
        const double den = nn - 0.14 * nume;
        const double xx = nume / den;
        // perturbations of the numerator and denominator of x from ln [M]
        // and ln F_cent, and the resulting bound on |x - x_FP64|; the
        // derivative of 1 / (1 + x^2) is at most 0.65 in magnitude
        const double elf = elnfc * (1.0 / SCR_LN10);
        const double dnum = SCR_LN_ERR * (1.0 / SCR_LN10) + 0.67 * elf;
        const double dden = 1.27 * elf + 0.14 * dnum;
        rok = rok && (zero_pr || (fabs(den) > 0.0625 && fabs(den) > 2.0 * dden));
        const double adn = fabs(den) > 2.0 * dden ? fabs(den) : 1.0;
        const double exx = (dnum * adn + fabs(nume) * dden) / (adn * (adn - dden));
        lkf = lkf + lnfc / (1.0 + xx * xx);
        ekf = ekf + fabs(lnfc) * 0.65 * exx + elnfc;
      }
      if (zero_pr) lkf = -INFINITY;
    }
    // forward and reverse terms of the mass action
    double xf = lkf, ef = ekf;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int sp = topo::left[r][c];
      if (c < topo::nl[r]) {
        xf = xf + ((sp == S) ? lm : lnC[sp < S ? sp : 0]);
        ef = ef + SCR_LN_ERR;
      }
    }
    ok = ok && (rok || !(fwd || rev));
    double F = 0.0, eF = 0.0;
    if (fwd) ok = exp_term(xf, ef, F, eF) && ok;
    double B = 0.0, eB = 0.0;
    if (has_reverse(r)) {
      double lnKp = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int sp = topo::right[r][c];
        if (c < topo::nr[r] && sp != S) lnKp = lnKp + G[sp < S ? sp : 0];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int sp = topo::left[r][c];
        if (c < topo::nl[r] && sp != S) lnKp = lnKp - G[sp < S ? sp : 0];
      }
      const double d_nu = (double)(topo::nr[r] - topo::nl[r]);
      double xb = (lkf - lnKp) - d_nu * ln_x, eb = ekf;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int sp = topo::right[r][c];
        if (c < topo::nr[r]) {
          xb = xb + ((sp == S) ? lm : lnC[sp < S ? sp : 0]);
          eb = eb + SCR_LN_ERR;
        }
      }
      if (rev) ok = exp_term(xb, eb, B, eB) && ok;
    }
#pragma unroll
    for (int i = 0; i < SA; ++i) {
      const int s = kSp.full[i];
      if (nu_net(r, s) != 0) acc[i] = acc[i] + (double)nu_net(r, s) * (F - B);
      if (nu_occ(r, s) != 0) bnd[i] = bnd[i] + (double)nu_occ(r, s) * (eF + eB);
    }
  }
  return ok;
}

enum { SCR_FROZEN = 0, SCR_PASS = 1, SCR_RECHECK = 2 };   // screen_kernel's outcome for a live cell

// The screen over every inner cell (a block per 256 cells, one cell per
// thread; a loop over slices would hold the expansions' FP64 constants in
// registers across it and halve the resident waves).  Outcomes: gated, frozen, passed on (screen_out), or
// listed for rate_screen_kernel (recheck).  With screen_fp64 every live cell
// is listed, so the FP64 estimate decides every cell.
__global__ void __launch_bounds__(256) screen_kernel(SurArgs a)
{
  __shared__ MechCoef mc;
  __shared__ double rW[S];
  __shared__ unsigned long long blk[SUR_SLOTS];
  if (threadIdx.x < SUR_SLOTS) blk[threadIdx.x] = 0ull;
  if (threadIdx.x < S) rW[threadIdx.x] = 1.0 / a.mech->W[threadIdx.x];
  stage_mech(mc, a.mech);
  const int lane = threadIdx.x & (WAVE - 1);
  const long long n = a.g.ninner;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = q < n;
  const long long off = valid ? cell_offset(a.g, q) : 0;
  // every input in one round trip to memory (the loads do not wait for the
  // gate test, as in gated_cell)
  double T = 0.0, rho = 1.0, y[S];
  int8_t bcv = 0;
  if (valid) {
    T = a.T[off];
    bcv = a.bc ? a.bc[off] : (int8_t)0;
    rho = a.rho[off];
#pragma unroll
    for (int s = 0; s < S; ++s) y[s] = a.Y[s][off];
  }
  const bool gated = valid && !(T >= a.t_gate && bcv == 0);
  const bool live = valid && !gated;
  int outcome = SCR_PASS;
  if (live) {
    double C[S], acc[SA], bnd[SA], H[S];
#pragma unroll
    for (int s = 0; s < S; ++s) C[s] = (y[s] * rho) * rW[s];
    const bool ok = screen_estimate(mc, T, C, acc, bnd, H);
    // the decision with the FP64 estimate's expressions, each side of each
    // threshold taken only where the bound puts the FP64 estimate there too
    const double rrho = 1.0 / rho;
    bool in = true, out = false;
    double de = 0.0, bde = 0.0, dabs = 0.0, cv = 0.0;
#pragma unroll
    for (int i = 0; i < SA; ++i) {
      const int s = kSp.full[i];
      const double est = ((acc[i] * a.dt) * mc.W[s]) * rrho;
      const double b = (((bnd[i] * a.dt) * mc.W[s]) * rrho) * (1.0 + 0x1p-40);
      in = in && fabs(est) + b <= a.dY_freeze;
      out = out || fabs(est) - b > a.dY_freeze;
      // e_s as e_molar forms it, from h_s with a5 * (1/T) where h_molar
      // divides: within 2^-50 (|a5| + |h_s|) of the FP64 estimate's e_s
      const double e = H[s] - R_GAS_J * T;
      de += e * (est * rW[s]);
      bde += fabs(e) * (b * rW[s]);
      const double ea = fabs(e) + 0x1p-10 * (fabs(mc.a[s][0][5]) + fabs(mc.a[s][1][5]) + fabs(H[s]));
      dabs += ea * ((fabs(est) + b) * rW[s]);
      asm volatile("" : "+v"(de), "+v"(bde), "+v"(dabs));   // in turn, as in screen_estimate
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      cv += (C[s] * rrho) * (cp_molar(mc, T, s) - R_GAS_J);
      asm volatile("" : "+v"(cv));
    }
    // bT bounds |dT - dT_FP64|: the bounds on the estimates, the two e_s
    // and the FP64 roundings of both sums (dabs)
    const double dT = -de / cv;
    const double bT = (bde + dabs * 0x1p-40) / cv + fabs(dT) * 0x1p-40;
    const bool sane = ok && cv > 0.0 && bT < INFINITY && fabs(dT) < INFINITY;
    in = in && fabs(dT) + bT <= a.dT_freeze;
    out = out || fabs(dT) - bT > a.dT_freeze;
    outcome = (!sane || a.screen_fp64) ? SCR_RECHECK : out ? SCR_PASS : in ? SCR_FROZEN : SCR_RECHECK;
    if (outcome == SCR_FROZEN) freeze_outputs(a, off);
  }
  if (gated) {
#pragma unroll
    for (int s = 0; s < S; ++s) a.Yp[s][off] = 0.0;
    a.Efp[off] = 0.0;
    if (a.status) a.status[off] = NRG_CELL_GATED;
  }
  wave_append(live && outcome == SCR_PASS, lane, q, a.screen_out, &a.acc[a.screen_slot]);
  wave_append(live && outcome == SCR_RECHECK, lane, q, a.recheck, &a.acc[SUR_RECHECK]);
  const unsigned long long c_frz = __popcll(__ballot(live && outcome == SCR_FROZEN ? 1 : 0));
  const unsigned long long c_gat = __popcll(__ballot(gated ? 1 : 0));
  if (lane == 0) {
    if (c_frz) {
      atomicAdd(&blk[SUR_ACCEPT], c_frz);
      atomicAdd(&blk[SUR_FROZEN], c_frz);
    }
    if (c_gat) atomicAdd(&blk[SUR_GATED], c_gat);
  }
  __syncthreads();
  if (threadIdx.x < SUR_SLOTS && (int)threadIdx.x != a.screen_slot && (int)threadIdx.x != SUR_RECHECK &&
      blk[threadIdx.x])
    atomicAdd(&a.acc[threadIdx.x], blk[threadIdx.x]);
}

template <bool RATE>
__global__ void __launch_bounds__(256) gate_kernel(SurArgs a)
{
  __shared__ MechCoef mc;
  __shared__ unsigned long long blk[SUR_SLOTS];
  if (threadIdx.x < SUR_SLOTS) blk[threadIdx.x] = 0ull;
  stage_mech(mc, a.mech);
  const int lane = threadIdx.x & (WAVE - 1);
  const long long pos = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = pos < a.npos;
  const long long q = valid ? (a.cand ? (long long)a.cand[pos] : pos) : 0;
  const long long off = valid ? cell_offset(a.g, q) : 0;
  const bool gated = valid && gated_cell(a, off);
  const bool live = valid && !gated;
  int reason = 0;   // 0 accept, else SUR_R_*
  if (live) {
    const double T = a.T[off], p = a.p[off], rho = a.rho[off];
    double y[S], dy[S];
#pragma unroll
    for (int s = 0; s < S; ++s) y[s] = a.Y[s][off];
    const float *yp = a.Ypred + pos * NOUT;
    const double dT = (double)yp[0] * a.out_scale[0] + a.out_shift[0];
    bool finite = isfinite(dT);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      dy[s] = (double)yp[1 + s] * a.out_scale[1 + s] + a.out_shift[1 + s];
      finite = finite && isfinite(dy[s]);
    }
    bool frozen = false;
    if (!finite || !(T >= a.t_min && T <= a.t_max && p >= a.p_min && p <= a.p_max) || !a.dt_ok) {
      reason = SUR_R_DOMAIN;
    } else {
      double dymax = 0.0;
#pragma unroll
      for (int s = 0; s < S; ++s) dymax = fmax(dymax, fabs(dy[s]));
      frozen = a.dT_freeze > 0.0 && fabs(dT) <= a.dT_freeze && dymax <= a.dY_freeze;
      if (!frozen && !(fabs(dT) <= a.dT_max)) reason = SUR_R_CONF;
      if (RATE && reason == 0 && a.rate_bad[pos]) {
        frozen = false;
        reason = SUR_R_RATE;
      }
    }
    if (frozen) {
      // inert: the FP64 step's outputs for a cell whose concentrations do not change
#pragma unroll
      for (int s = 0; s < S; ++s) {
        a.Yp[s][off] = 0.0;
        a.Yp2[s][off] = 0.0;
      }
      a.Efp[off] = 0.0;
      if (a.status) a.status[off] = NRG_CELL_SURROGATE;
    }
    if (frozen) reason = -1;
    // element balance of the raw increment: sum_s E_es dY_s / W_s  [mol/kg]
    double res[NRG_SUR_MAX_ELEMENTS] = {0.0, 0.0, 0.0, 0.0};
    double d[S];
    if (reason == 0) {
#pragma unroll
      for (int e = 0; e < NRG_SUR_MAX_ELEMENTS; ++e) {
        if (e >= a.ne) break;
        double r = 0.0, sc = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          r += a.E[e][s] * dy[s] / mc.W[s];
          sc += a.E[e][s] * fabs(dy[s]) / mc.W[s];
        }
        res[e] = r;
        if (fabs(r) > a.elem_tol * sc + a.elem_floor) reason = SUR_R_ELEM;
      }
    }
    if (reason == 0) {
      // project onto sum_s E_es dY_s / W_s = 0, weighting the correction by |dY_s|
      double M[NRG_SUR_MAX_ELEMENTS][NRG_SUR_MAX_ELEMENTS] = {};
#pragma unroll
      for (int s = 0; s < S; ++s) d[s] = fabs(dy[s]) + 1e-12;
#pragma unroll
      for (int e = 0; e < NRG_SUR_MAX_ELEMENTS; ++e)
#pragma unroll
        for (int f = 0; f < NRG_SUR_MAX_ELEMENTS; ++f) {
          if (e >= a.ne || f >= a.ne) continue;
          double m = 0.0;
#pragma unroll
          for (int s = 0; s < S; ++s) m += a.E[e][s] * a.E[f][s] * d[s] / (mc.W[s] * mc.W[s]);
          M[e][f] = m;
        }
      if (small_solve(M, res, a.ne)) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
          double c = 0.0;
#pragma unroll
          for (int e = 0; e < NRG_SUR_MAX_ELEMENTS; ++e)
            if (e < a.ne) c += a.E[e][s] * res[e];
          dy[s] -= d[s] * c / mc.W[s];
        }
      }
#pragma unroll
      for (int s = 0; s < S; ++s)
        if (y[s] + dy[s] < -a.y_neg_tol) reason = SUR_R_NEG;
    }
    double dc[S], efp = 0.0;
    if (reason == 0) {
      // chemistry sources of the projected increment, as the FP64 step writes them
#pragma unroll
      for (int s = 0; s < S; ++s) {
        dc[s] = (rho * dy[s]) / mc.W[s];
        efp = efp - (mc.h_ref[s] * dc[s]) / a.dt;
      }
      // energy: E_f(T, Y), source application and e->T, as nrg_chem_closure_dev
      double avg = 0.0, X[S], e0 = 0.0, hc = 0.0;
#pragma unroll
      for (int s = 0; s < S; ++s) avg = avg + y[s] / mc.W[s];
      double wmix = 1.0 / avg;
#pragma unroll
      for (int s = 0; s < S; ++s) X[s] = (y[s] / mc.W[s]) * wmix;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        e0 = e0 + e_molar(mc, T, s) * X[s];
        hc = hc + mc.h_ref[s] * X[s];
      }
      const double Ef = (e0 - hc) / wmix + (0.0 + efp * a.dt) / rho;
      double y1[S], sum = 0.0;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        y1[s] = y[s] + (0.0 + ((rho * dy[s]) / a.dt) * a.dt) / rho;
        sum = sum + y1[s];
      }
      avg = 0.0;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        y1[s] = dmax(y1[s], 0.0) / sum;
        avg = avg + y1[s] / mc.W[s];
      }
      wmix = 1.0 / avg;
#pragma unroll
      for (int s = 0; s < S; ++s) X[s] = (y1[s] / mc.W[s]) * wmix;
      const double T1 = calculate_temperature(mc, T, Ef * wmix, X);
      if (!(fabs(dT - (T1 - T)) <= a.energy_tol)) reason = SUR_R_ENERGY;
    }
    if (reason == 0) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        a.Yp[s][off] = (dc[s] * mc.W[s]) / a.dt;
        a.Yp2[s][off] = dc[s];
      }
      a.Efp[off] = efp;
      if (a.status) a.status[off] = NRG_CELL_SURROGATE;
    }
  }
  const bool rej = live && reason > 0;
  const unsigned long long mask = __ballot(rej ? 1 : 0);
  if (mask) {
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(&a.acc[SUR_FALLBACK], (unsigned long long)__popcll(mask));
    base = __shfl(base, leader, WAVE);
    if (rej) a.list[base + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)q;
  }
  const unsigned long long c_acc = __popcll(__ballot(live && reason <= 0 ? 1 : 0));
  const unsigned long long c_frz = __popcll(__ballot(live && reason < 0 ? 1 : 0));
  const unsigned long long c_gat = __popcll(__ballot(gated ? 1 : 0));
  unsigned long long c_r[SUR_NREASONS];
#pragma unroll
  for (int i = 0; i < SUR_NREASONS; ++i)
    c_r[i] = __popcll(__ballot(rej && reason == SUR_R_DOMAIN + i ? 1 : 0));
  if (lane == 0) {
    atomicAdd(&blk[SUR_ACCEPT], c_acc);
    atomicAdd(&blk[SUR_GATED], c_gat);
    atomicAdd(&blk[SUR_FROZEN], c_frz);
#pragma unroll
    for (int i = 0; i < SUR_NREASONS; ++i) atomicAdd(&blk[SUR_R_DOMAIN + i], c_r[i]);
  }
  __syncthreads();
  if (threadIdx.x < SUR_SLOTS && threadIdx.x != SUR_FALLBACK && blk[threadIdx.x])
    atomicAdd(&a.acc[threadIdx.x], blk[threadIdx.x]);
}

}  // namespace kern
using namespace kern;

// ===========================================================================
// Scratch of nrg_chem_step_dev_noalloc: the work accumulators and queue heads
// (ACC_SLOTS words, padded to 512 bytes), then the cell records of one chunk
// of NRG_NOALLOC_CHUNK_CELLS cells.  One copy per device, held until the
// process ends.  Host code reaches it only through the per-device address
// hipGetSymbolAddress returned (g_noalloc_scratch), never through its name.
// ===========================================================================
constexpr size_t NOALLOC_ACC_BYTES = 512;
#if NRG_TOPO_PRIMARY
static_assert(ACC_SLOTS * sizeof(unsigned long long) <= NOALLOC_ACC_BYTES, "accumulators");
static_assert(NOALLOC_ACC_BYTES + (size_t)NRG_NOALLOC_CHUNK_CELLS * REC_PAIRS * 2 * sizeof(double) <=
                  (size_t)NRG_NOALLOC_SCRATCH_BYTES, "cell records of one chunk");
static_assert((long long)NRG_NOALLOC_MAX_CELLS <= (long long)MAX_CHUNKS * NRG_NOALLOC_CHUNK_CELLS,
              "one queue head per chunk");
// (one build of the library holds the array; the others refuse the no-alloc entry point)
static __device__ __align__(256) unsigned char scratch[NRG_NOALLOC_SCRATCH_BYTES];
#endif

// ===========================================================================
// Host side
// ===========================================================================
// What a captured step graph depends on: its kernel arguments and the queue set-up.
struct StepKey {
  StepArgs args;
  const double *rates;         // buffer the graph's kernels write (reallocated when it grows)
  int use_hist, form;
  long long chunk_max;
  long long persistent_blocks; // grid of the persistent form (the autotuner may change it alone)
  int jacobian;                // NRG_JAC_*: which integrator kernel the graph launches
};

struct nrg_chem_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  MechCoef host_mech{};
  MechCoef *dev_mech = nullptr;
  BdfConst k{};
  nrg_chem_opts opts{};
  unsigned long long *dev_acc = nullptr;   // [16] accumulators + [1] queue head
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int form = NRG_FORM_PLAIN;
  int jac_pct = 50;
  int persistent_blocks = 0;               // resident one-wave blocks of the persistent form
  // staging buffers for the host-pointer entry point
  std::vector<void *> stage;
  // cell records of one chunk: REC_PAIRS pairs of doubles per position
  double *rates = nullptr;
  long long rates_cells = 0;
  long long chunk_max = 1ll << 23;
  size_t stage_cells = 0;
  // learned step: features, network outputs, fallback list, counters
  hipEvent_t ev2 = nullptr, ev3 = nullptr;
  nrg::bf16 *sur_X = nullptr;
  float *sur_Y = nullptr;
  int32_t *sur_list = nullptr;
  int32_t *sur_cand = nullptr;             // cells the explicit-rate screen hands to the network
  int32_t *sur_recheck = nullptr;          // cells its FP32 pass leaves to the FP64 estimate
  int8_t *sur_rate_bad = nullptr;
  long long recheck_blocks = 0;            // resident blocks of rate_screen_kernel
  unsigned long long *sur_acc = nullptr;
  long long sur_cells = 0;
  // stiffness-ordered queue
  int order = NRG_ORDER_NATURAL;
  int32_t *order_list = nullptr;
  int32_t *hist = nullptr;
  unsigned long long *bins = nullptr;      // count[NBIN], cursor[NBIN]
  long long order_cells = 0;
  bool hist_valid = false;
  GridIdx hist_grid{};
  // table-approximated mode: T_t[rows], then dC_t[S][rows]
  double *table = nullptr;
  int table_rows = 0;
  int table_ascending = 0;
  int num_cu = 1;
  // step graph (opts.use_graph)
  bool use_graph = false;
  hipGraphExec_t graph_exec = nullptr;
  StepKey graph_key{};
  long long graph_captures = 0;
};

namespace {

// Gear coefficient tables EL, TQ for MINT = 2 (slatec/ddcst.f90:39-86): the
// l-vectors of the BDF formulas of orders 1..5, normalised by l1.
void gear_tables(double el[MAXORD][MAXORD + 1], double tq[MAXORD][3])
{
  double E[14][13] = {};      // E[i][j] = EL(i, j), 1-based
  double fact[13] = {};
  fact[1] = 1.0;
  for (int i = 2; i <= MAXORD; ++i) fact[i] = (double)i * fact[i - 1];
  E[1][1] = 1.0;
  E[2][1] = 1.0;
  for (int j = 2; j <= MAXORD; ++j) {
    E[1][j] = fact[j];
    for (int i = 2; i <= j; ++i) E[i][j] = (double)j * E[i][j - 1] + E[i - 1][j - 1];
    E[j + 1][j] = 1.0;
  }
  double harm = 1.0;
  for (int j = 2; j <= MAXORD; ++j) {
    harm = harm + 1.0 / (double)j;
    for (int i = 1; i <= j + 1; ++i) E[i][j] = E[i][j] / (fact[j] * harm);
  }
  for (int nq = 1; nq <= MAXORD; ++nq) {
    for (int j = 1; j <= MAXORD + 1; ++j) el[nq - 1][j - 1] = E[j][nq];
    tq[nq - 1][0] = (nq > 1) ? 1.0 / fact[nq - 1] : 0.0;
    tq[nq - 1][1] = (double)(nq + 1) / E[1][nq];
    tq[nq - 1][2] = (double)(nq + 2) / E[1][nq];
  }
}

int build_mech(const nrg_chem_mech *m, MechCoef &mc)
{
  if (!m) return set_error(NRG_CHEM_E_ARG, "mech is NULL");
  if (m->species_number != S || m->reactions_number != R)
    return set_error(NRG_CHEM_E_MECH, "compiled for S=%d R=%d, got S=%d R=%d", S, R,
                     m->species_number, m->reactions_number);
  if (!m->reactions_type || !m->chemical_coeffs || !m->A || !m->beta || !m->E_act || !m->A_low ||
      !m->beta_low || !m->E_act_low || !m->Troe_coeffs || !m->enhanced_efficiencies ||
      !m->a_coeffs || !m->molar_masses)
    return set_error(NRG_CHEM_E_ARG, "mechanism array pointer is NULL");
  for (int r = 0; r < R; ++r)
    if (m->reactions_type[r] != topo::rtype[r])
      return set_error(NRG_CHEM_E_MECH, "reaction %d has type %d, kernels compiled for %d", r + 1,
                       m->reactions_type[r], topo::rtype[r]);
  for (int i = 0; i < 8 * R; ++i)
    if (m->chemical_coeffs[i] != topo::chemical_coeffs[i])
      return set_error(NRG_CHEM_E_MECH, "chemical_coeffs differ from the compiled topology at %d", i);
  for (int s = 0; s < S; ++s)
    if (!(m->molar_masses[s] > 0.0))
      return set_error(NRG_CHEM_E_MECH, "molar mass of species %d is not positive", s + 1);
  std::memset(&mc, 0, sizeof(mc));
  for (int r = 0; r < R; ++r) {
    mc.A[r] = m->A[r];
    mc.beta[r] = m->beta[r];
    mc.A_low[r] = m->A_low[r];
    mc.beta_low[r] = m->beta_low[r];
    mc.mE_R[r] = (-m->E_act[r]) / R_GAS_J;
    mc.mElow_R[r] = (-m->E_act_low[r]) / R_GAS_J;
    for (int q = 0; q < 4; ++q) mc.troe[q][r] = m->Troe_coeffs[r + R * q];
    mc.one_m_alpha[r] = 1.0 - mc.troe[0][r];
    for (int s = 0; s < S; ++s) mc.eff[r][s] = m->enhanced_efficiencies[r + R * s];
  }
  for (int s = 0; s < S; ++s) {
    for (int set = 0; set < 2; ++set) {
      const double *a = m->a_coeffs + 14 * s + 7 * set;
      for (int i = 0; i < 7; ++i) mc.a[s][set][i] = a[i];
      mc.hdiv[s][set][0] = a[4] / 5.0;
      mc.hdiv[s][set][1] = a[3] / 4.0;
      mc.hdiv[s][set][2] = a[2] / 3.0;
      mc.hdiv[s][set][3] = a[1] / 2.0;
      mc.hdiv[s][set][4] = a[0] / 1.0;
      mc.sdiv[s][set][0] = a[4] / 4.0;
      mc.sdiv[s][set][1] = a[3] / 3.0;
      mc.sdiv[s][set][2] = a[2] / 2.0;
    }
    mc.W[s] = m->molar_masses[s];
  }
  // h_s(T_ref): low-T set (T_ref < 1000 K), same Horner order as the kernel.
  for (int s = 0; s < S; ++s) {
    const double t = T_REF;
    const double *d = mc.hdiv[s][0];
    double hs = d[0];
    for (int i = 1; i < 5; ++i) hs = hs * t + d[i];
    hs = hs + mc.a[s][0][5] / t;
    mc.h_ref[s] = hs * t;
  }
  gear_tables(mc.el, mc.tq);
  mc.patm_r = 101325.0 / R_GAS_J;
  for (int r = 0; r < R; ++r) {
    mc.lnA[r] = std::log(mc.A[r]);
    mc.lnA_low[r] = std::log(mc.A_low[r]);
  }
  mc.ln_patm_r = std::log(mc.patm_r);
  return NRG_CHEM_OK;
}

BdfConst make_const(const nrg_chem_opts &o)
{
  BdfConst k{};
  k.eps = o.eps;
  k.ewt = o.ewt;
  k.hmax = o.hmax;
  k.uround = 2.220446049250313080847e-16;        // epsilon(1d0)
  k.sqrtn = std::sqrt((double)S);
  k.sqrt_uround = std::sqrt(k.uround);
  k.br = std::pow(k.uround, 0.875);
  k.bl = std::pow(k.uround, 0.75);
  k.bp = std::pow(k.uround, -0.15);
  k.facmin = std::pow(k.uround, 0.78);
  k.t_gate = o.t_gate;
  k.mxstep = o.mxstep;
  k.maxord = std::min(o.mxord, MAXORD);
  k.gather_pct = o.errtest_gather_pct;
  k.refill_idle = o.refill_min_idle > 0 ? o.refill_min_idle : REFILL_MIN_IDLE_DEFAULT;
  return k;
}

int make_grid(const nrg_grid *grid, GridIdx &g)
{
  if (!grid) return set_error(NRG_CHEM_E_ARG, "grid is NULL");
  for (int d = 0; d < 3; ++d) {
    if (grid->hi[d] < grid->lo[d]) return set_error(NRG_CHEM_E_ARG, "grid hi < lo in dim %d", d);
    if (grid->inner_lo[d] < grid->lo[d] || grid->inner_hi[d] > grid->hi[d])
      return set_error(NRG_CHEM_E_ARG, "inner bounds outside the field bounds in dim %d", d);
  }
  g.lo0 = grid->lo[0];
  g.lo1 = grid->lo[1];
  g.lo2 = grid->lo[2];
  g.nx = grid->hi[0] - grid->lo[0] + 1;
  g.ny = grid->hi[1] - grid->lo[1] + 1;
  g.ilo0 = grid->inner_lo[0];
  g.ilo1 = grid->inner_lo[1];
  g.ilo2 = grid->inner_lo[2];
  g.ix = std::max(0, grid->inner_hi[0] - grid->inner_lo[0] + 1);
  g.iy = std::max(0, grid->inner_hi[1] - grid->inner_lo[1] + 1);
  g.iz = std::max(0, grid->inner_hi[2] - grid->inner_lo[2] + 1);
  g.ninner = (long long)g.ix * g.iy * g.iz;
  g.pad0 = g.pad1 = 0;
  if (g.ninner > 0) {
    auto off = [&](int i, int j, int k) {
      return (long long)(i - g.lo0) + (long long)g.nx * ((j - g.lo1) + (long long)g.ny * (k - g.lo2));
    };
    g.pad0 = off(grid->inner_lo[0], grid->inner_lo[1], grid->inner_lo[2]) & ~63ll;
    g.pad1 = off(grid->inner_hi[0], grid->inner_hi[1], grid->inner_hi[2]) + 1;
  }
  g.idx32 = g.ninner < (1ll << 31) && g.pad1 < (1ll << 31);
  return NRG_CHEM_OK;
}

long long field_cells(const nrg_grid *grid)
{
  long long n = 1;
  for (int d = 0; d < 3; ++d) n *= (long long)(grid->hi[d] - grid->lo[d] + 1);
  return n;
}

int check_device(int dev)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return set_error(NRG_CHEM_E_NO_DEVICE, "no HIP device");
  if (dev < 0 || dev >= n) return set_error(NRG_CHEM_E_NO_DEVICE, "device %d out of range", dev);
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, dev));
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return set_error(NRG_CHEM_E_NO_DEVICE, "device %d is %s, kernels are built for gfx950", dev,
                     p.gcnArchName);
  return NRG_CHEM_OK;
}

}  // namespace

// Device address of `scratch` on each device, taken once per device (with that
// device current) when the first handle on it is created.
constexpr int NOALLOC_MAX_DEVICES = 64;
static void *g_noalloc_scratch[NOALLOC_MAX_DEVICES];
static std::mutex g_noalloc_mu;

static int noalloc_setup(int dev)
{
#if !NRG_TOPO_PRIMARY
  (void)dev;
  return NRG_CHEM_OK;
#else
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return set_error(NRG_CHEM_E_NO_DEVICE, "device %d beyond the %d the library serves", dev,
                     NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_scratch[dev]) return NRG_CHEM_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return set_error(NRG_CHEM_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                     hipGetErrorString(e));
  g_noalloc_scratch[dev] = ptr;
  return NRG_CHEM_OK;
#endif
}

// Blocks of 256 threads covering the aligned walk, one offset per thread.
static long long walk_blocks(const GridIdx &g) { return std::max<long long>(1, (g.pad1 - g.pad0 + 255) / 256); }

// Blocks for a kernel that strides over the aligned walk itself: its resident
// blocks per CU times the CUs (a block more than fits would run as a tail
// after the others), and no more blocks than the walk has 256-cell slices.
template <class K>
static long long resident_blocks(const nrg_chem_handle *h, K kernel, size_t dyn_lds, const GridIdx &g)
{
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, dyn_lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return std::min<long long>(walk_blocks(g), (long long)per_cu * h->num_cu);
}

// ---- entry points of this topology (forwarded from nrg_api.cpp) ----------

int nrg_chem_create(const nrg_chem_mech *mech, const nrg_chem_opts *opts, nrg_chem_t **out)
{
  if (!out) return set_error(NRG_CHEM_E_ARG, "out is NULL");
  *out = nullptr;
  nrg_chem_opts o;
  if (opts)
    o = *opts;
  else
    nrg_chem_default_opts(&o);
  if (!(o.eps > 0.0) || !(o.ewt > 0.0) || !(o.hmax > 0.0) || o.mxstep <= 0 || o.mxord <= 0)
    return set_error(NRG_CHEM_E_ARG, "invalid integrator options");
  if (o.errtest_gather_pct < 0 || o.errtest_gather_pct > 100)
    return set_error(NRG_CHEM_E_ARG, "errtest_gather_pct must be in 0..100");
  if (o.refill_min_idle < 0 || o.refill_min_idle > WAVE)
    return set_error(NRG_CHEM_E_ARG, "refill_min_idle must be in 0..64");
  if (o.jacobian != NRG_JAC_FD_SLATEC && o.jacobian != NRG_JAC_ANALYTIC)
    return set_error(NRG_CHEM_E_ARG, "jacobian mode %d not available", o.jacobian);
  if (o.form != NRG_FORM_PLAIN && o.form != NRG_FORM_PERSISTENT)
    return set_error(NRG_CHEM_E_ARG, "unknown form %d", o.form);
  if (o.jac_batch_pct < 0 || o.jac_batch_pct > 100)
    return set_error(NRG_CHEM_E_ARG, "jac_batch_pct must be in 0..100");
  if (o.waves_per_cu < 0 || o.waves_per_cu > 32)
    return set_error(NRG_CHEM_E_ARG, "waves_per_cu must be in 0..32");
  if (o.cell_order != NRG_ORDER_NATURAL && o.cell_order != NRG_ORDER_STIFFNESS)
    return set_error(NRG_CHEM_E_ARG, "unknown cell_order %d", o.cell_order);
  if (o.chunk_log2 != 0 && (o.chunk_log2 < 16 || o.chunk_log2 > 26))
    return set_error(NRG_CHEM_E_ARG, "chunk_log2 must be 0 or in 16..26");
  auto *h = new (std::nothrow) nrg_chem_handle();
  if (!h) return set_error(NRG_CHEM_E_ARG, "out of host memory");
  int st = build_mech(mech, h->host_mech);
  if (!st) st = check_device(o.device);
  if (st) {
    delete h;
    return st;
  }
  h->opts = o;
  h->device = o.device;
  h->k = make_const(o);
  h->form = o.form;
  h->order = o.cell_order;
  h->use_graph = o.use_graph != 0;
  if (o.chunk_log2) h->chunk_max = 1ll << o.chunk_log2;
  h->jac_pct = o.jac_batch_pct ? o.jac_batch_pct : 50;
  auto fail = [&](int s) {
    nrg_chem_destroy(h);
    return s;
  };
  if (hipSetDevice(h->device) != hipSuccess) return fail(set_error(NRG_CHEM_E_HIP, "hipSetDevice"));
  if (hipStreamCreate(&h->stream) != hipSuccess) return fail(set_error(NRG_CHEM_E_HIP, "hipStreamCreate"));
  if ((st = noalloc_setup(h->device))) return fail(st);
  if (hipMalloc(&h->dev_mech, sizeof(MechCoef)) != hipSuccess)
    return fail(set_error(NRG_CHEM_E_HIP, "hipMalloc mech"));
  if (hipMalloc(&h->dev_acc, ACC_SLOTS * sizeof(unsigned long long)) != hipSuccess)
    return fail(set_error(NRG_CHEM_E_HIP, "hipMalloc acc"));
  {
    hipDeviceProp_t prop;
    int per_cu = 0;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (chem_bdf_kernel<true, false>), WAVE, 0) !=
            hipSuccess)
      return fail(set_error(NRG_CHEM_E_HIP, "occupancy query"));
    if (o.waves_per_cu > 0) per_cu = o.waves_per_cu;
    if (per_cu < 1) per_cu = 1;
    h->persistent_blocks = per_cu * prop.multiProcessorCount;
    h->num_cu = std::max(prop.multiProcessorCount, 1);
  }
  if (hipMemcpy(h->dev_mech, &h->host_mech, sizeof(MechCoef), hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_error(NRG_CHEM_E_HIP, "hipMemcpy mech"));
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
      hipEventCreate(&h->ev2) != hipSuccess || hipEventCreate(&h->ev3) != hipSuccess)
    return fail(set_error(NRG_CHEM_E_HIP, "hipEventCreate"));
  *out = h;
  return NRG_CHEM_OK;
}

void nrg_chem_destroy(nrg_chem_t *h)
{
  if (!h) return;
  // teardown has no caller to report to: failures here are ignored
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void *p : h->stage)
    if (p) (void)hipFree(p);
  if (h->dev_mech) (void)hipFree(h->dev_mech);
  if (h->dev_acc) (void)hipFree(h->dev_acc);
  if (h->rates) (void)hipFree(h->rates);
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  for (hipEvent_t e : {h->ev0, h->ev1, h->ev2, h->ev3})
    if (e) (void)hipEventDestroy(e);
  for (void *p : {(void *)h->sur_X, (void *)h->sur_Y, (void *)h->sur_list, (void *)h->sur_rate_bad,
                  (void *)h->sur_cand, (void *)h->sur_recheck, (void *)h->sur_acc,
                  (void *)h->order_list, (void *)h->hist, (void *)h->bins, (void *)h->table})
    if (p) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

void *nrg_chem_stream(nrg_chem_t *h) { return h ? (void *)h->stream : nullptr; }

static bool same_grid(const GridIdx &x, const GridIdx &y)
{
  return std::memcmp(&x, &y, sizeof(GridIdx)) == 0;
}

// Build the stiffness-ordered cell list of this call into h->order_list and
// point a.list / a.hist at it (device work only, so it can be captured).
static int prepare_order(nrg_chem_t *h, long long n)
{
  if (n >= (1ll << 31)) return set_error(NRG_CHEM_E_ARG, "more than 2^31 inner cells");
  if (n > h->order_cells) {
    if (h->order_list) HIP_TRY(hipFree(h->order_list));
    if (h->hist) HIP_TRY(hipFree(h->hist));
    h->order_list = nullptr;
    h->hist = nullptr;
    h->order_cells = 0;
    h->hist_valid = false;
    HIP_TRY(hipMalloc(&h->order_list, (size_t)n * sizeof(int32_t)));
    HIP_TRY(hipMalloc(&h->hist, (size_t)n * sizeof(int32_t)));
    // no step counts yet (never inside a capture: callers prepare before it)
    HIP_TRY(hipMemsetAsync(h->hist, 0, (size_t)n * sizeof(int32_t), h->stream));
    h->order_cells = n;
  }
  if (!h->bins) HIP_TRY(hipMalloc(&h->bins, 2 * NBIN * sizeof(unsigned long long)));
  return NRG_CHEM_OK;
}

// Order the cells in_list[0..n) (all inner cells when in_list is NULL) into
// h->order_list, most expensive bin first.
static void enqueue_order(nrg_chem_t *h, StepArgs &a, bool use_hist, const int32_t *in_list = nullptr,
                          long long n = -1)
{
  if (n < 0) n = a.g.ninner;
  BinArgs b{};
  b.g = a.g;
  b.T = a.T;
  b.bc = a.bc;
  b.t_gate = a.k.t_gate;
  b.hist = use_hist ? h->hist : nullptr;
  b.count = h->bins;
  b.cursor = h->bins + NBIN;
  b.list = h->order_list;
  b.in_list = in_list;
  b.npos = n;
  const unsigned nb = (unsigned)((n + BIN_RUN - 1) / BIN_RUN);
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(2 * NBIN), 0, h->stream, h->bins, 2 * NBIN);
  hipLaunchKernelGGL(bin_count_kernel, dim3(nb), dim3(BIN_THREADS), 0, h->stream, b);
  hipLaunchKernelGGL(bin_offsets_kernel, dim3(1), dim3(64), 0, h->stream, b);
  hipLaunchKernelGGL(bin_scatter_kernel, dim3(nb), dim3(BIN_THREADS), 0, h->stream, b);
}

// Enqueue rates_kernel + chem_bdf_kernel over npos queue positions (all inner
// cells when a.list is NULL, else the cells a.list names), in chunks of at
// most chunk_max positions, each with its own queue head.
static long long chunk_for(const nrg_chem_t *h, long long npos)
{
  long long chunk = std::min<long long>(npos, h->chunk_max);
  const long long min_chunk = (npos + MAX_CHUNKS - 1) / MAX_CHUNKS;
  return std::max<long long>(std::max<long long>(chunk, min_chunk), 1);
}

static int prepare_rates(nrg_chem_t *h, long long npos)
{
  const long long chunk = chunk_for(h, npos);
  if (npos > 0 && h->rates_cells < chunk) {
    if (h->rates) HIP_TRY(hipFree(h->rates));
    h->rates = nullptr;
    h->rates_cells = 0;
    HIP_TRY(hipMalloc(&h->rates, (size_t)chunk * REC_PAIRS * 2 * sizeof(double)));
    h->rates_cells = chunk;
  }
  return NRG_CHEM_OK;
}

// Launch rates_kernel + chem_bdf_kernel per chunk of `chunk` positions, with
// the chunk's rate constants in `rates` and its queue head in acc[ACC_QUEUE0 + ci].
static void enqueue_chunks(nrg_chem_t *h, StepArgs &a, long long npos, long long chunk,
                           double *rates, unsigned long long *acc)
{
  const long long nchunks = npos > 0 ? (npos + chunk - 1) / chunk : 0;
  a.rates = rates;
  a.rates_out = rates;
  for (long long ci = 0; ci < nchunks; ++ci) {
    a.base = ci * chunk;
    a.nchunk = std::min<long long>(chunk, npos - a.base);
    a.queue = acc + ACC_QUEUE0 + ci;
    hipLaunchKernelGGL(rates_kernel, dim3((unsigned)((a.nchunk + RATES_CELLS - 1) / RATES_CELLS)),
                       dim3(256), 0, h->stream, a);
    const long long need = (a.nchunk + WAVE - 1) / WAVE;
    const bool an = h->opts.jacobian == NRG_JAC_ANALYTIC;
    if (h->form == NRG_FORM_PERSISTENT) {
      const long long nb = std::min<long long>(h->persistent_blocks, need);
      if (an)
        hipLaunchKernelGGL((chem_bdf_kernel<true, true>), dim3((unsigned)nb), dim3(WAVE), 0, h->stream, a);
      else
        hipLaunchKernelGGL((chem_bdf_kernel<true, false>), dim3((unsigned)nb), dim3(WAVE), 0, h->stream, a);
    } else {
      if (an)
        hipLaunchKernelGGL((chem_bdf_kernel<false, true>), dim3((unsigned)need), dim3(WAVE), 0, h->stream, a);
      else
        hipLaunchKernelGGL((chem_bdf_kernel<false, false>), dim3((unsigned)need), dim3(WAVE), 0, h->stream, a);
    }
  }
}

static int enqueue_integrator(nrg_chem_t *h, StepArgs &a, long long npos)
{
  int st = prepare_rates(h, npos);
  if (st) return st;
  enqueue_chunks(h, a, npos, chunk_for(h, npos), h->rates, h->dev_acc);
  HIP_TRY(hipGetLastError());
  return NRG_CHEM_OK;
}

static void fill_stats(const unsigned long long (&acc)[16], long long ninner, float ms,
                       nrg_chem_stats *stats)
{
  std::memset(stats, 0, sizeof(*stats));
  stats->cells_inner = ninner;
  stats->cells_integrated = (int64_t)acc[ACC_INTEG];
  stats->nstep_sum = (int64_t)acc[ACC_NSTEP];
  stats->nfe_sum = (int64_t)acc[ACC_NFE];
  stats->nje_sum = (int64_t)acc[ACC_NJE];
  stats->ncorr_sum = (int64_t)acc[ACC_NCORR];
  stats->natt_sum = (int64_t)acc[ACC_NATT];
  stats->nstep_max = (int32_t)acc[ACC_NSTEP_MAX];
  stats->nfe_max = (int32_t)acc[ACC_NFE_MAX];
  for (int i = 0; i < 6; ++i) stats->status_count[i] = (int64_t)acc[ACC_STATUS0 + i];
  stats->kernel_ms = ms;
}

// The reference STOPs on these (slatec/ddriv3.f90 IERFLG 41, 42); here they are statuses.
static int integrator_status(const unsigned long long (&acc)[16])
{
  if (acc[ACC_STATUS0 + NRG_CELL_STEP_ZERO])
    return set_error(NRG_CHEM_E_STEP_ZERO, "%llu cell(s): step size went to zero",
                     acc[ACC_STATUS0 + NRG_CELL_STEP_ZERO]);
  if (acc[ACC_STATUS0 + NRG_CELL_REPEATED_FAIL] || acc[ACC_STATUS0 + NRG_CELL_SINGULAR])
    return set_error(NRG_CHEM_E_REPEATED_FAIL, "%llu cell(s): repeated integrator failures",
                     acc[ACC_STATUS0 + NRG_CELL_REPEATED_FAIL] + acc[ACC_STATUS0 + NRG_CELL_SINGULAR]);
  return NRG_CHEM_OK;
}

// The device work of one chemistry step: counters reset, optional ordering,
// rate constants and integrator per chunk.  Launches only.
static int enqueue_step(nrg_chem_t *h, StepArgs &a, bool use_hist)
{
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(ACC_SLOTS), 0, h->stream, h->dev_acc, ACC_SLOTS);
  if (a.list) enqueue_order(h, a, use_hist);
  return enqueue_integrator(h, a, a.g.ninner);
}

// Replay the captured step when the arguments equal the captured ones;
// otherwise capture the step into a new graph first.
static int launch_step_graph(nrg_chem_t *h, StepArgs &a, bool use_hist)
{
  StepKey key;
  std::memset(&key, 0, sizeof(key));
  std::memcpy(&key.args, &a, sizeof(StepArgs));
  key.rates = h->rates;
  key.use_hist = use_hist;
  key.form = h->form;
  key.chunk_max = h->chunk_max;
  key.persistent_blocks = h->persistent_blocks;
  key.jacobian = h->opts.jacobian;
  if (!h->graph_exec || std::memcmp(&key, &h->graph_key, sizeof(key)) != 0) {
    if (h->graph_exec) HIP_TRY(hipGraphExecDestroy(h->graph_exec));
    h->graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    StepArgs c = a;
    const int st = enqueue_step(h, c, use_hist);
    const hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (st) {
      if (graph) (void)hipGraphDestroy(graph);
      return st;
    }
    if (e != hipSuccess) return set_error(NRG_CHEM_E_HIP, "graph capture: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&h->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      h->graph_exec = nullptr;
      return set_error(NRG_CHEM_E_HIP, "graph instantiate: %s", hipGetErrorString(ei));
    }
    h->graph_key = key;
    h->graph_captures++;
  }
  HIP_TRY(hipGraphLaunch(h->graph_exec, h->stream));
  return NRG_CHEM_OK;
}

// Step arguments from an entry point's pointers (checked); the work counters,
// accumulators and queue are left for the caller to set.
static int fill_step_args(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T,
                          const double *rho, const double *const *Y, const int8_t *bc,
                          double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                          int32_t *cell_status, StepArgs &a)
{
  if (!h) return set_error(NRG_CHEM_E_ARG, "handle is NULL");
  if (!(dt > 0.0) || !std::isfinite(dt)) return set_error(NRG_CHEM_E_ARG, "time_step must be > 0");
  if (!T || !rho || !Y || !Y_prod || !Y_prod2 || !E_f_prod)
    return set_error(NRG_CHEM_E_ARG, "required field pointer is NULL");
  int st = make_grid(grid, a.g);
  if (st) return st;
  for (int s = 0; s < S; ++s) {
    if (!Y[s] || !Y_prod[s] || !Y_prod2[s]) return set_error(NRG_CHEM_E_ARG, "species %d field is NULL", s + 1);
    a.Y[s] = Y[s];
    a.Yp[s] = Y_prod[s];
    a.Yp2[s] = Y_prod2[s];
  }
  a.dt = dt;
  a.k = h->k;
  a.mech = h->dev_mech;
  a.T = T;
  a.rho = rho;
  a.bc = bc;
  a.Efp = E_f_prod;
  a.status = cell_status;
  a.jac_pct = h->jac_pct;
  return NRG_CHEM_OK;
}

static int step_dev_impl(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T,
                         const double *rho, const double *const *Y, const int8_t *bc,
                         double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                         int32_t *cell_status, nrg_chem_stats *stats, int32_t *nstep,
                         int32_t *nfe, int32_t *nje)
{
  StepArgs a{};
  int st = fill_step_args(h, grid, dt, T, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status, a);
  if (st) return st;
  a.nstep = nstep;
  a.nfe = nfe;
  a.nje = nje;
  a.acc = h->dev_acc;
  HIP_TRY(hipSetDevice(h->device));
  unsigned long long acc[16] = {};
  float ms = 0.f;
  if (a.g.ninner <= 0) {
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(ACC_SLOTS), 0, h->stream, h->dev_acc, ACC_SLOTS);
    HIP_TRY(hipGetLastError());
  } else {
    // buffers first: nothing below allocates, so the sequence can be captured
    const bool ordered = h->order == NRG_ORDER_STIFFNESS;
    bool use_hist = false;
    if (ordered) {
      if ((st = prepare_order(h, a.g.ninner))) return st;
      use_hist = h->hist_valid && same_grid(h->hist_grid, a.g);
      a.list = h->order_list;
      a.hist = h->hist;
    }
    if ((st = prepare_rates(h, a.g.ninner))) return st;
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    if (h->use_graph) {
      if ((st = launch_step_graph(h, a, use_hist))) return st;
    } else {
      if ((st = enqueue_step(h, a, use_hist))) return st;
    }
    if (ordered) {
      h->hist_valid = true;
      h->hist_grid = a.g;
    }
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipMemcpyAsync(acc, h->dev_acc, sizeof(acc), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  }
  if (stats) fill_stats(acc, a.g.ninner, ms, stats);
  return integrator_status(acc);
}

int nrg_chem_step_dev(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T,
                      const double *p, const double *rho, const double *const *Y,
                      const int8_t *bc, double *const *Y_prod, double *const *Y_prod2,
                      double *E_f_prod, int32_t *cell_status, nrg_chem_stats *stats)
{
  (void)p;   // read but unused by the reference step (chemical_kinetics_solver.f90:307)
  return step_dev_impl(h, grid, dt, T, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status, stats,
                       nullptr, nullptr, nullptr);
}

int nrg_chem_step_dev_diag(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T,
                           const double *p, const double *rho, const double *const *Y,
                           const int8_t *bc, double *const *Y_prod, double *const *Y_prod2,
                           double *E_f_prod, int32_t *cell_status, nrg_chem_stats *stats,
                           int32_t *nstep, int32_t *nfe, int32_t *nje)
{
  (void)p;
  if ((nstep || nfe || nje) && !(nstep && nfe && nje))
    return set_error(NRG_CHEM_E_ARG, "nstep, nfe and nje must be all NULL or all set");
  return step_dev_impl(h, grid, dt, T, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status, stats,
                       nstep, nfe, nje);
}

// nrg_chem_step_dev with every device buffer beyond the caller's fields taken
// from `scratch` (through its per-device address): no allocation, no graph,
// natural cell order, chunks of NRG_NOALLOC_CHUNK_CELLS cells.  The kernels and
// their arithmetic are those of nrg_chem_step_dev.
int nrg_chem_step_dev_noalloc(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T,
                              const double *p, const double *rho, const double *const *Y,
                              const int8_t *bc, double *const *Y_prod, double *const *Y_prod2,
                              double *E_f_prod, int32_t *cell_status, nrg_chem_stats *stats)
{
  (void)p;
#if !NRG_TOPO_PRIMARY
  (void)h; (void)grid; (void)dt; (void)T; (void)rho; (void)Y; (void)bc; (void)Y_prod;
  (void)Y_prod2; (void)E_f_prod; (void)cell_status; (void)stats;
  return set_error(NRG_CHEM_E_MECH, "nrg_chem_step_dev_noalloc serves the KEROMNES kernels only, "
                                    "not " NRG_TOPO_NAME);
#else
  StepArgs a{};
  int st = fill_step_args(h, grid, dt, T, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status, a);
  if (st) return st;
  if (a.g.ninner > (long long)NRG_NOALLOC_MAX_CELLS)
    return set_error(NRG_CHEM_E_ARG,
                     "%lld inner cells: nrg_chem_step_dev_noalloc serves at most %lld "
                     "(NRG_NOALLOC_MAX_CELLS); use nrg_chem_step_dev",
                     a.g.ninner, (long long)NRG_NOALLOC_MAX_CELLS);
  void *ptr = (h->device >= 0 && h->device < NOALLOC_MAX_DEVICES) ? g_noalloc_scratch[h->device] : nullptr;
  if (!ptr) return set_error(NRG_CHEM_E_HIP, "no scratch address for device %d", h->device);
  auto *acc_dev = (unsigned long long *)ptr;
  auto *rates = (double *)((unsigned char *)ptr + NOALLOC_ACC_BYTES);
  a.acc = acc_dev;
  HIP_TRY(hipSetDevice(h->device));
  unsigned long long acc[16] = {};
  float ms = 0.f;
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(ACC_SLOTS), 0, h->stream, acc_dev, ACC_SLOTS);
  if (a.g.ninner > 0) enqueue_chunks(h, a, a.g.ninner, NRG_NOALLOC_CHUNK_CELLS, rates, acc_dev);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipMemcpyAsync(acc, acc_dev, sizeof(acc), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (stats) fill_stats(acc, a.g.ninner, ms, stats);
  return integrator_status(acc);
#endif
}

// Host fields: stage them through the handle's device buffers, run the FP64
// step (or the table step) there, copy the outputs back.
static int step_host(nrg_chem_t *h, bool table, const nrg_grid *grid, double dt, const double *T,
                     const double *rho, const double *const *Y, const int8_t *bc,
                     double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                     int32_t *cell_status, nrg_chem_stats *stats)
{
  if (!h) return set_error(NRG_CHEM_E_ARG, "handle is NULL");
  if (!grid || !T || !rho || !Y || !Y_prod || !Y_prod2 || !E_f_prod)
    return set_error(NRG_CHEM_E_ARG, "required pointer is NULL");
  const long long n = field_cells(grid);
  if (n <= 0) return set_error(NRG_CHEM_E_ARG, "empty field");
  HIP_TRY(hipSetDevice(h->device));
  // staging layout: T, rho, E_f_prod, Y[S], Y_prod[S], Y_prod2[S] (doubles), bc (int8), status (int32)
  const int nd = 3 + 3 * S;
  if (h->stage_cells < (size_t)n) {
    h->stage_cells = 0;
    for (void *&q : h->stage)
      if (q) {
        void *old = q;
        q = nullptr;   // never freed twice, even when this free fails
        HIP_TRY(hipFree(old));
      }
    h->stage.assign(nd + 2, nullptr);
    for (int i = 0; i < nd; ++i) HIP_TRY(hipMalloc(&h->stage[i], n * sizeof(double)));
    HIP_TRY(hipMalloc(&h->stage[nd], n * sizeof(int8_t)));
    HIP_TRY(hipMalloc(&h->stage[nd + 1], n * sizeof(int32_t)));
    h->stage_cells = n;
  }
  auto D = [&](int i) { return (double *)h->stage[i]; };
  const size_t bytes = n * sizeof(double);
  hipStream_t s = h->stream;
  HIP_TRY(hipMemcpyAsync(D(0), T, bytes, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(D(1), rho, bytes, hipMemcpyHostToDevice, s));
  const double *Yd[S];
  double *Ypd[S], *Yp2d[S];
  for (int q = 0; q < S; ++q) {
    if (!Y[q] || !Y_prod[q] || !Y_prod2[q]) return set_error(NRG_CHEM_E_ARG, "species %d field is NULL", q + 1);
    HIP_TRY(hipMemcpyAsync(D(3 + q), Y[q], bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(D(3 + 2 * S + q), Y_prod2[q], bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(D(3 + S + q), Y_prod[q], bytes, hipMemcpyHostToDevice, s));
    Yd[q] = D(3 + q);
    Ypd[q] = D(3 + S + q);
    Yp2d[q] = D(3 + 2 * S + q);
  }
  HIP_TRY(hipMemcpyAsync(D(2), E_f_prod, bytes, hipMemcpyHostToDevice, s));
  int8_t *bcd = nullptr;
  if (bc) {
    bcd = (int8_t *)h->stage[nd];
    HIP_TRY(hipMemcpyAsync(bcd, bc, n, hipMemcpyHostToDevice, s));
  }
  int32_t *std_ = cell_status ? (int32_t *)h->stage[nd + 1] : nullptr;
  if (std_) HIP_TRY(hipMemcpyAsync(std_, cell_status, n * sizeof(int32_t), hipMemcpyHostToDevice, s));
  int rc = table ? nrg_chem_table_step_dev(h, grid, dt, D(0), nullptr, D(1), Yd, bcd, Ypd, Yp2d, D(2),
                                           std_, stats)
                 : step_dev_impl(h, grid, dt, D(0), D(1), Yd, bcd, Ypd, Yp2d, D(2), std_, stats, nullptr,
                                 nullptr, nullptr);
  if (rc != NRG_CHEM_OK && rc != NRG_CHEM_E_STEP_ZERO && rc != NRG_CHEM_E_REPEATED_FAIL) return rc;
  for (int q = 0; q < S; ++q) {
    HIP_TRY(hipMemcpyAsync(Y_prod[q], Ypd[q], bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(Y_prod2[q], Yp2d[q], bytes, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipMemcpyAsync(E_f_prod, D(2), bytes, hipMemcpyDeviceToHost, s));
  if (std_) HIP_TRY(hipMemcpyAsync(cell_status, std_, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return rc;
}

int nrg_chem_step(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T,
                  const double *p, const double *rho, const double *const *Y, const int8_t *bc,
                  double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                  int32_t *cell_status, nrg_chem_stats *stats)
{
  (void)p;
  return step_host(h, false, grid, dt, T, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status, stats);
}

int nrg_chem_table_step(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T,
                        const double *p, const double *rho, const double *const *Y, const int8_t *bc,
                        double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                        int32_t *cell_status, nrg_chem_stats *stats)
{
  (void)p;
  if (h && !h->table) return set_error(NRG_CHEM_E_ARG, "no table (nrg_chem_set_table)");
  return step_host(h, true, grid, dt, T, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status, stats);
}

static int closure_impl(nrg_chem_t *h, const nrg_grid *grid, double dt, int apply,
                        const double *rho, const double *const *v, int32_t ndim,
                        const double *E_f_prod, const double *const *Y_prod, const int8_t *bc,
                        double *E_f, double *const *Y, double *T, double *p)
{
  if (!h) return set_error(NRG_CHEM_E_ARG, "handle is NULL");
  if (!rho || !E_f || !Y || !T) return set_error(NRG_CHEM_E_ARG, "required field pointer is NULL");
  if (ndim < 0 || ndim > 3 || (ndim > 0 && !v)) return set_error(NRG_CHEM_E_ARG, "bad velocity");
  if (apply && (!E_f_prod || !Y_prod || !(dt > 0.0)))
    return set_error(NRG_CHEM_E_ARG, "sources and dt > 0 required");
  ClosureArgs a{};
  int st = make_grid(grid, a.g);
  if (st) return st;
  a.dt = dt;
  a.apply_sources = apply;
  a.ndim = ndim;
  a.mech = h->dev_mech;
  a.rho = rho;
  for (int d = 0; d < ndim; ++d) {
    if (!v[d]) return set_error(NRG_CHEM_E_ARG, "velocity component %d is NULL", d);
    a.v[d] = v[d];
  }
  a.Efp = E_f_prod;
  for (int s = 0; s < S; ++s) {
    if (!Y[s] || (apply && !Y_prod[s])) return set_error(NRG_CHEM_E_ARG, "species %d field is NULL", s + 1);
    a.Y[s] = Y[s];
    a.Yp[s] = apply ? Y_prod[s] : nullptr;
  }
  a.bc = bc;
  a.E_f = E_f;
  a.T = T;
  a.p = p;
  HIP_TRY(hipSetDevice(h->device));
  if (a.g.ninner > 0) {
    const long long nb = walk_blocks(a.g);
    hipLaunchKernelGGL(closure_kernel, dim3((unsigned)nb), dim3(256), 0, h->stream, a);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return NRG_CHEM_OK;
}

int nrg_chem_closure_dev(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *rho,
                         const double *const *v, int32_t ndim, const double *E_f_prod,
                         const double *const *Y_prod, const int8_t *bc, double *E_f,
                         double *const *Y, double *T, double *p)
{
  return closure_impl(h, grid, dt, 1, rho, v, ndim, E_f_prod, Y_prod, bc, E_f, Y, T, p);
}

int nrg_state_equation_dev(nrg_chem_t *h, const nrg_grid *grid, const double *rho,
                           const double *const *v, int32_t ndim, const double *E_f,
                           const double *const *Y, const int8_t *bc, double *T, double *p)
{
  return closure_impl(h, grid, 0.0, 0, rho, v, ndim, nullptr, nullptr, bc, const_cast<double *>(E_f),
                      const_cast<double *const *>(Y), T, p);
}

int nrg_energy_from_temperature_dev(nrg_chem_t *h, const nrg_grid *grid, const double *T,
                                    const double *const *Y, const double *const *v, int32_t ndim,
                                    double *E_f)
{
  if (!h) return set_error(NRG_CHEM_E_ARG, "handle is NULL");
  if (!T || !Y || !E_f) return set_error(NRG_CHEM_E_ARG, "required field pointer is NULL");
  if (ndim < 0 || ndim > 3 || (ndim > 0 && !v)) return set_error(NRG_CHEM_E_ARG, "bad velocity");
  EnergyArgs a{};
  int st = make_grid(grid, a.g);
  if (st) return st;
  a.ndim = ndim;
  a.mech = h->dev_mech;
  a.T = T;
  for (int s = 0; s < S; ++s) {
    if (!Y[s]) return set_error(NRG_CHEM_E_ARG, "species %d field is NULL", s + 1);
    a.Y[s] = Y[s];
  }
  for (int d = 0; d < ndim; ++d) a.v[d] = v[d];
  a.E_f = E_f;
  HIP_TRY(hipSetDevice(h->device));
  if (a.g.ninner > 0) {
    const long long nb = (a.g.ninner + 255) / 256;
    hipLaunchKernelGGL(energy_kernel, dim3((unsigned)nb), dim3(256), 0, h->stream, a);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return NRG_CHEM_OK;
}

int nrg_chem_set_table(nrg_chem_t *h, int32_t rows, const double *T_t, const double *dC_t)
{
  if (!h) return set_error(NRG_CHEM_E_ARG, "handle is NULL");
  if (rows < 2 || !T_t || !dC_t) return set_error(NRG_CHEM_E_ARG, "a table needs >= 2 rows and both arrays");
  if (rows > (1 << 24)) return set_error(NRG_CHEM_E_ARG, "a table has at most 2^24 rows, got %d", rows);
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->table) HIP_TRY(hipFree(h->table));
  h->table = nullptr;
  h->table_rows = 0;
  std::vector<double> t((size_t)(S + 1) * rows);
  std::memcpy(t.data(), T_t, sizeof(double) * rows);
  std::memcpy(t.data() + rows, dC_t, sizeof(double) * S * rows);
  HIP_TRY(hipMalloc(&h->table, t.size() * sizeof(double)));
  HIP_TRY(hipMemcpy(h->table, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  h->table_rows = rows;
  h->table_ascending = 1;
  for (int32_t i = 0; i + 1 < rows; ++i)
    if (!(T_t[i] < T_t[i + 1])) h->table_ascending = 0;
  return NRG_CHEM_OK;
}

int nrg_chem_table_step_dev(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T,
                            const double *p, const double *rho, const double *const *Y,
                            const int8_t *bc, double *const *Y_prod, double *const *Y_prod2,
                            double *E_f_prod, int32_t *cell_status, nrg_chem_stats *stats)
{
  (void)p;
  (void)rho;
  if (!h) return set_error(NRG_CHEM_E_ARG, "handle is NULL");
  if (!h->table) return set_error(NRG_CHEM_E_ARG, "no table (nrg_chem_set_table)");
  if (!(dt > 0.0) || !std::isfinite(dt)) return set_error(NRG_CHEM_E_ARG, "time_step must be > 0");
  if (!T || !Y || !Y_prod || !Y_prod2 || !E_f_prod)
    return set_error(NRG_CHEM_E_ARG, "required field pointer is NULL");
  TableArgs a{};
  int st = make_grid(grid, a.g);
  if (st) return st;
  for (int s = 0; s < S; ++s) {
    if (!Y[s] || !Y_prod[s] || !Y_prod2[s]) return set_error(NRG_CHEM_E_ARG, "species %d field is NULL", s + 1);
    a.Y[s] = Y[s];
    a.Yp[s] = Y_prod[s];
    a.Yp2[s] = Y_prod2[s];
  }
  a.dt = dt;
  a.t_gate = h->k.t_gate;
  a.mech = h->dev_mech;
  a.T = T;
  a.bc = bc;
  a.Efp = E_f_prod;
  a.status = cell_status;
  a.table = h->table;
  a.rows = h->table_rows;
  a.ascending = h->table_ascending;
  a.acc = h->dev_acc;
  HIP_TRY(hipSetDevice(h->device));
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(ACC_SLOTS), 0, h->stream, h->dev_acc, ACC_SLOTS);
  unsigned long long acc[16] = {};
  float ms = 0.f;
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  if (a.g.ninner > 0) {
    const size_t tab_bytes = sizeof(double) * (size_t)(S + 1) * (size_t)a.rows;
    if (tab_bytes <= (size_t)TABLE_LDS_BYTES)
      hipLaunchKernelGGL(table_kernel<true>, dim3((unsigned)resident_blocks(h, table_kernel<true>, tab_bytes, a.g)),
                         dim3(TABLE_THREADS), tab_bytes, h->stream, a);
    else
      hipLaunchKernelGGL(table_kernel<false>, dim3((unsigned)resident_blocks(h, table_kernel<false>, 0, a.g)),
                         dim3(TABLE_THREADS), 0, h->stream, a);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipMemcpyAsync(acc, h->dev_acc, sizeof(acc), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (stats) fill_stats(acc, a.g.ninner, ms, stats);
  return NRG_CHEM_OK;
}

// ---------------------------------------------------------------------------
// Low-Mach coupling entry points (include/nrg_low_mach.h)
// ---------------------------------------------------------------------------
int nrg_low_mach_state_equation_dev(nrg_chem_t *h, const nrg_grid *grid, double time_step,
                                    int32_t predictor, const double *rho,
                                    const double *const *Y, const double *const *v, int32_t ndim,
                                    const double *dp_stat_dt, const int8_t *bc, double *p_stat,
                                    double *p_stat_old, double *T, double *mix_mol_mass, double *p,
                                    double *mixture_cp, double *gamma, double *h_s, double *v_s,
                                    double *e_i, double *E_f)
{
  if (!h) return set_error(NRG_CHEM_E_ARG, "handle is NULL");
  if (!rho || !Y || !dp_stat_dt || !p_stat || !p_stat_old || !T || !mix_mol_mass)
    return set_error(NRG_CHEM_E_ARG, "required field pointer is NULL");
  if (ndim < 0 || ndim > 3 || (ndim > 0 && !v)) return set_error(NRG_CHEM_E_ARG, "bad velocity");
  if (!std::isfinite(time_step)) return set_error(NRG_CHEM_E_ARG, "time_step is not finite");
  LowMachEosArgs a{};
  int st = make_grid(grid, a.g);
  if (st) return st;
  for (int s = 0; s < S; ++s) {
    if (!Y[s]) return set_error(NRG_CHEM_E_ARG, "species %d field is NULL", s + 1);
    a.Y[s] = Y[s];
  }
  for (int d = 0; d < ndim; ++d) {
    if (!v[d]) return set_error(NRG_CHEM_E_ARG, "velocity %d is NULL", d + 1);
    a.v[d] = v[d];
  }
  a.dt = time_step;
  a.predictor = predictor != 0;
  a.ndim = ndim;
  a.mech = h->dev_mech;
  a.rho = rho;
  a.dp = dp_stat_dt;
  a.bc = bc;
  a.p_stat = p_stat;
  a.p_stat_old = p_stat_old;
  a.T = T;
  a.mmm = mix_mol_mass;
  a.p = p;
  a.cp = mixture_cp;
  a.gamma = gamma;
  a.h_s = h_s;
  a.v_s = v_s;
  a.e_i = e_i;
  a.E_f = E_f;
  HIP_TRY(hipSetDevice(h->device));
  if (a.g.ninner > 0) {
    hipLaunchKernelGGL(low_mach_eos_kernel, dim3((unsigned)walk_blocks(a.g)), dim3(LM_THREADS), 0,
                       h->stream, a);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return NRG_CHEM_OK;
}

int nrg_low_mach_species_dev(nrg_chem_t *h, const nrg_grid *grid, double time_step,
                             int32_t corrector, const double *rho_a, const double *const *Y_a,
                             const double *rho_b, const double *const *Y_b,
                             const double *const *rhs, const double *const *Y_prod_diff,
                             const double *const *Y_prod_chem, const int8_t *bc, double *rho_out,
                             double *const *Y_out)
{
  if (!h) return set_error(NRG_CHEM_E_ARG, "handle is NULL");
  if (!rho_a || !Y_a || !rho_out || !Y_out) return set_error(NRG_CHEM_E_ARG, "required field pointer is NULL");
  if (corrector && (!rho_b || !Y_b)) return set_error(NRG_CHEM_E_ARG, "the corrector needs rho_b and Y_b");
  if (!std::isfinite(time_step)) return set_error(NRG_CHEM_E_ARG, "time_step is not finite");
  LowMachSpeciesArgs a{};
  int st = make_grid(grid, a.g);
  if (st) return st;
  for (int s = 0; s < S; ++s) {
    if (!Y_a[s] || !Y_out[s] || (corrector && !Y_b[s]) || (rhs && !rhs[s]) ||
        (Y_prod_diff && !Y_prod_diff[s]) || (Y_prod_chem && !Y_prod_chem[s]))
      return set_error(NRG_CHEM_E_ARG, "species %d field is NULL", s + 1);
    a.Y_a[s] = Y_a[s];
    a.Y_b[s] = corrector ? Y_b[s] : nullptr;
    a.rhs[s] = rhs ? rhs[s] : nullptr;
    a.Yd[s] = Y_prod_diff ? Y_prod_diff[s] : nullptr;
    a.Yc[s] = Y_prod_chem ? Y_prod_chem[s] : nullptr;
    a.Y_out[s] = Y_out[s];
  }
  a.dt = time_step;
  a.corrector = corrector != 0;
  a.rho_a = rho_a;
  a.rho_b = corrector ? rho_b : nullptr;
  a.bc = bc;
  a.rho_out = rho_out;
  HIP_TRY(hipSetDevice(h->device));
  if (a.g.ninner > 0) {
    hipLaunchKernelGGL(low_mach_species_kernel, dim3((unsigned)walk_blocks(a.g)), dim3(LM_THREADS), 0,
                       h->stream, a);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return NRG_CHEM_OK;
}

int nrg_low_mach_chem_divergence_dev(nrg_chem_t *h, const nrg_grid *grid, const double *rho,
                                     const double *T, const double *const *Y,
                                     const double *mix_mol_mass, const double *E_f_prod_chem,
                                     const double *const *Y_prod_chem, const int8_t *bc,
                                     double *div_v)
{
  if (!h) return set_error(NRG_CHEM_E_ARG, "handle is NULL");
  if (!rho || !T || !Y || !mix_mol_mass || !E_f_prod_chem || !Y_prod_chem || !div_v)
    return set_error(NRG_CHEM_E_ARG, "required field pointer is NULL");
  LowMachDivArgs a{};
  int st = make_grid(grid, a.g);
  if (st) return st;
  for (int s = 0; s < S; ++s) {
    if (!Y[s] || !Y_prod_chem[s]) return set_error(NRG_CHEM_E_ARG, "species %d field is NULL", s + 1);
    a.Y[s] = Y[s];
    a.Yc[s] = Y_prod_chem[s];
  }
  a.mech = h->dev_mech;
  a.rho = rho;
  a.T = T;
  a.mmm = mix_mol_mass;
  a.Efc = E_f_prod_chem;
  a.bc = bc;
  a.div = div_v;
  HIP_TRY(hipSetDevice(h->device));
  if (a.g.ninner > 0) {
    hipLaunchKernelGGL(low_mach_div_kernel, dim3((unsigned)walk_blocks(a.g)), dim3(LM_THREADS), 0,
                       h->stream, a);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return NRG_CHEM_OK;
}

int nrg_chem_autotune_dev(nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T,
                          const double *p, const double *rho, const double *const *Y,
                          const int8_t *bc, double *const *Y_prod, double *const *Y_prod2,
                          double *E_f_prod, nrg_chem_tuning *out)
{
  (void)p;
  if (!h || !out) return set_error(NRG_CHEM_E_ARG, "handle or out is NULL");
  std::memset(out, 0, sizeof(*out));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, h->device));
  int occ = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (chem_bdf_kernel<true, false>), WAVE, 0));
  occ = std::max(occ, 1);
  struct Cfg {
    int form, order, jac, wpc;
    long long chunk;
    int gather, refill;
  };
  auto apply = [&](const Cfg &c) {
    h->k.gather_pct = c.gather;
    h->k.refill_idle = c.refill;
    h->form = c.form;
    h->order = c.order;
    h->jac_pct = c.jac;
    h->persistent_blocks = (c.wpc > 0 ? c.wpc : occ) * prop.multiProcessorCount;
    h->chunk_max = c.chunk;
  };
  auto trial = [&](const Cfg &c, double &ms) -> int {
    apply(c);
    nrg_chem_stats s;
    const int st = step_dev_impl(h, grid, dt, T, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, nullptr,
                                 &s, nullptr, nullptr, nullptr);
    ms = s.kernel_ms;
    if (st == NRG_CHEM_OK && out->n_trials < NRG_TUNE_MAX_TRIALS) {
      nrg_chem_tune_trial &t = out->trials[out->n_trials++];
      t.form = c.form; t.cell_order = c.order; t.jac_batch_pct = c.jac;
      t.waves_per_cu = c.wpc > 0 ? c.wpc : occ; t.chunk_cells = c.chunk; t.kernel_ms = ms;
      t.errtest_gather_pct = c.gather;
      t.refill_min_idle = c.refill;
    }
    return st;
  };
  // trials run eagerly: a capture per trial set-up would add host latency to its time
  const bool graph = h->use_graph;
  h->use_graph = false;
  struct Restore {
    nrg_chem_t *h;
    bool g;
    ~Restore() { h->use_graph = g; }
  } restore{h, graph};
  const int wpc0 = h->persistent_blocks / std::max(prop.multiProcessorCount, 1);
  Cfg best{h->form, h->order, h->jac_pct, wpc0 == occ ? 0 : wpc0, h->chunk_max, h->k.gather_pct,
           h->k.refill_idle};
  double best_ms = 0.0;
  int st;
  // the first run sets up the stiffness history that the ordered trials use
  if ((st = trial(best, best_ms))) return st;
  if ((st = trial(best, best_ms))) return st;
  auto consider = [&](Cfg c) -> int {
    double ms = 0.0;
    const int s = trial(c, ms);
    if (s) return s;
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
    return NRG_CHEM_OK;
  };
  for (int order : {NRG_ORDER_NATURAL, NRG_ORDER_STIFFNESS})
    if (order != best.order) {
      Cfg c = best;
      c.order = order;
      if ((st = consider(c))) return st;
    }
  for (int form : {NRG_FORM_PLAIN, NRG_FORM_PERSISTENT})
    if (form != best.form) {
      Cfg c = best;
      c.form = form;
      if ((st = consider(c))) return st;
    }
  for (int gather : {0, 100})
    if (gather != best.gather) {
      Cfg c = best;
      c.gather = gather;
      if ((st = consider(c))) return st;
    }
  for (int jac : {10, 25, 50, 75, 100})
    if (jac != best.jac) {
      Cfg c = best;
      c.jac = jac;
      if ((st = consider(c))) return st;
    }
  if (best.form == NRG_FORM_PERSISTENT)
    for (int refill : {48, REFILL_MIN_IDLE_DEFAULT})
      if (refill != best.refill) {
        Cfg c = best;
        c.refill = refill;
        if ((st = consider(c))) return st;
      }
  if (best.form == NRG_FORM_PERSISTENT)
    for (int w : {2 * occ, 4 * occ}) {
      Cfg c = best;
      c.wpc = w;
      if ((st = consider(c))) return st;
    }
  for (long long ch : {1ll << 21, 1ll << 22, 1ll << 23, 1ll << 24})
    if (ch != best.chunk) {
      Cfg c = best;
      c.chunk = ch;
      if ((st = consider(c))) return st;
    }
  apply(best);
  out->best.form = best.form;
  out->best.cell_order = best.order;
  out->best.jac_batch_pct = best.jac;
  out->best.waves_per_cu = best.wpc > 0 ? best.wpc : occ;
  out->best.chunk_cells = best.chunk;
  out->best.kernel_ms = best_ms;
  out->best.errtest_gather_pct = best.gather;
  out->best.refill_min_idle = best.refill;
  return NRG_CHEM_OK;
}

int nrg_chem_surrogate_step_dev(nrg_chem_t *h, nrg_mlp_t *mlp, const nrg_surrogate_norm *norm,
                                const nrg_surrogate_gate *gate, const nrg_grid *grid, double dt,
                                const double *T, const double *p, const double *rho,
                                const double *const *Y, const int8_t *bc,
                                double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                                int32_t *cell_status, nrg_surrogate_stats *stats)
{
  if (!h || !gate) return set_error(NRG_CHEM_E_ARG, "handle or gate is NULL");
  const bool use_net = gate->screen != NRG_SUR_SCREEN_ONLY;
  if (use_net && (!mlp || !norm)) return set_error(NRG_CHEM_E_ARG, "network or norm is NULL");
  if (!(dt > 0.0) || !std::isfinite(dt)) return set_error(NRG_CHEM_E_ARG, "time_step must be > 0");
  if (!T || !rho || !Y || !Y_prod || !Y_prod2 || !E_f_prod || (use_net && !p))
    return set_error(NRG_CHEM_E_ARG, "required field pointer is NULL (the network needs p)");
  if (use_net && (mlp_n_in(mlp) != NIN || mlp_n_out(mlp) != NOUT))
    return set_error(NRG_CHEM_E_ARG, "network maps %d -> %d, the step needs %d -> %d", mlp_n_in(mlp),
                     mlp_n_out(mlp), NIN, NOUT);
  if (gate->n_elements < 0 || gate->n_elements > NRG_SUR_MAX_ELEMENTS)
    return set_error(NRG_CHEM_E_ARG, "n_elements must be in 0..%d", NRG_SUR_MAX_ELEMENTS);
  StepArgs a{};
  int st = make_grid(grid, a.g);
  if (st) return st;
  if (a.g.ninner >= (1ll << 31)) return set_error(NRG_CHEM_E_ARG, "more than 2^31 inner cells");
  SurArgs sa{};
  sa.g = a.g;
  for (int s = 0; s < S; ++s) {
    if (!Y[s] || !Y_prod[s] || !Y_prod2[s]) return set_error(NRG_CHEM_E_ARG, "species %d field is NULL", s + 1);
    a.Y[s] = sa.Y[s] = Y[s];
    a.Yp[s] = sa.Yp[s] = Y_prod[s];
    a.Yp2[s] = sa.Yp2[s] = Y_prod2[s];
  }
  a.dt = sa.dt = dt;
  a.k = h->k;
  a.mech = sa.mech = h->dev_mech;
  a.T = sa.T = T;
  a.rho = sa.rho = rho;
  sa.p = p;
  a.bc = sa.bc = bc;
  a.Efp = sa.Efp = E_f_prod;
  a.status = sa.status = cell_status;
  a.acc = h->dev_acc;
  a.jac_pct = h->jac_pct;
  sa.t_gate = h->k.t_gate;
  if (norm) {
    for (int i = 0; i < NIN; ++i) {
      sa.in_shift[i] = norm->in_shift[i];
      sa.in_scale[i] = norm->in_scale[i];
    }
    for (int i = 0; i < NOUT; ++i) {
      sa.out_shift[i] = norm->out_shift[i];
      sa.out_scale[i] = norm->out_scale[i];
    }
    sa.y_floor = norm->y_floor;
  }
  sa.t_min = gate->t_min; sa.t_max = gate->t_max;
  sa.p_min = gate->p_min; sa.p_max = gate->p_max;
  sa.dt_ok = dt >= gate->dt_min && dt <= gate->dt_max;
  sa.elem_tol = gate->elem_tol; sa.elem_floor = gate->elem_floor;
  sa.y_neg_tol = gate->y_neg_tol; sa.energy_tol = gate->energy_tol; sa.dT_max = gate->dT_max;
  sa.dT_freeze = gate->dT_freeze; sa.dY_freeze = gate->dY_freeze;
  if (!(gate->rate_tol >= 0.0)) return set_error(NRG_CHEM_E_ARG, "gate rate_tol must be >= 0");
  sa.rate_tol = gate->rate_tol;
  const int screen = gate->screen;
  if (screen != NRG_SUR_SCREEN_NONE && screen != NRG_SUR_SCREEN_RATES && screen != NRG_SUR_SCREEN_ONLY)
    return set_error(NRG_CHEM_E_ARG, "unknown gate screen %d", screen);
  if (screen != NRG_SUR_SCREEN_NONE && !(gate->dT_freeze > 0.0))
    return set_error(NRG_CHEM_E_ARG, "the explicit-rate screen needs dT_freeze > 0");
  sa.ne = gate->n_elements;
  for (int e = 0; e < sa.ne; ++e)
    for (int s = 0; s < S; ++s) sa.E[e][s] = gate->elements[e][s];

  HIP_TRY(hipSetDevice(h->device));
  const long long n = a.g.ninner;
  if (n > h->sur_cells) {
    for (void *q : {(void *)h->sur_X, (void *)h->sur_Y, (void *)h->sur_list, (void *)h->sur_rate_bad,
                    (void *)h->sur_cand, (void *)h->sur_recheck})
      if (q) HIP_TRY(hipFree(q));
    h->sur_X = nullptr; h->sur_Y = nullptr; h->sur_list = nullptr; h->sur_rate_bad = nullptr;
    h->sur_cand = nullptr; h->sur_recheck = nullptr;
    h->sur_cells = 0;
    HIP_TRY(hipMalloc(&h->sur_X, (size_t)n * MLP_IN_PAD * sizeof(bf16)));
    HIP_TRY(hipMalloc(&h->sur_Y, (size_t)n * NOUT * sizeof(float)));
    HIP_TRY(hipMalloc(&h->sur_list, (size_t)n * sizeof(int32_t)));
    HIP_TRY(hipMalloc(&h->sur_rate_bad, (size_t)n * sizeof(int8_t)));
    HIP_TRY(hipMalloc(&h->sur_cand, (size_t)n * sizeof(int32_t)));
    HIP_TRY(hipMalloc(&h->sur_recheck, (size_t)n * sizeof(int32_t)));
    h->sur_cells = n;
  }
  if (!h->sur_acc) HIP_TRY(hipMalloc(&h->sur_acc, SUR_SLOTS * sizeof(unsigned long long)));
  sa.X = h->sur_X;
  sa.Ypred = h->sur_Y;
  sa.list = h->sur_list;
  sa.rate_bad = h->sur_rate_bad;
  sa.acc = h->sur_acc;
  sa.cand = nullptr;
  sa.npos = n;

  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(ACC_SLOTS), 0, h->stream, h->dev_acc, ACC_SLOTS);
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(SUR_SLOTS), 0, h->stream, h->sur_acc, SUR_SLOTS);
  HIP_TRY(hipGetLastError());
  unsigned long long sur[SUR_SLOTS] = {}, acc[16] = {};
  float ms_scr = 0.f, ms_inf = 0.f, ms_gate = 0.f, ms_fb = 0.f;
  long long ncand = n;
  if (n > 0) {
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    if (screen != NRG_SUR_SCREEN_NONE) {
      // the explicit-rate screen freezes the inert cells; the others go to the
      // network (SCREEN_RATES) or straight to the FP64 integrator (SCREEN_ONLY)
      sa.screen_out = screen == NRG_SUR_SCREEN_RATES ? h->sur_cand : h->sur_list;
      sa.screen_slot = screen == NRG_SUR_SCREEN_RATES ? SUR_CAND : SUR_FALLBACK;
      sa.screen_fp64 = gate->screen_fp64 != 0;
      sa.recheck = h->sur_recheck;
      sa.recheck_n = h->sur_acc + SUR_RECHECK;
      if (!h->recheck_blocks) {
        // resident blocks of the FP64 pass, which strides over its list
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rate_screen_kernel, 256, 0) != hipSuccess ||
            per_cu < 1)
          per_cu = 1;
        h->recheck_blocks = (long long)per_cu * h->num_cu;
      }
      // the FP32 pass over every cell, then the FP64 estimate on the cells it
      // lists (a grid that reads their number on the device)
      hipLaunchKernelGGL(screen_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, sa);
      hipLaunchKernelGGL(rate_screen_kernel,
                         dim3((unsigned)std::min<long long>(h->recheck_blocks, (n + RCHECK_CELLS - 1) / RCHECK_CELLS)),
                         dim3(256), 0, h->stream, sa);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(sur, h->sur_acc, sizeof(sur), hipMemcpyDeviceToHost, h->stream));
      HIP_TRY(hipStreamSynchronize(h->stream));
      ncand = screen == NRG_SUR_SCREEN_RATES ? (long long)sur[SUR_CAND] : 0;
      if (ncand > n || (long long)sur[SUR_FALLBACK] > n || (long long)sur[SUR_RECHECK] > n)
        return set_error(NRG_CHEM_E_HIP, "screen counts exceed %lld cells", n);
      sa.cand = h->sur_cand;
      sa.npos = ncand;
    }
    HIP_TRY(hipEventRecord(h->ev2, h->stream));
    if (ncand > 0) {
      const unsigned nb = (unsigned)((ncand + 255) / 256);
      hipLaunchKernelGGL(featurize_kernel, dim3(nb), dim3(256), 0, h->stream, sa);
      HIP_TRY(hipGetLastError());
      if ((st = mlp_infer_bf16(mlp, ncand, h->sur_X, h->sur_Y, h->stream))) return st;
    }
    HIP_TRY(hipEventRecord(h->ev3, h->stream));
    HIP_TRY(hipEventSynchronize(h->ev3));
    HIP_TRY(hipEventElapsedTime(&ms_scr, h->ev0, h->ev2));
    HIP_TRY(hipEventElapsedTime(&ms_inf, h->ev2, h->ev3));
    if (ncand > 0) {
      const unsigned nb = (unsigned)((ncand + 255) / 256);
      if (sa.rate_tol > 0.0) {
        hipLaunchKernelGGL(rate_check_kernel, dim3((unsigned)((ncand + RCHECK_CELLS - 1) / RCHECK_CELLS)),
                           dim3(256), 0, h->stream, sa);
        hipLaunchKernelGGL(gate_kernel<true>, dim3(nb), dim3(256), 0, h->stream, sa);
      } else
        hipLaunchKernelGGL(gate_kernel<false>, dim3(nb), dim3(256), 0, h->stream, sa);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(h->ev2, h->stream));
    HIP_TRY(hipMemcpyAsync(sur, h->sur_acc, sizeof(sur), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipEventElapsedTime(&ms_gate, h->ev3, h->ev2));
    HIP_TRY(hipEventRecord(h->ev3, h->stream));
    const long long nfb = (long long)sur[SUR_FALLBACK];
    if (nfb > n) return set_error(NRG_CHEM_E_HIP, "fallback count %lld exceeds %lld cells", nfb, n);
    if (nfb > 0) {
      a.list = h->sur_list;
      // the fallback cells are the stiffest of the field: hand them out most
      // expensive first, by the step counts this handle recorded for them
      // (an FP64 step or an earlier fallback; temperature bins when it has
      // none for this grid), and record their step counts for the next call
      if (h->order == NRG_ORDER_STIFFNESS) {
        if ((st = prepare_order(h, n))) return st;
        const bool use_hist = h->hist_valid && same_grid(h->hist_grid, a.g);
        enqueue_order(h, a, use_hist, h->sur_list, nfb);
        a.list = h->order_list;
        a.hist = h->hist;
        h->hist_valid = true;
        h->hist_grid = a.g;
      }
      if ((st = enqueue_integrator(h, a, nfb))) return st;
    }
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipMemcpyAsync(acc, h->dev_acc, sizeof(acc), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipEventElapsedTime(&ms_fb, h->ev3, h->ev1));
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->cells_inner = n;
    stats->cells_gated = (int64_t)sur[SUR_GATED];
    stats->cells_accepted = (int64_t)sur[SUR_ACCEPT];
    stats->cells_fallback = (int64_t)sur[SUR_FALLBACK];
    stats->reject_domain = (int64_t)sur[SUR_R_DOMAIN];
    stats->reject_elements = (int64_t)sur[SUR_R_ELEM];
    stats->reject_negative = (int64_t)sur[SUR_R_NEG];
    stats->reject_energy = (int64_t)sur[SUR_R_ENERGY];
    stats->reject_confidence = (int64_t)sur[SUR_R_CONF];
    stats->reject_rate = (int64_t)sur[SUR_R_RATE];
    stats->cells_frozen = (int64_t)sur[SUR_FROZEN];
    stats->cells_network = ncand;
    stats->screen_ms = ms_scr;
    stats->cells_screen_fp64 = (int64_t)sur[SUR_RECHECK];
    stats->infer_ms = ms_inf;
    stats->gate_ms = ms_gate;
    stats->fallback_ms = ms_fb;
    fill_stats(acc, (long long)sur[SUR_FALLBACK], ms_fb, &stats->fallback);
  }
  return integrator_status(acc);
}

// ---- this topology's table for the C-ABI -------------------------------------
#define H(h) reinterpret_cast<nrg_chem_t *>(h)
#define NRG_TOPO_ADAPTER(name, params, args) static int op_##name params { return name args; }
NRG_TOPO_ENTRY_POINTS(NRG_TOPO_ADAPTER)
#undef NRG_TOPO_ADAPTER

static int op_check(const nrg_chem_mech *mech)
{
  MechCoef tmp;
  return build_mech(mech, tmp);
}
static int op_create(const nrg_chem_mech *mech, const nrg_chem_opts *opts, ::nrg_chem_t **out)
{
  nrg_chem_t *h = nullptr;
  const int st = nrg_chem_create(mech, opts, &h);
  *out = reinterpret_cast<::nrg_chem_t *>(h);
  return st;
}
static void op_destroy(::nrg_chem_t *h) { nrg_chem_destroy(H(h)); }
static void *op_stream(::nrg_chem_t *h) { return nrg_chem_stream(H(h)); }
#undef H

}  // namespace NRG_TOPO_NS
#pragma GCC visibility pop

__attribute__((visibility("hidden"))) const nrg_topo_ops *NRG_TOPO_OPS_FN()
{
  using namespace NRG_TOPO_NS;
#define NRG_TOPO_OP_INIT(name, params, args) t.name = op_##name;
  static const nrg_topo_ops table = [] {
    nrg_topo_ops t{};
    t.name = NRG_TOPO_NAME;
    t.check = op_check;
    t.create = op_create;
    t.destroy = op_destroy;
    t.stream = op_stream;
    NRG_TOPO_ENTRY_POINTS(NRG_TOPO_OP_INIT)
    return t;
  }();
#undef NRG_TOPO_OP_INIT
  return &table;
}
#endif  // NRG_PART != 2
