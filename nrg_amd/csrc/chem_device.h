// chem_device.h -- device code of the gfx950 chemistry step.
//
// One cell per lane.  A cell's Gear-BDF state (Nordsieck history, iteration
// matrix, work vectors) lives in VGPRs/AGPRs; its 2R rate constants live in a
// per-lane LDS slot; the mechanism's numeric coefficients are one LDS copy per
// block.  Every register-array index is a compile-time constant after
// unrolling; the few order- or pivot-dependent accesses are select chains.
//
// SLATEC's DDRIV3 is restated as a reverse-communication state machine
// (advance()): a lane runs until it needs a right-hand side f(y) or a
// Jacobian, then the wave evaluates f for all lanes that asked at one code
// site, or -- when enough lanes wait for one -- runs a batched
// finite-difference Jacobian + LU for those lanes (jac_batch()).  Decisions
// and arithmetic follow the reference operation by operation (this TU is
// built with -ffp-contract=off), so results agree with NRG's Fortran step to
// the rounding of the transcendental functions.
//
// Reference: computing_module/src/current_build/chemical_kinetics_solver.f90
// (rates :360-493, RHS :495-540, cell step :228-346) and SLATEC SDRIVE in
// computing_module/src/third_party/slatec/ (ddriv3.f90, ddstp.f90, ddntl.f90,
// ddpst.f90, ddcor.f90, ddpsc.f90, ddscl.f90, ddntp.f90, dgefa.f90,
// dgesl.f90, dnrm2.f90).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

// the reaction topology these kernels are built for (one build per topology)
#ifndef NRG_TOPOLOGY_HEADER
#define NRG_TOPOLOGY_HEADER "topology_keromnes.h"
#endif
#include NRG_TOPOLOGY_HEADER

#define NRG_DI __device__ __forceinline__

namespace nrg {

constexpr int S = topo::S;
constexpr int R = topo::R;
constexpr int MAXORD = 5;
constexpr int WAVE = 64;

constexpr double R_GAS_J = 8.31446261815324;   // package_library/src/global_data.f90:36
constexpr double T_REF = 298.15;               // :38
constexpr double T_CEIL = 6000.0;              // :44

// Mechanism coefficients, staged once per block in LDS.  Quotients by
// constants that the reference evaluates per call (a_i/i, (-E)/R_u, P_atm/R_u)
// are precomputed on the host with the same single IEEE division.
struct MechCoef {
  double A[R], beta[R], A_low[R], beta_low[R];
  double mE_R[R], mElow_R[R];    // (-E_act)/R_u, (-E_act_low)/R_u
  double troe[4][R];             // alpha, T***, T*, T**
  double one_m_alpha[R];         // 1 - alpha
  double eff[R][S];              // enhanced efficiencies
  double a[S][2][7];             // NASA-7 x R_u; [s][set][i], set 1 = T >= 1000 K
  double hdiv[S][2][5];          // a5/5, a4/4, a3/3, a2/2, a1/1
  double sdiv[S][2][3];          // a5/4, a4/3, a3/2
  double W[S];                   // molar masses
  double h_ref[S];               // h_s(T_ref), T_ref = 298.15 K
  double el[MAXORD][MAXORD + 1]; // el[nq-1][j-1] = EL(j, nq)   (slatec/ddcst.f90, MINT = 2)
  double tq[MAXORD][3];          // tq[nq-1][k-1] = TQ(k, nq)
  double patm_r;                 // P_atm / R_u
  double lnA[R], lnA_low[R];     // ln A, ln A_low (the explicit-rate screen's exponents)
  double ln_patm_r;              // ln(P_atm / R_u)
};

// The part of the mechanism the integrator kernel reads (one LDS copy per block).
struct BdfCoef {
  double el[MAXORD][MAXORD + 1];
  double tq[MAXORD][3];
  double W[S];
  double h_ref[S];
};

// Integrator constants of one call (reference literals + D1MACH-derived values).
struct BdfConst {
  double eps, ewt, hmax, uround, sqrtn, sqrt_uround;
  double br, bl, bp, facmin;     // u**.875, u**.75, u**-.15, u**.78 (slatec/ddpst.f90:54-57)
  double t_gate;
  int mxstep, maxord;
  int gather_pct;
  int refill_idle;               // persistent form: idle lanes (of 64) that trigger a refill
};

NRG_DI double dmax(double a, double b) { return (a >= b) ? a : b; }
NRG_DI double dmin(double a, double b) { return (a <= b) ? a : b; }

// One out-of-line copy of each transcendental for the integrator kernel,
// which calls pow a few times per step from several sites; inlined, every
// call site would carry its own 100+-instruction expansion.
__device__ __noinline__ double nl_pow(double x, double y) { return pow(x, y); }

// Three independent powers in one call: the order-selection candidates of a
// step, whose dependent chains interleave here instead of running back to back.
struct Pow3 {
  double a, b, c;
};
__device__ __noinline__ Pow3 nl_pow3(double xa, double ya, double xb, double yb, double xc, double yc)
{
  return Pow3{pow(xa, ya), pow(xb, yb), pow(xc, yc)};
}

// Math for the rate-constant kernel, where inlining is the better trade.
// Cases whose IEEE result is exact skip the expansion: pow(x, +-0) = 1,
// pow(x, 1) = x, exp(+-0) = 1; exp(x) rounds to 1 for |x| < 2^-54 and to +0
// below -746 (under half the least subnormal).  The mechanism's zero
// temperature exponents and activation energies and its disabled Troe terms
// (T*** = 0, T* = T** = 1e30) hit them for every cell alike.
struct InlineMath {
  static NRG_DI double pow_(double x, double y)
  {
    if (y == 0.0) return 1.0;
    if (y == 1.0) return x;
    return pow(x, y);
  }
  static NRG_DI double exp_(double x)
  {
    if (fabs(x) < 0x1p-60) return 1.0;
    if (x < -746.0) return 0.0;
    return exp(x);
  }
  static NRG_DI double log_(double x) { return log(x); }
  static NRG_DI double log10_(double x) { return log10(x); }
  // x / T and x / R_u, as the reference divides
  static NRG_DI double over_T(double x, double T) { return x / T; }
  static NRG_DI double over_R(double x) { return x / R_GAS_J; }
  // k_r = k_f / K_c with K_c = exp(lnKp) * x**d_nu (d_nu = 0: pow(x, 0) = 1 exactly)
  static NRG_DI double reverse_(double kf, double lnKp, double x, double d_nu)
  {
    const double Kp = exp_(lnKp);
    const double Kc = (d_nu == 0.0) ? Kp * 1.0 : Kp * pow_(x, d_nu);
    return kf / Kc;
  }
  // A T^b exp(mE_R / T), in the reference's order of operations
  static NRG_DI double arrh_(double A, double b, double mE_R, double T)
  {
    return (A * pow_(T, b)) * exp_(mE_R / T);
  }
};

// Cheaper transcendentals for estimates that no parity path depends on (the
// surrogate gate's rate-consistency test): x^y as exp(y ln x) on positive x,
// so ln T and ln(P_atm/(R_u T)) are computed once per cell and shared by every
// reaction instead of a full pow per rate constant, and an Arrhenius term as
// one exp.  Relative error ~1e-15 for the exponents of a mechanism
// (|y ln x| < 50).
struct EstimateMath {
  static NRG_DI double pow_(double x, double y)
  {
    if (y == 0.0) return 1.0;
    if (y == 1.0) return x;
    if (y == -1.0) return 1.0 / x;
    return exp(y * log(x));
  }
  static NRG_DI double exp_(double x) { return exp(x); }
  static NRG_DI double log_(double x) { return log(x); }
  static NRG_DI double log10_(double x) { return log10(x); }
  // divisions by T and R_u as products with one reciprocal per cell (1/T is
  // the same expression at every call, formed once) and a constant
  static NRG_DI double over_T(double x, double T) { return x * (1.0 / T); }
  static NRG_DI double over_R(double x) { return x * (1.0 / R_GAS_J); }
  // k_f / K_c as k_f exp(-lnKp) x**(-d_nu): no division
  static NRG_DI double reverse_(double kf, double lnKp, double x, double d_nu)
  {
    return kf * exp(-lnKp) * ((d_nu == 0.0) ? 1.0 : pow_(x, -d_nu));
  }
  static NRG_DI double arrh_(double A, double b, double mE_R, double T)
  {
    return A * exp(b * log(T) + mE_R * (1.0 / T));
  }
};

// A value pinned in a VGPR.  Select chains over register arrays go through it,
// so the optimiser cannot turn "select between loads" into one load through a
// selected address (which would keep the whole array in private memory).
NRG_DI double opaque(double x)
{
  asm("" : "+v"(x));
  return x;
}

// ---------------------------------------------------------------------------
// Compile-time facts about the topology
// ---------------------------------------------------------------------------
// Does fun (chemical_kinetics_solver.f90:509-527) multiply by [M] for reaction r?
constexpr bool m_in_rhs(int r)
{
  for (int c = 0; c < 3; ++c)
    if (c < topo::nl[r] && topo::left[r][c] == S) return true;
  for (int c = 0; c < 3; ++c)
    if (c < topo::nr[r] && topo::right[r][c] == S) return true;
  return false;
}
constexpr int m_slot(int r)
{
  int k = 0;
  for (int q = 0; q < r; ++q) k += m_in_rhs(q) ? 1 : 0;
  return m_in_rhs(r) ? k : -1;
}
struct MSlotTable {
  int v[R];
};
constexpr MSlotTable make_mslot()
{
  MSlotTable t{};
  for (int r = 0; r < R; ++r) t.v[r] = m_slot(r);
  return t;
}
constexpr MSlotTable kMSlot = make_mslot();   // [M] register slot of reaction r, or -1
constexpr int count_m()
{
  int k = 0;
  for (int r = 0; r < R; ++r) k += m_in_rhs(r) ? 1 : 0;
  return k;
}
constexpr int NM = count_m() > 0 ? count_m() : 1;
// [M] needed at setup: third-body and fall-off reactions.
constexpr bool uses_m(int r)
{
  return topo::rtype[r] == 1 || topo::rtype[r] == 2 || topo::rtype[r] == 3 ||
         topo::rtype[r] == 6 || topo::rtype[r] == 7;
}

// Cell record: what rates_kernel hands the integrator for one queue position
// t of a chunk of nrec positions, as pairs of doubles (16 bytes: one lane's
// share of a direct HBM -> LDS load, global_load_lds_dwordx4).
//   pairs [0, R)             (k_f(r), k_r(r))
//   pairs [REC_M, REC_C)     [M] slots 2j, 2j+1
//   pairs [REC_C, REC_PAIRS) C_s (mol/m^3) at slot s, then the gate flag at
//                            slot S (1: integrate, 0: gated)
// In HBM pair p of position t sits at rec[2 (p nrec + t)] (pair-major: a
// wave's lanes read 1 KiB runs); the integrator keeps the lane's copy in LDS,
// pair p of lane l at recl[2 (p WAVE + l)].
constexpr int REC_M = R;
constexpr int REC_C = REC_M + (NM + 1) / 2;
constexpr int REC_PAIRS = REC_C + (S + 2) / 2;
constexpr int REC_FLAG = S;
// Offset (doubles) of slot j of the pairs starting at pair p0, position 0.
NRG_DI long long rec_slot(int p0, int j, long long nrec)
{
  return 2 * ((long long)(p0 + j / 2) * nrec) + (j & 1);
}
// The lane's LDS copy (kl = recl + 2 * lane).
NRG_DI double rec_kf(const double *kl, int r) { return kl[r * 2 * WAVE]; }
NRG_DI double rec_kr(const double *kl, int r) { return kl[r * 2 * WAVE + 1]; }
NRG_DI double rec_m(const double *kl, int j) { return kl[(REC_M + j / 2) * 2 * WAVE + (j & 1)]; }
NRG_DI double rec_c(const double *kl, int s) { return kl[(REC_C + s / 2) * 2 * WAVE + (s & 1)]; }

// Species that no reaction consumes or produces (N2 in H2-air) enter only
// through the third-body concentrations, which are fixed per cell.  For them
// f_s = 0 and no f_i depends on y_s, so under DDRIV3 every correction, every
// YH column beyond the first and every norm term of theirs is an exact zero
// (Lawson's norm passes zeros through unchanged) and their row and column of
// P are those of the identity (the LU and the solves of the other species do
// not see them).  The integrator state therefore holds the SA active species;
// an inert species keeps its initial concentration.
constexpr bool reacts(int s)
{
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < 3; ++c)
      if ((c < topo::nl[r] && topo::left[r][c] == s) || (c < topo::nr[r] && topo::right[r][c] == s))
        return true;
  return false;
}
constexpr int count_active()
{
  int n = 0;
  for (int s = 0; s < S; ++s) n += reacts(s) ? 1 : 0;
  return n;
}
constexpr int SA = count_active();
constexpr int NI = S - SA;
struct SpeciesMap {
  int act[S];                    // active slot of species s, or -1
  int full[SA];                  // species of active slot a
  int islot[S];                  // inert slot of species s, or -1
  int inert[NI > 0 ? NI : 1];    // species of inert slot i
};
constexpr SpeciesMap make_species_map()
{
  SpeciesMap m{};
  int a = 0, i = 0;
  for (int s = 0; s < S; ++s) {
    m.act[s] = -1;
    m.islot[s] = -1;
    if (reacts(s)) {
      m.act[s] = a;
      m.full[a++] = s;
    } else {
      m.islot[s] = i;
      m.inert[i++] = s;
    }
  }
  return m;
}
constexpr SpeciesMap kSp = make_species_map();
constexpr int act_of(int s) { return (s >= 0 && s < S && kSp.act[s] >= 0) ? kSp.act[s] : 0; }

// ---------------------------------------------------------------------------
// NASA-7 thermochemistry (thermophysical_properties_class.f90:225-300)
// ---------------------------------------------------------------------------
NRG_DI double clip_t(double T) { return dmin(T, T_CEIL); }

template <class M = InlineMath>
NRG_DI double h_molar(const MechCoef &mc, double T, int s)
{
  const double temp = clip_t(T);
  const int set = (temp >= 1000.0) ? 1 : 0;
  const double *d = mc.hdiv[s][set];
  double hs = d[0];
  hs = hs * temp + d[1];
  hs = hs * temp + d[2];
  hs = hs * temp + d[3];
  hs = hs * temp + d[4];
  hs = hs + M::over_T(mc.a[s][set][5], temp);
  return hs * temp;
}

template <class M = InlineMath>
NRG_DI double s_molar(const MechCoef &mc, double T, int s)
{
  const double temp = clip_t(T);
  const int set = (temp >= 1000.0) ? 1 : 0;
  const double *d = mc.sdiv[s][set];
  const double *a = mc.a[s][set];
  double sm = d[0];
  sm = sm * temp + d[1];
  sm = sm * temp + d[2];
  sm = sm * temp + a[1];
  return (sm * temp + a[6]) + a[0] * M::log_(temp);
}

NRG_DI double cp_molar(const MechCoef &mc, double T, int s)
{
  const double temp = clip_t(T);
  const double *a = mc.a[s][(temp >= 1000.0) ? 1 : 0];
  double cp = a[4];
  cp = cp * temp + a[3];
  cp = cp * temp + a[2];
  cp = cp * temp + a[1];
  cp = cp * temp + a[0];
  return cp;
}

NRG_DI double e_molar(const MechCoef &mc, double T, int s)
{
  return h_molar(mc, T, s) - R_GAS_J * T;   // unclipped T (:298)
}

// ---------------------------------------------------------------------------
// Rate constants (chemical_kinetics_solver.f90:360-493) of one cell, written
// into its cell record (below): k_f(r) and k_r(r) as pair r, and [M] of the
// reactions whose RHS term uses it (:235-248) in the [M] pairs.
// ---------------------------------------------------------------------------
// Reactions [R0, R1) only (their k_f, k_r and [M] slots).
// k_f, k_r and [M] of reaction r at T (already capped at 1e4 K), from the
// species' entropies sv and enthalpies hv at T (:394-485).
template <class M>
NRG_DI void reaction_k(const MechCoef &mc, double T, const double (&sv)[S], const double (&hv)[S],
                       const double (&C)[S], int r, double &kf_out, double &kr_out, double &m_out)
{
  double m = 0.0;
  if (uses_m(r)) {
#pragma unroll
    for (int s = 0; s < S; ++s) m = m + C[s] * mc.eff[r][s];
  }
  double d_s = 0.0, d_h = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int sp = topo::right[r][c];
    if (c < topo::nr[r] && sp != S) {
      d_s = d_s + sv[sp < S ? sp : 0];
      d_h = d_h + hv[sp < S ? sp : 0];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int sp = topo::left[r][c];
    if (c < topo::nl[r] && sp != S) {
      d_s = d_s - sv[sp < S ? sp : 0];
      d_h = d_h - hv[sp < S ? sp : 0];
    }
  }
  const double d_nu = (double)(topo::nr[r] - topo::nl[r]);
  const int rt = topo::rtype[r];
  double kf = 0.0, kr = 0.0;
  if (rt == 0 || rt == 1 || rt == 5 || rt == 6) {
    kf = M::arrh_(mc.A[r], mc.beta[r], mc.mE_R[r], T);
  } else if (rt == 2 || rt == 7) {
    const double k0 = M::arrh_(mc.A_low[r], mc.beta_low[r], mc.mElow_R[r], T);
    const double kinf = M::arrh_(mc.A[r], mc.beta[r], mc.mE_R[r], T);
    const double pr = (k0 * m) / kinf;
    kf = kinf * (pr / (1.0 + pr));
  } else if (rt == 3) {
    // (A T^b) exp(-E/RT): the product the reference forms after its two exps
    const double k0 = M::arrh_(mc.A_low[r], mc.beta_low[r], mc.mElow_R[r], T);
    const double kinf = M::arrh_(mc.A[r], mc.beta[r], mc.mE_R[r], T);
    const double pr = (k0 * m) / kinf;
    const double fcent = (mc.one_m_alpha[r] * M::exp_(-(T / mc.troe[1][r])) +
                          mc.troe[0][r] * M::exp_(-(T / mc.troe[2][r]))) +
                         M::exp_(-M::over_T(mc.troe[3][r], T));
    const double lf = M::log10_(fcent);
    const double c = -0.4 - 0.67 * lf;
    const double n = 0.75 - 1.27 * lf;
    const double d = 0.14;
    double blend;
    if (pr != 0.0) {
      const double lp = M::log10_(pr);
      const double x = (lp + c) / (n - d * (lp + c));
      blend = M::pow_(fcent, 1.0 / (1.0 + x * x));
    } else {
      blend = M::pow_(fcent, -1.0 / d);
    }
    kf = (kinf * (pr / (1.0 + pr))) * blend;
  }
  if (rt == 0 || rt == 1 || rt == 2 || rt == 3) {
    // K_c = K_p * ((P_atm/R_u)/T)**d_nu, K_p = exp(dS/R - dH/(R T))
    kr = M::reverse_(kf, M::over_R(d_s) - M::over_T(M::over_R(d_h), T), M::over_T(mc.patm_r, T), d_nu);
  }
  kf_out = kf;
  kr_out = kr;
  m_out = m;
}

template <class M, int R0 = 0, int R1 = R>
NRG_DI void rate_constants(const MechCoef &mc, double Tin, const double (&C)[S], double *rec,
                           long long nrec)
{
  const double T = dmin(Tin, 10000.0);
  double sv[S], hv[S];
#pragma unroll
  for (int c = 0; c < S; ++c) {
    sv[c] = s_molar<M>(mc, T, c);
    hv[c] = h_molar(mc, T, c);
  }
#pragma unroll
  for (int r = R0; r < R1; ++r) {
    double kf, kr, m;
    reaction_k<M>(mc, T, sv, hv, C, r, kf, kr, m);
    if (uses_m(r) && kMSlot.v[r] >= 0) rec[rec_slot(REC_M, kMSlot.v[r], nrec)] = m;
    *reinterpret_cast<double2 *>(rec + rec_slot(r, 0, nrec)) = make_double2(kf, kr);
  }
}

// ---------------------------------------------------------------------------
// Mass-action right-hand side, the accumulation order of fun (:507-536), on
// the active species (an inert species' f is zero).
// ---------------------------------------------------------------------------
// Mass-action contribution of reaction r (:520-535) with rate constants kf,
// kr and third-body concentration m.
NRG_DI void rhs_reaction(int r, double kf, double kr, double m, const double (&Z)[SA], double (&Zd)[SA])
{
  double fwd = kf;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int sp = topo::left[r][c];
    if (c < topo::nl[r]) fwd = fwd * ((sp == S) ? m : Z[act_of(sp)]);
  }
  double rev = kr;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int sp = topo::right[r][c];
    if (c < topo::nr[r]) rev = rev * ((sp == S) ? m : Z[act_of(sp)]);
  }
#pragma unroll
  for (int a = 0; a < SA; ++a) {
    const int s = kSp.full[a];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (c < topo::nl[r] && topo::left[r][c] == s) Zd[a] = (Zd[a] - fwd) + rev;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (c < topo::nr[r] && topo::right[r][c] == s) Zd[a] = (Zd[a] - rev) + fwd;
  }
}

// Rate constants and [M] from the lane's cell record in LDS (kl).
NRG_DI void rhs(const double *kl, const double (&Z)[SA], double (&Zd)[SA])
{
#pragma unroll
  for (int s = 0; s < SA; ++s) Zd[s] = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r)
    rhs_reaction(r, rec_kf(kl, r), rec_kr(kl, r), kMSlot.v[r] >= 0 ? rec_m(kl, kMSlot.v[r]) : 0.0,
                 Z, Zd);
}

// Species (bit s = species s) on one side of reaction r, third body left out.
constexpr unsigned side_mask(int r, int side)
{
  unsigned m = 0;
  for (int c = 0; c < 3; ++c) {
    const int sp = side ? topo::right[r][c] : topo::left[r][c];
    const int n = side ? topo::nr[r] : topo::nl[r];
    if (c < n && sp >= 0 && sp < S) m |= 1u << sp;
  }
  return m;
}
constexpr bool has_reverse(int r)
{
  return topo::rtype[r] == 0 || topo::rtype[r] == 1 || topo::rtype[r] == 2 || topo::rtype[r] == 3;
}
static_assert(S <= 32, "species masks are 32-bit");

// The part of omega(C, T) of the active species that reactions [R0, R1)
// contribute, with the rate constants computed one reaction at a time (no
// array of all k_f, k_r is held).  A reaction whose left side holds a species
// the cell does not contain has a zero forward rate, and likewise for the
// right side and the reverse rate; when both are zero for every lane of the
// wave, the reaction adds nothing and is skipped (fresh gas, whose radicals
// are exact zeros, needs a handful of the reactions).
template <class Mth, int R0 = 0, int R1 = R>
NRG_DI void explicit_rates(const MechCoef &mc, double Tin, const double (&C)[S], double (&Zd)[SA])
{
  const double T = dmin(Tin, 10000.0);
  double sv[S], hv[S], Z[SA];
  unsigned absent = 0;
#pragma unroll
  for (int c = 0; c < S; ++c) {
    sv[c] = s_molar<Mth>(mc, T, c);
    hv[c] = h_molar<Mth>(mc, T, c);
    absent |= (C[c] == 0.0) ? 1u << c : 0u;
  }
#pragma unroll
  for (int i = 0; i < SA; ++i) {
    Z[i] = C[kSp.full[i]];
    Zd[i] = 0.0;
  }
#pragma unroll
  for (int r = R0; r < R1; ++r) {
    const bool fwd = (absent & side_mask(r, 0)) == 0;
    const bool rev = has_reverse(r) && (absent & side_mask(r, 1)) == 0;
    if (__any((fwd || rev) ? 1 : 0)) {
      double kf, kr, m;
      reaction_k<Mth>(mc, T, sv, hv, C, r, kf, kr, m);
      rhs_reaction(r, kf, kr, m, Z, Zd);
    }
  }
}

// ---------------------------------------------------------------------------
// DNRM2 (slatec/dnrm2.f90:84-160): Lawson's phases as a per-element state machine.
// st: 0 = start (label 30), 1 = phase 1 (50), 2 = phase 2 (70), 3 = phase 4 (110),
//     4 = phase 3 (the DO 95 loop, unscaled sum).
// One out-of-line copy serves the call sites; the vector arrives in argument
// VGPRs.  The inert species' terms, exact zeros, are left out: in every phase
// a zero leaves SUM, XMAX and the phase as they were.  HITEST keeps DDRIV3's
// N (all species).
// ---------------------------------------------------------------------------
template <int n>
NRG_DI double dnrm2_core(const double *x)
{
  const double cutlo = 8.232e-11;
  const double hitest = 1.304e19 / (double)S;
  double sum = 0.0, xmax = 0.0;
  int st = 0;
#pragma unroll
  for (int i = 0; i < n; ++i) {
    const double v = x[i], av = fabs(v);
    // the usual case first (phase 3, no overflow risk): what the state
    // machine below does for it, without walking through the other phases
    if (st == 4 && av < hitest) {
      sum = sum + v * v;
      continue;
    }
    bool done = false;
    if (st == 0) {
      if (av > cutlo) {
        st = 4;
      } else {
        st = 1;
        xmax = 0.0;
      }
    }
    if (st == 1) {
      if (v == 0.0) {
        done = true;
      } else if (av > cutlo) {
        st = 4;
      } else {
        st = 2;
        xmax = av;
        const double r = v / xmax;
        sum = sum + r * r;
        done = true;
      }
    }
    if (!done && st == 2) {
      if (av > cutlo) {
        sum = (sum * xmax) * xmax;
        st = 4;
      } else {
        if (av <= xmax) {
          const double r = v / xmax;
          sum = sum + r * r;
        } else {
          const double r = xmax / v;
          sum = 1.0 + sum * (r * r);
          xmax = av;
        }
        done = true;
      }
    }
    if (!done && st == 3) {
      if (av <= xmax) {
        const double r = v / xmax;
        sum = sum + r * r;
      } else {
        const double r = xmax / v;
        sum = 1.0 + sum * (r * r);
        xmax = av;
      }
      done = true;
    }
    if (!done && st == 4) {
      if (av >= hitest) {
        sum = (sum / v) / v;
        xmax = av;
        const double r = v / xmax;
        sum = sum + r * r;
        st = 3;
      } else {
        sum = sum + v * v;
      }
    }
  }
  return (st == 4) ? sqrt(sum) : xmax * sqrt(sum);
}

template <class... D>
__device__ __noinline__ double dnrm2_vals(D... xs)
{
  const double x[sizeof...(D)] = {xs...};
  return dnrm2_core<(int)sizeof...(D)>(x);
}

// Two norms in one call (independent state machines, interleaved).
struct Norm2 {
  double a, b;
};
template <class... D>
__device__ __noinline__ Norm2 dnrm2_pair_vals(D... xs)
{
  constexpr int n = (int)sizeof...(D) / 2;
  const double v[2 * n] = {xs...};
  return Norm2{dnrm2_core<n>(v), dnrm2_core<n>(v + n)};
}

template <int N, int... I>
NRG_DI double dnrm2_seq(const double (&x)[N], std::integer_sequence<int, I...>)
{
  return dnrm2_vals(x[I]...);
}

// Norm of a vector over the active species.
NRG_DI double dnrm2(const double (&x)[SA])
{
  return dnrm2_seq<SA>(x, std::make_integer_sequence<int, SA>{});
}

template <int... I>
NRG_DI Norm2 dnrm2_pair_seq(const double (&x)[SA], const double (&y)[SA], std::integer_sequence<int, I...>)
{
  return dnrm2_pair_vals(x[I]..., y[I]...);
}

// Norms of two vectors over the active species.
NRG_DI Norm2 dnrm2_pair(const double (&x)[SA], const double (&y)[SA])
{
  return dnrm2_pair_seq(x, y, std::make_integer_sequence<int, SA>{});
}

// Norm of a vector over all species.
NRG_DI double dnrm2_all(const double (&x)[S])
{
  return dnrm2_seq<S>(x, std::make_integer_sequence<int, S>{});
}

// ---------------------------------------------------------------------------
// 9x9 LU with partial pivoting (DGEFA) and solve (DGESL), column-major a[col][row].
// Row interchanges use select chains; a wave skips them when no lane pivots.
// ---------------------------------------------------------------------------
NRG_DI bool wave_any(bool p) { return __any(p ? 1 : 0) != 0; }

// y[i] += t * x[i] for i in [LO, SA), left alone on the lanes with `skip` set
// (DAXPY returns at once for DA == 0, slatec/daxpy.f90:68; adding 0*x is not
// the same when x is not finite or y is -0).  The wave takes the select-free
// form when no lane skips and drops the loop when every lane does; both
// happen for whole columns (a species that no rate depends on gives exact
// zeros in P and in the corrector's right-hand side).
template <int LO>
NRG_DI void daxpy_lanes(bool skip, double t, const double (&x)[SA], double (&y)[SA])
{
  if (!wave_any(!skip)) return;
  if (wave_any(skip)) {
#pragma unroll
    for (int i = LO; i < SA; ++i) y[i] = skip ? y[i] : y[i] + t * x[i];
  } else {
#pragma unroll
    for (int i = LO; i < SA; ++i) y[i] = y[i] + t * x[i];
  }
}

// As daxpy_lanes over [0, HI).
template <int HI>
NRG_DI void daxpy_head(bool skip, double t, const double (&x)[SA], double (&y)[SA])
{
  if (!wave_any(!skip)) return;
  if (wave_any(skip)) {
#pragma unroll
    for (int i = 0; i < HI; ++i) y[i] = skip ? y[i] : y[i] + t * x[i];
  } else {
#pragma unroll
    for (int i = 0; i < HI; ++i) y[i] = y[i] + t * x[i];
  }
}

template <int K>
NRG_DI void gefa_col(double (&a)[SA][SA], int (&ipvt)[SA], int &info)
{
  int l = K;
  double dmx = fabs(a[K][K]), piv = a[K][K];
#pragma unroll
  for (int i = K + 1; i < SA; ++i) {
    const double xm = fabs(a[K][i]);
    if (xm > dmx) {
      l = i;
      dmx = xm;
      piv = a[K][i];
    }
  }
  ipvt[K] = l;
  const bool zero = (piv == 0.0);
  if (zero) info = K + 1;
  const bool anyswap = wave_any(l != K);
  if (anyswap) {
#pragma unroll
    for (int i = K + 1; i < SA; ++i) a[K][i] = (l == i) ? a[K][K] : a[K][i];
    a[K][K] = piv;
  }
  // DSCAL of the column below the pivot (skipped for a zero pivot)
  const double t = -1.0 / a[K][K];
  if (wave_any(zero)) {
#pragma unroll
    for (int i = K + 1; i < SA; ++i) a[K][i] = zero ? a[K][i] : t * a[K][i];
  } else {
#pragma unroll
    for (int i = K + 1; i < SA; ++i) a[K][i] = t * a[K][i];
  }
#pragma unroll
  for (int j = K + 1; j < SA; ++j) {
    double tj = a[j][K];
    if (anyswap) {
#pragma unroll
      for (int i = K + 1; i < SA; ++i) tj = (l == i) ? a[j][i] : tj;
#pragma unroll
      for (int i = K + 1; i < SA; ++i) a[j][i] = (l == i) ? a[j][K] : a[j][i];
      a[j][K] = tj;
    }
    daxpy_lanes<K + 1>(zero || (tj == 0.0), tj, a[K], a[j]);
  }
}

template <int... K>
NRG_DI void gefa_cols(double (&a)[SA][SA], int (&ipvt)[SA], int &info, std::integer_sequence<int, K...>)
{
  (gefa_col<K>(a, ipvt, info), ...);
}

NRG_DI int gefa(double (&a)[SA][SA], int (&ipvt)[SA])
{
  int info = 0;
  gefa_cols(a, ipvt, info, std::make_integer_sequence<int, SA - 1>{});
  ipvt[SA - 1] = SA - 1;
  if (a[SA - 1][SA - 1] == 0.0) info = SA;
  return info;
}

template <int K>
NRG_DI void gesl_fwd(const double (&a)[SA][SA], const int (&ipvt)[SA], double (&b)[SA])
{
  const int l = ipvt[K];
  double t = b[K];
  if (wave_any(l != K)) {
#pragma unroll
    for (int i = K + 1; i < SA; ++i) t = (l == i) ? b[i] : t;
#pragma unroll
    for (int i = K + 1; i < SA; ++i) b[i] = (l == i) ? b[K] : b[i];
    b[K] = t;
  }
  daxpy_lanes<K + 1>(t == 0.0, t, a[K], b);
}

template <int K>
NRG_DI void gesl_back(const double (&a)[SA][SA], double (&b)[SA])
{
  b[K] = b[K] / a[K][K];
  const double t = -b[K];
  daxpy_head<K>(t == 0.0, t, a[K], b);
}

template <int... K>
NRG_DI void gesl_steps(const double (&a)[SA][SA], const int (&ipvt)[SA], double (&b)[SA],
                       std::integer_sequence<int, K...>)
{
  (gesl_fwd<K>(a, ipvt, b), ...);
  (gesl_back<SA - 1 - K>(a, b), ...);
  gesl_back<SA - 1 - (int)sizeof...(K)>(a, b);
}

NRG_DI void gesl(const double (&a)[SA][SA], const int (&ipvt)[SA], double (&b)[SA])
{
  gesl_steps(a, ipvt, b, std::make_integer_sequence<int, SA - 1>{});
}

// ---------------------------------------------------------------------------
// Gear BDF state of one cell (DDRIV3's WORK/IWORK, private to the lane) and
// the program counter of the reverse-communication form.
// ---------------------------------------------------------------------------
enum Pc : int {
  PC_260 = 0,      // DDRIV3 label 260: weights and guards, then DDSTP entry
  PC_NTL_F,        // DDNTL after its f(y)
  PC_ATTEMPT,      // DDSTP label 100: predict, then f(y_pred)
  PC_PRED_F,       // after f(y_pred) (label 110)
  PC_JAC_DONE,     // after the batched FD Jacobian + LU (DDPST)
  PC_CORR,         // label 130: corrector (DDCOR)
  PC_CORR_F,       // after f(y) of a corrector iteration
  PC_FAIL_CONV,    // label 160: corrector failure
  PC_ERRTEST,      // label 170: error test, accept, step/order selection
  PC_RETRACT,      // retract the prediction after a failed attempt
  PC_SCALE,        // apply a pending DDSCL rescale, then continue
  PC_DONE
};
enum Req : int { REQ_NONE = 0, REQ_F = 1, REQ_JAC = 2, REQ_DONE = 3 };
enum CellOutcome { OUT_OK = 0, OUT_MXSTEP = 2, OUT_STEP_ZERO = 3, OUT_REPEATED_FAIL = 4,
                   OUT_SINGULAR = 5, OUT_EPS = 6 };

struct Cell {
  double yh[MAXORD + 1][SA];   // yh[j][i] = YH(i, j+1)
  double y[SA], ywt[SA], save1[SA], save2[SA], fac[SA];
  double P[SA][SA];             // iteration matrix / LU factors, column-major
  int ipvt[SA];
  double yi[NI > 0 ? NI : 1];   // concentrations of the inert species (constant)
  double t, h, rc, rmax, trend, told, d1, ctest;
  int nq, nwait, jtask, jstepl, nstep, nfe, nje, ncorr, natt;
  int ntry, nfail, iter, pc, req, outcome;
  bool convrg, ier, evaljc, ntl_restart;
  // pending retraction / rescale (stages R and S)
  double rh, etest;
  bool fail_err, do_scale, rmax_after, after_accept;
};

NRG_DI double el_of(const BdfCoef &mc, int j, int nq) { return mc.el[nq - 1][j - 1]; }
NRG_DI double tq_of(const BdfCoef &mc, int k, int nq) { return mc.tq[nq - 1][k - 1]; }

// DDPSC (slatec/ddpsc.f90:26-38) for a compile-time order.
template <int NQ, bool ADD>
NRG_DI void psc_n(double (&yh)[MAXORD + 1][SA])
{
#pragma unroll
  for (int j1 = 1; j1 <= NQ; ++j1)
#pragma unroll
    for (int j2 = j1; j2 <= NQ; ++j2) {
      const int j = NQ - j2 + j1;
#pragma unroll
      for (int i = 0; i < SA; ++i) yh[j - 1][i] = ADD ? yh[j - 1][i] + yh[j][i] : yh[j - 1][i] - yh[j][i];
    }
}

template <bool ADD>
NRG_DI void psc(int nq, double (&yh)[MAXORD + 1][SA])
{
  switch (nq) {
    case 1: psc_n<1, ADD>(yh); break;
    case 2: psc_n<2, ADD>(yh); break;
    case 3: psc_n<3, ADD>(yh); break;
    case 4: psc_n<4, ADD>(yh); break;
    default: psc_n<5, ADD>(yh); break;
  }
}

// YH(:, col+1) read through a select chain (col in 1..MAXORD).
NRG_DI void yh_col(const Cell &c, int col, double (&out)[SA])
{
#pragma unroll
  for (int i = 0; i < SA; ++i) {
    double v = opaque(c.yh[1][i]);
#pragma unroll
    for (int j = 2; j <= MAXORD; ++j) v = (col == j) ? opaque(c.yh[j][i]) : v;
    out[i] = v;
  }
}

// DDSCL (slatec/ddscl.f90:24-35).
NRG_DI void scl(Cell &c, double rh, const BdfConst &k)
{
  const double ah = fabs(c.h);
  if (c.h < 1.0)
    rh = dmin(dmin(ah * rh, ah * c.rmax), k.hmax) / ah;
  else
    rh = dmin(dmin(rh, c.rmax), k.hmax / ah);
  double r1 = 1.0;
#pragma unroll
  for (int j = 1; j <= MAXORD; ++j) {
    r1 = r1 * rh;
    const bool on = (j <= c.nq);
#pragma unroll
    for (int i = 0; i < SA; ++i) c.yh[j][i] = on ? c.yh[j][i] * r1 : c.yh[j][i];
  }
  c.h = c.h * rh;
  c.rc = c.rc * rh;
}

// DDCOR, MITER = 2, IMPL = 0, IERROR = 3 (slatec/ddcor.f90:51-54,98-108).
NRG_DI double cor(Cell &c, const BdfConst &k)
{
#pragma unroll
  for (int i = 0; i < SA; ++i) c.save2[i] = ((c.h * c.save2[i]) - c.yh[1][i]) - c.save1[i];
  gesl(c.P, c.ipvt, c.save2);
#pragma unroll
  for (int i = 0; i < SA; ++i) {
    c.save1[i] = c.save1[i] + c.save2[i];
    c.save2[i] = c.save2[i] / dmax(fabs(c.y[i]), c.ywt[i]);
  }
  c.ncorr++;
  return dnrm2(c.save2) / k.sqrtn;
}

// DDNTP, K = 0 (slatec/ddntp.f90:25-32).
NRG_DI void ntp(Cell &c, double tout)
{
  double col[SA];
  yh_col(c, c.nq, col);
#pragma unroll
  for (int i = 0; i < SA; ++i) c.y[i] = col[i];
  const double r = (tout - c.t) / c.h;
#pragma unroll
  for (int j = MAXORD; j >= 1; --j) {
    const bool on = (j <= c.nq);
#pragma unroll
    for (int i = 0; i < SA; ++i) c.y[i] = on ? c.yh[j - 1][i] + r * c.y[i] : c.y[i];
  }
}

NRG_DI void copy_yh0_to_y(Cell &c)
{
#pragma unroll
  for (int i = 0; i < SA; ++i) c.y[i] = c.yh[0][i];
}

NRG_DI void finish(Cell &c, int outcome)
{
  c.outcome = outcome;
  c.pc = PC_DONE;
  c.req = REQ_DONE;
}

NRG_DI void request_f(Cell &c, int resume)
{
  c.pc = resume;
  c.req = REQ_F;
}

// DDNTL up to its f call (slatec/ddntl.f90:52-63); EL/TQ are precomputed.
NRG_DI void ntl_head(Cell &c)
{
  c.ier = false;
  if (c.jtask == 0) c.rmax = 10000.0;
  c.rc = 0.0;
  c.convrg = false;
  c.trend = 1.0;
  c.nq = 1;
  c.nwait = 3;
}

// DDRIV3 init for NSTATE = 1 (slatec/ddriv3.f90:930-982): y holds C_in.
NRG_DI void bdf_init(Cell &c, double tout, const BdfConst &k)
{
  c.t = 0.0;
  double h = (tout - c.t) * (1.0 - 4.0 * k.uround);
  h = copysign(dmin(fabs(h), k.hmax), h);
  c.h = h;
  c.jtask = 0;
  c.nstep = c.nfe = c.nje = c.ncorr = c.natt = 0;
  c.jstepl = 0;
  c.ier = false;
  c.rc = 0.0;
  c.rmax = 0.0;
  c.trend = 0.0;
  c.nq = 1;
  c.nwait = 0;
  c.convrg = false;
  c.iter = 0;
  c.ntry = 0;
  c.nfail = 0;
#pragma unroll
  for (int i = 0; i < SA; ++i) c.yh[0][i] = c.y[i];
  c.pc = PC_260;
  c.req = REQ_NONE;
  c.outcome = OUT_OK;
}

// Stage guards: a lane runs a stage when it sits at that program counter and
// has not yet asked for anything in this pass; a stage's code is skipped by
// the whole wave when no lane needs it.
NRG_DI bool at(const Cell &c, int pc) { return c.req == REQ_NONE && c.pc == pc; }

// Run the cell's program until it needs f(y) (req = REQ_F, y set), a
// Jacobian (REQ_JAC) or is finished (REQ_DONE).  On REQ_F resumption the
// caller has put f(y) into save2.  hsign is +1 throughout (tout > 0).
//
// DDRIV3/DDSTP's control flow is laid out as seven stages in a fixed order,
// so that every transition taken in one pass goes forward:
//   1 DDNTL tail   2 after f / Jacobian   3 corrector (label 130)
//   4 error test, accept, step/order choice (label 170)
//   5 corrector failure (label 160)   6 DDRIV3 label 260   7 attempt (label 100)
// Each stage's code exists once; a lane may pass several stages per call.
NRG_DI void advance(Cell &c, double tout, const BdfConst &k, const BdfCoef &mc)
{
  // ---- 1: rest of DDNTL after its f(y) (slatec/ddntl.f90:68,152-161)
  if (wave_any(at(c, PC_NTL_F)) && at(c, PC_NTL_F)) {
    c.nfe++;
#pragma unroll
    for (int i = 0; i < SA; ++i) c.save1[i] = c.save2[i] / dmax(1.0, c.ywt[i]);
    const double sum = dnrm2(c.save1) / k.sqrtn;
    if (sum > k.eps / fabs(c.h)) c.h = copysign(k.eps / sum, c.h);
#pragma unroll
    for (int i = 0; i < SA; ++i) c.yh[1][i] = c.h * c.save2[i];
#pragma unroll
    for (int i = 0; i < SA; ++i) c.fac[i] = k.sqrt_uround;
    if (c.ntl_restart) c.rmax = 10.0;
    if (c.h == 0.0) {
      copy_yh0_to_y(c);
      finish(c, OUT_STEP_ZERO);
    } else {
      c.pc = PC_ATTEMPT;
    }
  }
  // ---- 2: after f(y_pred), after a corrector f, after the Jacobian
  if (wave_any(c.req == REQ_NONE && (c.pc == PC_PRED_F || c.pc == PC_JAC_DONE || c.pc == PC_CORR_F))) {
    if (at(c, PC_PRED_F)) {
      c.nfe++;
      if (c.evaljc || c.ier) {
        c.pc = PC_JAC_DONE;
        c.req = REQ_JAC;
      } else {
#pragma unroll
        for (int i = 0; i < SA; ++i) c.save1[i] = 0.0;
        c.pc = PC_CORR;
      }
    } else if (at(c, PC_JAC_DONE)) {
      if (c.ier) {
        c.pc = PC_FAIL_CONV;
      } else {
        c.convrg = false;
        c.rc = 1.0;
        c.jstepl = c.nstep;
#pragma unroll
        for (int i = 0; i < SA; ++i) c.save1[i] = 0.0;
        c.pc = PC_CORR;
      }
    } else if (at(c, PC_CORR_F)) {
      c.nfe++;
      c.pc = PC_CORR;
    }
  }
  // ---- 3: corrector iteration, label 130 (slatec/ddstp.f90:152-205)
  if (wave_any(at(c, PC_CORR)) && at(c, PC_CORR)) {
    const double d = cor(c, k);
    if (c.iter > 0) c.trend = dmax(0.9 * c.trend, d / c.d1);
    c.d1 = d;
    c.ctest = dmin(2.0 * c.trend, 1.0) * d;
    if (c.ctest <= k.eps) {
      c.pc = PC_ERRTEST;
    } else {
      c.iter++;
      if (c.iter < 3) {
        const double l0 = el_of(mc, 1, c.nq);
#pragma unroll
        for (int i = 0; i < SA; ++i) c.y[i] = c.yh[0][i] + l0 * c.save1[i];
        request_f(c, PC_CORR_F);
      } else if (c.convrg) {                           // stale Jacobian: refresh (label 110)
        c.evaljc = true;
        c.iter = 0;
        copy_yh0_to_y(c);
        request_f(c, PC_PRED_F);
      } else {
        c.pc = PC_FAIL_CONV;
      }
    }
  }
  // ---- 4: error test, accept, step/order selection, label 170 (:222-432)
  // A converged lane may wait here (it asks for nothing, so it re-enters on
  // the next pass unchanged) until gather_pct % of the lanes stepping in this
  // pass have arrived: the stage's code then runs once for all of them.  The
  // others always make progress, so the share is reached.
  const bool here4 = at(c, PC_ERRTEST);
  const int n4 = __popcll(__ballot(here4 ? 1 : 0));
  const int nact = __popcll(__ballot(1));
  if (n4 > 0 && n4 * 100 >= k.gather_pct * nact && here4) {
    c.convrg = true;
#pragma unroll
    for (int i = 0; i < SA; ++i) c.save2[i] = c.save1[i] / dmax(fabs(c.y[i]), c.ywt[i]);
    const double etest = dnrm2(c.save2) / (tq_of(mc, 2, c.nq) * k.sqrtn);
    c.etest = etest;
    if (etest > k.eps) {
      c.fail_err = true;                               // -> stage R (:238-289)
      c.pc = PC_RETRACT;
    } else {
      // accepted step (slatec/ddstp.f90:291-300)
      c.nstep++;
#pragma unroll
      for (int j = 1; j <= MAXORD + 1; ++j) {
        if (j <= c.nq + 1) {
          const double e = el_of(mc, j, c.nq);
#pragma unroll
          for (int i = 0; i < SA; ++i) c.yh[j - 1][i] = c.yh[j - 1][i] + e * c.save1[i];
        }
      }
      copy_yh0_to_y(c);
      c.do_scale = false;
      c.rmax_after = false;
      c.after_accept = true;
      // step-size and order selection (slatec/ddstp.f90:360-432)
      if (c.jtask == 0 || c.jtask == 2) {
        const double rh = 1.0 / dmax(k.uround, 1.2 * nl_pow(etest / k.eps, 1.0 / (double)(c.nq + 1)));
        c.rh = rh;
        c.do_scale = rh > 1.0;
      } else if (c.nwait > 1) {
        c.nwait = c.nwait - 1;
        if (c.nwait == 1 && c.nq < k.maxord) {
#pragma unroll
          for (int i = 0; i < SA; ++i) c.yh[MAXORD][i] = c.save1[i];
        }
      } else {
        // The three candidates (order down, same, up) are independent: both
        // norms come from one call and the three powers from another.  A
        // candidate that does not apply (order 1 cannot go down, MAXORD not
        // up) is formed from whatever its vector holds and then not used.
        double dn[SA], up[SA];
        yh_col(c, c.nq, dn);
#pragma unroll
        for (int i = 0; i < SA; ++i) {
          const double w = dmax(fabs(c.y[i]), c.ywt[i]);
          dn[i] = dn[i] / w;
          up[i] = (c.save1[i] - c.yh[MAXORD][i]) / w;
        }
        const Norm2 nr = dnrm2_pair(dn, up);
        const double erdn = nr.a / (tq_of(mc, 1, c.nq) * k.sqrtn);
        const double erup = nr.b / (tq_of(mc, 3, c.nq) * k.sqrtn);
        const Pow3 pw = nl_pow3(erdn / k.eps, 1.0 / (double)c.nq, etest / k.eps,
                                1.0 / (double)(c.nq + 1), erup / k.eps, 1.0 / (double)(c.nq + 2));
        const double rh1 = (c.nq != 1) ? 1.0 / dmax(k.uround, 1.3 * pw.a) : 0.0;
        const double rh2 = 1.0 / dmax(k.uround, 1.2 * pw.b);
        const double rh3 = (c.nq != k.maxord) ? 1.0 / dmax(k.uround, 1.4 * pw.c) : 0.0;
        double rh;
        bool change = true;
        if (rh1 > rh2 && rh1 >= rh3) {
          rh = rh1;
          if (rh <= 1.0) {
            change = false;
          } else {
            c.nq = c.nq - 1;
            c.rc = (c.rc * el_of(mc, 1, c.nq)) / el_of(mc, 1, c.nq + 1);
          }
        } else if (rh2 >= rh1 && rh2 >= rh3) {
          rh = rh2;
          if (rh <= 1.0) change = false;
        } else {
          rh = rh3;
          if (rh <= 1.0) {
            change = false;
          } else {
            const double e = el_of(mc, c.nq + 1, c.nq);
            const double den = (double)(c.nq + 1);
            const int dst = c.nq + 1;
            double newc[SA];
#pragma unroll
            for (int i = 0; i < SA; ++i) newc[i] = (c.save1[i] * e) / den;
#pragma unroll
            for (int j = 2; j <= MAXORD; ++j) {
#pragma unroll
              for (int i = 0; i < SA; ++i) c.yh[j][i] = (dst == j) ? newc[i] : c.yh[j][i];
            }
            c.nq = c.nq + 1;
            c.rc = (c.rc * el_of(mc, 1, c.nq)) / el_of(mc, 1, c.nq - 1);
          }
        }
        c.rh = rh;
        c.do_scale = change;
        c.rmax_after = change;
        c.nwait = c.nq + 2;
      }
      c.pc = PC_SCALE;
    }
  }
  // ---- 5: corrector failure, label 160 (slatec/ddstp.f90:206-217)
  if (wave_any(at(c, PC_FAIL_CONV)) && at(c, PC_FAIL_CONV)) {
    c.fail_err = false;
    c.pc = PC_RETRACT;
  }
  // ---- R: retract the prediction, then the failure's step-size choice
  if (wave_any(at(c, PC_RETRACT)) && at(c, PC_RETRACT)) {
    c.t = c.told;
    psc<false>(c.nq, c.yh);
    c.after_accept = false;
    c.rmax_after = false;
    if (!c.fail_err) {                                 // label 160
      c.nwait = c.nq + 2;
      if (c.jtask != 0 && c.jtask != 2) c.rmax = 2.0;
      const double rh = (c.iter == 0) ? 0.3 : 0.9 * nl_pow(k.eps / c.ctest, 0.2);
      if (rh * c.h == 0.0) {
        copy_yh0_to_y(c);
        finish(c, OUT_STEP_ZERO);
      } else {
        c.rh = rh;
        c.do_scale = true;
        c.pc = PC_SCALE;
      }
    } else {                                           // error test failed (:238-289)
      c.nfail++;
      if (c.nfail < 3 || c.nq == 1) {
        if (c.jtask != 0 && c.jtask != 2) c.rmax = 2.0;
        const double rh2 = 1.0 / (1.2 * nl_pow(c.etest / k.eps, 1.0 / (double)(c.nq + 1)));
        double rh = rh2;
        if (c.nq > 1) {
          double col[SA];
          yh_col(c, c.nq, col);
#pragma unroll
          for (int i = 0; i < SA; ++i) c.save2[i] = col[i] / dmax(fabs(c.y[i]), c.ywt[i]);
          const double erdn = dnrm2(c.save2) / (tq_of(mc, 1, c.nq) * k.sqrtn);
          const double rh1 = 1.0 / dmax(1.0, 1.3 * nl_pow(erdn / k.eps, 1.0 / (double)c.nq));
          if (rh2 < rh1) {
            c.nq = c.nq - 1;
            c.rc = (c.rc * el_of(mc, 1, c.nq)) / el_of(mc, 1, c.nq + 1);
            rh = rh1;
          }
        }
        c.nwait = c.nq + 2;
        if (rh * c.h == 0.0) {
          copy_yh0_to_y(c);
          finish(c, OUT_STEP_ZERO);
        } else {
          c.rh = rh;
          c.do_scale = true;
          c.pc = PC_SCALE;
        }
      } else {
        c.nfail = 0;                                   // restart at order 1 (:275-288)
        c.jtask = 2;
        copy_yh0_to_y(c);
        ntl_head(c);
        c.ntl_restart = true;
        request_f(c, PC_NTL_F);
      }
    }
  }
  // ---- S: pending DDSCL, then the next attempt or DDRIV3's label 470
  if (wave_any(at(c, PC_SCALE)) && at(c, PC_SCALE)) {
    if (c.do_scale) scl(c, c.rh, k);
    if (c.rmax_after) c.rmax = 10.0;
    if (!c.after_accept) {
      c.pc = PC_ATTEMPT;
    } else {
      // DDSTP returned JSTATE = 1; DDRIV3 label 470 (:1371, NTASK = 1 test :1443-1449)
      c.jtask = 1;
      if (c.t < tout) {
        c.pc = PC_260;
      } else {
        ntp(c, tout);
        finish(c, OUT_OK);
      }
    }
  }
  // ---- 6: DDRIV3 label 260: weights and guards, DDSTP entry (slatec/ddriv3.f90:1292-1351)
  if (wave_any(at(c, PC_260)) && at(c, PC_260)) {
#pragma unroll
    for (int i = 0; i < SA; ++i) c.ywt[i] = dmax(k.ewt, fabs(c.y[i]));
    // |y_i / YWT_i| <= 1, so SUM = max(1, ||y/YWT|| / sqrt(N)) stays below 2
    // and EPS < SUM * UROUND cannot hold for EPS >= 2 * UROUND; the norm is
    // formed only when it can decide (its SAVE2 is dead: the next f overwrites it).
    bool eps_fail = false;
    if (k.eps < 2.0 * k.uround) {
      double v[S];
#pragma unroll
      for (int sp = 0; sp < S; ++sp) {
        if (kSp.act[sp] >= 0) {
          v[sp] = c.y[kSp.act[sp]] / c.ywt[kSp.act[sp]];
        } else {
          const double yv = c.yi[kSp.islot[sp]];
          v[sp] = yv / dmax(k.ewt, fabs(yv));
        }
      }
      double sum = dnrm2_all(v) / k.sqrtn;
      sum = dmax(1.0, sum);
      eps_fail = k.eps < sum * k.uround;
    }
    if (eps_fail) {
      copy_yh0_to_y(c);
      finish(c, OUT_EPS);
    } else if (c.nstep == k.mxstep) {
      copy_yh0_to_y(c);
      finish(c, OUT_MXSTEP);
    } else {
      c.told = c.t;                                    // DDSTP entry (:91-106)
      c.ntry = 0;
      c.nfail = 0;
      if (c.jtask <= 0) {
        ntl_head(c);
        c.ntl_restart = false;
        request_f(c, PC_NTL_F);
      } else {
        c.pc = PC_ATTEMPT;
      }
    }
  }
  // ---- 7: attempt, labels 100/110 (slatec/ddstp.f90:107-118)
  if (wave_any(at(c, PC_ATTEMPT)) && at(c, PC_ATTEMPT)) {
    c.ntry++;
    if (c.ntry > 50) {
      finish(c, OUT_REPEATED_FAIL);
    } else {
      c.natt++;
      c.t = c.t + c.h;
      psc<true>(c.nq, c.yh);
      c.evaljc = (fabs(c.rc - 1.0) > 0.3) || (c.nstep >= c.jstepl + 10);
      c.iter = 0;
      copy_yh0_to_y(c);
      request_f(c, PC_PRED_F);
    }
  }
}

// DDPST, MITER = 2 (slatec/ddpst.f90:39-114,142-143) for the lanes with
// `mine` set: finite-difference Jacobian with adaptive increments FAC,
// P = I - h*l0*J and its LU.  The column loop is unrolled: with a static
// column index the gathers of y_j, YWT_j, FAC_j and the scatter of the column
// into P are plain register moves instead of select chains over all nine
// columns (measured: 11-18 % less time per step than one shared loop body,
// despite nine RHS copies).
NRG_DI void jac_batch(Cell &c, bool mine, const BdfConst &k, const BdfCoef &mc, const double *kl)
{
  if (mine) {
    c.nje++;
    c.ier = false;
  }
  const double l0 = mine ? el_of(mc, 1, c.nq) : 0.0;
#pragma unroll
  for (int j = 0; j < SA; ++j) {
    double ywj = 0.0, yj = 0.0, fj = 0.0, s2j = 0.0, yh2j = 0.0;
#pragma unroll
    for (int q = 0; q < SA; ++q) {
      const bool hit = (q == j);
      ywj = hit ? opaque(c.ywt[q]) : ywj;
      yj = hit ? opaque(c.y[q]) : yj;
      fj = hit ? opaque(c.fac[q]) : fj;
      s2j = hit ? opaque(c.save2[q]) : s2j;
      yh2j = hit ? opaque(c.yh[2][q]) : yh2j;
    }
    double f1[SA];
    if (mine) {
      const double ys = dmax(fabs(ywj), fabs(yj));
      double dy = fj * ys;
      while (dy == 0.0) {
        if (fj < 0.5) {
          fj = dmin(100.0 * fj, 0.5);
          dy = fj * ys;
        } else {
          dy = ys;
          break;
        }
      }
      dy = (c.nq == 1) ? copysign(dy, s2j) : copysign(dy, yh2j);
      dy = (yj + dy) - yj;
      const double yp = yj + dy;
#pragma unroll
      for (int q = 0; q < SA; ++q) c.y[q] = (q == j) ? yp : c.y[q];
      rhs(kl, c.y, f1);
#pragma unroll
      for (int q = 0; q < SA; ++q) c.y[q] = (q == j) ? yj : c.y[q];
      const double factor = ((-l0) * c.h) / dy;
      double col[SA];
#pragma unroll
      for (int i = 0; i < SA; ++i) col[i] = (f1[i] - c.save2[i]) * factor;
      double diff = fabs(c.save2[0] - f1[0]);
      double s2m = c.save2[0], s1m = f1[0];
#pragma unroll
      for (int i = 1; i < SA; ++i) {
        const double di = fabs(c.save2[i] - f1[i]);
        if (di > diff) {
          diff = di;
          s2m = c.save2[i];
          s1m = f1[i];
        }
      }
      if (dmin(fabs(s2m), fabs(s1m)) > 0.0) {
        const double scale = dmax(fabs(s2m), fabs(s1m));
        if (diff > 0.5 * scale)
          fj = dmax(k.facmin, fj * 0.5);
        else if (k.br * scale <= diff && diff <= k.bl * scale)
          fj = dmin(fj * 2.0, 0.5);
        else if (diff < k.br * scale)
          fj = dmin(k.bp * fj, 0.5);
      }
#pragma unroll
      for (int q = 0; q < SA; ++q) {
        const bool hit = (q == j);
        c.fac[q] = hit ? fj : c.fac[q];
#pragma unroll
        for (int i = 0; i < SA; ++i) c.P[q][i] = hit ? col[i] : c.P[q][i];
      }
    }
  }
  if (mine) {
    c.nfe += S;                                        // DDPST counts N evaluations
#pragma unroll
    for (int i = 0; i < SA; ++i) c.P[i][i] = c.P[i][i] + 1.0;
  }
  // DGEFA for the lanes of the batch (the pivot ballots inside see only them)
  if (mine) {
    if (gefa(c.P, c.ipvt) != 0) c.ier = true;
  }
}

// Net stoichiometric coefficient of species s in reaction r (right - left terms).
constexpr int nu_of(int r, int s)
{
  int n = 0;
  for (int c = 0; c < 3; ++c) {
    if (c < topo::nr[r] && topo::right[r][c] == s) ++n;
    if (c < topo::nl[r] && topo::left[r][c] == s) --n;
  }
  return n;
}

// Fast mode (NRG_JAC_ANALYTIC): the iteration matrix P = I - h*l0*J with the
// exact J = d(omega)/dC of the law of mass action at the cell's fixed rate
// constants and [M] (the third-body concentration is frozen per cell, so it
// has no derivative), as DDPST forms it for a user Jacobian (MITER = 1).  No
// extra right-hand sides are evaluated; the step sequence therefore differs
// from the finite-difference (SLATEC MITER = 2) parity mode.
NRG_DI void jac_analytic(Cell &c, bool mine, const BdfCoef &mc, const double *kl)
{
  if (!mine) return;
  c.nje++;
  c.ier = false;
  const double fac = (-el_of(mc, 1, c.nq)) * c.h;
#pragma unroll
  for (int q = 0; q < SA; ++q)
#pragma unroll
    for (int i = 0; i < SA; ++i) c.P[q][i] = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double m = kMSlot.v[r] >= 0 ? rec_m(kl, kMSlot.v[r]) : 0.0;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int n = side == 0 ? topo::nl[r] : topo::nr[r];
      const double k = side == 0 ? rec_kf(kl, r) : -rec_kr(kl, r);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int sa = side == 0 ? topo::left[r][a] : topo::right[r][a];
        if (a >= n || sa == S) continue;
        // d/dC_sa of k * prod_terms: the product of the other terms
        double d = k;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const int sb = side == 0 ? topo::left[r][b] : topo::right[r][b];
          if (b < n && b != a) d = d * (sb == S ? m : c.y[act_of(sb)]);
        }
        const double fd = fac * d;
#pragma unroll
        for (int i = 0; i < SA; ++i) {
          const int nu = nu_of(r, kSp.full[i]);
          if (nu != 0) c.P[act_of(sa)][i] = c.P[act_of(sa)][i] + (double)nu * fd;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < SA; ++i) c.P[i][i] = c.P[i][i] + 1.0;
  if (gefa(c.P, c.ipvt) != 0) c.ier = true;
}

}  // namespace nrg
