// nrg_api.cpp -- the C-ABI of include/nrg_chem.h, nrg_low_mach.h and
// nrg_surrogate.h over the topologies this library is built for.
//
// The kernels unroll a mechanism's reaction lists at compile time, so the
// library holds one build of them per reaction topology (topo_api.h).  A handle
// remembers the build whose topology its mechanism has, and every entry point
// forwards to that build.  Status text and the last error message live here,
// shared by every build and by the loader and network entry points.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "nrg_common.h"
#include "topo_api.h"

namespace nrg {

thread_local char g_last_error[512] = "";

int set_error(int status, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return status;
}

}  // namespace nrg

using nrg::set_error;

// The C-ABI's handle: the build it belongs to and that build's own handle.
struct nrg_chem_handle {
  const nrg_topo_ops *ops;
  ::nrg_chem_t *impl;
};

static const nrg_topo_ops *topology(int i)
{
  switch (i) {
    case 0: return nrg_topo_ops_keromnes();
    case 1: return nrg_topo_ops_zhang();
    case 2: return nrg_topo_ops_tereza();
    case 3: return nrg_topo_ops_konnov();
    case 4: return nrg_topo_ops_acetylene();
    default: return nullptr;
  }
}

extern "C" {

int nrg_chem_abi_version(void) { return NRG_CHEM_ABI_VERSION; }

void nrg_chem_default_opts(nrg_chem_opts *o)
{
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->eps = 1.0e-9;                    // chemical_kinetics_solver.f90:274
  o->ewt = 1.0e-4;                    // :275
  o->hmax = (double)1.0e-7f;          // :287, default-real literal
  o->t_gate = 305.0;                  // :228
  o->mxstep = 10000;                  // :288
  o->mxord = 5;                       // :282
  o->jacobian = NRG_JAC_FD_SLATEC;
  o->device = 0;
  o->form = NRG_FORM_PLAIN;
  o->jac_batch_pct = 0;
  o->waves_per_cu = 0;
}

const char *nrg_chem_strerror(int st)
{
  switch (st) {
    case NRG_CHEM_OK: return "ok";
    case NRG_CHEM_E_ARG: return "invalid argument";
    case NRG_CHEM_E_MECH: return "mechanism not supported by the compiled kernels";
    case NRG_CHEM_E_HIP: return "HIP runtime error";
    case NRG_CHEM_E_NO_DEVICE: return "no gfx950 device";
    case NRG_CHEM_E_FORMAT: return "unreadable mechanism, thermo or molar-mass file";
    case NRG_CHEM_E_STEP_ZERO: return "integrator step size went to zero in a cell (SLATEC IERFLG 41)";
    case NRG_CHEM_E_REPEATED_FAIL: return "integrator failed 50 attempts in a cell (SLATEC IERFLG 42)";
    default: return "unknown status";
  }
}

const char *nrg_chem_last_error(void) { return nrg::g_last_error; }

int nrg_chem_create(const nrg_chem_mech *mech, const nrg_chem_opts *opts, nrg_chem_t **out)
{
  if (!out) return set_error(NRG_CHEM_E_ARG, "out is NULL");
  *out = nullptr;
  char why[480] = "";       // each build's reason, "NAME: detail; ..."
  for (int i = 0; const nrg_topo_ops *ops = topology(i); ++i) {
    int st = ops->check(mech);
    if (st == NRG_CHEM_E_MECH) {
      const size_t used = std::strlen(why);
      std::snprintf(why + used, sizeof(why) - used, "%s%s: %s", i ? "; " : "", ops->name,
                    nrg::g_last_error);
      continue;
    }
    if (st) return st;
    ::nrg_chem_t *impl = nullptr;
    if ((st = ops->create(mech, opts, &impl))) return st;
    auto *h = new (std::nothrow) nrg_chem_handle{ops, impl};
    if (!h) {
      ops->destroy(impl);
      return set_error(NRG_CHEM_E_ARG, "out of host memory");
    }
    *out = h;
    return NRG_CHEM_OK;
  }
  return set_error(NRG_CHEM_E_MECH,
                   "the mechanism (%d species, %d reactions) has the reaction lists of none of the "
                   "compiled kernels (%s)", mech->species_number, mech->reactions_number, why);
}

void nrg_chem_destroy(nrg_chem_t *h)
{
  if (!h) return;
  h->ops->destroy(h->impl);
  delete h;
}

void *nrg_chem_stream(nrg_chem_t *h) { return h ? h->ops->stream(h->impl) : nullptr; }

const char *nrg_chem_topology(nrg_chem_t *h) { return h ? h->ops->name : nullptr; }

#define H(h) (h)->impl
#define NRG_FORWARD(name, params, args)                                \
  int name params                                                      \
  {                                                                    \
    if (!h) return set_error(NRG_CHEM_E_ARG, "handle is NULL");        \
    return h->ops->name args;                                          \
  }
NRG_TOPO_ENTRY_POINTS(NRG_FORWARD)
#undef NRG_FORWARD
#undef H

}  // extern "C"
