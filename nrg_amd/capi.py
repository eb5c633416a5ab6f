"""ctypes view of the C-ABI declared in include/nrg_chem.h.

The structures here are byte-for-byte the C ones; ``load_library`` opens the
in-tree HIP build ``nrg_amd/libnrg_chem.so`` and fails loudly when it is
missing -- there is no CPU fallback behind this interface.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libnrg_chem.so")

# status codes (nrg_chem.h)
NRG_CHEM_OK = 0
NRG_CHEM_E_ARG = 1
NRG_CHEM_E_MECH = 2
NRG_CHEM_E_HIP = 3
NRG_CHEM_E_NO_DEVICE = 4
NRG_CHEM_E_FORMAT = 5
NRG_CHEM_E_STEP_ZERO = 41
NRG_CHEM_E_REPEATED_FAIL = 42

NRG_CELL_OK = 0
NRG_CELL_GATED = 1
NRG_CELL_MXSTEP = 2
NRG_CELL_STEP_ZERO = 3
NRG_CELL_REPEATED_FAIL = 4
NRG_CELL_SINGULAR = 5

NRG_JAC_FD_SLATEC = 0
NRG_JAC_ANALYTIC = 1

NRG_FORM_PLAIN = 0
NRG_FORM_PERSISTENT = 1

NRG_ORDER_NATURAL = 0
NRG_ORDER_STIFFNESS = 1

NRG_NOALLOC_SCRATCH_BYTES = 64 << 20    # nrg_chem_step_dev_noalloc
NRG_NOALLOC_CHUNK_CELLS = 1 << 17
NRG_NOALLOC_MAX_CELLS = 48 * NRG_NOALLOC_CHUNK_CELLS

NRG_CELL_SURROGATE = 6          # include/nrg_surrogate.h
NRG_SUR_MAX_IN = 16
NRG_SUR_MAX_OUT = 16
NRG_SUR_MAX_ELEMENTS = 4
NRG_SUR_SCREEN_NONE = 0         # gate.screen: the network sees every live cell
NRG_SUR_SCREEN_RATES = 1        # explicit-rate screen first, network on the rest
NRG_SUR_SCREEN_ONLY = 2         # explicit-rate screen, FP64 on the rest (no network)


class NrgChemMech(C.Structure):
    _fields_ = [
        ("species_number", C.c_int32),
        ("reactions_number", C.c_int32),
        ("reactions_type", C.POINTER(C.c_int32)),
        ("chemical_coeffs", C.POINTER(C.c_int32)),
        ("A", C.POINTER(C.c_double)),
        ("beta", C.POINTER(C.c_double)),
        ("E_act", C.POINTER(C.c_double)),
        ("A_low", C.POINTER(C.c_double)),
        ("beta_low", C.POINTER(C.c_double)),
        ("E_act_low", C.POINTER(C.c_double)),
        ("Troe_coeffs", C.POINTER(C.c_double)),
        ("enhanced_efficiencies", C.POINTER(C.c_double)),
        ("a_coeffs", C.POINTER(C.c_double)),
        ("molar_masses", C.POINTER(C.c_double)),
    ]


class NrgChemOpts(C.Structure):
    _fields_ = [
        ("eps", C.c_double),
        ("ewt", C.c_double),
        ("hmax", C.c_double),
        ("t_gate", C.c_double),
        ("mxstep", C.c_int32),
        ("mxord", C.c_int32),
        ("jacobian", C.c_int32),
        ("device", C.c_int32),
        ("form", C.c_int32),
        ("jac_batch_pct", C.c_int32),
        ("waves_per_cu", C.c_int32),
        ("cell_order", C.c_int32),
        ("use_graph", C.c_int32),
        ("chunk_log2", C.c_int32),
        ("errtest_gather_pct", C.c_int32),
        ("refill_min_idle", C.c_int32),
    ]


class NrgGrid(C.Structure):
    _fields_ = [
        ("lo", C.c_int32 * 3),
        ("hi", C.c_int32 * 3),
        ("inner_lo", C.c_int32 * 3),
        ("inner_hi", C.c_int32 * 3),
    ]


class NrgChemStats(C.Structure):
    _fields_ = [
        ("cells_inner", C.c_int64),
        ("cells_integrated", C.c_int64),
        ("nstep_sum", C.c_int64),
        ("nfe_sum", C.c_int64),
        ("nje_sum", C.c_int64),
        ("ncorr_sum", C.c_int64),
        ("natt_sum", C.c_int64),
        ("nstep_max", C.c_int32),
        ("nfe_max", C.c_int32),
        ("status_count", C.c_int64 * 8),
        ("kernel_ms", C.c_double),
    ]

    def as_dict(self):
        return {
            "cells_inner": self.cells_inner,
            "cells_integrated": self.cells_integrated,
            "nstep_sum": self.nstep_sum,
            "nfe_sum": self.nfe_sum,
            "nje_sum": self.nje_sum,
            "ncorr_sum": self.ncorr_sum,
            "natt_sum": self.natt_sum,
            "nstep_max": self.nstep_max,
            "nfe_max": self.nfe_max,
            "status_count": list(self.status_count),
            "kernel_ms": self.kernel_ms,
        }


class NrgSurrogateNorm(C.Structure):
    """struct nrg_surrogate_norm (include/nrg_surrogate.h)."""
    _fields_ = [("in_shift", C.c_float * NRG_SUR_MAX_IN), ("in_scale", C.c_float * NRG_SUR_MAX_IN),
                ("out_shift", C.c_float * NRG_SUR_MAX_OUT),
                ("out_scale", C.c_float * NRG_SUR_MAX_OUT),
                ("y_floor", C.c_float), ("reserved", C.c_int32 * 7)]


class NrgSurrogateGate(C.Structure):
    """struct nrg_surrogate_gate (include/nrg_surrogate.h)."""
    _fields_ = [("t_min", C.c_double), ("t_max", C.c_double), ("p_min", C.c_double),
                ("p_max", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
                ("elem_tol", C.c_double), ("elem_floor", C.c_double), ("y_neg_tol", C.c_double),
                ("energy_tol", C.c_double), ("dT_max", C.c_double),
                ("dT_freeze", C.c_double), ("dY_freeze", C.c_double),
                ("n_elements", C.c_int32), ("screen", C.c_int32), ("rate_tol", C.c_double),
                ("screen_fp64", C.c_int32), ("reserved", C.c_int32 * 3),
                ("elements", (C.c_double * 32) * NRG_SUR_MAX_ELEMENTS)]


class NrgSurrogateStats(C.Structure):
    """struct nrg_surrogate_stats (include/nrg_surrogate.h)."""
    _fields_ = [("cells_inner", C.c_int64), ("cells_gated", C.c_int64),
                ("cells_accepted", C.c_int64), ("cells_fallback", C.c_int64),
                ("reject_domain", C.c_int64), ("reject_elements", C.c_int64),
                ("reject_negative", C.c_int64), ("reject_energy", C.c_int64),
                ("reject_confidence", C.c_int64), ("cells_frozen", C.c_int64),
                ("infer_ms", C.c_double), ("gate_ms", C.c_double), ("fallback_ms", C.c_double),
                ("fallback", NrgChemStats), ("reject_rate", C.c_int64),
                ("cells_network", C.c_int64), ("screen_ms", C.c_double),
                ("cells_screen_fp64", C.c_int64)]


class NrgChemTuneTrial(C.Structure):
    _fields_ = [("form", C.c_int32), ("cell_order", C.c_int32), ("jac_batch_pct", C.c_int32),
                ("waves_per_cu", C.c_int32), ("chunk_cells", C.c_int64), ("kernel_ms", C.c_double),
                ("errtest_gather_pct", C.c_int32), ("refill_min_idle", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


NRG_TUNE_MAX_TRIALS = 24


class NrgChemTuning(C.Structure):
    _fields_ = [("best", NrgChemTuneTrial), ("n_trials", C.c_int32), ("reserved", C.c_int32),
                ("trials", NrgChemTuneTrial * NRG_TUNE_MAX_TRIALS)]


def default_opts(**kw):
    """Reference literals (chemical_kinetics_solver.f90:270-290), see nrg_chem_default_opts."""
    o = NrgChemOpts()
    o.eps = 1.0e-9
    o.ewt = 1.0e-4
    o.hmax = float(np.float32(1.0e-7))
    o.t_gate = 305.0
    o.mxstep = 10000
    o.mxord = 5
    o.jacobian = NRG_JAC_FD_SLATEC
    o.device = 0
    o.form = NRG_FORM_PLAIN
    o.jac_batch_pct = 0
    o.waves_per_cu = 0
    o.cell_order = NRG_ORDER_NATURAL
    o.use_graph = 0
    o.chunk_log2 = 0
    o.errtest_gather_pct = 0
    o.refill_min_idle = 0
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def grid_for_shape(shape, ghost=1):
    """Grid of NRG fields with `ghost` layers in active dimensions (extent > 1).

    Mirrors computational_domain bounds (0:n+1 in active dims, 1:1 otherwise).
    """
    g = NrgGrid()
    for d in range(3):
        n = shape[d] if d < len(shape) else 1
        gh = ghost if n > 1 else 0
        g.lo[d] = 1 - gh
        g.hi[d] = n + gh
        g.inner_lo[d] = 1
        g.inner_hi[d] = n
    return g


def flat_grid(ncells):
    """A 1-D list of cells with no ghosts: offset == cell index."""
    g = NrgGrid()
    g.lo[0], g.hi[0], g.inner_lo[0], g.inner_hi[0] = 1, ncells, 1, ncells
    for d in (1, 2):
        g.lo[d] = g.hi[d] = g.inner_lo[d] = g.inner_hi[d] = 1
    return g


class MechStruct:
    """Owns contiguous copies of a Mechanism's arrays and the C struct pointing at them."""

    def __init__(self, mech):
        self.mech = mech
        self._keep = {}

        def dptr(name, arr, dtype=np.float64):
            a = np.ascontiguousarray(arr, dtype=dtype)
            self._keep[name] = a
            ctype = C.c_double if dtype == np.float64 else C.c_int32
            return a.ctypes.data_as(C.POINTER(ctype))

        s = NrgChemMech()
        s.species_number = mech.species_number
        s.reactions_number = mech.reactions_number
        s.reactions_type = dptr("rt", mech.reactions_type, np.int32)
        s.chemical_coeffs = dptr("cc", mech.chemical_coeffs, np.int32)
        s.A = dptr("A", mech.A)
        s.beta = dptr("beta", mech.beta)
        s.E_act = dptr("E", mech.E_act)
        s.A_low = dptr("Al", mech.A_low)
        s.beta_low = dptr("bl", mech.beta_low)
        s.E_act_low = dptr("El", mech.E_act_low)
        s.Troe_coeffs = dptr("troe", mech.Troe_coeffs)
        s.enhanced_efficiencies = dptr("eff", mech.enhanced_efficiencies)
        s.a_coeffs = dptr("a", mech.a_coeffs)
        s.molar_masses = dptr("W", mech.molar_masses)
        self.struct = s

    def ref(self):
        return C.byref(self.struct)


def ptr_array(ptrs):
    """C array of `double*` (device or host addresses given as ints)."""
    arr = (C.c_void_p * len(ptrs))(*[C.c_void_p(int(p)) for p in ptrs])
    return arr


class NrgChemError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


_lib = None


def load_library(path=None):
    """Open libnrg_chem.so (the HIP build).  Raises if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            "libnrg_chem.so is not built (%s); run __graft_entry__.build() "
            "or `make -C nrg_amd/csrc`" % p)
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
    # Load torch first so this library binds to the runtime torch uses: with
    # the library loaded first, torch finds ROCm's runtime already in the
    # process and reports no GPUs.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(p)
    vp = C.c_void_p
    lib.nrg_chem_default_opts.argtypes = [C.POINTER(NrgChemOpts)]
    lib.nrg_chem_default_opts.restype = None
    lib.nrg_chem_strerror.argtypes = [C.c_int]
    lib.nrg_chem_strerror.restype = C.c_char_p
    lib.nrg_chem_last_error.argtypes = []
    lib.nrg_chem_last_error.restype = C.c_char_p
    lib.nrg_chem_abi_version.argtypes = []
    lib.nrg_chem_abi_version.restype = C.c_int
    lib.nrg_chem_create.argtypes = [C.POINTER(NrgChemMech), C.POINTER(NrgChemOpts), C.POINTER(vp)]
    lib.nrg_chem_create.restype = C.c_int
    lib.nrg_chem_destroy.argtypes = [vp]
    lib.nrg_chem_destroy.restype = None
    step_args = [vp, C.POINTER(NrgGrid), C.c_double, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                 C.POINTER(NrgChemStats)]
    lib.nrg_chem_step_dev.argtypes = step_args
    lib.nrg_chem_step_dev.restype = C.c_int
    lib.nrg_chem_step.argtypes = step_args
    lib.nrg_chem_step.restype = C.c_int
    lib.nrg_chem_step_dev_diag.argtypes = step_args + [vp, vp, vp]
    lib.nrg_chem_step_dev_diag.restype = C.c_int
    lib.nrg_chem_step_dev_noalloc.argtypes = step_args
    lib.nrg_chem_step_dev_noalloc.restype = C.c_int
    lib.nrg_chem_closure_dev.argtypes = [vp, C.POINTER(NrgGrid), C.c_double, vp, vp, C.c_int32,
                                         vp, vp, vp, vp, vp, vp, vp]
    lib.nrg_chem_closure_dev.restype = C.c_int
    lib.nrg_state_equation_dev.argtypes = [vp, C.POINTER(NrgGrid), vp, vp, C.c_int32, vp, vp, vp,
                                           vp, vp]
    lib.nrg_state_equation_dev.restype = C.c_int
    lib.nrg_energy_from_temperature_dev.argtypes = [vp, C.POINTER(NrgGrid), vp, vp, vp, C.c_int32,
                                                    vp]
    lib.nrg_energy_from_temperature_dev.restype = C.c_int
    lib.nrg_chem_stream.argtypes = [vp]
    lib.nrg_chem_stream.restype = vp
    lib.nrg_chem_topology.argtypes = [vp]
    lib.nrg_chem_topology.restype = C.c_char_p
    lib.nrg_chem_autotune_dev.argtypes = step_args[:-2] + [C.POINTER(NrgChemTuning)]
    lib.nrg_chem_autotune_dev.restype = C.c_int
    lib.nrg_chem_set_table.argtypes = [vp, C.c_int32, vp, vp]
    lib.nrg_chem_set_table.restype = C.c_int
    lib.nrg_chem_table_step_dev.argtypes = step_args
    lib.nrg_chem_table_step.argtypes = step_args
    lib.nrg_chem_table_step_dev.restype = C.c_int
    lib.nrg_chem_table_step.restype = C.c_int
    # include/nrg_mech.h
    lib.nrg_mech_load.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_char_p,
                                  C.POINTER(vp)]
    lib.nrg_mech_load.restype = C.c_int
    lib.nrg_mech_load_task_setup.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.nrg_mech_load_task_setup.restype = C.c_int
    lib.nrg_mech_view.argtypes = [vp, C.POINTER(NrgChemMech)]
    lib.nrg_mech_view.restype = C.c_int
    lib.nrg_mech_species_name.argtypes = [vp, C.c_int32]
    lib.nrg_mech_species_name.restype = C.c_char_p
    lib.nrg_mech_reaction_string.argtypes = [vp, C.c_int32]
    lib.nrg_mech_reaction_string.restype = C.c_char_p
    lib.nrg_mech_free.argtypes = [vp]
    lib.nrg_mech_free.restype = None
    # include/nrg_low_mach.h
    lib.nrg_low_mach_state_equation_dev.argtypes = [vp, C.POINTER(NrgGrid), C.c_double, C.c_int32,
                                                    vp, vp, vp, C.c_int32, vp, vp, vp, vp, vp, vp,
                                                    vp, vp, vp, vp, vp, vp, vp]
    lib.nrg_low_mach_state_equation_dev.restype = C.c_int
    lib.nrg_low_mach_species_dev.argtypes = [vp, C.POINTER(NrgGrid), C.c_double, C.c_int32, vp, vp,
                                             vp, vp, vp, vp, vp, vp, vp, vp]
    lib.nrg_low_mach_species_dev.restype = C.c_int
    lib.nrg_low_mach_chem_divergence_dev.argtypes = [vp, C.POINTER(NrgGrid), vp, vp, vp, vp, vp,
                                                     vp, vp, vp]
    lib.nrg_low_mach_chem_divergence_dev.restype = C.c_int
    # include/nrg_mlp.h
    lib.nrg_mlp_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   C.POINTER(vp)]
    lib.nrg_mlp_create.restype = C.c_int
    lib.nrg_mlp_destroy.argtypes = [vp]
    lib.nrg_mlp_destroy.restype = None
    lib.nrg_mlp_num_params.argtypes = [vp]
    lib.nrg_mlp_num_params.restype = C.c_int64
    lib.nrg_mlp_set_params.argtypes = [vp, vp]
    lib.nrg_mlp_set_params.restype = C.c_int
    lib.nrg_mlp_get_params.argtypes = [vp, vp]
    lib.nrg_mlp_get_params.restype = C.c_int
    lib.nrg_mlp_forward_dev.argtypes = [vp, C.c_int64, vp, vp]
    lib.nrg_mlp_forward_dev.restype = C.c_int
    lib.nrg_mlp_grad_dev.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp, vp]
    lib.nrg_mlp_grad_dev.restype = C.c_int
    lib.nrg_mlp_train_step_dev.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_float, vp]
    lib.nrg_mlp_train_step_dev.restype = C.c_int
    lib.nrg_mlp_stream.argtypes = [vp]
    lib.nrg_mlp_stream.restype = vp
    # include/nrg_surrogate.h
    lib.nrg_chem_surrogate_step_dev.argtypes = [
        vp, vp, C.POINTER(NrgSurrogateNorm), C.POINTER(NrgSurrogateGate), C.POINTER(NrgGrid),
        C.c_double, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(NrgSurrogateStats)]
    lib.nrg_chem_surrogate_step_dev.restype = C.c_int
    if path is None:
        _lib = lib
    return lib


def check(lib, status):
    if status != NRG_CHEM_OK:
        detail = lib.nrg_chem_last_error().decode(errors="replace")
        raise NrgChemError(status, "%s: %s" % (lib.nrg_chem_strerror(status).decode(), detail))
    return status
