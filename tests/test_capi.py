"""The C-ABI library loads and exports every entry point that include/*.h
declares, and the Python bindings agree with the declared signatures.  No GPU
needed: nothing here launches device work."""
import ctypes as C
import os
import re

import pytest

from nrg_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def _declarations():
    txt = "".join(open(h).read() for h in HEADERS)
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    decls = {}
    for m in re.finditer(r"\b(?:int|void|const char|int64_t)\s*\*?\s*(nrg_\w+)\s*\(([^;]*?)\)\s*;",
                         txt):
        name, params = m.group(1), m.group(2).strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        decls[name] = n
    return decls


def test_header_declares_entry_points():
    d = _declarations()
    for name in ("nrg_chem_create", "nrg_chem_destroy", "nrg_chem_step", "nrg_chem_step_dev",
                 "nrg_chem_step_dev_diag", "nrg_chem_closure_dev", "nrg_state_equation_dev",
                 "nrg_energy_from_temperature_dev", "nrg_chem_default_opts", "nrg_chem_strerror",
                 "nrg_chem_last_error", "nrg_chem_abi_version", "nrg_chem_stream",
                 "nrg_mlp_create", "nrg_mlp_destroy", "nrg_mlp_num_params", "nrg_mlp_set_params",
                 "nrg_mlp_get_params", "nrg_mlp_forward_dev", "nrg_mlp_grad_dev",
                 "nrg_mlp_train_step_dev", "nrg_mlp_stream", "nrg_chem_surrogate_step_dev"):
        assert name in d, name


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    for name in _declarations():
        assert hasattr(lib, name), "libnrg_chem.so does not export %s" % name


def test_bindings_match_declared_arity():
    lib = capi.load_library()
    for name, n in _declarations().items():
        fn = getattr(lib, name)
        assert fn.argtypes is not None, "%s has no argtypes" % name
        assert len(fn.argtypes) == n, "%s: binding has %d params, header %d" % (
            name, len(fn.argtypes), n)


def test_struct_layouts():
    # nrg_chem_mech: 2 int32 + 12 pointers; opts: 4 doubles + 12 int32
    assert C.sizeof(capi.NrgChemMech) == 8 + 12 * 8
    assert C.sizeof(capi.NrgChemOpts) == 4 * 8 + 12 * 4
    assert C.sizeof(capi.NrgGrid) == 12 * 4
    assert C.sizeof(capi.NrgChemStats) == 7 * 8 + 2 * 4 + 8 * 8 + 8
    assert C.sizeof(capi.NrgSurrogateNorm) == 4 * 16 * 4 + 4 + 7 * 4
    assert C.sizeof(capi.NrgSurrogateGate) == 13 * 8 + 4 + 7 * 4 + 4 * 32 * 8
    assert C.sizeof(capi.NrgSurrogateStats) == 10 * 8 + 3 * 8 + C.sizeof(capi.NrgChemStats) + 4 * 8


def test_default_opts_are_reference_literals():
    lib = capi.load_library()
    o = capi.NrgChemOpts()
    lib.nrg_chem_default_opts(C.byref(o))
    ref = capi.default_opts()
    for f in ("eps", "ewt", "hmax", "t_gate", "mxstep", "mxord", "jacobian"):
        assert getattr(o, f) == getattr(ref, f), f
    # HMAX = 1.0E-7 is a default-real literal in chemical_kinetics_solver.f90:287
    assert o.hmax == float.fromhex("0x1.ad7f2a0000000p-24")
    assert o.eps == 1e-9 and o.ewt == 1e-4 and o.mxstep == 10000


def test_errors_are_reported_not_fatal(mech):
    lib = capi.load_library()
    assert lib.nrg_chem_strerror(capi.NRG_CHEM_E_STEP_ZERO)
    h = C.c_void_p()
    ms = capi.MechStruct(mech)
    # invalid options are rejected before any device work
    bad = capi.default_opts(eps=-1.0)
    assert lib.nrg_chem_create(C.byref(ms.struct), C.byref(bad), C.byref(h)) == capi.NRG_CHEM_E_ARG
    assert h.value is None
    for g in (-1, 101):
        bad = capi.default_opts(errtest_gather_pct=g)
        assert lib.nrg_chem_create(C.byref(ms.struct), C.byref(bad), C.byref(h)) == capi.NRG_CHEM_E_ARG
        assert b"errtest_gather_pct" in lib.nrg_chem_last_error()
        assert h.value is None
    for r in (-1, 65):
        bad = capi.default_opts(refill_min_idle=r)
        assert lib.nrg_chem_create(C.byref(ms.struct), C.byref(bad), C.byref(h)) == capi.NRG_CHEM_E_ARG
        assert b"refill_min_idle" in lib.nrg_chem_last_error()
        assert h.value is None
    # a NULL handle is an argument error, not a crash
    st = capi.NrgChemStats()
    rc = lib.nrg_chem_step_dev(None, None, 1e-7, *([None] * 9), C.byref(st))
    assert rc == capi.NRG_CHEM_E_ARG


def test_mechanism_topology_is_checked(mech):
    """A mechanism whose reaction list differs from the compiled kernels is refused."""
    import numpy as np
    lib = capi.load_library()
    m2 = type(mech)(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                       for k, v in mech.__dict__.items()})
    m2.chemical_coeffs[0, 0, 1] = 3       # H -> N2 on the left of reaction 1
    ms = capi.MechStruct(m2)
    h = C.c_void_p()
    rc = lib.nrg_chem_create(C.byref(ms.struct), None, C.byref(h))
    # the topology is checked against every compiled build before any device work
    assert rc == capi.NRG_CHEM_E_MECH and h.value is None
    msg = lib.nrg_chem_last_error()
    assert b"KEROMNES: chemical_coeffs" in msg, msg
    for name in (b"ZHANG: ", b"TEREZA: ", b"KONNOV: ", b"ACETYLENE: "):
        assert name in msg, msg
