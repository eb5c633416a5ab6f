"""ctypes bindings of the oracle (liboracle.so) and of SLATEC built from the
reference sources (_ref/libslatec_ref.so).  TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from nrg_amd.capi import MechStruct, NrgChemOpts, NrgChemStats, NrgGrid, default_opts

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libslatec_ref.so")

RHS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double),
                     C.POINTER(C.c_double))

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build_oracle(force=False):
    """Compile the oracle (and the SLATEC reference build when the sources exist)."""
    if force or not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", ORACLE_DIR, "-j8", "oracle"], check=True,
                       stdout=subprocess.DEVNULL)
    return ORACLE_SO


def _as_dp(a):
    return a.ctypes.data_as(_dp)


class BdfOpts(C.Structure):
    _fields_ = [("eps", C.c_double), ("ewt", C.c_double), ("hmax", C.c_double),
                ("mxstep", C.c_int), ("mxord", C.c_int)]


class BdfOut(C.Structure):
    _fields_ = [("nstate", C.c_int), ("ierflg", C.c_int), ("nstep", C.c_int), ("nfe", C.c_int),
                ("nje", C.c_int), ("ncorr", C.c_int), ("natt", C.c_int), ("nqused", C.c_int),
                ("hused", C.c_double), ("t", C.c_double)]


def _ptrs(arrs):
    for a in arrs:
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("field rows handed to the oracle must be C-contiguous")
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class Oracle:
    """Thin wrapper over liboracle.so for one mechanism."""

    def __init__(self, mech, opts=None):
        build_oracle()
        self.lib = C.CDLL(ORACLE_SO)
        L = self.lib
        L.or_dnrm2.argtypes = [C.c_int, _dp]
        L.or_dnrm2.restype = C.c_double
        L.or_dgefa.argtypes = [_dp, C.c_int, C.c_int, _ip]
        L.or_dgefa.restype = C.c_int
        L.or_dgesl.argtypes = [_dp, C.c_int, C.c_int, _ip, _dp]
        L.or_ddcst.argtypes = [C.c_int, _dp, _dp]
        L.or_ddriv3.argtypes = [C.c_int, _dp, C.c_double, C.POINTER(BdfOpts), RHS_FN, C.c_void_p,
                                C.POINTER(BdfOut)]
        for fn in ("or_cp_molar", "or_h_molar", "or_s_molar", "or_e_molar"):
            getattr(L, fn).argtypes = [C.c_void_p, C.c_double, C.c_int]
            getattr(L, fn).restype = C.c_double
        L.or_calculate_temperature.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
        L.or_calculate_temperature.restype = C.c_double
        L.or_mixture_concentration.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.or_rate_constants.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
        L.or_rhs.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.or_chem_step.argtypes = [C.c_void_p, C.POINTER(NrgChemOpts), C.POINTER(NrgGrid),
                                   C.c_double, _dp, _dp, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, _dp, C.c_void_p, C.POINTER(NrgChemStats), C.c_int]
        L.or_chem_step.restype = C.c_int
        L.or_chem_cells.argtypes = [C.c_void_p, C.POINTER(NrgChemOpts), C.c_int64, C.c_double,
                                    _dp, _dp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _dp,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.or_chem_cells.restype = C.c_int
        L.or_closure_cells.argtypes = [C.c_void_p, C.c_int64, C.c_double, _dp, _dp, C.c_void_p,
                                       C.c_void_p, _dp, C.c_void_p, _dp, C.c_void_p, C.c_int]
        L.or_energy_from_temperature_cells.argtypes = [C.c_void_p, C.c_int64, _dp, C.c_void_p, _dp]
        L.or_table_cells.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.c_double, C.c_int32, _dp,
                                     _dp, _dp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _dp,
                                     C.c_void_p]
        vp = C.c_void_p
        L.or_low_mach_eos_cells.argtypes = [vp, C.c_int64, C.c_double, C.c_int, _dp, vp, vp, C.c_int,
                                            _dp, vp, _dp, _dp, _dp, vp, _dp, vp, vp, vp, vp, vp, vp]
        L.or_low_mach_species_cells.argtypes = [vp, C.c_int64, C.c_double, C.c_int, _dp, vp, vp, vp,
                                                vp, vp, vp, vp, _dp, vp]
        L.or_low_mach_chem_divergence_cells.argtypes = [vp, C.c_int64, _dp, _dp, vp, _dp, _dp, vp,
                                                        vp, _dp]
        L.or_chem_ctx_size.restype = C.c_size_t
        L.or_chem_ctx_init.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, _dp, _dp]
        self.chem_fun = RHS_FN(C.cast(L.or_chem_fun, C.c_void_p).value)
        self.mech = mech
        self.ms = MechStruct(mech)
        self.mp = C.cast(C.pointer(self.ms.struct), C.c_void_p)
        self.opts = opts if opts is not None else default_opts()
        self.S = mech.species_number
        self.R = mech.reactions_number

    # -- linear algebra / integrator pieces -----------------------------------
    def dnrm2(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return self.lib.or_dnrm2(len(x), _as_dp(x))

    def dgefa(self, a):
        """LU of a (n x n, numpy); returns (lu column-major as numpy (n,n) view, ipvt, info)."""
        n = a.shape[0]
        f = np.asfortranarray(a, dtype=np.float64).copy(order="F")
        ip = np.zeros(n, dtype=np.int32)
        info = self.lib.or_dgefa(f.ctypes.data_as(_dp), n, n, ip.ctypes.data_as(_ip))
        return f, ip, info

    def dgesl(self, lu, ipvt, b):
        n = lu.shape[0]
        bb = np.ascontiguousarray(b, dtype=np.float64).copy()
        self.lib.or_dgesl(lu.ctypes.data_as(_dp), n, n, ipvt.ctypes.data_as(_ip), _as_dp(bb))
        return bb

    def ddcst(self, maxord=5):
        el = np.zeros((12, 13))
        tq = np.zeros((12, 3))
        self.lib.or_ddcst(maxord, _as_dp(el), _as_dp(tq))
        return el, tq          # el[j-1, i-1] = EL(i,j); tq[j-1, k-1] = TQ(k,j)

    def ddriv3(self, rhs, y0, tout, eps=1e-9, ewt=1e-4, hmax=None, mxstep=10000, mxord=5):
        """Integrate with the C restatement; rhs(t, y) -> ydot (numpy)."""
        if hmax is None:
            hmax = float(np.float32(1.0e-7))
        y = np.ascontiguousarray(y0, dtype=np.float64).copy()
        n = len(y)

        def cb(ctx, nn, t, yp, ydp):
            yy = np.ctypeslib.as_array(yp, shape=(nn,))
            out = np.ctypeslib.as_array(ydp, shape=(nn,))
            out[:] = rhs(t, yy.copy())

        f = RHS_FN(cb)
        o = BdfOpts(eps, ewt, hmax, mxstep, mxord)
        out = BdfOut()
        self.lib.or_ddriv3(n, _as_dp(y), tout, C.byref(o), f, None, C.byref(out))
        return y, out

    def cell_context(self, T, rho, Y):
        """(ctx, C_in) of one cell: rate constants and [M] frozen at T, as the reference."""
        ctx = C.create_string_buffer(self.lib.or_chem_ctx_size())
        cin = np.zeros(self.S)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        self.lib.or_chem_ctx_init(ctx, self.mp, T, rho, _as_dp(Y), _as_dp(cin))
        return ctx, cin

    def ddriv3_cell(self, T, rho, Y, dt, eps=1e-9, ewt=1e-4, hmax=None, mxstep=10000):
        """One cell's DDRIV3 call through the C restatement; returns (C_out, C_in, BdfOut)."""
        if hmax is None:
            hmax = float(np.float32(1.0e-7))
        ctx, cin = self.cell_context(T, rho, Y)
        y = cin.copy()
        o = BdfOpts(eps, ewt, hmax, mxstep, 5)
        out = BdfOut()
        self.lib.or_ddriv3(self.S, _as_dp(y), dt, C.byref(o), self.chem_fun,
                           C.cast(ctx, C.c_void_p), C.byref(out))
        return y, cin, out

    # -- thermo / kinetics -----------------------------------------------------
    def h(self, T, s):
        return self.lib.or_h_molar(self.mp, T, s)

    def s(self, T, s):
        return self.lib.or_s_molar(self.mp, T, s)

    def cp(self, T, s):
        return self.lib.or_cp_molar(self.mp, T, s)

    def e(self, T, s):
        return self.lib.or_e_molar(self.mp, T, s)

    def calculate_temperature(self, T, e_molar, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        return self.lib.or_calculate_temperature(self.mp, T, e_molar, _as_dp(X))

    def mixture_concentration(self, Y, rho):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        M = np.zeros(self.R)
        self.lib.or_mixture_concentration(self.mp, _as_dp(Y), rho, _as_dp(M))
        return M

    def rate_constants(self, T, Mc):
        Mc = np.ascontiguousarray(Mc, dtype=np.float64)
        kf = np.zeros(self.R)
        kr = np.zeros(self.R)
        self.lib.or_rate_constants(self.mp, T, _as_dp(Mc), _as_dp(kf), _as_dp(kr))
        return kf, kr

    def rhs(self, kf, kr, Mc, Z):
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        out = np.zeros(self.S)
        self.lib.or_rhs(self.mp, _as_dp(np.ascontiguousarray(kf)), _as_dp(np.ascontiguousarray(kr)),
                        _as_dp(np.ascontiguousarray(Mc)), _as_dp(Z), _as_dp(out))
        return out

    def chem_cells(self, dt, T, rho, Y, bc=None, Y_prod2=None, nthreads=1):
        """Chemistry step over a flat list of cells.  Y: (S, n).  Returns dict of outputs."""
        n = len(T)
        T = np.ascontiguousarray(T, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        yp = np.zeros((self.S, n))
        yp2 = (np.zeros((self.S, n)) if Y_prod2 is None
               else np.array(Y_prod2, dtype=np.float64, order="C"))
        ef = np.zeros(n)
        st = np.zeros(n, dtype=np.int32)
        ns = np.zeros(n, dtype=np.int32)
        nf = np.zeros(n, dtype=np.int32)
        nj = np.zeros(n, dtype=np.int32)
        bcp = None
        if bc is not None:
            bc = np.ascontiguousarray(bc, dtype=np.int8)
            bcp = bc.ctypes.data
        rc = self.lib.or_chem_cells(self.mp, C.byref(self.opts), n, dt, _as_dp(T), _as_dp(rho),
                                    _ptrs(list(Y)), bcp, _ptrs(list(yp)), _ptrs(list(yp2)),
                                    _as_dp(ef), st.ctypes.data, ns.ctypes.data, nf.ctypes.data,
                                    nj.ctypes.data, nthreads)
        return dict(rc=rc, Y_prod=yp, Y_prod2=yp2, E_f_prod=ef, status=st, nstep=ns, nfe=nf,
                    nje=nj)

    def table_cells(self, dt, T_t, dC_t, T, Y, bc=None, Y_prod2=None, t_gate=305.0):
        """Table-approximated step over a flat list of cells.  dC_t: (S, rows)."""
        n = len(T)
        T = np.ascontiguousarray(T, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        T_t = np.ascontiguousarray(T_t, dtype=np.float64)
        dC_t = np.ascontiguousarray(dC_t, dtype=np.float64)
        yp = np.zeros((self.S, n))
        yp2 = (np.zeros((self.S, n)) if Y_prod2 is None
               else np.array(Y_prod2, dtype=np.float64, order="C"))
        ef = np.zeros(n)
        st = np.zeros(n, dtype=np.int32)
        bcp = None
        if bc is not None:
            bc = np.ascontiguousarray(bc, dtype=np.int8)
            bcp = bc.ctypes.data
        self.lib.or_table_cells(self.mp, n, dt, t_gate, len(T_t), _as_dp(T_t), _as_dp(dC_t),
                                _as_dp(T), _ptrs(list(Y)), bcp, _ptrs(list(yp)), _ptrs(list(yp2)),
                                _as_dp(ef), st.ctypes.data)
        return dict(Y_prod=yp, Y_prod2=yp2, E_f_prod=ef, status=st)

    # -- low-Mach (FDS) coupling -----------------------------------------------
    def low_mach_eos(self, dt, predictor, rho, Y, p_stat, p_stat_old, dp_stat_dt, v=None,
                     E_f=None, bc=None, fill=0.0):
        """apply_state_equation_low_mach_fds over a flat list of cells.  Returns a
        dict of the fields after the call (in/out fields start from the inputs,
        pure outputs from `fill`)."""
        n = len(rho)
        f64 = lambda a: np.array(a, dtype=np.float64, order="C")  # noqa: E731
        out = dict(p_stat=f64(p_stat), p_stat_old=f64(p_stat_old),
                   E_f=f64(E_f) if E_f is not None else np.full(n, fill))
        for k in ("T", "p", "mix_mol_mass", "mixture_cp", "gamma", "h_s", "v_s", "e_i"):
            out[k] = np.full(n, fill)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        dp = np.ascontiguousarray(dp_stat_dt, dtype=np.float64)
        ndim = 0
        vptr = None
        if v is not None:
            v = np.ascontiguousarray(v, dtype=np.float64)
            ndim = v.shape[0]
            vptr = _ptrs(list(v))
        bcp = None
        if bc is not None:
            bc = np.ascontiguousarray(bc, dtype=np.int8)
            bcp = bc.ctypes.data
        o = out
        self.lib.or_low_mach_eos_cells(self.mp, n, dt, int(bool(predictor)), _as_dp(rho),
                                       _ptrs(list(Y)), vptr, ndim, _as_dp(dp), bcp,
                                       _as_dp(o["p_stat"]), _as_dp(o["p_stat_old"]), _as_dp(o["T"]),
                                       o["p"].ctypes.data, _as_dp(o["mix_mol_mass"]),
                                       o["mixture_cp"].ctypes.data, o["gamma"].ctypes.data,
                                       o["h_s"].ctypes.data, o["v_s"].ctypes.data,
                                       o["e_i"].ctypes.data, o["E_f"].ctypes.data)
        return out

    def low_mach_species(self, dt, corrector, rho_a, Y_a, rho_b=None, Y_b=None, rhs=None,
                         Y_prod_diff=None, Y_prod_chem=None, bc=None, rho_out=None, Y_out=None):
        """Species update of calculate_interm_Y_predictor / _corrector.  Returns
        (rho_out, Y_out); cells left untouched keep rho_out / Y_out's inputs
        (zeros by default)."""
        n = len(rho_a)
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64) if a is not None else None  # noqa: E731
        Y_a, Y_b, rhs, Y_prod_diff, Y_prod_chem = (c(a) for a in (Y_a, Y_b, rhs, Y_prod_diff,
                                                                   Y_prod_chem))
        pl = lambda a: _ptrs(list(a)) if a is not None else None  # noqa: E731
        ro = np.array(rho_out, dtype=np.float64, order="C") if rho_out is not None else np.zeros(n)
        yo = (np.array(Y_out, dtype=np.float64, order="C") if Y_out is not None
              else np.zeros((self.S, n)))
        bcp = None
        if bc is not None:
            bc = np.ascontiguousarray(bc, dtype=np.int8)
            bcp = bc.ctypes.data
        rb = c(rho_b) if rho_b is not None else None
        self.lib.or_low_mach_species_cells(self.mp, n, dt, int(bool(corrector)), _as_dp(c(rho_a)),
                                           pl(Y_a), rb.ctypes.data if rb is not None else None,
                                           pl(Y_b), pl(rhs), pl(Y_prod_diff), pl(Y_prod_chem), bcp,
                                           _as_dp(ro), _ptrs(list(yo)))
        return ro, yo

    def low_mach_chem_divergence(self, rho, T, Y, mix_mol_mass, E_f_prod_chem, Y_prod_chem,
                                 bc=None, div_v=None):
        n = len(rho)
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        d = np.array(div_v, dtype=np.float64, order="C") if div_v is not None else np.zeros(n)
        bcp = None
        if bc is not None:
            bc = np.ascontiguousarray(bc, dtype=np.int8)
            bcp = bc.ctypes.data
        Y = c(Y)
        Yc = c(Y_prod_chem)
        self.lib.or_low_mach_chem_divergence_cells(self.mp, n, _as_dp(c(rho)), _as_dp(c(T)),
                                                   _ptrs(list(Y)), _as_dp(c(mix_mol_mass)),
                                                   _as_dp(c(E_f_prod_chem)), _ptrs(list(Yc)), bcp,
                                                   _as_dp(d))
        return d

    def closure_cells(self, dt, rho, E_f_prod, Y_prod, E_f, Y, T, bc=None, nthreads=1):
        """Split closure + e->T.  Returns (E_f, Y, T, p) new arrays."""
        n = len(T)
        E_f = np.array(E_f, dtype=np.float64, order="C")
        Y = np.array(Y, dtype=np.float64, order="C")
        T = np.array(T, dtype=np.float64, order="C")
        p = np.zeros(n)
        Y_prod = np.ascontiguousarray(Y_prod, dtype=np.float64)
        bcp = None
        if bc is not None:
            bc = np.ascontiguousarray(bc, dtype=np.int8)
            bcp = bc.ctypes.data
        self.lib.or_closure_cells(self.mp, n, dt, _as_dp(np.ascontiguousarray(rho)),
                                  _as_dp(np.ascontiguousarray(E_f_prod)), _ptrs(list(Y_prod)),
                                  bcp, _as_dp(E_f), _ptrs(list(Y)), _as_dp(T), _as_dp(p), nthreads)
        return E_f, Y, T, p

    def energy_from_temperature(self, T, Y):
        n = len(T)
        T = np.ascontiguousarray(T, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        E = np.zeros(n)
        self.lib.or_energy_from_temperature_cells(self.mp, n, _as_dp(T), _ptrs(list(Y)), _as_dp(E))
        return E


class RefSlatec:
    """The reference's SLATEC, compiled from its own sources (oracle/_ref)."""

    def __init__(self):
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO)
        self.lib = C.CDLL(REF_SO)
        L = self.lib
        L.ref_ddriv3.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_int, RHS_FN, C.c_void_p, _ip, _ip, _dp]
        L.ref_ddriv3.restype = C.c_int
        for fn in ("dnrm2_", "dgefa_", "dgesl_", "ddcst_"):
            getattr(L, fn).restype = None
        L.dnrm2_.restype = C.c_double
        vp = C.c_void_p
        L.ref_chem_cells.argtypes = [vp, vp, C.c_int64, C.c_double, vp, vp, vp, vp, vp, vp, vp, vp,
                                     vp, vp, C.c_int]
        L.ref_chem_cells.restype = C.c_int

    def chem_cells(self, oracle, dt, T, rho, Y, nthreads=1):
        """The chemistry step over a flat list of cells (no markers), each cell
        integrated by the reference's DDRIV3 with the oracle's rate constants
        and right-hand side; same outputs as Oracle.chem_cells."""
        n = len(T)
        S = oracle.S
        T = np.ascontiguousarray(T, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        yp = np.zeros((S, n))
        yp2 = np.zeros((S, n))
        ef = np.zeros(n)
        st = np.zeros(n, dtype=np.int32)
        ns, nf, nj = (np.zeros(n, dtype=np.int32) for _ in range(3))
        rc = self.lib.ref_chem_cells(oracle.mp, C.cast(C.pointer(oracle.opts), C.c_void_p), n, dt,
                                     T.ctypes.data, rho.ctypes.data, _ptrs(list(Y)),
                                     _ptrs(list(yp)), _ptrs(list(yp2)), ef.ctypes.data,
                                     st.ctypes.data, ns.ctypes.data, nf.ctypes.data,
                                     nj.ctypes.data, int(nthreads))
        if rc != 0:
            raise RuntimeError("ref_chem_cells failed (%d)" % rc)
        return dict(Y_prod=yp, Y_prod2=yp2, E_f_prod=ef, status=st, nstep=ns, nfe=nf, nje=nj)

    def dnrm2(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = C.c_int(len(x))
        inc = C.c_int(1)
        return self.lib.dnrm2_(C.byref(n), _as_dp(x), C.byref(inc))

    def dgefa(self, a):
        n = a.shape[0]
        f = np.asfortranarray(a, dtype=np.float64).copy(order="F")
        ip = np.zeros(n, dtype=np.int32)
        info = C.c_int(0)
        nn = C.c_int(n)
        self.lib.dgefa_(f.ctypes.data_as(_dp), C.byref(nn), C.byref(nn), ip.ctypes.data_as(_ip),
                        C.byref(info))
        return f, ip - 1, info.value

    def dgesl(self, lu, ipvt0, b):
        n = lu.shape[0]
        bb = np.ascontiguousarray(b, dtype=np.float64).copy()
        ip = (np.asarray(ipvt0, dtype=np.int32) + 1).copy()
        nn = C.c_int(n)
        job = C.c_int(0)
        self.lib.dgesl_(lu.ctypes.data_as(_dp), C.byref(nn), C.byref(nn), ip.ctypes.data_as(_ip),
                        _as_dp(bb), C.byref(job))
        return bb

    def ddcst(self, maxord=5):
        el = np.zeros((12, 13))
        tq = np.zeros((12, 3))
        mo = C.c_int(maxord)
        mint = C.c_int(2)
        isw = C.c_int(2)
        self.lib.ddcst_(C.byref(mo), C.byref(mint), C.byref(isw), _as_dp(el), _as_dp(tq))
        return el, tq

    def ddriv3_cell(self, oracle, T, rho, Y, dt, eps=1e-9, ewt=1e-4, hmax=None, mxstep=10000):
        """One cell through the reference DDRIV3 with the oracle's chemistry RHS."""
        ctx, cin = oracle.cell_context(T, rho, Y)
        y, info = self._ddriv3(oracle.chem_fun, C.cast(ctx, C.c_void_p), cin, dt, eps, ewt, hmax,
                               mxstep, 5)
        return y, cin, info

    def ddriv3(self, rhs, y0, tout, eps=1e-9, ewt=1e-4, hmax=None, mxstep=10000, mxord=5):
        def cb(ctx, nn, t, yp, ydp):
            yy = np.ctypeslib.as_array(yp, shape=(nn,))
            out = np.ctypeslib.as_array(ydp, shape=(nn,))
            out[:] = rhs(t, yy.copy())

        return self._ddriv3(RHS_FN(cb), None, y0, tout, eps, ewt, hmax, mxstep, mxord)

    def _ddriv3(self, f, ctx, y0, tout, eps, ewt, hmax, mxstep, mxord):
        if hmax is None:
            hmax = float(np.float32(1.0e-7))
        y = np.ascontiguousarray(y0, dtype=np.float64).copy()
        n = len(y)
        ierflg = C.c_int(0)
        iwork = np.zeros(n + 50, dtype=np.int32)
        work = np.zeros(n * n + 10 * n + 250)
        nstate = self.lib.ref_ddriv3(n, _as_dp(y), tout, eps, ewt, hmax, mxstep, mxord, f, ctx,
                                     C.byref(ierflg), iwork.ctypes.data_as(_ip), _as_dp(work))
        return y, dict(nstate=nstate, ierflg=ierflg.value, nqused=int(iwork[1]),
                       nstep=int(iwork[2]), nfe=int(iwork[3]), nje=int(iwork[4]))
