"""Host-side mirror of NRG's chemistry-solver interface over the C-ABI.

NRG's flow solvers own named fields in a ``data_manager`` and drive chemistry
through the ``chemical_kinetics_solver`` type::

    chem_kin_solver = chemical_kinetics_solver_c(manager)       ! cabaret_solver.f90:621
    call chem_kin_solver%solve_chemical_kinetics(time_step)     ! cabaret_solver.f90:1362

This module keeps those names and that argument meaning:

* ``DataManager`` holds device-resident FP64 fields in NRG's layout
  (Fortran column-major with ghost layers, one allocation per species) and
  creates fields by name (data_manager%create_*_field);
* ``chemical_kinetics_solver_c(manager)`` looks up ``temperature``,
  ``pressure``, ``density``, ``specie_mass_fraction`` and creates
  ``energy_production_chemistry``, ``specie_production_chemistry`` and
  ``specie_production_chemistry2`` (chemical_kinetics_solver.f90:121-138);
* ``solve_chemical_kinetics(time_step)`` runs the gfx950 step in place.

Where the reference would STOP (SLATEC fatal errors, slatec/xerhlt.f90:38)
this raises ``NrgChemError``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import NrgChemStats, check, load_library
from .mechanism import load_mechanism


def _torch():
    import torch
    return torch


def _order_after_torch():
    """The library launches on its own HIP stream; work torch queued on its
    current stream (field fills, copies) must finish before the library reads
    those fields.  Every library call returns with its own stream drained."""
    torch = _torch()
    if torch.cuda.is_available() and torch.cuda.is_initialized():
        torch.cuda.current_stream().synchronize()


class Field:
    """One FP64 (or int8) field with NRG's bounds: Fortran (lo:hi) per dim, column-major.

    ``data`` is a torch tensor of shape (nz, ny, nx) (C order) == Fortran (nx, ny, nz).
    """

    def __init__(self, name, short_name, grid, device, dtype=None, fill=0.0):
        torch = _torch()
        self.name = name
        self.name_short = short_name
        shp = tuple(grid.hi[d] - grid.lo[d] + 1 for d in (2, 1, 0))
        dt = dtype if dtype is not None else torch.float64
        self.data = torch.full(shp, fill, dtype=dt, device=device)

    def ptr(self):
        return self.data.data_ptr()

    def inner(self, grid):
        """View of the inner cells, shape (iz, iy, ix)."""
        sl = tuple(slice(grid.inner_lo[d] - grid.lo[d], grid.inner_hi[d] - grid.lo[d] + 1)
                   for d in (2, 1, 0))
        return self.data[sl]


class VectorField:
    """Per-species (or per-dimension) list of scalar fields: NRG's field_vector (Y%pr(s)%cells)."""

    def __init__(self, name, short_name, n, grid, device, fill=0.0):
        self.name = name
        self.name_short = short_name
        self.pr = [Field("%s_%d" % (name, i + 1), "%s_%d" % (short_name, i + 1), grid, device,
                         fill=fill) for i in range(n)]

    def ptrs(self):
        return [f.ptr() for f in self.pr]


class DataManager:
    """Named device fields over one NRG grid (package_library/src/data_manager_class.f90)."""

    def __init__(self, shape, device="cuda:0", ghost=1, mech=None):
        torch = _torch()
        self.grid = capi.grid_for_shape(shape, ghost)
        self.shape = tuple(shape) + (1,) * (3 - len(shape))
        self.device = torch.device(device)
        self.chemistry = mech if mech is not None else load_mechanism()
        self.fields = {}
        self.bc_markers = Field("bc_markers", "bc", self.grid, self.device, dtype=torch.int8, fill=0)
        S = self.chemistry.species_number
        self.create_scalar_field("temperature", "T")
        self.create_scalar_field("pressure", "p")
        self.create_scalar_field("density", "rho")
        self.create_scalar_field("full_energy", "E_f")
        self.create_vector_field("specie_mass_fraction", "Y", S)

    def create_scalar_field(self, name, short_name, fill=0.0):
        f = Field(name, short_name, self.grid, self.device, fill=fill)
        self.fields[name] = f
        return f

    def create_vector_field(self, name, short_name, n, fill=0.0):
        f = VectorField(name, short_name, n, self.grid, self.device, fill=fill)
        self.fields[name] = f
        return f

    def get_cons_field_pointer_by_name(self, name):
        try:
            return self.fields[name]
        except KeyError:
            raise KeyError("field %r is not registered in the data manager" % name) from None

    def inner_cells(self):
        g = self.grid
        return int(np.prod([g.inner_hi[d] - g.inner_lo[d] + 1 for d in range(3)]))


class ChemistryStep:
    """Owns one nrg_chem handle (constructor / destructor of the C-ABI)."""

    def __init__(self, mech=None, opts=None, device_index=0):
        self.lib = load_library()
        self.mech = mech if mech is not None else load_mechanism()
        self.ms = capi.MechStruct(self.mech)
        o = opts if opts is not None else capi.default_opts()
        o.device = device_index
        self.opts = o
        h = C.c_void_p()
        check(self.lib, self.lib.nrg_chem_create(C.byref(self.ms.struct), C.byref(o), C.byref(h)))
        self.h = h
        self.last_stats = None

    def close(self):
        if self.h:
            self.lib.nrg_chem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step_dev(self, grid, dt, T, rho, Y, Y_prod, Y_prod2, E_f_prod, bc=None, status=None,
                 counters=None, p=None):
        """Device-pointer step.  Fields are ints (device addresses); Y* are lists of addresses."""
        _order_after_torch()
        st = NrgChemStats()
        Yp = capi.ptr_array(Y)
        Ypp = capi.ptr_array(Y_prod)
        Yp2 = capi.ptr_array(Y_prod2)
        if counters is None:
            rc = self.lib.nrg_chem_step_dev(self.h, C.byref(grid), float(dt), T, p, rho, Yp, bc,
                                            Ypp, Yp2, E_f_prod, status, C.byref(st))
        else:
            rc = self.lib.nrg_chem_step_dev_diag(self.h, C.byref(grid), float(dt), T, p, rho, Yp,
                                                 bc, Ypp, Yp2, E_f_prod, status, C.byref(st),
                                                 *counters)
        self.last_stats = st
        check(self.lib, rc)
        return st

    def set_table(self, T_t, dC_t):
        """Load a table dC(T) for the table-approximated step: T_t[rows] and
        dC_t[S, rows] (host arrays; copied to the device)."""
        T_t = np.ascontiguousarray(T_t, dtype=np.float64)
        dC_t = np.ascontiguousarray(dC_t, dtype=np.float64)
        if T_t.ndim != 1 or dC_t.shape != (self.mech.species_number, T_t.shape[0]):
            raise ValueError("table shapes: T_t (rows,), dC_t (S, rows); got %s, %s"
                             % (T_t.shape, dC_t.shape))
        check(self.lib, self.lib.nrg_chem_set_table(self.h, int(T_t.shape[0]),
                                                    T_t.ctypes.data, dC_t.ctypes.data))

    def table_step_dev(self, grid, dt, T, rho, Y, Y_prod, Y_prod2, E_f_prod, bc=None,
                       status=None, p=None):
        """Table-approximated step on device pointers (same fields as step_dev)."""
        _order_after_torch()
        st = NrgChemStats()
        rc = self.lib.nrg_chem_table_step_dev(self.h, C.byref(grid), float(dt), T, p, rho,
                                              capi.ptr_array(Y), bc, capi.ptr_array(Y_prod),
                                              capi.ptr_array(Y_prod2), E_f_prod, status,
                                              C.byref(st))
        self.last_stats = st
        check(self.lib, rc)
        return st

    # -- low-Mach (FDS) coupling, include/nrg_low_mach.h -------------------------
    def low_mach_state_equation_dev(self, grid, dt, predictor, rho, Y, dp_stat_dt, p_stat,
                                    p_stat_old, T, mix_mol_mass, v=None, bc=None, p=None,
                                    mixture_cp=None, gamma=None, h_s=None, v_s=None, e_i=None,
                                    E_f=None):
        _order_after_torch()
        vv = capi.ptr_array(v) if v else None
        rc = self.lib.nrg_low_mach_state_equation_dev(
            self.h, C.byref(grid), float(dt), int(bool(predictor)), rho, capi.ptr_array(Y), vv,
            len(v) if v else 0, dp_stat_dt, bc, p_stat, p_stat_old, T, mix_mol_mass, p, mixture_cp,
            gamma, h_s, v_s, e_i, E_f)
        check(self.lib, rc)

    def low_mach_species_dev(self, grid, dt, corrector, rho_a, Y_a, rho_out, Y_out, rho_b=None,
                             Y_b=None, rhs=None, Y_prod_diff=None, Y_prod_chem=None, bc=None):
        _order_after_torch()
        pa = lambda y: capi.ptr_array(y) if y else None  # noqa: E731
        rc = self.lib.nrg_low_mach_species_dev(
            self.h, C.byref(grid), float(dt), int(bool(corrector)), rho_a, pa(Y_a), rho_b, pa(Y_b),
            pa(rhs), pa(Y_prod_diff), pa(Y_prod_chem), bc, rho_out, pa(Y_out))
        check(self.lib, rc)

    def low_mach_chem_divergence_dev(self, grid, rho, T, Y, mix_mol_mass, E_f_prod_chem,
                                     Y_prod_chem, div_v, bc=None):
        _order_after_torch()
        rc = self.lib.nrg_low_mach_chem_divergence_dev(
            self.h, C.byref(grid), rho, T, capi.ptr_array(Y), mix_mol_mass, E_f_prod_chem,
            capi.ptr_array(Y_prod_chem), bc, div_v)
        check(self.lib, rc)

    def closure_dev(self, grid, dt, rho, E_f_prod, Y_prod, E_f, Y, T, p=None, bc=None, v=None):
        _order_after_torch()
        vv = capi.ptr_array(v) if v else None
        nd = len(v) if v else 0
        rc = self.lib.nrg_chem_closure_dev(self.h, C.byref(grid), float(dt), rho, vv, nd, E_f_prod,
                                           capi.ptr_array(Y_prod), bc, E_f, capi.ptr_array(Y), T, p)
        check(self.lib, rc)

    def state_equation_dev(self, grid, rho, E_f, Y, T, p=None, bc=None, v=None):
        _order_after_torch()
        vv = capi.ptr_array(v) if v else None
        nd = len(v) if v else 0
        rc = self.lib.nrg_state_equation_dev(self.h, C.byref(grid), rho, vv, nd, E_f,
                                             capi.ptr_array(Y), bc, T, p)
        check(self.lib, rc)

    def energy_from_temperature_dev(self, grid, T, Y, E_f, v=None):
        _order_after_torch()
        vv = capi.ptr_array(v) if v else None
        nd = len(v) if v else 0
        rc = self.lib.nrg_energy_from_temperature_dev(self.h, C.byref(grid), T, capi.ptr_array(Y),
                                                      vv, nd, E_f)
        check(self.lib, rc)


class chemical_kinetics_solver:  # noqa: N801  (reference type name)
    """Mirror of NRG's ``chemical_kinetics_solver`` type (chemical_kinetics_solver.f90:81-103)."""

    def __init__(self, manager, opts=None, ODE_solver="slatec", table_file=None):  # noqa: N803
        self.manager = manager
        self.T = manager.get_cons_field_pointer_by_name("temperature")
        self.p = manager.get_cons_field_pointer_by_name("pressure")
        self.rho = manager.get_cons_field_pointer_by_name("density")
        self.Y = manager.get_cons_field_pointer_by_name("specie_mass_fraction")
        S = manager.chemistry.species_number
        self.E_f_prod = manager.create_scalar_field("energy_production_chemistry", "E_f_prod_chem")
        self.Y_prod = manager.create_vector_field("specie_production_chemistry", "Y_prod_chem", S)
        self.Y_prod2 = manager.create_vector_field("specie_production_chemistry2", "Y_prod_chem2", S)
        self.boundary = manager.bc_markers
        if ODE_solver not in ("slatec", "table_approximated"):
            raise ValueError("ODE_solver must be 'slatec' or 'table_approximated', not %r"
                             % (ODE_solver,))
        self.ODE_solver = ODE_solver
        dev = manager.device.index if manager.device.index is not None else 0
        self.engine = ChemistryStep(manager.chemistry, opts, dev)
        self.cell_status = Field("chemistry_cell_status", "chem_status", manager.grid,
                                 manager.device, dtype=_torch().int32, fill=0)
        self.T_t = self.Y_t = None
        if ODE_solver == "table_approximated":                          # :168-170
            if table_file is None:
                raise ValueError("ODE_solver='table_approximated' needs table_file")
            self.read_chemical_kinetics_table(table_file)

    def solve_chemical_kinetics(self, time_step):
        """One chemistry step of length ``time_step`` over all inner cells (:175-358):
        the SLATEC integration or, with ODE_solver 'table_approximated', the
        table lookup (:250-346)."""
        g = self.manager.grid
        step = (self.engine.table_step_dev if self.ODE_solver == "table_approximated"
                else self.engine.step_dev)
        return step(g, time_step, self.T.ptr(), self.rho.ptr(), self.Y.ptrs(),
                    self.Y_prod.ptrs(), self.Y_prod2.ptrs(), self.E_f_prod.ptr(),
                    bc=self.boundary.ptr(), status=self.cell_status.ptr(), p=self.p.ptr())

    def read_chemical_kinetics_table(self, table_file):
        """Load dC(T) from a table file (:608-648) into the library."""
        from . import chem_table
        self.T_t, self.Y_t = chem_table.read_table(table_file,
                                                   self.manager.chemistry.species_number)
        self.engine.set_table(self.T_t, self.Y_t)

    def write_chemical_kinetics_table(self, table_file):
        """Write T and Y_prod2 of the grid line j = k = 1 (:542-606), from the first
        cell with 300.5 K < T < 310 K (Fortran index 0 if none) to the hottest."""
        from . import chem_table
        g = self.manager.grid
        j, k = 1 - g.lo[1], 1 - g.lo[2]
        i0 = g.inner_lo[0] - g.lo[0]                        # array index of the first inner cell
        T_line = self.T.data[k, j, :].cpu().numpy()
        lo, hi = chem_table.table_rows(T_line[i0:g.inner_hi[0] - g.lo[0] + 1],
                                       none_index=-g.inner_lo[0])
        lo, hi = max(i0 + lo, 0), i0 + hi
        dC = np.stack([f.data[k, j, lo:hi + 1].cpu().numpy() for f in self.Y_prod2.pr])
        chem_table.write_table(table_file, T_line[lo:hi + 1], dC,
                               self.manager.chemistry.species_names)
        return lo, hi

    def autotune(self, time_step):
        """Run the library's autotuner (nrg_chem_autotune_dev) on the current
        fields; the chosen queue set-up stays in the handle.  Writes the
        chemistry outputs (as a step would)."""
        _order_after_torch()
        tun = capi.NrgChemTuning()
        e = self.engine
        rc = e.lib.nrg_chem_autotune_dev(
            e.h, C.byref(self.manager.grid), float(time_step), self.T.ptr(), self.p.ptr(),
            self.rho.ptr(), capi.ptr_array(self.Y.ptrs()), self.boundary.ptr(),
            capi.ptr_array(self.Y_prod.ptrs()), capi.ptr_array(self.Y_prod2.ptrs()),
            self.E_f_prod.ptr(), C.byref(tun))
        check(e.lib, rc)
        return tun

    def apply_split_sources(self, time_step):
        """Legacy CABARET source application + e->T (cabaret_solver.f90:3316-3416 and
        table_approximated_real_gas.f90:410-440), device-resident."""
        m = self.manager
        E_f = m.get_cons_field_pointer_by_name("full_energy")
        self.engine.closure_dev(m.grid, time_step, self.rho.ptr(), self.E_f_prod.ptr(),
                                self.Y_prod.ptrs(), E_f.ptr(), self.Y.ptrs(), self.T.ptr(),
                                p=self.p.ptr(), bc=self.boundary.ptr())


def chemical_kinetics_solver_c(manager, opts=None, ODE_solver="slatec", table_file=None):  # noqa: N803
    """Constructor with the reference's name (chemical_kinetics_solver.f90:105-111)."""
    return chemical_kinetics_solver(manager, opts, ODE_solver, table_file)
