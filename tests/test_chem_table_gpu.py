"""Table-approximated chemistry on the GPU (nrg_chem_set_table /
nrg_chem_table_step_dev) against the oracle's restatement of
chemical_kinetics_solver.f90:250-346, 650-685: bit for bit, on flat and
ghosted grids, and the reference's own round trip (a SLATEC run writes the
table, a table run reads it back)."""
import numpy as np
import pytest

from nrg_amd import capi, inputs
from nrg_amd import chem_table as ct

pytestmark = pytest.mark.gpu
S = 9


def _report(name, **kw):
    print("[table] %s " % name + " ".join("%s=%s" % (k, v) for k, v in kw.items()), flush=True)


def _tables():
    rng = np.random.default_rng(21)
    T_up = np.concatenate([[320.0], np.sort(rng.uniform(330.0, 2500.0, 62)), [2600.0]])
    dC_up = rng.standard_normal((S, T_up.size)) * 10.0 ** rng.uniform(-9, -3, (S, 1))
    # not monotonic: overlapping intervals, the last one that brackets T wins
    T_mix = np.array([400.0, 900.0, 1100.0, 800.0, 1200.0, 700.0, 2000.0])
    dC_mix = rng.standard_normal((S, T_mix.size)) * 1e-4
    # a table that ends below its start: both range rules fire between the ends
    T_down = np.array([2400.0, 1500.0, 600.0])
    dC_down = rng.standard_normal((S, T_down.size)) * 1e-5
    # more rows than the kernel stages in LDS (32 KB): read from global memory
    T_long = np.sort(rng.uniform(300.0, 2700.0, 600))
    dC_long = rng.standard_normal((S, T_long.size)) * 1e-4
    return {"ascending": (T_up, dC_up), "overlapping": (T_mix, dC_mix),
            "descending": (T_down, dC_down), "long": (T_long, dC_long)}


TABLES = _tables()


def _state(mech, n=5000, seed=4):
    st = inputs.to_numpy(inputs.ignition_mix(mech, n, seed=seed, frac_window=0.5))
    rng = np.random.default_rng(seed)
    T = rng.uniform(200.0, 2900.0, n)         # below, inside and above every table
    # the gate edge, the table ends and interior nodes (a node sits in two
    # intervals; the later one wins)
    T[:9] = [305.0, 304.999, 320.0, 2600.0, 1100.0, 900.0, 800.0, 1500.0, 2400.0]
    T_up = TABLES["ascending"][0]
    if n >= 200:
        T[9:9 + T_up.size] = T_up
        T[100:100 + T_up.size - 1] = np.nextafter(T_up[1:], 0.0)  # just below each node
    bc = np.zeros(n, dtype=np.int8)
    bc[7:40:3] = 1
    return T, st["rho"], np.ascontiguousarray(st["Y"]), bc


def _gpu_table_step(engine, dt, T, rho, Y, bc, fill2=-7.0):
    import torch
    n = T.shape[0]
    cu = lambda a, dt_=torch.float64: torch.as_tensor(a, dtype=dt_, device="cuda:0")  # noqa: E731
    Tg, rg = cu(T), cu(rho)
    Yg = [cu(Y[s]) for s in range(S)]
    bcg = cu(bc, torch.int8)
    Yp = torch.full((S, n), 3.0, dtype=torch.float64, device="cuda:0")
    Yp2 = torch.full((S, n), fill2, dtype=torch.float64, device="cuda:0")
    Ef = torch.full((n,), 5.0, dtype=torch.float64, device="cuda:0")
    status = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
    stats = engine.table_step_dev(capi.flat_grid(n), dt, Tg.data_ptr(), rg.data_ptr(),
                                  [y.data_ptr() for y in Yg], [Yp[s].data_ptr() for s in range(S)],
                                  [Yp2[s].data_ptr() for s in range(S)], Ef.data_ptr(),
                                  bc=bcg.data_ptr(), status=status.data_ptr())
    torch.cuda.synchronize()
    return dict(Y_prod=Yp.cpu().numpy(), Y_prod2=Yp2.cpu().numpy(), E_f_prod=Ef.cpu().numpy(),
                status=status.cpu().numpy(), stats=stats)


def _bits(a):
    return np.ascontiguousarray(a).view(np.int64)


@pytest.mark.parametrize("name", sorted(TABLES))
def test_table_step_bitwise_vs_oracle(engine, oracle, mech, name):
    T_t, dC = TABLES[name]
    T, rho, Y, bc = _state(mech)
    dt = 1e-7
    print("[table] case=%s rows=%d cells=%d" % (name, T_t.size, T.size), flush=True)
    engine.set_table(T_t, dC)
    g = _gpu_table_step(engine, dt, T, rho, Y, bc)
    o = oracle.table_cells(dt, T_t, dC, T, Y, bc=bc, Y_prod2=np.full((S, T.size), -7.0))
    assert np.array_equal(g["status"], o["status"])
    assert np.array_equal(_bits(g["Y_prod"]), _bits(o["Y_prod"]))
    assert np.array_equal(_bits(g["Y_prod2"]), _bits(o["Y_prod2"]))
    assert np.array_equal(_bits(g["E_f_prod"]), _bits(o["E_f_prod"]))
    integ = int((o["status"] == capi.NRG_CELL_OK).sum())
    assert g["stats"].cells_inner == T.size and g["stats"].cells_integrated == integ
    _report(name, integrated=integ, gated=T.size - integ, kernel_ms=g["stats"].kernel_ms)


def test_host_array_table_step(engine, oracle, mech):
    """nrg_chem_table_step on host arrays (staged through the handle's device
    buffers) against the oracle, bit for bit."""
    import ctypes as C
    T_t, dC = TABLES["ascending"]
    T, rho, Y, bc = _state(mech, n=3000, seed=6)
    n = T.size
    dt = 1e-7
    engine.set_table(T_t, dC)
    Yp = np.full((S, n), 3.0)
    Yp2 = np.full((S, n), -7.0)
    Ef = np.zeros(n)
    status = np.full(n, -1, dtype=np.int32)
    st = capi.NrgChemStats()
    ptrs = lambda a: capi.ptr_array([a[s].ctypes.data for s in range(S)])  # noqa: E731
    rc = engine.lib.nrg_chem_table_step(engine.h, C.byref(capi.flat_grid(n)), dt, T.ctypes.data,
                                        None, rho.ctypes.data, ptrs(Y), bc.ctypes.data, ptrs(Yp),
                                        ptrs(Yp2), Ef.ctypes.data, status.ctypes.data, C.byref(st))
    capi.check(engine.lib, rc)
    o = oracle.table_cells(dt, T_t, dC, T, Y, bc=bc, Y_prod2=np.full((S, n), -7.0))
    assert np.array_equal(status, o["status"])
    assert np.array_equal(_bits(Yp), _bits(o["Y_prod"]))
    assert np.array_equal(_bits(Yp2), _bits(o["Y_prod2"]))
    assert np.array_equal(_bits(Ef), _bits(o["E_f_prod"]))
    assert st.cells_integrated == int((o["status"] == capi.NRG_CELL_OK).sum())


def test_table_step_argument_errors(engine, mech):
    from nrg_amd.capi import NrgChemError
    from nrg_amd.solver import ChemistryStep
    fresh = ChemistryStep(mech)
    try:
        T, rho, Y, bc = _state(mech, n=64)
        with pytest.raises(NrgChemError, match="no table"):
            _gpu_table_step(fresh, 1e-7, T, rho, Y, bc)
        import ctypes as C
        st = capi.NrgChemStats()
        with pytest.raises(NrgChemError, match="no table"):
            capi.check(fresh.lib, fresh.lib.nrg_chem_table_step(
                fresh.h, C.byref(capi.flat_grid(T.size)), 1e-7, T.ctypes.data, None, rho.ctypes.data,
                None, None, None, None, None, None, C.byref(st)))
        with pytest.raises(NrgChemError, match="2 rows"):
            fresh.set_table(np.array([300.0]), np.zeros((S, 1)))
        fresh.set_table(*TABLES["descending"])
        with pytest.raises(NrgChemError, match="time_step"):
            _gpu_table_step(fresh, 0.0, T, rho, Y, bc)
    finally:
        fresh.close()


def _fill_manager(dm, mech, stn, shape):
    import torch
    nx, ny, nz = shape
    g = dm.grid
    for name, key in (("temperature", "T"), ("density", "rho"), ("pressure", "p")):
        dm.fields[name].inner(g).copy_(torch.as_tensor(stn[key]).view(nz, ny, nx))
    Yf = dm.fields["specie_mass_fraction"]
    for s in range(mech.species_number):
        Yf.pr[s].inner(g).copy_(torch.as_tensor(stn["Y"][s]).view(nz, ny, nx))


def _inner(field, g):
    return field.inner(g).cpu().numpy().reshape(-1)


def test_solver_table_mode_on_ghosted_grid(oracle, mech, tmp_path):
    """chemical_kinetics_solver with ODE_solver='table_approximated' reading a
    table file: ghost layers, markers and the gate, against the oracle on the
    table as read back from the file."""
    import torch
    from nrg_amd.solver import DataManager, chemical_kinetics_solver_c
    shape = (16, 9, 5)
    nx, ny, nz = shape
    dm = DataManager(shape, device="cuda:0", mech=mech)
    stn = inputs.to_numpy(inputs.hot_kernels_3d(mech, shape, seed=9, n_kernels=2,
                                                r_range=(2.0, 5.0)))
    stn["T"][::13] = 299.0
    _fill_manager(dm, mech, stn, shape)
    bc = np.zeros(stn["T"].size, dtype=np.int8)
    bc[::9] = 1
    dm.bc_markers.inner(dm.grid).copy_(torch.as_tensor(bc).view(nz, ny, nx))
    path = str(tmp_path / "15.0_pcnt_H2-Air_table(T).dat")
    T_t, dC = TABLES["ascending"]
    ct.write_table(path, T_t, dC, mech.species_names)
    solver = chemical_kinetics_solver_c(dm, ODE_solver="table_approximated", table_file=path)
    for s in range(S):
        solver.Y_prod2.pr[s].data.fill_(-7.0)
    dt = 1e-7
    stats = solver.solve_chemical_kinetics(dt)
    torch.cuda.synchronize()
    T_r, dC_r = ct.read_table(path, S)
    assert np.allclose(T_r, T_t, rtol=5e-7, atol=0) and np.allclose(dC_r, dC, rtol=5e-7, atol=0)
    o = oracle.table_cells(dt, T_r, dC_r, stn["T"], stn["Y"], bc=bc,
                           Y_prod2=np.full((S, stn["T"].size), -7.0))
    g = dm.grid
    Yp = np.stack([_inner(solver.Y_prod.pr[s], g) for s in range(S)])
    Yp2 = np.stack([_inner(solver.Y_prod2.pr[s], g) for s in range(S)])
    assert np.array_equal(_inner(solver.cell_status, g), o["status"])
    assert np.array_equal(_bits(Yp), _bits(o["Y_prod"]))
    assert np.array_equal(_bits(Yp2), _bits(o["Y_prod2"]))
    assert np.array_equal(_bits(_inner(solver.E_f_prod, g)), _bits(o["E_f_prod"]))
    assert stats.cells_integrated == int((o["status"] == capi.NRG_CELL_OK).sum())
    full = solver.Y_prod.pr[0].data.cpu().numpy()
    inner_mask = np.zeros_like(full, dtype=bool)
    inner_mask[1:-1, 1:-1, 1:-1] = True
    assert np.all(full[~inner_mask] == 0.0)          # ghost cells untouched
    _report("ghosted", integrated=stats.cells_integrated, cells=stats.cells_inner)


def test_reference_round_trip_flame_line(oracle, mech, tmp_path):
    """The reference's workflow on a 1-D flame line: a SLATEC step, then
    write_chemical_kinetics_table, then a table step that reads the file.  The
    table step equals the oracle's on the table read back, bit for bit; on the
    table's own cells it reproduces the SLATEC Y_prod and E_f_prod to the
    file's 7 digits (the rounding of the row temperatures moves each cell off
    its node by up to 5e-7 relative: 1.5e-6 of each species' scale on this
    line)."""
    import torch
    from nrg_amd.solver import DataManager, chemical_kinetics_solver_c
    n = 256
    dm = DataManager((n,), device="cuda:0", mech=mech)
    st = inputs.to_numpy(ct.flame_line(mech, n, width=16.0))
    _fill_manager(dm, mech, st, (n, 1, 1))
    dt = 1e-7
    fp64 = chemical_kinetics_solver_c(dm)
    fp64.solve_chemical_kinetics(dt)
    torch.cuda.synchronize()
    g = dm.grid
    Yp_ref = np.stack([_inner(fp64.Y_prod.pr[s], g) for s in range(S)])
    Ef_ref = _inner(fp64.E_f_prod, g)
    path = str(tmp_path / "flame_table.dat")
    lo, hi = fp64.write_chemical_kinetics_table(path)
    # array indices with one ghost cell: inner cell q sits at q + 1
    lo_q, hi_q = lo - 1, hi - 1
    assert (lo_q, hi_q) == ct.table_rows(st["T"])
    assert hi_q == int(np.argmax(st["T"])) and 300.5 < st["T"][lo_q] < 310.0
    fp64.engine.close()

    tab = chemical_kinetics_solver_c(dm, ODE_solver="table_approximated", table_file=path)
    assert tab.T_t.size == hi_q - lo_q + 1
    tab.solve_chemical_kinetics(dt)
    torch.cuda.synchronize()
    Yp = np.stack([_inner(tab.Y_prod.pr[s], g) for s in range(S)])
    Ef = _inner(tab.E_f_prod, g)
    o = oracle.table_cells(dt, tab.T_t, tab.Y_t, st["T"], st["Y"])
    assert np.array_equal(_bits(Yp), _bits(o["Y_prod"]))
    assert np.array_equal(_bits(Ef), _bits(o["E_f_prod"]))
    rows = slice(lo_q, hi_q + 1)
    scale = np.abs(Yp_ref[:, rows]).max(axis=1, keepdims=True)
    err_y = (np.abs(Yp[:, rows] - Yp_ref[:, rows]) / np.maximum(scale, 1e-300)).max()
    err_e = (np.abs(Ef[rows] - Ef_ref[rows]) / np.abs(Ef_ref[rows]).max()).max()
    _report("round_trip", rows=hi_q - lo_q + 1, rel_Yprod=err_y, rel_Efprod=err_e)
    assert err_y <= 1e-5 and err_e <= 1e-5
    # the cells before the table's first row (T <= 300.5 K) are below the gate
    cold = st["T"] < 305.0
    assert cold[:lo_q].all() and np.all(Yp[:, cold] == 0.0) and np.all(Ef[cold] == 0.0)
    tab.engine.close()
