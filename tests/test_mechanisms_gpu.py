"""Mechanisms other than KEROMNES end to end: ZHANG (9 species, 28 reactions),
TEREZA (10 species, 24 reactions, another species order), KONNOV (13
species, 71 reactions) and ACETYLENE (17 species, 25 irreversible reactions,
in C2H2-air), all shipped with NRG's task set-up, run on their own builds of
the kernels, picked by
nrg_chem_create from the mechanism's reaction lists, and match the oracle
(a restatement of the reference step for any mechanism) to the parity bars;
a task set-up folder whose chemical_data.inf names ZHANG gives the same
handle and the same step."""
import os
import shutil

import numpy as np
import pytest

from gpu_util import TOL_T_REL, TOL_Y_ABS, gpu_flat_step
from nrg_amd import capi, inputs
from nrg_amd import mechanism as M

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "nrg_amd", "data")


@pytest.fixture(scope="module")
def zhang():
    return M.load_mechanism("ZHANG.txt", "ZHANG_THERMO.txt")


MECHS = {"ZHANG": ("ZHANG.txt", "ZHANG_THERMO.txt"), "TEREZA": ("TEREZA.txt", "TEREZA_THERMO.txt"),
         "KONNOV": ("KONNOV.txt", "KONNOV_THERMO.txt"),
         "ACETYLENE": ("ACETYLENE_Varatharajan.txt", "FFCM-1_thermo.txt")}


def _acetylene_front(mech, nx, ny, seed):
    """A 2-D tanh flame front in C2H2-air (the acetylene mechanism's own fuel):
    fresh C2H2 : O2 : N2 = phi : 2.5 : 2.5*3.762, burnt = complete-combustion
    products (CO2, H2O, the O2 left over) with small OH/H/O/CO fractions,
    T 300-2600 K with 2 % noise, p = 1 atm, phi ~ U[0.5, 1]."""
    import torch

    from nrg_amd.mechanism import P_ATM, R_GAS_J
    names = list(mech.species_names)
    ix = {n: names.index(n) for n in ("C2H2", "O2", "N2", "CO2", "H2O", "OH", "H", "O", "CO")}
    g = torch.Generator().manual_seed(seed)
    n = nx * ny
    x = (torch.arange(n) % nx).to(torch.float64)
    c = 0.5 * (1.0 + torch.tanh((x - 0.5 * nx) / 6.0))
    T = (300.0 + 2300.0 * c) * (1.0 + 0.02 * torch.randn(n, generator=g, dtype=torch.float64))
    phi = 0.5 + 0.5 * torch.rand(n, generator=g, dtype=torch.float64)
    S = len(names)
    Xu = torch.zeros((S, n), dtype=torch.float64)
    Xu[ix["C2H2"]], Xu[ix["O2"]], Xu[ix["N2"]] = phi, 2.5, 2.5 * 3.762
    Xb = torch.zeros_like(Xu)
    Xb[ix["CO2"]], Xb[ix["H2O"]] = 2.0 * phi, phi
    Xb[ix["O2"]], Xb[ix["N2"]] = 2.5 * (1.0 - phi), 2.5 * 3.762
    Xu, Xb = Xu / Xu.sum(0), Xb / Xb.sum(0)
    for k, v in (("OH", 0.01), ("H", 0.005), ("O", 0.003), ("CO", 0.005)):
        Xb[ix[k]] += v
    Xb = Xb / Xb.sum(0)
    X = (1.0 - c) * Xu + c * Xb
    X = X / X.sum(0)
    W = torch.tensor(mech.molar_masses, dtype=torch.float64).view(-1, 1)
    Wmix = (X * W).sum(0)
    Y = X * W / Wmix
    Y = Y / Y.sum(0, keepdim=True)
    p = torch.full((n,), P_ATM, dtype=torch.float64)
    return dict(T=T, p=p, rho=p * Wmix / (R_GAS_J * T), Y=Y.contiguous())


def _closure(engine, state, out, dt):
    import torch
    n = state["T"].shape[0]
    S = state["Y"].shape[0]
    grid = capi.flat_grid(n)
    T, Y = state["T"].clone(), state["Y"].clone()
    E = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    engine.energy_from_temperature_dev(grid, T.data_ptr(), [Y[s].data_ptr() for s in range(S)],
                                       E.data_ptr())
    E0 = E.clone()
    Yp = torch.as_tensor(out["Y_prod"], device="cuda:0").contiguous()
    Ef = torch.as_tensor(out["E_f_prod"], device="cuda:0").contiguous()
    engine.closure_dev(grid, dt, state["rho"].data_ptr(), Ef.data_ptr(),
                       [Yp[s].data_ptr() for s in range(S)], E.data_ptr(),
                       [Y[s].data_ptr() for s in range(S)], T.data_ptr())
    torch.cuda.synchronize()
    return E0.cpu().numpy(), T.cpu().numpy(), Y.cpu().numpy()


@pytest.mark.parametrize("name", sorted(MECHS))
def test_step_and_closure_match_the_oracle(name):
    from nrg_amd.solver import ChemistryStep
    from oracle import Oracle
    mech = M.load_mechanism(*MECHS[name])
    eng = ChemistryStep(mech, capi.default_opts(form=capi.NRG_FORM_PERSISTENT,
                                                cell_order=capi.NRG_ORDER_STIFFNESS, use_graph=1))
    try:
        assert eng.lib.nrg_chem_topology(eng.h) == name.encode()
        dt = 1.0e-7
        st = (_acetylene_front(mech, 128, 64, 4321) if name == "ACETYLENE"
              else inputs.flame_front_2d(mech, 128, 64, seed=4321))
        stn = inputs.to_numpy(st)
        cu = {k: v.to("cuda:0") for k, v in st.items()}
        orc = Oracle(mech)
        g = gpu_flat_step(eng, cu, dt)
        o = orc.chem_cells(dt, stn["T"], stn["rho"], stn["Y"], nthreads=8)
        integ = o["status"] == capi.NRG_CELL_OK
        assert np.array_equal(g["status"], o["status"]) and integ.sum() > 4000
        assert o["nfe"][integ].mean() > 20          # the chemistry is not idle
        dY = np.abs(g["Y_prod"] - o["Y_prod"]) * dt / stn["rho"]
        same = (g["nstep"] == o["nstep"]) & (g["nfe"] == o["nfe"]) & (g["nje"] == o["nje"])
        E0, T_g, Y_g = _closure(eng, cu, g, dt)
        _, Y_o, T_o, _ = orc.closure_cells(dt, stn["rho"], o["E_f_prod"], o["Y_prod"], E0, stn["Y"],
                                           stn["T"], nthreads=8)
        relT = np.abs(T_g - T_o) / T_o
        print("[%s] cells %d integrated %d: max |dY| %.3e, max rel T %.3e, work identical %.4f, "
              "mean NFE %.1f" % (name, stn["T"].size, integ.sum(), dY.max(), relT.max(),
                                 same[integ].mean(), o["nfe"][integ].mean()), flush=True)
        assert dY.max() <= TOL_Y_ABS
        assert relT.max() <= TOL_T_REL
        assert np.abs(Y_g - Y_o).max() <= TOL_Y_ABS
        assert same[integ].mean() >= 0.99
        if name == "ZHANG":
            # the same state under KEROMNES (same species) integrates differently
            from nrg_amd.mechanism import load_mechanism
            k = ChemistryStep(load_mechanism())
            try:
                assert k.lib.nrg_chem_topology(k.h) == b"KEROMNES"
                gk = gpu_flat_step(k, cu, dt)
                assert not np.array_equal(gk["Y_prod"], g["Y_prod"])
            finally:
                k.close()
    finally:
        eng.close()


def test_task_setup_naming_zhang_runs_on_its_build(zhang, tmp_path):
    """chemical_data.inf and thermophysical_setup.inf (the namelists NRG's main
    program reads) name ZHANG: the C++ loader reads the folder, nrg_chem_create
    picks the ZHANG build, and the step equals the one from the direct load."""
    import ctypes as C

    from nrg_amd.solver import ChemistryStep
    d = tmp_path / "task"
    os.makedirs(d / "chemical_mechanisms")
    os.makedirs(d / "thermophysical_data")
    shutil.copy(os.path.join(DATA, "ZHANG.txt"), d / "chemical_mechanisms")
    shutil.copy(os.path.join(DATA, "ZHANG_THERMO.txt"), d / "thermophysical_data")
    shutil.copy(os.path.join(DATA, "molar_masses.dat"), d / "thermophysical_data")
    (d / "chemical_data.inf").write_text(
        "&chemical_properties\n chemical_mechanism_file_name = 'ZHANG.txt'\n"
        " default_enhanced_efficiencies = 1.0\n E_act_units = 'cal.mol'\n/\n")
    (d / "thermophysical_setup.inf").write_text(
        "&thermo_properties\n thermo_data_file_name = 'ZHANG_THERMO.txt'\n"
        " molar_masses_data_file_name = 'molar_masses.dat'\n/\n")
    native = M.load_task_setup_native(str(d))
    for a in ("reactions_type", "chemical_coeffs", "A", "E_act", "a_coeffs", "molar_masses"):
        assert np.array_equal(np.asarray(getattr(native, a)), np.asarray(getattr(zhang, a))), a
    a, b = ChemistryStep(native), ChemistryStep(zhang)
    try:
        assert a.lib.nrg_chem_topology(a.h) == b"ZHANG"
        st = inputs.flame_front_2d(zhang, 48, 16, seed=9)
        cu = {k: v.to("cuda:0") for k, v in st.items()}
        ga, gb = gpu_flat_step(a, cu, 1e-7), gpu_flat_step(b, cu, 1e-7)
        for k in ("Y_prod", "E_f_prod", "status", "nfe"):
            assert np.array_equal(ga[k], gb[k]), k
    finally:
        a.close()
        b.close()


def _screen_only(eng, cu, dt, S, ytol, ttol, fp64):
    """One learned step with gate.screen = ONLY (no network) on eng."""
    import ctypes as C

    import torch
    n = cu["T"].shape[0]
    g = capi.NrgSurrogateGate()
    g.screen = capi.NRG_SUR_SCREEN_ONLY
    g.dY_freeze, g.dT_freeze = ytol, ttol
    g.screen_fp64 = 1 if fp64 else 0
    g.n_elements = 0
    Yp = torch.zeros((S, n), dtype=torch.float64, device="cuda:0")
    Yp2 = torch.full((S, n), 7.0, dtype=torch.float64, device="cuda:0")
    Ef = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    status = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
    Y = cu["Y"].contiguous()
    stats = capi.NrgSurrogateStats()
    torch.cuda.synchronize()
    rc = eng.lib.nrg_chem_surrogate_step_dev(
        eng.h, None, None, C.byref(g), C.byref(capi.flat_grid(n)), dt,
        cu["T"].data_ptr(), cu["p"].data_ptr(), cu["rho"].data_ptr(),
        capi.ptr_array([Y[s].data_ptr() for s in range(S)]), None,
        capi.ptr_array([Yp[s].data_ptr() for s in range(S)]),
        capi.ptr_array([Yp2[s].data_ptr() for s in range(S)]), Ef.data_ptr(), status.data_ptr(),
        C.byref(stats))
    capi.check(eng.lib, rc)
    return dict(Y_prod=Yp.cpu().numpy(), Y_prod2=Yp2.cpu().numpy(), E_f_prod=Ef.cpu().numpy(),
                status=status.cpu().numpy(), stats=stats)


@pytest.mark.parametrize("name", sorted(MECHS))
def test_screen_fp32_pass_keeps_the_fp64_decisions(name):
    """The explicit-rate screen on each mechanism's own build (third-body,
    Troe and irreversible reactions; 9 to 17 species): the FP32 pass decides
    only where its bound settles the decision, so outputs, statuses and
    counters equal those of the screen with the FP64 estimate deciding every
    live cell."""
    from nrg_amd.solver import ChemistryStep
    mech = M.load_mechanism(*MECHS[name])
    S = mech.species_number
    eng = ChemistryStep(mech, capi.default_opts())
    try:
        st = (_acetylene_front(mech, 128, 64, 4321) if name == "ACETYLENE"
              else inputs.flame_front_2d(mech, 128, 64, seed=4321))
        cu = {k: v.to("cuda:0") for k, v in st.items()}
        for ytol, ttol in ((1e-6, 0.01), (1e-4, 1.0)):
            a = _screen_only(eng, cu, 1e-7, S, ytol, ttol, False)
            b = _screen_only(eng, cu, 1e-7, S, ytol, ttol, True)
            sa, sb = a["stats"], b["stats"]
            live = int(sa.cells_inner - sa.cells_gated)
            print("[%s] screen dY_freeze %g: live %d frozen %d decided by the FP64 estimate %d"
                  % (name, ytol, live, sa.cells_frozen, sa.cells_screen_fp64), flush=True)
            assert sb.cells_screen_fp64 == live and sa.cells_screen_fp64 < live
            assert 0 < sa.cells_frozen < live
            for k in ("Y_prod", "Y_prod2", "E_f_prod", "status"):
                assert np.array_equal(a[k], b[k]), k
            for k in ("cells_accepted", "cells_fallback", "cells_gated", "cells_frozen"):
                assert getattr(sa, k) == getattr(sb, k), k
    finally:
        eng.close()
