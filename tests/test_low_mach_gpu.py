"""Low-Mach (FDS) coupling kernels (include/nrg_low_mach.h) against the
oracle on the same cells: bit for bit (the sound speed to 1 ulp: the device
and glibc square roots), on a ghosted grid with markers, and chained after
the FP64 chemistry step in the FDS solver's order (chemistry, reactive
divergence, species predictor, equation of state)."""
import numpy as np
import pytest

from nrg_amd import capi, inputs

pytestmark = pytest.mark.gpu
S = 9
R = 8.31446261815324
SHAPE = (20, 12, 6)


def _report(name, **kw):
    print("[low_mach] %s " % name + " ".join("%s=%s" % (k, v) for k, v in kw.items()), flush=True)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def _ulps(a, b):
    return int(np.abs(_bits(a) - _bits(b)).max()) if np.size(a) else 0


class Ghosted:
    """Device fields in NRG's ghosted layout over SHAPE; host views of the inner cells."""

    def __init__(self):
        import torch
        self.torch = torch
        self.grid = capi.grid_for_shape(SHAPE)
        nx, ny, nz = SHAPE
        self.full = (nz + 2, ny + 2, nx + 2)
        self.n = nx * ny * nz

    def field(self, inner=None, fill=0.0, dtype=None):
        t = self.torch.full(self.full, fill, dtype=dtype or self.torch.float64, device="cuda:0")
        if inner is not None:
            nx, ny, nz = SHAPE
            t[1:-1, 1:-1, 1:-1] = self.torch.as_tensor(np.ascontiguousarray(inner)).view(nz, ny, nx)
        return t

    def inner(self, t):
        return t[1:-1, 1:-1, 1:-1].reshape(-1).cpu().numpy()


def _state(mech, n, seed):
    st = inputs.to_numpy(inputs.ignition_mix(mech, n, seed=seed, frac_window=0.5))
    return st["rho"], np.ascontiguousarray(st["Y"]), st["T"]


@pytest.mark.parametrize("predictor", [True, False])
def test_eos_bitwise_vs_oracle(engine, oracle, mech, predictor):
    gh = Ghosted()
    n = gh.n
    rho, Y, _ = _state(mech, n, seed=21)
    W = np.asarray(mech.molar_masses)
    rng = np.random.default_rng(3)
    Wm = 1.0 / (Y / W[:, None]).sum(0)
    p0 = rho * R * rng.uniform(320.0, 3000.0, n) / Wm
    p_old = p0 * rng.uniform(0.97, 1.03, n)
    dp = rng.uniform(-2e6, 2e6, n)
    v = rng.uniform(-40.0, 40.0, (2, n))
    bc = np.zeros(n, dtype=np.int8)
    bc[::7] = 1
    dt = 2e-6
    print("[low_mach] eos predictor=%d cells=%d" % (predictor, n), flush=True)
    f = {k: gh.field(fill=-1.0) for k in ("T", "mmm", "p", "cp", "gamma", "h_s", "v_s", "e_i")}
    f["p_stat"], f["p_old"] = gh.field(p0), gh.field(p_old)
    f["E_f"] = gh.field(np.full(n, -5.0))
    rho_d, dp_d = gh.field(rho), gh.field(dp)
    Y_d = [gh.field(Y[s]) for s in range(S)]
    v_d = [gh.field(v[d]) for d in range(2)]
    bc_d = gh.field(bc, dtype=gh.torch.int8)
    engine.low_mach_state_equation_dev(
        gh.grid, dt, predictor, rho_d.data_ptr(), [y.data_ptr() for y in Y_d],
        dp_d.data_ptr(), f["p_stat"].data_ptr(), f["p_old"].data_ptr(), f["T"].data_ptr(),
        f["mmm"].data_ptr(), v=[x.data_ptr() for x in v_d], bc=bc_d.data_ptr(),
        p=f["p"].data_ptr(), mixture_cp=f["cp"].data_ptr(), gamma=f["gamma"].data_ptr(),
        h_s=f["h_s"].data_ptr(), v_s=f["v_s"].data_ptr(), e_i=f["e_i"].data_ptr(),
        E_f=f["E_f"].data_ptr())
    gh.torch.cuda.synchronize()
    o = oracle.low_mach_eos(dt, predictor, rho, Y, p0, p_old, dp, v=v, E_f=np.full(n, -5.0),
                            bc=bc, fill=-1.0)
    names = dict(T="T", mmm="mix_mol_mass", p="p", cp="mixture_cp", gamma="gamma", h_s="h_s",
                 e_i="e_i", E_f="E_f", p_stat="p_stat", p_old="p_stat_old")
    for k, ok in names.items():
        assert np.array_equal(_bits(gh.inner(f[k])), _bits(o[ok])), k
    ulp = _ulps(gh.inner(f["v_s"]), o["v_s"])
    _report("eos", predictor=predictor, v_s_max_ulp=ulp)
    assert ulp <= 1
    full = f["T"].cpu().numpy()
    mask = np.ones_like(full, dtype=bool)
    mask[1:-1, 1:-1, 1:-1] = False
    assert np.all(full[mask] == -1.0)                   # ghost cells untouched


def test_eos_optional_outputs_and_no_velocity(engine, oracle, mech):
    import torch
    n = 300
    rho, Y, _ = _state(mech, n, seed=5)
    W = np.asarray(mech.molar_masses)
    p0 = rho * R * np.linspace(400.0, 2600.0, n) * (Y / W[:, None]).sum(0)
    cu = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda:0")  # noqa: E731
    rho_d, dp_d, ps_d, po_d = cu(rho), cu(np.zeros(n)), cu(p0), cu(np.zeros(n))
    T_d, m_d, Ef_d = cu(np.zeros(n)), cu(np.zeros(n)), cu(np.full(n, 9.0))
    Y_d = [cu(Y[s]) for s in range(S)]
    engine.low_mach_state_equation_dev(capi.flat_grid(n), 1e-6, True, rho_d.data_ptr(),
                                       [y.data_ptr() for y in Y_d], dp_d.data_ptr(),
                                       ps_d.data_ptr(), po_d.data_ptr(), T_d.data_ptr(),
                                       m_d.data_ptr(), E_f=Ef_d.data_ptr())
    torch.cuda.synchronize()
    o = oracle.low_mach_eos(1e-6, True, rho, Y, p0, np.zeros(n), np.zeros(n))
    assert np.array_equal(_bits(T_d.cpu().numpy()), _bits(o["T"]))
    assert np.all(Ef_d.cpu().numpy() == 9.0)            # ndim = 0: E_f untouched


@pytest.mark.parametrize("corrector", [False, True])
def test_species_bitwise_vs_oracle(engine, oracle, mech, corrector):
    gh = Ghosted()
    n = gh.n
    rho_a, Y_a, _ = _state(mech, n, seed=31)
    rho_b, Y_b, _ = _state(mech, n, seed=32)
    rng = np.random.default_rng(4)
    rhs = rng.standard_normal((S, n)) * 2e-3
    Yd = rng.standard_normal((S, n)) * 1e-3
    Yc = rng.standard_normal((S, n)) * 1e-2
    rhs[:, 5] = -1e12                                   # clamps to nothing: outputs untouched
    bc = np.zeros(n, dtype=np.int8)
    bc[3::11] = 1
    dt = 1e-4
    print("[low_mach] species corrector=%d cells=%d" % (corrector, n), flush=True)
    dev = lambda a: [gh.field(a[s]) for s in range(S)]  # noqa: E731
    Ya_d, Yb_d, rhs_d, Yd_d, Yc_d = dev(Y_a), dev(Y_b), dev(rhs), dev(Yd), dev(Yc)
    ra_d, rb_d = gh.field(rho_a), gh.field(rho_b)
    ro_d = gh.field(fill=7.0)
    Yo_d = [gh.field(fill=7.0) for _ in range(S)]
    bc_d = gh.field(bc, dtype=gh.torch.int8)
    ptr = lambda L: [x.data_ptr() for x in L]  # noqa: E731
    engine.low_mach_species_dev(gh.grid, dt, corrector, ra_d.data_ptr(), ptr(Ya_d),
                                ro_d.data_ptr(), ptr(Yo_d), rho_b=rb_d.data_ptr(), Y_b=ptr(Yb_d),
                                rhs=ptr(rhs_d), Y_prod_diff=ptr(Yd_d), Y_prod_chem=ptr(Yc_d),
                                bc=bc_d.data_ptr())
    gh.torch.cuda.synchronize()
    ro, yo = oracle.low_mach_species(dt, corrector, rho_a, Y_a, rho_b, Y_b, rhs=rhs,
                                     Y_prod_diff=Yd, Y_prod_chem=Yc, bc=bc,
                                     rho_out=np.full(n, 7.0), Y_out=np.full((S, n), 7.0))
    assert np.array_equal(_bits(gh.inner(ro_d)), _bits(ro))
    for s in range(S):
        assert np.array_equal(_bits(gh.inner(Yo_d[s])), _bits(yo[s])), s
    assert gh.inner(ro_d)[5] == 7.0
    # switched-off sources and in-place update (outputs alias rho_a / Y_a)
    engine.low_mach_species_dev(gh.grid, dt, False, ra_d.data_ptr(), ptr(Ya_d), ra_d.data_ptr(),
                                ptr(Ya_d), Y_prod_chem=ptr(Yc_d))
    gh.torch.cuda.synchronize()
    ro2, yo2 = oracle.low_mach_species(dt, False, rho_a, Y_a, Y_prod_chem=Yc, rho_out=rho_a,
                                       Y_out=Y_a)
    assert np.array_equal(_bits(gh.inner(ra_d)), _bits(ro2))
    assert np.array_equal(_bits(np.stack([gh.inner(y) for y in Ya_d])), _bits(yo2))
    _report("species", corrector=corrector, cells=n, marked=int(bc.sum()))


def test_chemistry_then_fds_coupling_chain(engine, oracle, mech):
    """The FDS order on one field: FP64 chemistry step, reactive divergence
    (with the EOS's mixture molar mass), species predictor with Y_prod_chem,
    equation of state (predictor); every stage against the oracle chain."""
    import torch
    gh = Ghosted()
    n = gh.n
    st = inputs.to_numpy(inputs.hot_kernels_3d(mech, SHAPE, seed=13, n_kernels=2,
                                               r_range=(2.0, 5.0)))
    rho, Y, T, p = st["rho"], np.ascontiguousarray(st["Y"]), st["T"], st["p"]
    bc = np.zeros(n, dtype=np.int8)
    bc[::13] = 1
    dt = 1e-7
    print("[low_mach] chain cells=%d" % n, flush=True)
    ptr = lambda L: [x.data_ptr() for x in L]  # noqa: E731
    rho_d, T_d, bc_d = gh.field(rho), gh.field(T), gh.field(bc, dtype=torch.int8)
    Y_d = [gh.field(Y[s]) for s in range(S)]
    Yp_d = [gh.field() for _ in range(S)]
    Yp2_d = [gh.field() for _ in range(S)]
    Ef_d = gh.field()
    engine.step_dev(gh.grid, dt, T_d.data_ptr(), rho_d.data_ptr(), ptr(Y_d), ptr(Yp_d),
                    ptr(Yp2_d), Ef_d.data_ptr(), bc=bc_d.data_ptr())
    o = oracle.chem_cells(dt, T, rho, Y, bc=bc)
    Yp = np.stack([gh.inner(y) for y in Yp_d])
    dY = np.abs(Yp - o["Y_prod"]) * dt / rho
    assert dY.max() <= 1e-9                              # the FP64 step's parity bar
    # mixture molar mass from the EOS at the current pressure (no pressure change)
    W = np.asarray(mech.molar_masses)
    ps_d, po_d, mmm_d, T2_d = gh.field(p), gh.field(), gh.field(), gh.field()
    dp_d = gh.field()
    engine.low_mach_state_equation_dev(gh.grid, 0.0, True, rho_d.data_ptr(), ptr(Y_d),
                                       dp_d.data_ptr(), ps_d.data_ptr(), po_d.data_ptr(),
                                       T2_d.data_ptr(), mmm_d.data_ptr(), bc=bc_d.data_ptr())
    oe = oracle.low_mach_eos(0.0, True, rho, Y, p, np.zeros(n), np.zeros(n), bc=bc)
    assert np.array_equal(_bits(gh.inner(mmm_d)[bc == 0]), _bits(oe["mix_mol_mass"][bc == 0]))
    T2 = gh.inner(T2_d)
    # reactive divergence from the device chemistry outputs, compared with the
    # oracle on the same inputs
    div_d = gh.field(fill=3.0)
    engine.low_mach_chem_divergence_dev(gh.grid, rho_d.data_ptr(), T2_d.data_ptr(), ptr(Y_d),
                                        mmm_d.data_ptr(), Ef_d.data_ptr(), ptr(Yp_d),
                                        div_d.data_ptr(), bc=bc_d.data_ptr())
    Ef = gh.inner(Ef_d)
    mmm = gh.inner(mmm_d)
    od = oracle.low_mach_chem_divergence(rho, T2, Y, mmm, Ef, Yp, bc=bc, div_v=np.full(n, 3.0))
    assert np.array_equal(_bits(gh.inner(div_d)), _bits(od))
    # species predictor with the chemistry source
    ri_d = gh.field()
    Yi_d = [gh.field() for _ in range(S)]
    engine.low_mach_species_dev(gh.grid, dt, False, rho_d.data_ptr(), ptr(Y_d), ri_d.data_ptr(),
                                ptr(Yi_d), Y_prod_chem=ptr(Yp_d), bc=bc_d.data_ptr())
    ro, yo = oracle.low_mach_species(dt, False, rho, Y, Y_prod_chem=Yp, bc=bc)
    assert np.array_equal(_bits(gh.inner(ri_d)), _bits(ro))
    assert np.array_equal(_bits(np.stack([gh.inner(y) for y in Yi_d])), _bits(yo))
    # the chemistry source conserves mass up to the shipped molar masses, which
    # are not element-consistent (2 W_H != W_H2): 7.6e-8 relative on this field
    live = bc == 0
    assert np.allclose(ro[live], rho[live], rtol=1e-6, atol=0)
    _report("chain", max_dY=float(dY.max()), max_div=float(np.abs(od[live]).max()))
