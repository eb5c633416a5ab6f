"""The learned step's gate and FP64 fallback (include/nrg_surrogate.h) on
networks whose answer is known exactly: with all weights zero the network
returns out_shift, so every gate decision and every projected increment can
be checked, and rejected cells must reproduce nrg_chem_step_dev bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from nrg_amd import capi, inputs
from nrg_amd.mechanism import element_matrix

pytestmark = pytest.mark.gpu
S = 9


def _report(name, **kw):
    print("[surrogate] %s " % name + " ".join("%s=%s" % (k, v) for k, v in kw.items()), flush=True)


@pytest.fixture(scope="module")
def zero_net():
    from nrg_amd.surrogate import MLP
    net = MLP(S + 3, S + 1)
    net.set_params(np.zeros(net.num_params, dtype=np.float32))
    yield net
    net.close()


def _norm(shift):
    from nrg_amd.surrogate import Normalizer
    return Normalizer(np.zeros(S + 3), np.ones(S + 3), shift, np.ones(S + 1))


def _gate(mech, **kw):
    from nrg_amd.surrogate import gate_struct
    base = dict(t_range=(0.0, 1e4), p_range=(0.0, 1e9), dt_range=(1e-12, 1.0), elem_tol=1e9,
                elem_floor=1e9, y_neg_tol=1.0, energy_tol=1e9)
    base.update(kw)
    return gate_struct(mech, **base)


def _state(mech, n=3000, seed=5):
    import torch
    st = inputs.ignition_mix(mech, n, seed=seed, frac_window=0.3, device="cuda:0")
    st["T"][:7] = 250.0                     # a few cold (gated) cells
    bc = torch.zeros(n, dtype=torch.int8, device="cuda:0")
    bc[7:12] = 1                            # and boundary cells
    return st, bc


def _run(engine, net, norm, gate, st, bc, dt):
    import torch
    n = st["T"].shape[0]
    Yp = torch.zeros((S, n), dtype=torch.float64, device="cuda:0")
    Yp2 = torch.full((S, n), 7.0, dtype=torch.float64, device="cuda:0")
    Ef = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    status = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
    Y = st["Y"].contiguous()
    stats = capi.NrgSurrogateStats()
    torch.cuda.synchronize()
    rc = engine.lib.nrg_chem_surrogate_step_dev(
        engine.h, net.h if net is not None else None,
        C.byref(norm.struct()) if norm is not None else None, C.byref(gate),
        C.byref(capi.flat_grid(n)), dt,
        st["T"].data_ptr(), st["p"].data_ptr(), st["rho"].data_ptr(),
        capi.ptr_array([Y[s].data_ptr() for s in range(S)]), bc.data_ptr(),
        capi.ptr_array([Yp[s].data_ptr() for s in range(S)]),
        capi.ptr_array([Yp2[s].data_ptr() for s in range(S)]), Ef.data_ptr(), status.data_ptr(),
        C.byref(stats))
    capi.check(engine.lib, rc)
    return dict(Y_prod=Yp.cpu().numpy(), Y_prod2=Yp2.cpu().numpy(), E_f_prod=Ef.cpu().numpy(),
                status=status.cpu().numpy(), stats=stats)


def _fp64(engine, st, bc, dt):
    import torch
    n = st["T"].shape[0]
    Yp = torch.zeros((S, n), dtype=torch.float64, device="cuda:0")
    Yp2 = torch.full((S, n), 7.0, dtype=torch.float64, device="cuda:0")
    Ef = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    status = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
    Y = st["Y"].contiguous()
    engine.step_dev(capi.flat_grid(n), dt, st["T"].data_ptr(), st["rho"].data_ptr(),
                    [Y[s].data_ptr() for s in range(S)], [Yp[s].data_ptr() for s in range(S)],
                    [Yp2[s].data_ptr() for s in range(S)], Ef.data_ptr(), bc=bc.data_ptr(),
                    status=status.data_ptr())
    torch.cuda.synchronize()
    return dict(Y_prod=Yp.cpu().numpy(), Y_prod2=Yp2.cpu().numpy(), E_f_prod=Ef.cpu().numpy(),
                status=status.cpu().numpy())


def test_accept_all_zero_increment(engine, mech, zero_net):
    st, bc = _state(mech)
    o = _run(engine, zero_net, _norm(np.zeros(S + 1)), _gate(mech), st, bc, 1e-7)
    T = st["T"].cpu().numpy()
    gated = (T < 305.0) | (bc.cpu().numpy() != 0)
    s = o["stats"]
    _report("accept_all", accepted=s.cells_accepted, gated=s.cells_gated, fallback=s.cells_fallback)
    assert s.cells_gated == gated.sum() and s.cells_accepted == (~gated).sum()
    assert s.cells_fallback == 0
    assert np.all(o["status"][~gated] == capi.NRG_CELL_SURROGATE)
    assert np.all(o["status"][gated] == capi.NRG_CELL_GATED)
    assert np.all(o["Y_prod"] == 0) and np.all(o["E_f_prod"] == 0)
    assert np.all(o["Y_prod2"][:, ~gated] == 0) and np.all(o["Y_prod2"][:, gated] == 7.0)


def test_reject_all_is_the_fp64_step(engine, mech, zero_net):
    st, bc = _state(mech, seed=8)
    o = _run(engine, zero_net, _norm(np.zeros(S + 1)), _gate(mech, energy_tol=-1.0), st, bc, 1e-7)
    f = _fp64(engine, st, bc, 1e-7)
    s = o["stats"]
    _report("reject_all", fallback=s.cells_fallback, reject_energy=s.reject_energy,
            nfe=s.fallback.nfe_sum)
    assert s.cells_accepted == 0 and s.cells_fallback == s.reject_energy > 0
    for k in ("Y_prod", "Y_prod2", "E_f_prod", "status"):
        assert np.array_equal(o[k], f[k]), k


def test_domain_split(engine, mech, zero_net):
    st, bc = _state(mech, seed=9)
    o = _run(engine, zero_net, _norm(np.zeros(S + 1)), _gate(mech, t_range=(1500.0, 1e4)), st, bc,
             1e-7)
    f = _fp64(engine, st, bc, 1e-7)
    T = st["T"].cpu().numpy()
    gated = (T < 305.0) | (bc.cpu().numpy() != 0)
    rej = ~gated & (T < 1500.0)
    acc = ~gated & (T >= 1500.0)
    s = o["stats"]
    assert s.reject_domain == rej.sum() and s.cells_accepted == acc.sum()
    assert np.array_equal(o["Y_prod"][:, rej], f["Y_prod"][:, rej])
    assert np.array_equal(o["status"][rej], f["status"][rej])
    assert np.all(o["Y_prod"][:, acc] == 0)


def _reaction_increment(mech, a, off=1e-3):
    """dY of a mol/kg of H2 + 1/2 O2 -> H2O, with H2O overshooting by `off` (not conserving)."""
    W = mech.molar_masses
    d = np.zeros(S + 1)
    d[1 + 0] = -a * W[0]
    d[1 + 1] = -0.5 * a * W[1]
    d[1 + 6] = a * W[6] * (1.0 + off)
    return d


def test_projection_conserves_elements(engine, mech, zero_net):
    st, bc = _state(mech, seed=10)
    shift = _reaction_increment(mech, 1e-3)
    o = _run(engine, zero_net, _norm(shift), _gate(mech), st, bc, 1e-7)
    acc = o["status"] == capi.NRG_CELL_SURROGATE
    assert acc.sum() > 100
    E = element_matrix(mech.species_names)
    dC = o["Y_prod2"][:, acc]                      # mol/m3
    res = np.abs(E @ dC)
    scale = np.abs(E) @ np.abs(dC)
    # the N row of E is zero for every species the increment moves (N2 is
    # inert), so its scale is 0; report the rows that carry a balance
    live = scale > 0
    _report("projection", max_rel_residual=float((res[live] / scale[live]).max()),
            rows=int(live.any(axis=1).sum()))
    assert np.all(res <= 1e-12 * scale)
    # the correction is small: the raw increment is off by 1e-3 in H2O only
    rho = st["rho"].cpu().numpy()[acc]
    dY = o["Y_prod"][:, acc] * 1e-7 / rho
    raw = shift[1:][:, None]
    assert np.all(np.abs(dY - raw) <= 2e-3 * np.abs(raw).max())
    # energy released by the accepted chemistry: E_f_prod = -sum h_s(T_ref) dC_s / dt > 0
    assert np.all(o["E_f_prod"][acc] > 0)


def test_element_and_energy_gates_reject(engine, mech, zero_net):
    st, bc = _state(mech, seed=12)
    shift = _reaction_increment(mech, 1e-3, off=0.2)     # 20 % H2O overshoot
    o = _run(engine, zero_net, _norm(shift), _gate(mech, elem_tol=0.05, elem_floor=0.0), st, bc,
             1e-7)
    s = o["stats"]
    live = s.cells_inner - s.cells_gated
    assert s.reject_elements == live and s.cells_accepted == 0
    # conserving increment but a network dT of 0 while heat is released
    o = _run(engine, zero_net, _norm(_reaction_increment(mech, 1e-3, off=0.0)),
             _gate(mech, energy_tol=1e-3), st, bc, 1e-7)
    s = o["stats"]
    _report("gates", reject_energy=s.reject_energy, live=live)
    assert s.reject_energy == live


def test_fallback_in_stiffness_order_is_bitwise_identical(engine, mech, zero_net):
    """With NRG_ORDER_STIFFNESS and a step-count history for the grid, the
    fallback cells are handed out most expensive first; outputs and work
    counts equal those of the unordered fallback bit for bit, over two calls
    (the second uses the counts the first recorded)."""
    from nrg_amd.solver import ChemistryStep
    st, bc = _state(mech, seed=13)
    eng = ChemistryStep(mech, capi.default_opts(cell_order=capi.NRG_ORDER_STIFFNESS))
    try:
        _fp64(eng, st, bc, 1e-7)                       # history for this grid
        gate = _gate(mech, t_range=(1500.0, 1e4))
        for call in range(2):
            a = _run(engine, zero_net, _norm(np.zeros(S + 1)), gate, st, bc, 1e-7)
            b = _run(eng, zero_net, _norm(np.zeros(S + 1)), gate, st, bc, 1e-7)
            _report("fallback_order", call=call, fallback=b["stats"].cells_fallback,
                    nfe=b["stats"].fallback.nfe_sum)
            assert b["stats"].cells_fallback > 0
            for k in ("Y_prod", "Y_prod2", "E_f_prod", "status"):
                assert np.array_equal(a[k], b[k]), (call, k)
            assert a["stats"].fallback.nfe_sum == b["stats"].fallback.nfe_sum
    finally:
        eng.close()


def test_rate_consistency_rejects_what_the_explicit_step_contradicts(engine, oracle, mech, zero_net):
    """rate_tol: with a network that predicts no change, a cell is kept only if
    the explicit estimate of its step, omega(C, T) dt W_s / rho from the oracle's
    rate constants and mass action, stays within rate_tol in every species;
    the rejected cells are the FP64 step's, bit for bit."""
    st, bc = _state(mech, seed=12)
    tol, dt = 1e-6, 1e-7
    o = _run(engine, zero_net, _norm(np.zeros(S + 1)), _gate(mech, rate_tol=tol), st, bc, dt)
    f = _fp64(engine, st, bc, dt)
    T, rho, Y = st["T"].cpu().numpy(), st["rho"].cpu().numpy(), st["Y"].cpu().numpy()
    live = (T >= 305.0) & (bc.cpu().numpy() == 0)
    W = np.asarray(mech.molar_masses)
    est = np.zeros(T.size)
    for q in np.flatnonzero(live):
        Mc = oracle.mixture_concentration(Y[:, q], rho[q])
        kf, kr = oracle.rate_constants(T[q], Mc)
        w = oracle.rhs(kf, kr, Mc, Y[:, q] * rho[q] / W)
        est[q] = np.abs(w * dt * W / rho[q]).max()
    s = o["stats"]
    rej = live & (o["status"] != capi.NRG_CELL_SURROGATE)
    _report("rate_consistency", rejected=int(rej.sum()), live=int(live.sum()),
            reject_rate=s.reject_rate)
    assert s.reject_rate == rej.sum() == s.cells_fallback and 0 < rej.sum() < live.sum()
    assert np.all(rej[live & (est > 2.0 * tol)]) and not np.any(rej[live & (est < 0.5 * tol)])
    for k in ("Y_prod", "Y_prod2", "E_f_prod", "status"):
        assert np.array_equal(o[k][..., rej], f[k][..., rej]), k
    assert np.all(o["Y_prod"][:, live & ~rej] == 0)


def _explicit_estimate(oracle, mech, st, dt):
    """Per cell: max_s |omega_s(C, T) dt W_s / rho| and the temperature change the
    estimate implies at frozen cv, -sum_s e_s(T) dY_s / W_s / cv, from the oracle."""
    T, rho, Y = st["T"].cpu().numpy(), st["rho"].cpu().numpy(), st["Y"].cpu().numpy()
    W = np.asarray(mech.molar_masses)
    R = 8.31446261815324
    est, dT = np.zeros(T.size), np.zeros(T.size)
    for q in range(T.size):
        if T[q] < 305.0:
            continue
        Mc = oracle.mixture_concentration(Y[:, q], rho[q])
        kf, kr = oracle.rate_constants(T[q], Mc)
        dY = oracle.rhs(kf, kr, Mc, Y[:, q] * rho[q] / W) * dt * W / rho[q]
        est[q] = np.abs(dY).max()
        cv = sum(Y[s, q] / W[s] * (oracle.cp(T[q], s) - R) for s in range(S))
        dT[q] = -sum(oracle.e(T[q], s) * dY[s] / W[s] for s in range(S)) / cv
    return est, dT


def test_rate_screen_without_network_is_the_freeze_test_plus_fp64(engine, oracle, mech):
    """gate.screen = ONLY, the method without the network: a live cell whose
    explicit estimate stays within dY_freeze in every species and dT_freeze in
    temperature gets the zero increment; every other live cell is the FP64
    step's, bit for bit; the network is never run."""
    st, bc = _state(mech, seed=21)
    dt, ytol, ttol = 1e-7, 1e-6, 0.05
    gate = _gate(mech, dY_freeze=ytol, dT_freeze=ttol, screen=capi.NRG_SUR_SCREEN_ONLY)
    o = _run(engine, None, None, gate, st, bc, dt)
    f = _fp64(engine, st, bc, dt)
    est, dT = _explicit_estimate(oracle, mech, st, dt)
    live = (st["T"].cpu().numpy() >= 305.0) & (bc.cpu().numpy() == 0)
    frozen = live & (o["status"] == capi.NRG_CELL_SURROGATE)
    fb = live & ~frozen
    s = o["stats"]
    _report("screen_only", live=int(live.sum()), frozen=int(frozen.sum()), fallback=s.cells_fallback,
            network=s.cells_network)
    assert s.cells_network == 0 and s.cells_fallback == fb.sum() and s.cells_frozen == frozen.sum()
    assert 0 < frozen.sum() < live.sum()
    # the oracle's estimate decides, away from the thresholds
    clear_in = live & (est < 0.5 * ytol) & (np.abs(dT) < 0.5 * ttol)
    clear_out = live & ((est > 2.0 * ytol) | (np.abs(dT) > 2.0 * ttol))
    assert np.all(frozen[clear_in]) and not np.any(frozen[clear_out])
    for k in ("Y_prod", "Y_prod2", "E_f_prod", "status"):
        assert np.array_equal(o[k][..., ~frozen], f[k][..., ~frozen]), k
    assert np.all(o["Y_prod"][:, frozen] == 0) and np.all(o["Y_prod2"][:, frozen] == 0)
    assert np.all(o["E_f_prod"][frozen] == 0)


def _kernels_state(mech, n=48, seed=41):
    import torch
    st = inputs.hot_kernels_3d(mech, (n, n, n), seed=seed, device="cuda:0", n_kernels=6)
    for k in ("T", "p", "rho"):
        st[k] = st[k].reshape(-1)
    st["Y"] = st["Y"].reshape(S, -1)
    return st, torch.zeros(n ** 3, dtype=torch.int8, device="cuda:0")


@pytest.mark.parametrize("field", ["ignition", "hot_kernels"])
@pytest.mark.parametrize("screen", ["only", "rates"])
@pytest.mark.parametrize("tol", [(1e-6, 0.01), (1e-8, 1e-4)])
def test_screen_fp32_pass_keeps_the_fp64_decisions(engine, mech, zero_net, field, screen, tol):
    """The screen's first pass (FP32 exponentials with a bound on its distance
    from the FP64 estimate) decides only the cells whose side of each threshold
    the bound settles, and leaves the others to the FP64 estimate: outputs,
    statuses and counters equal those of the screen with the FP64 estimate
    deciding every live cell, on an igniting mixture and on hot kernels in fresh
    gas, at the bench's thresholds and at 100x tighter ones."""
    st, bc = _state(mech, n=20000, seed=31) if field == "ignition" else _kernels_state(mech)
    dt = 1e-7
    ytol, ttol = tol
    scr = capi.NRG_SUR_SCREEN_ONLY if screen == "only" else capi.NRG_SUR_SCREEN_RATES
    net, norm = (None, None) if screen == "only" else (zero_net, _norm(np.zeros(S + 1)))
    kw = dict(dY_freeze=ytol, dT_freeze=ttol, screen=scr, rate_tol=1e-5 if screen == "rates" else 0.0)
    a = _run(engine, net, norm, _gate(mech, **kw), st, bc, dt)
    b = _run(engine, net, norm, _gate(mech, screen_fp64=True, **kw), st, bc, dt)
    sa, sb = a["stats"], b["stats"]
    live = int(sa.cells_inner - sa.cells_gated)
    _report("screen_fp32_pass", field=field, screen=screen, dY_freeze=ytol, live=live,
            frozen=sa.cells_frozen, decided_fp64=sa.cells_screen_fp64,
            fp64_share="%.5f" % (sa.cells_screen_fp64 / max(live, 1)))
    assert sb.cells_screen_fp64 == live
    assert sa.cells_screen_fp64 < live and 0 < sa.cells_frozen < live
    for k in ("Y_prod", "Y_prod2", "E_f_prod", "status"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("cells_accepted", "cells_fallback", "cells_gated", "cells_frozen", "cells_network",
              "reject_rate"):
        assert getattr(sa, k) == getattr(sb, k), k


def test_screen_defers_a_cell_on_the_threshold(engine, oracle, mech):
    """A cell whose explicit estimate sits on dY_freeze (the oracle's value of
    it) is one the FP32 pass cannot place on either side: it goes to the FP64
    estimate, and outputs, statuses and counters equal those of the screen that
    sends every cell there.  Three cells, one at a time."""
    st, bc = _state(mech, seed=37)
    dt = 1e-7
    est, _ = _explicit_estimate(oracle, mech, st, dt)
    live = (st["T"].cpu().numpy() >= 305.0) & (bc.cpu().numpy() == 0)
    cand = np.flatnonzero(live & (est > 1e-10) & (est < 1e-3))
    assert cand.size >= 3
    for q0 in cand[np.linspace(0, cand.size - 1, 3).astype(int)]:
        kw = dict(dY_freeze=float(est[q0]), dT_freeze=1e9, screen=capi.NRG_SUR_SCREEN_ONLY)
        a = _run(engine, None, None, _gate(mech, **kw), st, bc, dt)
        b = _run(engine, None, None, _gate(mech, screen_fp64=True, **kw), st, bc, dt)
        sa, sb = a["stats"], b["stats"]
        _report("screen_defers", cell=int(q0), dY_freeze="%.6e" % est[q0], decided_fp64=sa.cells_screen_fp64,
                frozen=sa.cells_frozen)
        assert 1 <= sa.cells_screen_fp64 < int(live.sum()) == sb.cells_screen_fp64
        for k in ("Y_prod", "Y_prod2", "E_f_prod", "status"):
            assert np.array_equal(a[k], b[k]), k
        for k in ("cells_accepted", "cells_fallback", "cells_gated", "cells_frozen"):
            assert getattr(sa, k) == getattr(sb, k), k


def test_fallback_takes_its_handles_integrator_options(engine, mech):
    """The learned step on a handle created with a looser eps and the analytic
    Jacobian: the screen freezes the same cells as on the reference handle (its
    test does not involve the integrator), and every other live cell is that
    handle's own FP64 step, bit for bit."""
    from nrg_amd.solver import ChemistryStep
    st, bc = _state(mech, seed=23)
    dt = 1e-7
    gate = _gate(mech, dY_freeze=1e-6, dT_freeze=0.05, screen=capi.NRG_SUR_SCREEN_ONLY)
    loose = ChemistryStep(mech, capi.default_opts(eps=1e-6, jacobian=capi.NRG_JAC_ANALYTIC))
    try:
        a = _run(engine, None, None, gate, st, bc, dt)
        b = _run(loose, None, None, gate, st, bc, dt)
        f = _fp64(loose, st, bc, dt)
        frozen = b["status"] == capi.NRG_CELL_SURROGATE
        _report("loose_fallback", frozen=int(frozen.sum()), fallback=b["stats"].cells_fallback,
                fallback_nfe=b["stats"].fallback.nfe_sum, reference_nfe=a["stats"].fallback.nfe_sum)
        assert np.array_equal(frozen, a["status"] == capi.NRG_CELL_SURROGATE)
        assert b["stats"].cells_fallback == a["stats"].cells_fallback > 0
        for k in ("Y_prod", "Y_prod2", "E_f_prod", "status"):
            assert np.array_equal(b[k][..., ~frozen], f[k][..., ~frozen]), k
        # the looser integrator is the one in use: fewer right-hand sides, other answers
        assert b["stats"].fallback.nfe_sum < a["stats"].fallback.nfe_sum
        assert not np.array_equal(a["Y_prod"], b["Y_prod"])
    finally:
        loose.close()


def test_rate_screen_then_network_equals_network_on_every_cell(engine, mech, zero_net):
    """gate.screen = RATES runs the network only on the cells the screen does not
    freeze.  With a network that answers zero, the gate freezes a cell exactly
    when its explicit estimate is within rate_tol, and the screen (dY_freeze <=
    rate_tol) only takes cells the gate would freeze anyway: outputs, statuses
    and the FP64 fallback equal those of the network run on every cell."""
    st, bc = _state(mech, seed=22)
    dt = 1e-7
    kw = dict(dY_freeze=1e-6, dT_freeze=0.05, rate_tol=1e-5)
    a = _run(engine, zero_net, _norm(np.zeros(S + 1)),
             _gate(mech, screen=capi.NRG_SUR_SCREEN_RATES, **kw), st, bc, dt)
    b = _run(engine, zero_net, _norm(np.zeros(S + 1)),
             _gate(mech, screen=capi.NRG_SUR_SCREEN_NONE, **kw), st, bc, dt)
    live = int(a["stats"].cells_inner - a["stats"].cells_gated)
    _report("screen_rates", live=live, network=a["stats"].cells_network,
            accepted=a["stats"].cells_accepted, fallback=a["stats"].cells_fallback)
    # without the screen the network runs on every inner cell (gated ones included)
    assert b["stats"].cells_network == b["stats"].cells_inner
    assert 0 < a["stats"].cells_network < live
    for k in ("Y_prod", "Y_prod2", "E_f_prod", "status"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("cells_accepted", "cells_fallback", "cells_gated", "cells_frozen"):
        assert getattr(a["stats"], k) == getattr(b["stats"], k), k
    assert a["stats"].fallback.nfe_sum == b["stats"].fallback.nfe_sum
