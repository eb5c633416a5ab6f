"""Pins of the oracle's thermochemistry, kinetics and split-step closure.

Each check is independent of the oracle's own code: published thermochemical
tables, textbook formulas evaluated in 50-digit mpmath from the raw mechanism
arrays, conservation laws, and scipy's stiff integrators.
"""
import mpmath as mp
import numpy as np
import pytest

from nrg_amd import inputs
from nrg_amd.mechanism import P_ATM, R_GAS_J, RT_THIRD_BODY, RT_TROE, T_REF, element_matrix

mp.mp.dps = 50

# NIST-JANAF (Chase 1998) / Active Thermochemical Tables values at 298.15 K:
# formation enthalpy kJ/mol, entropy J/mol/K, cp J/mol/K.  The tolerances allow
# for the spread between table editions (OH, HO2, H2O2) and the polynomial fit;
# a wrong coefficient, power of T or temperature range is off by far more.
JANAF_298 = {
    "H2": (0.0, 130.68, 28.84), "O2": (0.0, 205.15, 29.38), "N2": (0.0, 191.61, 29.12),
    "O": (249.18, 161.06, 21.91), "H": (218.00, 114.72, 20.79), "OH": (37.3, 183.74, 29.89),
    "H2O": (-241.83, 188.83, 33.59), "HO2": (12.0, 229.0, 34.9), "H2O2": (-136.1, 233.0, 43.1),
}
TOL_298 = {"H": 1.0, "S": 2.0, "cp": 1.0}
TOL_H_WIDE = {"OH": 1.8, "HO2": 2.5}

# JANAF cp (J/mol/K) and H(T) - H(298.15) (kJ/mol) at 1000 K and 2000 K
JANAF_HIGH = {
    ("H2", 1000): (30.20, 20.68), ("H2", 2000): (34.28, 52.95),
    ("O2", 1000): (34.87, 22.70), ("O2", 2000): (37.74, 59.18),
    ("N2", 1000): (32.70, 21.46), ("N2", 2000): (36.00, 56.14),
    ("H2O", 1000): (41.27, 26.00), ("H2O", 2000): (51.18, 72.79),
}


def test_thermo_at_298_matches_janaf(mech, oracle):
    for s, nm in enumerate(mech.species_names):
        h, S, cp = JANAF_298[nm]
        assert oracle.h(T_REF, s) / 1e3 == pytest.approx(h, abs=TOL_H_WIDE.get(nm, TOL_298["H"])), nm
        assert oracle.s(T_REF, s) == pytest.approx(S, abs=TOL_298["S"]), nm
        assert oracle.cp(T_REF, s) == pytest.approx(cp, abs=TOL_298["cp"]), nm


def test_thermo_high_temperature_branch_matches_janaf(mech, oracle):
    for (nm, T), (cp, dh) in JANAF_HIGH.items():
        s = mech.species_names.index(nm)
        assert oracle.cp(T, s) == pytest.approx(cp, abs=0.3), (nm, T)
        assert (oracle.h(T, s) - oracle.h(T_REF, s)) / 1e3 == pytest.approx(dh, abs=0.25), (nm, T)


def test_thermo_derivative_identities(mech, oracle):
    """cp = dh/dT, cp/T = ds/dT and e = h - R T on both polynomial ranges."""
    d = 1e-3
    for s in range(mech.species_number):
        for T in (400.0, 800.0, 1500.0, 2500.0):
            dh = (oracle.h(T + d, s) - oracle.h(T - d, s)) / (2 * d)
            ds = (oracle.s(T + d, s) - oracle.s(T - d, s)) / (2 * d)
            cp = oracle.cp(T, s)
            assert dh == pytest.approx(cp, rel=1e-6, abs=1e-4), (s, T)
            assert ds == pytest.approx(cp / T, rel=1e-6), (s, T)
            assert oracle.e(T, s) == pytest.approx(oracle.h(T, s) - R_GAS_J * T, abs=1e-9), (s, T)


def test_thermo_continuous_at_range_switch(mech, oracle):
    lo = np.nextafter(1000.0, 0.0)
    for s in range(mech.species_number):
        assert oracle.h(lo, s) == pytest.approx(oracle.h(1000.0, s), abs=40.0)
        assert oracle.cp(lo, s) == pytest.approx(oracle.cp(1000.0, s), rel=2e-3)
        assert oracle.s(lo, s) == pytest.approx(oracle.s(1000.0, s), abs=0.05)


# ---- kinetics ------------------------------------------------------------------------

def _nasa(mech, s, T):
    """h/(R T) and s/R of species s from the NASA 7-term polynomials (Gordon &
    McBride, NASA RP-1311), high set at T >= 1000 K, in mpmath."""
    a = [mp.mpf(float(x)) / mp.mpf(R_GAS_J) for x in mech.a_coeffs[s, 1 if T >= 1000 else 0]]
    T = mp.mpf(T)
    h = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
    S = a[0] * mp.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
    return h, S


def _sides(mech, r):
    """Species (0-based, M excluded) of each side, with repeats; and whether M is a term."""
    S = mech.species_number
    out = []
    for side in (0, 1):
        n = mech.chemical_coeffs[side, r, 0]
        out.append([int(sp) - 1 for sp in mech.chemical_coeffs[side, r, 1:n + 1] if sp != S + 1])
    has_m = (S + 1) in mech.chemical_coeffs[0, r, 1:]
    return out[0], out[1], has_m


def _textbook_rates(mech, T, Mc):
    """k_f: modified Arrhenius; fall-off (+M): Troe form (Gilbert, Luther & Troe
    1983, F_cent from alpha, T***, T*, T**); k_r = k_f / K_c with
    K_c = exp(dS/R - dH/RT) (p_atm / R T)^dnu, dnu over species only."""
    R = mp.mpf(R_GAS_J)
    Tm = mp.mpf(T)
    hs = [_nasa(mech, s, T) for s in range(mech.species_number)]
    kf, kr = [], []
    for r in range(mech.reactions_number):
        arr = lambda A, b, E: mp.mpf(float(A)) * Tm ** mp.mpf(float(b)) * mp.exp(-mp.mpf(float(E)) / (R * Tm))
        kinf = arr(mech.A[r], mech.beta[r], mech.E_act[r])
        if mech.reactions_type[r] == RT_TROE and Mc[r] == 0.0:
            k = mp.mpf(0)          # low-pressure limit k0 [M] with no collider (+AR, +HE)
        elif mech.reactions_type[r] == RT_TROE:
            k0 = arr(mech.A_low[r], mech.beta_low[r], mech.E_act_low[r])
            pr = k0 * mp.mpf(float(Mc[r])) / kinf
            al, t3, t1, t2 = [mp.mpf(float(x)) for x in mech.Troe_coeffs[:, r]]
            fc = (1 - al) * mp.exp(-Tm / t3) + al * mp.exp(-Tm / t1) + mp.exp(-t2 / Tm)
            c = -mp.mpf("0.4") - mp.mpf("0.67") * mp.log10(fc)
            n = mp.mpf("0.75") - mp.mpf("1.27") * mp.log10(fc)
            x = (mp.log10(pr) + c) / (n - mp.mpf("0.14") * (mp.log10(pr) + c))
            k = kinf * pr / (1 + pr) * mp.power(10, mp.log10(fc) / (1 + x * x))
        else:
            k = kinf
        left, right, _ = _sides(mech, r)
        dH = sum(hs[s][0] for s in right) - sum(hs[s][0] for s in left)
        dS = sum(hs[s][1] for s in right) - sum(hs[s][1] for s in left)
        dnu = len(right) - len(left)
        Kc = mp.exp(dS - dH) * (mp.mpf(P_ATM) / (R * Tm)) ** dnu
        kf.append(k)
        kr.append(k / Kc)
    return kf, kr


def _cells(mech, n, seed):
    st = inputs.to_numpy(inputs.ignition_mix(mech, n, seed=seed, frac_window=0.5))
    return st["T"], st["rho"], st["Y"]


@pytest.mark.parametrize("T", [700.0, 999.0, 1000.0, 1350.0, 2200.0, 3100.0])
def test_rate_constants_match_textbook(mech, oracle, T):
    _, rho, Y = _cells(mech, 4, seed=5)
    Mc = oracle.mixture_concentration(np.ascontiguousarray(Y[:, 1]), rho[1])
    kf, kr = oracle.rate_constants(T, Mc)
    kf_t, kr_t = _textbook_rates(mech, T, Mc)
    for r in range(mech.reactions_number):
        assert kf[r] == pytest.approx(float(kf_t[r]), rel=2e-12), ("kf", r + 1, T)
        assert kr[r] == pytest.approx(float(kr_t[r]), rel=5e-12), ("kr", r + 1, T)


def test_third_body_concentration(mech, oracle):
    """[M]_r = sum_s eff(s, r) rho Y_s / W_s."""
    _, rho, Y = _cells(mech, 3, seed=9)
    y = np.ascontiguousarray(Y[:, 2])
    C = rho[2] * y / mech.molar_masses
    Mc = oracle.mixture_concentration(y, rho[2])
    np.testing.assert_allclose(Mc, mech.enhanced_efficiencies.T @ C, rtol=1e-14)


def _mass_action(mech, kf, kr, Mc, Z):
    """omega_s = sum_r nu_sr (k_f prod C^nu' - k_r prod C^nu''); a third body
    that is a term (X+M) multiplies by [M]_r."""
    w = [mp.mpf(0)] * mech.species_number
    for r in range(mech.reactions_number):
        left, right, has_m = _sides(mech, r)
        f = mp.mpf(float(kf[r]))
        b = mp.mpf(float(kr[r]))
        for s in left:
            f *= mp.mpf(float(Z[s]))
        for s in right:
            b *= mp.mpf(float(Z[s]))
        if has_m:
            f *= mp.mpf(float(Mc[r]))
            b *= mp.mpf(float(Mc[r]))
        for s in left:
            w[s] -= f - b
        for s in right:
            w[s] += f - b
    return w


def test_rhs_is_mass_action_and_conserves_elements(mech, oracle):
    T, rho, Y = _cells(mech, 8, seed=13)
    E = element_matrix(mech.species_names)
    for i in range(8):
        y = np.ascontiguousarray(Y[:, i])
        Mc = oracle.mixture_concentration(y, rho[i])
        kf, kr = oracle.rate_constants(T[i], Mc)
        Z = rho[i] * y / mech.molar_masses
        Z[3:] += 1e-3 * Z[:3].max()            # radicals present: every reaction active
        w = oracle.rhs(kf, kr, Mc, Z)
        wt = _mass_action(mech, kf, kr, Mc, Z)
        scale = max(abs(float(x)) for x in wt)
        for s in range(mech.species_number):
            assert abs(w[s] - float(wt[s])) <= 1e-13 * scale, (i, s)
        # atoms are neither made nor destroyed
        assert np.all(np.abs(E @ w) <= 1e-12 * np.abs(E) @ np.abs(w))


def test_chemistry_step_agrees_with_scipy_radau(mech, oracle):
    """One cell's DDRIV3 integration against scipy Radau at rtol 1e-12 on the
    same frozen-T right-hand side; DDRIV3 runs at EPS = 1e-9."""
    from scipy.integrate import solve_ivp
    T, rho, Y = _cells(mech, 16, seed=21)
    hot = [i for i in range(16) if 1000.0 <= T[i] <= 1300.0][:3]
    assert hot
    for dt in (1e-7, 1e-6):
        for i in hot:
            y = np.ascontiguousarray(Y[:, i])
            c_out, c_in, out = oracle.ddriv3_cell(T[i], rho[i], y, dt)
            assert out.nstate == 2
            Mc = oracle.mixture_concentration(y, rho[i])
            kf, kr = oracle.rate_constants(T[i], Mc)
            sol = solve_ivp(lambda t, z: oracle.rhs(kf, kr, Mc, z), (0.0, dt), c_in,
                            method="Radau", rtol=1e-12, atol=1e-16)
            ref = sol.y[:, -1]
            err = np.abs(c_out - ref) / np.maximum(np.abs(ref), 1e-4)
            assert err.max() < 1e-6, (i, dt, err)


def test_chemistry_step_gating_and_mass_conservation(mech, oracle):
    T, rho, Y = _cells(mech, 64, seed=4)
    T = T.copy()
    T[:4] = [304.999, 305.0, 200.0, 1200.0]
    bc = np.zeros(64, dtype=np.int8)
    bc[3] = 1
    sentinel = np.full((mech.species_number, 64), 7.0)
    o = oracle.chem_cells(1e-7, T, rho, Y, bc=bc, Y_prod2=sentinel)
    gated = (T < 305.0) | (bc != 0)
    assert np.all(o["Y_prod"][:, gated] == 0) and np.all(o["E_f_prod"][gated] == 0)
    assert np.all(o["Y_prod2"][:, gated] == 7.0)           # left untouched
    assert not gated[1] and np.all(o["Y_prod2"][:, 1] != 7.0)
    # Y_prod2 is the concentration change C_out - C_in: it conserves atoms to
    # the rounding of C_in
    E = element_matrix(mech.species_names)
    dC = o["Y_prod2"][:, ~gated]
    C_in = (rho * Y / mech.molar_masses[:, None])[:, ~gated]
    assert np.all(np.abs(E @ dC) <= 1e-13 * (E @ C_in))


# molar_masses.dat is rounded per species (2 W_H = 2.0160e-3 but W_H2 = 2.0159e-3),
# so mass and energy balances hold only to ~1e-5 with it.  The exact invariants
# below use the same mechanism with element-consistent molar masses.
ATOMIC_MASS = np.array([1.00794e-3, 15.9994e-3, 14.0067e-3])      # H, O, N (kg/mol)


@pytest.fixture(scope="module")
def consistent(mech):
    import dataclasses

    from oracle import Oracle
    m = dataclasses.replace(mech, molar_masses=ATOMIC_MASS @ element_matrix(mech.species_names))
    return m, Oracle(m)


def test_chemistry_step_conserves_mass(consistent):
    m, orc = consistent
    T, rho, Y = _cells(m, 48, seed=4)
    dt = 1e-6
    o = orc.chem_cells(dt, T, rho, Y)
    yp = o["Y_prod"]
    assert np.abs(yp).max() * dt > 1e-3 * rho.min()          # something reacted
    # sum_s W_s (C_out - C_in) / dt = 0 to the rounding of rho Y_s
    assert np.all(np.abs(yp.sum(axis=0)) <= 1e-14 * rho / dt)


def test_split_step_conserves_internal_energy(consistent):
    """Chemistry + closure at fixed density is an adiabatic constant-volume
    reactor: U = sum_s C_s e_s(T) (formation energies included) is unchanged,
    and p = rho R T / W_mix."""
    m, orc = consistent
    T, rho, Y = _cells(m, 32, seed=8)
    E_f = orc.energy_from_temperature(T, Y)
    dt = 1e-6
    o = orc.chem_cells(dt, T, rho, Y)
    E2, Y2, T2, p2 = orc.closure_cells(dt, rho, o["E_f_prod"], o["Y_prod"], E_f, Y, T)
    S = m.species_number
    moved = 0
    for i in range(32):
        U0 = sum(rho[i] * Y[s, i] / m.molar_masses[s] * orc.e(T[i], s) for s in range(S))
        U1 = sum(rho[i] * Y2[s, i] / m.molar_masses[s] * orc.e(T2[i], s) for s in range(S))
        cv = sum(rho[i] * Y2[s, i] / m.molar_masses[s] * (orc.cp(T2[i], s) - R_GAS_J)
                 for s in range(S))
        # the e -> T Newton stops at |dT| < 1e-8 K
        assert abs(U1 - U0) <= 1e-11 * abs(U0) + 1e-7 * cv, (i, U1 - U0)
        moved += abs(T2[i] - T[i]) > 1e-3
    assert moved > 0                       # some cells actually released heat
    assert np.allclose(Y2.sum(axis=0), 1.0, rtol=0, atol=1e-14)
    Wmix = 1.0 / (Y2 / m.molar_masses[:, None]).sum(axis=0)
    np.testing.assert_allclose(p2, rho * R_GAS_J * T2 / Wmix, rtol=1e-14)


def test_temperature_inversion_against_bisection(mech, oracle):
    T, rho, Y = _cells(mech, 12, seed=30)
    for i in range(12):
        y = Y[:, i]
        X = (y / mech.molar_masses) / (y / mech.molar_masses).sum()
        href = sum(X[s] * oracle.h(T_REF, s) for s in range(mech.species_number))
        e_mix = lambda t: sum(X[s] * oracle.e(t, s) for s in range(mech.species_number)) - href
        target = e_mix(T[i])
        lo, hi = 200.0, 6000.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if e_mix(mid) < target else (lo, mid)
        got = oracle.calculate_temperature(T[i] * 0.8 + 100.0, target, np.ascontiguousarray(X))
        assert got == pytest.approx(0.5 * (lo + hi), abs=1e-6), i


def test_energy_temperature_round_trip(mech, oracle):
    T, rho, Y = _cells(mech, 16, seed=31)
    E = oracle.energy_from_temperature(T, Y)
    zero_y = np.zeros_like(Y)
    E2, Y2, T2, _ = oracle.closure_cells(1e-7, rho, np.zeros_like(T), zero_y, E, Y,
                                         T * 0.9 + 50.0)
    np.testing.assert_allclose(T2, T, rtol=0, atol=1e-6)
    np.testing.assert_allclose(Y2, Y / Y.sum(axis=0), rtol=0, atol=1e-15)
