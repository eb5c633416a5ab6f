"""Table-approximated chemistry mode (cm/chemical_kinetics_solver.f90:250-346,
542-685): the oracle's restatement against the closed forms of linear
interpolation and of the output assembly, and the table file format."""
import numpy as np
import pytest

from nrg_amd import chem_table as ct

S = 9


@pytest.fixture(scope="module")
def W(mech):
    return np.asarray(mech.molar_masses, dtype=np.float64)


def _table():
    T_t = np.array([320.0, 700.0, 1000.0, 1600.0, 2400.0])
    dC = np.outer(np.arange(1, S + 1) * 1e-3, np.array([0.0, 1.0, 4.0, -2.0, 0.5]))
    return T_t, dC


def test_interpolation_closed_forms(oracle, mech, W):
    T_t, dC = _table()
    T = np.array([250.0, 310.0, 700.0, 850.0, 1000.0, 1300.0, 2400.0, 3000.0, 400.0])
    Y = np.zeros((S, T.size))                 # C_in = Y = 0, so Y_prod2 = C_out exactly
    dt = 1e-7
    o = oracle.table_cells(dt, T_t, dC, T, Y, Y_prod2=np.full((S, T.size), 7.0))
    # gated (T < 305 K): Y_prod = 0, E_f_prod = 0, Y_prod2 untouched
    assert o["status"][0] == 1 and np.all(o["Y_prod"][:, 0] == 0) and o["E_f_prod"][0] == 0
    assert np.all(o["Y_prod2"][:, 0] == 7.0)
    c = o["Y_prod2"]
    assert np.all(c[:, 1] == 0.0)              # below the table: 0
    assert np.allclose(c[:, 2], dC[:, 1], rtol=0, atol=1e-18)          # a node
    assert np.allclose(c[:, 3], 0.5 * (dC[:, 1] + dC[:, 2]), rtol=1e-14)  # midpoint
    # T = 1000 sits on two intervals; the later one wins (its left node)
    assert np.array_equal(c[:, 4], dC[:, 2] * (1000.0 - 1600.0) / (1000.0 - 1600.0))
    assert np.allclose(c[:, 5], 0.5 * (dC[:, 2] + dC[:, 3]), rtol=1e-14)
    assert np.allclose(c[:, 6], dC[:, 4], rtol=1e-15)
    assert np.array_equal(c[:, 7], dC[:, 4])   # above the table: the last row
    frac = (400.0 - 320.0) / (700.0 - 320.0)
    assert np.allclose(c[:, 8], dC[:, 0] + frac * (dC[:, 1] - dC[:, 0]), rtol=1e-13, atol=1e-20)
    # outputs of an integrated cell
    live = np.arange(1, T.size)
    assert np.allclose(o["Y_prod"][:, live], c[:, live] * W[:, None] / dt, rtol=1e-15, atol=0)
    href = np.array([oracle.h(298.15, s) for s in range(S)])
    ef = -(href[:, None] * c[:, live]).sum(0) / dt
    assert np.allclose(o["E_f_prod"][live], ef, rtol=1e-12, atol=1e-9)
    assert np.all(o["status"][live] == 0)
    # Y_prod2 = C_out - C_in with C_in the mass fractions themselves
    Y1 = np.full((S, T.size), 0.05)
    o1 = oracle.table_cells(dt, T_t, dC, T, Y1)
    assert np.array_equal(o1["Y_prod2"][:, live], c[:, live] - 0.05)
    assert np.array_equal(o1["Y_prod"], o["Y_prod"])


def test_unordered_tables_follow_the_reference_rules(oracle):
    """Every bracketing interval overwrites C_out (the last one wins); with a
    table that ends below its start, T between the ends triggers both range
    rules and the second (the last row) wins."""
    Y = np.full((S, 1), 0.1)
    T_t = np.array([400.0, 900.0, 1100.0, 800.0, 1200.0])     # 1000 K: [900,1100] and [800,1200]
    dC = np.vstack([np.array([1.0, 2.0, 3.0, 4.0, 5.0])] * S)
    o = oracle.table_cells(1e-7, T_t, dC, np.array([1000.0]), Y)
    assert np.allclose(o["Y_prod2"][:, 0], 4.0 + 0.5 * (5.0 - 4.0) - 0.1, rtol=1e-15)
    T_d = np.array([1200.0, 600.0])                            # descending
    o = oracle.table_cells(1e-7, T_d, np.vstack([np.array([3.0, 7.0])] * S), np.array([900.0]), Y)
    assert np.array_equal(o["Y_prod2"][:, 0], np.full(S, 7.0 - 0.1))


def test_file_format_round_trip(tmp_path):
    T_t, dC = _table()
    dC = dC * np.pi
    f = tmp_path / "t.dat"
    names = ["H2", "O2", "N2", "O", "H", "OH", "H2O", "HO2", "H2O2"]
    ct.write_table(f, T_t, dC, names)
    lines = f.read_text().splitlines()
    assert lines[0] == "T" + "".join("  r" + n for n in names)
    assert all(len(ln) == 14 * (S + 1) for ln in lines[1:])
    T2, dC2 = ct.read_table(f, S)
    assert np.array_equal(T2, T_t)
    assert np.allclose(dC2, dC, rtol=5e-7, atol=0)             # 7 significant digits
    assert lines[1][:14] == " 0.3200000E+03"


def test_fortran_fields_are_read():
    assert ct._fortran_float(" 0.1000000-299") == 1e-300
    assert ct._fortran_float("-0.2500000D+01") == -2.5
    assert ct.fortran_e14_7(-0.5) == "-0.5000000E+00"


def test_row_range_is_the_references():
    T = np.array([300.0, 300.2, 300.7, 305.0, 900.0, 2500.0, 2600.0, 2550.0])
    assert ct.table_rows(T) == (2, 6)
    assert ct.table_rows(np.array([300.0, 320.0, 1000.0])) == (0, 2)   # no 300.5..310 cell
