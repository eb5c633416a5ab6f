"""Seeded synthetic chemistry states for the workloads the step is measured on.

All generators return float64 torch tensors in NRG inner-cell order (i fastest)
on the requested device: ``T`` [K], ``p`` [Pa], ``rho`` [kg/m3] and ``Y`` of
shape (S, n).  Compositions are H2-air: fresh X_H2 : X_O2 : X_N2 = 2*phi : 1 :
3.762 (the reference's air ratio, package_interface/src/tests/classic_tests/
1D_laminar_velocity.f90:478-480), burnt gas = complete-combustion products with
small OH/H/O/HO2/H2O2 fractions, blended by a progress variable.  Density
follows the ideal-gas law rho = p*W/(R_u*T).

* ``flame_front_2d``: 2-D tanh flame front 300-2600 K with 2 % noise, phi ~ U[0.5, 2],
  p = 1 atm (the 128 x 128 CPU-reference case);
* ``hot_kernels_3d``: 3-D field of spherical hot kernels (T 1000-2600 K, radius
  4-20 cells) in a 300-900 K fresh mixture, p ~ U[0.5, 5] atm (the 256^3 case);
* ``ignition_mix``: a fraction of cells in the 1000-1300 K ignition window at
  phi ~ 1, the rest burnt or inert (uneven stiffness).
"""
from __future__ import annotations

import math

import numpy as np

from .mechanism import P_ATM, R_GAS_J

# species order of KEROMNES: H2 O2 N2 O H OH H2O HO2 H2O2
I_H2, I_O2, I_N2, I_O, I_H, I_OH, I_H2O, I_HO2, I_H2O2 = range(9)
RADICALS_BURNT = {I_OH: 0.010, I_H: 0.005, I_O: 0.003, I_HO2: 1.0e-5, I_H2O2: 1.0e-6}
_RADICALS_BURNT_BY_NAME = {"OH": 0.010, "H": 0.005, "O": 0.003, "HO2": 1.0e-5, "H2O2": 1.0e-6}


def _index(mech):
    """Positions of the H2-air species in the mechanism's own species list (a
    mechanism with another order, such as TEREZA's, gets the same mixtures)."""
    names = [n.upper() for n in mech.species_names]
    at = {n: names.index(n) for n in ("H2", "O2", "N2", "H2O") if n in names}
    missing = {"H2", "O2", "N2", "H2O"} - set(at)
    if missing:
        raise ValueError("H2-air states need species %s" % sorted(missing))
    rad = {names.index(k): v for k, v in _RADICALS_BURNT_BY_NAME.items() if k in names}
    return at, rad


def _torch():
    import torch
    return torch


def _compose(phi, c, W, idx=None):
    """Mole fractions (S, n) blending fresh and burnt gas by progress variable c in [0, 1].
    idx: species positions from _index (default: KEROMNES order)."""
    torch = _torch()
    n = phi.shape[0]
    S = W.shape[0]
    if idx is None:
        at = {"H2": I_H2, "O2": I_O2, "N2": I_N2, "H2O": I_H2O}
        rad = RADICALS_BURNT
    else:
        at, rad = idx
    Xu = torch.zeros((S, n), dtype=torch.float64, device=phi.device)
    Xu[at["H2"]] = 2.0 * phi
    Xu[at["O2"]] = 1.0
    Xu[at["N2"]] = 3.762
    Xb = torch.zeros_like(Xu)
    lean = phi <= 1.0
    Xb[at["H2O"]] = torch.where(lean, 2.0 * phi, torch.full_like(phi, 2.0))
    Xb[at["O2"]] = torch.where(lean, 1.0 - phi, torch.zeros_like(phi))
    Xb[at["H2"]] = torch.where(lean, torch.zeros_like(phi), 2.0 * phi - 2.0)
    Xb[at["N2"]] = 3.762
    Xu = Xu / Xu.sum(0, keepdim=True)
    Xb = Xb / Xb.sum(0, keepdim=True)
    for k, v in rad.items():
        Xb[k] = Xb[k] + v
    Xb = Xb / Xb.sum(0, keepdim=True)
    X = (1.0 - c) * Xu + c * Xb
    return X / X.sum(0, keepdim=True)


def _finish(T, p, X, W):
    Wcol = W.view(-1, 1)
    Wmix = (X * Wcol).sum(0)
    Y = X * Wcol / Wmix
    Y = Y / Y.sum(0, keepdim=True)
    rho = p * Wmix / (R_GAS_J * T)
    return dict(T=T.contiguous(), p=p.contiguous(), rho=rho.contiguous(), Y=Y.contiguous())


def _gen(seed, device):
    torch = _torch()
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def flame_front_2d(mech, nx=128, ny=128, seed=1234, device="cpu", noise=0.02, width=6.0):
    torch = _torch()
    W = torch.tensor(mech.molar_masses, dtype=torch.float64, device=device)
    g = _gen(seed, device)
    n = nx * ny
    i = torch.arange(n, device=device) % nx
    x = i.to(torch.float64)
    c = 0.5 * (1.0 + torch.tanh((x - 0.5 * nx) / width))
    eta = torch.randn(n, generator=g, dtype=torch.float64, device=device)
    T = (300.0 + 2300.0 * c) * (1.0 + noise * eta)
    phi = 0.5 + 1.5 * torch.rand(n, generator=g, dtype=torch.float64, device=device)
    p = torch.full((n,), P_ATM, dtype=torch.float64, device=device)
    X = _compose(phi, c, W, _index(mech))
    return _finish(T, p, X, W)


def hot_kernels_3d(mech, n=(256, 256, 256), seed=2024, device="cpu", n_kernels=None,
                   r_range=(4.0, 20.0), t_kernel=(1000.0, 2600.0), t_fresh=(300.0, 900.0),
                   p_atm=(0.5, 5.0)):
    """Spherical hot kernels in a fresh mixture; n = (nx, ny, nz)."""
    torch = _torch()
    nx, ny, nz = n
    W = torch.tensor(mech.molar_masses, dtype=torch.float64, device=device)
    g = _gen(seed, device)
    N = nx * ny * nz
    if n_kernels is None:
        # about 5 % of the volume inside kernels (E[r^3] for r ~ U[4, 20] is ~2496)
        n_kernels = max(1, int(round(0.05 * N / (4.0 / 3.0 * math.pi * 2496.0))))
    T = t_fresh[0] + (t_fresh[1] - t_fresh[0]) * torch.rand(N, generator=g, dtype=torch.float64,
                                                            device=device)
    p = P_ATM * (p_atm[0] + (p_atm[1] - p_atm[0]) * torch.rand(N, generator=g, dtype=torch.float64,
                                                               device=device))
    phi = 0.5 + 1.5 * torch.rand(N, generator=g, dtype=torch.float64, device=device)
    c = torch.zeros(N, dtype=torch.float64, device=device)
    kc = torch.rand((n_kernels, 3), generator=g, dtype=torch.float64, device=device)
    kc = kc * torch.tensor([nx, ny, nz], dtype=torch.float64, device=device)
    kr = r_range[0] + (r_range[1] - r_range[0]) * torch.rand(n_kernels, generator=g,
                                                             dtype=torch.float64, device=device)
    kt = t_kernel[0] + (t_kernel[1] - t_kernel[0]) * torch.rand(n_kernels, generator=g,
                                                                dtype=torch.float64, device=device)
    zi = torch.arange(nz, device=device, dtype=torch.float64)
    yi = torch.arange(ny, device=device, dtype=torch.float64)
    xi = torch.arange(nx, device=device, dtype=torch.float64)
    Tv = T.view(nz, ny, nx)
    cv = c.view(nz, ny, nx)
    for q in range(n_kernels):
        cx, cy, cz = (float(v) for v in kc[q].tolist())
        r = float(kr[q])
        lo = [max(0, int(math.floor(cz - r - 2))), max(0, int(math.floor(cy - r - 2))),
              max(0, int(math.floor(cx - r - 2)))]
        hi = [min(nz, int(math.ceil(cz + r + 2)) + 1), min(ny, int(math.ceil(cy + r + 2)) + 1),
              min(nx, int(math.ceil(cx + r + 2)) + 1)]
        if lo[0] >= hi[0] or lo[1] >= hi[1] or lo[2] >= hi[2]:
            continue
        dz = (zi[lo[0]:hi[0]] - cz).view(-1, 1, 1)
        dy = (yi[lo[1]:hi[1]] - cy).view(1, -1, 1)
        dx = (xi[lo[2]:hi[2]] - cx).view(1, 1, -1)
        d = torch.sqrt(dx * dx + dy * dy + dz * dz)
        prof = 0.5 * (1.0 - torch.tanh((d - 0.8 * r) / (0.1 * r + 0.5)))
        prof = torch.where(d <= r, prof, torch.zeros_like(prof))   # fresh gas outside r
        sub_T = Tv[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        sub_c = cv[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        newT = sub_T + (float(kt[q]) - sub_T) * prof
        Tv[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = torch.maximum(sub_T, newT)
        cv[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = torch.maximum(sub_c, 0.7 * prof)
    X = _compose(phi, c, W, _index(mech))
    return _finish(T, p, X, W)


def ignition_mix(mech, n, seed=77, device="cpu", frac_window=0.05):
    """Uneven stiffness: frac_window of cells at T ~ U[1000, 1300] K, phi ~ 1 (fresh),
    the rest burnt (T ~ U[1800, 2600]) or cold inert (T ~ U[300, 700])."""
    torch = _torch()
    W = torch.tensor(mech.molar_masses, dtype=torch.float64, device=device)
    g = _gen(seed, device)
    u = torch.rand(n, generator=g, dtype=torch.float64, device=device)
    win = u < frac_window
    burnt = (u >= frac_window) & (u < 0.5 + 0.5 * frac_window)
    T = torch.where(win, 1000.0 + 300.0 * torch.rand(n, generator=g, dtype=torch.float64, device=device),
                    torch.where(burnt, 1800.0 + 800.0 * torch.rand(n, generator=g, dtype=torch.float64,
                                                                     device=device),
                                300.0 + 400.0 * torch.rand(n, generator=g, dtype=torch.float64,
                                                           device=device)))
    phi = torch.where(win, 0.9 + 0.2 * torch.rand(n, generator=g, dtype=torch.float64, device=device),
                      0.5 + 1.5 * torch.rand(n, generator=g, dtype=torch.float64, device=device))
    c = torch.where(burnt, torch.ones_like(T), torch.zeros_like(T))
    p = P_ATM * (0.5 + 4.5 * torch.rand(n, generator=g, dtype=torch.float64, device=device))
    X = _compose(phi, c, W, _index(mech))
    return _finish(T, p, X, W)


def to_numpy(state):
    return {k: v.detach().cpu().numpy() for k, v in state.items()}


def training_states(mech, n, seed=11, device="cpu", T_range=(300.0, 2800.0), phi_range=(0.3, 3.0),
                    p_atm_range=(0.5, 5.0), burnt_fraction=0.5):
    """Initial states of the surrogate's training trajectories: T ~ U[T_range],
    phi log-uniform in phi_range, p ~ U[p_atm_range] atm; half fresh mixture
    (progress 0), half blended towards burnt gas with progress ~ U[0, 1]."""
    torch = _torch()
    W = torch.tensor(mech.molar_masses, dtype=torch.float64, device=device)
    g = _gen(seed, device)
    u = lambda: torch.rand(n, generator=g, dtype=torch.float64, device=device)
    T = T_range[0] + (T_range[1] - T_range[0]) * u()
    phi = torch.exp(math.log(phi_range[0]) + (math.log(phi_range[1]) - math.log(phi_range[0])) * u())
    p = P_ATM * (p_atm_range[0] + (p_atm_range[1] - p_atm_range[0]) * u())
    c = torch.where(u() < burnt_fraction, u(), torch.zeros(n, dtype=torch.float64, device=device))
    X = _compose(phi, c, W, _index(mech))
    return _finish(T, p, X, W)
