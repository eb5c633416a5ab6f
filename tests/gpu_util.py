"""Helpers shared by the GPU tests: run the HIP step and the oracle on one state."""
import numpy as np

from nrg_amd import capi

# Parity bars, stated once (north star): FP64 integrator vs the reference
# integrator: 1e-10 relative in temperature after the split step, 1e-9 absolute
# in mass fractions.
TOL_T_REL = 1e-10
TOL_Y_ABS = 1e-9


def gpu_flat_step(engine, state, dt, counters=True, bc=None):
    """Run nrg_chem_step_dev on a flat list of cells.  state: dict of torch tensors on cuda."""
    import torch
    T, rho, Y = state["T"], state["rho"], state["Y"]
    n = T.shape[0]
    S = Y.shape[0]
    dev = T.device
    Yp = torch.zeros((S, n), dtype=torch.float64, device=dev)
    Yp2 = torch.full((S, n), -7.0, dtype=torch.float64, device=dev)
    Ef = torch.zeros(n, dtype=torch.float64, device=dev)
    status = torch.full((n,), -1, dtype=torch.int32, device=dev)
    cnt = [torch.zeros(n, dtype=torch.int32, device=dev) for _ in range(3)] if counters else None
    bct = None
    if bc is not None:
        bct = torch.as_tensor(bc, dtype=torch.int8, device=dev)
    grid = capi.flat_grid(n)
    Ys = [Y[s].contiguous() for s in range(S)]
    stats = engine.step_dev(grid, dt, T.data_ptr(), rho.data_ptr(), [y.data_ptr() for y in Ys],
                            [Yp[s].data_ptr() for s in range(S)],
                            [Yp2[s].data_ptr() for s in range(S)], Ef.data_ptr(),
                            bc=bct.data_ptr() if bct is not None else None,
                            status=status.data_ptr(),
                            counters=[c.data_ptr() for c in cnt] if cnt else None)
    torch.cuda.synchronize()
    out = dict(Y_prod=Yp.cpu().numpy(), Y_prod2=Yp2.cpu().numpy(), E_f_prod=Ef.cpu().numpy(),
               status=status.cpu().numpy(), stats=stats)
    if cnt:
        out["nstep"], out["nfe"], out["nje"] = (c.cpu().numpy() for c in cnt)
    return out


def compare_step(gpu, orc, T, rho, dt):
    """Per-cell mass-fraction increments and energy sources of two step outputs."""
    dY_g = gpu["Y_prod"] * dt / rho
    dY_o = orc["Y_prod"] * dt / rho
    err_y = np.abs(dY_g - dY_o).max(axis=0)
    return err_y
