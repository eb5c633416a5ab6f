"""Long runs of the learned split step with a running FP64 audit and online
retraining (the 512x512x256, 200-step configuration).

Each split step runs the learned chemistry step (network, gate, FP64
fallback; include/nrg_surrogate.h) and the closure on device-resident fields
in NRG's layout.  Before the step a random `audit_frac` of the inner cells is
copied out, and after it those cells are integrated again from the same
state by the FP64 kernels (chemistry step + closure).  The learned and FP64
results give the step's audit error, and the audited cell-steps are kept as
training pairs.  Every `retrain_every` steps the network is trained further
on 80 % of the pairs kept since the last retraining (Adam, small learning
rate, normaliser unchanged).  The new weights are kept only if the learned
step on the other 20 % of the states falls back less with no larger error;
the run continues with the network it kept.  The report holds per-step
fallback fractions, audit errors, retraining losses, the HBM in use and the
times.

    python -m nrg_amd.learned_run --shape 512 512 256 --steps 200 --out run.json
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAJOR = (0, 1, 6)          # H2, O2, H2O
BAR_REL_T, BAR_ABS_Y = 1e-2, 1e-3


def _gate_kwargs(path):
    z = np.load(path, allow_pickle=False)
    return {k: float(z["gate_" + k]) for k in ("dT_max", "dT_freeze", "dY_freeze", "energy_tol",
                                               "elem_tol", "elem_floor", "rate_tol")
            if ("gate_" + k) in z.files}


def run(shape=(512, 512, 256), steps=200, dt=1e-7, retrain_every=50, audit_frac=0.01,
        model=None, seed=2024, retrain_epochs=2, retrain_lr=1e-4, retrain_batch=4096,
        max_pairs=1 << 22, gate_overrides=None, log=print, device="cuda:0"):
    import torch

    from . import capi, inputs
    from . import surrogate as sg
    from .mechanism import load_mechanism
    from .solver import DataManager, chemical_kinetics_solver_c

    t_start = time.time()
    mech = load_mechanism()
    S = mech.species_number
    model = model or os.path.join(ROOT, "nrg_amd", "data", "surrogate_keromnes.npz")
    dev = torch.device(device)
    nx, ny, nz = shape
    view = (nz, ny, nx)
    st = inputs.hot_kernels_3d(mech, shape, seed=seed, device=dev)
    dm = DataManager(shape, device=dev, mech=mech)
    g = dm.grid
    for name, key in (("temperature", "T"), ("pressure", "p"), ("density", "rho")):
        dm.fields[name].inner(g).copy_(st[key].view(view))
    Yf = dm.fields["specie_mass_fraction"]
    for s in range(S):
        Yf.pr[s].inner(g).copy_(st["Y"][s].view(view))
    del st
    solver = chemical_kinetics_solver_c(dm)
    eng = solver.engine
    E_f = dm.fields["full_energy"]
    eng.energy_from_temperature_dev(g, dm.fields["temperature"].ptr(), Yf.ptrs(), E_f.ptr())
    mlp, norm = sg.load(model)
    gkw = _gate_kwargs(model)
    gkw.update(gate_overrides or {})
    chem = sg.SurrogateChemistry(solver, mlp, norm, sg.gate_struct(mech, **gkw))
    ncells = nx * ny * nz
    n_aud = max(1, int(round(ncells * audit_frac)))
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    flat = capi.flat_grid(n_aud)
    NX, NY = g.hi[0] - g.lo[0] + 1, g.hi[1] - g.lo[1] + 1

    def offsets(q):
        """Allocation offsets of inner cells q (i fastest), in NRG's layout."""
        i = q % nx + (g.inner_lo[0] - g.lo[0])
        j = (q // nx) % ny + (g.inner_lo[1] - g.lo[1])
        kk = q // (nx * ny) + (g.inner_lo[2] - g.lo[2])
        return i + NX * (j + NY * kk)

    def at(field, off):
        return field.data.view(-1)[off]

    Tf, pf, rf = dm.fields["temperature"], dm.fields["pressure"], dm.fields["density"]
    pairs_f, pairs_d, pairs_s = [], [], []
    per_step_pairs = max(1, max_pairs // max(1, retrain_every))
    steps_log, retrains = [], []
    hbm_peak = 0
    torch.cuda.synchronize()
    log("learned run: %s = %d cells, %d steps, dt %.3g, audit %d cells/step, retrain every %d"
        % (shape, ncells, steps, dt, n_aud, retrain_every))
    t_loop = time.time()
    for k in range(steps):
        off = offsets(torch.randint(0, ncells, (n_aud,), generator=gen, device=dev))
        T0, p0, r0, E0 = (at(f_, off).clone() for f_ in (Tf, pf, rf, E_f))
        Y0 = torch.stack([at(Yf.pr[s], off) for s in range(S)]).contiguous()
        torch.cuda.synchronize()
        t0 = time.time()
        sst = chem.solve_chemical_kinetics(dt)
        chem.apply_split_sources(dt)
        torch.cuda.synchronize()
        t_step = time.time() - t0
        T1 = at(Tf, off)
        Y1 = torch.stack([at(Yf.pr[s], off) for s in range(S)])
        # FP64 from the same state: chemistry step + closure on the audited cells
        Ta, pa, Ea, Ya = T0.clone(), p0.clone(), E0.clone(), Y0.clone()
        Yp = torch.zeros((S, n_aud), dtype=torch.float64, device=dev)
        Yp2 = torch.zeros_like(Yp)
        Ef = torch.zeros(n_aud, dtype=torch.float64, device=dev)
        ys = [Ya[s].data_ptr() for s in range(S)]
        eng.step_dev(flat, dt, Ta.data_ptr(), r0.data_ptr(), ys, [Yp[s].data_ptr() for s in range(S)],
                     [Yp2[s].data_ptr() for s in range(S)], Ef.data_ptr())
        eng.closure_dev(flat, dt, r0.data_ptr(), Ef.data_ptr(), [Yp[s].data_ptr() for s in range(S)],
                        Ea.data_ptr(), ys, Ta.data_ptr(), p=pa.data_ptr())
        torch.cuda.synchronize()
        relT = ((T1 - Ta).abs() / Ta)
        dY = torch.stack([(Y1[s] - Ya[s]).abs() for s in MAJOR]).max(dim=0).values
        # the worst audited cell: its state, how the learned step treated it
        # (status 6 = learned answer; all-zero Y_prod = frozen as inert)
        w = int(torch.argmax(dY))
        ow = off[w:w + 1]
        stat_w = int(at(solver.cell_status, ow)[0])
        yp_w = torch.stack([at(solver.Y_prod.pr[s], ow)[0] for s in range(S)])
        worst = {"T": float(T0[w]), "p": float(p0[w]), "status": stat_w,
                 "frozen": bool(stat_w == capi.NRG_CELL_SURROGATE and bool((yp_w == 0).all())),
                 "Y": [float(v) for v in Y0[:, w]],
                 "dY_fp64": [float(Ya[s, w] - Y0[s, w]) for s in MAJOR],
                 "dY_learned": [float(Y1[s, w] - Y0[s, w]) for s in MAJOR]}
        live = T0 >= eng.opts.t_gate
        nl = int(live.sum())
        if nl:
            keep = torch.nonzero(live).view(-1)
            if keep.numel() > per_step_pairs:
                keep = keep[torch.randperm(keep.numel(), generator=gen, device=dev)[:per_step_pairs]]
            f = sg.raw_features(T0[keep], p0[keep], Y0[:, keep], dt)
            d = torch.cat([(Ta - T0)[keep].view(-1, 1), (Ya - Y0)[:, keep].t()], dim=1)
            pairs_f.append(f)
            pairs_d.append(d)
            # the states themselves, to run the learned step on held-out pairs
            pairs_s.append(torch.cat([T0[keep].view(1, -1), p0[keep].view(1, -1),
                                      r0[keep].view(1, -1), Y0[:, keep]], dim=0))
        free, total = torch.cuda.mem_get_info()
        hbm_peak = max(hbm_peak, total - free)
        rec = {
            "step": k, "ms": round(t_step * 1e3, 3),
            "fallback_frac": sst.cells_fallback / ncells, "accepted_frac": sst.cells_accepted / ncells,
            "frozen_frac": sst.cells_frozen / ncells,
            "audit_max_rel_T": float(relT.max()), "audit_p99_rel_T": float(torch.quantile(relT.float(), 0.99)),
            "audit_max_abs_Y_major": float(dY.max()),
            "audit_p99_abs_Y_major": float(torch.quantile(dY.float(), 0.99)),
            "audit_cells_over_bar": int((dY > BAR_ABS_Y).sum()), "audit_worst": worst,
        }
        steps_log.append(rec)
        log("step %d: %.1f ms, fallback %.4f, audit max rel T %.2e, max |dY| %.2e (%d over; worst "
            "T %.0f K status %d%s)" % (k, rec["ms"], rec["fallback_frac"], rec["audit_max_rel_T"],
                                        rec["audit_max_abs_Y_major"], rec["audit_cells_over_bar"],
                                        worst["T"], worst["status"],
                                        " frozen" if worst["frozen"] else ""))
        if retrain_every > 0 and (k + 1) % retrain_every == 0 and (k + 1) < steps and pairs_f:
            fcat = torch.cat(pairs_f)
            dcat = torch.cat(pairs_d)
            x, y = norm.x(fcat), norm.y(dcat)
            # Retrain on 80 % of the kept pairs.  The other 20 % decide whether the
            # new weights stay: the learned step (network, gate, FP64 fallback) runs
            # on those held-out states with the old and the new network, and the
            # new one is kept only if it falls back on fewer cells with no larger
            # major-species error against the FP64 increments.  A lower training
            # loss alone is not enough: a network that fits the increments better
            # on average can predict small nonzero changes for quiet cells, which
            # the gate's freeze and confidence thresholds then reject.
            perm = torch.randperm(x.shape[0], generator=gen, device=dev)
            n_tr = max(1, int(0.8 * x.shape[0]))
            tr, va = perm[:n_tr], perm[n_tr:][:65536]
            if va.numel() == 0:
                va = tr[:65536]
            scat = torch.cat(pairs_s, dim=1)[:, va].contiguous()
            dv = dcat[va]
            Tv, pv, rv = scat[0].contiguous(), scat[1].contiguous(), scat[2].contiguous()
            Yv = scat[3:].contiguous()
            nv = Tv.shape[0]

            def validate():
                out = sg._flat_outputs(nv, S, dev)
                stv = sg.surrogate_step_flat(eng, mlp, chem.norm_s, chem.gate, Tv.clone(), pv, rv,
                                             Yv.clone(), dt, out)
                torch.cuda.synchronize()
                dYl = out["Yp"] * dt / rv
                err = max(float((dYl[s] - dv[:, 1 + s]).abs().max()) for s in MAJOR)
                return stv.cells_fallback / nv, err

            fb0, err0 = validate()
            params0 = mlp.get_params()
            t0 = time.time()
            xt, yt = x[tr].contiguous(), y[tr].contiguous()
            hist = sg.train(mlp, xt, yt, epochs=retrain_epochs, batch=min(retrain_batch, xt.shape[0]),
                            lr=retrain_lr, lr_min=retrain_lr * 0.1, seed=seed + k, log=None)
            torch.cuda.synchronize()
            fb1, err1 = validate()
            kept = fb1 < fb0 and err1 <= max(err0, 0.5 * BAR_ABS_Y)
            if not kept:
                mlp.set_params(params0)
            retrains.append({"after_step": k, "pairs": int(x.shape[0]), "train_pairs": int(n_tr),
                             "val_states": int(nv), "epochs": retrain_epochs,
                             "val_fallback_before": fb0, "val_fallback_after": fb1,
                             "val_max_abs_dY_major_before": err0, "val_max_abs_dY_major_after": err1,
                             "kept": bool(kept), "train_steps": len(hist) and hist[-1][0],
                             "s": round(time.time() - t0, 2)})
            log("retrain after step %d: %d pairs; held-out fallback %.4f -> %.4f, max |dY| %.2e -> "
                "%.2e: %s (%.1f s)" % (k, x.shape[0], fb0, fb1, err0, err1,
                                      "kept" if kept else "reverted", time.time() - t0))
            pairs_f, pairs_d, pairs_s = [], [], []
    t_total = time.time() - t_loop
    field_bytes = sum(t.data.numel() * t.data.element_size()
                      for f in dm.fields.values() for t in (getattr(f, "pr", None) or [f]))
    rep = {
        "shape": list(shape), "cells": ncells, "steps": steps, "dt": dt, "model": os.path.relpath(model, ROOT),
        "gate": gkw, "audit_cells_per_step": n_aud, "retrain_every": retrain_every,
        "cell_steps_per_s_learned": ncells * steps / sum(r["ms"] for r in steps_log) * 1e3,
        "wall_s": round(t_total, 1), "setup_s": round(t_loop - t_start, 1),
        "fallback_frac_mean": float(np.mean([r["fallback_frac"] for r in steps_log])),
        "fallback_frac_max": float(np.max([r["fallback_frac"] for r in steps_log])),
        "audit_max_rel_T": float(np.max([r["audit_max_rel_T"] for r in steps_log])),
        "audit_max_abs_Y_major": float(np.max([r["audit_max_abs_Y_major"] for r in steps_log])),
        "bars": {"rel_T": BAR_REL_T, "abs_Y_major": BAR_ABS_Y},
        "hbm_in_use_peak_GB": round(hbm_peak / 1e9, 2),
        "field_GB": round(field_bytes / 1e9, 2),
        "retrains": retrains, "per_step": steps_log,
    }
    mlp.close()
    eng.close()
    return rep


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--shape", type=int, nargs=3, default=[512, 512, 256])
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--dt", type=float, default=1e-7)
    ap.add_argument("--retrain-every", type=int, default=50)
    ap.add_argument("--audit-frac", type=float, default=0.01)
    ap.add_argument("--retrain-epochs", type=int, default=2)
    ap.add_argument("--retrain-lr", type=float, default=1e-4)
    ap.add_argument("--model", default=None)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--gate", action="append", default=[], metavar="K=V",
                    help="override a gate threshold of the model file, e.g. rate_tol=1e-4")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    over = dict((k, float(v)) for k, v in (g.split("=", 1) for g in a.gate))
    rep = run(tuple(a.shape), a.steps, a.dt, a.retrain_every, a.audit_frac, a.model, a.seed,
              a.retrain_epochs, a.retrain_lr, gate_overrides=over,
              log=lambda m: print(m, flush=True))
    line = json.dumps({k: v for k, v in rep.items() if k != "per_step"})
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(rep, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
