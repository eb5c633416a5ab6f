"""Long runs of the learned split step beside an FP64 shadow trajectory, with
a seeded synthetic advection shift and online retraining (the 512x512x256,
200-step configuration).

Two copies of the same field are advanced side by side on the device, in
NRG's layout:

* the learned trajectory: the learned chemistry step (explicit-rate screen,
  network, gate, FP64 fallback; include/nrg_surrogate.h) and the closure.
  Its fallback integrates with `fallback_eps` and `fallback_jacobian`: by
  default eps 1e-6 and the analytic Jacobian, a tolerance the run's error
  bars leave room for (the shadow keeps the reference's eps 1e-9 and its
  finite-difference Jacobian; None / 'fd' give the fallback those too);
* the FP64 shadow: the FP64 chemistry step (nrg_chem_step_dev) and the
  same closure.

After every split step both fields are shifted by the same seeded periodic
integer offset along x and y.  The shift stands in for the gas-dynamics
update between two chemistry calls (computing_module/src/current_build/
cabaret_solver.f90:1243-1268 between the chemistry calls at :1362): it
carries burning and igniting gas into fresh mixture, so the set of stiff
cells keeps changing.  The error of the learned field against the shadow is
the accumulated drift of the learned trajectory: every step records its
maximum over all cells and its 99th percentile over a random 1 % sample, in
relative temperature and in absolute major-species (H2, O2, H2O) mass
fraction.

Training pairs, (state, dt) -> (dT, dY) of one split step, come from two
places, both computed by the FP64 kernels anyway: the cells the learned step
sent to its FP64 fallback (the cells the network could not take), and random
live cells of the shadow.  Every `retrain_every` steps the network is trained
further on 80 % of the pairs kept since the last retraining (Adam, small
learning rate, normaliser unchanged).  The new weights are kept only if the
learned step on the other 20 % of the states falls back less with no larger
major-species error; the run continues with the network it kept.

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


class _Trajectory:
    """One field in NRG's layout with its chemistry solver."""

    def __init__(self, shape, st, mech, dev, opts):
        from .solver import DataManager, chemical_kinetics_solver_c
        nx, ny, nz = shape
        self.dm = DataManager(shape, device=dev, mech=mech)
        g = self.dm.grid
        view = (nz, ny, nx)
        for name, key in (("temperature", "T"), ("pressure", "p"), ("density", "rho")):
            self.dm.fields[name].inner(g).copy_(st[key].view(view))
        self.Yf = self.dm.fields["specie_mass_fraction"]
        for s in range(mech.species_number):
            self.Yf.pr[s].inner(g).copy_(st["Y"][s].view(view))
        self.solver = chemical_kinetics_solver_c(self.dm, opts)
        self.E_f = self.dm.fields["full_energy"]
        self.solver.engine.energy_from_temperature_dev(g, self.dm.fields["temperature"].ptr(),
                                                       self.Yf.ptrs(), self.E_f.ptr())
        f = self.dm.fields
        self.T, self.p, self.rho = f["temperature"], f["pressure"], f["density"]
        self.state = [self.T, self.p, self.rho, self.E_f] + list(self.Yf.pr)

    def shift(self, sy, sx):
        """Periodic shift of every state field's inner cells by (sy, sx) cells."""
        import torch
        g = self.dm.grid
        for f in self.state:
            v = f.inner(g)
            v.copy_(torch.roll(v, shifts=(sy, sx), dims=(1, 2)))

    def gather(self, off):
        """T, p, rho (n,) and Y (S, n) at allocation offsets `off`."""
        import torch
        at = lambda f: f.data.view(-1)[off]  # noqa: E731
        return (at(self.T), at(self.p), at(self.rho),
                torch.stack([at(y) for y in self.Yf.pr]))


def run(shape=(512, 512, 256), steps=200, dt=1e-7, retrain_every=50, audit_frac=0.01,
        model=None, seed=2024, shift_max=(3, 2), fallback_pairs=1 << 15, shadow_pairs=1 << 15,
        retrain_epochs=2, retrain_lr=1e-4, retrain_batch=4096, gate_overrides=None,
        screen="rates", autotune=True, log=print, device="cuda:0", fallback_eps=1e-6,
        fallback_jacobian="analytic"):
    import torch

    from . import capi, inputs
    from . import surrogate as sg
    from .mechanism import load_mechanism

    t_start = time.time()
    mech = load_mechanism()
    S = mech.species_number
    model = model or os.path.join(ROOT, "nrg_amd", "data", "surrogate_keromnes.npz")
    dev = torch.device(device)
    nx, ny, nz = shape
    ncells = nx * ny * nz
    st = inputs.hot_kernels_3d(mech, shape, seed=seed, device=dev)
    opts = dict(form=capi.NRG_FORM_PERSISTENT, cell_order=capi.NRG_ORDER_STIFFNESS, use_graph=1)
    lopts = capi.default_opts(**opts)
    if fallback_eps is not None:
        lopts.eps = fallback_eps
    if fallback_jacobian == "analytic":
        lopts.jacobian = capi.NRG_JAC_ANALYTIC
        lopts.jac_batch_pct = 10   # no right-hand sides per Jacobian: batch it early
    lrn = _Trajectory(shape, st, mech, dev, lopts)
    shd = _Trajectory(shape, st, mech, dev, capi.default_opts(**opts))
    del st
    g = lrn.dm.grid
    tuning = None
    if autotune:
        # each handle keeps the queue set-up its autotuner picks (the learned
        # step's FP64 fallback runs on the learned handle); the trial steps
        # write the chemistry outputs only, not the state
        tuning = {"learned": lrn.solver.autotune(dt).best.as_dict(),
                  "fp64": shd.solver.autotune(dt).best.as_dict()}
    # screen "only": the method without the network (explicit-rate freeze, FP64 for the rest)
    mlp, norm = sg.load(model) if screen != "only" else (None, None)
    gkw = _gate_kwargs(model)
    gkw.update(gate_overrides or {})
    scr = {"none": capi.NRG_SUR_SCREEN_NONE, "rates": capi.NRG_SUR_SCREEN_RATES,
           "only": capi.NRG_SUR_SCREEN_ONLY}[screen]
    chem = sg.SurrogateChemistry(lrn.solver, mlp, norm, sg.gate_struct(mech, screen=scr, **gkw))
    n_aud = max(1, int(round(ncells * audit_frac)))
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    NX, NY = g.hi[0] - g.lo[0] + 1, g.hi[1] - g.lo[1] + 1
    i0, j0, k0 = (g.inner_lo[d] - g.lo[d] for d in range(3))

    def offsets(q):
        """Allocation offsets of inner cells q (i fastest), in NRG's layout."""
        return (q % nx + i0) + NX * ((q // nx) % ny + j0 + NY * (q // (nx * ny) + k0))

    def fallback_cells():
        """Inner indices of the cells the learned step integrated with FP64."""
        s_in = lrn.solver.cell_status.inner(g)
        m = (s_in == capi.NRG_CELL_OK) | (s_in == capi.NRG_CELL_MXSTEP)
        return torch.nonzero(m.reshape(-1)).view(-1)

    def sample(q, n):
        if q.numel() > n:
            q = q[torch.randperm(q.numel(), generator=gen, device=dev)[:n]]
        return q

    def pair(pre, post):
        T0, p0, r0, Y0 = pre
        T1, _, _, Y1 = post
        f = sg.raw_features(T0, p0, Y0, dt)
        d = torch.cat([(T1 - T0).view(-1, 1), (Y1 - Y0).t()], dim=1)
        s = torch.cat([T0.view(1, -1), p0.view(1, -1), r0.view(1, -1), Y0], dim=0)
        return f, d, s

    pairs = {"f": [], "d": [], "s": [], "src": []}
    steps_log, retrains = [], []
    hbm_peak = 0
    torch.cuda.synchronize()
    log("learned run: %s = %d cells, %d steps, dt %.3g, shift <= %s cells/step, audit sample %d "
        "cells, retrain every %d" % (shape, ncells, steps, dt, shift_max, n_aud, retrain_every))
    t_loop = time.time()
    for k in range(steps):
        # ---- learned split step; pairs from the cells its FP64 fallback took
        torch.cuda.synchronize()
        t0 = time.time()
        sst = chem.solve_chemical_kinetics(dt)
        torch.cuda.synchronize()
        t_chem = time.time() - t0
        off_fb = offsets(sample(fallback_cells(), fallback_pairs))
        pre_fb = lrn.gather(off_fb)
        t0 = time.time()
        chem.apply_split_sources(dt)
        torch.cuda.synchronize()
        t_step = t_chem + time.time() - t0
        if off_fb.numel():
            f, d, s = pair(pre_fb, lrn.gather(off_fb))
            for key, v in (("f", f), ("d", d), ("s", s)):
                pairs[key].append(v)
            pairs["src"].append(torch.zeros(f.shape[0], dtype=torch.int8, device=dev))
        # ---- FP64 shadow split step; pairs from random live cells
        q_sh = torch.randint(0, ncells, (shadow_pairs,), generator=gen, device=dev)
        off_sh = offsets(q_sh)
        pre_sh = shd.gather(off_sh)
        torch.cuda.synchronize()
        t0 = time.time()
        fst = shd.solver.solve_chemical_kinetics(dt)
        shd.solver.apply_split_sources(dt)
        torch.cuda.synchronize()
        t_shadow = time.time() - t0
        live = pre_sh[0] >= shd.solver.engine.opts.t_gate
        if bool(live.any()):
            post_sh = shd.gather(off_sh[live])
            f, d, s = pair(tuple(x[..., live] for x in pre_sh), post_sh)
            for key, v in (("f", f), ("d", d), ("s", s)):
                pairs[key].append(v)
            pairs["src"].append(torch.ones(f.shape[0], dtype=torch.int8, device=dev))
        # ---- accumulated drift of the learned field against the shadow
        Tl, Ts = lrn.T.inner(g), shd.T.inner(g)
        relT_max = float(((Tl - Ts).abs() / Ts).max())
        dY_max = max(float((lrn.Yf.pr[s].inner(g) - shd.Yf.pr[s].inner(g)).abs().max())
                     for s in MAJOR)
        off_a = offsets(torch.randint(0, ncells, (n_aud,), generator=gen, device=dev))
        Ta, _, _, Ya = lrn.gather(off_a)
        Tb, _, _, Yb = shd.gather(off_a)
        relT_a = (Ta - Tb).abs() / Tb
        dY_a = torch.stack([(Ya[s] - Yb[s]).abs() for s in MAJOR]).max(dim=0).values
        # ---- the flow: the same seeded shift of both fields
        sx = int(rng.integers(-shift_max[0], shift_max[0] + 1))
        sy = int(rng.integers(-shift_max[1], shift_max[1] + 1))
        lrn.shift(sy, sx)
        shd.shift(sy, sx)
        free, total = torch.cuda.mem_get_info()
        hbm_peak = max(hbm_peak, total - free)
        rec = {
            "step": k, "ms": round(t_step * 1e3, 3), "fp64_ms": round(t_shadow * 1e3, 3),
            "shift": [sx, sy],
            "fallback_frac": sst.cells_fallback / ncells, "accepted_frac": sst.cells_accepted / ncells,
            "frozen_frac": sst.cells_frozen / ncells, "network_frac": sst.cells_network / ncells,
            "fallback_nfe": int(sst.fallback.nfe_sum), "fp64_nfe": int(fst.nfe_sum),
            "max_rel_T": relT_max, "max_abs_Y_major": dY_max,
            "sample_max_rel_T": float(relT_a.max()),
            "sample_p99_rel_T": float(torch.quantile(relT_a.float(), 0.99)),
            "sample_max_abs_Y_major": float(dY_a.max()),
            "sample_p99_abs_Y_major": float(torch.quantile(dY_a.float(), 0.99)),
            "sample_cells_over_bar": int(((dY_a > BAR_ABS_Y) | (relT_a > BAR_REL_T)).sum()),
        }
        steps_log.append(rec)
        log("step %d: learned %.1f ms (FP64 %.1f ms), fallback %.4f, network %.4f, drift max rel T "
            "%.2e, max |dY| %.2e, p99 %.1e / %.1e" % (
                k, rec["ms"], rec["fp64_ms"], rec["fallback_frac"], rec["network_frac"],
                relT_max, dY_max, rec["sample_p99_rel_T"], rec["sample_p99_abs_Y_major"]))
        if (retrain_every > 0 and (k + 1) % retrain_every == 0 and (k + 1) < steps and pairs["f"]
                and mlp is not None):
            fcat, dcat = torch.cat(pairs["f"]), torch.cat(pairs["d"])
            src = torch.cat(pairs["src"])
            x, y = norm.x(fcat), norm.y(dcat)
            # Retrain on 80 % of the kept pairs.  The other 20 % decide whether the
            # new weights stay: the learned step (screen, network, gate, FP64
            # fallback) runs on those held-out states with the old and the new
            # network, and the new one is kept only if it falls back on fewer
            # cells with no larger major-species error against the FP64
            # increments.  A lower training loss alone is not enough: a network
            # that fits the increments better on average can predict small
            # nonzero changes for quiet cells, which the gate then rejects.
            perm = torch.randperm(x.shape[0], generator=gen, device=dev)
            n_tr = max(1, int(0.8 * x.shape[0]))
            tr, va = perm[:n_tr], perm[n_tr:][:65536]
            if va.numel() == 0:
                va = tr[:65536]
            scat = torch.cat(pairs["s"], dim=1)[:, va].contiguous()
            dv = dcat[va]
            Tv, pv, rv = scat[0].contiguous(), scat[1].contiguous(), scat[2].contiguous()
            Yv = scat[3:].contiguous()
            nv = Tv.shape[0]

            def validate():
                out = sg._flat_outputs(nv, S, dev)
                stv = sg.surrogate_step_flat(lrn.solver.engine, mlp, chem.norm_s, chem.gate,
                                             Tv.clone(), pv, rv, Yv.clone(), dt, out)
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
            # no slack on the error: a network kept with a larger one-step error
            # drifts further from the FP64 trajectory step after step
            kept = fb1 < fb0 and err1 <= err0
            if not kept:
                mlp.set_params(params0)
            n_fb = int((src == 0).sum())
            retrains.append({"after_step": k, "pairs": int(x.shape[0]), "fallback_pairs": n_fb,
                             "fallback_share": n_fb / int(x.shape[0]), "train_pairs": int(n_tr),
                             "val_states": int(nv), "epochs": retrain_epochs,
                             "val_fallback_before": fb0, "val_fallback_after": fb1,
                             "val_max_abs_dY_major_before": err0, "val_max_abs_dY_major_after": err1,
                             "kept": bool(kept), "train_steps": len(hist) and hist[-1][0],
                             "s": round(time.time() - t0, 2)})
            log("retrain after step %d: %d pairs (%d from fallback cells); held-out fallback "
                "%.4f -> %.4f, max |dY| %.2e -> %.2e: %s (%.1f s)" % (
                    k, x.shape[0], n_fb, fb0, fb1, err0, err1, "kept" if kept else "reverted",
                    time.time() - t0))
            pairs = {"f": [], "d": [], "s": [], "src": []}
    t_total = time.time() - t_loop
    state_bytes = sum(f.data.numel() * f.data.element_size() for f in lrn.state)
    field_bytes = sum(t.data.numel() * t.data.element_size()
                      for tr in (lrn, shd) for f in tr.dm.fields.values()
                      for t in (getattr(f, "pr", None) or [f]))
    ms_l = sum(r["ms"] for r in steps_log)
    ms_f = sum(r["fp64_ms"] for r in steps_log)
    rep = {
        "shape": list(shape), "cells": ncells, "steps": steps, "dt": dt,
        "model": os.path.relpath(model, ROOT), "gate": gkw, "screen": screen,
        "shift_max_cells": list(shift_max), "audit_cells_per_step": n_aud,
        "retrain_every": retrain_every, "autotune": tuning,
        "fallback_integrator": {"eps": lopts.eps, "jacobian": fallback_jacobian},
        "cell_steps_per_s_learned": ncells * steps / ms_l * 1e3,
        "cell_steps_per_s_fp64": ncells * steps / ms_f * 1e3,
        "speedup_vs_fp64": ms_f / ms_l,
        "wall_s": round(t_total, 1), "setup_s": round(t_loop - t_start, 1),
        "fallback_frac_mean": float(np.mean([r["fallback_frac"] for r in steps_log])),
        "fallback_frac_max": float(np.max([r["fallback_frac"] for r in steps_log])),
        "network_frac_mean": float(np.mean([r["network_frac"] for r in steps_log])),
        "max_rel_T": float(np.max([r["max_rel_T"] for r in steps_log])),
        "max_abs_Y_major": float(np.max([r["max_abs_Y_major"] for r in steps_log])),
        "final_max_rel_T": steps_log[-1]["max_rel_T"],
        "final_max_abs_Y_major": steps_log[-1]["max_abs_Y_major"],
        "bars": {"rel_T": BAR_REL_T, "abs_Y_major": BAR_ABS_Y},
        "steps_within_bars": int(sum(r["max_rel_T"] <= BAR_REL_T and
                                     r["max_abs_Y_major"] <= BAR_ABS_Y for r in steps_log)),
        "hbm_in_use_peak_GB": round(hbm_peak / 1e9, 2),
        "state_GB": round(state_bytes / 1e9, 2),
        "field_GB": round(field_bytes / 1e9, 2),
        "retrains": retrains, "per_step": steps_log,
    }
    if mlp is not None:
        mlp.close()
    for tr in (lrn, shd):
        tr.solver.engine.close()
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
    ap.add_argument("--shift-max", type=int, nargs=2, default=[3, 2],
                    help="largest |shift| per step along x and y, in cells")
    ap.add_argument("--screen", choices=["rates", "none", "only"], default="rates",
                    help="only: the method without the network")
    ap.add_argument("--model", default=None)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--gate", action="append", default=[], metavar="K=V",
                    help="override a gate threshold of the model file, e.g. rate_tol=1e-4")
    ap.add_argument("--fallback-eps", type=float, default=1e-6,
                    help="integrator eps of the learned step's FP64 fallback (0: the "
                         "reference's 1e-9)")
    ap.add_argument("--fallback-jacobian", choices=["fd", "analytic"], default="analytic")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    over = dict((k, float(v)) for k, v in (g.split("=", 1) for g in a.gate))
    rep = run(tuple(a.shape), a.steps, a.dt, a.retrain_every, a.audit_frac, a.model, a.seed,
              shift_max=tuple(a.shift_max), retrain_epochs=a.retrain_epochs,
              retrain_lr=a.retrain_lr, gate_overrides=over, screen=a.screen,
              log=lambda m: print(m, flush=True), fallback_eps=a.fallback_eps or None,
              fallback_jacobian=a.fallback_jacobian)
    line = json.dumps({k: v for k, v in rep.items() if k != "per_step"})
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(rep, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
