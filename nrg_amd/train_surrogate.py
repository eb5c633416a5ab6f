"""Train the chemistry surrogate on the GPU and save it.

    python -m nrg_amd.train_surrogate --out <dir> [--traj 1048576 --steps 20 --epochs 8]

1. pairs: trajectories of seeded H2-air states (T 300-2800 K, phi 0.3-3,
   p 0.5-5 atm; half fresh, half partly burnt) advanced split step by split
   step by the library's FP64 integrator (chemistry step + closure) with
   dt log-uniform in [1e-8, 1e-6] s per group of trajectories;
2. training: the library's BF16 MFMA forward/backward/Adam kernels on the
   normalised pairs (90 % of them), cosine learning-rate decay;
3. report: errors of the network on the held-out 10 % against the FP64
   increments and against the method without it (zero increment).
Writes <dir>/surrogate_keromnes.npz and <dir>/surrogate_report.json.
"""
import argparse
import json
import os
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--traj", type=int, default=1 << 20)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--groups", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--calibrate", default=None,
                    help="only calibrate the gate of this trained file and write it to --out")
    ap.add_argument("--target-dy", type=float, default=1e-4,
                    help="max one-step error of accepted major species against FP64")
    ap.add_argument("--set-gate", default=None, metavar="FILE",
                    help="only rewrite the gate thresholds of this trained file (--gate K=V ...) "
                         "into --out")
    ap.add_argument("--gate", action="append", default=[], metavar="K=V",
                    help="gate threshold for --set-gate, e.g. dT_max=10")
    a = ap.parse_args(argv)
    if a.set_gate:
        from nrg_amd import surrogate as sg
        kw = dict((k, float(v)) for k, v in (g.split("=", 1) for g in a.gate))
        os.makedirs(a.out, exist_ok=True)
        sg.set_gate(a.set_gate, os.path.join(a.out, "surrogate_keromnes.npz"), **kw)
        return 0

    import torch

    from nrg_amd import surrogate as sg
    from nrg_amd.mechanism import load_mechanism
    from nrg_amd.solver import ChemistryStep

    def log(msg):
        print(msg, flush=True)

    os.makedirs(a.out, exist_ok=True)
    mech = load_mechanism()
    eng = ChemistryStep(mech)
    if a.calibrate:
        net, norm = sg.load(a.calibrate)
        rep = calibrate(sg, eng, mech, net, norm, a, log)
        sg.save_with_gate(os.path.join(a.out, "surrogate_keromnes.npz"), net, norm, rep["gate"])
        with open(os.path.join(a.out, "surrogate_calibration.json"), "w") as fh:
            json.dump(rep, fh, indent=1)
        return 0
    t0 = time.time()
    pairs = sg.generate_pairs(eng, mech, a.traj, a.steps, seed=a.seed, groups=a.groups, log=log)
    t_gen = time.time() - t0
    f, d = sg.pair_tensors(pairs)
    del pairs
    n = f.shape[0]
    g = torch.Generator(device=f.device)
    g.manual_seed(a.seed)
    perm = torch.randperm(n, generator=g, device=f.device)
    tr, va = perm[: int(0.9 * n)], perm[int(0.9 * n):]
    norm = sg.Normalizer.fit(f[tr], d[tr])
    net = sg.MLP(mech.species_number + 3, mech.species_number + 1)
    net.set_params(sg.init_params(net.n_in, net.n_out, seed=a.seed))
    x, y = norm.x(f[tr]), norm.y(d[tr])
    t1 = time.time()
    hist = sg.train(net, x, y, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t1
    rep = sg.evaluate(net, norm, f[va], d[va])
    rep.update(pairs=n, generate_s=t_gen, train_s=t_train, epochs=a.epochs, batch=a.batch,
               train_steps=len(tr) // a.batch * a.epochs, final_loss=hist[-1][1],
               samples_per_s=len(tr) * a.epochs / t_train)
    log(json.dumps(rep))
    cal = calibrate(sg, eng, mech, net, norm, a, log)
    rep["gate"] = cal["gate"]
    sg.save_with_gate(os.path.join(a.out, "surrogate_keromnes.npz"), net, norm, cal["gate"],
                      meta=dict(traj=a.traj, steps=a.steps, epochs=a.epochs, seed=a.seed))
    with open(os.path.join(a.out, "surrogate_calibration.json"), "w") as fh:
        json.dump(cal, fh, indent=1)
    with open(os.path.join(a.out, "surrogate_report.json"), "w") as fh:
        json.dump(rep, fh, indent=1)
    net.close()
    eng.close()
    return 0


def calibrate(sg, eng, mech, net, norm, a, log):
    """Gate thresholds for the trained network (see surrogate.calibrate_gate)."""
    states = sg.calibration_states(eng, mech, 65536, seed=a.seed + 77)
    best, results = sg.calibrate_gate(eng, mech, net, norm, states, target_dY=a.target_dy, log=log)
    log("calibrate: chosen %s" % json.dumps(best))
    return {"gate": {k: best[k] for k in ("dT_max", "energy_tol", "elem_tol", "elem_floor")},
            "chosen": best, "target_dY": a.target_dy, "candidates": results}


if __name__ == "__main__":
    sys.exit(main())
