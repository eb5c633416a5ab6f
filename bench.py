#!/usr/bin/env python3
"""Throughput of NRG's operator-split chemistry step on one MI355X.

Workload: a 256^3 H2-air field (KEROMNES, 9 species / 24 reactions) of seeded
spherical hot kernels (T 1000-2600 K, radius 4-20 cells) in a 300-900 K fresh
mixture, p ~ U[0.5, 5] atm, FP64 structure-of-arrays fields with NRG's ghost
layers, dt = 1e-7 s.  One "step" is one split step: the chemistry step
(nrg_chem_step_dev: per-cell Gear-BDF with SLATEC DDRIV3 semantics) followed by
the legacy CABARET source application and e->T inversion (nrg_chem_closure_dev).
Fields stay resident in HBM; the timed region holds exactly --steps split steps.

Prints one JSON line: metric, value in cell-steps/s, the roofline of the
chemistry kernels (HIP events on the library's stream; HBM bytes and lane
share from the committed rocprofv3 summaries), the CPU baseline (the
reference's own SLATEC integrator built with flang, on the host cores, over a
sample of the same state) with an audit of the sample against the oracle, and
legs run from the same initial state: the learned step and the method without
the network (their FP64 fallback integrating at eps 1e-6 with the analytic
Jacobian, on a handle of its own, and again at the reference's eps 1e-9 with
the FD Jacobian), the table-approximated mode, the low-Mach coupling kernels,
and the uneven-stiffness field (persistent queue vs one cell per lane at
three dt).
"""
import argparse
import json
import re
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "chem cell-steps/s at 256^3: FP64 integrator vs MLP surrogate @≤1e-2 rel T err"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: the configuration the committed rocprofv3 summaries were taken on
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="cells per dimension")
    ap.add_argument("--dt", type=float, default=1.0e-7)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-sample", type=int, default=1 << 20,
                    help="cells of the initial state timed on the host cores (the reference's "
                         "SLATEC integrator, and the oracle) and audited (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: every core this process may run on, at most 16 (one GPU's share "
                         "of the host)")
    ap.add_argument("--form", choices=["plain", "persistent"], default="persistent",
                    help="integrator launch form (same results)")
    ap.add_argument("--jacobian", choices=["fd", "analytic"], default="fd",
                    help="fd: SLATEC finite differences (reference parity); analytic: fast mode")
    ap.add_argument("--order", choices=["natural", "stiffness"], default="stiffness",
                    help="queue order of the cells (same results)")
    ap.add_argument("--graph", type=int, default=1, help="capture the step in a HIP graph (same results)")
    ap.add_argument("--chunk-log2", type=int, default=0,
                    help="queue chunk of 2^k cells (0: library default 2^23; the autotuner may change it)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="run the library's autotuner on the initial state before warmup")
    ap.add_argument("--jac-pct", type=int, default=0,
                    help="batched-Jacobian threshold in percent of pending lanes (0: default)")
    ap.add_argument("--gather-pct", type=int, default=0,
                    help="error-test gathering threshold in percent of stepping lanes (same results)")
    ap.add_argument("--refill-idle", type=int, default=0,
                    help="persistent form: idle lanes of 64 that trigger a refill (0: default 63; "
                         "same results)")
    ap.add_argument("--surrogate", default="auto",
                    help="trained surrogate (.npz) timed beside the FP64 integrator from the same "
                         "initial state; 'auto': nrg_amd/data/surrogate_keromnes.npz if present; "
                         "'none': skip")
    ap.add_argument("--gate-energy-tol", type=float, default=None, help="K (default: trained file's)")
    ap.add_argument("--fallback-eps", type=float, default=1e-6,
                    help="learned-step legs: integrator eps of the FP64 fallback, on a handle of "
                         "its own (0: the reference's 1e-9 on the FP64 leg's handle)")
    ap.add_argument("--fallback-jacobian", choices=["fd", "analytic"], default="analytic",
                    help="learned-step legs: Jacobian of the FP64 fallback")
    ap.add_argument("--reference-fallback", type=int, default=1,
                    help="also time both learned-step legs with the reference-eps FD fallback")
    ap.add_argument("--gate-elem-tol", type=float, default=None)
    ap.add_argument("--gate-elem-floor", type=float, default=None, help="mol/kg")
    ap.add_argument("--gate-dt-max", type=float, default=None, help="K")
    ap.add_argument("--gate-dt-freeze", type=float, default=None, help="K")
    ap.add_argument("--gate-dy-freeze", type=float, default=None)
    ap.add_argument("--gate-rate-tol", type=float, default=None,
                    help="mass fraction; 0 turns the rate-consistency test off")
    ap.add_argument("--screen", choices=["none", "rates"], default="rates",
                    help="cells the network sees: 'rates': those the explicit-rate screen does "
                         "not freeze; 'none': every live cell")
    ap.add_argument("--no-network", type=int, default=1,
                    help="also time the method without the network (explicit-rate screen, FP64 "
                         "for every cell it does not freeze) from the same initial state")
    ap.add_argument("--table", type=int, default=1,
                    help="time the table_approximated mode beside the FP64 integrator (0: skip)")
    ap.add_argument("--low-mach", type=int, default=1,
                    help="time the low-Mach coupling kernels on the FP64 run's outputs (0: skip)")
    ap.add_argument("--uneven", type=int, default=1,
                    help="time the uneven-stiffness 256^3 field (5%% of cells in the 1000-1300 K "
                         "ignition window) at dt 1e-8, 1e-7, 1e-6 s: persistent graph queue vs "
                         "one cell per lane (0: skip)")
    ap.add_argument("--table-line", type=int, default=512,
                    help="cells of the flame line the table is written from")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device copy of this size before the timed region, as the "
                         "known-size reference for HBM counters (FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def effective_setup(args, tuning):
    """Queue set-up the timed steps ran with: the autotuner's choice when it ran."""
    from nrg_amd import capi
    if tuning is None:
        return {"form": args.form, "order": args.order, "jac_batch_pct": args.jac_pct or 50,
                "errtest_gather_pct": args.gather_pct, "refill_min_idle": args.refill_idle or 63,
                "chunk_log2": args.chunk_log2 or 23}
    b = tuning["best"]
    return {"form": "persistent" if b["form"] == capi.NRG_FORM_PERSISTENT else "plain",
            "order": "stiffness" if b["cell_order"] == capi.NRG_ORDER_STIFFNESS else "natural",
            "jac_batch_pct": b["jac_batch_pct"], "errtest_gather_pct": b["errtest_gather_pct"],
            "refill_min_idle": b["refill_min_idle"],
            "chunk_log2": int(b["chunk_cells"]).bit_length() - 1}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only (the chemistry step is cell-local; "
              "see DESIGN.md for how it shards); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from nrg_amd import capi, flop_model, inputs
    from nrg_amd.mechanism import load_mechanism
    from nrg_amd.solver import DataManager, chemical_kinetics_solver_c

    mech = load_mechanism()
    S = mech.species_number
    n = args.n
    shape = (n, n, n)
    dev = torch.device("cuda:0")

    # ---- state in HBM ---------------------------------------------------------------
    st = inputs.hot_kernels_3d(mech, shape, seed=args.seed, device=dev)
    dm = DataManager(shape, device=dev, mech=mech)
    g = dm.grid
    view = (n, n, n)
    dm.fields["temperature"].inner(g).copy_(st["T"].view(view))
    dm.fields["pressure"].inner(g).copy_(st["p"].view(view))
    dm.fields["density"].inner(g).copy_(st["rho"].view(view))
    Yf = dm.fields["specie_mass_fraction"]
    for s in range(S):
        Yf.pr[s].inner(g).copy_(st["Y"][s].view(view))
    form = capi.NRG_FORM_PERSISTENT if args.form == "persistent" else capi.NRG_FORM_PLAIN
    order = capi.NRG_ORDER_STIFFNESS if args.order == "stiffness" else capi.NRG_ORDER_NATURAL
    jac = capi.NRG_JAC_ANALYTIC if args.jacobian == "analytic" else capi.NRG_JAC_FD_SLATEC
    solver = chemical_kinetics_solver_c(dm, capi.default_opts(form=form, jac_batch_pct=args.jac_pct,
                                                              cell_order=order, jacobian=jac,
                                                              use_graph=args.graph,
                                                              chunk_log2=args.chunk_log2,
                                                              errtest_gather_pct=args.gather_pct,
                                                              refill_min_idle=args.refill_idle))
    E_f = dm.fields["full_energy"]
    solver.engine.energy_from_temperature_dev(g, dm.fields["temperature"].ptr(), Yf.ptrs(),
                                              E_f.ptr())
    ncells = n * n * n

    # CPU-baseline / audit sample: fixed random cells of the initial state
    sample = None
    if args.cpu_sample > 0:
        rng = np.random.default_rng(args.seed)
        idx = np.sort(rng.choice(ncells, size=min(args.cpu_sample, ncells), replace=False))
        sample = dict(idx=idx, T=st["T"].cpu().numpy()[idx], rho=st["rho"].cpu().numpy()[idx],
                      Y=st["Y"].cpu().numpy()[:, idx],
                      E_f=E_f.inner(g).reshape(-1).cpu().numpy()[idx])
    del st
    # initial state, kept for the surrogate leg (which starts from the same fields)
    fields0 = {k: dm.fields[k].data.clone() for k in ("temperature", "pressure", "density",
                                                      "full_energy")}
    Y0 = [Yf.pr[s].data.clone() for s in range(S)]
    torch.cuda.synchronize()

    tuning = None
    if args.autotune:
        t_tune = time.perf_counter()
        tun = solver.autotune(args.dt)
        tuning = {"best": tun.best.as_dict(), "trials": tun.n_trials,
                  "wall_s": round(time.perf_counter() - t_tune, 2)}

    def split_step():
        stats = solver.solve_chemical_kinetics(args.dt)
        solver.apply_split_sources(args.dt)
        return stats

    # ---- warmup (the first warmup step is also the audit step) ---------------------------
    audit_out = None
    for w in range(args.warmup):
        stats = split_step()
        if w == 0 and sample is not None:
            idx = sample["idx"]
            audit_out = dict(
                Y_prod=np.stack([solver.Y_prod.pr[s].inner(g).reshape(-1)[idx].cpu().numpy()
                                 for s in range(S)]),
                T=dm.fields["temperature"].inner(g).reshape(-1)[idx].cpu().numpy(),
                Y=np.stack([Yf.pr[s].inner(g).reshape(-1)[idx].cpu().numpy() for s in range(S)]))
    torch.cuda.synchronize()

    if args.copy_calibration_mib > 0:
        nel = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)   # float64 copy: the only 'direct_copy' of this size
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---- timed region ------------------------------------------------------------------
    kernel_ms, flops, nfe, nstep_max, integ = [], [], 0, 0, 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats = split_step()
        kernel_ms.append(stats.kernel_ms)
        flops.append(flop_model.step_flops(stats))
        nfe += stats.nfe_sum
        nstep_max = max(nstep_max, stats.nstep_max)
        integ += stats.cells_integrated
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    value = ncells * args.steps / wall

    mean_ms = float(np.mean(kernel_ms))
    mean_flops = float(np.mean(flops))
    achieved_tflops = mean_flops / (mean_ms * 1e-3) / 1e12
    chunk_cells = (tuning["best"]["chunk_cells"] if tuning else
                   (1 << (args.chunk_log2 or 23)))
    nchunks = -(-ncells // chunk_cells)
    tr = profiled_traffic(ncells)
    roofline = {
        "kernel": "rates_kernel + chem_bdf_kernel (one chemistry step)",
        "bound": "fp64_valu",
        "compute": "FP64 vector ALU (the 9x9 LU and mass-action work is per-lane vector code; "
                   "MI355X's FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
        "achieved": round(achieved_tflops, 4),
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tflops / PEAK_FP64_TFLOPS, 6),
        "traffic": round(tr["bytes_per_call"]) if tr else None,
        "traffic_detail": tr,
        "avg_kernel_ms": round(mean_ms, 3),
        "dispatches_per_call": nchunks,
        "algorithmic_flops_per_launch": mean_flops,
        "hbm_boundary_GBs": round(flop_model.BYTES_BOUNDARY * ncells / (mean_ms * 1e-3) / 1e9, 2),
        "lane_utilisation": profiled_lanes(),
        "lane_utilisation_source": os.path.relpath(SQ_PROFILE, ROOT),
        "chem_bdf_kernel_resources": kernel_resources(
            "chem_bdf_kernelILb1ELb0E" if effective_setup(args, tuning)["form"] == "persistent"
            else "chem_bdf_kernelILb0ELb0E"),
    }

    # ---- CPU baseline and audit on the sample ---------------------------------------------
    cpu_baseline = None
    audit = None
    if sample is not None:
        cpu_baseline, o, (E2, Y2, T2) = cpu_leg(args, mech, sample)
        m = len(sample["idx"])
        if audit_out is not None:
            dY = np.abs(audit_out["Y"] - Y2).max()
            relT = float((np.abs(audit_out["T"] - T2) / T2).max())
            audit = {"cells": m, "max_abs_Y": float(dY), "max_rel_T": relT,
                     "bar_Y": 1e-9, "bar_T": 1e-10}

    # ---- the FP64 run's final state, against which the other legs are audited ----------
    T_ref = dm.fields["temperature"].inner(g).clone()
    Y_ref = [Yf.pr[s].inner(g).clone() for s in range(S)]

    # ---- surrogate + FP64 fallback from the same initial state ---------------------------
    sur = None
    path = args.surrogate
    if path == "auto":
        path = os.path.join(ROOT, "nrg_amd", "data", "surrogate_keromnes.npz")
        path = path if os.path.exists(path) else "none"
    if path != "none":
        fb = (args.fallback_eps or None, args.fallback_jacobian)
        sur = surrogate_leg(args, path, mech, dm, solver, fields0, Y0, T_ref, Y_ref, value, S,
                            screen=args.screen, tuning=tuning, fallback=fb)
        if args.no_network:
            base = surrogate_leg(args, path, mech, dm, solver, fields0, Y0, T_ref, Y_ref, value, S,
                                 screen="only", tuning=tuning, fallback=fb)
            sur["no_network"] = base
            sur["network_gain"] = round(sur["value"] / base["value"], 4)
        if args.reference_fallback and fb != (None, "fd"):
            # the same legs with the fallback at the reference's eps and FD Jacobian
            keep = ("ms_per_step", "speedup_vs_fp64", "fallback_ms", "fallback_nfe_per_step",
                    "fallback_integrator", "audit_vs_fp64")
            ref = {}
            for name, scr in (("learned", args.screen), ("no_network", "only")):
                if name == "no_network" and not args.no_network:
                    continue
                r = surrogate_leg(args, path, mech, dm, solver, fields0, Y0, T_ref, Y_ref, value, S,
                                  screen=scr, tuning=tuning, fallback=(None, "fd"))
                ref[name] = {k: r[k] for k in keep}
            sur["reference_fallback"] = ref
        sur["fallback_nfe_share_of_fp64"] = round(sur["fallback_nfe_per_step"] /
                                                  (nfe / args.steps), 4)

    # ---- low-Mach (FDS) coupling kernels on the FP64 run's outputs ------------------------
    lm = low_mach_leg(args, dm, solver, S) if args.low_mach else None

    # ---- table-approximated mode from the same initial state -----------------------------
    tab = table_leg(args, mech, dm, solver, fields0, Y0, T_ref, Y_ref, value, S) if args.table else None

    # ---- uneven stiffness: persistent graph queue vs one cell per lane --------------------
    uneven = None
    if args.uneven:
        del fields0, Y0, T_ref, Y_ref
        torch.cuda.empty_cache()
        uneven = uneven_leg(args, mech, tuning)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded hot-kernel field)",
        "config": {
            "workload": "FP64 Gear-BDF split step (SLATEC DDRIV3 semantics, FD Jacobian), "
                        "256^3 hot kernels in fresh H2-air, KEROMNES 9sp/24r, dt=1e-7 s",
            "cells": ncells, "dt": args.dt, "mechanism": "KEROMNES",
            "path": "fp64_integrator", "jacobian": args.jacobian, "graph": bool(args.graph),
            **effective_setup(args, tuning), "autotune": tuning,
        },
        "roofline": roofline,
        "cpu_baseline": cpu_baseline,
        "audit_vs_oracle": audit,
        "work": {"nfe_per_step": nfe / args.steps, "integrated_cells_per_step": integ / args.steps,
                 "nstep_max": nstep_max},
        "surrogate": sur,
        "table_mode": tab,
        "low_mach": lm,
        "uneven_stiffness": uneven,
    }
    print(json.dumps(line), flush=True)


def host_cpu():
    """Model name and core counts of the host this process runs on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count() or 1, "usable": usable}


def cpu_leg(args, mech, sample):
    """One split step of the sampled cells on the host cores, timed two ways:
    'reference': the reference's own SLATEC DDRIV3 (its Fortran sources built
    with flang, oracle/_ref) integrating every cell, with the oracle's rate
    constants, right-hand side and closure around it; 'port': the oracle's C
    restatement of the whole step.  Both give the same numbers bit for bit.
    Returns the baseline record, the chemistry outputs and the closed state."""
    import numpy as np

    from oracle import Oracle
    from oracle.pyoracle import REF_SO, RefSlatec
    orc = Oracle(mech)
    cpu = host_cpu()
    threads = args.cpu_threads or min(16, cpu["usable"])
    m = len(sample["idx"])

    def timed(chem):
        c0 = time.perf_counter()
        o = chem()
        closed = orc.closure_cells(args.dt, sample["rho"], o["E_f_prod"], o["Y_prod"],
                                   sample["E_f"], sample["Y"], sample["T"], nthreads=threads)
        return o, closed[:3], time.perf_counter() - c0

    o_port, closed, t_port = timed(lambda: orc.chem_cells(args.dt, sample["T"], sample["rho"],
                                                          sample["Y"], nthreads=threads))
    rec = {"unit": "cell-steps/s", "cores": threads, "cpu_model": cpu["model"],
           "nproc": cpu["nproc"], "usable_cores": cpu["usable"],
           "port_value": round(m / t_port, 1)}
    if os.path.exists(REF_SO):
        ref = RefSlatec()
        o_ref, closed_ref, t_ref = timed(lambda: ref.chem_cells(orc, args.dt, sample["T"],
                                                                sample["rho"], sample["Y"],
                                                                nthreads=threads))
        same = all(np.array_equal(o_ref[k], o_port[k]) for k in ("Y_prod", "E_f_prod", "nfe"))
        rec.update(value=round(m / t_ref, 1), kind="reference", same_as_port=bool(same),
                   sample="%d uniformly sampled cells of the initial state, one split step on %d "
                          "threads: the reference's SLATEC DDRIV3 (flang build of its sources) "
                          "per cell, with the C restatement's rate constants, right-hand side "
                          "and closure" % (m, threads))
    else:
        rec.update(value=rec["port_value"], kind="port",
                   sample="%d uniformly sampled cells of the initial state, one split step "
                          "(chemistry + closure), C restatement on %d threads" % (m, threads))
    rec["cpu_seconds"] = round((t_ref if rec["kind"] == "reference" else t_port) * threads, 1)
    return rec, o_port, closed


PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
TABLE_BYTES_INTEGRATED = 8 + 1 + 9 * 8 + (9 + 9) * 8 + 8 + 4
TABLE_BYTES_GATED = 8 + 1 + 9 * 8 + 8 + 4
HBM_PROFILE = os.path.join(ROOT, "profiles", "chem_step_256cube_hbm_bytes.txt")
SQ_PROFILE = os.path.join(ROOT, "profiles", "chem_step_256cube_sq_counters.txt")
RESOURCE_USAGE = os.path.join(ROOT, "nrg_amd", "csrc", "resource_usage.txt")


def profiled_lanes(kernel="chem_bdf_kernel"):
    """Share of a wave's 64 lanes busy per VALU cycle of `kernel`, from the
    committed SQ counter summary (LANES lines), or None."""
    if not os.path.exists(SQ_PROFILE):
        return None
    for line in open(SQ_PROFILE):
        if line.startswith("LANES"):
            _, k, v = line.rstrip("\n").split("\t")
            if k == kernel:
                return float(v)
    return None


def kernel_resources(name="chem_bdf_kernelILb1ELb0E"):
    """VGPRs, AGPRs, scratch bytes per lane and occupancy of a kernel, from the
    compiler's resource report written when the library was built, or None."""
    if not os.path.exists(RESOURCE_USAGE):
        return None
    out, cur = {}, False
    for line in open(RESOURCE_USAGE):
        if "Function Name:" in line:
            cur = name in line
            continue
        if cur:
            m = re.search(r"(VGPRs|AGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)",
                          line)
            if m:
                key = {"VGPRs": "vgprs", "AGPRs": "agprs", "ScratchSize [bytes/lane]":
                       "scratch_bytes_per_lane", "Occupancy [waves/SIMD]": "waves_per_simd"}[m.group(1)]
                out[key] = int(m.group(2))
    return out or None


def profiled_traffic(cells_per_call, kernels=("chem_bdf_kernel", "rates_kernel")):
    """HBM bytes of one chemistry call from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary: per-dispatch bytes (copy-calibrated) of the profiled
    run, whose dispatches covered 2^k cells (its command's --chunk-log2 k),
    scaled to the cells of this call."""
    if not os.path.exists(HBM_PROFILE):
        return None
    per = {}
    k_log2 = 23
    for line in open(HBM_PROFILE):
        if line.startswith("#"):
            m = re.search(r"--chunk-log2 (\d+)", line)
            if m:
                k_log2 = int(m.group(1))
        if line.startswith("TRAFFIC"):
            _, k, c, b = line.rstrip("\n").split("\t")
            for want in kernels:
                if k.startswith(want):
                    per[(want, c)] = float(b)
    if not per:
        return None
    prof_cells = 1 << k_log2
    scale = cells_per_call / prof_cells
    return {"bytes_per_call": sum(per.values()) * scale,
            "per_dispatch": {"%s %s" % kc: v for kc, v in sorted(per.items())},
            "profiled_dispatch_cells": prof_cells, "cells_per_call": cells_per_call,
            "source": os.path.relpath(HBM_PROFILE, ROOT)}


def restore_state(dm, fields0, Y0, S):
    for k, v in fields0.items():
        dm.fields[k].data.copy_(v)
    Yf = dm.fields["specie_mass_fraction"]
    for s in range(S):
        Yf.pr[s].data.copy_(Y0[s])


def audit_against(dm, T_ref, Y_ref, S, steps):
    """Final state of a leg against the FP64 run's final state (all inner cells)."""
    import torch
    g = dm.grid
    Yf = dm.fields["specie_mass_fraction"]
    T_s = dm.fields["temperature"].inner(g)
    relT = ((T_s - T_ref).abs() / T_ref).reshape(-1)
    major = (0, 1, 6)       # H2, O2, H2O
    dY = torch.stack([(Yf.pr[s].inner(g) - Y_ref[s]).abs().reshape(-1) for s in major])
    dY_all = torch.stack([(Yf.pr[s].inner(g) - Y_ref[s]).abs().reshape(-1) for s in range(S)])
    q = lambda t, p: float(torch.quantile(t[::max(1, t.numel() // 4_000_000)].float(), p))  # noqa: E731
    return {
        "steps": steps, "cells": int(T_s.numel()),
        "max_rel_T": float(relT.max()), "p99_rel_T": q(relT, 0.99),
        "max_abs_Y_major": float(dY.max()), "p99_abs_Y_major": q(dY.max(dim=0).values, 0.99),
        "max_abs_Y_any": float(dY_all.max()),
        "bar_rel_T": 1e-2, "bar_abs_Y_major": 1e-3,
    }


# algorithmic bytes per inner cell of the low-Mach coupling kernels as bench.py
# calls them (no markers set, three velocity components, all optional outputs):
# EOS predictor reads rho, 9 Y, dp_stat_dt, p_stat, 3 v, marker (121 B) and writes
# 11 fields (88 B); the species predictor with the chemistry source reads rho,
# 9 Y, 9 Y_prod, marker (153 B) and writes rho_int, 9 Y_int (80 B); the reactive
# divergence reads rho, T, 9 Y, W_mix, E_f_prod, 9 Y_prod, marker (177 B) and
# writes div_v (8 B)
LOW_MACH_BYTES = {"low_mach_eos_kernel": 121 + 88, "low_mach_species_kernel": 153 + 80,
                  "low_mach_div_kernel": 177 + 8}


def low_mach_leg(args, dm, solver, S, reps=10):
    """Time the three low-Mach coupling kernels (include/nrg_low_mach.h) on
    the 256^3 fields the FP64 run left (its Y_prod and E_f_prod are the
    chemistry sources), each with HIP events on the library's stream."""
    import numpy as np
    import torch
    e = solver.engine
    g = dm.grid
    stream = torch.cuda.ExternalStream(e.lib.nrg_chem_stream(e.h))
    shp = dm.fields["temperature"].data.shape
    new = lambda fill=0.0: torch.full(shp, fill, dtype=torch.float64, device="cuda:0")  # noqa: E731
    rho = dm.fields["density"]
    Yf = dm.fields["specie_mass_fraction"]
    p_stat = dm.fields["pressure"].data.clone()
    p_old, dp = new(), new(1.0e3)
    T, mmm, p, cp, gam, hs, vs, ei, Ef = (new() for _ in range(9))
    v = [new(1.0), new(2.0), new(3.0)]
    div, rho_i = new(), new()
    Y_i = [new() for _ in range(S)]
    ptr = lambda L: [x.data_ptr() for x in L]  # noqa: E731
    calls = {
        "low_mach_eos_kernel": lambda: e.low_mach_state_equation_dev(
            g, args.dt, True, rho.ptr(), Yf.ptrs(), dp.data_ptr(), p_stat.data_ptr(),
            p_old.data_ptr(), T.data_ptr(), mmm.data_ptr(), v=ptr(v), bc=dm.bc_markers.ptr(),
            p=p.data_ptr(), mixture_cp=cp.data_ptr(), gamma=gam.data_ptr(), h_s=hs.data_ptr(),
            v_s=vs.data_ptr(), e_i=ei.data_ptr(), E_f=Ef.data_ptr()),
        "low_mach_div_kernel": lambda: e.low_mach_chem_divergence_dev(
            g, rho.ptr(), T.data_ptr(), Yf.ptrs(), mmm.data_ptr(), solver.E_f_prod.ptr(),
            solver.Y_prod.ptrs(), div.data_ptr(), bc=dm.bc_markers.ptr()),
        "low_mach_species_kernel": lambda: e.low_mach_species_dev(
            g, args.dt, False, rho.ptr(), Yf.ptrs(), rho_i.data_ptr(), ptr(Y_i),
            Y_prod_chem=solver.Y_prod.ptrs(), bc=dm.bc_markers.ptr()),
    }
    ncells = dm.inner_cells()
    out = {}
    for name, fn in calls.items():
        fn()                                               # warm-up
        ms = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()                                           # returns with the stream drained
            b.record(stream)
            b.synchronize()
            ms.append(a.elapsed_time(b))
        t = float(np.median(ms))
        nbytes = LOW_MACH_BYTES[name] * ncells
        gbs = nbytes / (t * 1e-3) / 1e9
        out[name] = {"ms": round(t, 4), "algorithmic_bytes": nbytes, "achieved_GBs": round(gbs, 1),
                     "frac_of_hbm_peak": round(gbs / PEAK_HBM_GBS, 4)}
    return {"cells": ncells, "bound": "hbm", "peak_GBs": PEAK_HBM_GBS, "kernels": out,
            "note": "median of %d calls, HIP events on the library's stream" % reps}


def uneven_leg(args, mech, tuning, dts=(1e-8, 1e-7, 1e-6), reps=2):
    """The uneven-stiffness workload: n^3 cells, 5 % of them fresh mixture in the
    1000-1300 K ignition window (tens to hundreds of steps each), the rest burnt
    or cold (a few steps).  For each dt, the chemistry step (rates_kernel +
    chem_bdf_kernel, device time from HIP events) in two launch forms from the
    same state: the persistent work queue in a HIP graph with stiffness order and
    the main leg's autotuned queue set-up, and the plain one-cell-per-lane launch
    in natural order.  Both give the same outputs bit for bit (checked here)."""
    import numpy as np
    import torch

    from nrg_amd import capi, flop_model, inputs
    from nrg_amd.solver import ChemistryStep
    S = mech.species_number
    n = args.n ** 3
    st = inputs.ignition_mix(mech, n, seed=args.seed, device="cuda:0", frac_window=0.05)
    T, rho = st["T"], st["rho"]
    Y = st["Y"].contiguous()
    grid = capi.flat_grid(n)
    outs = [[torch.zeros(n, dtype=torch.float64, device="cuda:0") for _ in range(2 * S + 1)]
            for _ in range(2)]
    b = tuning["best"] if tuning else {}
    queued = capi.default_opts(form=capi.NRG_FORM_PERSISTENT, cell_order=capi.NRG_ORDER_STIFFNESS,
                               use_graph=1, jac_batch_pct=b.get("jac_batch_pct", 0),
                               waves_per_cu=b.get("waves_per_cu", 0),
                               chunk_log2=(int(b["chunk_cells"]).bit_length() - 1) if b else 0,
                               errtest_gather_pct=b.get("errtest_gather_pct", 0),
                               refill_min_idle=b.get("refill_min_idle", 0))
    plain = capi.default_opts(form=capi.NRG_FORM_PLAIN, cell_order=capi.NRG_ORDER_NATURAL,
                              use_graph=0, chunk_log2=(n - 1).bit_length() if n > 1 else 16)
    res = {"cells": n, "window_frac": 0.05, "reps": reps, "per_dt": []}
    engines = {"persistent_graph_queue": ChemistryStep(mech, queued),
               "one_cell_per_lane": ChemistryStep(mech, plain)}
    try:
        for dt in dts:
            row = {"dt": dt}
            for slot, (name, eng) in enumerate(engines.items()):
                o = outs[slot]
                call = lambda: eng.step_dev(  # noqa: E731
                    grid, dt, T.data_ptr(), rho.data_ptr(), [Y[s].data_ptr() for s in range(S)],
                    [o[s].data_ptr() for s in range(S)], [o[S + s].data_ptr() for s in range(S)],
                    o[2 * S].data_ptr())
                call()                                   # warm-up (and step-count history)
                recs = [call() for _ in range(reps)]
                ms = float(np.mean([r.kernel_ms for r in recs]))
                r = recs[-1]
                fl = flop_model.step_flops(r)
                row[name] = {
                    "kernel_ms": round(ms, 3),
                    "cell_steps_per_s": round(n / (ms * 1e-3), 1),
                    "fp64_tflops": round(fl / (ms * 1e-3) / 1e12, 4),
                    "frac_fp64_peak": round(fl / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 5),
                    "integrated_cells": int(r.cells_integrated),
                    "nfe_per_integrated_cell": round(r.nfe_sum / max(1, r.cells_integrated), 2),
                    "nstep_max": int(r.nstep_max)}
            row["queue_speedup"] = round(row["one_cell_per_lane"]["kernel_ms"] /
                                         row["persistent_graph_queue"]["kernel_ms"], 3)
            row["identical_outputs"] = bool(all(torch.equal(a, c) for a, c in zip(*outs)))
            res["per_dt"].append(row)
    finally:
        for eng in engines.values():
            eng.close()
    return res


def table_leg(args, mech, dm, solver, fields0, Y0, T_ref, Y_ref, fp64_value, S):
    """NRG's table_approximated mode as a non-learned baseline: a table dC(T)
    written from this library's FP64 step on a 1-D 15 % H2-air flame line at
    1 atm (the reference's table is named for that mixture), then the table
    split step (lookup + closure) over the same steps from the same initial
    state, audited against the FP64 run."""
    import numpy as np
    import torch

    from nrg_amd import chem_table as ct
    e = solver.engine
    T_t, dC_t = ct.build_table(e, mech, args.dt, n=args.table_line)
    e.set_table(T_t, dC_t)
    restore_state(dm, fields0, Y0, S)
    mode0 = solver.ODE_solver
    solver.ODE_solver = "table_approximated"

    def step():
        st = solver.solve_chemical_kinetics(args.dt)
        solver.apply_split_sources(args.dt)
        return st

    try:
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        recs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            recs.append(step())
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    finally:
        solver.ODE_solver = mode0
    ncells = dm.inner_cells()
    value = ncells * args.steps / wall
    k_ms = float(np.mean([r.kernel_ms for r in recs]))
    integ = float(np.mean([r.cells_integrated for r in recs]))
    # table_kernel's algorithmic bytes: an integrated cell reads T, marker and 9 Y
    # (81 B) and writes 9 Y_prod, 9 Y_prod2, E_f_prod and its status (156 B); a
    # gated one reads T and marker and writes 9 Y_prod, E_f_prod, status (93 B)
    nbytes = integ * TABLE_BYTES_INTEGRATED + (ncells - integ) * TABLE_BYTES_GATED
    gbs = nbytes / (k_ms * 1e-3) / 1e9
    return {
        "value": round(value, 1),
        "unit": "cell-steps/s",
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "speedup_vs_fp64": round(value / fp64_value, 3),
        "dtype": "f64",
        "table_rows": int(T_t.size), "table_T_range": [float(T_t[0]), float(T_t[-1])],
        "table_source": "FP64 step on a %d-cell 15%% H2-air flame line at 1 atm" % args.table_line,
        "kernel_ms": round(k_ms, 4),
        "roofline": {"kernel": "table_kernel", "bound": "hbm", "achieved": round(gbs, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
                     "algorithmic_bytes_per_launch": nbytes},
        "audit_vs_fp64": audit_against(dm, T_ref, Y_ref, S, args.warmup + args.steps),
    }


def surrogate_leg(args, path, mech, dm, solver, fields0, Y0, T_ref, Y_ref, fp64_value, S,
                  screen="rates", tuning=None, fallback=(None, "fd")):
    """Time the learned split step (explicit-rate screen, network, gate, FP64
    fallback, closure) over the same number of steps from the same initial
    state, and audit its final state against the FP64 integrator's.  With
    screen='only' the network is not used: every cell the explicit-rate screen
    does not freeze goes to the FP64 integrator (the method without the
    network).  fallback = (eps, jacobian) of the FP64 fallback: (None, 'fd')
    runs it on the FP64 leg's own handle (the reference's eps 1e-9), anything
    else on a handle of its own with the FP64 leg's queue set-up."""
    import numpy as np
    import torch

    from nrg_amd import capi
    from nrg_amd import surrogate as sg
    restore_state(dm, fields0, Y0, S)
    z = np.load(path, allow_pickle=False)
    mlp, norm = sg.load(path) if screen != "only" else (None, None)
    gkw = {}
    if args.gate_energy_tol is not None:
        gkw["energy_tol"] = args.gate_energy_tol
    if args.gate_elem_tol is not None:
        gkw["elem_tol"] = args.gate_elem_tol
    if args.gate_elem_floor is not None:
        gkw["elem_floor"] = args.gate_elem_floor
    for k, v in (("dT_max", args.gate_dt_max), ("dT_freeze", args.gate_dt_freeze),
                 ("dY_freeze", args.gate_dy_freeze), ("rate_tol", args.gate_rate_tol)):
        if v is not None:
            gkw[k] = v
    for k in ("dT_max", "dT_freeze", "dY_freeze", "energy_tol", "elem_tol", "elem_floor", "rate_tol"):
        if k not in gkw and ("gate_" + k) in z.files:
            gkw[k] = float(z["gate_" + k])
    scr = {"none": capi.NRG_SUR_SCREEN_NONE, "rates": capi.NRG_SUR_SCREEN_RATES,
           "only": capi.NRG_SUR_SCREEN_ONLY}[screen]
    gate = sg.gate_struct(mech, screen=scr, **gkw)
    fb_engine = None
    fb_eps, fb_jac = fallback
    if fb_eps is not None or fb_jac != "fd":
        # a handle of its own for the fallback, with the FP64 leg's queue set-up
        from nrg_amd.solver import ChemistryStep
        b = tuning["best"] if tuning else {}
        o = capi.default_opts(
            form=b.get("form", capi.NRG_FORM_PERSISTENT if args.form == "persistent" else capi.NRG_FORM_PLAIN),
            cell_order=b.get("cell_order", capi.NRG_ORDER_STIFFNESS if args.order == "stiffness"
                             else capi.NRG_ORDER_NATURAL),
            jac_batch_pct=b.get("jac_batch_pct", args.jac_pct), waves_per_cu=b.get("waves_per_cu", 0),
            chunk_log2=(int(b["chunk_cells"]).bit_length() - 1) if b else args.chunk_log2,
            errtest_gather_pct=b.get("errtest_gather_pct", args.gather_pct),
            refill_min_idle=b.get("refill_min_idle", args.refill_idle), use_graph=args.graph)
        if fb_eps is not None:
            o.eps = fb_eps
        if fb_jac == "analytic":
            # an analytic Jacobian costs no right-hand sides: batch it at 10 % of the
            # pending lanes (measured on the fallback: 100 % 7.2 ms, 50 % 6.8, 10 % 6.6)
            o.jacobian = capi.NRG_JAC_ANALYTIC
            o.jac_batch_pct = 10
        fb_engine = ChemistryStep(mech, o)
    chem = sg.SurrogateChemistry(solver, mlp, norm, gate, engine=fb_engine)

    def step():
        st = chem.solve_chemical_kinetics(args.dt)
        chem.apply_split_sources(args.dt)
        return st

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    recs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        recs.append(step())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ncells = dm.inner_cells()
    value = ncells * args.steps / wall
    mean = lambda f: float(np.mean([f(r) for r in recs]))
    infer_ms = mean(lambda r: r.infer_ms)
    net_cells = mean(lambda r: r.cells_network)
    flops = 2 * (12 * 256 + 3 * 256 * 256 + 256 * 10) * net_cells   # 404,480 per inferred cell
    tflops = flops / (infer_ms * 1e-3) / 1e12 if infer_ms > 0 else 0.0
    res = {
        "value": round(value, 1),
        "unit": "cell-steps/s",
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "speedup_vs_fp64": round(value / fp64_value, 3),
        "screen": screen,
        "dtype": ("f64 screen and fallback (no network)" if screen == "only" else
                  "bf16 (fp32 accumulate) network, f64 screen, gate and fallback"),
        "network_frac": round(net_cells / ncells, 6),
        "fallback_frac": round(mean(lambda r: r.cells_fallback) / ncells, 6),
        "accepted_frac": round(mean(lambda r: r.cells_accepted) / ncells, 6),
        "frozen_frac": round(mean(lambda r: r.cells_frozen) / ncells, 6),
        "fallback_nfe_per_step": mean(lambda r: r.fallback.nfe_sum),
        "fallback_nstep_max": max(r.fallback.nstep_max for r in recs),
        "rejects_per_step": {k: mean(lambda r, k=k: getattr(r, k)) for k in
                             ("reject_domain", "reject_confidence", "reject_elements",
                              "reject_negative", "reject_energy", "reject_rate")},
        "gate": gkw,
        "screen_ms": round(mean(lambda r: r.screen_ms), 3),
        "screen_fp64_frac": round(mean(lambda r: r.cells_screen_fp64) / ncells, 6),
        "infer_ms": round(infer_ms, 3), "gate_ms": round(mean(lambda r: r.gate_ms), 3),
        "fallback_ms": round(mean(lambda r: r.fallback_ms), 3),
        "fallback_integrator": {"eps": (fb_engine or solver.engine).opts.eps,
                                "jacobian": "analytic" if (fb_engine or solver.engine).opts.jacobian
                                == capi.NRG_JAC_ANALYTIC else "fd"},
        "audit_vs_fp64": audit_against(dm, T_ref, Y_ref, S, args.warmup + args.steps),
    }
    if fb_engine is not None:
        fb_engine.close()
    if mlp is not None:
        res["roofline_infer"] = {
            "bound": "mfma", "achieved": round(tflops, 2), "peak": PEAK_BF16_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tflops / PEAK_BF16_TFLOPS, 4),
            "kernel": "featurize_kernel + mlp_infer_kernel (404,480 FLOP per cell the network "
                      "runs on)"}
        res["model"] = os.path.relpath(path, ROOT)
        mlp.close()
    return res


if __name__ == "__main__":
    main()
