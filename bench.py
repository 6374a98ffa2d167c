"""Benchmark: batched E+F throughput (atoms/s) of the Behler-Parrinello hot path on one MI355X.

Workload (one step = NNP::compute of the whole batch: neighbour-list build, symmetry functions,
networks, adjoint forces, FP64 energy/virial reductions): 2 elements (50/50), 1000 atoms on a
jittered (0.15 A) bcc lattice at 0.085 atoms/A^3, periodic cell, cutoff 6.0 A, 16 G2 per neighbour
species + 16 G4 per species pair = 80 inputs, 80-32-32-1 tanh networks, FP32 arithmetic with FP64
reductions.  Inputs are resident in HBM before the timed region.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atoms/s E+F eval and train structs/s: batched adjoint vs plain FP64 GPU"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def work_counts(model, strucs, off, j, typ):
    """Pairs P, triplets T, radial evaluations E_R and angular evaluations E_A of one E+F pass,
    counted from the exported neighbour list (skin 0: every listed pair is inside rc)."""
    from ace_amd.model import tri_index
    nt = model.ntypes
    base = np.repeat(np.concatenate([[0], np.cumsum([s.natoms for s in strucs])[:-1]]),
                     [s.natoms for s in strucs])
    P = T = ER = EA = 0
    for i in range(len(typ)):
        I = typ[i]
        n = int(off[i + 1] - off[i])
        cnt = np.bincount(typ[base[i] + j[off[i]:off[i + 1]]], minlength=nt)
        P += n
        T += n * (n - 1) // 2
        ER += sum(int(cnt[J]) * len(model.radial[I][J].rs) for J in range(nt))
        for J in range(nt):
            for K in range(J, nt):
                tr = cnt[J] * (cnt[J] - 1) // 2 if J == K else cnt[J] * cnt[K]
                EA += int(tr) * len(model.angular[I][tri_index(J, K)].eta)
    return P, T, ER, EA


def mlp_flops(model, typ):
    f = 0
    for I in range(model.ntypes):
        w = model.nets[I].widths
        f += 4 * np.count_nonzero(typ == I) * sum(w[l] * w[l + 1] for l in range(len(w) - 1))
    return f


def cpu_baseline(model, s, samples):
    """The FP64 oracle (reference loop structure, brute-force neighbour list) on one host thread."""
    from oracle import oracle as O
    orc = O.Oracle(model)
    t0 = time.perf_counter()
    out = None
    for _ in range(samples):
        out = orc.compute(s)
    dt = (time.perf_counter() - t0) / samples
    return s.natoms / dt, out


def train_bench(steps):
    """Force-matching training throughput (structures/s): 64 structures of 32-1024 atoms (log-uniform),
    3 elements, random packing at 0.08 atoms/A^3, 24 G2 + 32 G4, 56-64-64-1; targets from a seeded teacher
    network of the same architecture (evaluated on the GPU in FP64-basis mode).  One step = loss and
    gradient (energy MSE + 0.1 force MSE, d(F)/dW by forward-over-reverse) replayed from a captured HIP graph,
    then the ADAM update on the GPU (FP64 master parameters, device layouts refreshed in place)."""
    from ace_amd import PREC_FP64_BASIS, NNPDevice, configs
    from ace_amd.training import Trainer
    model, teacher, strucs = configs.config3()
    tdev = NNPDevice(teacher)
    tdev.upload(strucs)
    tdev.build_neighbors()
    tdev.set_precision(PREC_FP64_BASIS)
    G = tdev.compute_symm()
    configs.set_input_scaling(teacher, strucs, G)
    tdev.sync_model_scaling()
    E, F, _ = tdev.compute(PREC_FP64_BASIS)
    off = np.concatenate([[0], np.cumsum([s.natoms for s in strucs])])
    for k, s in enumerate(strucs):
        s.energy, s.forces = E[k], F[off[k]:off[k + 1]]
    tdev.close()
    for I in range(model.ntypes):
        model.nets[I].inpw, model.nets[I].inpb = teacher.nets[I].inpw, teacher.nets[I].inpb
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    tr = Trainer(dev, model, strucs, gamma=1e-3, w_force=0.1, graph=True, device_opt=True)
    for _ in range(3):
        tr.step()
    dev.sync()
    t0 = time.perf_counter()
    first = last = None
    for k in range(steps):
        out = tr.step()
        first = out if k == 0 else first
        last = out
    dev.sync()
    dt = (time.perf_counter() - t0) / steps
    dev.close()
    return {"metric": "train structs/s", "value": len(strucs) / dt, "unit": "structures/s",
            "ms_per_step": dt * 1e3, "structures": len(strucs), "atoms": int(off[-1]), "steps": steps, "graph": True,
            "optimizer": "ADAM on the GPU",
            "loss_first": float(first[0]), "loss_last": float(last[0]), "dtype": "fp32",
            "workload": "64 structures, N log-uniform [32,1024], 3 elements, 56-64-64-1, energy + 0.1 force MSE"}


def md_bench(steps):
    """Molecular dynamics throughput: 108,000-atom fcc cell (30^3 conventional cells, 0.085 atoms/A^3), 2
    elements 70/30 with masses 27/16 amu, Maxwell-Boltzmann velocities at 300 K, list radius 6 + 1 A rebuilt
    every 20 steps, 80-32-32-1 networks (output layers scaled to an rms force component of 0.1 eV/A),
    velocity Verlet with dt = 1 fs, steps between rebuilds replayed from HIP graphs.  The FP32 run's energy
    drift is reported next to the same trajectory length with FP64 symmetry functions."""
    from ace_amd import PREC_FP32, PREC_FP64_BASIS, NNPDevice, configs
    model, s, masses = configs.config4()
    v0 = configs.maxwell_boltzmann(masses[s.type], 300.0, np.random.default_rng(0xACE4))
    dev = NNPDevice(model)
    dev.upload([s])
    dev.build_neighbors()
    configs.set_input_scaling(model, [s], dev.compute_symm())
    dev.sync_model_scaling()
    _, F, _ = dev.compute()
    configs.calibrate_forces(model, F, rms=0.1)
    for I in range(model.ntypes):
        dev.set_params(I, model.nets[I].params)
    out = {}
    for prec, name in ((PREC_FP32, "fp32"), (PREC_FP64_BASIS, "fp64_basis")):
        dev.update_positions(s.pos)
        dev.md_init(masses, v0)
        dev.md_run(20, precision=prec)                 # warm-up: list sizing, graph capture
        dev.update_positions(s.pos)
        dev.md_init(masses, v0)
        dev.sync()
        t0 = time.perf_counter()
        e = dev.md_run(steps, precision=prec)
        dt = time.perf_counter() - t0
        tot = e.sum(axis=1)
        out[name] = {"steps_per_s": steps / dt, "ms_per_step": dt / steps * 1e3,
                     "atom_steps_per_s": s.natoms * steps / dt,
                     "temperature_end_K": float(2 * e[-1, 1] / (3 * s.natoms * configs.KB_EV)),
                     "drift": (tot - tot[0]) / s.natoms}
    dev.close()
    d32, d64 = out["fp32"].pop("drift"), out["fp64_basis"].pop("drift")
    for name, d in (("fp32", d32), ("fp64_basis", d64)):
        out[name]["max_abs_energy_drift_per_atom"] = float(np.abs(d).max())
        out[name]["end_energy_drift_per_atom"] = float(d[-1])
    return {"metric": "MD steps/s", "value": out["fp32"]["steps_per_s"], "unit": "steps/s", "atoms": int(s.natoms),
            "steps": steps, "dt_fs": 1.0, "rebuild_every": 20, "skin_A": 1.0, "dtype": "fp32", "graph": True,
            "fp32": out["fp32"], "fp64_basis": out["fp64_basis"],
            "max_abs_drift_difference_fp32_vs_fp64_basis": float(np.abs(d32 - d64).max()),
            "workload": "fcc 108,000 atoms, 2 elements 70/30, 300 K, velocity Verlet 1 fs, list 7 A rebuilt every 20 steps"}


def c5_bench(steps):
    """Mixed-precision batched inference: 256 structures of 256-8192 atoms (log-uniform, 568,089 atoms), 4
    elements, rc 6.5 A, 32 G2 + 96 G4, 128-128-64-1; FP32, and the first layer on BF16 matrix cores.
    Per-structure errors against the FP64 oracle for the two smallest structures."""
    from ace_amd import PREC_BF16L1, PREC_FP32, PREC_FP64_BASIS, NNPDevice, configs
    from oracle import oracle as O
    model, strucs = configs.config5()
    N = sum(x.natoms for x in strucs)
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    dev.set_precision(PREC_FP64_BASIS)
    configs.set_input_scaling(model, strucs, dev.compute_symm())
    dev.sync_model_scaling()
    dev.set_precision(PREC_FP32)
    res = {"metric": "atoms/s E+F", "unit": "atoms/s", "structures": len(strucs), "atoms": int(N), "steps": steps,
           "workload": "256 structures, N log-uniform [256,8192], 4 elements, rc 6.5, 128-128-64-1"}
    off = np.concatenate([[0], np.cumsum([x.natoms for x in strucs])])
    small = [int(k) for k in np.argsort([x.natoms for x in strucs])[:2]]
    orc = O.Oracle(model)
    ref = {k: orc.compute(strucs[k]) for k in small}
    for prec, name in ((PREC_FP32, "fp32"), (PREC_BF16L1, "bf16_layer1")):
        dev.compute_step(prec)
        dev.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            dev.compute_step(prec)
        dev.sync()
        ms = (time.perf_counter() - t0) / steps * 1e3
        E, F, _ = dev.compute(prec)
        err = [{"structure": k, "atoms": strucs[k].natoms,
                "energy_per_atom": abs(E[k] - ref[k][0]) / strucs[k].natoms,
                "force_max_abs": float(np.max(np.abs(F[off[k]:off[k + 1]] - ref[k][1])))} for k in small]
        res[name] = {"value": N / (ms * 1e-3), "ms_per_step": ms, "error_vs_fp64_oracle": err}
    dev.close()
    res["value"] = res["fp32"]["value"]
    return res


def copy_calibration(mib):
    import torch
    a = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    log(f"copy calibration: 5 x neg over {mib} MiB (read {mib} MiB, write {mib} MiB per launch)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-samples", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-steps", type=int, default=20)
    ap.add_argument("--md-steps", type=int, default=1000)
    ap.add_argument("--c5-steps", type=int, default=5)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 elementwise copies of MIB MiB (reads MIB, writes MIB): the reference point "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU path (structures over GPUs for training and "
              "batched inference, spatial bricks with a ghost layer for MD) is not built", file=sys.stderr)
        sys.exit(2)

    from ace_amd import PREC_FP32, PREC_FP64_BASIS, NNPDevice, configs

    model, [s] = configs.config2()
    dev = NNPDevice(model)
    dev.upload([s])
    dev.build_neighbors()
    # PreBias MEAN / PreScale DEV (nnptefr.cpp:786-880) from FP64 symmetry functions, as the reference
    # computes its input statistics in FP64
    dev.set_precision(PREC_FP64_BASIS)
    G = dev.compute_symm()
    dev.set_precision(PREC_FP32)
    configs.set_input_scaling(model, [s], G)
    dev.sync_model_scaling()
    tune = {"fp32": dev.autotune(PREC_FP32), "fp64_basis": dev.autotune(PREC_FP64_BASIS)}

    def run(prec, steps, events):
        """K steps of NNP::compute (neighbour build + energy/forces, one ace_compute call each); returns ms per
        step and per-kernel event times."""
        dev.profile(events)
        dev.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            dev.compute_step(prec)
        dev.sync()
        t1 = time.perf_counter()
        kern = {}
        if events:
            sfx = "_f64" if prec == PREC_FP64_BASIS else ""
            for k in ("nb_bin", "nb_search", "basis_fused" + sfx, "basis_symm" + sfx, "mlp_energy_grad",
                      "basis_adjoint" + sfx):
                v = dev.profile_read(k)
                if v[1]:
                    kern[k.replace(sfx, "") if sfx else k] = v
            dev.profile(False)
        return (t1 - t0) / steps * 1e3, kern

    for prec in (PREC_FP32, PREC_FP64_BASIS):
        run(prec, args.warmup, False)
    # timed region: K steps, no per-kernel events
    ms, _ = run(PREC_FP32, args.steps, False)
    value = s.natoms / (ms * 1e-3)
    # the same K steps again with HIP events around every kernel (per-kernel durations)
    ms_ev, kern = run(PREC_FP32, args.steps, True)
    ms64, _ = run(PREC_FP64_BASIS, args.steps, False)
    _, kern64 = run(PREC_FP64_BASIS, args.steps, True)

    # algorithmic work of the dominant kernel
    off, j, img = dev.neighbors()
    P, T, ER, EA = work_counts(model, [s], off, j, s.type)
    flops = {
        "basis_symm": 10 * P + 6 * ER + 15 * T + 7 * EA,
        "basis_adjoint": 22 * P + 6 * ER + 45 * T + 7 * EA,
        "mlp_energy_grad": mlp_flops(model, s.type),
    }
    # the fused kernel does all three: symmetry functions, networks and the adjoint
    flops["basis_fused"] = flops["basis_symm"] + flops["basis_adjoint"] + flops["mlp_energy_grad"]
    avg = {k: v[0] / v[1] for k, v in kern.items()}
    dom = max((k for k in ("basis_fused", "basis_symm", "basis_adjoint") if k in avg), key=lambda k: avg[k])
    achieved = flops[dom] / (avg[dom] * 1e-3) / 1e12
    traffic = None   # HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary, when present
    tf = os.path.join(ROOT, "profiles", f"{dom}_c2_traffic.txt")
    if os.path.exists(tf):
        for line in open(tf):
            if line.startswith("bytes_per_launch:"):
                traffic = int(float(line.split(":")[1]))

    # plain unbatched FP64 GPU port (the 5x denominator), same step shape
    for _ in range(3):
        dev.build_neighbors()
        dev.plain_compute()
    nplain = max(3, args.steps // 20)
    t0 = time.perf_counter()
    for _ in range(nplain):
        dev.build_neighbors()
        dev.plain_compute()
    plain_ms = (time.perf_counter() - t0) / nplain * 1e3
    plain_value = s.natoms / (plain_ms * 1e-3)

    E, F, _ = dev.compute(PREC_FP32)
    E64, F64, _ = dev.compute(PREC_FP64_BASIS)
    train = train_bench(args.train_steps) if args.train_steps > 0 else None
    md = md_bench(args.md_steps) if args.md_steps > 0 else None
    c5 = c5_bench(args.c5_steps) if args.c5_steps > 0 else None
    avg64 = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in kern64.items()}
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded jittered bcc, random-init networks)",
        "config": {"workload": "bcc 1000 atoms, 2 elements 50/50, rc 6.0 A, 80 inputs (16 G2 + 16 G4 per "
                               "species pair), 80-32-32-1 tanh; step = neighbour build + E + F + virial",
                   "global_batch": 1, "atoms": int(s.natoms), "pairs": int(P), "triplets": int(T)},
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "flops_per_launch": int(flops[dom]), "avg_launch_ms": avg[dom],
                     "note": "FP32 vector-ALU kernel priced against the FP32 dense peak"},
        "kernels_avg_us": {k: round(v * 1e3, 2) for k, v in avg.items()},
        "autotune": tune,
        "ms_per_step_with_kernel_events": ms_ev,
        "fp64_basis": {"value": s.natoms / (ms64 * 1e-3), "unit": "atoms/s", "ms_per_step": ms64,
                       "kernels_avg_us": {k: round(v * 1e3, 2) for k, v in avg64.items()}},
        "plain_fp64_port": {"value": plain_value, "unit": "atoms/s", "ms_per_step": plain_ms,
                            "speedup_batched_vs_plain": value / plain_value},
        "train": train,
        "md": md,
        "mixed_precision_inference": c5,
    }
    if not args.no_cpu:
        cpu_value, (E_ref, F_ref, _) = cpu_baseline(model, s, args.cpu_samples)
        res["cpu_baseline"] = {"value": cpu_value, "unit": "atoms/s", "cores": 1, "kind": "port",
                               "sample": f"{args.cpu_samples} full E+F evaluations of the same 1000-atom structure "
                                         "(brute-force neighbour list, compute_symm, compute_force), FP64, 1 thread"}
        res["speedup_vs_cpu"] = value / cpu_value
        res["error_vs_fp64_oracle"] = {"energy_per_atom": abs(E[0] - E_ref) / s.natoms,
                                       "force_max_abs": float(np.max(np.abs(F - F_ref)))}
        res["fp64_basis"]["error_vs_fp64_oracle"] = {"energy_per_atom": abs(E64[0] - E_ref) / s.natoms,
                                                     "force_max_abs": float(np.max(np.abs(F64 - F_ref)))}
    dev.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
