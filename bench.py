"""Benchmark of the Behler-Parrinello energy/force/training hot path on one MI355X.  Prints one JSON line.

Headline workload (`value`): the mixed-precision batched inference batch -- 256 structures of 256-8192
atoms (log-uniform; 568,089 atoms), 4 elements, rc 6.5 A, 32 G2 + 96 G4 per element network, 128-128-64-1
tanh, FP32 arithmetic with FP64 geometry, sums and reductions.  One step = NNP::compute of the whole batch:
neighbour-list build, symmetry functions, networks, adjoint forces, per-structure FP64 energies and virials
(ace_compute), inputs resident in HBM, results left in HBM.  The same batch with the first network layer
on BF16 matrix cores is a sub-line.  Secondary lines: the 1000-atom bcc structure (FP32 and FP64 symmetry
functions), the 64-structure force-matching training step, the 108,000-atom MD run, the plain unbatched
FP64 GPU port, and the reference algorithm on the host's cores (the FP64 oracle, structures block-sharded
over worker processes as thread::Dist does over MPI ranks).
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
CPU_WORKERS_MAX = 16        # host cores of one GPU's share on the GPU boxes


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---- algorithmic work -------------------------------------------------------------------------------

def work_counts(model, strucs, off, j):
    """Pairs P, triplets T, radial evaluations E_R and angular evaluations E_A of one E+F pass, from the
    exported neighbour list (skin 0: every listed pair is inside rc); vectorised over atoms."""
    from ace_amd.model import tri_index
    nt = model.ntypes
    nat = np.array([s.natoms for s in strucs])
    typ = np.concatenate([s.type for s in strucs]).astype(np.int64)
    base = np.repeat(np.concatenate([[0], np.cumsum(nat)[:-1]]), nat)
    n_i = np.diff(off)
    atom_of_pair = np.repeat(np.arange(len(typ)), n_i)
    spec = typ[base[atom_of_pair] + j.astype(np.int64)]
    cnt = np.bincount(atom_of_pair * nt + spec, minlength=len(typ) * nt).reshape(len(typ), nt).astype(np.int64)
    P = int(n_i.sum())
    T = int((n_i * (n_i - 1) // 2).sum())
    nR = np.array([[len(model.radial[I][J].rs) for J in range(nt)] for I in range(nt)])
    ER = int((cnt * nR[typ]).sum())
    EA = 0
    for J in range(nt):
        for K in range(J, nt):
            pairs = cnt[:, J] * (cnt[:, J] - 1) // 2 if J == K else cnt[:, J] * cnt[:, K]
            nA = np.array([len(model.angular[I][tri_index(J, K)].eta) for I in range(nt)])
            EA += int((pairs * nA[typ]).sum())
    return P, T, ER, EA


def mlp_flops(model, typ):
    f = 0
    for I in range(model.ntypes):
        w = model.nets[I].widths
        f += 4 * int(np.count_nonzero(typ == I)) * sum(w[l] * w[l + 1] for l in range(len(w) - 1))
    return f


def kernel_flops(model, strucs, dev):
    """SURVEY's FLOP model per kernel of one E+F pass over the uploaded batch (adds, muls, compares)."""
    off, j, _ = dev.neighbors()
    P, T, ER, EA = work_counts(model, strucs, off, j)
    typ = np.concatenate([s.type for s in strucs])
    fl = {"basis_symm": 10 * P + 6 * ER + 15 * T + 7 * EA, "basis_adjoint": 22 * P + 6 * ER + 45 * T + 7 * EA,
          "mlp_energy_grad": mlp_flops(model, typ)}
    fl["basis_fused"] = fl["basis_symm"] + fl["basis_adjoint"] + fl["mlp_energy_grad"]
    fl["mlp_energy_grad_bf16"] = fl["mlp_energy_grad"]
    return fl, {"pairs": P, "triplets": T, "radial_evals": ER, "angular_evals": EA}


# ---- timing helpers ------------------------------------------------------------------------------------

def timed_steps(dev, prec, steps):
    """K steps of ace_compute (neighbour build + energies/forces/virials), one synchronisation before the
    first and after each step: (ms per step over the K steps, per-step milliseconds)."""
    dev.sync()
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        dev.compute_step(prec)
        dev.sync()
        per.append((time.perf_counter() - a) * 1e3)
    total = (time.perf_counter() - t0) * 1e3
    return total / steps, per


KERNELS = ("nb_bin", "nb_search", "basis_fused", "basis_symm", "mlp_energy_grad", "mlp_energy_grad_bf16",
           "basis_adjoint", "force_gather", "finalize")


def kernel_times(dev, prec, steps):
    """Average HIP-event duration (ms) of each kernel over `steps` more steps, on the library's stream."""
    sfx = "_f64" if prec == 2 else ""
    dev.profile(True)
    for _ in range(steps):
        dev.compute_step(prec)
    dev.sync()
    out = {}
    for k in KERNELS:
        name = k + sfx if k.startswith("basis") else k
        t, n = dev.profile_read(name)
        if n:
            out[k] = t / n
    dev.profile(False)
    return out


def roofline(flops, kern, workload):
    dom = max((k for k in ("basis_fused", "basis_symm", "basis_adjoint") if k in kern), key=lambda k: kern[k])
    achieved = flops[dom] / (kern[dom] * 1e-3) / 1e12
    traffic = None     # HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary, when present
    tf = os.path.join(ROOT, "profiles", f"{dom}_{workload}_traffic.txt")
    if os.path.exists(tf):
        for line in open(tf):
            if line.startswith("bytes_per_launch:"):
                traffic = int(float(line.split(":")[1]))
    return {"bound": "valu", "kernel": dom, "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic, "flops_per_launch": int(flops[dom]),
            "avg_launch_ms": kern[dom],
            "note": "FP32/FP64 vector-ALU kernel (no matrix cores) priced against MI355X's FP32 peak; the build "
                    "disables packed FP32, so half that peak is this kernel's vector-ALU ceiling"}


# ---- the reference algorithm on the host's cores (cpu_baseline leg: the FP64 oracle) --------------------

def _oracle_worker(args):
    """One worker = one MPI rank of the reference: its contiguous block of the sample's structures
    (thread::Dist::init, dist.cpp:29-41), each through NeighborList::build, compute_symm and compute_force;
    returns (structure index, atoms, E, F) per structure and the worker's busy time."""
    cfg, block = args
    from oracle import oracle as O
    model, strucs = cfg_load(cfg)
    orc = O.Oracle(model)
    done = []
    t0 = time.perf_counter()
    for k in block:
        E, F, _ = orc.compute(strucs[k])
        done.append((k, strucs[k].natoms, E, F))
    return done, time.perf_counter() - t0


def cfg_load(cfg):
    from ace_amd import configs
    name, scaling = cfg
    model, strucs = configs.config5() if name == "c5" else configs.config2()
    o = 0
    for I in range(model.ntypes):
        K = model.input_width(I)
        model.nets[I].inpw, model.nets[I].inpb = scaling[0][o:o + K], scaling[1][o:o + K]
        o += K
    return model, strucs


def dist_blocks(items, workers):
    """thread::Dist::init (dist.cpp:29-41): contiguous blocks, the first len % workers one longer."""
    q, r = divmod(len(items), workers)
    out, s = [], 0
    for w in range(workers):
        c = q + (1 if w < r else 0)
        out.append(list(items[s:s + c]))
        s += c
    return out


def cpu_baseline_sharded(cfg, sample, note):
    """The FP64 oracle on the host's cores (capped at one GPU's share), the sample's structures
    block-sharded over worker processes; atoms/s = atoms of the sample / the longest worker's busy time."""
    from multiprocessing import get_context
    ncpu = os.cpu_count() or 1
    workers = max(1, min(ncpu, CPU_WORKERS_MAX, len(sample)))
    blocks = dist_blocks(sample, workers)
    t0 = time.perf_counter()
    with get_context("spawn").Pool(workers) as pool:
        res = pool.map(_oracle_worker, [(cfg, b) for b in blocks])
    wall = time.perf_counter() - t0
    done = [d for r in res for d in r[0]]
    busy = max(r[1] for r in res)
    atoms = sum(d[1] for d in done)
    return {"value": atoms / busy, "unit": "atoms/s", "cores": workers, "kind": "port", "nproc": ncpu,
            "structures": len(done), "atoms": int(atoms), "busy_s": busy, "wall_s": wall,
            "sample": f"the FP64 oracle (reference loops: brute-force NeighborList::build, compute_symm, "
                      f"compute_force) on {note}, block-sharded over {workers} worker processes "
                      f"(thread::Dist::init); atoms / the longest worker's busy time"}, {d[0]: (d[2], d[3]) for d in done}


def cpu_baseline_one(model, s, samples):
    """The FP64 oracle on one host thread (the reference has no intra-structure parallelism)."""
    from oracle import oracle as O
    orc = O.Oracle(model)
    t0 = time.perf_counter()
    out = None
    for _ in range(samples):
        out = orc.compute(s)
    dt = (time.perf_counter() - t0) / samples
    return s.natoms / dt, out


# ---- workloads --------------------------------------------------------------------------------------------

def c5_headline(args):
    from ace_amd import PREC_BF16L1, PREC_FP32, PREC_FP64_BASIS, NNPDevice, configs
    model, strucs = configs.config5()
    N = sum(x.natoms for x in strucs)
    log(f"c5: {len(strucs)} structures, {N} atoms")
    dev = NNPDevice(model)
    dev.upload(strucs)
    dev.build_neighbors()
    # PreBias MEAN / PreScale DEV (nnptefr.cpp:786-880) over the batch, from FP64 symmetry functions
    dev.set_precision(PREC_FP64_BASIS)
    configs.set_input_scaling(model, strucs, dev.compute_symm())
    dev.sync_model_scaling()
    dev.set_precision(PREC_FP32)
    tune = {"fp32": dev.autotune(PREC_FP32), "bf16_layer1": dev.autotune(PREC_BF16L1)}
    for _ in range(args.warmup):
        dev.compute_step(PREC_FP32)
    ms, per = timed_steps(dev, PREC_FP32, args.steps)
    kern = kernel_times(dev, PREC_FP32, max(3, args.steps // 2))
    for _ in range(2):
        dev.compute_step(PREC_BF16L1)
    ms_bf, per_bf = timed_steps(dev, PREC_BF16L1, args.steps)
    kern_bf = kernel_times(dev, PREC_BF16L1, max(3, args.steps // 2))
    flops, counts = kernel_flops(model, strucs, dev)
    rl = roofline(flops, kern, "c5")
    E32, F32, _ = dev.compute(PREC_FP32)
    Ebf, Fbf, _ = dev.compute(PREC_BF16L1)
    dev.close()
    scaling = (np.concatenate([model.nets[I].inpw for I in range(model.ntypes)]),
               np.concatenate([model.nets[I].inpb for I in range(model.ntypes)]))
    mlp_ms = kern.get("mlp_energy_grad", 0.0)
    out = {
        "value": N / (ms * 1e-3), "ms_per_step": ms, "median_ms_per_step": float(np.median(per)),
        "min_ms_per_step": float(np.min(per)), "max_ms_per_step": float(np.max(per)),
        "config": {"workload": "256 structures, N log-uniform [256,8192], 4 elements 25/25/25/25, rc 6.5 A, 32 G2 + "
                               "96 G4 per element, 128-128-64-1 tanh; step = neighbour build + E + F + virial",
                   "global_batch": len(strucs), "atoms": int(N), **{k: int(v) for k, v in counts.items()}},
        "roofline": rl,
        "kernels_avg_ms": {k: round(v, 4) for k, v in kern.items()},
        "network_mfma": {"kernel": "mlp_energy_grad", "flops_per_launch": int(flops["mlp_energy_grad"]),
                         "avg_launch_ms": mlp_ms,
                         "achieved_tflops": flops["mlp_energy_grad"] / (mlp_ms * 1e-3) / 1e12 if mlp_ms else None,
                         "peak_tflops": PEAK_FP32_TFLOPS},
        "autotune": tune,
        "bf16_layer1": {"value": N / (ms_bf * 1e-3), "ms_per_step": ms_bf, "median_ms_per_step": float(np.median(per_bf)),
                        "kernels_avg_ms": {k: round(v, 4) for k, v in kern_bf.items()}},
    }
    return out, (model, strucs, scaling, E32, F32, Ebf, Fbf)


def c2_line(args):
    from ace_amd import PREC_FP32, PREC_FP64_BASIS, NNPDevice, configs
    model, [s] = configs.config2()
    dev = NNPDevice(model)
    dev.upload([s])
    dev.build_neighbors()
    dev.set_precision(PREC_FP64_BASIS)
    configs.set_input_scaling(model, [s], dev.compute_symm())
    dev.sync_model_scaling()
    dev.set_precision(PREC_FP32)
    tune = {"fp32": dev.autotune(PREC_FP32), "fp64_basis": dev.autotune(PREC_FP64_BASIS)}
    for prec in (PREC_FP32, PREC_FP64_BASIS):
        for _ in range(20):
            dev.compute_step(prec)
    dev.sync()
    t0 = time.perf_counter()
    for _ in range(args.c2_steps):
        dev.compute_step(PREC_FP32)
    dev.sync()
    ms = (time.perf_counter() - t0) / args.c2_steps * 1e3
    t0 = time.perf_counter()
    for _ in range(args.c2_steps):
        dev.compute_step(PREC_FP64_BASIS)
    dev.sync()
    ms64 = (time.perf_counter() - t0) / args.c2_steps * 1e3
    kern = kernel_times(dev, PREC_FP32, args.c2_steps)
    kern64 = kernel_times(dev, PREC_FP64_BASIS, args.c2_steps)
    flops, counts = kernel_flops(model, [s], dev)
    E, F, _ = dev.compute(PREC_FP32)
    E64, F64, _ = dev.compute(PREC_FP64_BASIS)
    # the plain unbatched FP64 GPU port on the same structure, same step shape
    for _ in range(3):
        dev.build_neighbors()
        dev.plain_compute()
    npl = 5
    t0 = time.perf_counter()
    for _ in range(npl):
        dev.build_neighbors()
        dev.plain_compute()
    plain_ms = (time.perf_counter() - t0) / npl * 1e3
    dev.close()
    scaling = (np.concatenate([model.nets[I].inpw for I in range(model.ntypes)]),
               np.concatenate([model.nets[I].inpb for I in range(model.ntypes)]))
    out = {"value": s.natoms / (ms * 1e-3), "unit": "atoms/s", "ms_per_step": ms,
           "config": {"workload": "bcc 1000 atoms, 2 elements 50/50, rc 6.0 A, 80 inputs (16 G2 + 16 G4 per "
                                  "species pair), 80-32-32-1 tanh; step = neighbour build + E + F + virial",
                      "atoms": int(s.natoms), **{k: int(v) for k, v in counts.items()}},
           "roofline": roofline(flops, kern, "c2"),
           "kernels_avg_ms": {k: round(v, 5) for k, v in kern.items()}, "autotune": tune,
           "fp64_basis": {"value": s.natoms / (ms64 * 1e-3), "ms_per_step": ms64,
                          "kernels_avg_ms": {k: round(v, 5) for k, v in kern64.items()}},
           "plain_fp64_port": {"value": s.natoms / (plain_ms * 1e-3), "ms_per_step": plain_ms}}
    return out, (model, s, scaling, E, F, E64, F64)


def plain_port_c5_sample(strucs_model, stride=16):
    """The plain unbatched FP64 GPU port (one structure per launch, one thread per atom) on every `stride`-th
    structure of the 568k-atom batch (a bounded sample: the port needs minutes for the whole batch)."""
    from ace_amd import NNPDevice
    model, strucs = strucs_model
    sub = strucs[::stride]
    n = sum(x.natoms for x in sub)
    dev = NNPDevice(model)
    dev.upload(sub)
    dev.build_neighbors()
    dev.plain_compute()
    dev.sync()
    t0 = time.perf_counter()
    dev.build_neighbors()
    dev.plain_compute()
    ms = (time.perf_counter() - t0) * 1e3
    dev.close()
    return {"value": n / (ms * 1e-3), "unit": "atoms/s", "ms": ms, "structures": len(sub), "atoms": int(n),
            "sample": f"every {stride}th structure of the batch ({len(sub)} structures, {n} atoms), one pass"}


def train_line(steps):
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
    configs.set_input_scaling(teacher, strucs, tdev.compute_symm())
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
    flops, counts = kernel_flops(model, strucs, dev)
    typ = np.concatenate([s.type for s in strucs])
    # SURVEY's training-step model: twice the basis forward + adjoint (the tangent recomputes dG/dr) plus
    # forward, tangent-forward, weight-gradient and backward network products
    flop_train = 2 * (flops["basis_symm"] + flops["basis_adjoint"]) + 4 * mlp_flops(model, typ)
    tr = Trainer(dev, model, strucs, gamma=1e-3, w_force=0.1, graph=True, device_opt=True)
    for _ in range(3):
        tr.step()
    dev.sync()
    per = []
    t0 = time.perf_counter()
    first = last = None
    for k in range(steps):
        a = time.perf_counter()
        out = tr.step()
        per.append((time.perf_counter() - a) * 1e3)
        first = out if k == 0 else first
        last = out
    dev.sync()
    dt = (time.perf_counter() - t0) / steps
    dev.close()
    achieved = flop_train / dt / 1e12
    return {"metric": "train structs/s", "value": len(strucs) / dt, "unit": "structures/s",
            "ms_per_step": dt * 1e3, "median_ms_per_step": float(np.median(per)), "structures": len(strucs),
            "atoms": int(off[-1]), "steps": steps, "graph": True, "optimizer": "ADAM on the GPU",
            "loss_first": float(first[0]), "loss_last": float(last[0]), "dtype": "fp32",
            "roofline": {"bound": "valu", "kernel": "whole training step", "achieved": achieved,
                         "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                         "flops_per_step": int(flop_train)},
            "workload": "64 structures, N log-uniform [32,1024], 3 elements, 56-64-64-1, energy + 0.1 force MSE"}


def md_line(steps):
    """Molecular dynamics throughput: 108,000-atom fcc cell (30^3 conventional cells, 0.085 atoms/A^3), 2
    elements 70/30 with masses 27/16 amu, Maxwell-Boltzmann velocities at 300 K, list radius 6 + 1 A rebuilt
    every 20 steps, 80-32-32-1 networks (output layers scaled to an rms force component of 0.1 eV/A),
    velocity Verlet with dt = 1 fs, steps between rebuilds replayed from HIP graphs.  The FP32 run's energy
    drift is reported next to the same trajectory with FP64 symmetry functions."""
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


def errors(ref, E, F, off, strucs):
    out = []
    for k, (Er, Fr) in sorted(ref.items()):
        n = strucs[k].natoms
        out.append({"structure": int(k), "atoms": int(n), "energy_per_atom": float(abs(E[k] - Er) / n),
                    "force_max_abs": float(np.max(np.abs(F[off[k]:off[k] + n] - Fr)))})
    return out


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
    ap.add_argument("--steps", type=int, default=10, help="timed steps of the headline batch")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--c2-steps", type=int, default=200)
    ap.add_argument("--train-steps", type=int, default=20)
    ap.add_argument("--md-steps", type=int, default=1000)
    ap.add_argument("--cpu-stride", type=int, default=2,
                    help="the CPU baseline's sample: every n-th structure of at most 2048 atoms")
    ap.add_argument("--cpu-samples", type=int, default=5, help="one-thread evaluations of the 1000-atom structure")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-headline", action="store_true", help="the headline batch only (profiling runs)")
    ap.add_argument("--only", choices=("headline", "c2", "train", "md"), default=None,
                    help="one leg only, its JSON object printed (profiling runs)")
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

    if args.only_headline:
        args.only = "headline"
    if args.only in ("c2", "train", "md"):
        leg = {"c2": lambda: c2_line(args)[0], "train": lambda: train_line(args.train_steps),
               "md": lambda: md_line(args.md_steps)}[args.only]()
        print(json.dumps(leg))
        return
    head, (m5, s5, sc5, E32, F32, Ebf, Fbf) = c5_headline(args)
    log(f"c5: {head['value'] / 1e6:.2f} M atoms/s ({head['ms_per_step']:.2f} ms per step)")
    res = {
        "metric": METRIC, "value": head["value"], "unit": "atoms/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded jittered fcc blocks, random-init networks with batch input statistics)",
        "config": head["config"], "roofline": head["roofline"],
        "median_ms_per_step": head["median_ms_per_step"], "min_ms_per_step": head["min_ms_per_step"],
        "max_ms_per_step": head["max_ms_per_step"], "kernels_avg_ms": head["kernels_avg_ms"],
        "network_mfma": head["network_mfma"], "autotune": head["autotune"], "bf16_layer1": head["bf16_layer1"],
    }
    if args.only == "headline":
        print(json.dumps(res))
        return
    res["plain_fp64_port"] = plain_port_c5_sample((m5, s5))
    res["plain_fp64_port"]["speedup_batched_vs_plain"] = head["value"] / res["plain_fp64_port"]["value"]
    c2, (m2, s2, sc2, E2, F2, E264, F264) = c2_line(args)
    res["c2"] = c2
    res["train"] = train_line(args.train_steps) if args.train_steps > 0 else None
    res["md"] = md_line(args.md_steps) if args.md_steps > 0 else None
    if not args.no_cpu:
        # cpu_baseline leg: the reference algorithm (FP64 oracle) on the host, timed on the headline batch;
        # its results double as the FP64 reference for the error report
        # sample: every 2nd of the structures with at most 2048 atoms (the oracle's time per atom grows with N
        # through its O(N^2) brute-force list, so this sample favours the CPU); ~10 s on 16 workers
        small = [k for k in range(len(s5)) if s5[k].natoms <= 2048]
        sample = small[::args.cpu_stride]
        cpu, ref = cpu_baseline_sharded(("c5", sc5), sample, f"every {args.cpu_stride}th of the batch's structures of at "
                                        f"most 2048 atoms ({len(sample)} structures, "
                                        f"{sum(s5[k].natoms for k in sample)} atoms)")
        res["cpu_baseline"] = cpu
        res["speedup_vs_cpu"] = head["value"] / cpu["value"]
        off = np.concatenate([[0], np.cumsum([x.natoms for x in s5])])
        res["error_vs_fp64_oracle"] = {"fp32": errors(ref, E32, F32, off, s5), "bf16_layer1": errors(ref, Ebf, Fbf, off, s5)}
        v1, (Er, Fr, _) = cpu_baseline_one(m2, s2, args.cpu_samples)
        res["c2"]["cpu_one_thread"] = {"value": v1, "unit": "atoms/s", "cores": 1, "kind": "port",
                                       "sample": f"{args.cpu_samples} evaluations of the 1000-atom structure, one thread "
                                                 "(the reference has no intra-structure parallelism)"}
        res["c2"]["error_vs_fp64_oracle"] = {
            "fp32": {"energy_per_atom": float(abs(E2[0] - Er) / s2.natoms), "force_max_abs": float(np.max(np.abs(F2 - Fr)))},
            "fp64_basis": {"energy_per_atom": float(abs(E264[0] - Er) / s2.natoms),
                           "force_max_abs": float(np.max(np.abs(F264 - Fr)))}}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
