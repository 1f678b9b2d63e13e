#!/usr/bin/env python3
"""Throughput of the GP-MOOD step on one MI355X.

Workload (the headline configuration): 2048 x 2048 cells on [0,1]^2, FP64,
GP-R3 reconstruction (Mord 7, 4 Gauss points per face), HLLC, SSP-RK3,
CFL 0.8, reflecting walls, MOOD with PAD + DMP + U2 and the order cascade
GP-R3 -> GP-R2 -> GP-R1 -> first order (the reference's own cascade, 7 -> 3 ->
1, is timed beside it);
seeded blast initial condition (rho = 1 +/- 1% white noise, u = v = 0,
p = 1e-5, total energy 1.0 added as a Gaussian of width 2 dx at a seeded centre
in [0.3, 0.7]^2).  The state stays resident in HBM; a step is dt reduction +
three RK stages, each a first-pass kernel plus the on-device MOOD passes.

Prints one JSON line: metric, value (cell-updates/s per RK stage), the
roofline of the dominant kernel (the first-pass stage kernel, timed with HIP
events on its own stream), a CPU baseline (the reference Fortran, serial,
timed on a one-step sample of the same workload on this host), the plain vs
NN-GP-MOOD comparison at 2048^2 GP-R2 with the committed predictor weights,
the host-array drop-in rate (PCIe round trip per step included), the
throughput of the reference's other options (fluxes, integrators, formula
boundaries, the RT source, FOG) and the cost of the 512^2 initial-state
sensitivity run (frozen-mask adjoint).
"""
import argparse
import importlib.util
import json
import os
import re
import resource
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "cell-updates/s per RK stage at 2048² FP64, plain vs NN-MOOD; % HBM roofline"
HBM_PEAK = 8.0e12        # B/s, MI355X HBM3E
FP64_PEAK = 78.6e12      # FLOP/s, MI355X FP64 (vector and matrix)
BYTES_PER_CELL_STAGE = 256.0 / 3  # 5 state reads + 3 writes of 32 B per step
# Kernel durations come from HIP events around the stage and MOOD launches of
# every TIMING_EVERY-th step of the timed region: each event pair costs the
# stream a few microseconds, which events on every launch would add to the
# measured step.
TIMING_EVERY = 5


def flops_per_cell_stage(mord, ngp):
    """Algorithmic FP64 operations per cell and RK stage (divide, sqrt, min,
    max count as one): reconstruction 32*ngp*sz, HLLC 2*ngp*100, quadrature
    8*(2*ngp-1), update 32, detection 80, RK combination 12, dt 7."""
    sz = {1: 1, 3: 5, 5: 13, 7: 25}[mord]
    return 32 * ngp * sz + 2 * ngp * 100 + 8 * (2 * ngp - 1) + 32 + 80 + 12 + 7


def load_package():
    pkg_dir = os.path.join(ROOT, "gp-mood_amd")
    spec = importlib.util.spec_from_file_location("gpmood_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gpmood_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _raise_stack():
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    try:
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))
    except (ValueError, OSError):
        pass


def cpu_baseline_reference(U0, n, steps=1, cascade="full"):
    """The reference Fortran (built from its sources by oracle/ref/build_ref.py;
    for the cascade 7 -> 5 -> 3 -> 1 the build with DETECTION's one-line cascade
    change) on the same initial state: serial, `steps` steps, its own cpu_time
    around the step loop."""
    exe = os.path.join(ROOT, "oracle", "_ref", f"r3_n{n}_ref" + ("_full" if cascade == "full" else ""), "gpm_ref.exe")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory(dir=os.path.dirname(exe)) as d:
        U0.tofile(os.path.join(d, "ic.bin"))
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([exe, os.path.join(d, "ic.bin"), str(steps), d, "0"], capture_output=True, text=True,
                           preexec_fn=_raise_stack, env=env, timeout=600)
        if r.returncode != 0:
            sys.stderr.write(f"reference CPU baseline failed (rc {r.returncode}): {r.stderr[-500:]}\n")
            return None
        m = re.search(r"cpu_s=\s*([0-9.Ee+-]+)\s+cell_stage_per_s=\s*([0-9.Ee+-]+)", r.stdout)
        if not m:
            return None
        return {"value": float(m.group(2)), "unit": "cell-updates/s per RK stage", "cores": 1,
                "kind": "reference", "seconds": float(m.group(1)),
                "sample": f"{steps} SSP-RK3 step(s) of the same {n}^2 R3 blast workload ({cascade} cascade) from "
                          f"the same initial state, reference Fortran (amdflang -O3, serial) on this host"}


def cpu_baseline_port(U0, n, steps=1, cascade="full"):
    """Fallback: the C restatement of the reference (oracle/gpm_oracle.c), one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle as O
        gpm = sys.modules["gpmood_amd"]
        t = gpm.gp_tables(7, 4, n)
        o = O.Oracle(n, 7, 4, "reflective", t, cascade=cascade)
        U = o.full(U0)
        t0 = time.perf_counter()
        o.run(U, steps)
        el = time.perf_counter() - t0
        return {"value": 3.0 * n * n * steps / el, "unit": "cell-updates/s per RK stage",
                "cores": int(os.environ.get("OMP_NUM_THREADS", "1")), "kind": "port", "seconds": el,
                "sample": f"{steps} step(s) of the same workload, C restatement of the reference"}
    except Exception as e:  # pragma: no cover - reported, not fatal
        sys.stderr.write(f"port CPU baseline failed: {e}\n")
        return None


def adjoint_measure(gpm, K, n=512, m=20, w=64):
    """Initial-state sensitivity of J = sum over a 64x64 subregion of
    rho (u^2 + v^2) after K steps of 512^2 GP-R2 (four-quadrant IC, outflow),
    masks and dt frozen, a checkpoint every m steps."""
    import torch
    s = gpm.Solver(gpm.Config(lf=n, Mord=5, ngp=3, BC_type="neumann"))
    U0 = torch.from_numpy(gpm.ic.make("four_quadrant", n, 3)).cuda().requires_grad_(True)
    # warm-up, forward and backward over two checkpoint segments: the first
    # backward pays one-time costs (the adjoint's work buffers, the kernels'
    # first loads), 0.07-0.3 s on top of the steady-state cost
    Uw = gpm.adjoint.steps(s, U0, min(K, 2 * m), checkpoint_every=m)
    gpm.adjoint.kinetic_energy(Uw, 224, 224, w).backward()
    U0.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    UK = gpm.adjoint.steps(s, U0, K, checkpoint_every=m)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    J = gpm.adjoint.kinetic_energy(UK, 224, 224, w)
    J.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    s.close()
    return {"workload": f"{n}x{n} GP-R2 four-quadrant seed 3, outflow, K={K} SSP-RK3 steps, checkpoint every {m}, "
                        f"J over a {w}x{w} subregion", "forward_s": t1 - t0, "backward_s": t2 - t1,
            "stage_adjoints_per_s": 3.0 * n * n * K / (t2 - t1),
            "note": "backward = frozen-mask recompute of each checkpoint segment + adjoint kernels; "
                    "unit: cell-stages per second"}


def large_run(gpm, steps, n=8192, seed=2000):
    """The large NN-GP-MOOD run: 8192^2 FP64, GP-R3 with the full cascade
    7 -> 5 -> 3 -> 1, HLLC, SSP-RK3, periodic, shear-layer IC (a seed no
    training run of the GP-R3 predictor generated); plain GP-MOOD and
    NN-GP-MOOD with the committed GP-R3 predictor from the same state.  Reports
    throughput, MOOD activity, the density L1 distance between the two and the
    drift of total mass and energy (periodic: conserved to round-off)."""
    import torch
    wfile = os.path.join(ROOT, "gp-mood_amd", "weights", "nn_r3.npz")
    U0 = torch.from_numpy(gpm.ic.make("shear_layer", n, seed)).cuda()
    tot0 = U0.sum(dim=(0, 1))
    out, final = {}, {}
    m = gpm.MLP.load(wfile) if os.path.exists(wfile) else None
    for label in ("plain", "nn"):
        if label == "nn" and m is None:
            continue
        cfg = gpm.Config(lf=n, Mord=7, ngp=4, BC_type="periodic", cascade="full",
                         method=gpm.capi.NN_GP_MOOD if label == "nn" else gpm.capi.GP_MOOD)
        s = gpm.Solver(cfg)
        if label == "nn":
            s.set_predictor(m, gate=True)
        torch.cuda.synchronize()
        s.set_state_device(U0.data_ptr())
        s.set_time(0.0)
        s.reset_stats()
        print(f"[bench] large run {label}: {steps} steps at {n}^2", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            k = min(100, steps - done)
            s.run(k)
            done += k
            torch.cuda.synchronize()
            print(f"[bench]   {label} {done}/{steps} steps, {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        el = time.perf_counter() - t0
        st = s.stats()
        U = torch.empty_like(U0)
        s.get_state_device(U.data_ptr())
        torch.cuda.synchronize()
        tot = U.sum(dim=(0, 1))
        final[label] = U[..., 0].clone()
        out[label] = {"cell_updates_per_s": 3.0 * n * n * steps / el, "ms_per_step": 1e3 * el / steps,
                      "second_pass_fraction": st["second_pass_cells"] / (3.0 * n * n * steps),
                      "max_passes": st["max_passes"], "a_priori_lowered_fraction":
                          st["a_priori_lowered"] / (3.0 * n * n * steps),
                      "mass_drift": float(abs(tot[0] - tot0[0]) / abs(tot0[0])),
                      "energy_drift": float(abs(tot[3] - tot0[3]) / abs(tot0[3]))}
        del U
        s.close()
    if "nn" in final:
        out["speedup"] = out["nn"]["cell_updates_per_s"] / out["plain"]["cell_updates_per_s"]
        out["density_l1_distance"] = float((final["nn"] - final["plain"]).abs().sum() / final["plain"].abs().sum())
    out["workload"] = (f"{n}x{n} FP64 GP-R3 (Mord 7, ngp 4), full cascade, HLLC, SSP-RK3, CFL 0.8, periodic, "
                       f"shear layer seed {seed}, {steps} steps; NN: gp-mood_amd/weights/nn_r3.npz")
    if m is not None:
        m.close()
    return out


# The reference's other options at GPU-scale grids: label, Config fields,
# initial condition (ic.problem kind) and RK stages per step.
OPTION_RUNS = [
    ("hll_r3_2048", dict(lf=2048, Mord=7, ngp=4, BC_type="reflective", numFlux="hll"), "blast", 3),
    ("unsplit_r3_2048", dict(lf=2048, Mord=7, ngp=4, BC_type="reflective", numFlux="unsplit"), "blast", 3),
    ("llf_r3_2048", dict(lf=2048, Mord=7, ngp=4, BC_type="reflective", numFlux="llf"), "blast", 3),
    ("ssp_rk2_r3_2048", dict(lf=2048, Mord=7, ngp=4, BC_type="reflective", integrator="ssp_rk2"), "blast", 2),
    ("ssp_rk4_r3_2048", dict(lf=2048, Mord=7, ngp=4, BC_type="reflective", integrator="ssp_rk4"), "blast", 5),
    ("dmr_r1_4096x1024", dict(lf=4096, nf=1024, Mord=3, ngp=2, BC_type="dmr", Lx=4.0, Ly=1.0), "dmr", 3),
    ("mach800_r1_2048", dict(lf=2048, Mord=3, ngp=2, BC_type="mach800", numFlux="hll", Lx=1.5, Ly=1.5), "mach800", 3),
    ("rt_r1_1024x3072", dict(lf=1024, nf=3072, Mord=3, ngp=2, BC_type="rt", numFlux="unsplit", Lx=0.5, Ly=1.5,
                             problem=51), "rayleigh_taylor", 3),
    ("fog_2048", dict(lf=2048, Mord=3, ngp=1, BC_type="kh", numFlux="unsplit", integrator="fe", method=1),
     "quadrants", 1),
    # the reference's own network (NN_GP_MOOD, a seeded network of the problem
    # files' L = 57, length = 60; niter past nsteps_with_no_NN) and plain
    # GP-MOOD on the same R1 workload
    ("gp_mood_r1_2048_quadrants", dict(lf=2048, Mord=3, ngp=2), "quadrants", 3),
    ("refnn_r1_2048_quadrants", dict(lf=2048, Mord=3, ngp=2, method=203), "quadrants", 3),
    # the other reconstructions: polynomial MOOD, unlimited GP and unlimited
    # polynomials (no limiter: a smooth periodic shear layer)
    ("pol_mood_r1_2048_quadrants", dict(lf=2048, Mord=3, ngp=2, method=200), "quadrants", 3),
    ("gp_se_r3_2048_shear", dict(lf=2048, Mord=7, ngp=4, BC_type="periodic", method=202), "shear", 3),
    ("unlim_pol_2048_shear", dict(lf=2048, Mord=3, ngp=2, BC_type="periodic", method=201), "shear", 3),
]


def options_measure(gpm, steps, warmup=3):
    """Throughput of the reference's other options (fluxes, integrators,
    formula boundaries, the RT source, FOG) at GPU-scale grids: cell-updates/s
    per RK stage over `steps` device-resident steps, and the first-pass
    kernel's average launch."""
    import torch
    out = {}
    for label, kw, kind, stages in OPTION_RUNS:
        cfg = gpm.Config(**kw)
        lf, nf = cfg.lf, cfg.nf or cfg.lf
        if kind == "blast":
            U0 = gpm.ic.blast(lf, 0)
        elif kind == "shear":
            U0 = gpm.ic.shear_layer(lf, 0)
        else:
            U0 = gpm.ic.problem(kind, lf, nf, cfg.Lx, cfg.Ly, seed=0, noise=0.0)
        s = gpm.Solver(cfg)
        if label.startswith("refnn"):
            s.set_reference_nn(**gpm.refnn.seeded_network(57, 60, seed=0, U=U0))
            s.set_niter(10)
        s.set_state(U0)
        s.set_time(0.0)
        s.run(warmup)
        torch.cuda.synchronize()
        s.reset_stats()
        s.set_timing(True, every=TIMING_EVERY)
        t0 = time.perf_counter()
        s.run(steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        kt = s.kernel_time()
        st = s.stats()
        out[label] = {"cell_updates_per_s": stages * lf * nf * steps / el, "ms_per_step": 1e3 * el / steps,
                      "stage_kernel_avg_ms": kt["stage_ms"], "mood_kernel_avg_ms": kt["mood_ms"],
                      "max_passes": st["max_passes"], "grid": [lf, nf], "stages_per_step": stages}
        if label.startswith("refnn"):
            out[label]["a_priori_fraction"] = st["count_detected_cell_a_priori"] / (stages * lf * nf)
        s.close()
        print(f"[bench]   {label}: {out[label]['cell_updates_per_s'] / 1e9:.3f} G cell-updates/s per stage",
              file=sys.stderr, flush=True)
    return out


def nn_comparison(gpm, n, steps, seeds=(2000, 2001, 2002, 2003)):
    """Plain GP-MOOD vs NN-GP-MOOD (committed predictor weights) on the
    2048^2 GP-R2 four-quadrant workload: throughput, second-pass cells and the
    density L1 distance after `steps` steps, on four-quadrant states whose
    seeds no training, held-out or threshold-choice run generated (the first
    seed in detail, every seed in `seeds`); held-out precision/recall and the
    threshold choice from the predictor's training report."""
    wdir = os.path.join(ROOT, "gp-mood_amd", "weights")
    wfile = os.path.join(wdir, "nn_r2.npz")
    if not os.path.exists(wfile):
        return None
    from gpmood_amd import nn_train
    m = gpm.MLP.load(wfile)
    out = nn_train.in_step(m, n=n, steps=steps, kind="four_quadrant", seed=seeds[0], bc="neumann",
                           log=lambda msg: print(msg, file=sys.stderr, flush=True))
    out.pop("_plain")
    # the spread over the other states (the gain is bounded by the share of the
    # plain step spent in MOOD passes, which varies by state)
    spread = {str(seeds[0]): {"speedup": out["speedup"], "density_l1_distance": out["density_l1_distance"],
                              "plain_ms_per_step": out["plain"]["ms_per_step"]}}
    for sd in seeds[1:]:
        o = nn_train.in_step(m, n=n, steps=steps, kind="four_quadrant", seed=sd, bc="neumann", log=lambda msg: None)
        o.pop("_plain")
        spread[str(sd)] = {"speedup": o["speedup"], "density_l1_distance": o["density_l1_distance"],
                           "plain_ms_per_step": o["plain"]["ms_per_step"]}
        print(f"[bench]   NN seed {sd}: {o['speedup']:.3f}x, L1 {o['density_l1_distance']:.2e}", file=sys.stderr,
              flush=True)
    out["seeds"] = spread
    sp = [v["speedup"] for v in spread.values()]
    out["speedup_geomean_over_seeds"] = float(np.exp(np.mean(np.log(sp))))
    out["seeds_meeting_target"] = sum(1 for v in spread.values()
                                      if v["speedup"] >= 1.5 and v["density_l1_distance"] <= 1e-3)
    rep = os.path.join(wdir, "nn_r2_report.json")
    if os.path.exists(rep):
        r = json.load(open(rep))
        ev = r.get("eval_test", {})
        out["held_out"] = {"troubled_precision": ev.get("troubled", {}).get("precision"),
                           "troubled_recall": ev.get("troubled", {}).get("recall"),
                           "accuracy": ev.get("accuracy"), "samples": ev.get("samples")}
        th = r.get("threshold", {})
        out["threshold"] = {"mord_logit_shift": th.get("mord_logit_shift"), "chosen_on": th.get("chosen_on"),
                            "curve": [{k: c[k] for k in ("shift", "speedup_geomean", "max_density_l1")}
                                      for c in th.get("curve", [])]}
        out["training_seeds"] = r.get("seeds")
    out["weights"] = "gp-mood_amd/weights/nn_r2.npz"
    m.close()
    return out


BF16_PEAK = 2.5e15        # FLOP/s, MI355X dense BF16 matrix


def mlp_train_measure(gpm, steps, batch=8192, warmup=5):
    """AdamW training steps of the order predictor (forward, cross-entropy,
    backward on the matrix cores, AdamW) at batch 8192 on resident random
    features, for the GP-R2 (100 -> 3 classes) and GP-R3 (196 -> 4) networks.
    Loss is not read back inside the timed loop.  FLOPs are the algorithmic
    ones of the unpadded layers: forward 2*MACs, backward weights 2*MACs,
    backward data 2*MACs of all layers but the first."""
    import torch
    out = {}
    for tag, in_f, C in (("r2", 100, 3), ("r3", 196, 4)):
        dims = (in_f, 128, 128, 64, C)
        macs = [a * b for a, b in zip(dims[:-1], dims[1:])]
        flop_sample = 2 * (2 * sum(macs) + sum(macs[1:]))
        m = gpm.MLP(in_f, C, max_batch=batch, seed=1)
        g = torch.Generator(device="cuda").manual_seed(7)
        x = torch.rand((batch, in_f), generator=g, device="cuda", dtype=torch.float32) * 2 - 1
        y = torch.randint(0, C, (batch,), generator=g, device="cuda", dtype=torch.int32)
        for _ in range(warmup):
            m.train_step(x, y, report=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            m.train_step(x, y, report=False)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        loss, _ = m.train_step(x, y)
        m.close()
        fl = flop_sample * batch * steps / el
        out[tag] = {"value": batch * steps / el, "unit": "samples/s", "ms_per_step": 1e3 * el / steps,
                    "batch": batch, "layers": list(dims), "flop_per_sample": flop_sample,
                    "achieved_tflops": fl / 1e12, "bf16_peak_tflops": BF16_PEAK / 1e12, "frac": fl / BF16_PEAK,
                    "final_loss": loss}
    out["note"] = ("each step is 32 dependent launches (forward 4, loss 1, backward 15, AdamW 8, BF16 copies 4) "
                   "of a few microseconds over a few MB: launch- and latency-bound, not matrix-core-bound")
    return out


def plain_run(gpm, n, steps, warmup, seed, cascade):
    """The headline workload (2048^2 GP-R3 reflective blast) with the given
    order cascade: `warmup` untimed steps, then `steps` timed device-resident
    steps with HIP-event timing of the kernels."""
    import torch
    U0 = gpm.ic.blast(n, seed)
    cfg = gpm.Config(lf=n, Mord=7, ngp=4, BC_type="reflective", CFL=0.8, cascade=cascade)
    s = gpm.Solver(cfg)
    s.set_state(U0)
    s.set_time(0.0)
    print(f"[bench] {cascade} cascade: warmup {warmup} steps at {n}^2", file=sys.stderr, flush=True)
    s.run(warmup)
    torch.cuda.synchronize()
    s.reset_stats()
    s.set_timing(True, every=TIMING_EVERY)
    print(f"[bench] {cascade} cascade: timing {steps} steps", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return s, U0, el, s.kernel_time(), s.stats()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--nn-steps", type=int, default=100, help="steps of the plain vs NN-MOOD comparison (0: skip)")
    ap.add_argument("--nn-seeds", default="2000,2001,2002,2003",
                    help="four-quadrant seeds of the NN comparison (comma list; none of them is a training, "
                         "held-out or threshold-choice seed)")
    ap.add_argument("--dropin-steps", type=int, default=3, help="host-array drop-in steps timed (0: skip)")
    ap.add_argument("--large-steps", type=int, default=20,
                    help="steps of the 8192^2 GP-R3 periodic shear-layer run, plain and NN-GP-MOOD (0: skip)")
    ap.add_argument("--options-steps", type=int, default=20,
                    help="steps of each run of the other flux, integrator, boundary and method options (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="K of the 512^2 GP-R2 initial-state sensitivity run (0: skip)")
    ap.add_argument("--mlp-train-steps", type=int, default=200,
                    help="AdamW steps of the order predictor at batch 8192, GP-R2 and GP-R3 networks (0: skip)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this size with a plain device copy kernel "
                         "(reference point for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); the multi-GPU "
              "decomposition is a design note, not implemented", file=sys.stderr)
        sys.exit(2)

    import torch
    gpm = load_package()
    n, mord, ngp = args.n, 7, 4
    s, U0, el, kt, st = plain_run(gpm, n, args.steps, args.warmup, args.seed, "full")
    if args.calib_copy_mib > 0:
        nbytes = args.calib_copy_mib << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    # the reference's own cascade (7 -> 3 -> 1) on the same workload
    s_ref, _, el_ref, kt_ref, st_ref = plain_run(gpm, n, args.steps, args.warmup, args.seed, "reference")
    s_ref.close()
    ref_cascade = {"value": 3.0 * n * n * args.steps / el_ref, "ms_per_step": 1e3 * el_ref / args.steps,
                   "stage_kernel_avg_ms": kt_ref["stage_ms"], "mood_kernel_avg_ms": kt_ref["mood_ms"],
                   "max_passes": st_ref["max_passes"],
                   "detected_per_step": st_ref["detected_total"] / max(1, args.steps),
                   "workload": f"the same {n}x{n} GP-R3 blast run with the reference's cascade 7 -> 3 -> 1 "
                               "(mod_subroutine_MOOD.f90:159-163)"}

    value = 3.0 * n * n * args.steps / el
    fl = flops_per_cell_stage(mord, ngp)
    avg_s = kt["stage_ms"] * 1e-3
    bytes_launch = BYTES_PER_CELL_STAGE * n * n
    flops_launch = fl * n * n
    fp64_ach = flops_launch / avg_s if avg_s > 0 else 0.0
    hbm_ach = bytes_launch / avg_s if avg_s > 0 else 0.0
    attainable = min(HBM_PEAK * fl / BYTES_PER_CELL_STAGE, FP64_PEAK)
    # HBM bytes per launch need hardware counters (rocprofv3 --pmc), which a
    # program cannot read about itself: the value is the one stored from the
    # profiling run of this kernel on this workload
    traffic, traffic_source = None, None
    tf = os.path.join(ROOT, "profiles", "traffic_k_stage_r3_2048.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
            traffic_source = ("stored: profiles/traffic_k_stage_r3_2048.json (rocprofv3 --pmc FETCH_SIZE and "
                              "WRITE_SIZE of this kernel on this workload, scaled by a copy-kernel calibration)")
        except Exception:
            traffic = None
    roofline = {
        "bound": "fp64-valu", "achieved": fp64_ach / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
        "frac": fp64_ach / FP64_PEAK, "traffic": traffic, "traffic_source": traffic_source,
        "kernel": "k_stage<7,4,*> (first MOOD pass: GP-R3 reconstruction + HLLC + update + detection)",
        "avg_launch_ms": kt["stage_ms"], "launches": kt["stage_launches"],
        "timing": f"HIP events on the solver's stream around the launches of every {TIMING_EVERY}th timed step",
        "flop_per_cell_stage": fl, "bytes_per_cell_stage": BYTES_PER_CELL_STAGE,
        "hbm": {"achieved": hbm_ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": hbm_ach / HBM_PEAK},
        "attainable_frac": fp64_ach / attainable,
        "step_hbm_frac": value * BYTES_PER_CELL_STAGE / HBM_PEAK,
        # bit-identical results forbid fused multiply-adds, so each multiply and
        # each add is its own instruction: the unfused FP64 rate is half the peak
        "unfused_fp64_peak": FP64_PEAK / 2 / 1e12,
        "frac_of_unfused_peak": fp64_ach / (FP64_PEAK / 2),
        "note": "FP64-bound (AI %.1f FLOP/B > ridge 9.8); runs on the FP64 VALU, whose peak equals the "
                "FP64 matrix peak; the reference's rounding needs unfused multiplies and adds, whose "
                "rate is half that peak" % (fl / BYTES_PER_CELL_STAGE),
    }
    dropin = None
    if args.dropin_steps > 0:
        # gpm_time_stepping with host arrays in the reference layout: the
        # per-step drop-in of time_stepping(Uin, Uout), PCIe copies included
        U = s.get_state()
        t_host = 0.0
        el_d = 0.0
        for _ in range(args.dropin_steps):
            dt = s.setdt()
            t0 = time.perf_counter()
            U, _, _ = s.time_stepping(U, t_host, dt)
            el_d += time.perf_counter() - t0
            t_host += dt
        dropin = {"value": 3.0 * n * n * args.dropin_steps / el_d, "unit": "cell-updates/s per RK stage",
                  "ms_per_step": 1e3 * el_d / args.dropin_steps,
                  "note": "gpm_time_stepping on host arrays (nf+8)x(lf+8)x4 FP64: one host-to-device and one "
                          "device-to-host copy of the state per step, plus ghost filling on the host"}
    s.close()
    nn = None
    if args.nn_steps > 0:
        print(f"[bench] plain vs NN-GP-MOOD, {args.nn_steps} steps at {n}^2 GP-R2", file=sys.stderr, flush=True)
        nn = nn_comparison(gpm, n, args.nn_steps, [int(v) for v in args.nn_seeds.split(",") if v.strip()])
    large = None
    if args.large_steps > 0:
        large = large_run(gpm, args.large_steps)
    opts = None
    if args.options_steps > 0:
        print(f"[bench] other options, {args.options_steps} steps each", file=sys.stderr, flush=True)
        opts = options_measure(gpm, args.options_steps)
    adj = None
    if args.adjoint_steps > 0:
        print(f"[bench] adjoint, K={args.adjoint_steps} at 512^2 GP-R2", file=sys.stderr, flush=True)
        adj = adjoint_measure(gpm, args.adjoint_steps)
    mlp = None
    if args.mlp_train_steps > 0:
        print(f"[bench] predictor training, {args.mlp_train_steps} steps", file=sys.stderr, flush=True)
        mlp = mlp_train_measure(gpm, args.mlp_train_steps)
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        cpu = (cpu_baseline_reference(U0, n, args.cpu_steps, "full") or
               cpu_baseline_reference(U0, n, args.cpu_steps, "reference") or cpu_baseline_port(U0, n, args.cpu_steps))
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s per RK stage", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": f"synthetic: seeded blast initial condition (seed {args.seed})",
        "config": {"workload": f"plain GP-MOOD step {n}x{n}, GP-R3 (Mord 7, ngp 4), cascade R3->R2->R1->first "
                               "order, HLLC, SSP-RK3, CFL 0.8, reflective walls, blast IC", "n": n, "mord": mord,
                   "ngp": ngp, "cascade": "full", "parallelism": "single GPU"},
        "reference_cascade": ref_cascade,
        "mood": {"max_passes": st["max_passes"], "detected_per_step": st["detected_total"] / max(1, args.steps),
                 "second_pass_cells_per_step": st["second_pass_cells"] / max(1, args.steps),
                 "mood_kernel_avg_ms": kt["mood_ms"]},
        "roofline": roofline, "cpu_baseline": cpu, "nn_mood": nn, "host_array_drop_in": dropin, "adjoint": adj,
        "large_run": large, "options": opts, "predictor_training": mlp,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
