#!/usr/bin/env python3
"""NN-GP-MOOD pipeline: dataset from the plain GPU solver, training on the
GPU, held-out evaluation and the in-step comparison with plain GP-MOOD.

Dataset: seeded initial conditions at 512^2 (four-quadrant problems with
outflow boundaries and periodic shear layers), run by the plain GP-R2 solver
(Mord 5, 3 Gauss points, HLLC, SSP-RK3, CFL 0.8); every `dump_every` steps the
first RK stage's (normalised 5x5x4 stencil, final MOOD order) pairs of the
cells the gate would send to the network are dumped on the device (cells
below Mord always, cells at Mord with probability keep_mord), as the
reference does at every first RK stage (mod_FE.f90:176-181).

Training: MLP 100 -> 128 -> 128 -> 64 -> 3 (ReLU), BF16 weights with FP32
accumulation, AdamW lr 1e-3, batch 8192, cross-entropy.

Seeds: the training ICs are seeds 0..n_ics/2-1 of each family.  The held-out
samples (precision / recall) come from seeds HOLDOUT_SEEDS, the decision
threshold is chosen on the in-step runs of VALIDATION_SEEDS, and the in-step
evaluation (the report, the bench and the tests) uses EVAL_SEEDS: no IC of the
last three sets is ever generated for training (an IC depends on its seed
only, not on the grid).

Decision threshold: the order is the argmax of the logits after subtracting
`mord_logit_shift` from the Mord class's logit (folded into that class's
bias), so a cell keeps Mord a priori only when the network prefers it by that
margin.  Missing a troubled cell costs MOOD passes, while predicting a lower
order than needed adds diffusion.  The shift is the one of SHIFTS with the
largest geometric-mean speed-up over the validation ICs among those whose
density L1 distance from plain GP-MOOD stays within 1e-3 on every one of them
(the whole sweep is kept in the report as a speed-up / distance curve).  The
held-out evaluation uses the shifted rule.

GP-R3 predictor (--mord 7): the same pipeline with Mord 7, 4 Gauss points,
the 7x7x4 stencil (196 inputs) and 4 classes {1, 3, 5, 7}, run with the full
cascade 7 -> 5 -> 3 -> 1 so that every class occurs.

Usage: python gp-mood_amd/nn_train.py --out DIR [--mord 5|7 --n-ics 128 --n 512 --steps 400]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _package():
    if "gpmood_amd" in sys.modules:
        return sys.modules["gpmood_amd"]
    spec = importlib.util.spec_from_file_location("gpmood_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gpmood_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class Family:
    """Order family of a predictor: Mord, Gauss points, inputs, classes, cascade."""

    def __init__(self, mord=5):
        if mord not in (5, 7):
            raise ValueError("predictors exist for Mord 5 and 7")
        self.mord = mord
        self.ngp = {5: 3, 7: 4}[mord]
        self.stencil = mord  # box stencil of side 2R+1 = Mord
        self.inputs = 4 * self.stencil * self.stencil
        self.classes = (mord + 1) // 2
        self.cascade = "reference" if mord == 5 else "full"
        self.tag = {5: "r2", 7: "r3"}[mord]

    def config(self, gpm, n, bc, **kw):
        return gpm.Config(lf=n, Mord=self.mord, ngp=self.ngp, BC_type=bc, cascade=self.cascade, **kw)


R2 = Family(5)

# IC seeds never generated for training (training uses 0 .. n_ics/2 - 1)
HOLDOUT_SEEDS = tuple(range(1000, 1008))   # held-out samples
VALIDATION_SEEDS = (1000, 1001, 1002, 1003)  # choice of the decision threshold (in-step runs)
EVAL_SEEDS = (2000, 2001, 2002, 2003)      # in-step evaluation: report, bench, tests
SHIFTS = (0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0)


def ic_list(n_ics):
    """Half four-quadrant (outflow), half shear layers (periodic), seeds 0.."""
    half = n_ics // 2
    return [("four_quadrant", s, "neumann") for s in range(half)] + \
           [("shear_layer", s, "periodic") for s in range(n_ics - half)]


def holdout_list():
    """The held-out sample ICs: both families at HOLDOUT_SEEDS."""
    return [("four_quadrant", s, "neumann") for s in HOLDOUT_SEEDS] + \
           [("shear_layer", s, "periodic") for s in HOLDOUT_SEEDS]


def generate(ics, n=512, steps=400, dump_every=8, keep_mord=0.3, cap=20_000_000, log=print, fam=R2):
    """(X, Y, owner): samples of the plain solver's runs of `ics` ((kind,
    seed, bc) triples) in device memory, and the sample range of each IC."""
    import torch
    gpm = _package()
    X = torch.empty((cap, fam.inputs), dtype=torch.float32, device="cuda")
    Y = torch.empty(cap, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    owner = []  # (ic index, first sample, end)
    t0 = time.time()
    n_ics = len(ics)
    for k, (kind, seed, bc) in enumerate(ics):
        s = gpm.Solver(fam.config(gpm, n, bc))
        s.set_state(gpm.ic.make(kind, n, seed))
        first = int(cnt.item())
        done = 0
        while done < steps:
            quiet = min(dump_every - 1, steps - done)
            if quiet:
                s.run(quiet)
                done += quiet
            if done < steps:
                s.set_dump(X.data_ptr(), Y.data_ptr(), cap, cnt.data_ptr(), keep_mord=keep_mord,
                           seed=(0x9E37 * (seed + 1) + (kind == "shear_layer")) * 100003 + done, gated_only=True,
                           stage_mask=1)
                s.run(1)
                s.set_dump(0, 0, 0, 0)
                done += 1
        end = min(int(cnt.item()), cap)
        owner.append((k, first, end))
        if k % 16 == 0 or k == n_ics - 1:
            log(f"[data] ic {k + 1}/{n_ics} ({kind} seed {seed}): {end} samples, {time.time() - t0:.1f}s")
        s.close()
        if end >= cap:
            log("[data] capacity reached")
            break
    total = min(int(cnt.item()), cap)
    return X[:total], Y[:total], owner


def gather(X, Y, ranges):
    import torch
    idx = torch.cat([torch.arange(a, b, device="cuda") for a, b in ranges if b > a])
    return X[idx].contiguous(), Y[idx].contiguous()


def train(Xtr, Ytr, epochs=4, batch=8192, lr=1e-3, weight_decay=1e-2, seed=0, log=print, fam=R2):
    import torch
    gpm = _package()
    m = gpm.MLP(fam.inputs, fam.classes, max_batch=batch, seed=seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n = Xtr.shape[0]
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=g)
        ls, ac, nb = 0.0, 0.0, 0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            loss, acc = m.train_step(Xtr[idx].contiguous(), Ytr[idx].contiguous(), lr=lr, weight_decay=weight_decay)
            ls += loss
            ac += acc
            nb += 1
        hist.append({"epoch": ep + 1, "loss": ls / max(nb, 1), "accuracy": ac / max(nb, 1), "batches": nb})
        log(f"[train] epoch {ep + 1}: loss {ls / max(nb, 1):.4f} acc {ac / max(nb, 1):.4f}")
    return m, hist


def shift_mord_logit(m, shift):
    """Fold the decision threshold into the Mord class's output bias (relative
    to the current bias)."""
    if shift:
        p = m.get_params()
        p[-1] -= shift  # last layer bias, class C-1 = order Mord
        m.set_params(p)


def select_shift(m, seeds=VALIDATION_SEEDS, shifts=SHIFTS, n=2048, steps=100, kind="four_quadrant", bc="neumann",
                 l1_bar=1e-3, log=print, fam=R2):
    """The decision threshold from in-step runs on validation ICs: for each
    Mord-logit shift, speed-up over plain GP-MOOD and density L1 distance per
    seed.  Returns (chosen shift, curve): the shift with the largest
    geometric-mean speed-up whose distance stays within l1_bar on every seed
    (the smallest worst-case distance if none does).  m is left unshifted."""
    curve, plain = [], {}
    for sh in shifts:
        shift_mord_logit(m, sh)
        per = {}
        for sd in seeds:
            o = in_step(m, n=n, steps=steps, kind=kind, seed=sd, bc=bc, log=lambda msg: None, fam=fam,
                        plain=plain.get(sd))
            plain[sd] = o["_plain"]
            per[str(sd)] = {"speedup": o["speedup"], "density_l1_distance": o["density_l1_distance"]}
        shift_mord_logit(m, -sh)
        sp = [v["speedup"] for v in per.values()]
        d = [v["density_l1_distance"] for v in per.values()]
        row = {"shift": sh, "speedup_geomean": float(np.exp(np.mean(np.log(sp)))), "max_density_l1": float(max(d)),
               "seeds": per}
        curve.append(row)
        log(f"[threshold] shift {sh}: speed-up {row['speedup_geomean']:.3f}, max L1 {row['max_density_l1']:.2e}")
    ok = [r for r in curve if r["max_density_l1"] <= l1_bar]
    best = max(ok, key=lambda r: r["speedup_geomean"]) if ok else min(curve, key=lambda r: r["max_density_l1"])
    return best["shift"], curve


def evaluate(m, X, Y, fam=R2):
    """Confusion matrix over the order classes and the troubled-cell view
    (predicted below Mord vs. final order below Mord)."""
    C = fam.classes
    pred = np.argmax(m.logits(X).cpu().numpy()[:, :C], axis=1)
    y = Y.cpu().numpy()
    cm = np.zeros((C, C), dtype=np.int64)
    np.add.at(cm, (y, pred), 1)
    per_class = {}
    for c in range(C):
        tp = cm[c, c]
        per_class[f"order_{2 * c + 1}"] = {
            "precision": float(tp / max(cm[:, c].sum(), 1)), "recall": float(tp / max(cm[c, :].sum(), 1)),
            "support": int(cm[c, :].sum())}
    top = C - 1
    t_true, t_pred = y < top, pred < top
    tp = int(np.sum(t_true & t_pred))
    return {"accuracy": float(np.mean(pred == y)), "confusion": cm.tolist(), "per_class": per_class,
            "troubled": {"precision": tp / max(int(t_pred.sum()), 1), "recall": tp / max(int(t_true.sum()), 1),
                         "support": int(t_true.sum())}, "samples": int(y.size)}


def in_step(m, n=2048, steps=100, kind="four_quadrant", seed=0, bc="neumann", gate=True, log=print, fam=R2,
            plain=None):
    """Plain vs NN-GP-MOOD on the same initial state: throughput, second-pass
    cells, a posteriori detections, and the density L1 distance after `steps`.
    `plain`: the "_plain" entry of an earlier call on the same state (its
    plain run is then not repeated)."""
    import torch
    gpm = _package()
    U0 = gpm.ic.make(kind, n, seed)
    out = {}
    states = {}
    if plain is not None:
        out["plain"], states["plain"] = plain
    for label, use_nn in (("plain", False), ("nn", True)):
        if label in out:
            continue
        s = gpm.Solver(fam.config(gpm, n, bc, method=gpm.capi.NN_GP_MOOD if use_nn else gpm.capi.GP_MOOD))
        if use_nn:
            s.set_predictor(m, gate=gate)
        s.set_state(U0)
        s.run(3)  # warm-up
        s.set_state(U0)
        s.reset_stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run(steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        st = s.stats()
        states[label] = s.get_state()[4:-4, 4:-4]
        out[label] = {"cell_updates_per_s": 3.0 * n * n * steps / el, "ms_per_step": 1e3 * el / steps,
                      "second_pass_fraction": st["second_pass_cells"] / (3.0 * n * n * steps),
                      "a_posteriori_detected_fraction": st["detected_total"] / (3.0 * n * n * steps),
                      "a_priori_lowered_fraction": st["a_priori_lowered"] / (3.0 * n * n * steps),
                      "max_passes": st["max_passes"]}
        log(f"[in-step] {label}: {out[label]}")
        s.close()
    rp, rn = states["plain"][..., 0], states["nn"][..., 0]
    out["_plain"] = (out["plain"], states["plain"])
    out["density_l1_distance"] = float(np.abs(rn - rp).sum() / np.abs(rp).sum())
    out["speedup"] = out["nn"]["cell_updates_per_s"] / out["plain"]["cell_updates_per_s"]
    out["workload"] = (f"{n}x{n} GP-R{(fam.mord - 1) // 2} (Mord {fam.mord}, ngp {fam.ngp}, cascade {fam.cascade}) "
                       f"{kind} seed {seed} {bc}, {steps} steps, gate {int(gate)}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--mord", type=int, default=5, choices=(5, 7))
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--n-ics", type=int, default=128)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--dump-every", type=int, default=4)
    ap.add_argument("--keep-mord", type=float, default=0.3)
    ap.add_argument("--cap", type=int, default=24_000_000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-n", type=int, default=2048)
    ap.add_argument("--eval-steps", type=int, default=100)
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    fam = Family(args.mord)

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    Xtr, Ytr, _ = generate(ic_list(args.n_ics), args.n, args.steps, args.dump_every, args.keep_mord, args.cap, log,
                           fam=fam)
    Xte, Yte, _ = generate(holdout_list(), args.n, args.steps, args.dump_every, args.keep_mord, 4_000_000, log,
                           fam=fam)
    counts = np.bincount(Ytr.cpu().numpy(), minlength=fam.classes).tolist()
    log(f"[data] train {Xtr.shape[0]} samples (class counts {counts}), held-out {Xte.shape[0]}, "
        f"{time.time() - t0:.1f}s")
    m, hist = train(Xtr, Ytr, args.epochs, args.batch, args.lr, log=log, fam=fam)
    ev_train = evaluate(m, Xtr[:2_000_000], Ytr[:2_000_000], fam=fam)
    del Xtr, Ytr
    if fam.mord == 5:
        sel_kind, sel_bc = "four_quadrant", "neumann"
        evals = (("in_step", "four_quadrant", "neumann"), ("in_step_blast", "blast", "reflective"))
    else:
        sel_kind, sel_bc = "shear_layer", "periodic"
        evals = (("in_step", "shear_layer", "periodic"), ("in_step_four_quadrant", "four_quadrant", "neumann"))
    shift, curve = select_shift(m, n=args.eval_n, steps=args.eval_steps, kind=sel_kind, bc=sel_bc, log=log, fam=fam)
    log(f"[threshold] chosen Mord-logit shift {shift}")
    shift_mord_logit(m, shift)
    m.save(os.path.join(args.out, f"nn_{fam.tag}.npz"))
    ev_test = evaluate(m, Xte, Yte, fam=fam)
    log(f"[eval] held-out: {ev_test}")
    report = {"family": {"mord": fam.mord, "ngp": fam.ngp, "inputs": fam.inputs, "classes": fam.classes,
                         "cascade": fam.cascade},
              "seeds": {"train": f"0..{args.n_ics // 2 - 1} of each family", "holdout": list(HOLDOUT_SEEDS),
                        "validation": list(VALIDATION_SEEDS), "eval": list(EVAL_SEEDS)},
              "dataset": {"n": args.n, "n_ics": args.n_ics, "steps": args.steps, "dump_every": args.dump_every,
                          "keep_mord": args.keep_mord, "train": int(sum(counts)), "class_counts": counts,
                          "holdout_samples": int(Xte.shape[0])},
              "training": {"epochs": args.epochs, "batch": args.batch, "lr": args.lr, "history": hist},
              "threshold": {"mord_logit_shift": shift, "chosen_on": f"{sel_kind} seeds {list(VALIDATION_SEEDS)} at "
                            f"{args.eval_n}^2, {args.eval_steps} steps", "curve": curve},
              "eval_train": ev_train, "eval_test": ev_test}
    del Xte, Yte
    for key, kind, bc in evals:
        per = {}
        for sd in EVAL_SEEDS:
            o = in_step(m, n=args.eval_n, steps=args.eval_steps, kind=kind, seed=sd, bc=bc, log=log, fam=fam)
            o.pop("_plain")
            per[str(sd)] = o
        sp = [v["speedup"] for v in per.values()]
        report[key] = {"seeds": per, "speedup_geomean": float(np.exp(np.mean(np.log(sp)))),
                       "max_density_l1": float(max(v["density_l1_distance"] for v in per.values()))}
    report["seconds"] = time.time() - t0
    json.dump(report, open(os.path.join(args.out, f"nn_{fam.tag}_report.json"), "w"), indent=1)
    log(f"[done] {report['seconds']:.1f}s")


if __name__ == "__main__":
    main()
