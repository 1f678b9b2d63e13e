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
accumulation, AdamW lr 1e-3, batch 8192, cross-entropy; the last ICs of each
family are held out.

Decision threshold: the order is the argmax of the logits after subtracting
`mord_logit_shift` from the Mord class's logit (folded into that class's
bias), so a cell keeps Mord a priori only when the network prefers it by that
margin.  Missing a troubled cell costs MOOD passes, while predicting a lower
order than needed only adds a little diffusion.  The held-out evaluation uses
the shifted rule.

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


def ic_list(n_ics):
    """Half four-quadrant (outflow), half shear layers (periodic), seeds 0.."""
    half = n_ics // 2
    return [("four_quadrant", s, "neumann") for s in range(half)] + \
           [("shear_layer", s, "periodic") for s in range(n_ics - half)]


def generate(n=512, n_ics=128, steps=400, dump_every=8, keep_mord=0.3, cap=20_000_000, log=print, fam=R2):
    import torch
    gpm = _package()
    X = torch.empty((cap, fam.inputs), dtype=torch.float32, device="cuda")
    Y = torch.empty(cap, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    owner = []  # (ic index, first sample, end)
    t0 = time.time()
    for k, (kind, seed, bc) in enumerate(ic_list(n_ics)):
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
                           seed=1000 * k + done, gated_only=True, stage_mask=1)
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


def split(owner, n_ics, holdout_per_family=8):
    half = n_ics // 2
    test_ics = set(range(half - holdout_per_family, half)) | set(range(n_ics - holdout_per_family, n_ics))
    tr, te = [], []
    for k, a, b in owner:
        (te if k in test_ics else tr).append((a, b))
    return tr, te


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
    """Fold the decision threshold into the Mord class's output bias."""
    if shift:
        p = m.get_params()
        p[-1] -= shift  # last layer bias, class C-1 = order Mord
        m.set_params(p)


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


def in_step(m, n=2048, steps=100, kind="four_quadrant", seed=0, bc="neumann", gate=True, log=print, fam=R2):
    """Plain vs NN-GP-MOOD on the same initial state: throughput, second-pass
    cells, a posteriori detections, and the density L1 distance after `steps`."""
    import torch
    gpm = _package()
    U0 = gpm.ic.make(kind, n, seed)
    out = {}
    states = {}
    for label, use_nn in (("plain", False), ("nn", True)):
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
    ap.add_argument("--dump-every", type=int, default=8)
    ap.add_argument("--keep-mord", type=float, default=0.3)
    ap.add_argument("--cap", type=int, default=20_000_000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--mord-logit-shift", type=float, default=1.0)
    ap.add_argument("--eval-n", type=int, default=2048)
    ap.add_argument("--eval-steps", type=int, default=100)
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    fam = Family(args.mord)

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    X, Y, owner = generate(args.n, args.n_ics, args.steps, args.dump_every, args.keep_mord, args.cap, log, fam=fam)
    tr, te = split(owner, args.n_ics, holdout_per_family=max(1, args.n_ics // 16))
    Xtr, Ytr = gather(X, Y, tr)
    Xte, Yte = gather(X, Y, te)
    counts = np.bincount(Y.cpu().numpy(), minlength=fam.classes).tolist()
    log(f"[data] {X.shape[0]} samples (train {Xtr.shape[0]}, test {Xte.shape[0]}), class counts {counts}, "
        f"{time.time() - t0:.1f}s")
    del X, Y
    m, hist = train(Xtr, Ytr, args.epochs, args.batch, args.lr, log=log, fam=fam)
    shift_mord_logit(m, args.mord_logit_shift)
    m.save(os.path.join(args.out, f"nn_{fam.tag}.npz"))
    ev_train = evaluate(m, Xtr[:2_000_000], Ytr[:2_000_000], fam=fam)
    ev_test = evaluate(m, Xte, Yte, fam=fam)
    log(f"[eval] test: {ev_test}")
    report = {"family": {"mord": fam.mord, "ngp": fam.ngp, "inputs": fam.inputs, "classes": fam.classes,
                         "cascade": fam.cascade},
              "dataset": {"n": args.n, "n_ics": args.n_ics, "steps": args.steps, "dump_every": args.dump_every,
                          "keep_mord": args.keep_mord, "samples": int(Xtr.shape[0] + Xte.shape[0]),
                          "train": int(Xtr.shape[0]), "test": int(Xte.shape[0]), "class_counts": counts},
              "training": {"epochs": args.epochs, "batch": args.batch, "lr": args.lr, "history": hist,
                           "mord_logit_shift": args.mord_logit_shift},
              "eval_train": ev_train, "eval_test": ev_test}
    del Xtr, Ytr, Xte, Yte
    if fam.mord == 5:
        evals = (("in_step", "four_quadrant", "neumann"), ("in_step_blast", "blast", "reflective"))
    else:
        evals = (("in_step", "shear_layer", "periodic"), ("in_step_four_quadrant", "four_quadrant", "neumann"))
    for key, kind, bc in evals:
        report[key] = in_step(m, n=args.eval_n, steps=args.eval_steps, kind=kind, bc=bc, log=log, fam=fam)
    report["seconds"] = time.time() - t0
    json.dump(report, open(os.path.join(args.out, f"nn_{fam.tag}_report.json"), "w"), indent=1)
    log(f"[done] {report['seconds']:.1f}s")


if __name__ == "__main__":
    main()
