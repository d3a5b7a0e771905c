#!/usr/bin/env python3
"""Train the learned pass scheduler on traces of the plain GPU presolve.

1. Traces: the plain presolve (parity parameters, collect_trace=1) runs on the
   GPU over `--seeds` instances, alternating two families: the Adams-Johnson
   QAP LP with n cycling through [qap-min, qap-max] and the synthetic
   generator (uniform column degree 6, 40% equality rows, planted structure)
   at `--rows` x 1.5 rows.  Instances are generated on host threads ahead of
   the GPU.
2. Samples: one per scheduler round (learned.round_samples).
3. Training: Adam on the GPU through gp_mlp_loss_grad / gp_mlp_adam_step,
   minibatches of `--batch`, instances split 90/10 into train / validation.
4. Thresholds: per group, keep `--keep` of the productive training rounds.
5. Output: the weight blob and a JSON summary in --out (and, with
   --install, in the package's learned/ directory).
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import gpupre as G  # noqa: E402
import learned as L  # noqa: E402


def log(msg):
    print(f"[train] {msg}", flush=True)


def instance(seed, args):
    if seed % 2 == 0:
        span = args.qap_max - args.qap_min + 1
        n = args.qap_min + (seed // 2) % span
        return "qap", n, G.generate_qap(n, seed=seed)
    rows = args.rows
    return "gen", rows, G.generate(rows, rows * 3 // 2, seed=seed, degree_mode=0, mean_degree=6.0)


def collect(args):
    X, Y, M, inst = [], [], [], []
    stats = {"qap": [0, 0.0, 0], "gen": [0, 0.0, 0]}  # instances, presolve seconds, rounds
    t0 = time.time()
    params = G.parity_params(collect_trace=1)
    with cf.ThreadPoolExecutor(max_workers=args.threads) as pool:
        futs = [pool.submit(instance, s, args) for s in range(min(args.seeds, args.threads * 2))]
        nxt = len(futs)
        for s in range(args.seeds):
            fam, size, lp = futs[s].result()
            futs[s] = None
            if nxt < args.seeds:
                futs.append(pool.submit(instance, nxt, args))
                nxt += 1
            with G.DeviceLP(lp) as d:
                sm = G.preprocess(d, params)
                try:
                    trace = sm.trace()
                    secs = sm.elapsed_seconds
                finally:
                    sm.free()
            x, y, m = L.round_samples(trace, lp.m, lp.n)
            X.append(x)
            Y.append(y)
            M.append(m)
            inst.append(np.full(x.shape[0], s, dtype=np.int32))
            st = stats[fam]
            st[0] += 1
            st[1] += secs
            st[2] += len(trace)
            if (s + 1) % 50 == 0 or s + 1 == args.seeds:
                log(f"traces {s + 1}/{args.seeds} ({time.time() - t0:.0f}s): {sum(a.shape[0] for a in X)} rounds")
    return np.concatenate(X), np.concatenate(Y), np.concatenate(M), np.concatenate(inst), stats


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=2000)
    ap.add_argument("--rows", type=int, default=200_000)
    ap.add_argument("--qap-min", type=int, default=12)
    ap.add_argument("--qap-max", type=int, default=24)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--keep", type=float, default=0.98)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--data", default=None, help="reuse a samples .npz written by an earlier run")
    ap.add_argument("--out", default="scheduler_training")
    ap.add_argument("--install", action="store_true")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    if G.device_count() < 1:
        sys.exit("no HIP device visible")

    if args.data:
        z = np.load(args.data)
        X, Y, M, inst = z["X"], z["Y"], z["M"], z["inst"]
        stats = json.loads(str(z["stats"]))
    else:
        X, Y, M, inst, stats = collect(args)
        np.savez_compressed(os.path.join(args.out, "samples.npz"), X=X, Y=Y, M=M, inst=inst,
                            stats=json.dumps(stats))
    log(f"samples: {X.shape[0]} rounds from {len(np.unique(inst))} instances; positives per group "
        f"{Y.sum(0).astype(int).tolist()} of {M.sum(0).astype(int).tolist()}")

    rng = np.random.default_rng(args.init_seed)
    ids = np.unique(inst)
    val_ids = set(rng.choice(ids, size=max(1, len(ids) // 10), replace=False).tolist())
    is_val = np.array([i in val_ids for i in inst])
    Xt, Yt, Mt = X[~is_val], Y[~is_val], M[~is_val]
    Xv, Yv, Mv = X[is_val], Y[is_val], M[is_val]

    tr = L.GpuTrainer(L.init_params(args.init_seed))
    n = Xt.shape[0]
    t0 = time.time()
    history = []
    for ep in range(args.epochs):
        perm = rng.permutation(n)
        tr.set_data(Xt[perm], Yt[perm], Mt[perm])
        lr = args.lr if ep < args.epochs * 2 // 3 else args.lr * 0.25
        losses = []
        for start in range(0, n, args.batch):
            losses.append(tr.step(start, min(args.batch, n - start), lr))
        history.append(float(np.mean(losses)))
        if ep % 5 == 0 or ep + 1 == args.epochs:
            log(f"epoch {ep}: mean minibatch loss {history[-1]:.4f} ({time.time() - t0:.0f}s)")

    params = tr.params()
    lt = tr.logits(Xt)
    lv = tr.logits(Xv)
    thr = L.choose_thresholds(lt, Yt, Mt, args.keep)
    tr.free()

    def rates(lg, Yy, Mm):
        out = {}
        for g in range(G.GP_MLP_OUT):
            sel = Mm[:, g] > 0
            if not sel.any():
                continue
            run = lg[sel, g] >= thr[g]
            pos = Yy[sel, g] > 0
            out[L.GROUP_NAMES[g]] = {"rounds": int(sel.sum()), "productive": int(pos.sum()),
                                     "skipped": round(float(1.0 - run.mean()), 4),
                                     "productive_kept": round(float(run[pos].mean()), 4) if pos.any() else None}
        return out

    blob = L.make_blob(params, thr)
    meta = {
        "network": "MLP 32-64-64-7, FP32, ReLU; BCE-with-logits masked to the round's phase groups",
        "families": {"qap": f"Adams-Johnson QAP LP, n in [{args.qap_min}, {args.qap_max}]",
                     "gen": f"synthetic generator, {args.rows} x {args.rows * 3 // 2}, uniform degree 6"},
        "seeds": args.seeds, "instances_per_family": {k: v[0] for k, v in stats.items()},
        "plain_presolve_seconds": {k: round(v[1], 3) for k, v in stats.items()},
        "rounds_per_family": {k: v[2] for k, v in stats.items()},
        "samples": int(X.shape[0]), "train_samples": int(Xt.shape[0]), "val_samples": int(Xv.shape[0]),
        "epochs": args.epochs, "batch": args.batch, "lr": args.lr, "keep": args.keep, "init_seed": args.init_seed,
        "loss_first_epoch": history[0], "loss_last_epoch": history[-1],
        "thresholds": [float(t) for t in thr], "train": rates(lt, Yt, Mt), "validation": rates(lv, Yv, Mv),
    }
    L.save(blob, meta, os.path.join(args.out, "scheduler_mlp.f32"), os.path.join(args.out, "scheduler_mlp.json"))
    if args.install:
        L.save(blob, meta)
    log(json.dumps(meta["validation"]))
    log("done")


if __name__ == "__main__":
    main()
