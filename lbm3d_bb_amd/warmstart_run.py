"""The learned warm-start workload end to end on one GPU.

  python -m lbm3d_bb_amd.warmstart_run --size 128 --train 2000 --eval 200 --out DIR
  python -m lbm3d_bb_amd.warmstart_run --size 128 --train 1000 --train-from 1000 --params P --eval 0 --out DIR
  python -m lbm3d_bb_amd.warmstart_run --size 128 --train 0 --params P --eval 200 --out DIR
  python -m lbm3d_bb_amd.warmstart_run --size 64 --batch 4 --timing 10   (network timing only)

1. generates steady FP32 fields of seeded sphere packs with the solver from
   uniform starts (training seeds 0..train-1, held-out seeds from 1_000_000);
2. trains the 4-level 16/32/64/128-channel U-Net on them with this library's
   BF16 MFMA convolution kernels (Adam, cosine decay);
3. runs every held-out pack to a converged drag twice, from the uniform start
   and from the network's prediction, with the same residual rule, and reports
   the steps of both and the difference of the converged drags.

Training can continue over several runs (--params from the last run,
--train-from to take the next seeds), each generating its own fields; with
--train 0 only the evaluation runs.  Writes DIR/summary.json and
DIR/unet_params.npy; prints one progress line per sample and epoch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from .packs import U_IN, re_max
from .warmstart import PACK_NET, evaluate, generate_dataset, pack_net, train

HELD_OUT_BASE = 1_000_000


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--train", type=int, default=2000, help="training packs (seeds --train-from ...)")
    ap.add_argument("--train-from", type=int, default=0, help="first training seed")
    ap.add_argument("--gen-budget", type=float, default=None, metavar="S",
                    help="stop generating training fields after S seconds")
    ap.add_argument("--eval", type=int, default=200, help="held-out packs to evaluate (0: none)")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=2)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--max-steps", type=int, default=100000)
    ap.add_argument("--out", default="warmstart_out")
    ap.add_argument("--params", default=None, help="start from these network parameters (.npy)")
    ap.add_argument("--timing", type=int, default=0, metavar="K",
                    help="only time K forward calls and K train steps of the network at --size / --batch on "
                         "synthetic inputs (no data, no solver) and print one JSON line")
    args = ap.parse_args(argv)
    if args.timing > 0:
        return timing(args.size, args.batch, args.timing)
    os.makedirs(args.out, exist_ok=True)

    def log(msg):
        print(msg, flush=True)

    L = args.size
    kw = dict(tol=args.tol, every=args.every, max_steps=args.max_steps)
    net = pack_net((L, L, L), batch=args.batch, seed=0, lr=args.lr)
    if args.params:
        net.load(args.params)
    summary = {
        "workload": f"{L}^3 sphere packs (4-32 spheres, r in [5,12]*{L}/128), Re uniform in "
                    f"[1,{re_max(L):.0f}], inlet u={U_IN}, zero-gradient outlet, periodic y/z, TRT 3/16, FP32; "
                    f"drag residual {args.tol} over {args.every}-step window means, 3 in a row",
        "network": PACK_NET, "params_in": args.params, "batch": args.batch, "epochs": args.epochs, "lr": args.lr,
        "seconds": {},
    }
    if args.train > 0:
        t0 = time.perf_counter()
        data = generate_dataset(range(args.train_from, args.train_from + args.train), L, log=log,
                                budget_s=args.gen_budget, **kw)
        summary["seconds"]["generate"] = time.perf_counter() - t0
        used = int(data["X"].shape[0])
        summary.update(train_seeds=[int(args.train_from), int(args.train_from + len(data["seeds"]))],
                       train_samples_used=used, uniform_steps_median=float(np.median(data["steps"])))
        t0 = time.perf_counter()
        ckpt = os.path.join(args.out, "unet_params.npy")
        hist = train(net, data["X"], data["Y"], args.epochs, args.batch, lr=args.lr, log=log,
                     on_epoch=lambda ep, loss: net.save(ckpt))
        summary["seconds"]["train"] = time.perf_counter() - t0
        ms, flops = net.last_timing()
        summary.update(loss_first_epoch=hist[0], loss_last_epoch=hist[-1], loss_per_epoch=hist, train_step_ms=ms,
                       train_step_tflops=flops / ms / 1e9 if ms > 0 else None)
        del data
    if args.eval > 0:
        t0 = time.perf_counter()
        ev = evaluate(net, range(HELD_OUT_BASE, HELD_OUT_BASE + args.eval), L, log=log, **kw)
        summary["seconds"]["evaluate"] = time.perf_counter() - t0
        summary.update({k: v for k, v in ev.items() if k != "rows"})
        summary["rows"] = ev["rows"]
    with open(os.path.join(args.out, "summary.json"), "w", encoding="utf-8") as fh:
        json.dump(summary, fh, indent=1)
    log(json.dumps({k: v for k, v in summary.items() if k not in ("rows", "loss_per_epoch")}))
    return 0


def timing(L: int, batch: int, k: int) -> int:
    """Device time (HIP events) of the network's forward and train step at
    L^3 and the given batch, on a random mask and random targets; the best of
    k calls after one untimed call of each."""
    rng = np.random.default_rng(0)
    shape = (L, L, L)
    net = pack_net(shape, batch=batch, seed=0)
    x = np.zeros((batch, net.cin) + shape, dtype=np.float32)
    x[:, 0] = rng.random((batch,) + shape) < 0.2
    x[:, 1:] = 0.5
    t = (0.1 * rng.standard_normal((batch, 4) + shape)).astype(np.float32)
    out = {"workload": f"U-Net 4 levels 16/32/64/128 channels, {net.cin} input channels, y/z wrap, {L}^3, "
                       f"batch {batch}", "calls": k}
    for name, fn in (("forward", lambda: net.forward(x)), ("train_step", lambda: net.train_step(x, t))):
        fn()
        best, flops = float("inf"), 0.0
        for _ in range(k):
            fn()
            ms, flops = net.last_timing()
            best = min(best, ms)
        out[name] = {"ms": round(best, 4), "gflop": round(flops / 1e9, 2), "tflops": round(flops / best / 1e9, 2)}
    net.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
