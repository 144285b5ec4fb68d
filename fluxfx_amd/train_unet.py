"""Train the learned pressure preconditioner on one GPU.

    python -m fluxfx_amd.train_unet --n 128 --scenes 128 --steps-per-scene 32 --epochs 3 --out unet.npz

Harvests scenes x steps-per-scene (divergence, solid -> impulse) pairs from
plain FP32 tolerance-multigrid solves of seeded random scenes (seeds from
10000) -- or, with --increments, the warm-started solves' increment problems
of (2n)^3 scenes restricted to n^3, for the network run one level down --
trains with Adam (lr, betas 0.9/0.999, eps 1e-8, one pair per step,
cosine decay to lr/20), reports the masked MSE on held-out scenes (seeds from
20000) and writes the fp32 parameters (.npz) plus a JSON summary beside them.
"""
import argparse
import json
import math
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--scenes", type=int, default=128)
    ap.add_argument("--steps-per-scene", type=int, default=32)
    ap.add_argument("--held-out", type=int, default=4)
    ap.add_argument("--epochs", type=float, default=3.0)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    ap.add_argument("--increments", action="store_true",
                    help="pairs from warm-started increment solves at (2n)^3, restricted to n^3 (for "
                         "learned_level = 1 on a (2n)^3 grid)")
    args = ap.parse_args(argv)
    from .backend.device import HipDevice
    from .learned import UNetPreconditioner
    from . import training

    def log(msg):
        print(msg, flush=True)

    device = HipDevice(0)
    shape = (args.n,) * 3
    t0 = time.time()
    if args.increments:
        train_set = training.harvest_increments(device, 2 * args.n, args.scenes, args.steps_per_scene, seed0=10000,
                                                log=log)
        held_out = training.harvest_increments(device, 2 * args.n, args.held_out, 2, seed0=20000, log=log)
    else:
        train_set = training.harvest(device, args.n, args.scenes, args.steps_per_scene, seed0=10000, log=log)
        held_out = training.harvest(device, args.n, args.held_out, 2, seed0=20000, log=log)
    t_harvest = time.time() - t0
    net = UNetPreconditioner(device, seed=args.seed)
    before = training.evaluate(net, held_out, shape)
    log(f"held-out loss before training {before:.4e}")
    steps = int(math.ceil(args.epochs * len(train_set)))
    hist = []
    t1 = time.time()
    chunk = max(1, steps // 20)
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        # cosine schedule, constant within a chunk
        lr = args.lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * done / steps)))
        h = training.train(net, train_set, shape, k, lr=lr, seed=args.seed + done, log=None, log_every=k)
        done += k
        hist.append((done, lr, h[-1][1]))
        log(f"train: {done}/{steps} steps, lr {lr:.2e}, last loss {h[-1][1]:.4e}, {time.time() - t1:.1f}s")
    t_train = time.time() - t1
    after = training.evaluate(net, held_out, shape)
    log(f"held-out loss after training {after:.4e}")
    net.save(args.out)
    summary = {"n": args.n, "pairs": len(train_set), "data": "increments" if args.increments else "solutions",
               "held_out_pairs": len(held_out), "steps": steps,
               "lr": args.lr, "betas": [0.9, 0.999], "eps": 1e-8, "held_out_loss_before": before,
               "held_out_loss_after": after, "harvest_s": t_harvest, "train_s": t_train,
               "ms_per_step": 1e3 * t_train / max(steps, 1), "history": hist}
    with open(args.out.rsplit(".", 1)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(json.dumps({k: v for k, v in summary.items() if k != "history"}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
