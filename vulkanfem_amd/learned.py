"""Training pipeline of the learned coarse-grid correction.

    python -m vulkanfem_amd.learned --geometries 100 --out DIR

1. Pairs from the solver: seeded 64^3 ellipsoid-union geometries (the
   generator of the measured workloads at a quarter of the 256^3 size:
   40 ellipsoids, semi-axes 3-15 voxels), each with one top-layer load and
   three patch loads; 6 levels, so level 3 is the 9^3-node grid (8^3 cells).
   For every load, up to `iters` FP64 PCG iterations of the damped cycle run
   and each iteration contributes the residual the cycle restricted to level 3,
   the exact level-3 correction A_3^-1 r_3 (dense Cholesky) and the cycle's own
   correction (the plain cycle's answer, which judges the network's).
2. Training on the GPU with the library's HIP kernels (Adam, cosine-decayed
   learning rate), on all geometries but the last few, which are held out.
3. The held-out error of the network next to the plain cycle's, and the
   weights, written to DIR (weights.npy, meta.json).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import api
from . import problems

LEVEL = 3
LEVELS = 6


def training_geometry(seed, n=64):
    vox, info = problems.ellipsoid_union(n, 40, 3.0 * n / 64, 15.0 * n / 64, seed)
    if vox.sum() == 0:
        raise ValueError("no solid component touches z = 0")
    return vox, info


def pairs_for_geometry(seed, iters, n=64, log=None):
    vox, info = training_geometry(seed, n)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    A = api.MatrixFreeSparse(vox, problems.ke0(), fixed, LEVELS)
    if A.num_levels <= LEVEL + 1:
        raise ValueError("hierarchy too shallow")
    zs = np.nonzero((ids >= 0).any(axis=(1, 2)))[0]
    top = ids[max(zs.max() - 1, 0):zs.max() + 1]
    loaded = np.sort(top[top >= 0])
    lv = np.zeros((loaded.size, 3))
    lv[:, 2] = -1.0 / loaded.size
    loads = [problems.apply_boundary_conditions(A.n, loaded, lv)]
    loads += list(problems.patch_load_cases(vox, 3, seed=seed + 7919))
    out = [A.collect_coarse_pairs(f, LEVEL, iters) for f in loads]
    return tuple(np.concatenate([o[k] for o in out]) for k in range(4)), info


def rel_err(pred, y, m):
    d = ((pred - y) * m[:, None]).reshape(len(y), -1)
    t = (y * m[:, None]).reshape(len(y), -1)
    return float(np.mean(np.linalg.norm(d, axis=1) / np.maximum(np.linalg.norm(t, axis=1), 1e-30)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--geometries", type=int, default=100)
    ap.add_argument("--held-out", type=int, default=5)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    t0 = time.time()
    train, held = [], []
    seed = args.seed
    while len(train) + len(held) < args.geometries:
        try:
            (x, y, yv, m), info = pairs_for_geometry(seed, args.iters)
        except ValueError:
            seed += 1
            continue
        (held if len(held) < args.held_out else train).append((x, y, yv, m))
        print(f"[learned] geometry {seed}: fill {info['fill']:.3f}, {len(x)} pairs "
              f"({sum(len(t[0]) for t in train)} training, {sum(len(t[0]) for t in held)} held out), "
              f"{time.time() - t0:.0f} s", flush=True)
        seed += 1
    X, Y, YV, M = (np.concatenate([t[k] for t in train]) for k in range(4))
    Xh, Yh, YVh, Mh = (np.concatenate([t[k] for t in held]) for k in range(4))
    net = api.UNet(seed=args.seed)
    before = rel_err(net.forward(Xh), Yh, Mh)
    t1 = time.time()
    losses = []
    for e in range(args.epochs):  # one library call per epoch, so progress is visible
        lr = args.lr * 0.5 * (1 + np.cos(np.pi * e / args.epochs))
        losses += list(net.train(X, Y, M, epochs=1, batch=args.batch, lr=lr, seed=args.seed + e))
        print(f"[learned] epoch {e + 1}/{args.epochs}: loss {losses[-1]:.4f}, {time.time() - t1:.0f} s", flush=True)
    after = rel_err(net.forward(Xh), Yh, Mh)
    cycle = rel_err(YVh, Yh, Mh)
    w = net.get_weights()
    np.save(os.path.join(args.out, "weights.npy"), w)
    meta = {"level": LEVEL, "levels": LEVELS, "grid": list(X.shape[2:]), "training_pairs": int(len(X)),
            "held_out_pairs": int(len(Xh)), "geometry_seeds": [args.seed, seed], "epochs": args.epochs,
            "batch": args.batch, "lr": args.lr, "epoch_loss": [float(v) for v in losses],
            "held_out_rel_err_untrained": before, "held_out_rel_err_trained": after,
            "held_out_rel_err_plain_cycle": cycle, "data_s": round(t1 - t0, 1), "train_s": round(time.time() - t1, 1),
            "params": int(w.size)}
    with open(os.path.join(args.out, "meta.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print("[learned]", json.dumps({k: v for k, v in meta.items() if k != "epoch_loss"}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
