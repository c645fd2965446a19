"""Trains the learned correction of the pressure-history warm start and evaluates it.

    python -m vflip_amd.train_correction [--n 256] [--seeds 160] [--epochs 10] [--out PATH] [--report PATH]

Pairs come from the plume workload's own solver at n^3 (vflip_amd.learned.
generate_correction_pairs): the divergence the history guess leaves and that
guess's error against the converged pressure, on the grid coarsened to 64^3.
Training runs the library's U-Net kernels (forward, backward, Adam) with the
energy-norm loss under the hierarchy's Galerkin operator of that grid;
evaluation runs held-out seeds with and without the correction. Weights go to
an .npz (float16).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "unet_correction256.npz")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seeds", type=int, default=160)
    ap.add_argument("--steps", type=int, default=15)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--report", default="")
    a = ap.parse_args(argv)

    import torch
    from . import learned
    from .unet import UNet

    def log(msg):
        print(msg, flush=True)

    factor = a.n // 64
    t0 = time.time()
    data = learned.generate_correction_pairs(list(range(1, a.seeds + 1)), n=a.n, factor=factor, steps=a.steps,
                                             log=log)
    t_gen = time.time() - t0
    S = data["rhs"].shape[0]
    its = [s["iterations"] for s in data["stats"]]
    log(f"pairs: {S} in {t_gen:.1f} s; history-guess solves: mean {sum(its) / len(its):.2f} iterations")
    ratio = (data["p"].abs().amax(dim=(1, 2, 3)) / data["rhs"].abs().amax(dim=(1, 2, 3)).clamp_min(1e-30)).cpu()
    gain = float(ratio.median())
    log(f"correction gain max|e|/max|r0|: median {gain:.4e}")
    net = UNet(seed=a.seed)
    net.gain = gain
    t1 = time.time()
    hist = learned.train(net, data, epochs=a.epochs, batch=a.batch, lr=a.lr, loss="energy", log=log)
    t_train = time.time() - t1
    del data
    torch.cuda.empty_cache()
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out)
    log(f"saved {a.out}")
    held = list(range(a.seeds + 1, a.seeds + 1 + a.eval))
    rows = learned.evaluate_history(net, held, n=a.n, factor=factor, steps=a.steps)
    mean = lambda k: sum(r[k] for r in rows) / len(rows)
    summary = {k: mean(k) for k in ("its_history", "its_learned", "ms_history", "ms_learned")}
    summary["max_true_res_learned"] = max(r["true_res_learned"] for r in rows)
    log(f"eval {a.n}^3: {json.dumps(summary)}")
    report = dict(n=a.n, factor=factor, seeds=a.seeds, steps=a.steps, pairs=S, epochs=a.epochs, batch=a.batch,
                  lr=a.lr, gain=gain, seconds_generate=t_gen, seconds_train=t_train, loss_per_epoch=hist,
                  eval=dict(summary=summary, rows=rows))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
