"""Trains the learned warm start on solver-generated pairs and evaluates it.

    python -m vflip_amd.train_warm_start [--pairs 2000] [--epochs 8] [--out PATH] [--report PATH]

Pairs: seeds 1..pairs of the 64^3 plume workload (vflip_amd.learned); held-out
evaluation on seeds pairs+1..pairs+eval at 64^3 (and at --eval-n with the
network on the grid coarsened to 64^3). Everything runs on the GPU through
libvflip; the weights go to an .npz (float16).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "unet_plume64.npz")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=16)
    ap.add_argument("--eval-n", type=int, nargs="*", default=[], help="also evaluate on these grids (multiples of 64)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--loss", default="energy", choices=["energy", "mse"])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--report", default="")
    a = ap.parse_args(argv)

    import torch
    from . import learned
    from .unet import UNet

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    data = learned.generate_pairs(list(range(1, a.pairs + 1)), log=log)
    t_gen = time.time() - t0
    its = [s["iterations"] for s in data["stats"]]
    log(f"pairs: {a.pairs} in {t_gen:.1f} s; mixed-precision iterations per pair: mean {sum(its) / len(its):.2f}")
    # pressure unit of the normalisation: median over the pairs of max|p| / max|rhs|
    ratio = (data["p"].abs().amax(dim=(1, 2, 3)) / data["rhs"].abs().amax(dim=(1, 2, 3))).cpu()
    gain = float(ratio.median())
    log(f"pressure gain max|p|/max|rhs|: median {gain:.4e} (range {float(ratio.min()):.3e}..{float(ratio.max()):.3e})")
    net = UNet(seed=a.seed)
    net.gain = gain
    t1 = time.time()
    hist = learned.train(net, data, epochs=a.epochs, batch=a.batch, lr=a.lr, loss=a.loss, log=log)
    t_train = time.time() - t1
    del data
    torch.cuda.empty_cache()
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out)
    log(f"saved {a.out}")
    held = list(range(a.pairs + 1, a.pairs + 1 + a.eval))
    report = dict(pairs=a.pairs, epochs=a.epochs, batch=a.batch, lr=a.lr, loss=a.loss, gain=gain,
                  seconds_generate=t_gen,
                  seconds_train=t_train, loss_per_epoch=hist)
    for n in [64] + list(a.eval_n):
        rows = learned.evaluate(net, held, n=n, factor=n // 64)
        mean = lambda k: sum(r[k] for r in rows) / len(rows)
        summary = {k: mean(k) for k in ("its_plain", "its_learned", "ms_plain", "ms_learned", "guess_rel_err")}
        summary["max_true_res_learned"] = max(r["true_res_learned"] for r in rows)
        report[f"eval_{n}"] = dict(summary=summary, rows=rows)
        log(f"eval {n}^3: {json.dumps(summary)}")
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
