"""Training and evaluation of the learned warm start, on the GPU end to end.

Data: pairs rendered from seeds on host threads (pdf_synth_pairs_float) are solved by the
batched plain path with the reference schedule; pdf_batch_learn_samples turns every solved
pair into the network's input (its 60x80 level: I1, Z1, I2, Z2) and label (its final flow
averaged over each 60x80 pixel's block).  Samples stay in device memory.  Training: L1 loss,
Adam, cosine learning-rate decay, shuffled mini-batches gathered on the device.

Evaluation (``evaluate``): for held-out seeds, the plain solve with the reference schedule,
the learned solve (network at the 60x80 level, every remaining level's iterations divided by
``factor``), and the ablation (the same cut schedule from the coarsest level without the
network); end-point errors against the plain solve, energies and residuals, and the fallback
rule (energy or residual more than ``tol`` above the plain solver's).

    python -m pdflow.train --pairs 50000 --steps 20000 --batch 64 --out weights.bin
"""
import argparse
import threading
import time

import numpy as np

from . import schedule
from .lib import Batch, DeviceBuffer
from .net import Net
from .synth import synth_pairs_float

NET_H, NET_W = 60, 80
TRAIN_SEED0 = 1_000_000       # training seeds: TRAIN_SEED0 + i
EVAL_SEED0 = 9_000_000        # held-out seeds: EVAL_SEED0 + i (disjoint from training)


class Dataset:
    """n samples in device memory: x [n][60][80][4], t [n][60][80][3], m [n][60][80]"""

    def __init__(self, n):
        self.n = n
        px = NET_H * NET_W
        self.x = DeviceBuffer(n * px * 4 * 4)
        self.t = DeviceBuffer(n * px * 3 * 4)
        self.m = DeviceBuffer(n * px * 4)

    def ptrs(self, first=0):
        px = NET_H * NET_W
        return (self.x.ptr.value + first * px * 16, self.t.ptr.value + first * px * 12,
                self.m.ptr.value + first * px * 4)


def _render_async(seed0, n, rows, threads):
    out = {}

    def run():
        out["planes"] = synth_pairs_float(seed0, n, rows, rows * 4 // 3, threads)

    th = threading.Thread(target=run)
    th.start()
    return th, out


def make_dataset(n_pairs, rows=240, seed0=TRAIN_SEED0, chunk=256, threads=None, log=print):
    """renders, solves (plain, reference schedule) and samples n_pairs pairs"""
    data = Dataset(n_pairs)
    chunk = min(chunk, n_pairs)
    batch = Batch([rows] * chunk)
    t0 = time.time()
    last = t0
    done = 0
    th, pending = _render_async(seed0, min(chunk, n_pairs), rows, threads)
    while done < n_pairs:
        cnt = min(chunk, n_pairs - done)
        th.join()
        planes = pending["planes"]
        nxt = done + cnt
        if nxt < n_pairs:       # render the next chunk while this one is solved
            th, pending = _render_async(seed0 + nxt, min(chunk, n_pairs - nxt), rows, threads)
        if cnt < chunk:         # last partial chunk: a batch of its own size
            batch = Batch([rows] * cnt)
        for i in range(cnt):
            batch.set_inputs(i, *(p[i] for p in planes))
        batch.solve(graph=True)
        batch.synchronize()
        batch.learn_samples(NET_H, NET_W, *data.ptrs(done))
        done = nxt
        if time.time() - last > 30 or done == n_pairs:
            last = time.time()
            log(f"data: {done}/{n_pairs} pairs solved and sampled, {last - t0:.0f} s", flush=True)
    return data


def train(net, data, steps=20000, batch=64, lr=1e-3, lr_min=1e-5, seed=0, log=print, log_every=30.0,
          checkpoint=None):
    rng = np.random.default_rng(seed)
    x, t, m = data.ptrs()
    perm = rng.permutation(data.n)
    pos = 0
    t0 = last = time.time()
    losses = []
    for step in range(steps):
        if pos + batch > data.n:
            perm = rng.permutation(data.n)
            pos = 0
        idx = perm[pos:pos + batch]
        pos += batch
        cur = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * step / steps))
        want = (step % 100 == 0) or step == steps - 1
        loss = net.train_step_indexed(x, t, m, data.n, idx, cur, want_loss=want)
        if want:
            losses.append((step, loss))
        if time.time() - last > log_every or step == steps - 1:
            last = time.time()
            log(f"train: step {step + 1}/{steps} loss {losses[-1][1]:.4f} lr {cur:.2e} "
                f"{(last - t0) / (step + 1) * 1e3:.1f} ms/step", flush=True)
            if checkpoint:
                net.save(checkpoint)
    net.synchronize()
    return losses


def cut_schedule(rows, factor):
    """the reference schedule with every level's count divided by factor (rounded down, >= 1)"""
    return [max(1, int(n / factor)) for n in schedule.num_max_iter(schedule.ctf_levels_for_rows(rows))]


def _gt_epe(batch, i, gt, Z1):
    ok = Z1 > 0
    d = np.sqrt(sum((f.astype(np.float64) - g) ** 2 for f, g in zip(batch.flow(i), gt)))[ok]
    return float(d.mean())


def evaluate(net, n_pairs=256, rows=240, factor=4, seed0=EVAL_SEED0, tol=0.05, chunk=256, threads=None):
    """per held-out pair: EPE of the learned and ablation solves against the plain solve and of
    all three against the rendered ground truth, energy and residual ratios to the plain
    solve's, and the fallback rule's decision (then the EPE after the fallback)"""
    keys = ("epe_learned", "epe_ablation", "gt_plain", "gt_learned", "gt_ablation", "energy_ratio",
            "residual_ratio", "energy_ratio_ablation", "residual_ratio_ablation", "fallback",
            "epe_after_fallback", "gt_after_fallback")
    res = {k: [] for k in keys}
    done = 0
    while done < n_pairs:
        cnt = min(chunk, n_pairs - done)
        *planes, gt = synth_pairs_float(seed0 + done, cnt, rows, rows * 4 // 3, threads, with_gt=True)
        cut = [cut_schedule(rows, factor)] * cnt
        plain, learned, ablation = Batch([rows] * cnt), Batch([rows] * cnt, iters=cut), Batch([rows] * cnt, iters=cut)
        for b in (plain, learned, ablation):
            for i in range(cnt):
                b.set_inputs(i, *(p[i] for p in planes))
        learned.set_learned(net)
        for b in (plain, learned, ablation):
            b.solve(graph=True)
            b.synchronize()
        pe, pr = plain.metrics()
        le, lr_ = learned.metrics()
        ae, ar = ablation.metrics()
        res["epe_learned"] += list(learned.epe(plain))
        res["epe_ablation"] += list(ablation.epe(plain))
        res["energy_ratio"] += list(le / pe)
        res["residual_ratio"] += list(lr_ / pr)
        res["energy_ratio_ablation"] += list(ae / pe)
        res["residual_ratio_ablation"] += list(ar / pr)
        for i in range(cnt):
            res["gt_plain"].append(_gt_epe(plain, i, gt[i], planes[1][i]))
            res["gt_learned"].append(_gt_epe(learned, i, gt[i], planes[1][i]))
            res["gt_ablation"].append(_gt_epe(ablation, i, gt[i], planes[1][i]))
        flags = learned.fallback(pe, pr, tol)
        res["fallback"] += list(flags)
        res["epe_after_fallback"] += list(learned.epe(plain))
        for i in range(cnt):
            res["gt_after_fallback"].append(_gt_epe(learned, i, gt[i], planes[1][i]))
        done += cnt
    return {k: np.asarray(v, np.float64) for k, v in res.items()}


def summary(res):
    def stats(a):
        return {"mean": float(a.mean()), "median": float(np.median(a)), "p90": float(np.percentile(a, 90)),
                "p99": float(np.percentile(a, 99)), "max": float(a.max())}
    out = {"pairs": int(res["epe_learned"].size)}
    for k, v in res.items():
        if k == "fallback":
            out["fallback_fraction"] = float(v.mean())
        else:
            out[k] = stats(v)
    return out


def main(argv=None):
    import json
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--pairs", type=int, default=50000)
    ap.add_argument("--rows", type=int, default=240)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--eval-pairs", type=int, default=2000)
    ap.add_argument("--factor", type=float, default=4)
    ap.add_argument("--threads", type=int, default=None)
    ap.add_argument("--init", default=None, help="weights to start from (resume)")
    ap.add_argument("--out", default="warmstart.bin")
    ap.add_argument("--report", default=None, help="write the evaluation summary (JSON) here")
    a = ap.parse_args(argv)
    net = Net(max_batch=max(a.batch, 256), h=NET_H, w=NET_W, seed=a.seed)
    if a.init:
        net.load(a.init)
    t0 = time.time()
    data = make_dataset(a.pairs, a.rows, threads=a.threads)
    t1 = time.time()
    losses = train(net, data, a.steps, a.batch, a.lr, seed=a.seed, checkpoint=a.out)
    t2 = time.time()
    net.save(a.out)
    del data
    rep = summary(evaluate(net, a.eval_pairs, a.rows, a.factor, threads=a.threads))
    rep.update({"train_pairs": a.pairs, "steps": a.steps, "batch": a.batch, "rows": a.rows,
                "data_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1), "eval_s": round(time.time() - t2, 1),
                "loss_first": losses[0][1], "loss_last": losses[-1][1]})
    print(json.dumps(rep), flush=True)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()
