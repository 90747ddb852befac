"""Calibration of the learned path's acceptance rule, and its evaluation, on the GPU.

The benchmark configuration states the learned path's quality rule relative to the plain solver:
a pair is re-solved by the plain path when its final energy or warped-intensity residual exceeds
the plain solver's by more than 5%.  Applying that rule needs the plain solve of every pair, which
costs more than the learned path saves.  This module replaces it with a rule computed from the
learned solve alone, with constants taken from training pairs only:

  * the finest level's iteration budget: the smallest budget, in steps of ``step`` iterations
    from the cut (4 by default, one blocked launch), at which at least ``target`` of the training
    pairs meet the plain-referenced rule.
    Per pair, the energy ratio to the plain solve has a spread of about 14% (log scale) that no
    quantity of the learned solve predicts (the two paths linearise the data terms around
    different coarse solutions), so the budget, not a per-pair test, is what brings the learned
    path to the plain rule's level;
  * per-pair bounds (pdf_batch_fallback_bounds): the finest-level energy per pixel, and the
    warped residual as a fraction of the zero-motion residual mean |I2(x) - I1(x)|, may not exceed
    1.05 times the largest values the plain solver itself reaches on the training pairs.  A flow
    that explains the intensity change worse than any plain solution did (a failed prediction)
    is rejected.

``evaluate`` then applies the calibrated rule to held-out pairs and counts how often it agrees with
the plain-referenced rule, with the end-point errors of the pairs on which they disagree, and how
many pairs of a deliberately corrupted network it rejects.

    python -m pdflow.calibrate --out pd-flow_amd/pdflow/weights/warmstart_qvga_rule.json
"""
import argparse
import json
import time

import numpy as np

from .lib import Batch
from .net import Net
from .synth import synth_pairs_float
from .train import EVAL_SEED0, TRAIN_SEED0, cut_schedule

ROWS, COLS = 480, 640
PLAIN_TOL = 0.05          # the plain-referenced rule's tolerance (benchmark configuration)
BOUND_MARGIN = 1.05       # the bounds' margin over the largest training value
CALIB_SEED0 = TRAIN_SEED0            # training pairs for the constants
EVAL_SEED_OFFSET = 20000             # held-out pairs: EVAL_SEED0 + EVAL_SEED_OFFSET + i


def _log(msg):
    print(msg, flush=True)


def _fill(batches, planes, n):
    for i in range(n):
        fr = [p[i] for p in planes]
        for b in batches:
            b.set_inputs(i, *fr)


def _plain_rule(e, r, pe, pr):
    return (e <= (1 + PLAIN_TOL) * pe) & (r <= (1 + PLAIN_TOL) * pr)


def calibrate(net, n_pairs=2048, factor=2.5, target=0.95, max_finest=96, step=4, chunk=256, log=_log):
    """the rule's constants from training pairs (see the module text); finest-level budgets are
    tried in steps of `step` iterations from the cut"""
    cut = cut_schedule(ROWS, factor)
    budgets = list(range(cut[-1], max_finest + 1, step))
    plain, learned = Batch([ROWS] * chunk), Batch([ROWS] * chunk, iters=[cut] * chunk)
    learned.set_learned(net)
    allp = list(range(chunk))
    passed = {b: [] for b in budgets}
    e_pix, r_ratio = [], []
    for done in range(0, n_pairs, chunk):
        planes = synth_pairs_float(CALIB_SEED0 + done, chunk, ROWS, COLS)
        _fill((plain, learned), planes, chunk)
        plain.solve(graph=True)
        learned.solve(graph=True)
        plain.synchronize()
        pe, pr, p0 = plain.quality()
        e_pix.append(pe / (ROWS * COLS))
        r_ratio.append(pr / p0)
        for k, b in enumerate(budgets):
            if k:
                learned.continue_finest(allp, step)   # equals a solve whose finest level ran b
            e, r, _ = learned.quality()
            passed[b].append(_plain_rule(e, r, pe, pr))
        log(f"calibrate: {done + chunk} of {n_pairs} training pairs")
    rate = {b: float(np.concatenate(v).mean()) for b, v in passed.items()}
    finest = next((b for b in budgets if rate[b] >= target), budgets[-1])
    e_pix, r_ratio = np.concatenate(e_pix), np.concatenate(r_ratio)
    return {
        "factor": factor,
        "schedule": cut[:-1] + [finest],
        "finest_iters": finest,
        "target_plain_rule_pass": target,
        "energy_per_pixel_max": float(BOUND_MARGIN * e_pix.max()),
        "residual_ratio_max": float(BOUND_MARGIN * r_ratio.max()),
        "training": {"pairs": n_pairs, "seeds": f"{CALIB_SEED0}..{CALIB_SEED0 + n_pairs - 1}",
                     "plain_rule_pass_by_finest_iters": {str(b): round(v, 4) for b, v in rate.items()},
                     "plain_energy_per_pixel": {"median": float(np.median(e_pix)), "max": float(e_pix.max())},
                     "plain_residual_ratio": {"median": float(np.median(r_ratio)), "max": float(r_ratio.max())}},
    }


def _gt_epe(b, i, gt, z1):
    ok = z1 > 0
    f = b.flow(i)
    return float(np.sqrt(sum((f[k].astype(np.float64) - gt[k]) ** 2 for k in range(3)))[ok].mean())


def evaluate(net, rule, n_pairs=2048, chunk=256, corrupt_scale=3.0, log=_log):
    """the calibrated rule on held-out pairs against the plain-referenced rule"""
    sched = rule["schedule"]
    plain, learned = Batch([ROWS] * chunk), Batch([ROWS] * chunk, iters=[sched] * chunk)
    learned.set_learned(net)
    emax, rmax = rule["energy_per_pixel_max"], rule["residual_ratio_max"]
    cols = {k: [] for k in ("self_ok", "plain_ok", "epe_plain", "gt_learned", "gt_plain", "e_ratio")}
    for done in range(0, n_pairs, chunk):
        *planes, gt = synth_pairs_float(EVAL_SEED0 + EVAL_SEED_OFFSET + done, chunk, ROWS, COLS, with_gt=True)
        _fill((plain, learned), planes, chunk)
        plain.solve(graph=True)
        learned.solve(graph=True)
        plain.synchronize()
        learned.synchronize()
        pe, pr, _ = plain.quality()
        e, r, r0 = learned.quality()
        cols["self_ok"].append((e <= emax * ROWS * COLS) & (r <= rmax * r0))
        cols["plain_ok"].append(_plain_rule(e, r, pe, pr))
        cols["epe_plain"].append(learned.epe(plain).astype(np.float64))
        cols["e_ratio"].append((e / pe).astype(np.float64))
        cols["gt_learned"].append(np.array([_gt_epe(learned, i, gt[i], planes[1][i]) for i in range(chunk)]))
        cols["gt_plain"].append(np.array([_gt_epe(plain, i, gt[i], planes[1][i]) for i in range(chunk)]))
        log(f"evaluate: {done + chunk} of {n_pairs} held-out pairs")
    c = {k: np.concatenate(v) for k, v in cols.items()}
    s, p = c["self_ok"], c["plain_ok"]

    def mm(a):
        return {"mean": round(float(a.mean()) * 1e3, 3), "p90": round(float(np.percentile(a, 90)) * 1e3, 3)} if a.size else None

    # a corrupted network (every parameter scaled, as a failed prediction would look) on one chunk
    bad = Net(max_batch=chunk, h=60, w=80)
    bad.set_params(net.params() * np.float32(corrupt_scale))
    learned.set_learned(bad)
    planes = synth_pairs_float(EVAL_SEED0 + EVAL_SEED_OFFSET + n_pairs, chunk, ROWS, COLS)
    _fill((learned,), planes, chunk)
    learned.solve(graph=True)
    learned.synchronize()
    e, r, r0 = learned.quality()
    rejected_bad = float(np.mean(~((e <= emax * ROWS * COLS) & (r <= rmax * r0))))
    learned.set_learned(net)
    return {
        "pairs": int(s.size), "seeds": f"{EVAL_SEED0 + EVAL_SEED_OFFSET}..{EVAL_SEED0 + EVAL_SEED_OFFSET + n_pairs - 1}",
        "rule_accepts": round(float(s.mean()), 4),
        "plain_rule_accepts": round(float(p.mean()), 4),
        "agree": round(float((s == p).mean()), 4),
        "accepted_by_rule_rejected_by_plain_rule": {"fraction": round(float((s & ~p).mean()), 4),
                                                     "epe_vs_plain_mm": mm(c["epe_plain"][s & ~p]),
                                                     "epe_vs_truth_mm": mm(c["gt_learned"][s & ~p]),
                                                     "plain_epe_vs_truth_mm": mm(c["gt_plain"][s & ~p])},
        "rejected_by_rule_accepted_by_plain_rule": round(float((~s & p).mean()), 4),
        "epe_vs_plain_mm": mm(c["epe_plain"]),
        "epe_vs_truth_mm": {"learned": mm(c["gt_learned"]), "plain": mm(c["gt_plain"])},
        "energy_ratio_to_plain": {"median": round(float(np.median(c["e_ratio"])), 4),
                                  "p90": round(float(np.percentile(c["e_ratio"], 90)), 4)},
        "corrupted_network": {"scale": corrupt_scale, "pairs": chunk, "rejected": round(rejected_bad, 4)},
    }


def main(argv=None):
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--weights", default=os.path.join(here, "weights", "warmstart_qvga.bin"))
    ap.add_argument("--factor", type=float, default=2.5)
    ap.add_argument("--target", type=float, default=0.95)
    ap.add_argument("--step", type=int, default=4, help="finest-level budgets tried, in iterations")
    ap.add_argument("--train-pairs", type=int, default=2048)
    ap.add_argument("--eval-pairs", type=int, default=2048)
    ap.add_argument("--out", default=os.path.join(here, "weights", "warmstart_qvga_rule.json"))
    a = ap.parse_args(argv)
    net = Net(max_batch=256, h=60, w=80)
    net.load(a.weights)
    t0 = time.time()
    rule = calibrate(net, a.train_pairs, a.factor, a.target, step=a.step)
    t1 = time.time()
    rule["heldout"] = evaluate(net, rule, a.eval_pairs)
    rule["weights"] = os.path.basename(a.weights)
    rule["calibrate_s"], rule["evaluate_s"] = round(t1 - t0, 1), round(time.time() - t1, 1)
    with open(a.out, "w") as f:
        json.dump(rule, f, indent=1)
    print(json.dumps(rule), flush=True)


if __name__ == "__main__":
    main()
