"""Learned warm start: data generation from ensembles, operator training and
the warm-versus-cold time-to-steady evaluation.

    python -m cudadsmc_amd.warmstart data  --out DIR [--runs 512 --batch 64 ...]
    python -m cudadsmc_amd.warmstart train --out DIR [--iters 3000 ...]
    python -m cudadsmc_amd.warmstart eval  --out DIR [--cases 32 ...]

Time-to-steady of one arm (cold: uniform freestream loading; warm: local
Maxwellians from the operator's predicted fields) is measured in chunks of
`chunk` steps:

* start rule (both arms, no reference needed): SteadyDetector compares the
  block fields of consecutive chunks with those of the two halves of a
  chunk; sampling restarts from zero at the end of the first chunk whose
  remaining drift is within its noise;
* steady: the fields accumulated since then, averaged over blocks of
  `block`^3 cells, are within `tol` relative L2 (n, |u| and T) of a long
  cold reference run of the same case, and the sampled C_d is within
  3 sigma (batch means, combined with the reference's) of the reference C_d.

The warm arm falls back to a cold start when the prediction is not finite or
not physical. Wall time counts the operator's inference, the loading and the
read-outs of both rules.
"""
from __future__ import annotations

import argparse
import json
import math
import time
from dataclasses import asdict
from pathlib import Path

import numpy as np

from . import Batch
from . import ensemble as ens
from . import neural

BOX = (*ens.BOX_MIN, *ens.BOX_MAX)


# ------------------------------------------------------------------- data
def case_inputs(cases: list[ens.Case]) -> tuple[np.ndarray, np.ndarray]:
    p = neural.normalise_params([c.U for c in cases], [c.n for c in cases], [c.T_inf for c in cases],
                                [c.T_w for c in cases], [c.r for c in cases])
    s = np.array([c.sphere() for c in cases], np.float32)
    return p, s


def generate(out: Path, runs: int, batch: int, transient: int, sampling: int, seed: int, N_range) -> dict:
    """Ensemble runs -> dataset.npz (targets, masks, parameters, C_d)."""
    cases = ens.draw_cases(runs, seed, N_range=N_range, seed0=seed * 1000)
    X, Y, Z = ens.GRID
    tg = np.zeros((runs, Z, Y, X, 5), np.float32)
    mk = np.zeros((runs, Z, Y, X), np.float32)
    cd = np.zeros(runs)
    cds = np.zeros(runs)
    t0 = time.perf_counter()
    steps_total = 0
    np_total = 0
    for b0 in range(0, runs, batch):
        chunk = cases[b0:b0 + batch]
        r = ens.run_batch(chunk, transient, sampling)
        for i, c in enumerate(chunk):
            f = ens.stack_fields(r["fields"][i])
            tg[b0 + i] = neural.fields_to_targets(f[..., 0], f[..., 1], f[..., 2], f[..., 3], f[..., 4], c.n, c.U,
                                                  c.T_inf)
            ff = np.asarray(r["fluid_fraction"][i]).reshape(Z, Y, X)
            cnt = np.asarray(r["counts"][i]).reshape(Z, Y, X)
            # fluid cells with enough sampled particles for a usable temperature
            mk[b0 + i] = ((ff > 0.5) & (cnt >= 20)).astype(np.float32)
            cd[b0 + i] = r["cd"][i]
            cds[b0 + i] = r["cd_sigma"][i]
            np_total += r["n_particles"][i]
        steps_total += len(chunk) * (transient + sampling)
        print(f"[data] {b0 + len(chunk)}/{runs} runs, {time.perf_counter() - t0:.1f} s", flush=True)
    p, s = case_inputs(cases)
    np.savez(out / "dataset.npz", targets=tg, masks=mk, params=p, spheres=s, cd=cd, cd_sigma=cds,
             raw=np.array([[c.U, c.n, c.T_inf, c.T_w, c.r, c.N, c.seed] for c in cases]))
    info = {"runs": runs, "batch": batch, "transient": transient, "sampling": sampling,
            "wall_s": time.perf_counter() - t0, "seed": seed, "N_range": list(N_range)}
    (out / "dataset.json").write_text(json.dumps(info, indent=1))
    return info


# ---------------------------------------------------------------- training
def train(out: Path, iters: int, batch: int, lr: float, width: int, seed: int, holdout: int = 0) -> dict:
    import torch

    d = np.load(out / "dataset.npz", allow_pickle=False)
    tg, mk, p, s = d["targets"], d["masks"], d["params"], d["spheres"]
    n = len(tg) - holdout
    X, Y, Z = ens.GRID
    shape = (Z, Y, X)
    op = neural.FieldOperator(shape=shape, width=width, seed=seed, max_batch=batch)
    print(f"[train] operator with {op.n_params()} parameters, {n} training fields", flush=True)
    g = np.random.default_rng(seed)
    dev_t = torch.from_numpy(tg[:n].reshape(n, -1, 5)).cuda()
    dev_m = torch.from_numpy(mk[:n].reshape(n, -1)).cuda()
    M1 = Z * Y * X
    a_all = op.make_inputs(p[:n], s[:n], BOX).reshape(n, M1, neural.C_IN)
    hist = []
    t0 = time.perf_counter()
    torch.cuda.synchronize()
    for it in range(iters):
        idx = torch.from_numpy(g.choice(n, size=batch, replace=False)).cuda()
        a0 = a_all[idx].reshape(batch * M1, neural.C_IN).contiguous()
        tgt = dev_t[idx].reshape(batch * M1, 5).contiguous()
        msk = dev_m[idx].reshape(batch * M1).contiguous()
        # linear warm-up over 100 iterations, then cosine decay to lr/20
        warm = min(1.0, (it + 1) / 100.0)
        lr_t = warm * lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * it / max(1, iters))))
        loss = op.train_step(a0, tgt, msk, batch, lr_t)
        if it % 50 == 0 or it == iters - 1:
            lv = float(loss.item())
            hist.append((it, lv))
            print(f"[train] it {it} loss {lv:.5f} ({time.perf_counter() - t0:.1f} s)", flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    op.save(out / "operator.npz")
    info = {"iters": iters, "batch": batch, "lr": lr, "width": width, "params": op.n_params(), "wall_s": wall,
            "ms_per_iter": 1e3 * wall / max(iters, 1), "loss": hist}
    (out / "train.json").write_text(json.dumps(info, indent=1))
    return info


# -------------------------------------------------------------- evaluation
def predicted_fields(op: neural.FieldOperator, cases: list[ens.Case]) -> list[dict | None]:
    """Operator prediction -> loadable fields per case (None = fall back to cold)."""
    p, s = case_inputs(cases)
    pred = op.predict(p, s, BOX)
    out = []
    for c, t in zip(cases, pred):
        n, u, v, w, T = neural.targets_to_fields(t, c.n, c.U, c.T_inf)
        ok = all(np.all(np.isfinite(a)) for a in (n, u, v, w, T))
        ok = ok and float(np.max(n)) < 100.0 * c.n and float(np.max(T)) < 100.0 * c.T_inf
        ok = ok and float(np.min(T)) > 0.01 * c.T_inf
        if not ok:
            out.append(None)
            continue
        # clip the few non-physical cells (negative density or temperature)
        out.append({"n": np.maximum(n, 0.0).astype(np.float32).ravel(), "u": u.astype(np.float32).ravel(),
                    "v": v.astype(np.float32).ravel(), "w": w.astype(np.float32).ravel(),
                    "T": np.maximum(T, 0.05 * c.T_inf).astype(np.float32).ravel()})
    return out


def block_mean(a: np.ndarray, b: int) -> np.ndarray:
    X, Y, Z = ens.GRID
    a = a.reshape(Z // b, b, Y // b, b, X // b, b)
    return a.mean(axis=(1, 3, 5))


def field_error(f: dict, ref: dict, frac: np.ndarray, U: float, block: int) -> float:
    """max over n, |u|, T of the relative L2 difference of block averages
    over blocks that are fully fluid."""
    X, Y, Z = ens.GRID
    fr = block_mean(frac.reshape(Z, Y, X), block)
    m = fr > 0.999
    errs = []
    for key in ("n", "T"):
        a = block_mean(np.asarray(f[key], np.float64).reshape(Z, Y, X), block)[m]
        r = block_mean(np.asarray(ref[key], np.float64).reshape(Z, Y, X), block)[m]
        errs.append(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30))
    sp = lambda d: np.sqrt(sum(np.asarray(d[k], np.float64) ** 2 for k in ("u", "v", "w")))  # noqa: E731
    a = block_mean(sp(f).reshape(Z, Y, X), block)[m]
    r = block_mean(sp(ref).reshape(Z, Y, X), block)[m]
    errs.append(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30))
    return float(max(errs))


def block_moments(stats: np.ndarray, block: int) -> np.ndarray:
    """Per-cell moment sums (C, 5: count, sum v_x, v_y, v_z, sum |v|^2;
    X fastest) -> per-block sums (n_blocks, 5)."""
    X, Y, Z = ens.GRID
    a = np.asarray(stats, np.float64).reshape(Z // block, block, Y // block, block, X // block, block, 5)
    return a.sum(axis=(1, 3, 5)).reshape(-1, 5)


def moment_fields(m: np.ndarray, steps: int) -> np.ndarray:
    """Block moment sums over `steps` sampled steps -> (n_blocks, 3): particles
    per step, bulk speed, and the thermal part of <|v|^2> (3kT/m)."""
    N = m[:, 0]
    safe = np.maximum(N, 1.0)
    u = m[:, 1:4] / safe[:, None]
    th = m[:, 4] / safe - (u * u).sum(1)
    return np.stack([N / max(steps, 1), np.linalg.norm(u, axis=1), th], 1)


def moment_change(a: np.ndarray, b: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """Per field (density, speed, temperature): relative L2 difference of
    block fields a and b over the masked blocks."""
    a, b = a[mask], b[mask]
    return np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(b, axis=0), 1e-30)


def block_mask(frac: np.ndarray, block: int) -> np.ndarray:
    """Blocks whose cells are all fluid (flattened like block_moments)."""
    X, Y, Z = ens.GRID
    return (block_mean(np.asarray(frac, np.float64).reshape(Z, Y, X), block) > 0.999).ravel()


class SteadyDetector:
    """Sampling-start rule that needs no reference and treats warm and cold
    starts alike. The run is sampled in windows of `half` steps; two windows
    make a chunk. With block fields F, noise sigma per chunk and a systematic
    drift D per chunk,

        |F(chunk j) - F(chunk j-1)|^2            ~ 2 sigma^2 + D^2
        |F(2nd half of j) - F(1st half of j)|^2  ~ 4 sigma^2 + D^2 / 4,

    so the ratio R = first / second is 1/2 for independent noise and tends
    to 4 while the fields drift. Fluctuations of a block persist for about
    the time particles take to cross it (tens of steps), which makes
    adjacent windows correlated and raises the stationary R towards 1. The
    test passes when R averaged over the three fields (density, speed,
    temperature) is at most 1: D below about the noise of one chunk.
    Sampling starts at the end of that chunk.
    """

    def __init__(self, mask: np.ndarray, block: int, half: int):
        self.mask, self.block, self.half = mask, block, half
        self.prev_cum = None
        self.halves: list[np.ndarray] = []
        self.prev_chunk = None
        self.last = None  # (chunk change, half change) of the last test

    def feed(self, cum_stats: np.ndarray) -> bool:
        """Cumulative per-cell sums at the end of a half window -> True when
        the chunk that just ended passes the test."""
        cum = block_moments(cum_stats, self.block)
        win = cum if self.prev_cum is None else cum - self.prev_cum
        self.prev_cum = cum
        self.halves.append(win)
        if len(self.halves) < 2:
            return False
        A, B = self.halves
        self.halves = []
        chunk = A + B
        fire = False
        if self.prev_chunk is not None:
            fc = moment_fields(chunk, 2 * self.half)
            d_c = moment_change(fc, moment_fields(self.prev_chunk, 2 * self.half), self.mask)
            d_h = moment_change(moment_fields(B, self.half), moment_fields(A, self.half), self.mask)
            self.last = (d_c, d_h)
            fire = bool(np.mean(d_c ** 2 / np.maximum(d_h ** 2, 1e-30)) <= 1.0)
        self.prev_chunk = chunk
        return fire


def freestream_fields(c: ens.Case) -> dict:
    C_ = ens.GRID[0] * ens.GRID[1] * ens.GRID[2]
    one = np.ones(C_, np.float32)
    return {"n": c.n * one, "u": c.U * one, "v": 0 * one, "w": 0 * one, "T": c.T_inf * one}


def reference_runs(cases, transient, sampling, windows=20, batch=32) -> list[dict]:
    refs = []
    for b0 in range(0, len(cases), batch):
        r = ens.run_batch(cases[b0:b0 + batch], transient, sampling, windows=windows, keep_stats=True)
        for i in range(len(r["fields"])):
            refs.append({"fields": r["fields"][i], "cd": r["cd"][i], "cd_sigma": r["cd_sigma"][i],
                         "frac": r["fluid_fraction"][i], "stats": r["stats"][i], "steps": sampling})
        print(f"[eval] reference {b0 + len(r['fields'])}/{len(cases)}", flush=True)
    return refs


def time_to_steady(cases, refs, fields, chunk: int, max_steps: int, tol: float, block: int) -> dict:
    """Advance all cases as one batch and record, per case, the step and wall
    time at which the fields accumulated since the detector's sampling start
    meet the steady criterion (block fields within `tol` of the reference and
    C_d within 3 sigma of it, with at least three drag windows of `chunk`/2
    steps). Also records, per chunk before the sampling start, the
    instantaneous block-field error (density, speed, temperature) against
    the reference (diagnostic, not used by either rule)."""
    half = chunk // 2
    assert half * 2 == chunk and half > 0
    t0 = time.perf_counter()
    sols = [ens.make_solver(c, fields=None if fields is None else fields[i]) for i, c in enumerate(cases)]
    load_s = time.perf_counter() - t0
    b = Batch(sols)
    b.set_sampling(True)  # from step 0: the detector reads differences of the accumulators
    b.synchronize()
    k = len(cases)
    masks = [block_mask(r["frac"], block) for r in refs]
    ref_bf = [moment_fields(block_moments(r["stats"], block), r["steps"]) for r in refs]
    det = [SteadyDetector(masks[i], block, half) for i in range(k)]
    state = [{"sampling_from": None, "steady_step": None, "drag": None, "err": None, "trace": [], "tests": [],
              "cum0": None, "chunk_cum": None} for _ in range(k)]
    wall = 0.0
    step = 0
    while step < max_steps and any(st["steady_step"] is None for st in state):
        ts = time.perf_counter()
        b.run(half)
        b.synchronize()
        step += half
        chunk_end = step % chunk == 0
        for i, (c, s, st) in enumerate(zip(cases, sols, state)):
            if st["steady_step"] is not None:
                continue
            if st["sampling_from"] is None:
                cum = s.cell_stats()
                fire = det[i].feed(cum)
                if chunk_end:
                    bm = block_moments(cum, block)
                    prev = st["chunk_cum"]
                    win = bm if prev is None else bm - prev
                    st["chunk_cum"] = bm
                    st["trace"].append(moment_change(moment_fields(win, chunk), ref_bf[i], masks[i]).tolist())
                    if det[i].last is not None:
                        st["tests"].append([det[i].last[0].tolist(), det[i].last[1].tolist()])
                if fire:
                    st["sampling_from"] = step
                    st["drag"] = ens.DragWindows(c, 1e-6)
                    s.set_sampling(True, reset=True)
                continue
            sampled = step - st["sampling_from"]
            st["drag"].update(s, sampled)
            if not chunk_end:
                continue
            err = field_error(s.fields(), refs[i]["fields"], refs[i]["frac"], c.U, block)
            st["err"] = err
            m, e = st["drag"].cd()
            sig = math.sqrt((e if np.isfinite(e) else 0.0) ** 2 + refs[i]["cd_sigma"] ** 2)
            ok_cd = len(st["drag"].means) >= 3 and abs(m - refs[i]["cd"]) <= 3.0 * sig
            if err <= tol and ok_cd:
                st["steady_step"] = step
                st["steady_wall_s"] = wall + (time.perf_counter() - ts)
        wall += time.perf_counter() - ts
    b.close()
    rows = []
    for st, s in zip(state, sols):
        m, e = st["drag"].cd() if st["drag"] else (float("nan"), float("nan"))
        rows.append({"steady_step": st["steady_step"], "steady_wall_s": st.get("steady_wall_s"),
                     "sampling_from": st["sampling_from"], "final_err": st["err"], "cd": m, "cd_sigma": e,
                     "chunk_err": st["trace"], "detector_tests": st["tests"]})
        s.close()
    done = all(r["steady_step"] is not None for r in rows)
    # the ensemble is steady when its last member is
    return {"rows": rows, "load_s": load_s, "steps": step, "all_steady": done,
            "wall_to_all_steady_s": load_s + wall if done else None, "stepping_wall_s": wall}


def evaluate(out: Path, n_cases: int, seed: int, N: int, ref_transient: int, ref_sampling: int, chunk: int,
             max_steps: int, tol: float, block: int, tag: str = "") -> dict:
    cases = ens.draw_cases(n_cases, seed, N_range=(N, N), seed0=seed * 1000)
    op = neural.FieldOperator.load(out / "operator.npz")
    t0 = time.perf_counter()
    warm = predicted_fields(op, cases)
    import torch

    torch.cuda.synchronize()
    infer_s = time.perf_counter() - t0
    fallback = sum(1 for w in warm if w is None)
    refs = reference_runs(cases, ref_transient, ref_sampling)
    res_c = time_to_steady(cases, refs, None, chunk, max_steps, tol, block)
    print(f"[eval] cold arm: {res_c['steps']} steps, all steady: {res_c['all_steady']}", flush=True)
    res_w = time_to_steady(cases, refs, warm, chunk, max_steps, tol, block)
    print(f"[eval] warm arm: {res_w['steps']} steps, all steady: {res_w['all_steady']}", flush=True)
    rows = []
    for c, ref, w, rc, rw in zip(cases, refs, warm, res_c["rows"], res_w["rows"]):
        # error of the loaded fields themselves, before any step
        e_pred = None if w is None else field_error(w, ref["fields"], ref["frac"], c.U, block)
        e_free = field_error(freestream_fields(c), ref["fields"], ref["frac"], c.U, block)
        rows.append({"case": asdict(c), "pred_err": e_pred, "freestream_err": e_free, "cold": rc, "warm": rw})
    sc = [r["cold"]["steady_step"] for r in rows]
    sw = [r["warm"]["steady_step"] for r in rows]
    both = [(a, b) for a, b in zip(sc, sw) if a is not None and b is not None]
    med = lambda xs: float(np.median(xs)) if xs else None  # noqa: E731
    wc = res_c["wall_to_all_steady_s"]
    ww = None if res_w["wall_to_all_steady_s"] is None else res_w["wall_to_all_steady_s"] + infer_s
    summary = {"cases": n_cases, "N": N, "tol": tol, "block": block, "chunk": chunk, "max_steps": max_steps,
               "ref_steps": [ref_transient, ref_sampling],
               "fallback_to_cold": fallback, "infer_s": infer_s,
               "median_pred_err": med([r["pred_err"] for r in rows if r["pred_err"] is not None]),
               "median_freestream_err": med([r["freestream_err"] for r in rows]),
               "cold_reached": sum(a is not None for a in sc), "warm_reached": sum(b is not None for b in sw),
               "median_start_cold": med([r["cold"]["sampling_from"] for r in rows
                                         if r["cold"]["sampling_from"] is not None]),
               "median_start_warm": med([r["warm"]["sampling_from"] for r in rows
                                         if r["warm"]["sampling_from"] is not None]),
               "median_steps_cold": med([a for a, _ in both]),
               "median_steps_warm": med([b for _, b in both]),
               "steps_ratio_median": med([a / b for a, b in both]),
               "steps_all_cold": max(sc) if all(a is not None for a in sc) else None,
               "steps_all_warm": max(sw) if all(b is not None for b in sw) else None,
               "wall_to_all_steady_cold_s": wc, "wall_to_all_steady_warm_s": ww,
               "wall_ratio": (wc / ww) if (wc and ww) else None,
               "load_cold_s": res_c["load_s"], "load_warm_s": res_w["load_s"]}
    res = {"summary": summary, "rows": rows}
    (out / f"eval{tag}.json").write_text(json.dumps(res, indent=1, default=float))
    print(json.dumps(summary), flush=True)
    return res


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("stage", choices=["data", "train", "eval", "all"])
    ap.add_argument("--out", default="warmstart_out", help="folder for the data set, operator and results")
    ap.add_argument("--runs", type=int, default=512)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--transient", type=int, default=2000)
    ap.add_argument("--sampling", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--nmin", type=float, default=1e4)
    ap.add_argument("--nmax", type=float, default=5e5)
    ap.add_argument("--iters", type=int, default=6000)
    ap.add_argument("--train-batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--width", type=int, default=96)
    ap.add_argument("--cases", type=int, default=32)
    ap.add_argument("--eval-seed", type=int, default=99)
    ap.add_argument("--eval-n", type=int, default=500000)
    ap.add_argument("--ref-transient", type=int, default=2000)
    ap.add_argument("--ref-sampling", type=int, default=8000)
    ap.add_argument("--chunk", type=int, default=40)
    ap.add_argument("--max-steps", type=int, default=4000)
    ap.add_argument("--tol", type=float, default=0.02)
    ap.add_argument("--block", type=int, default=4)
    ap.add_argument("--tag", default="")
    a = ap.parse_args(argv)
    out = Path(a.out)
    out.mkdir(parents=True, exist_ok=True)
    if a.stage in ("data", "all"):
        generate(out, a.runs, a.batch, a.transient, a.sampling, a.seed, (a.nmin, a.nmax))
    if a.stage in ("train", "all"):
        train(out, a.iters, a.train_batch, a.lr, a.width, a.seed)
    if a.stage in ("eval", "all"):
        evaluate(out, a.cases, a.eval_seed, a.eval_n, a.ref_transient, a.ref_sampling, a.chunk, a.max_steps, a.tol,
                 a.block, a.tag)


if __name__ == "__main__":
    main()
