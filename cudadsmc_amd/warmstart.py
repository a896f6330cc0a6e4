"""Learned warm start: data generation from ensembles, operator training and
the warm-versus-cold time-to-steady evaluation.

    python -m cudadsmc_amd.warmstart data  --out DIR [--runs 512 --batch 64 ...]
    python -m cudadsmc_amd.warmstart train --out DIR [--iters 3000 ...]
    python -m cudadsmc_amd.warmstart eval  --out DIR [--cases 32 ...]

Time-to-steady of one arm (cold: uniform freestream loading; warm: local
Maxwellians from the operator's predicted fields) is measured in steps of
`chunk`:

* start rule (both arms): sampling starts at the first chunk boundary where
  the simulated particle count changed by less than `settle` (relative) over
  each of the last two chunks;
* steady: the accumulated fields n, |u| and T, averaged over blocks of
  `block`^3 cells, are within `tol` relative L2 of a long cold reference
  run of the same case, and the sampled C_d is within 3 sigma (batch means,
  combined with the reference's) of the reference C_d.

The warm arm falls back to a cold start when the prediction is not finite or
not physical. Wall time counts the operator's inference and the loading.
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


def reference_runs(cases, transient, sampling, windows=20, batch=32) -> list[dict]:
    refs = []
    for b0 in range(0, len(cases), batch):
        r = ens.run_batch(cases[b0:b0 + batch], transient, sampling, windows=windows)
        for i in range(len(r["fields"])):
            refs.append({"fields": r["fields"][i], "cd": r["cd"][i], "cd_sigma": r["cd_sigma"][i],
                         "frac": r["fluid_fraction"][i]})
        print(f"[eval] reference {b0 + len(r['fields'])}/{len(cases)}", flush=True)
    return refs


def time_to_steady(cases, refs, fields, chunk: int, max_steps: int, tol: float, block: int, settle: float,
                   window: int) -> list[dict]:
    """Advance all cases as one batch in chunks and record, per case, the step
    and wall time at which its accumulated fields meet the steady criterion."""
    t0 = time.perf_counter()
    sols = [ens.make_solver(c, fields=None if fields is None else fields[i]) for i, c in enumerate(cases)]
    load_s = time.perf_counter() - t0
    b = Batch(sols)
    b.synchronize()
    k = len(cases)
    state = [{"sampling_from": None, "steady_step": None, "n_hist": [sols[i].counts()[1]],
              "drag": None, "err": None} for i in range(k)]
    wall = 0.0
    step = 0
    while step < max_steps and any(st["steady_step"] is None for st in state):
        ts = time.perf_counter()
        b.run(chunk)
        b.synchronize()
        wall += time.perf_counter() - ts
        step += chunk
        for i, (c, s, st) in enumerate(zip(cases, sols, state)):
            if st["steady_step"] is not None:
                continue
            if st["sampling_from"] is None:
                st["n_hist"].append(s.counts()[1])
                h = st["n_hist"]
                if len(h) >= 3 and all(abs(h[-j] - h[-j - 1]) <= settle * h[-j] for j in (1, 2)):
                    st["sampling_from"] = step
                    st["drag"] = ens.DragWindows(c, 1e-6)
                    s.set_sampling(True)
                continue
            sampled = step - st["sampling_from"]
            if sampled % window == 0:
                st["drag"].update(s, sampled)
            err = field_error(s.fields(), refs[i]["fields"], refs[i]["frac"], c.U, block)
            st["err"] = err
            m, e = st["drag"].cd()
            sig = math.sqrt((e if np.isfinite(e) else 0.0) ** 2 + refs[i]["cd_sigma"] ** 2)
            ok_cd = len(st["drag"].means) >= 3 and abs(m - refs[i]["cd"]) <= 3.0 * sig
            if err <= tol and ok_cd:
                st["steady_step"] = step
                st["steady_wall_s"] = wall
    b.close()
    rows = []
    for st, s in zip(state, sols):
        m, e = st["drag"].cd() if st["drag"] else (float("nan"), float("nan"))
        rows.append({"steady_step": st["steady_step"], "steady_wall_s": st.get("steady_wall_s"),
                     "sampling_from": st["sampling_from"], "final_err": st["err"], "cd": m, "cd_sigma": e,
                     "n_history": st["n_hist"]})
        s.close()
    done = all(r["steady_step"] is not None for r in rows)
    # the ensemble is steady when its last member is
    return {"rows": rows, "load_s": load_s, "steps": step, "all_steady": done,
            "wall_to_all_steady_s": load_s + wall if done else None, "stepping_wall_s": wall}


def evaluate(out: Path, n_cases: int, seed: int, N: int, ref_transient: int, ref_sampling: int, chunk: int,
             max_steps: int, tol: float, block: int, settle: float) -> dict:
    cases = ens.draw_cases(n_cases, seed, N_range=(N, N), seed0=seed * 1000)
    op = neural.FieldOperator.load(out / "operator.npz")
    t0 = time.perf_counter()
    warm = predicted_fields(op, cases)
    import torch

    torch.cuda.synchronize()
    infer_s = time.perf_counter() - t0
    fallback = sum(1 for w in warm if w is None)
    refs = reference_runs(cases, ref_transient, ref_sampling)
    res_c = time_to_steady(cases, refs, None, chunk, max_steps, tol, block, settle, chunk)
    print(f"[eval] cold arm: {res_c['steps']} steps, all steady: {res_c['all_steady']}", flush=True)
    res_w = time_to_steady(cases, refs, warm, chunk, max_steps, tol, block, settle, chunk)
    print(f"[eval] warm arm: {res_w['steps']} steps, all steady: {res_w['all_steady']}", flush=True)
    rows = []
    for c, rc, rw in zip(cases, res_c["rows"], res_w["rows"]):
        rows.append({"case": asdict(c), "cold": rc, "warm": rw})
    sc = [r["cold"]["steady_step"] for r in rows]
    sw = [r["warm"]["steady_step"] for r in rows]
    both = [(a, b) for a, b in zip(sc, sw) if a is not None and b is not None]
    wc = res_c["wall_to_all_steady_s"]
    ww = None if res_w["wall_to_all_steady_s"] is None else res_w["wall_to_all_steady_s"] + infer_s
    summary = {"cases": n_cases, "N": N, "tol": tol, "block": block, "chunk": chunk, "max_steps": max_steps,
               "fallback_to_cold": fallback, "infer_s": infer_s,
               "cold_reached": sum(a is not None for a in sc), "warm_reached": sum(b is not None for b in sw),
               "median_steps_cold": float(np.median([a for a, _ in both])) if both else None,
               "median_steps_warm": float(np.median([b for _, b in both])) if both else None,
               "steps_ratio_median": float(np.median([a / b for a, b in both])) if both else None,
               "steps_all_cold": max(sc) if all(a is not None for a in sc) else None,
               "steps_all_warm": max(sw) if all(b is not None for b in sw) else None,
               "wall_to_all_steady_cold_s": wc, "wall_to_all_steady_warm_s": ww,
               "wall_ratio": (wc / ww) if (wc and ww) else None,
               "load_cold_s": res_c["load_s"], "load_warm_s": res_w["load_s"]}
    res = {"summary": summary, "rows": rows}
    (out / "eval.json").write_text(json.dumps(res, indent=1, default=float))
    print(json.dumps(summary), flush=True)
    return res


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("stage", choices=["data", "train", "eval", "all"])
    ap.add_argument("--out", default="gpurun_out/warmstart")
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
    ap.add_argument("--chunk", type=int, default=50)
    ap.add_argument("--max-steps", type=int, default=4000)
    ap.add_argument("--tol", type=float, default=0.02)
    ap.add_argument("--block", type=int, default=4)
    ap.add_argument("--settle", type=float, default=0.005)
    a = ap.parse_args(argv)
    out = Path(a.out)
    out.mkdir(parents=True, exist_ok=True)
    if a.stage in ("data", "all"):
        generate(out, a.runs, a.batch, a.transient, a.sampling, a.seed, (a.nmin, a.nmax))
    if a.stage in ("train", "all"):
        train(out, a.iters, a.train_batch, a.lr, a.width, a.seed)
    if a.stage in ("eval", "all"):
        evaluate(out, a.cases, a.eval_seed, a.eval_n, a.ref_transient, a.ref_sampling, a.chunk, a.max_steps, a.tol,
                 a.block, a.settle)


if __name__ == "__main__":
    main()
