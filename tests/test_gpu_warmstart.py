"""What the trained field operator gains over the cold start (GPU): the
whole pipeline in small, from ensemble data generation through training on
the GPU kernels to loading held-out cases from the prediction."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_trained_operator_beats_the_freestream_start(tmp_path):
    from cudadsmc_amd import Batch
    from cudadsmc_amd import ensemble as ens
    from cudadsmc_amd import neural
    from cudadsmc_amd import warmstart as ws

    print("warm start: data", flush=True)
    ws.generate(tmp_path, runs=48, batch=48, transient=500, sampling=1000, seed=5, N_range=(1.0e5, 1.0e5))
    print("warm start: training", flush=True)
    info = ws.train(tmp_path, iters=400, batch=8, lr=2e-3, width=32, seed=0)
    assert np.isfinite(info["loss"][-1][1])
    op = neural.FieldOperator.load(tmp_path / "operator.npz")
    cases = ens.draw_cases(4, 77, N_range=(2.0e5, 2.0e5), seed0=77000)
    print("warm start: references", flush=True)
    refs = ws.reference_runs(cases, 500, 1500, batch=4)
    warm = ws.predicted_fields(op, cases)
    assert all(w is not None for w in warm)
    e_pred = [ws.field_error(w, r["fields"], r["frac"], c.U, 4) for c, r, w in zip(cases, refs, warm)]
    e_free = [ws.field_error(ws.freestream_fields(c), r["fields"], r["frac"], c.U, 4) for c, r in zip(cases, refs)]
    print("loaded-field error, prediction vs freestream:", np.round(e_pred, 3), np.round(e_free, 3), flush=True)
    assert np.mean(e_pred) < 0.5 * np.mean(e_free)
    # the same after 40 steps of DSMC from each start (fields sampled over those steps)
    errs = {}
    for arm, fields in (("cold", None), ("warm", warm)):
        sols = [ens.make_solver(c, fields=None if fields is None else fields[i]) for i, c in enumerate(cases)]
        b = Batch(sols)
        b.set_sampling(True)
        b.run(40)
        b.synchronize()
        errs[arm] = [ws.field_error(s.fields(), r["fields"], r["frac"], c.U, 4) for c, s, r in zip(cases, sols, refs)]
        b.close()
        for s in sols:
            s.close()
    print("error after 40 steps, cold vs warm:", np.round(errs["cold"], 3), np.round(errs["warm"], 3), flush=True)
    assert np.mean(errs["warm"]) < 0.5 * np.mean(errs["cold"])
