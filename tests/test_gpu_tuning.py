"""The autotuner and its table: per size bucket, rows per lane K, waves W and
workgroups per CU, written as "capacity K W wpc" and read back by a new batch.
The choice changes only speed, never the arithmetic of an instance: a batch
built from the table reproduces the tuned batch's densities bit for bit, and
both match the untuned (default-variant) batch to the parity bar."""
import os

import numpy as np
import pytest

import oracle as O
from tests.helpers import instance

pytestmark = pytest.mark.gpu


def test_autotune_table_round_trip(tmp_path):
    import scft_amd as S
    rng = np.random.default_rng(31)
    ms = [40, 100, 200, 300, 600, 900, 1500, 2500]
    Ns = [int(rng.choice([256, 512])) for _ in ms]
    insts = [instance(m, 3.0 + 0.5 * i, N, bc=O.BC_DN, seed=500 + i)
             for i, (m, N) in enumerate(zip(ms, Ns))]
    xs = [it["x"] for it in insts]
    pts = [it["pt"][1:] for it in insts]
    ws = [it["w"][1:] for it in insts]
    table = str(tmp_path / "tune.txt")
    base = S.Batch(xs, Ns, pts, bc=S.BC_DN)
    phi0 = base.eval(ws)["phi"]
    base.close()
    b = S.Batch(xs, Ns, pts, bc=S.BC_DN, tune_file=table)
    b.autotune(reps=1)
    tuned = b.tuning()
    phi1 = b.eval(ws)["phi"]
    b.close()
    rows = [l.split() for l in open(table).read().splitlines() if l.strip()]
    assert len(rows) == len(tuned) and all(len(r) == 4 for r in rows), rows
    for r in rows:
        cap, K, W, wpc = map(int, r)
        assert 64 * K * W >= cap and wpc >= 0
    b2 = S.Batch(xs, Ns, pts, bc=S.BC_DN, tune_file=table)
    assert b2.tuning() == tuned
    phi2 = b2.eval(ws)["phi"]
    b2.close()
    for k in range(len(ms)):
        assert np.array_equal(phi1[k], phi2[k]), k
        np.testing.assert_allclose(phi1[k], phi0[k], rtol=0, atol=1e-10)
    # a three-column table (no workgroups-per-CU entry) is still accepted
    with open(table, "w") as f:
        f.writelines(" ".join(r[:3]) + "\n" for r in rows)
    b3 = S.Batch(xs, Ns, pts, bc=S.BC_DN, tune_file=table)
    assert b3.tuning() == tuned
    b3.close()
    os.remove(table)


def test_scratch_placement_choice_keeps_results():
    """scft_batch_tune_placement swaps the slab scratch for the fastest of
    several fresh allocations: it returns normally, keeps the tuning, and the
    batch evaluates and solves to the same bits afterwards (also after a
    second call, and with tries = 0)."""
    import scft_amd as S
    rng = np.random.default_rng(37)
    ms = [50, 150, 400, 700, 1200, 2000]
    Ns = [int(rng.choice([256, 512])) for _ in ms]
    insts = [instance(m, 2.5 + 0.7 * i, N, bc=O.BC_DN, seed=700 + i)
             for i, (m, N) in enumerate(zip(ms, Ns))]
    xs = [it["x"] for it in insts]
    pts = [it["pt"][1:] for it in insts]
    ws = [it["w"][1:] for it in insts]
    opts = S.anderson_opts(history=16, beta=1.0, tol=1e-6, max_iter=80, step_cap=0.0,
                           transform=S.MIX_LOG, precond=0.5, backtrack=1e30, check_every=4)
    b = S.Batch(xs, Ns, pts, bc=S.BC_DN)
    phi0 = b.eval(ws)["phi"]
    sol0 = b.solve(ws, opts)
    tuning = b.tuning()
    for tries in (3, 0, 2):
        b.upload(ws)
        b.tune_placement(tries=tries, reps=1)
        assert b.tuning() == tuning
        phi = b.eval(ws)["phi"]
        sol = b.solve(ws, opts)
        for k in range(len(ms)):
            assert np.array_equal(phi[k], phi0[k]), (tries, k)
            assert np.array_equal(sol["w"][k], sol0["w"][k]), (tries, k)
        assert np.array_equal(sol["iters"], sol0["iters"])
    b.close()
