"""Mesh continuation (the reference's warm start: refine, transfer the field
by spline, re-solve) against solving the refined meshes from the cold start."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_refined_meshes_start_from_the_transferred_fields():
    import scft_amd as S
    from scft_amd import continuation as CT, workloads as WL
    rng = np.random.default_rng(17)
    ms = rng.integers(48, 200, 32)
    Ls = rng.uniform(2.5, 6.0, 32)
    xs = [WL.uniform_mesh(int(m), float(L)) for m, L in zip(ms, Ls)]
    Ns = [1024] * 32
    w0s = [WL.w0_field(x, 100 + k) for k, x in enumerate(xs)]
    opts = S.anderson_opts(history=10, beta=1.0, tol=1e-6, max_iter=300, step_cap=0.0,
                           transform=S.MIX_LOG, precond=0.5, backtrack=1e30, check_every=4)
    xf, wf, stats = CT.solve_levels(S, xs, Ns, w0s, WL.exp_target, opts, levels=1)
    assert all(s["converged"] == 32 for s in stats)
    assert all(len(a) > len(b) for a, b in zip(xf, xs))  # every mesh refined
    # the same refined meshes from the cold start
    pts = [WL.exp_target(x) for x in xf]
    b = S.Batch(xf, Ns, [p[1:] for p in pts], bc=S.BC_DN)
    cold = b.solve([WL.w0_field(x, 100 + k)[1:] for k, x in enumerate(xf)], opts)
    b.close()
    warm_it, cold_it = stats[1]["iters"].mean(), cold["iters"].mean()
    print(f"refined meshes: transferred start {warm_it:.2f} evaluations, cold {cold_it:.2f}")
    assert cold["converged"] == 32 and warm_it < 0.9 * cold_it  # measured: 10.8 vs 13.3
    for k in range(0, 32, 8):  # the refined solutions solve the refined problems
        r = O.evaluate(xf[k], wf[k], pts[k], 1024, bc=O.BC_DN)["rnorm"]
        assert r <= 1e-6 + 1e-10
