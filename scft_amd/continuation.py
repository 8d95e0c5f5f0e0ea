"""Mesh continuation: the reference's own warm-start mechanism.

DEALII_SCFT/drivescft.cc:266-320 solves on a coarse mesh, refines it
(scft.cc:132-169 / Matlab_files/refine_mesh.m:1-36) and restarts the solver
on the finer mesh from the old solution carried over by spline_chen.  Here
every instance of a batch is refined by the gradient rule of refine_mesh.m
(scft_refine_mesh: wall cells and cells whose |dw/dx| reaches `factor`
times the median are bisected, the field moved by the not-a-knot spline)
and the refined batch is solved from the transferred fields.
"""
from __future__ import annotations

import numpy as np


def refine_batch(S, xs, ws_nodes, factor=10.0, bc=None):
    bc = S.BC_DN if bc is None else bc
    out = [S.refine_mesh(x, w, factor=factor, bc=bc) for x, w in zip(xs, ws_nodes)]
    return [o[0] for o in out], [o[1] for o in out]


def solve_levels(S, xs, Ns, w0s, target, opts, levels=1, factor=10.0, bc=None):
    """Solve, then `levels` times refine and re-solve from the transferred
    fields.  target(x) gives phi_t on all nodes.  Returns the final meshes,
    fields (all nodes) and per-level statistics."""
    bc = S.BC_DN if bc is None else bc
    cur_x, cur_w, stats = list(xs), list(w0s), []
    for lev in range(levels + 1):
        pts = [target(x) for x in cur_x]
        b = S.Batch(cur_x, Ns, [p[1:] for p in pts], bc=bc)
        out = b.solve([w[1:] for w in cur_w], opts)
        b.close()
        w_nodes = [np.concatenate([[0.0], v]) for v in out["w"]]
        stats.append(dict(level=lev, nodes=[len(x) for x in cur_x], iters=out["iters"].copy(),
                          converged=int(out["converged"]), seconds=out["seconds"]))
        if lev < levels:
            cur_x, cur_w = refine_batch(S, cur_x, w_nodes, factor=factor, bc=bc)
        else:
            cur_w = w_nodes
    return cur_x, cur_w, stats
