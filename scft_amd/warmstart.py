"""Learned warm start for the batched solver (host side of the MLP path).

A dense network (scft_amd.Mlp: FP32, trained on the device by the library's
own kernels) maps an instance's parameters and cold-start field to its
converged field w(x):

  inputs   L, log N, log m, log h_first, log h_last (domain length, contour
           steps, elements, first and last element: the two ends of a mesh
           graded toward the wall, equal on a uniform one) and the cold-start
           field w0 sampled on the 128-point grid             -> 133 values
  outputs  w on the 128-point grid, and w at the nodes of index NODE_IDX in
           the scale asinh(w / W_SCALE)                     -> 136 values
  grid     96 points uniform on the wall layer 0 <= x <= 1.2 (absolute: the
           target's wall width tau = 0.53 does not scale with L) and 32
           uniform on (1.2, L]; L >= 2 in every workload

Fields move between a mesh and the grid by the not-a-knot cubic spline
(scft_resample_spline_batch, the semantics of
DEALII_SCFT/src/spline_chen.c:13-107) in both directions: a converged field
mapped to the grid and back solves the evaluator to ~1e-5 (against ~1e-2 for
a grid uniform in x/L or for linear sampling).

The first unknowns carry a node-scale feature: the implicit-Euler march
cannot make phi as small as phi_t ~ (2 x / tau)^2 at x_1, so the converged
field there is large and grows as the mesh is refined (w_1 ~ 70 at h = 1.3e-2,
~400 at 7e-3, ~1e3-8e3 at 2.7e-3, depending on N), decaying within x ~ 0.03.
No smooth profile on a fixed grid carries it, so nodes with x < X_SPIKE are
predicted node by node (samples at NODE_IDX, asinh values interpolated
linearly in the node index), and the grid target is the field with those
nodes replaced by a linear extrapolation of the next two.  Coarse meshes
(h >= X_SPIKE) use the grid alone.  The wall node is always replaced by that
extrapolation before sampling: its stored value (0) is a convention, not an
unknown, and a jump there would ring through the spline.  Targets are the
fields the batched solver itself converged.

Instances can be split by element count into ranges, each with a network of
its own (WarmStart's cuts): coarse meshes, whose nodal fields vary with m in
ways the grid carries poorly, then no longer take the capacity the fine
meshes need.

The solver keeps the cold-start field as a fallback: an instance whose first
residual exceeds a factor of the batch median restarts cold.
"""
from __future__ import annotations

import numpy as np

GRID = 128
GRID_WALL = 96      # points on the wall layer
X_WALL = 1.2        # extent of the wall layer
NODE_IDX = np.array([1, 2, 3, 4, 6, 8, 12, 16])  # node samples near the wall
X_SPIKE = 0.04      # nodes closer to the wall than this use the node samples
W_SCALE = 10.0      # node-sample scale: asinh(w / W_SCALE)
CLIP = 50.0         # guard on the smooth part of the targets
N_SCALARS = 5       # L, log N, log m, log h_first, log h_last
DIMS = [N_SCALARS + GRID, 256, 256, 256, 256, GRID + len(NODE_IDX)]

# node k = 1 .. NODE_IDX[-1] from the samples, linear in the index (an
# exponential decay in x is linear in the asinh scale)
_NODE_W = np.array([[np.interp(k, NODE_IDX, e) for e in np.eye(len(NODE_IDX))]
                    for k in range(1, int(NODE_IDX[-1]) + 1)])


def _extend_to_wall(x: np.ndarray, w: np.ndarray, k: int) -> np.ndarray:
    """w with nodes 0..k replaced by the line through nodes k+1, k+2.  Node 0
    (the wall) is not an unknown: its stored value is a convention, and a
    jump there would ring through the spline."""
    w = np.array(w, dtype=np.float64)
    slope = (w[k + 2] - w[k + 1]) / (x[k + 2] - x[k + 1])
    w[:k + 1] = w[k + 1] + slope * (x[:k + 1] - x[k + 1])
    return w


def grid(L) -> np.ndarray:
    """The 128 sample points of a domain of length L (> X_WALL); an array of
    lengths gives one row per length."""
    L = np.asarray(L, dtype=np.float64)
    wall = X_WALL * np.arange(GRID_WALL, dtype=np.float64) / (GRID_WALL - 1)
    t = np.arange(1, GRID - GRID_WALL + 1, dtype=np.float64) / (GRID - GRID_WALL)
    rest = X_WALL + (L[..., None] - X_WALL) * t
    return np.concatenate([np.broadcast_to(wall, rest.shape[:-1] + wall.shape), rest], axis=-1)


def spike_nodes(x: np.ndarray) -> int:
    """Number of leading unknowns (nodes 1..k) handled by the node samples."""
    k = int(np.searchsorted(x, X_SPIKE)) - 1  # nodes 1.. with x < X_SPIKE
    return max(0, min(k, int(NODE_IDX[-1]), len(x) - 4))


def features(S, xs, Ns, w0s) -> np.ndarray:
    """Network inputs (n x 133) of instances with nodes xs, contour steps Ns
    and cold-start fields w0s (all nodes)."""
    Ls = np.array([x[-1] for x in xs], dtype=np.float64)
    F = np.empty((len(xs), N_SCALARS + GRID))
    F[:, 0] = Ls
    F[:, 1] = np.log(np.asarray(Ns, dtype=np.float64))
    F[:, 2] = np.log(np.array([len(x) - 1 for x in xs], dtype=np.float64))
    F[:, 3] = np.log(np.array([x[1] - x[0] for x in xs], dtype=np.float64))
    F[:, 4] = np.log(np.array([x[-1] - x[-2] for x in xs], dtype=np.float64))
    F[:, N_SCALARS:] = S.resample_spline_batch(xs, [_extend_to_wall(x, w, 0) for x, w in zip(xs, w0s)],
                                       grid(Ls)).reshape(len(xs), GRID)
    return F


def targets(S, xs, ws) -> np.ndarray:
    """Network outputs (n x 136) of converged fields ws (all nodes)."""
    smooth = []
    Y = np.empty((len(xs), GRID + len(NODE_IDX)))
    for i, (x, w) in enumerate(zip(xs, ws)):
        k = spike_nodes(x)
        w = np.asarray(w, dtype=np.float64)
        smooth.append(np.clip(_extend_to_wall(x, w, k), -CLIP, CLIP))
        Y[i, GRID:] = np.arcsinh(w[np.minimum(NODE_IDX, len(x) - 1)] / W_SCALE)
    Ls = np.array([x[-1] for x in xs], dtype=np.float64)
    Y[:, :GRID] = S.resample_spline_batch(xs, smooth, grid(Ls)).reshape(len(xs), GRID)
    return Y


def to_mesh(S, xs, Y) -> list:
    """Predictions (n x 136) on the instances' nodes (w = 0 at the wall)."""
    Y = np.asarray(Y, dtype=np.float64)
    Ls = np.array([x[-1] for x in xs], dtype=np.float64)
    flat, off = S.resample_spline_batch(grid(Ls), Y[:, :GRID], xs, flat=True)
    spike = W_SCALE * np.sinh(Y[:, GRID:] @ _NODE_W.T)  # n x NODE_IDX[-1]
    flat[off] = 0.0  # wall nodes
    for i, x in enumerate(xs):
        k = spike_nodes(x)
        if k > 0:
            flat[off[i] + 1:off[i] + 1 + k] = spike[i, :k]
    return [flat[o:o + len(x)] for o, x in zip(off, xs)]


def dataset(S, cfg: dict, w_solved: list, converged: np.ndarray):
    """(X, Y) rows of the converged instances of a solved configuration."""
    keep = [k for k, ok in enumerate(converged) if ok]
    xs = [cfg["xs"][k] for k in keep]
    X = features(S, xs, [cfg["Ns"][k] for k in keep], [cfg["w0"][k] for k in keep])
    Y = targets(S, xs, [w_solved[k] for k in keep])
    return X.astype(np.float32), Y.astype(np.float32)


def normalisation(X: np.ndarray, Y: np.ndarray, node_weight: float = 1.0):
    """Per-feature mean and std for the three scalars and the outputs; the
    sampled cold-start field keeps unit scale (its deterministic part is the
    same profile in every instance, so what varies is the 1e-3 noise, which
    a per-feature std would blow up to unit size).  The node samples next to
    the wall get their scale divided by `node_weight`, which weights their
    errors by node_weight^2 in the training loss: the first residual of a
    prediction on a graded mesh is set mostly there."""
    xm, xs = X.mean(axis=0), X.std(axis=0)
    ym, ys = Y.mean(axis=0), Y.std(axis=0)
    xs[N_SCALARS:] = 1.0
    xs[xs < 1e-6] = 1.0
    ys[ys < 1e-6] = 1.0
    ys[GRID:] /= node_weight
    return xm, xs, ym, ys


class Prepared:
    """What a warm start needs of a configuration that does not change from
    solve to solve: network inputs, the unknowns' coordinates in the solver's
    layout (instance after instance, nodes 1..), each instance's grid, its
    element count, and the instances with node samples (offset in the
    layout, count)."""

    def __init__(self, S, cfg: dict):
        xs = cfg["xs"]
        self.F = features(S, xs, cfg["Ns"], cfg["w0"]).astype(np.float32)
        self.m = np.array([len(x) - 1 for x in xs], dtype=np.int64)
        self.grids = grid(np.array([x[-1] for x in xs], dtype=np.float64))
        self.xu, self.nu, self.uoff = S._tables([x[1:] for x in xs])
        self.gx, self.gn, self.goff = S._tables(self.grids)
        self.wall = np.array([spike_nodes(x) for x in xs], dtype=np.int32)
        self.spikes = [(i, int(self.uoff[i]), int(k)) for i, k in enumerate(self.wall) if k > 0]
        # the node samples as one scatter: out[sp_dst] = spike[sp_inst, sp_node]
        self.sp_inst = np.concatenate([np.full(k, i) for i, _, k in self.spikes] or [[]]).astype(np.int64)
        self.sp_node = np.concatenate([np.arange(k) for _, _, k in self.spikes] or [[]]).astype(np.int64)
        self.sp_dst = np.concatenate([o + np.arange(k) for _, o, k in self.spikes] or [[]]).astype(np.int64)


def elements_of(X: np.ndarray) -> np.ndarray:
    """Element count m of each input row (its third feature is log m)."""
    return np.rint(np.exp(np.asarray(X, dtype=np.float64)[:, 2])).astype(np.int64)


class WarmStart:
    """Trains the network on solved fields and predicts initial fields.

    cuts: element counts splitting the instances into mesh-size ranges
    [0, c1), [c1, c2), ..., each with a network of its own (the same shape,
    trained on the fields of its range alone).  One network learning coarse
    and fine meshes together fits the fine ones worse: on 4096 uniform
    meshes with m log-uniform in 32-4096 a separate network for m >= 256
    lowers the median first residual of the m >= 512 instances from 1.3e-4
    to 3.7e-5."""

    def __init__(self, S, seed: int = 1, cuts=(), range_epochs=None):
        self.S = S
        self.cuts = tuple(sorted(int(c) for c in cuts))
        if any(c < 1 for c in self.cuts) or len(set(self.cuts)) != len(self.cuts):
            raise ValueError(f"cuts {cuts}: distinct positive element counts")
        # epochs of each range's network as a multiple of the schedule's
        self.range_epochs = tuple(int(e) for e in (range_epochs or (1,) * (len(self.cuts) + 1)))
        if len(self.range_epochs) != len(self.cuts) + 1 or min(self.range_epochs) < 1:
            raise ValueError(f"range_epochs {range_epochs}: one positive factor per range")
        self.nets = [S.Mlp(DIMS, seed=seed) for _ in range(len(self.cuts) + 1)]

    @property
    def mlp(self):
        """The network of the first range (the only one without cuts)."""
        return self.nets[0]

    def group(self, m) -> np.ndarray:
        """Index of the mesh-size range of each element count."""
        return np.searchsorted(np.asarray(self.cuts, dtype=np.int64), np.asarray(m), side="right")

    def fit(self, X, Y, epochs=50, batch=256, lr=1e-3, seed=2, schedule=None,
            node_weight=1.0):
        """Adam on the mean squared error of the normalised targets: `epochs`
        at `lr`, or the stages of `schedule` ((epochs, lr), ...) one after
        another with the optimiser state carried over; the wall node samples
        weighted by node_weight^2 (normalisation).  Each range's network
        trains on the rows of its range, every stage's epochs multiplied by
        the range's factor (range_epochs).  Returns the loss of every epoch:
        one array, or with cuts a list of one array per range."""
        g = self.group(elements_of(X))
        stages = schedule if schedule else ((epochs, lr),)
        hist = []
        for k, net in enumerate(self.nets):
            rows = np.flatnonzero(g == k)
            if len(rows) == 0:
                raise ValueError(f"no training fields in mesh-size range {k} (cuts {self.cuts})")
            Xk, Yk = X[rows], Y[rows]
            net.set_normalisation(*normalisation(Xk, Yk, node_weight))
            f = self.range_epochs[k]
            hist.append(np.concatenate([net.train(Xk, Yk, epochs=f * int(e), batch=batch,
                                                  lr=float(r), seed=seed + j)
                                        for j, (e, r) in enumerate(stages)]))
        return hist[0] if len(hist) == 1 else hist

    def _forward(self, F, m) -> np.ndarray:
        g = self.group(m)
        Y = None
        for k, net in enumerate(self.nets):
            rows = np.flatnonzero(g == k)
            if len(rows) == 0:
                continue
            Yk = net.forward(F[rows])
            if Y is None:
                Y = np.empty((len(F), Yk.shape[1]), dtype=Yk.dtype)
            Y[rows] = Yk
        return Y

    def predict(self, cfg: dict) -> list:
        """Initial fields (all nodes) for every instance of a configuration."""
        F = features(self.S, cfg["xs"], cfg["Ns"], cfg["w0"]).astype(np.float32)
        m = np.array([len(x) - 1 for x in cfg["xs"]])
        return to_mesh(self.S, cfg["xs"], self._forward(F, m))

    def predict_unknowns(self, prep: Prepared) -> np.ndarray:
        """Initial fields on the unknowns, concatenated in the solver's
        layout: the networks' forward passes and one batched resampling
        call."""
        S = self.S
        Y = self._forward(prep.F, prep.m).astype(np.float64)
        ys = np.ascontiguousarray(Y[:, :GRID]).reshape(-1)
        out = np.empty(int(prep.nu.sum()))
        S.check(S.lib.scft_resample_spline_batch(len(prep.gn), prep.gn, prep.goff, prep.gx, ys,
                                                 prep.nu, prep.uoff, prep.xu, out),
                "scft_resample_spline_batch")
        spike = W_SCALE * np.sinh(Y[:, GRID:] @ _NODE_W.T)
        out[prep.sp_dst] = spike[prep.sp_inst, prep.sp_node]
        return out

    def device_plan(self, batch, prep: Prepared):
        """A plan that writes the prediction for `prep`'s configuration
        straight into `batch`'s resident field (scft_amd.WarmPlan, one per
        mesh-size range): the same mapping as predict_unknowns, on the
        device.  Create it after training; close it before the batch or the
        networks."""
        return DevicePlan(self, batch, prep)

    def save(self, path: str, **meta):
        """Cuts, and each network's parameters, normalisation and shape, in
        an .npz (no pickled objects); `meta` adds scalar entries (training
        record)."""
        arrs = {"dims": np.asarray(DIMS), "cuts": np.asarray(self.cuts, dtype=np.int64)}
        for k, net in enumerate(self.nets):
            xm, xs, ym, ys = net._norm
            arrs.update({f"params_{k}": net.params(), f"x_mean_{k}": xm, f"x_std_{k}": xs,
                         f"y_mean_{k}": ym, f"y_std_{k}": ys})
        np.savez(path, **arrs, **{k: np.asarray(v) for k, v in meta.items()})

    def load(self, path: str) -> dict:
        """The networks saved by `save` (their cuts must be this object's);
        returns the file's scalar entries."""
        z = np.load(path, allow_pickle=False)
        if list(z["dims"]) != list(DIMS):
            raise ValueError(f"{path}: network shape {list(z['dims'])}, expected {DIMS}")
        ranged = "cuts" in z.files  # a file of one network has no cuts and no suffixes
        cuts = tuple(int(c) for c in z["cuts"]) if ranged else ()
        if cuts != self.cuts:
            raise ValueError(f"{path}: cuts {list(cuts)}, expected {list(self.cuts)}")
        key = (lambda name, k: f"{name}_{k}") if ranged else (lambda name, k: name)
        for k, net in enumerate(self.nets):
            net.set_params(z[key("params", k)])
            net.set_normalisation(z[key("x_mean", k)], z[key("x_std", k)], z[key("y_mean", k)],
                                  z[key("y_std", k)])
        return {k: z[k].item() for k in z.files if z[k].ndim == 0}

    def close(self):
        for net in self.nets:
            net.close()


class DevicePlan:
    """The warm start's prediction on the device for a batch: one
    scft_amd.WarmPlan per mesh-size range that has instances."""

    def __init__(self, ws: WarmStart, batch, prep: Prepared):
        g = ws.group(prep.m)
        self.plans = []
        try:
            for k, net in enumerate(ws.nets):
                rows = np.flatnonzero(g == k)
                if len(rows) == 0:
                    continue
                sel = None if len(rows) == len(g) else rows
                self.plans.append(ws.S.WarmPlan(batch, net, prep.F[rows], prep.grids[rows],
                                                NODE_IDX, W_SCALE, prep.wall[rows], sel=sel))
        except Exception:
            self.close()
            raise

    def predict(self):
        for p in self.plans:
            p.predict()

    def close(self):
        for p in self.plans:
            p.close()
        self.plans = []


def solve(S, cfg: dict, opts, w_init=None, w_cold=None, tune_file=None):
    """Batched solve of a configuration (DN instances); returns the solver
    output with fields on all nodes."""
    b = S.Batch(cfg["xs"], cfg["Ns"], [p[1:] for p in cfg["pts"]], bc=S.BC_DN,
                tune_file=tune_file)
    w = [v[1:] for v in (w_init if w_init is not None else cfg["w0"])]
    wc = [v[1:] for v in w_cold] if w_cold is not None else None
    out = b.solve(w, opts, w_cold_parts=wc)
    out["w_nodes"] = [np.concatenate([[0.0], v]) for v in out["w"]]
    b.close()
    return out
