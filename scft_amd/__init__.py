"""scft_amd -- host-side mirror of libscft_hip.so, the MI355X (gfx950) SCFT field
evaluation.

The compute lives in the C ABI (include/scft.h); this module only loads the
in-tree library, marshals numpy arrays through ctypes and mirrors the
reference's operator interface for the path:

* ``simple_FEM_1D_transient(w, out)`` -- the 1D_FEM.c:47-48 evaluator
  (NR 1-based arrays, out = phi - phi_t);
* ``Batch`` -- many independent instances evaluated / solved in one launch per
  size bucket;
* ``MLP`` -- the FP32 warm-start network with hand-written kernels.

There is no fallback: if the shared library is missing the import raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscft_hip.so")

SCFT_OK = 0
BC_DD, BC_DN = 0, 1
C_ROWSCALED, C_CONSISTENT = 0, 1
STEP_IE, STEP_IRK4 = 0, 1
ST_CONVERGED, ST_NAN, ST_PIVOT, ST_MAXITER, ST_FALLBACK, ST_FAILED = 1, 2, 4, 8, 16, 32
ST_PLAIN_FORM = 64  # left the ground-state form of the preconditioner
ST_RATIONAL = 128   # reached the rational form's threshold (gs_shift_below)
MIX_PLAIN, MIX_RELATIVE, MIX_LOG = 0, 1, 2
L_REF = 3.72374357332160
TAU = 0.530252230020752

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class ScftError(RuntimeError):
    pass


class Config(C.Structure):
    _fields_ = [("device", C.c_int), ("N", C.c_int), ("L", C.c_double), ("tau", C.c_double),
                ("bc", C.c_int), ("cmat", C.c_int), ("beta", C.c_double), ("verbose", C.c_int),
                ("stepper", C.c_int)]


class BatchDesc(C.Structure):
    _fields_ = [("n_inst", C.c_int), ("M", C.POINTER(C.c_int)), ("N", C.POINTER(C.c_int)),
                ("xoff", C.POINTER(C.c_longlong)), ("x", C.POINTER(C.c_double)),
                ("phi_t", C.POINTER(C.c_double)), ("bc", C.c_int), ("cmat", C.c_int),
                ("tune_file", C.c_char_p), ("stepper", C.c_int)]


class AndersonOpts(C.Structure):
    _fields_ = [("history", C.c_int), ("beta", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_int), ("step_cap", C.c_double),
                ("restart_factor", C.c_double), ("regularization", C.c_double),
                ("transform", C.c_int), ("floor", C.c_double),
                ("fallback_factor", C.c_double), ("check_every", C.c_int),
                ("progress_seconds", C.c_double), ("precond", C.c_double),
                ("backtrack", C.c_double), ("max_halvings", C.c_int),
                ("field_floor", C.c_double), ("gs_precond", C.c_double),
                ("gs_field_weight", C.c_double), ("gs_switch", C.c_int),
                ("gs_switch_growth", C.c_double), ("gs_shift", C.c_double),
                ("gs_shift_below", C.c_double), ("inner_sweeps", C.c_int)]


class BroydenStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("evaluations", C.c_long), ("sweeps", C.c_long),
                ("final_err", C.c_double), ("check", C.c_int)]


class ResHeader(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("iters", C.c_int), ("l", C.c_double),
                ("cp", C.c_double), ("mphi", C.c_double), ("t", C.c_double), ("Z", C.c_double),
                ("f", C.c_double), ("fmax", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("sweeps", C.c_int), ("converged", C.c_int), ("seconds", C.c_double),
                ("eval_seconds", C.c_double), ("node_steps", C.c_double), ("bytes", C.c_double)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built: run `python scft_amd/build.py` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    L.scft_default_config.argtypes = [C.POINTER(Config)]
    L.scft_init.argtypes = [C.POINTER(Config)]
    L.scft_strerror.restype = C.c_char_p
    L.scft_strerror.argtypes = [C.c_int]
    L.simple_FEM_1D_transient.argtypes = [C.c_int, _dp, _dp]
    L.scft_set_target.argtypes = [C.c_int, C.c_void_p]
    L.scft_fixed_point.argtypes = [C.c_int, _dp, _dp]
    L.scft_fixed_point.restype = None
    L.scft_fixed_point_nr.argtypes = [C.c_int, _dp, _dp]
    L.scft_fixed_point_nr.restype = None
    L.scft_adm_chen_f.argtypes = [C.c_int, _dp, _dp]
    L.scft_adm_chen_f.restype = None
    L.scft_batch_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(BatchDesc)]
    L.scft_batch_destroy.argtypes = [C.c_void_p]
    L.scft_batch_layout.argtypes = [C.c_void_p, _lp, C.POINTER(C.c_longlong)]
    L.scft_batch_eval.argtypes = [C.c_void_p, _dp, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]
    L.scft_default_anderson.argtypes = [C.POINTER(AndersonOpts)]
    L.scft_default_anderson.restype = None
    L.scft_batch_solve.argtypes = [C.c_void_p, C.POINTER(AndersonOpts), _dp, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.POINTER(SolveStats)]
    L.scft_batch_upload.argtypes = [C.c_void_p, _dp, C.c_void_p]
    L.scft_batch_solve_resident.argtypes = [C.c_void_p, C.POINTER(AndersonOpts),
                                            C.POINTER(SolveStats)]
    L.scft_batch_download.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(SolveStats)]
    L.scft_batch_time_eval.argtypes = [C.c_void_p, _dp, C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
    L.scft_batch_time_buckets.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int,
                                          C.POINTER(C.c_int), _dp, _dp, _ip]
    L.scft_batch_autotune.argtypes = [C.c_void_p, C.c_int]
    L.scft_batch_tune_placement.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.scft_batch_tuning.argtypes = [C.c_void_p, C.POINTER(C.c_int), _ip, _ip, _ip, C.c_int]
    L.scft_debug_slabs.argtypes = [C.c_void_p, C.c_int, _dp]
    L.scft_batch_set_walls.argtypes = [C.c_void_p, _dp, _dp]
    _fp = np.ctypeslib.ndpointer(np.float32, flags="C")
    L.scft_mlp_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, _ip, C.c_ulonglong]
    L.scft_mlp_destroy.argtypes = [C.c_void_p]
    L.scft_mlp_param_count.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    L.scft_mlp_get_params.argtypes = [C.c_void_p, _fp]
    L.scft_mlp_set_params.argtypes = [C.c_void_p, _fp]
    L.scft_mlp_set_normalisation.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p]
    L.scft_mlp_forward.argtypes = [C.c_void_p, C.c_int, _fp, _fp]
    L.scft_warm_plan_create.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_int,
                                        _fp, C.c_int, _dp, C.c_int, _ip, C.c_double, _ip]
    L.scft_warm_plan_create_sel.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p,
                                            C.c_int, _ip, C.c_int, _fp, C.c_int, _dp, C.c_int,
                                            _ip, C.c_double, _ip]
    L.scft_warm_plan_predict.argtypes = [C.c_void_p]
    L.scft_warm_plan_destroy.argtypes = [C.c_void_p]
    L.scft_mlp_loss_grad.argtypes = [C.c_void_p, C.c_int, _fp, _fp, C.POINTER(C.c_double),
                                     C.c_void_p]
    L.scft_mlp_adam_step.argtypes = [C.c_void_p, C.c_float]
    L.scft_mlp_train.argtypes = [C.c_void_p, C.c_int, _fp, _fp, C.c_int, C.c_int, C.c_float,
                                 C.c_ulonglong, C.c_void_p]
    L.scft_resample_spline.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    L.scft_read_res.argtypes = [C.c_char_p, C.c_int, _dp, _dp, _dp, _dp, _dp,
                                C.POINTER(C.c_int), C.POINTER(ResHeader)]
    L.scft_write_solution.argtypes = [C.c_char_p, C.c_int, _dp, _dp, C.c_double, C.c_double]
    L.scft_read_solution.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), _dp, _dp,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.scft_free_energy.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                   C.POINTER(C.c_double)]
    L.scft_refine_mesh.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int,
                                   C.POINTER(C.c_int), C.c_void_p, C.c_void_p]
    L.scft_broyden_solve.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_double,
                                     _dp, C.POINTER(BroydenStats)]
    L.scft_broyden_solve_fn.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_double, _dp,
                                        C.POINTER(BroydenStats)]
    L.scft_eval_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 8
    L.scft_noalloc_error.restype = C.c_char_p
    L.scft_resample_spline_batch.argtypes = [C.c_int, _ip, _lp, _dp, _dp, _ip, _lp, _dp, _dp]
    return L


lib = _load()


def check(rc: int, what: str = "") -> None:
    if rc != SCFT_OK:
        raise ScftError(f"{what}: {lib.scft_strerror(rc).decode()} (code {rc})")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------------ drop-in
def init(N=2048, L=L_REF, tau=TAU, bc=BC_DD, cmat=C_ROWSCALED, beta=0.1, device=0,
         verbose=False, stepper=STEP_IE):
    cfg = Config()
    lib.scft_default_config(C.byref(cfg))
    cfg.device, cfg.N, cfg.L, cfg.tau = device, N, L, tau
    cfg.bc, cfg.cmat, cfg.beta, cfg.verbose = bc, cmat, beta, 1 if verbose else 0
    cfg.stepper = stepper
    check(lib.scft_init(C.byref(cfg)), "scft_init")


def finalize():
    check(lib.scft_finalize(), "scft_finalize")


def set_target(phi_t_nodes):
    if phi_t_nodes is None:
        check(lib.scft_set_target(0, None), "scft_set_target")
    else:
        a = _f64(phi_t_nodes)
        set_target._keep = a
        check(lib.scft_set_target(len(a), a.ctypes.data), "scft_set_target")


def simple_FEM_1D_transient(w_nr, out_nr):
    """The reference evaluator signature (1D_FEM.c:47-48): w[1..n] -> out[1..n]."""
    n = len(w_nr) - 1
    return lib.simple_FEM_1D_transient(n, w_nr, out_nr)


def funcerr() -> int:
    return C.c_int.in_dll(lib, "funcerr").value


def local_iteration() -> int:
    return C.c_int.in_dll(lib, "local_iteration").value


# ------------------------------------------------------------------ batches
def anderson_opts(history=10, beta=0.1, tol=1e-6, max_iter=5000, step_cap=0.5,
                  restart_factor=1e30, regularization=1e-14, transform=MIX_PLAIN, floor=1e-8,
                  fallback_factor=0.0, check_every=8, progress_seconds=0.0, precond=0.0,
                  backtrack=0.0, max_halvings=30, field_floor=0.0, gs_precond=0.0,
                  gs_field_weight=0.25, gs_switch=0, gs_switch_growth=0.0, gs_shift=0.0,
                  gs_shift_below=0.0, inner_sweeps=1):
    return AndersonOpts(history, beta, tol, max_iter, step_cap, restart_factor, regularization,
                        transform, floor, fallback_factor, check_every, progress_seconds,
                        precond, backtrack, max_halvings, field_floor, gs_precond,
                        gs_field_weight, gs_switch, gs_switch_growth, gs_shift,
                        gs_shift_below, inner_sweeps)


class Batch:
    """Independent SCFT instances on one device context.

    xs: list of node-coordinate arrays (M_k each); Ns: contour steps; phits:
    list of targets on each instance's unknowns.
    """

    def __init__(self, xs, Ns, phits, bc=BC_DN, cmat=C_ROWSCALED, tune_file=None,
                 stepper=STEP_IE):
        self.bc, self.cmat, self.stepper = bc, cmat, stepper
        self.M = np.array([len(x) for x in xs], dtype=np.int32)
        self.N = np.ascontiguousarray(Ns, dtype=np.int32)
        self.n = self.M - (2 if bc == BC_DD else 1)
        xoff = np.zeros(len(xs), dtype=np.int64)
        xoff[1:] = np.cumsum(self.M[:-1])
        self._x = _f64(np.concatenate(xs))
        self._pt = _f64(np.concatenate(phits))
        if len(Ns) != len(xs) or len(phits) != len(xs):
            raise ValueError(f"{len(xs)} meshes, {len(Ns)} step counts, {len(phits)} targets")
        for k, p in enumerate(phits):
            if len(p) != self.n[k]:
                raise ValueError(f"instance {k}: target has {len(p)} values, the mesh "
                                 f"{self.n[k]} unknowns")
        self._xoff = xoff
        d = BatchDesc(len(xs), self.M.ctypes.data_as(C.POINTER(C.c_int)),
                      self.N.ctypes.data_as(C.POINTER(C.c_int)),
                      xoff.ctypes.data_as(C.POINTER(C.c_longlong)),
                      self._x.ctypes.data_as(C.POINTER(C.c_double)),
                      self._pt.ctypes.data_as(C.POINTER(C.c_double)), bc, cmat,
                      tune_file.encode() if tune_file else None, stepper)
        h = C.c_void_p()
        check(lib.scft_batch_create(C.byref(h), C.byref(d)), "scft_batch_create")
        self.h = h
        self.uoff = np.zeros(len(xs), dtype=np.int64)
        tot = C.c_longlong()
        check(lib.scft_batch_layout(self.h, self.uoff, C.byref(tot)), "scft_batch_layout")
        self.total = tot.value

    def close(self):
        if getattr(self, "h", None):
            lib.scft_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def split(self, flat):
        return [flat[o:o + n] for o, n in zip(self.uoff, self.n)]

    def join(self, parts):
        """Concatenated unknowns of all instances (a flat array passes through).
        Every part must hold exactly its instance's unknowns: the library
        copies `total` doubles from the result."""
        if isinstance(parts, np.ndarray) and parts.ndim == 1:
            if parts.size != self.total:
                raise ValueError(f"flat field has {parts.size} values, the batch "
                                 f"{self.total} unknowns")
            return _f64(parts)
        if len(parts) != len(self.n):
            raise ValueError(f"{len(parts)} parts for {len(self.n)} instances")
        for k, p in enumerate(parts):
            if len(p) != self.n[k]:
                raise ValueError(f"instance {k}: {len(p)} values for {self.n[k]} unknowns")
        return _f64(np.concatenate(parts))

    def eval(self, w_parts):
        w = self.join(w_parts)
        phi = np.empty(self.total)
        r = np.empty(self.total)
        rn = np.empty(len(self.n))
        st = np.zeros(len(self.n), dtype=np.uint32)
        check(lib.scft_batch_eval(self.h, w, phi.ctypes.data, r.ctypes.data, rn.ctypes.data,
                                  st.ctypes.data), "scft_batch_eval")
        return dict(phi=self.split(phi), r=self.split(r), rnorm=rn, status=st)

    def solve(self, w_parts, opts=None, w_cold_parts=None):
        opts = opts or anderson_opts()
        w = self.join(w_parts).copy()
        wc = self.join(w_cold_parts) if w_cold_parts is not None else None
        it = np.zeros(len(self.n), dtype=np.int32)
        st = np.zeros(len(self.n), dtype=np.uint32)
        stats = SolveStats()
        rc = lib.scft_batch_solve(self.h, C.byref(opts), w, wc.ctypes.data if wc is not None
                                  else None, it.ctypes.data, st.ctypes.data, C.byref(stats))
        if rc not in (0, 7):
            check(rc, "scft_batch_solve")
        return dict(w=self.split(w), iters=it, status=st, sweeps=stats.sweeps,
                    converged=stats.converged, seconds=stats.seconds,
                    node_steps=stats.node_steps, bytes=stats.bytes, rc=rc)

    def upload(self, w_parts, w_cold_parts=None):
        w = self.join(w_parts)
        wc = self.join(w_cold_parts) if w_cold_parts is not None else None
        self._up = (w, wc)
        check(lib.scft_batch_upload(self.h, w, wc.ctypes.data if wc is not None else None),
              "scft_batch_upload")

    def solve_resident(self, opts):
        st = SolveStats()
        check(lib.scft_batch_solve_resident(self.h, C.byref(opts), C.byref(st)),
              "scft_batch_solve_resident")
        return dict(sweeps=st.sweeps, seconds=st.seconds)

    def download(self):
        w = np.empty(self.total)
        it = np.zeros(len(self.n), dtype=np.int32)
        stv = np.zeros(len(self.n), dtype=np.uint32)
        st = SolveStats()
        rc = lib.scft_batch_download(self.h, w.ctypes.data, it.ctypes.data, stv.ctypes.data,
                                     C.byref(st))
        if rc not in (0, 7):
            check(rc, "scft_batch_download")
        return dict(w=self.split(w), iters=it, status=stv, converged=st.converged,
                    node_steps=st.node_steps, bytes=st.bytes, sweeps=st.sweeps)

    def time_eval(self, w_parts, reps=5):
        w = self.join(w_parts)
        ms = C.c_double()
        mm = C.c_double()
        check(lib.scft_batch_time_eval(self.h, w, reps, C.byref(ms), C.byref(mm)),
              "scft_batch_time_eval")
        return ms.value

    def time_buckets(self, w_parts, reps=3):
        """[(n_items, ms per launch, algorithmic bytes per launch)] per bucket."""
        w = self.join(w_parts)
        nb = C.c_int()
        ms, by = np.zeros(16), np.zeros(16)
        ni = np.zeros(16, dtype=np.int32)
        check(lib.scft_batch_time_buckets(self.h, w, reps, 16, C.byref(nb), ms, by, ni),
              "scft_batch_time_buckets")
        return [(int(ni[i]), float(ms[i]), float(by[i])) for i in range(nb.value)]

    def autotune(self, reps=2):
        check(lib.scft_batch_autotune(self.h, reps), "scft_batch_autotune")

    def tune_placement(self, tries=5, reps=2):
        """Keep the fastest of 1 + tries allocations of the slab scratch
        (scft_batch_tune_placement); autotune() ends with it."""
        check(lib.scft_batch_tune_placement(self.h, tries, reps), "scft_batch_tune_placement")

    def tuning(self):
        nb = C.c_int()
        nm = np.zeros(16, dtype=np.int32)
        K = np.zeros(16, dtype=np.int32)
        W = np.zeros(16, dtype=np.int32)
        check(lib.scft_batch_tuning(self.h, C.byref(nb), nm, K, W, 16), "scft_batch_tuning")
        return [(int(nm[i]), int(K[i]), int(W[i])) for i in range(nb.value)]

    def set_walls(self, wall0, wallL=None):
        """Per-instance field values at the walls for the consistent w-mass
        (default 0)."""
        for v in (wall0, wallL):
            if v is not None and np.ndim(v) > 0 and np.size(v) != len(self.M):
                raise ValueError(f"{np.size(v)} wall values for {len(self.M)} instances")
        w0 = _f64(np.broadcast_to(np.asarray(wall0, dtype=np.float64), (len(self.M),)))
        wl = _f64(np.broadcast_to(np.asarray(0.0 if wallL is None else wallL, dtype=np.float64),
                                  (len(self.M),)))
        check(lib.scft_batch_set_walls(self.h, w0, wl), "scft_batch_set_walls")

    def debug_slabs(self, inst=0):
        """Stored slabs q_1..q_{N/2-1} of a single-instance batch (node order)."""
        n, N = int(self.n[inst]), int(self.N[inst])
        out = np.empty((N // 2 - 1) * n)
        check(lib.scft_debug_slabs(self.h, inst, out), "scft_debug_slabs")
        return out.reshape(N // 2 - 1, n)


# ---------------------------------------------------------------- warm plan
class WarmPlan:
    """The warm start's prediction written straight into a batch's resident
    field (scft_warm_plan_*): the network's forward on inputs kept in HBM,
    then one kernel mapping each instance's outputs onto its unknowns.

    inputs: one row of raw features per instance; grid_x: one row of
    n_grid points per instance; node_idx: the 1-based node numbers of the
    wall samples, node_scale their asinh scale; wall_count: per instance,
    the leading unknowns taken from the samples.  sel (optional): the
    batch's instances the rows belong to (increasing), when the plan covers
    a subset; without it, every instance in order."""

    def __init__(self, batch, mlp, inputs, grid_x, node_idx, node_scale, wall_count, sel=None):
        inputs = _f32(inputs)
        grid_x = _f64(grid_x)
        node_idx = np.ascontiguousarray(node_idx, dtype=np.int32)
        wall_count = np.ascontiguousarray(wall_count, dtype=np.int32)
        rows = len(batch.n) if sel is None else len(sel)
        if inputs.ndim != 2 or inputs.shape[0] != rows or grid_x.ndim != 2 or \
                grid_x.shape[0] != rows or len(wall_count) != rows:
            raise ValueError("one row of inputs, grid points and wall count per instance")
        self._keep = (batch, mlp)
        h = C.c_void_p()
        if sel is None:
            check(lib.scft_warm_plan_create(C.byref(h), batch.h, mlp.h, inputs.shape[1], inputs,
                                            grid_x.shape[1], grid_x, len(node_idx), node_idx,
                                            float(node_scale), wall_count),
                  "scft_warm_plan_create")
        else:
            sel = np.ascontiguousarray(sel, dtype=np.int32)
            check(lib.scft_warm_plan_create_sel(C.byref(h), batch.h, mlp.h, len(sel), sel,
                                                inputs.shape[1], inputs, grid_x.shape[1], grid_x,
                                                len(node_idx), node_idx, float(node_scale),
                                                wall_count), "scft_warm_plan_create_sel")
        self.h = h

    def predict(self):
        check(lib.scft_warm_plan_predict(self.h), "scft_warm_plan_predict")

    def close(self):
        if getattr(self, "h", None):
            lib.scft_warm_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- resampling
def resample_spline(x_src, y_src, x_dst):
    """Not-a-knot cubic spline through (x_src, y_src) evaluated at x_dst
    (DEALII_SCFT/src/spline_chen.c:13-107 semantics)."""
    xs, ys, xd = _f64(x_src), _f64(y_src), _f64(x_dst)
    out = np.empty(len(xd))
    check(lib.scft_resample_spline(len(xs), xs, ys, len(xd), xd, out), "scft_resample_spline")
    return out


def _tables(arrs):
    """(concatenated values, lengths, offsets) of a list of 1-D arrays or of
    the rows of a 2-D array."""
    if isinstance(arrs, np.ndarray) and arrs.ndim == 2:
        n = np.full(arrs.shape[0], arrs.shape[1], dtype=np.int32)
        return _f64(arrs.reshape(-1)), n, np.arange(arrs.shape[0], dtype=np.int64) * arrs.shape[1]
    n = np.array([len(v) for v in arrs], dtype=np.int32)
    off = np.zeros(len(n), dtype=np.int64)
    off[1:] = np.cumsum(n[:-1])
    return _f64(np.concatenate(arrs)), n, off


def resample_spline_batch(xs_src, ys_src, xs_dst, flat=False):
    """resample_spline for many tables in one library call.  Each argument is
    a list of 1-D arrays or a 2-D array (one table per row).  Returns the
    concatenated result (n_total,) when the destinations form a 2-D array or
    flat=True (with the offsets for flat=True), else a list of arrays."""
    xs, ns, so = _tables(xs_src)
    ys, _, _ = _tables(ys_src)
    xd, nd, do = _tables(xs_dst)
    out = np.empty(int(nd.sum()))
    check(lib.scft_resample_spline_batch(len(ns), ns, so, xs, ys, nd, do, xd, out),
          "scft_resample_spline_batch")
    if flat:
        return out, do
    if isinstance(xs_dst, np.ndarray) and xs_dst.ndim == 2:
        return out
    return [out[o:o + n] for o, n in zip(do, nd)]


# ---------------------------------------------------------------- MLP
def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Mlp:
    """Dense FP32 network on the device (tanh hidden layers, linear output),
    trained with the library's own forward / backward / Adam kernels."""

    def __init__(self, dims, seed=1):
        self.dims = [int(d) for d in dims]
        h = C.c_void_p()
        d = np.ascontiguousarray(self.dims, dtype=np.int32)
        check(lib.scft_mlp_create(C.byref(h), len(self.dims) - 1, d, seed), "scft_mlp_create")
        self.h = h
        n = C.c_longlong()
        check(lib.scft_mlp_param_count(self.h, C.byref(n)), "scft_mlp_param_count")
        self.n_params = n.value

    def close(self):
        if self.h:
            lib.scft_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        check(lib.scft_mlp_get_params(self.h, p), "scft_mlp_get_params")
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.n_params
        check(lib.scft_mlp_set_params(self.h, p), "scft_mlp_set_params")

    def set_normalisation(self, x_mean=None, x_std=None, y_mean=None, y_std=None):
        if x_mean is None:
            check(lib.scft_mlp_set_normalisation(self.h, None, None, None, None), "norm")
            return
        self._norm = [_f32(v) for v in (x_mean, x_std, y_mean, y_std)]
        check(lib.scft_mlp_set_normalisation(self.h, *[v.ctypes.data for v in self._norm]),
              "scft_mlp_set_normalisation")

    def forward(self, X):
        X = _f32(X)
        Y = np.empty((X.shape[0], self.dims[-1]), dtype=np.float32)
        check(lib.scft_mlp_forward(self.h, X.shape[0], X, Y), "scft_mlp_forward")
        return Y

    def loss_grad(self, X, T, want_grad=True):
        X, T = _f32(X), _f32(T)
        loss = C.c_double()
        g = np.empty(self.n_params, dtype=np.float32) if want_grad else None
        check(lib.scft_mlp_loss_grad(self.h, X.shape[0], X, T, C.byref(loss),
                                     g.ctypes.data if g is not None else None),
              "scft_mlp_loss_grad")
        return loss.value, g

    def adam_step(self, lr=1e-3):
        check(lib.scft_mlp_adam_step(self.h, lr), "scft_mlp_adam_step")

    def train(self, X, Y, epochs=50, batch=256, lr=1e-3, seed=1):
        X, Y = _f32(X), _f32(Y)
        hist = np.zeros(epochs)
        check(lib.scft_mlp_train(self.h, X.shape[0], X, Y, epochs, batch, lr, seed,
                                 hist.ctypes.data), "scft_mlp_train")
        return hist


# ---------------------------------------------------------------- Broyden
def broyden_solve(x, phi_t_unknowns, N, w_unknowns, tolf=1e-8, bc=BC_DN, cmat=C_ROWSCALED):
    """Broyden (broydn.c semantics) on one instance with the batched
    finite-difference Jacobian; returns the field on the unknowns and the
    statistics (rc 0 = converged)."""
    x, pt, w = _f64(x), _f64(phi_t_unknowns), _f64(w_unknowns).copy()
    st = BroydenStats()
    rc = lib.scft_broyden_solve(len(x), int(N), x, pt, bc, cmat, tolf, w, C.byref(st))
    return dict(rc=rc, w=w, iterations=st.iterations, evaluations=st.evaluations,
                sweeps=st.sweeps, final_err=st.final_err, check=st.check)


BROYDEN_FN = C.CFUNCTYPE(C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                         C.c_void_p)


def broyden_solve_fn(fn, w, tolf=1e-8, ctx=None):
    """The same Broyden iteration on a residual the caller computes:
    fn is a C function pointer (or a BROYDEN_FN) F(n, w, f, ctx) -> nonzero
    on failure; the Jacobian columns are n further calls (host only)."""
    w = _f64(w).copy()
    st = BroydenStats()
    rc = lib.scft_broyden_solve_fn(len(w), C.cast(fn, C.c_void_p), ctx, tolf, w, C.byref(st))
    return dict(rc=rc, w=w, iterations=st.iterations, evaluations=st.evaluations,
                sweeps=st.sweeps, final_err=st.final_err, check=st.check)


# ---------------------------------------------------------------- no allocation
def eval_noalloc(M, N, d_x, d_w, d_phi_t, d_phi, d_r, d_rnorm, d_status, bc=BC_DN,
                 cmat=C_ROWSCALED, stream=None):
    """scft_eval_noalloc on caller-owned device buffers (integer addresses);
    returns the status code (SCFT_E_ARG with scft_noalloc_error() for an
    instance beyond the fixed scratch)."""
    return lib.scft_eval_noalloc(int(M), int(N), bc, cmat, d_x, d_w, d_phi_t, d_phi, d_r,
                                 d_rnorm, d_status, stream)


def noalloc_error() -> str:
    return (lib.scft_noalloc_error() or b"").decode()


# ---------------------------------------------------------------- file formats
def read_res(path, max_rows=1 << 16):
    """Columns and header of a .res file (1D_FEM.c:321-342 format)."""
    cols = [np.empty(max_rows) for _ in range(5)]
    rows, h = C.c_int(), ResHeader()
    check(lib.scft_read_res(str(path).encode(), max_rows, *cols, C.byref(rows), C.byref(h)),
          "scft_read_res")
    r = rows.value
    out = {k: c[:r].copy() for k, c in zip(("x_over_l", "phi", "eta", "phie", "phij"), cols)}
    out["header"] = {k: getattr(h, k) for k, _ in ResHeader._fields_}
    return out


def write_solution(path, x, w, err, free_energy):
    x, w = _f64(x), _f64(w)
    check(lib.scft_write_solution(str(path).encode(), len(x), x, w, err, free_energy),
          "scft_write_solution")


def read_solution(path, cap=1 << 20):
    x, w = np.empty(cap), np.empty(cap)
    M, e, fe = C.c_int(), C.c_double(), C.c_double()
    check(lib.scft_read_solution(str(path).encode(), cap, C.byref(M), x, w, C.byref(e),
                                 C.byref(fe)), "scft_read_solution")
    return dict(x=x[:M.value].copy(), w=w[:M.value].copy(), err=e.value, free_energy=fe.value)


def free_energy(x, w_nodes, L=None, tau=TAU, mirror=True, log2_points=18):
    """Mean-field free energy of DEALII_SCFT/src/scft.cc:404-450."""
    x, w = _f64(x), _f64(w_nodes)
    F = C.c_double()
    check(lib.scft_free_energy(len(x), x, w, float(L if L is not None else x[-1]), tau,
                               1 if mirror else 0, log2_points, C.byref(F)), "scft_free_energy")
    return F.value


def refine_mesh(x, w_nodes, factor=10.0, bc=BC_DN):
    """One gradient-rule refinement with spline transfer of the field."""
    x, w = _f64(x), _f64(w_nodes)
    Mn = C.c_int()
    check(lib.scft_refine_mesh(len(x), x, w, factor, bc, 0, C.byref(Mn), None, None),
          "scft_refine_mesh")
    xn, wn = np.empty(Mn.value), np.empty(Mn.value)
    check(lib.scft_refine_mesh(len(x), x, w, factor, bc, Mn.value, C.byref(Mn),
                               xn.ctypes.data, wn.ctypes.data), "scft_refine_mesh")
    return xn, wn
