"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker.  The product library
(scft_amd/) never imports it.

The C sources cite the reference lines they restate; this module only
marshals numpy arrays.  `ref` gives access to the reference's own
Numerical-Recipes routines compiled under oracle/_ref/ (present only where
/root/reference existed at build time).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SCFT_ORACLE_LIB: another build of the same sources (the sanitizer build of
# the CPU test suite loads oracle/_san/liboracle_san.so this way)
LIB_PATH = os.environ.get("SCFT_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REFERENCE_ROOT = "/root/reference"

BC_DD, BC_DN = 0, 1
C_ROWSCALED, C_CONSISTENT = 0, 1
STEP_IE, STEP_IRK4 = 0, 1

# Reference constants: tau and L at full precision (DEALII_SCFT/drivescft.cc:
# 269-271, Exp_m32_n2048_IE.res:3) and as typed in 1D_FEM.c:303 / drive_SCFT.m.
TAU = 0.530252230020752
L_REF = 3.72374357332160
TAU_SHORT = 0.5302
L_SHORT = 3.72374
F0BAR_KAT = 0.892581217773656  # scft.cc:368, testFiBar.cc:50

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir(REFERENCE_ROOT):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


class OrcProblem(C.Structure):
    _fields_ = [("M", C.c_int), ("N", C.c_int), ("bc", C.c_int), ("cmat", C.c_int),
                ("x", C.POINTER(C.c_double)), ("phi_t", C.POINTER(C.c_double))]


class AndersonOpts(C.Structure):
    _fields_ = [("history", C.c_int), ("beta", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_int), ("step_cap", C.c_double),
                ("restart_factor", C.c_double), ("regularization", C.c_double),
                ("transform", C.c_int), ("floor", C.c_double), ("precond", C.c_double),
                ("pstencil", C.c_void_p), ("backtrack", C.c_double), ("max_halvings", C.c_int),
                ("field_floor", C.c_double), ("field_lower", C.c_double),
                ("gs_precond", C.c_double), ("gs_field_weight", C.c_double),
                ("gs_switch", C.c_int), ("gs_switch_growth", C.c_double),
                ("gstencil", C.c_void_p), ("gs_shift", C.c_double), ("gs_shift_below", C.c_double)]


class AndersonStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("restarts", C.c_int), ("final_err", C.c_double),
                ("err_trace", C.POINTER(C.c_double)), ("trace_len", C.c_int)]


class BroydnStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("evaluations", C.c_long), ("final_err", C.c_double),
                ("err_trace", C.POINTER(C.c_double)), ("trace_len", C.c_int)]


FIXED_POINT_FN = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)
RESIDUAL_FN = C.CFUNCTYPE(C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_exp_profile.restype = C.c_double
        L.orc_exp_profile.argtypes = [C.c_double, C.c_double]
        L.orc_target_position.argtypes = [C.c_int, _dp, C.c_double, C.c_int, _dp]
        L.orc_target_count.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_int, _dp]
        L.orc_mesh_uniform.argtypes = [C.c_int, C.c_double, _dp]
        L.orc_graded_elements.restype = C.c_int
        L.orc_graded_elements.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_mesh_graded.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.orc_assemble.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.orc_system.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int, _dp, _dp]
        L.orc_thomas.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_march.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_march_ld.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_march_irk4.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_eval_stepper.argtypes = [C.POINTER(OrcProblem), C.c_int, _dp, C.c_void_p, _dp,
                                       _dp, C.POINTER(C.c_double)]
        L.orc_romint.restype = C.c_double
        L.orc_romint.argtypes = [_dp, C.c_int, C.c_double]
        L.orc_romberg_weights.argtypes = [C.c_int, C.c_double, _dp]
        L.orc_trapezoid.restype = C.c_double
        L.orc_trapezoid.argtypes = [_dp, C.c_int, C.c_double]
        L.orc_phi_romint.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_phi_weights.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_eval.argtypes = [C.POINTER(OrcProblem), _dp, C.c_void_p, _dp, _dp,
                               C.POINTER(C.c_double)]
        L.orc_eval_matlab.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, C.c_double, _dp]
        L.orc_sine_mode_exact.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        L.orc_gaussj.argtypes = [_dp, C.c_int, _dp, C.c_int, C.c_int]
        L.orc_adm.argtypes = [_dp, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_int, _dp,
                              C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_anderson.argtypes = [_dp, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(AndersonOpts),
                                   C.c_void_p, C.POINTER(AndersonStats)]
        L.orc_broydn.argtypes = [_dp, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                 C.POINTER(C.c_int), _dp, _dp, _dp, C.POINTER(BroydnStats)]
        L.orc_solve_anderson.argtypes = [C.POINTER(OrcProblem), _dp, C.POINTER(AndersonOpts),
                                         C.POINTER(AndersonStats)]
        L.orc_batch_solve.argtypes = [C.c_int, _ip, _ip, _lp, _dp, _dp, _dp, C.c_int, C.c_int,
                                      C.POINTER(AndersonOpts), C.c_int, _ip, _ip, _dp]
        L.orc_precond_stencil.argtypes = [_dp, C.c_int, C.c_int, _dp]
        L.orc_gs_stencil.argtypes = [_dp, C.c_int, C.c_int, _dp]
        L.orc_batch_eval.argtypes = [C.c_int, _ip, _ip, _lp, _dp, _dp, _dp, C.c_int, C.c_int,
                                     C.c_int, _dp, _dp, _dp]
        L.orc_global_setup.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_long]
        L.orc_global_calls.restype = C.c_long
        L.orc_ctx_init.argtypes = [C.c_void_p, C.POINTER(OrcProblem), C.c_double]
        L.orc_ctx_free.argtypes = [C.c_void_p]
        L.orc_mlp_param_count.restype = C.c_long
        L.orc_mlp_param_count.argtypes = [C.c_int, _ip]
        L.orc_mlp_forward.argtypes = [C.c_int, _ip, _dp, C.c_int, _dp, _dp]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_mlp_loss_grad.argtypes = [C.c_int, _ip, _dp, C.c_int, _dp, _dp, _dp]
        L.orc_adam_step.argtypes = [C.c_long, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_double]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- targets/mesh
def exp_profile(x, tau=TAU):
    return lib().orc_exp_profile(float(x), float(tau))


def target_position(x, tau=TAU, mirror=True):
    x = _f64(x)
    out = np.empty_like(x)
    lib().orc_target_position(len(x), x, tau, 1 if mirror else 0, out)
    return out


def target_count(x, tau=TAU_SHORT, L=L_SHORT, nan_guard=True):
    x = _f64(x)
    out = np.empty_like(x)
    lib().orc_target_count(len(x), x, tau, L, 1 if nan_guard else 0, out)
    return out


def mesh_uniform(M, L):
    x = np.empty(M)
    lib().orc_mesh_uniform(M, L, x)
    return x


def graded_elements(L, h0, ell, m_min=63, m_max=4095):
    return lib().orc_graded_elements(L, h0, ell, m_min, m_max)


def mesh_graded(M, L, ell):
    x = np.empty(M)
    lib().orc_mesh_graded(M, L, ell, x)
    return x


# ---------------------------------------------------------------- assembly/march
def assemble(x, w, cmat=C_ROWSCALED):
    x, w = _f64(x), _f64(w)
    M = len(x)
    A, B, Cm = np.empty(3 * M), np.empty(3 * M), np.empty(3 * M)
    lib().orc_assemble(M, x, w, cmat, A, B, Cm)
    return A.reshape(M, 3), B.reshape(M, 3), Cm.reshape(M, 3)


def system(x, w, dt, bc=BC_DD, cmat=C_ROWSCALED):
    x, w = _f64(x), _f64(w)
    M = len(x)
    D, R = np.empty(3 * M), np.empty(3 * M)
    lib().orc_system(M, x, w, dt, bc, cmat, D, R)
    return D.reshape(M, 3), R.reshape(M, 3)


def march(x, w, N, bc=BC_DD, cmat=C_ROWSCALED):
    x, w = _f64(x), _f64(w)
    M = len(x)
    q = np.empty((N + 1) * M)
    rc = lib().orc_march(M, N, x, w, bc, cmat, q)
    if rc != 0:
        raise RuntimeError(f"orc_march failed rc={rc}")
    return q.reshape(N + 1, M)


def march_irk4(x, w, N, bc=BC_DD, cmat=C_ROWSCALED):
    """Two-stage Gauss-Legendre march (the stage system as a 2x2-block
    tridiagonal solve); slabs (N+1, M)."""
    x, w = _f64(x), _f64(w)
    M = len(x)
    q = np.empty((N + 1) * M)
    rc = lib().orc_march_irk4(M, N, x, w, bc, cmat, q)
    if rc != 0:
        raise RuntimeError(f"orc_march_irk4 failed rc={rc}")
    return q.reshape(N + 1, M)


def march_ld(x, w, N, bc=BC_DD, cmat=C_ROWSCALED):
    """The march in x87 extended precision (the higher-precision truth)."""
    x, w = _f64(x), _f64(w)
    M = len(x)
    q = np.empty((N + 1) * M)
    rc = lib().orc_march_ld(M, N, x, w, bc, cmat, q)
    if rc != 0:
        raise RuntimeError(f"orc_march_ld failed rc={rc}")
    return q.reshape(N + 1, M)


def romint(f, m, hh):
    return lib().orc_romint(_f64(f), m, hh)


def romberg_weights(m, hh=None):
    hh = 1.0 / m if hh is None else hh
    c = np.empty(m + 1)
    rc = lib().orc_romberg_weights(m, hh, c)
    if rc != 0:
        raise ValueError(f"romberg weights need m = 2^k >= 16 (m={m})")
    return c


def phi_romint(q):
    q = _f64(q)
    N1, M = q.shape
    phi = np.empty(M)
    lib().orc_phi_romint(M, N1 - 1, q, phi)
    return phi


def phi_weights(q, c):
    q = _f64(q)
    N1, M = q.shape
    phi = np.empty(M)
    lib().orc_phi_weights(M, N1 - 1, q, _f64(c), phi)
    return phi


def make_problem(x, phi_t, N, bc=BC_DD, cmat=C_ROWSCALED):
    x, phi_t = _f64(x), _f64(phi_t)
    p = OrcProblem(len(x), N, bc, cmat, x.ctypes.data_as(C.POINTER(C.c_double)),
                   phi_t.ctypes.data_as(C.POINTER(C.c_double)))
    p._keep = (x, phi_t)
    return p


def evaluate(x, w, phi_t, N, bc=BC_DD, cmat=C_ROWSCALED, want_q=False, stepper=STEP_IE):
    """Mode-R evaluation (implicit Euler, or IRK4 with stepper=STEP_IRK4);
    returns dict(phi, r, rnorm[, q])."""
    p = make_problem(x, phi_t, N, bc, cmat)
    M = len(x)
    w = _f64(w)
    phi, r = np.empty(M), np.empty(M)
    nrm = C.c_double()
    q = np.empty((N + 1) * M) if want_q else None
    rc = lib().orc_eval_stepper(C.byref(p), stepper, w, q.ctypes.data if want_q else None, phi, r,
                                C.byref(nrm))
    out = dict(phi=phi, r=r, rnorm=nrm.value, rc=rc)
    if want_q:
        out["q"] = q.reshape(N + 1, M)
    return out


def eval_matlab(x, w, phi_t, steps=2049, dt=1.0 / 2048):
    x, w, phi_t = _f64(x), _f64(w), _f64(phi_t)
    M = len(x)
    out = np.empty(M - 2)
    rc = lib().orc_eval_matlab(M, x, w, phi_t, steps, dt, out)
    if rc != 0:
        raise RuntimeError(f"orc_eval_matlab rc={rc}")
    return out


def sine_mode_exact(M, L, c, N, j):
    q = np.empty(M)
    lib().orc_sine_mode_exact(M, L, c, N, j, q)
    return q


# ---------------------------------------------------------------- solvers
def gaussj(a, b, mode=0):
    a = _f64(a).copy()
    b = _f64(b).copy()
    n = a.shape[0]
    m = b.shape[1] if b.ndim == 2 else 1
    rc = lib().orc_gaussj(a.ravel(), n, b.ravel(), m, mode)
    return rc, a, b


MIX_PLAIN, MIX_RELATIVE, MIX_LOG = 0, 1, 2


def anderson_opts(history=10, beta=0.1, tol=1e-6, max_iter=2000, step_cap=0.5,
                  restart_factor=1e30, regularization=1e-14, transform=MIX_PLAIN, floor=1e-8,
                  precond=0.0, backtrack=0.0, max_halvings=30, field_floor=0.0, field_lower=0.0,
                  gs_precond=0.0, gs_field_weight=0.25, gs_switch=0, gs_switch_growth=0.0,
                  gs_shift=0.0, gs_shift_below=0.0):
    return AndersonOpts(history, beta, tol, max_iter, step_cap, restart_factor, regularization,
                        transform, floor, precond, None, backtrack, max_halvings, field_floor,
                        field_lower, gs_precond, gs_field_weight, gs_switch,
                        gs_switch_growth, None, gs_shift, gs_shift_below)


def solve_anderson(x, phi_t, N, w_full, opts=None, bc=BC_DN, cmat=C_ROWSCALED, trace_len=0):
    p = make_problem(x, phi_t, N, bc, cmat)
    w = _f64(w_full).copy()
    opts = opts or anderson_opts()
    tr = np.zeros(max(trace_len, 1))
    st = AndersonStats(0, 0, 0.0, tr.ctypes.data_as(C.POINTER(C.c_double)), trace_len)
    rc = lib().orc_solve_anderson(C.byref(p), w, C.byref(opts), C.byref(st))
    return dict(rc=rc, w=w, iterations=st.iterations, restarts=st.restarts,
                final_err=st.final_err, trace=tr[:min(trace_len, st.iterations)])


def batch_solve(Ms, Ns, xs, phits, ws, opts, bc=BC_DN, cmat=C_ROWSCALED, threads=0):
    """Solve independent instances over OpenMP threads (the CPU baseline)."""
    Ms = np.ascontiguousarray(Ms, dtype=np.int32)
    Ns = np.ascontiguousarray(Ns, dtype=np.int32)
    off = np.zeros(len(Ms), dtype=np.int64)
    off[1:] = np.cumsum(Ms[:-1])
    x = _f64(np.concatenate(xs))
    pt = _f64(np.concatenate(phits))
    w = _f64(np.concatenate(ws)).copy()
    iters = np.zeros(len(Ms), dtype=np.int32)
    status = np.zeros(len(Ms), dtype=np.int32)
    ferr = np.zeros(len(Ms))
    lib().orc_batch_solve(len(Ms), Ms, Ns, off, x, pt, w, bc, cmat, C.byref(opts), threads,
                          iters, status, ferr)
    ws_out = [w[o:o + m] for o, m in zip(off, Ms)]
    return dict(w=ws_out, iters=iters, status=status, final_err=ferr)


def batch_eval(Ms, Ns, xs, phits, ws, bc=BC_DN, cmat=C_ROWSCALED, threads=0):
    Ms = np.ascontiguousarray(Ms, dtype=np.int32)
    Ns = np.ascontiguousarray(Ns, dtype=np.int32)
    off = np.zeros(len(Ms), dtype=np.int64)
    off[1:] = np.cumsum(Ms[:-1])
    x = _f64(np.concatenate(xs))
    pt = _f64(np.concatenate(phits))
    w = _f64(np.concatenate(ws))
    phi = np.empty_like(w)
    r = np.empty_like(w)
    rn = np.empty(len(Ms))
    rc = lib().orc_batch_eval(len(Ms), Ms, Ns, off, x, pt, w, bc, cmat, threads, phi, r, rn)
    return dict(rc=rc, phi=[phi[o:o + m] for o, m in zip(off, Ms)],
                r=[r[o:o + m] for o, m in zip(off, Ms)], rnorm=rn)


class EvalCtx:
    """An oracle evaluation context (orc_eval_ctx) for a fixed problem."""

    SIZE = 256  # bytes, larger than sizeof(orc_eval_ctx)

    def __init__(self, x, phi_t, N, bc=BC_DD, cmat=C_ROWSCALED, beta=0.1):
        self.p = make_problem(x, phi_t, N, bc, cmat)
        self.buf = C.create_string_buffer(self.SIZE)
        rc = lib().orc_ctx_init(self.buf, C.byref(self.p), beta)
        if rc != 0:
            raise RuntimeError("orc_ctx_init failed")

    @property
    def ptr(self):
        return C.cast(self.buf, C.c_void_p)

    def close(self):
        if self.buf is not None:
            lib().orc_ctx_free(self.buf)
            self.buf = None


def global_callbacks(ctx, n, trace_len=0, limit=0):
    """Install ctx behind the reference-signature callbacks; returns
    (fixed_point_fn_ptr, vecfunc_nr_fn_ptr, trace_array)."""
    trace = np.zeros(max(1, trace_len * n))
    lib().orc_global_setup(ctx.ptr, trace.ctypes.data, len(trace), limit)
    fp = C.cast(lib().orc_global_fixed_point, C.c_void_p)
    vf = C.cast(lib().orc_global_vecfunc_nr, C.c_void_p)
    return fp, vf, trace


# ---------------------------------------------------------------- MLP
def mlp_param_count(dims):
    dims = np.ascontiguousarray(dims, dtype=np.int32)
    return lib().orc_mlp_param_count(len(dims) - 1, dims)


def mlp_forward(dims, params, X):
    dims = np.ascontiguousarray(dims, dtype=np.int32)
    X = _f64(X)
    Y = np.empty((X.shape[0], dims[-1]))
    lib().orc_mlp_forward(len(dims) - 1, dims, _f64(params), X.shape[0], X, Y)
    return Y


def mlp_loss_grad(dims, params, X, T):
    dims = np.ascontiguousarray(dims, dtype=np.int32)
    X, T = _f64(X), _f64(T)
    g = np.empty(len(params))
    loss = lib().orc_mlp_loss_grad(len(dims) - 1, dims, _f64(params), X.shape[0], X, T, g)
    return loss, g


def adam_step(params, grad, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    lib().orc_adam_step(len(params), params, _f64(grad), m, v, t, lr, b1, b2, eps)


# ---------------------------------------------------------------- reference NR
class _Ref:
    """The reference's compiled NR routines (oracle/_ref), when built."""

    def __init__(self):
        self._dealii = self._top = self._chen = None

    def available(self):
        return all(os.path.exists(os.path.join(REF_DIR, f)) for f in
                   ("libnrref_dealii.so", "libnrref_top.so", "libnrref_chen.so"))

    @property
    def dealii(self):
        if self._dealii is None:
            L = C.CDLL(os.path.join(REF_DIR, "libnrref_dealii.so"))
            L.romint.restype = C.c_double
            L.romint.argtypes = [_dp, C.c_int, C.c_double]
            L.nrref_broydn_setup.argtypes = [C.c_int, C.c_double]
            L.nrref_get_err.restype = C.c_double
            L.gaussj.restype = C.c_int
            self._dealii = L
        return self._dealii

    @property
    def top(self):
        if self._top is None:
            L = C.CDLL(os.path.join(REF_DIR, "libnrref_top.so"))
            L.adm.restype = C.POINTER(C.c_double)
            L.adm.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int), C.c_void_p,
                              C.c_int]
            self._top = L
        return self._top

    @property
    def chen(self):
        if self._chen is None:
            L = C.CDLL(os.path.join(REF_DIR, "libnrref_chen.so"))
            L.nrref_adm_chen.restype = C.c_int
            L.nrref_adm_chen.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int, C.c_int,
                                         C.c_double, C.c_int, C.c_int]
            L.nrref_spline_chen.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int]
            self._chen = L
        return self._chen


ref = _Ref()
