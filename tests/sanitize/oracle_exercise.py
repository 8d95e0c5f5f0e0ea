"""Exercises the CPU oracle's C functions (oracle/*.c) in a process where
SCFT_ORACLE_LIB points at the AddressSanitizer + UndefinedBehaviorSanitizer
build and libasan is preloaded (tests/test_sanitizers.py starts it).  Any
sanitizer report aborts the process; the checks here are light, the point
is to run every code path on small inputs."""
import ctypes as C
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
import oracle as O  # noqa: E402

assert O.LIB_PATH == os.environ["SCFT_ORACLE_LIB"], O.LIB_PATH

for bc in (O.BC_DD, O.BC_DN):
    for cmat in (O.C_ROWSCALED, O.C_CONSISTENT):
        x = O.mesh_uniform(41, 3.0)
        pt = O.target_position(x, O.TAU, bc == O.BC_DD)
        w = 0.2 * np.cos(x)
        for stepper in (O.STEP_IE, O.STEP_IRK4):
            r = O.evaluate(x, w, pt, 64, bc=bc, cmat=cmat, want_q=True, stepper=stepper)
            assert r["rc"] == 0 and np.isfinite(r["rnorm"])
        O.march_ld(x, w, 32, bc=bc, cmat=cmat)
xg = O.mesh_graded(O.graded_elements(2.0, 1e-3, 0.1, 63, 4095) + 1, 2.0, 0.1)
r = O.evaluate(xg, 0.1 * np.exp(-xg), O.target_position(xg, O.TAU, False), 256, bc=O.BC_DN)
assert r["rc"] == 0
O.eval_matlab(O.mesh_uniform(33, O.L_SHORT), np.zeros(33), O.target_count(O.mesh_uniform(33, O.L_SHORT)),
              steps=129, dt=1.0 / 128)
O.sine_mode_exact(33, 3.0, 0.5, 64, 3)
assert abs(O.romint(np.ones(65), 64, 1.0 / 64) - 1.0) < 1e-12
O.romberg_weights(256)
rc, inv, xs = O.gaussj(np.array([[4.0, 1.0], [2.0, 3.0]]), np.array([[1.0], [2.0]]), mode=1)
assert rc == 0

# mixers: plain and preconditioned log Anderson with the field bound, adm, broydn
x = O.mesh_uniform(65, 3.72)
pt = O.target_position(x, O.TAU, False)
o = O.anderson_opts(history=16, beta=1.0, tol=1e-6, max_iter=60, step_cap=0.0,
                    transform=O.MIX_LOG, precond=0.5, backtrack=1e30, field_floor=0.5)
res = O.solve_anderson(x, pt, 256, np.zeros(65), o, bc=O.BC_DN, trace_len=8)
assert res["rc"] == 0, res
out = O.batch_solve([65, 33], [256, 128], [x, O.mesh_uniform(33, 3.0)],
                    [pt, O.target_position(O.mesh_uniform(33, 3.0), O.TAU, False)],
                    [np.zeros(65), np.zeros(33)], o, bc=O.BC_DN, threads=2)
assert (out["status"] == 0).all()


def toy(n, xv, xn, ctx):
    xn[0] = xv[0] * 0.5 - 2.0
    xn[1] = xv[1] * 0.5 + 3.0


cb = O.FIXED_POINT_FN(toy)
xv, xn = np.array([1.0, 2.0]), np.zeros(2)
its, err = C.c_int(), C.c_double()
assert O.lib().orc_adm(xv, 2, C.cast(cb, C.c_void_p), None, 1e-10, 100, xn, C.byref(its),
                       C.byref(err)) == 0

m = 16
x = O.mesh_uniform(m + 1, O.L_REF)
pt = O.target_position(x, O.TAU, True)
ctx = O.EvalCtx(x, pt, 128, bc=O.BC_DD)
O.global_callbacks(ctx, m - 1)
xm = np.zeros(m - 1)
jc = C.c_int(0)
n = m - 1
st = O.BroydnStats()
chk = O.lib().orc_broydn(xm, n, C.cast(O.lib().orc_global_residual_ctx, C.c_void_p), None, 1e-8,
                         C.byref(jc), np.zeros(n * n), np.zeros(n * n), np.zeros(n), C.byref(st))
ctx.close()

# MLP: forward, loss and gradient, Adam
dims = np.array([5, 7, 3], dtype=np.int32)
npar = O.mlp_param_count(dims)
rng = np.random.default_rng(1)
p = rng.standard_normal(npar) * 0.3
X, T = rng.standard_normal((4, 5)), rng.standard_normal((4, 3))
O.mlp_forward(dims, p, X)
loss, g = O.mlp_loss_grad(dims, p, X, T)
O.adam_step(p, g, np.zeros(npar), np.zeros(npar), 1)
print("oracle_exercise: done", flush=True)
