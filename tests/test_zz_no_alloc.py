"""scft_eval_noalloc (the evaluation that allocates nothing) against the
existing entry point it varies, scft_batch_eval on a one-instance batch: two
calls in a row, the second with a smaller instance, then an instance beyond
the fixed scratch, which is rejected.  Device buffers come from torch
(plumbing only: the library call under test allocates nothing)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from tests.helpers import instance, unknowns

pytestmark = pytest.mark.gpu


def _run(S, torch, x, w_unk, pt_unk, N, bc):
    dev = "cuda:0"
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    d_x, d_w, d_pt = t(x), t(w_unk), t(pt_unk)
    n = len(w_unk)
    d_phi = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    d_r = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    d_rn = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
    d_st = torch.full((1,), 12345, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    rc = S.eval_noalloc(len(x), N, d_x.data_ptr(), d_w.data_ptr(), d_pt.data_ptr(),
                        d_phi.data_ptr(), d_r.data_ptr(), d_rn.data_ptr(), d_st.data_ptr(),
                        bc=bc, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    return rc, dict(phi=d_phi.cpu().numpy(), r=d_r.cpu().numpy(), rnorm=float(d_rn.cpu()[0]),
                    status=int(d_st.cpu()[0]))


def _reference(S, x, w_unk, pt_unk, N, bc):
    b = S.Batch([x], [N], [pt_unk], bc=bc)
    out = b.eval([w_unk])
    b.close()
    return out


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_matches_batch_eval(order):
    import torch
    import scft_amd as S
    cases = [(1000, 2048, 3.72, O.BC_DN, 3), (300, 512, 5.0, O.BC_DD, 4)]
    for m, N, L, bc, seed in cases:
        inst = instance(m, L, N, bc=bc, seed=seed)
        x, pt, w = inst["x"], inst["pt"], inst["w"].copy()
        w[0] = 0.0
        if bc == O.BC_DD:
            w[-1] = 0.0
        w_unk, pt_unk = unknowns(w, bc), unknowns(pt, bc)
        print(f"noalloc case m={m} N={N} bc={bc}", flush=True)
        rc, got = _run(S, torch, x, w_unk, pt_unk, N, bc)
        assert rc == 0, S.noalloc_error()
        ref = _reference(S, x, w_unk, pt_unk, N, bc)
        same = np.array_equal(got["phi"], ref["phi"][0]) and np.array_equal(got["r"], ref["r"][0])
        print(f"  max|dphi| {np.abs(got['phi'] - ref['phi'][0]).max():.2e}, bit-identical {same}")
        np.testing.assert_allclose(got["phi"], ref["phi"][0], rtol=0, atol=1e-10)
        np.testing.assert_allclose(got["r"], ref["r"][0], rtol=0, atol=1e-10)
        assert abs(got["rnorm"] - ref["rnorm"][0]) <= 1e-10
        assert got["status"] == 0 == int(ref["status"][0])
        # and the oracle agrees (the parity bar of every evaluation)
        orc = O.evaluate(x, w, pt, N, bc=bc)
        np.testing.assert_allclose(got["phi"], unknowns(orc["phi"], bc), rtol=0, atol=1e-10)


def test_noalloc_rejects_instances_beyond_the_scratch():
    import torch
    import scft_amd as S
    m, N = 4096, 8192  # n = 4096 -> 4096 * 4096 * 8 B = 128 MiB of slabs > 64 MiB
    inst = instance(m, 6.0, N, bc=O.BC_DN, seed=5)
    x, pt, w = inst["x"], inst["pt"], inst["w"]
    rc, _ = _run(S, torch, x, unknowns(w, O.BC_DN), unknowns(pt, O.BC_DN), N, O.BC_DN)
    assert rc == 1  # SCFT_E_ARG
    assert "fixed scratch" in S.noalloc_error()
