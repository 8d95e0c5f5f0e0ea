"""Extract the reference's own data files into small fixtures under tests/golden.

Reads (text only, nothing executed):
  Exp_m32_n2048_IE.res, Exp_m1024_n2048_IE.res   Wang-solver fields: x/l, phi, eta
  Matlab_files/inputFiles/solution_matlab_N=33   converged MATLAB IE field
  DEALII_SCFT/inputFiles/N=33_for_read.txt        deal.II IRK field
and the constants quoted in the reference sources (f0bar KAT, scft.cc:368).
Also records the reference's own compiled NR routines' outputs (oracle/_ref,
built from /root/reference by oracle/Makefile) on fixed inputs, so tests that
run where the reference tree is absent still pin the oracle against them.

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def read_res(path):
    lines = open(path).read().splitlines()
    head = " ".join(lines[:6])
    rows = []
    for line in lines[9:]:
        parts = line.split()
        if len(parts) == 5:
            rows.append([float(p) for p in parts])
    a = np.array(rows)
    return dict(header=head, x_over_l=a[:, 0].tolist(), phi=a[:, 1].tolist(),
                eta=a[:, 2].tolist())


def read_csv_field(path, skip=1):
    rows = []
    for line in open(path).read().splitlines()[skip:]:
        if line.strip():
            rows.append([float(v) for v in line.split(",")])
    a = np.array(rows)
    return dict(i=a[:, 0].astype(int).tolist(), x=a[:, 1].tolist(), w=a[:, 2].tolist())


def ref_vectors():
    """Outputs of the reference's compiled romint / adm / gaussj on fixed inputs."""
    import oracle as O
    O.build()
    out = {}
    rng = np.random.default_rng(20261202)
    rom = []
    for m in (16, 32, 64, 128, 1024, 2048):
        f = rng.standard_normal(m + 1)
        rom.append(dict(m=m, f=f.tolist(), value=O.ref.dealii.romint(f, m, 1.0 / m)))
    out["romint"] = rom
    # gaussj (DEALII version) on a fixed 4x4 system, solution and inverse
    n = 4
    A = rng.standard_normal((n, n)) + 3 * np.eye(n)
    b = rng.standard_normal(n)
    L = O.ref.dealii
    L.dmatrix.restype = C.POINTER(C.POINTER(C.c_double))
    L.dmatrix.argtypes = [C.c_long] * 4
    a = L.dmatrix(1, n, 1, n)
    bb = L.dmatrix(1, n, 1, 1)
    for i in range(n):
        for j in range(n):
            a[i + 1][j + 1] = A[i, j]
        bb[i + 1][1] = b[i]
    L.gaussj.argtypes = [C.POINTER(C.POINTER(C.c_double)), C.c_int,
                         C.POINTER(C.POINTER(C.c_double)), C.c_int]
    rc = L.gaussj(a, n, bb, 1)
    out["gaussj"] = dict(A=A.tolist(), b=b.tolist(), rc=rc,
                         x=[bb[i + 1][1] for i in range(n)],
                         inv=[[a[i + 1][j + 1] for j in range(n)] for i in range(n)])
    # adm on the linear toy of 1D_FEM.c:372-379 (flag 0)
    top = O.ref.top
    FP = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double))

    def toy(n_, x, xn):
        xn[0] = x[0] * 0.5 - 2.0
        xn[1] = x[1] * 0.5 + 3.0

    cb = FP(toy)
    x = (C.c_double * 2)(1.0, 2.0)
    chk = C.c_int(1)
    res = top.adm(x, 2, C.byref(chk), C.cast(cb, C.c_void_p), 0)
    out["adm_toy"] = dict(x0=[1.0, 2.0], check=chk.value, xnew=[res[0], res[1]])
    return out


def main():
    fx = {}
    fx["res_m32"] = read_res(os.path.join(REF, "Exp_m32_n2048_IE.res"))
    fx["res_m1024"] = read_res(os.path.join(REF, "Exp_m1024_n2048_IE.res"))
    fx["matlab_N33"] = read_csv_field(
        os.path.join(REF, "Matlab_files/inputFiles/solution_matlab_N=33"))
    fx["dealii_N33"] = read_csv_field(os.path.join(REF, "DEALII_SCFT/inputFiles/N=33_for_read.txt"),
                                      skip=2)
    fx["constants"] = dict(f0bar=0.892581217773656, tau=0.530252230020752, L=3.72374357332160,
                           tau_short=0.5302, L_short=3.72374,
                           res_m32_fmax=7.628679e-11, res_m1024_fmax=2.700742e-12,
                           dealii_N33_error=1.422819e-09)
    fx["ref_nr"] = ref_vectors()
    path = os.path.join(HERE, "reference_fixtures.json")
    with open(path, "w") as f:
        json.dump(fx, f)
    print(path, os.path.getsize(path), "bytes")
    # data files the readers are tested on, byte for byte (no spaces in names)
    import shutil
    files = os.path.join(HERE, "files")
    os.makedirs(files, exist_ok=True)
    for src, dst in (("Exp_m32_n2048_IE.res", "Exp_m32_n2048_IE.res"),
                     ("DEALII_SCFT/inputFiles/N=33_for_read.txt", "N=33_for_read.txt"),
                     ("Matlab_files/inputFiles/solution_yita_1D_N= 141.txt",
                      "solution_yita_1D_N=141.txt")):
        shutil.copyfile(os.path.join(REF, src), os.path.join(files, dst))
        print(os.path.join(files, dst))


if __name__ == "__main__":
    main()
