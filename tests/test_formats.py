"""Reference file formats, the free-energy diagnostic and mesh refinement
(host code of the library, no device): checked on data files the reference
ships (copied under tests/golden/files by tests/golden/make_fixtures.py) and
on the values the reference printed into them."""
import os

import numpy as np
import pytest

from tests.helpers import fixtures
from tests.test_capi_symbols import _lib

FILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "files")


@pytest.fixture(scope="module")
def S():
    _lib()
    import scft_amd
    return scft_amd


def test_read_res_header_and_columns(S):
    r = S.read_res(os.path.join(FILES, "Exp_m32_n2048_IE.res"))
    h = r["header"]
    assert (h["m"], h["n"], h["iters"]) == (32, 2048, 1000)
    assert h["l"] == 3.72374357332160 and h["t"] == 5.30252230020752e-01
    assert h["mphi"] == 8.92581217765433e-01 and h["fmax"] == 7.628679e-11
    assert h["Z"] == 1.002289e+00 and h["f"] == 2.01673546641703e-03
    assert len(r["eta"]) == 33
    assert np.array_equal(r["eta"], np.array(fixtures()["res_m32"]["eta"]))
    assert r["x_over_l"][0] == 0.0 and r["x_over_l"][-1] == 1.0
    # 1D_FEM.c:331-340 skips 10 lines: its yita_middle is rows 1..M-2
    assert r["eta"][1] == -8.66338687955388e+00


def test_read_solution_files(S):
    s = S.read_solution(os.path.join(FILES, "N=33_for_read.txt"))
    assert len(s["w"]) == 33
    assert s["err"] == 1.422819e-09 and s["free_energy"] == 0.001945037280931
    assert np.array_equal(s["w"], np.array(fixtures()["dealii_N33"]["w"]))
    old = S.read_solution(os.path.join(FILES, "solution_yita_1D_N=141.txt"))
    assert len(old["w"]) == 141 and np.isnan(old["free_energy"])
    assert old["w"][1] == 14.625567952600205 and old["x"][1] == 0.007272929687500


def test_write_read_round_trip(S, tmp_path):
    x = np.linspace(0.0, 3.7, 17)
    w = np.sin(x) * 12.345678901234567
    p = tmp_path / "solution_yita_1D_N= 17.txt"
    S.write_solution(p, x, w, 1.5e-9, 0.00123456789012345)
    txt = p.read_text().splitlines()
    assert txt[0] == "N= 17, ERROR= 1.500000e-09 "
    assert txt[1] == "mean_field_free_energy, 0.001234567890123 "
    assert txt[2] == "0,0.000000000000000,0.000000000000000"
    s = S.read_solution(p)
    assert np.abs(s["w"] - w).max() <= 5e-16 * 100 and np.abs(s["x"] - x).max() <= 5e-16 * 10
    assert s["err"] == 1.5e-9


def test_free_energy_reproduces_the_reference_value(S):
    """The deal.II run's printed mean-field free energy of the field in
    N=33_for_read.txt (scft.cc:404-450, 2^18 + 1 points)."""
    s = S.read_solution(os.path.join(FILES, "N=33_for_read.txt"))
    F = S.free_energy(s["x"], s["w"], L=s["x"][-1], mirror=True, log2_points=18)
    print(f"free energy {F:.15f} vs reference {s['free_energy']:.15f}")
    assert abs(F - s["free_energy"]) <= 1e-12


def test_free_energy_closed_form(S):
    """w = c (constant): F = (c * int phi_t / f0bar / L + ln f0bar) / -1000,
    and int_0^L phi_t = f0bar L (scft.cc:368 defines f0bar that way)."""
    x = np.linspace(0.0, 3.72374357332160, 9)
    F = S.free_energy(x, np.full(9, 2.0), mirror=True, log2_points=16)
    f0 = 0.892581217773656
    assert abs(F - (2.0 + np.log(f0)) / -1000.0) <= 1e-10


def test_refine_rule_and_spline_transfer(S):
    """refine_mesh.m: wall cells always (inf error), cells at >= 10x the
    median gradient bisected; the transferred field is the not-a-knot
    spline, exact for a cubic."""
    x = np.linspace(0.0, 4.0, 41)
    w = 0.1 * x.copy()
    w[20] += 5.0  # steep cells 19 and 20
    xn, wn = S.refine_mesh(x, w, factor=10.0, bc=S.BC_DD)
    added = sorted(set(np.round(xn, 12)) - set(np.round(x, 12)))
    assert added == [0.05, 1.95, 2.05, 3.95]
    # DN: only the wall cell at x = 0 is forced
    xn2, _ = S.refine_mesh(x, 0.1 * x, factor=10.0, bc=S.BC_DN)
    assert sorted(set(np.round(xn2, 12)) - set(np.round(x, 12))) == [0.05]
    p = lambda t: 1.0 - 0.5 * t + 0.3 * t ** 2 - 0.05 * t ** 3
    wc = p(x)
    wc[0] = wc[-1] = 0.0  # wall values are conventions, not interpolated
    xn3, wn3 = S.refine_mesh(x, wc, factor=10.0, bc=S.BC_DD)
    assert np.abs(wn3[1:-1] - p(xn3[1:-1])).max() <= 1e-12
    assert wn3[0] == 0.0 and wn3[-1] == 0.0
