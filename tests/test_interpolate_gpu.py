"""interpolate.x on the GPU (rpa_interpolate, interpolate_hip.x) against the
oracle (orc_kk_optabs + orc_interpolate, Chi_RPA/interpolate.f90:103-186).

Values are FP64 throughout; the device KK sums the stencil on the FP64
matrix cores in another order than the oracle's loop, so results agree to
1e-12 of the largest |Pi| (the KK tests of the Optabs chain use the same
bound).  Driver outputs are list-directed doubles, compared to 1e-10.
"""

import importlib
import os
import re
import subprocess

import numpy as np
import pytest

import helpers as H
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
abi = importlib.import_module("2dmaterialsopticalprops_amd._abi")
build = importlib.import_module("2dmaterialsopticalprops_amd.build")

TOL = 1e-12
HARTREE = 2.0 * 13.6056923


def _inputs(no, nlf, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    s0 = scale * (rng.standard_normal((2 * no + 1, nlf, nlf)) + 1j * rng.standard_normal((2 * no + 1, nlf, nlf)))
    qeff = scale * (rng.standard_normal((nlf, nlf)) + 1j * rng.standard_normal((nlf, nlf)))
    return np.asfortranarray(s0), np.asfortranarray(qeff)


def _oracle(s0, qeff, no_interp, dw, dwt, eta):
    return orc.interpolate(orc.kk_optabs(s0, dw), qeff, no_interp, dwt, eta)


@pytest.mark.parametrize("no,nlf,no_interp", [(9, 3, 10), (40, 5, 7), (101, 2, 3), (33, 1, 2), (70, 33, 4)])
@pytest.mark.parametrize("all_pairs", [True, False])
def test_interpolate_matches_oracle(no, nlf, no_interp, all_pairs):
    print(f"[interp] no={no} nlf={nlf} no_interp={no_interp} all={all_pairs}", flush=True)
    s0, qeff = _inputs(no, nlf, seed=no * 7 + nlf)
    dw, dwt, eta = 0.011, 0.011 * (no - 1) / ((no_interp - 1) * (no - 2)), 0.0009
    got, intra = abi.interpolate(s0, qeff, no_interp, dw, dwt, eta, all_pairs=all_pairs)
    want, want_intra = _oracle(s0, qeff, no_interp, dw, dwt, eta)
    if not all_pairs:
        want = want[:, 0, 0]
    assert got.shape == want.shape
    assert np.isfinite(got).all()
    assert H.rel_max(got, want) < TOL
    assert H.rel_max(intra, want_intra) < 1e-15


def test_interpolate_one_pair_equals_the_all_pairs_column():
    """The (1,1)-only call (write_GiGj false) runs the KK over one column; the
    pairs are independent, so its result is the all-pairs call's (1,1)
    column, bit for bit."""
    s0, qeff = _inputs(60, 6, seed=3)
    a, ia = abi.interpolate(s0, qeff, 5, 0.02, 0.004, 0.001, all_pairs=True)
    b, ib = abi.interpolate(s0, qeff, 5, 0.02, 0.004, 0.001, all_pairs=False)
    assert np.array_equal(a[:, 0, 0], b)
    assert np.array_equal(ia, ib)


def test_interpolate_full_size_against_the_oracle_on_a_pair_subset():
    """No = 2001 coarse frequencies, No_interp = 10 (17991 dense points),
    Nlf = 65, every pair: the pairs are independent, so the oracle on the
    sub-block of local fields {1, 33, 65} must give the device's values of
    those pairs; the rest are checked finite and the intraband column
    against its closed form."""
    no, nlf, nint = 2001, 65, 10
    print(f"[interp] full size no={no} nlf={nlf} no_interp={nint}", flush=True)
    s0, qeff = _inputs(no, nlf, seed=11, scale=1e-2)
    omax = 50.0 / HARTREE
    dw, dwt, eta = omax / (no - 1), omax / ((nint - 1) * (no - 2)), 0.025 / HARTREE
    got, intra = abi.interpolate(s0, qeff, nint, dw, dwt, eta, all_pairs=True)
    assert got.shape == ((nint - 1) * (no - 2), nlf, nlf)
    assert np.isfinite(got).all()
    idx = np.array([0, 32, 64])
    sub = np.asfortranarray(s0[:, idx][:, :, idx])
    want, want_intra = _oracle(sub, np.asfortranarray(qeff[np.ix_(idx, idx)]), nint, dw, dwt, eta)
    assert H.rel_max(got[:, idx][:, :, idx], want) < TOL
    oi = np.arange(1, got.shape[0] + 1) * dwt
    np.testing.assert_allclose(intra, qeff[0, 0] * oi / (oi + 1j * eta), rtol=1e-14)


# ---------------------------------------------------------------------------
# the drop-in program

def _floats(line):
    pat = r"[-+]?(?:\d+\.\d*|\.\d+)(?:[eEdD][-+]?\d+)?|[-+]?\d+[eEdD][-+]?\d+"
    return [float(x.replace("D", "E").replace("d", "e")) for x in re.findall(pat, line)]


def _write_inputs(run, pol, s0, qeff, dw):
    no = (s0.shape[0] - 1) // 2
    nlf = s0.shape[1]
    with open(os.path.join(run, f"Corrfun_{pol}"), "w") as f:
        for io in range(-no, no + 1):
            f.write(f" omega= {io * dw:.15f} Hartree\n")
            for i in range(nlf):
                for j in range(nlf):
                    v = s0[io + no, i, j]
                    f.write(f"{v.real:15.10f}{v.imag:15.10f}\n")
    with open(os.path.join(run, f"Qeff_{pol}"), "w") as f:
        for i in range(nlf):
            for j in range(nlf):
                f.write(f"{qeff[i, j].real:15.10f}{qeff[i, j].imag:15.10f}\n")


def _read_back(run, pol, no, nlf):
    """The inputs as the program reads them (2F15.10)."""
    s0 = np.zeros((2 * no + 1, nlf, nlf), dtype=np.complex128, order="F")
    with open(os.path.join(run, f"Corrfun_{pol}")) as f:
        lines = f.read().splitlines()
    k = 0
    for w in range(2 * no + 1):
        k += 1
        for i in range(nlf):
            for j in range(nlf):
                s0[w, i, j] = float(lines[k][:15]) + 1j * float(lines[k][15:30])
                k += 1
    qeff = np.zeros((nlf, nlf), dtype=np.complex128, order="F")
    with open(os.path.join(run, f"Qeff_{pol}")) as f:
        lines = f.read().splitlines()
    for i in range(nlf):
        for j in range(nlf):
            ln = lines[i * nlf + j]
            qeff[i, j] = float(ln[:15]) + 1j * float(ln[15:30])
    return s0, qeff


@pytest.mark.parametrize("write_gigj", [True, False])
def test_interpolate_driver_matches_oracle(tmp_path, write_gigj):
    exe = os.path.join(build.BIN, "interpolate_hip.x")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (run __graft_entry__.build())")
    no, nlf, nint, pol = 30, 3, 5, "yy"
    omin, omax, eta_ev = 1.0e-5, 20.0, 0.025
    run = str(tmp_path)
    s0, qeff = _inputs(no, nlf, seed=21, scale=0.3)
    omin_h = omin / HARTREE
    omax_h = omax / HARTREE + omin_h
    dw = (omax_h - omin_h) / (no - 1)
    no_tot = (nint - 1) * (no - 2)
    dwt = (omax_h - omin_h) / no_tot
    eta = eta_ev / HARTREE
    _write_inputs(run, pol, s0, qeff, dw)
    flag = ".true." if write_gigj else ".false."
    with open(os.path.join(run, "config.interp.in"), "w") as f:
        f.write(f"&config\n pol = '{pol}'\n Nlf = {nlf}\n No_interp = {nint}\n No = {no}\n omin = {omin}\n"
                f" omax = {omax}\n eta_intra = {eta_ev}\n c0 = 32.5411\n write_GiGj = {flag}\n/\n")
    print(f"[interp] driver write_GiGj={write_gigj} in {run}", flush=True)
    r = subprocess.run([exe], cwd=run, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    s0r, qr = _read_back(run, pol, no, nlf)
    want, want_intra = _oracle(s0r, qr, nint, dw, dwt, eta)
    oi = np.arange(1, no_tot + 1) * dwt * HARTREE
    for name, ref in ((f"Pi_RPA_{pol}_dense", want[:, 0, 0]), (f"Pi_RPA_{pol}_dense_inter", want[:, 0, 0]),
                      (f"Pi_RPA_{pol}_dense_intra", want_intra)):
        rows = np.array([_floats(ln) for ln in open(os.path.join(run, name)).read().splitlines()])
        assert rows.shape == (no_tot, 3), name
        np.testing.assert_allclose(rows[:, 0], oi, rtol=1e-12)
        assert H.rel_max(rows[:, 1] + 1j * rows[:, 2], ref) < 1e-10, name
    path = os.path.join(run, f"Pi_RPA_{pol}_GiGj_dense")
    assert os.path.exists(path) == write_gigj
    if write_gigj:
        rows = np.array([_floats(ln) for ln in open(path).read().splitlines()])
        assert rows.shape == (no_tot * nlf * nlf, 3)
        got = (rows[:, 1] + 1j * rows[:, 2]).reshape(no_tot, nlf, nlf)
        np.testing.assert_allclose(rows[:, 0], np.repeat(oi, nlf * nlf), rtol=1e-12)
        assert H.rel_max(got, want) < 1e-10
