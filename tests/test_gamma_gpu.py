"""GW propagators on the GPU against the oracle: the Gamma+- matrices of the
resident W for every (G, G') pair (rpa_gamma_w, Sloss.f90:489-512 with
genW1W2) and genGammaPropagator for batches of series (rpa_gamma_propagator,
Modules/propagators.f90:11-108).  The device sums the stencils on the FP64
matrix cores in another order than the oracle's loops: 1e-12 of the largest
value."""

import importlib

import numpy as np
import pytest

import helpers as H
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
abi = importlib.import_module("2dmaterialsopticalprops_amd._abi")
host = importlib.import_module("2dmaterialsopticalprops_amd.host")


@pytest.mark.parametrize("no,nser", [(5, 1), (40, 7), (701, 300)])
def test_gamma_propagator_matches_oracle(no, nser):
    print(f"[gamma] propagator no={no} nser={nser}", flush=True)
    rng = np.random.default_rng(no + nser)
    dw = 0.011
    D = np.asfortranarray(rng.standard_normal((no, nser)) + 1j * rng.standard_normal((no, nser)))
    gp, gm, kks, skk = abi.gamma_propagator(D, dw)
    for s in range(nser):
        wp, wm, a, b = orc.gamma_propagator(D[:, s], dw)
        assert H.rel_max(gp[:, s], wp) < 1e-12
        assert H.rel_max(gm[:, s], wm) < 1e-12
        assert kks[s] == pytest.approx(a, rel=1e-10)
        assert skk[s] == pytest.approx(b, rel=1e-13)


def test_gamma_of_w_matches_oracle_genw1w2_for_every_pair():
    s = H.system("tiny")
    cfg = s.cfg
    r = host.Runner(s, host.RunConfig(program="sloss"))
    r.ctx.run()
    W = r.ctx.result("W")
    S0 = r.ctx.result("S0")
    gp, gm = r.ctx.gamma_w()
    nlf = W.shape[1]
    print(f"[gamma] W of tiny: no={cfg.no} nlf={nlf}", flush=True)
    SW = S0.copy()
    SW[:cfg.no - 1] = -(1.0 / np.pi) * W.imag
    wp = np.zeros_like(gp)
    wm = np.zeros_like(gm)
    for io in range(1, cfg.no):
        for i in range(nlf):
            for j in range(nlf):
                a, b = orc.w1w2(io, i + 1, j + 1, SW, r.domega)
                wp[io - 1, i, j] = a - 1j * np.pi * SW[io - 1, i, j]
                wm[io - 1, i, j] = -b
    assert H.rel_max(gp, wp) < 1e-12
    assert H.rel_max(gm, wm) < 1e-12
    # (1,1) equals the W Kramers-Kronig check's sums
    w1, w2 = r.ctx.w_kk()
    assert H.rel_max(gp[:, 0, 0].real, w1) < 1e-12
    assert H.rel_max(gm[:, 0, 0].real, -w2) < 1e-12
    r.close()
