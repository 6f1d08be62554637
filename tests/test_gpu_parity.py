"""Parity of the HIP path (through the C ABI) with the CPU oracle.

FP64 tolerances (max-norm relative, |a-b|_max / |b|_max):
  vertices 1e-12; S0, chi0/Pi0, V, eps^-1, chi, W, loss, Pi 1e-10.
The oracle follows the reference loops; the GPU uses matrix-core
contractions, so agreement is to rounding, not bitwise.
"""

import importlib

import numpy as np
import pytest

import helpers as H
from oracle import oracle as orc

host = importlib.import_module("2dmaterialsopticalprops_amd.host")

pytestmark = pytest.mark.gpu

TOL_V = 1e-12
TOL = 1e-10
# Mixed vertices: FP32 matrix-core forms 2e-6; the column form of the charge
# vertices splits FP32 inputs into BF16 hi + lo and multiplies hi.hi + hi.lo
# + lo.hi, each product within 2^-16 of the FP32 one (measured 4.8e-6 on
# cfg0), so 2e-5 there.
TOL_V_MIXED = 2e-6
TOL_V_SPLIT = 2e-5


def _runner(sys_, **kw):
    return host.Runner(sys_, host.RunConfig(**kw))


def _oracle_problem(sys_, r, map_mode):
    pb = orc.Problem(sys_, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=map_mode,
                     ngd=r.ngd, E_hartree=r.E, eps=r.rc.eps)
    return pb


@pytest.mark.parametrize("name", ["tiny", "cfg0"])
@pytest.mark.parametrize("mode", ["direct", "igk"])
def test_setup_matches_oracle(name, mode):
    s = H.system(name)
    r = _runner(s, program="sloss", vertex_map=mode)
    pb = H.setup(s, "sloss", map_mode=0 if mode == "direct" else 1)
    np.testing.assert_array_equal(r.K2, pb.K2)
    np.testing.assert_array_equal(r.iG0, pb.iG0)
    np.testing.assert_array_equal(r.lf_idx, pb.lf_idx)
    np.testing.assert_array_equal(r.q, pb.q)
    r.close()


@pytest.mark.parametrize("name", ["tiny", "cfg0"])
@pytest.mark.parametrize("mode", ["direct", "igk"])
def test_charge_vertices(name, mode):
    s = H.system(name)
    r = _runner(s, program="sloss", vertex_map=mode)
    mm = 0 if mode == "direct" else 1
    pb = _oracle_problem(s, r, mm)
    nv, nb = s.cfg.nval, s.cfg.nband
    ref = orc.vertices(pb, 0, s.gcar, 0, nv, nv, nb)
    got, _ = r.vertices()
    assert got.shape == (r.nlf, ref.shape[0])
    assert H.rel_max(got.T, ref) < TOL_V
    r.close()


@pytest.mark.parametrize("pol", ["xx", "yy", "zz", "yz"])
def test_current_vertices(pol):
    s = H.system("tiny")
    r = _runner(s, program="optabs", vertex_map="direct", pol=pol)
    pb = _oracle_problem(s, r, 0)
    nv, nb = s.cfg.nval, s.cfg.nband
    kind = {"xx": 1, "yy": 2, "zz": 3, "yz": 4}[pol]
    a1, a2 = orc.vertices(pb, kind, s.gcar, 0, nv, nv, nb, both=True)
    b1, b2 = orc.vertices(pb, kind, s.gcar, nv, nb, 0, nv, both=True)
    g1, g2 = r.vertices(both=True)
    nk = len(r.K1)
    tv = nv * (nb - nv)
    g1 = g1.T.reshape(nk, 2 * tv, r.nlf)
    g2 = g2.T.reshape(nk, 2 * tv, r.nlf)
    assert H.rel_max(g1[:, :tv], a1.reshape(nk, tv, r.nlf)) < TOL_V
    assert H.rel_max(g1[:, tv:], b1.reshape(nk, tv, r.nlf)) < TOL_V
    assert H.rel_max(g2[:, :tv], a2.reshape(nk, tv, r.nlf)) < TOL_V
    assert H.rel_max(g2[:, tv:], b2.reshape(nk, tv, r.nlf)) < TOL_V
    r.close()


@pytest.mark.parametrize("name,over", [("tiny", {}), ("cfg0", {}), ("tiny", {"nlf_target": 17})])
@pytest.mark.parametrize("mode", ["direct", "igk"])
def test_sloss_chain(name, over, mode):
    s = H.system(name, **over)
    cfg = s.cfg
    r = _runner(s, program="sloss", vertex_map=mode)
    pb = _oracle_problem(s, r, 0 if mode == "direct" else 1)
    out = r.full()
    S0 = orc.sloss_s0(pb, cfg.nval, cfg.no, r.domega, cfg.eta_ev / H.HARTREE, s.vcell)
    assert H.rel_max(out["S0"], S0) < TOL
    chi0 = orc.kk_sloss(S0, r.domega)
    assert H.rel_max(out["chi0"], chi0) < TOL
    # KK alone, on the GPU's own S0
    assert H.rel_max(out["chi0"], orc.kk_sloss(out["S0"], r.domega)) < 1e-12
    V = orc.genV(r.rc.eps, r.q[0] * s.gcar, r.q[1] * s.gcar, cfg.c, r.parg, r.glf, s.gcar)
    assert H.rel_max(out["V"], V) < 1e-13
    E, C, W = [], [], []
    for f in range(cfg.no - 1):
        e, c, w = orc.dyson_w(out["chi0"][f], V)
        E.append(e), C.append(c), W.append(w)
    assert H.rel_max(out["epsinv"], np.array(E)) < TOL
    assert H.rel_max(out["chi"], np.array(C)) < TOL
    assert H.rel_max(out["W"], np.array(W)) < TOL
    assert H.rel_max(out["loss"], -np.array(E)[:, 0, 0].imag) < TOL
    r.close()


@pytest.mark.parametrize("pol", ["xx", "yz"])
@pytest.mark.parametrize("temp", [0.0, 0.5])
def test_optabs_chain(pol, temp):
    s = H.system("tiny")
    cfg = s.cfg
    r = _runner(s, program="optabs", vertex_map="direct", pol=pol, temp_ev=temp)
    pb = _oracle_problem(s, r, 0)
    out = r.full()
    S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, host._abi.POL[pol], r.domega, cfg.eta_ev / H.HARTREE,
                          r.rc.lor_cut, r.rc.df_cut, r.rc.efermi_ev / H.HARTREE,
                          temp / H.HARTREE, s.vcell, s.gcar)
    assert H.rel_max(out["S0"], S0) < TOL
    if temp > 0:
        assert np.abs(Q).max() > 0
        assert H.rel_max(out["Qeff"], Q) < TOL
    Pi0 = orc.kk_optabs(S0, r.domega)
    assert H.rel_max(out["chi0"], Pi0) < TOL
    pi11, si, sa = orc.pi_assembly(Pi0, Q, r.domega, r.rc.eta_intra_ev / H.HARTREE, cfg.c)
    assert H.rel_max(out["pi11"], pi11) < TOL
    assert H.rel_max(out["sig"][0], si) < TOL
    if temp > 0:
        assert H.rel_max(out["sig"][1], sa) < TOL
    r.close()


def test_optabs_cfg0_xx():
    s = H.system("cfg0")
    cfg = s.cfg
    r = _runner(s, program="optabs", vertex_map="igk", pol="xx")
    pb = _oracle_problem(s, r, 1)
    out = r.full()
    S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, 1, r.domega, cfg.eta_ev / H.HARTREE, r.rc.lor_cut,
                          r.rc.df_cut, 0.0, 0.0, s.vcell, s.gcar)
    assert H.rel_max(out["S0"], S0) < TOL
    assert H.rel_max(out["chi0"], orc.kk_optabs(S0, r.domega)) < TOL
    r.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("program", ["sloss", "optabs"])
def test_graph_replay_matches_eager(program, precision):
    """Replays of the captured chain, with eager runs and host reads between
    them, reproduce the eager results bit for bit (the chain is
    deterministic and its graph holds no memset nodes)."""
    s = H.system("cfg0")
    r = _runner(s, program=program, precision=precision)
    eager = r.full()
    for _ in range(3):
        r.ctx.run(graph=True)
        loss, pi11, sig = r.ctx.epilogue()
        if program == "sloss":
            np.testing.assert_array_equal(loss, eager["loss"])
            np.testing.assert_array_equal(r.ctx.result("W"), eager["W"])
        else:
            np.testing.assert_array_equal(pi11, eager["pi11"])
            np.testing.assert_array_equal(r.ctx.result("S0"), eager["S0"])
        r.ctx.run(graph=False)
    r.close()


def test_rerun_is_deterministic():
    s = H.system("cfg0")
    r = _runner(s, program="sloss")
    a = r.full()
    b = r.full()
    np.testing.assert_array_equal(a["S0"], b["S0"])
    np.testing.assert_array_equal(a["W"], b["W"])
    r.close()


def test_get_result_matches_stage_outputs_sloss():
    s = H.system("tiny")
    r = _runner(s, program="sloss")
    out = r.full()
    r.ctx.run()
    for name, key in (("S0", "S0"), ("CHI0", "chi0"), ("V", "V"), ("EPSINV", "epsinv"),
                      ("CHI", "chi"), ("W", "W"), ("LOSS", "loss")):
        np.testing.assert_array_equal(r.ctx.result(name), out[key])
    r.close()


def test_get_result_matches_stage_outputs_optabs():
    s = H.system("tiny")
    r = _runner(s, program="optabs", temp_ev=0.5)
    out = r.full()
    r.ctx.run()
    np.testing.assert_array_equal(r.ctx.result("S0"), out["S0"])
    np.testing.assert_array_equal(r.ctx.result("QEFF"), out["Qeff"])
    np.testing.assert_array_equal(r.ctx.result("CHI0"), out["chi0"])
    np.testing.assert_array_equal(r.ctx.result("PI11"), out["pi11"])
    np.testing.assert_array_equal(r.ctx.result("SIG").T, out["sig"])
    r.close()


@pytest.mark.parametrize("program", ["sloss", "optabs"])
def test_load_s0_then_kk_matches_oracle(program):
    """The calc = 2 path: an S0 (here the oracle's, quantised like the
    Corrfun text file) is placed on the device and transformed there."""
    s = H.system("tiny")
    cfg = s.cfg
    r = _runner(s, program=program, temp_ev=0.5 if program == "optabs" else 0.0)
    pb = _oracle_problem(s, r, 0)
    if program == "sloss":
        S0 = orc.sloss_s0(pb, cfg.nval, cfg.no, r.domega, cfg.eta_ev / H.HARTREE, s.vcell)
        Q = None
        ref = orc.kk_sloss(S0, r.domega)
    else:
        S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, 1, r.domega, cfg.eta_ev / H.HARTREE, r.rc.lor_cut,
                              r.rc.df_cut, 0.0, 0.5 / H.HARTREE, s.vcell, s.gcar)
        S0 = np.round(S0.real, 10) + 1j * np.round(S0.imag, 10)
        Q = np.round(Q.real, 10) + 1j * np.round(Q.imag, 10)
        ref = orc.kk_optabs(S0, r.domega)
    r.ctx.load_s0(S0, Q)
    chi0 = r.ctx.kk()
    assert H.rel_max(chi0, ref) < 1e-12
    if program == "optabs":
        _, pi11, sig = r.ctx.epilogue()
        p11, si, sa = orc.pi_assembly(ref, Q, r.domega, r.rc.eta_intra_ev / H.HARTREE, cfg.c)
        assert H.rel_max(pi11, p11) < 1e-12
        assert H.rel_max(sig[1], sa) < 1e-12
    # a full run afterwards starts again from its own S0 and a cleared Qeff
    r.ctx.run()
    np.testing.assert_array_equal(r.ctx.result("S0").shape, S0.shape)
    r.close()


def test_w_kk_check_matches_oracle_genw1w2():
    s = H.system("tiny")
    cfg = s.cfg
    r = _runner(s, program="sloss")
    r.ctx.run()
    W = r.ctx.result("W")
    S0 = r.ctx.result("S0")
    w1, w2 = r.ctx.w_kk()
    # Sloss.f90:489: S0(1:no-1) = -(1/pi) Im WT; S0(no) keeps its value
    SW = S0.copy()
    SW[:cfg.no - 1] = -(1.0 / np.pi) * W.imag
    for io in range(1, cfg.no):
        a, b = orc.w1w2(io, 1, 1, SW, r.domega)
        assert w1[io - 1] == pytest.approx(a, rel=1e-12, abs=1e-12 * np.abs(w1).max())
        assert w2[io - 1] == pytest.approx(b, rel=1e-12, abs=1e-12 * np.abs(w2).max())
    r.close()


@pytest.mark.parametrize("program,pol", [("sloss", "xx"), ("optabs", "yz")])
def test_staged_vertex_form_matches_oracle(program, pol, monkeypatch):
    """The LDS-staged vertex kernel (used for more than two band blocks) on
    cases other kernels normally take."""
    monkeypatch.setenv("RPA_VERTEX_FORM", "staged")
    s = H.system("cfg0")
    r = _runner(s, program=program, vertex_map="direct", pol=pol)
    pb = _oracle_problem(s, r, 0)
    nv, nb = s.cfg.nval, s.cfg.nband
    if program == "sloss":
        ref = orc.vertices(pb, 0, s.gcar, 0, nv, nv, nb)
        got, _ = r.vertices()
        assert H.rel_max(got.T, ref) < TOL_V
    else:
        a1, a2 = orc.vertices(pb, 4, s.gcar, 0, nv, nv, nb, both=True)
        g1, g2 = r.vertices(both=True)
        nk, tv = len(r.K1), nv * (nb - nv)
        g1 = g1.T.reshape(nk, 2 * tv, r.nlf)
        g2 = g2.T.reshape(nk, 2 * tv, r.nlf)
        assert H.rel_max(g1[:, :tv], a1.reshape(nk, tv, r.nlf)) < TOL_V
        assert H.rel_max(g2[:, :tv], a2.reshape(nk, tv, r.nlf)) < TOL_V
    r.close()


def test_many_band_blocks_optabs_full_pairs():
    """Finite temperature keeps every band pair: nband = 20 gives more than
    two 8 x 16 blocks, so the staged kernel runs (cfg0 bands widened)."""
    s = H.system("tiny", nval=6, ncond=14)
    cfg = s.cfg
    r = _runner(s, program="optabs", vertex_map="direct", pol="xx", temp_ev=0.5)
    pb = _oracle_problem(s, r, 0)
    out = r.full()
    S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, 1, r.domega, cfg.eta_ev / H.HARTREE, r.rc.lor_cut,
                          r.rc.df_cut, r.rc.efermi_ev / H.HARTREE, 0.5 / H.HARTREE, s.vcell, s.gcar)
    assert H.rel_max(out["S0"], S0) < TOL
    assert H.rel_max(out["Qeff"], Q) < TOL
    r.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("form", ["col", "col-fp32", "lds", "gather", "staged"])
@pytest.mark.parametrize("mode", ["direct", "igk"])
def test_vertex_kernel_forms(precision, form, mode, monkeypatch):
    """Every vertex kernel form on the Sloss charge vertices of cfg0 (the
    column form applies to the direct map; with igk the shape rule picks).
    col-fp32 is the column form on the FP32 matrix cores instead of the
    split BF16 products (same as col in FP64)."""
    monkeypatch.setenv("RPA_VERTEX_FORM", form.split("-")[0])
    monkeypatch.setenv("RPA_VERTEX_COL_SPLIT", "0" if form == "col-fp32" else "1")
    s = H.system("cfg0")
    r = host.Runner(s, host.RunConfig(program="sloss", vertex_map=mode, precision=precision))
    pb = _oracle_problem(s, r, 0 if mode == "direct" else 1)
    nv, nb = s.cfg.nval, s.cfg.nband
    ref = orc.vertices(pb, 0, s.gcar, 0, nv, nv, nb)
    got, _ = r.vertices()
    split = precision == "mixed" and form == "col" and mode == "direct"
    tol = TOL_V if precision == "fp64" else TOL_V_SPLIT if split else TOL_V_MIXED
    err = H.rel_max(got.T, ref)
    print(f"{precision} {form} {mode} vertices rel max err {err:.2e} (tol {tol:.0e})")
    assert err < tol
    r.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pol", ["xx", "zz", "yz"])
def test_column_form_current_vertices(precision, pol, monkeypatch):
    """Optabs current vertices through the column-correlation form
    (j = alpha(G) rho + rho_beta) against the oracle's direct sums with the
    reference's J(G, i) per term."""
    monkeypatch.setenv("RPA_VERTEX_FORM", "col")
    s = H.system("cfg0")
    r = host.Runner(s, host.RunConfig(program="optabs", vertex_map="direct", pol=pol, precision=precision))
    pb = _oracle_problem(s, r, 0)
    nv, nb = s.cfg.nval, s.cfg.nband
    a1, a2 = orc.vertices(pb, host._abi.POL[pol], s.gcar, 0, nv, nv, nb, both=True)
    g1, g2 = r.vertices(both=True)
    nk, tv = len(r.K1), nv * (nb - nv)
    g1 = g1.T.reshape(nk, 2 * tv, r.nlf)
    g2 = g2.T.reshape(nk, 2 * tv, r.nlf)
    tol = TOL_V if precision == "fp64" else TOL_V_MIXED
    e1 = H.rel_max(g1[:, :tv], a1.reshape(nk, tv, r.nlf))
    e2 = H.rel_max(g2[:, :tv], a2.reshape(nk, tv, r.nlf))
    print(f"{precision} {pol} column-form current vertices rel max err {e1:.2e} {e2:.2e}")
    assert e1 < tol and e2 < tol
    r.close()
