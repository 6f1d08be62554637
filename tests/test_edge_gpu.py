"""Edge cases and error paths of the HIP path through the C ABI.

Covers what the reference's own checks imply: a single local field, q = 0
(iq = 1), NGd beyond a k-point's plane waves (the reference stops with
"NGd is bigger than NG1", Sloss.f90:339-342), calls out of order, compiled
limits, and a second q on the same context.  Tolerances as in
test_gpu_parity.py.
"""

import importlib

import numpy as np
import pytest

import helpers as H
from oracle import oracle as orc

host = importlib.import_module("2dmaterialsopticalprops_amd.host")
abi = importlib.import_module("2dmaterialsopticalprops_amd._abi")
pytestmark = pytest.mark.gpu
TOL = 1e-10


def _runner(s, **kw):
    return host.Runner(s, host.RunConfig(**kw))


def _sloss_ref(s, r):
    cfg = s.cfg
    pb = orc.Problem(s, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=0, ngd=r.ngd, E_hartree=r.E)
    S0 = orc.sloss_s0(pb, cfg.nval, cfg.no, r.domega, cfg.eta_ev / H.HARTREE, s.vcell)
    chi0 = orc.kk_sloss(S0, r.domega)
    V = orc.genV(r.rc.eps, r.q[0] * s.gcar, r.q[1] * s.gcar, cfg.c, r.parg, r.glf, s.gcar)
    E = np.array([orc.dyson_w(chi0[f], V)[0] for f in range(cfg.no - 1)])
    return S0, chi0, E


def test_single_local_field():
    s = H.system("tiny", nlf_target=1)
    r = _runner(s, program="sloss")
    assert r.nlf == 1
    out = r.full()
    S0, chi0, E = _sloss_ref(s, r)
    assert H.rel_max(out["S0"], S0) < TOL
    assert H.rel_max(out["epsinv"], E) < TOL
    # N_G = 1: chi = chi0 / (1 - V chi0)
    V = out["V"][0, 0]
    np.testing.assert_allclose(out["chi"][:, 0, 0], chi0[:, 0, 0] / (1 - V * chi0[:, 0, 0]), rtol=1e-9)
    r.close()


def test_q_zero_is_an_optical_limit_case():
    """iq = 1 gives q = 0: K2 = K1 and G0 = 0 for every k."""
    s = H.system("tiny")
    r = _runner(s, program="optabs", iq=1, pol="xx")
    assert np.all(r.K2 == r.K1)
    out = r.full()
    cfg = s.cfg
    pb = orc.Problem(s, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=0, ngd=r.ngd, E_hartree=r.E)
    S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, 1, r.domega, cfg.eta_ev / H.HARTREE, r.rc.lor_cut,
                          r.rc.df_cut, 0.0, 0.0, s.vcell, s.gcar)
    assert H.rel_max(out["S0"], S0) < TOL
    r.close()


def test_ngd_beyond_a_k_point_is_an_error():
    s = H.system("tiny")
    r = _runner(s, program="sloss")
    p = {f: getattr(r.ctx.params, f) for f, _ in abi.RpaParams._fields_}
    p["ngd"] = int(s.ngk.min()) + 1
    r.ctx.set_params(**p)
    with pytest.raises(abi.RpaError) as e:
        r.ctx.run()
    assert e.value.code == abi.RPA_E_ARG and "NGd" in str(e.value)
    r.close()


def test_calls_out_of_order_report_state_errors():
    s = H.system("tiny")
    cfg = s.cfg
    ctx = abi.Context(len(s.ngk), cfg.nband, s.npwx, 10, cfg.no, abi.RPA_SLOSS)
    with pytest.raises(abi.RpaError) as e:
        ctx.run()
    assert e.value.code == abi.RPA_E_STATE
    with pytest.raises(abi.RpaError) as e:
        ctx.kk()
    assert e.value.code == abi.RPA_E_STATE
    ctx.close()


def test_compiled_limits():
    s = H.system("tiny")
    cfg = s.cfg
    with pytest.raises(abi.RpaError) as e:
        abi.Context(len(s.ngk), cfg.nband, s.npwx, 1000, cfg.no, abi.RPA_SLOSS)   # Nlfd > 128
    assert e.value.code == abi.RPA_E_LIMIT
    with pytest.raises(abi.RpaError) as e:
        abi.Context(len(s.ngk), cfg.nband, s.npwx, 10, 3, abi.RPA_SLOSS)          # no < 4
    assert e.value.code == abi.RPA_E_ARG


def test_second_q_on_the_same_context():
    """The Sloss q loop: set_q again on a live context gives the fresh q's
    result (buffers are reused, nothing of the first q leaks)."""
    s = H.system("tiny")
    r = _runner(s, program="sloss")
    first = r.full()
    q3 = abi.find_min_q(r.ktot, 3)
    K1, K2, iG0 = abi.find_kq(r.ktot, s.kcart, r.g, q3, 1e-4)
    r.ctx.set_q(K1, K2, iG0, r.ktot, q3)
    r.ctx.run()
    S0 = r.ctx.result("S0")
    cfg = s.cfg
    pb = orc.Problem(s, K1, K2, iG0, r.ktot, q3, r.lf_idx, map_mode=0, ngd=r.ngd, E_hartree=r.E)
    ref = orc.sloss_s0(pb, cfg.nval, cfg.no, r.domega, cfg.eta_ev / H.HARTREE, s.vcell)
    assert H.rel_max(S0, ref) < TOL
    assert H.rel_max(first["S0"], ref) > 1e-3  # a different q
    r.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_q_loop_with_per_q_local_fields(precision):
    """The batched-W q loop on one context: every q brings a local-field set of
    its own size (the nearest |q+G|) and its own k+q map.  Each q gives
    bitwise what a fresh context gives for that q alone, and matches the
    oracle: FP64 to 1e-10; mixed to 2e-5 here, where a frequency bin holds
    only a few of the 64 transitions, so the 2^-16 relative error of a BF16
    hi/lo product barely averages out (the 1e-6 target is checked at
    workload sizes by the mixed chain tests)."""
    s = H.system("tiny", lf="xyz", nlf_target=9)
    r = _runner(s, program="sloss", precision=precision)
    for iq, nlf in [(2, 5), (3, 9), (5, 3), (2, 7)]:
        pt = r.q_point(iq, nlf)
        r.set_q_point(pt)
        r.ctx.run()
        assert r.nlf == nlf and r.ctx.nlf == nlf
        got = {k: r.ctx.result(k) for k in ("S0", "EPSINV", "W")}
        fresh = _runner(s, program="sloss", precision=precision)
        fresh.set_q_point(pt)
        fresh.ctx.run()
        for k, v in got.items():
            np.testing.assert_array_equal(v, fresh.ctx.result(k), err_msg=f"iq={iq} {k}")
        fresh.close()
        S0o, chi0o, Eo = _sloss_ref(s, r)
        V = orc.genV(r.rc.eps, r.q[0] * s.gcar, r.q[1] * s.gcar, s.cfg.c, r.parg, r.glf, s.gcar)
        Wo = np.array([orc.dyson_w(c, V)[2] for c in chi0o])
        errs = {"S0": H.rel_fro(got["S0"], S0o), "epsinv": H.rel_fro(got["EPSINV"], Eo),
                "W": H.rel_fro(got["W"], Wo)}
        print(f"{precision} iq={iq} N_G={nlf}", {k: f"{v:.2e}" for k, v in errs.items()})
        if precision == "fp64":
            assert H.rel_max(got["S0"], S0o) < TOL and H.rel_max(got["EPSINV"], Eo) < TOL
            assert H.rel_max(got["W"], Wo) < TOL
        else:
            assert max(errs.values()) < 2e-5
    r.close()


@pytest.mark.parametrize("pipe", ["0", "1"])
@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed-fp32"])
def test_column_form_multi_class(precision, pipe, monkeypatch):
    """The column-correlation vertex form, asked for on local fields in all
    directions (several in-plane classes, uneven shift sets) and on q points
    whose k+q wraps with G0, with the next chunk loaded during the products
    or between the barriers: vertices against the oracle (FP64 1e-10; mixed
    on the split BF16 hi/lo products 2e-5, on the FP32 matrix cores 2e-6
    max-norm as for the other FP32 forms)."""
    monkeypatch.setenv("RPA_VERTEX_FORM", "col")
    monkeypatch.setenv("RPA_VERTEX_COL_PIPE", pipe)
    monkeypatch.setenv("RPA_VERTEX_COL_SPLIT", "0" if precision == "mixed-fp32" else "1")
    tol = {"fp64": 1e-10, "mixed": 2e-5, "mixed-fp32": 2e-6}[precision]
    s = H.system("tiny", lf="xyz", nlf_target=9)
    r = _runner(s, program="sloss", precision=precision.split("-")[0])
    nv, nb = s.cfg.nval, s.cfg.nband
    for iq, nlf in [(2, 9), (4, 14), (6, 5)]:
        r.set_q_point(r.q_point(iq, nlf))
        pb = orc.Problem(s, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=0, ngd=r.ngd, E_hartree=r.E)
        ref = orc.vertices(pb, 0, s.gcar, 0, nv, nv, nb)
        got, _ = r.vertices()
        err = H.rel_max(got.T, ref)
        print(f"{precision} iq={iq} N_G={nlf} column-form vertices rel max err {err:.2e}")
        assert err < tol
    r.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed-fp32"])
def test_column_form_widest_tiles(precision, monkeypatch):
    """The column form at the largest tile count per wave: 16 conduction
    bands (4 groups of matrix-core columns) and 61 local fields along z
    (16 shift groups), so each wave runs 16 tiles, the largest template
    instance of every precision.  Vertices against the oracle (FP64 1e-10;
    mixed split BF16 2e-5, FP32 matrix cores 2e-6 max-norm)."""
    monkeypatch.setenv("RPA_VERTEX_FORM", "col")
    monkeypatch.setenv("RPA_VERTEX_COL_SPLIT", "0" if precision == "mixed-fp32" else "1")
    tol = {"fp64": 1e-10, "mixed": 2e-5, "mixed-fp32": 2e-6}[precision]
    s = H.system("tiny", ncond=16, nlf_target=64, ngk_target=250.0)
    r = _runner(s, program="sloss", precision=precision.split("-")[0])
    lf = s.mill_g[np.asarray(r.lf_idx) - 1]
    assert r.nlf == 61 and len({(int(x[0]), int(x[1])) for x in lf}) == 1
    nv, nb = s.cfg.nval, s.cfg.nband
    pb = orc.Problem(s, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=0, ngd=r.ngd, E_hartree=r.E)
    ref = orc.vertices(pb, 0, s.gcar, 0, nv, nv, nb)
    got, _ = r.vertices()
    err = H.rel_max(got.T, ref)
    print(f"{precision} N_G={r.nlf} column-form vertices, 16 tiles per wave: rel max err {err:.2e}")
    assert err < tol
    r.close()


@pytest.mark.parametrize("nlf_target", [60, 70, 90, 118])
def test_dyson_lds_and_global_forms(nlf_target):
    """N_G up to 64 runs the column form (eps in LDS, chi0 and V from L2);
    up to 100 the general form with eps in LDS; beyond (to the 128 fields
    the kernels take; the reference's gjel allows 1000) the general form
    with eps in place in the eps^-1 output.  All against the oracle."""
    s = H.system("tiny", lf="xyz", nlf_target=nlf_target)
    r = _runner(s, program="sloss")
    form = "col" if r.nlf <= 64 else "lds" if r.nlf <= 100 else "global"
    assert form == {60: "col", 70: "lds", 90: "lds", 118: "global"}[nlf_target], r.nlf
    out = r.full()
    S0, chi0, E = _sloss_ref(s, r)
    V = orc.genV(r.rc.eps, r.q[0] * s.gcar, r.q[1] * s.gcar, s.cfg.c, r.parg, r.glf, s.gcar)
    W = np.array([orc.dyson_w(c, V)[2] for c in chi0])
    assert H.rel_max(out["S0"], S0) < TOL
    assert H.rel_max(out["epsinv"], E) < TOL
    assert H.rel_max(out["W"], W) < TOL
    r.close()


@pytest.mark.parametrize("ncond", [2, 12])
def test_gather_form_paired_and_wide_blocks(ncond, monkeypatch):
    """The FP32 gather-direct vertex form with two local fields per wave
    (band blocks at most 8 bands m wide; odd N_G leaves the last wave half
    used) and with one field per wave (12 bands m): Sloss charge vertices on
    local fields in all directions against the oracle, 2e-6 max-norm as for
    the other FP32 forms."""
    monkeypatch.setenv("RPA_VERTEX_FORM", "gather")
    s = H.system("tiny", lf="xyz", nlf_target=9, ncond=ncond)
    r = _runner(s, program="sloss", precision="mixed")
    nv, nb = s.cfg.nval, s.cfg.nband
    for iq, nlf in [(2, 9), (4, 14), (6, 5), (3, 1)]:
        r.set_q_point(r.q_point(iq, nlf))
        pb = orc.Problem(s, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=0, ngd=r.ngd, E_hartree=r.E)
        ref = orc.vertices(pb, 0, s.gcar, 0, nv, nv, nb)
        got, _ = r.vertices()
        err = H.rel_max(got.T, ref)
        print(f"ncond={ncond} iq={iq} N_G={nlf} gather-form vertices rel max err {err:.2e}")
        assert err < 2e-6
    r.close()


@pytest.mark.parametrize("pol", ["xx", "yz"])
def test_gather_form_paired_current_vertices(pol, monkeypatch):
    """Optabs current vertices (J weights shuffled from the lanes of each
    field) through the paired gather-direct form: both band directions of
    the tiny system are 2-wide blocks."""
    monkeypatch.setenv("RPA_VERTEX_FORM", "gather")
    s = H.system("tiny", lf="xyz", nlf_target=9)
    r = _runner(s, program="optabs", precision="mixed", pol=pol)
    nv, nb = s.cfg.nval, s.cfg.nband
    for iq, nlf in [(2, 9), (4, 6)]:
        r.set_q_point(r.q_point(iq, nlf))
        pb = orc.Problem(s, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=0, ngd=r.ngd, E_hartree=r.E)
        a1, a2 = orc.vertices(pb, host._abi.POL[pol], s.gcar, 0, nv, nv, nb, both=True)
        r.ctx.vertices()
        g1, g2 = r.vertices(both=True)
        nk, tv = len(r.K1), nv * (nb - nv)
        g1 = g1.T.reshape(nk, 2 * tv, r.nlf)
        g2 = g2.T.reshape(nk, 2 * tv, r.nlf)
        e1 = H.rel_max(g1[:, :tv], a1.reshape(nk, tv, r.nlf))
        e2 = H.rel_max(g2[:, :tv], a2.reshape(nk, tv, r.nlf))
        print(f"{pol} iq={iq} N_G={nlf} paired gather current vertices rel max err {e1:.2e} {e2:.2e}")
        assert e1 < 2e-6 and e2 < 2e-6
    r.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("ncond", [2, 12])
def test_staged_form_paired_and_wide_blocks(precision, ncond, monkeypatch):
    """The staged vertex form with two local fields per wave (band blocks at
    most 8 bands m wide; odd N_G leaves the last wave half used) and with one
    (12 bands m): Sloss charge vertices on local fields in all directions
    against the oracle (FP64 1e-10, FP32 2e-6 max-norm)."""
    monkeypatch.setenv("RPA_VERTEX_FORM", "staged")
    s = H.system("tiny", lf="xyz", nlf_target=9, ncond=ncond)
    r = _runner(s, program="sloss", precision=precision)
    nv, nb = s.cfg.nval, s.cfg.nband
    for iq, nlf in [(2, 9), (4, 14), (6, 5), (3, 1)]:
        r.set_q_point(r.q_point(iq, nlf))
        pb = orc.Problem(s, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=0, ngd=r.ngd, E_hartree=r.E)
        ref = orc.vertices(pb, 0, s.gcar, 0, nv, nv, nb)
        got, _ = r.vertices()
        err = H.rel_max(got.T, ref)
        print(f"{precision} ncond={ncond} iq={iq} N_G={nlf} staged-form vertices rel max err {err:.2e}")
        assert err < (1e-10 if precision == "fp64" else 2e-6)
    r.close()


@pytest.mark.parametrize("program,precision", [("sloss", "fp64"), ("sloss", "mixed"), ("optabs", "fp64"),
                                               ("optabs", "mixed")])
def test_column_box_basis_is_bitwise_the_global_order(program, precision, monkeypatch):
    """The column vertex form reads the basis in column-box order (Miller
    columns contiguous, bands split at Nval, zeros past each k's ngk) instead
    of the global order with a per-row ngk test: the staged values are the
    same, so the vertices are bitwise equal (RPA_VERTEX_BOX=0 reads the
    global order)."""
    monkeypatch.setenv("RPA_VERTEX_FORM", "col")
    s = H.system("cfg0")
    got = {}
    for box in ("1", "0"):
        monkeypatch.setenv("RPA_VERTEX_BOX", box)
        r = _runner(s, program=program, precision=precision, pol="xx")
        got[box] = r.vertices()[0]
        r.close()
    assert np.array_equal(got["1"], got["0"])


@pytest.mark.parametrize("case", ["cfg0", "multi-class", "widest"])
def test_split_column_form_dma_staging_is_bitwise_the_register_staging(case, monkeypatch):
    """Mixed charge vertices of the split column form stage the k+q rows by
    LDS DMA from a BF16 hi/lo copy of the box basis; RPA_VERTEX_COL_DMA=0
    stages them through registers and splits them there.  The chunks are the
    same and so are the staged words, so the vertices are bitwise equal: on
    the benchmark-like system, on several in-plane classes with G0 wraps, and
    at 16 tiles per wave (16 conduction bands, the widest staged rows)."""
    monkeypatch.setenv("RPA_VERTEX_FORM", "col")
    if case == "cfg0":
        s, qs = H.system("cfg0"), [None]
    elif case == "multi-class":
        s, qs = H.system("tiny", lf="xyz", nlf_target=9), [(2, 9), (4, 14), (6, 5)]
    else:
        s, qs = H.system("tiny", ncond=16, nlf_target=64, ngk_target=250.0), [None]
    got = {}
    for dma in ("1", "0"):
        monkeypatch.setenv("RPA_VERTEX_COL_DMA", dma)
        r = _runner(s, program="sloss", precision="mixed")
        got[dma] = []
        for q in qs:
            if q is not None:
                r.set_q_point(r.q_point(*q))
            got[dma].append(r.vertices()[0])
        r.close()
    for a, b in zip(got["1"], got["0"]):
        assert np.isfinite(a).all()
        assert np.array_equal(a, b)
