"""Point-group unfolding (Nsym > 1) on the device.

An irreducible k list with point operations (C2, C3, C6, C3v of the
hexagonal lattice) is unfolded on the fly: each FBZ k carries R1 and R2
(findKinIBZ / findKQinBZ, Optabs.f90:776-869), the vertices take C2 at
RI(R2) (R(R1) G'' + Glf + G0) and the current vertex J has R(R1) G''
(Optabs.f90:1399-1449, Sloss.f90:907-946, Modules/vertices.f90:88-394).

Each case is checked two ways:
  * against the oracle on the same irreducible inputs (to 1e-10 in FP64);
  * against the same system written out on the whole zone, C_Rk(R G) = C_k(G)
    and E(Rk) = E(k), run with Nsym = 1 on the device: the unfolding must
    give the whole zone's results (to 1e-10).  The oracle itself is pinned
    the same way on the CPU.
Mixed precision is checked against the FP64 oracle at the mixed tolerance
(5e-6 relative max, as the other mixed chain tests).
"""

import dataclasses
import importlib

import numpy as np
import pytest

import helpers as H
from oracle import oracle as orc

host = importlib.import_module("2dmaterialsopticalprops_amd.host")
abi = importlib.import_module("2dmaterialsopticalprops_amd._abi")
synth = H.synth
pytestmark = pytest.mark.gpu
TOL = 1e-10
TOL_MIXED = 5e-6


def _pair(kind, nk=6):
    cfg = dataclasses.replace(synth.CONFIGS["tiny"], nk=nk, lf="xyz", nlf_target=9)
    ibz, fz, R, nsym, kinfo, ngd = synth.make_symmetric_pair(cfg, kind, 90, 60)
    return cfg, ibz, fz, R, nsym, ngd


def _runners(kind, **kw):
    cfg, ibz, fz, R, nsym, ngd = _pair(kind)
    print(f"[symmetry] {kind} nsym={nsym} nki={len(ibz.ngk)} ntot={len(fz.ngk)} {kw}", flush=True)
    rs = host.Runner(ibz, host.RunConfig(ngd=ngd, **kw), ops=(R, nsym))
    rf = host.Runner(fz, host.RunConfig(ngd=ngd, **kw))
    assert len(rs.K1) == len(rf.K1) == len(fz.ngk)
    assert (rs.R1 != 1).any() and (rs.R2 != 1).any()
    return cfg, ibz, fz, R, nsym, rs, rf


def _oracle_problem(s, r, R, nsym, map_mode=0):
    RI = orc.point_inverse(R, nsym)
    return orc.Problem(s, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=map_mode, ngd=r.ngd,
                       E_hartree=r.E, R=R, RI=RI, R1=r.R1, R2=r.R2)


@pytest.mark.parametrize("kind", ["c2", "c3", "c6", "c3v"])
def test_sloss_unfolded_equals_oracle_and_whole_zone(kind):
    cfg, ibz, fz, R, nsym, rs, rf = _runners(kind, program="sloss")
    out_s, out_f = rs.full(), rf.full()
    pb = _oracle_problem(ibz, rs, R, nsym)
    v = orc.vertices(pb, 0, ibz.gcar, 0, cfg.nval, cfg.nval, cfg.nband)
    S0 = orc.sloss_s0(pb, cfg.nval, cfg.no, rs.domega, cfg.eta_ev / H.HARTREE, ibz.vcell)
    assert H.rel_max(rs.vertices()[0].T, v) < TOL
    assert H.rel_max(out_s["S0"], S0) < TOL
    for k in ("S0", "chi0", "V", "epsinv", "chi", "W", "loss"):
        assert H.rel_max(out_s[k], out_f[k]) < TOL, k
    rs.close()
    rf.close()


@pytest.mark.parametrize("pol", ["xx", "yy", "yz"])
def test_optabs_unfolded_equals_oracle_and_whole_zone(pol):
    cfg, ibz, fz, R, nsym, rs, rf = _runners("c3v", program="optabs", pol=pol)
    out_s, out_f = rs.full(), rf.full()
    pb = _oracle_problem(ibz, rs, R, nsym)
    polk = {"xx": 1, "yy": 2, "yz": 4}[pol]
    S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, polk, rs.domega, cfg.eta_ev / H.HARTREE, rs.rc.lor_cut,
                          rs.rc.df_cut, 0.0, 0.0, ibz.vcell, ibz.gcar)
    assert H.rel_max(out_s["S0"], S0) < TOL
    assert H.rel_max(out_s["Qeff"], Q) < TOL
    for k in ("S0", "Qeff", "chi0", "pi11", "sig"):
        assert H.rel_max(out_s[k], out_f[k]) < TOL, k
    rs.close()
    rf.close()


def test_igk_map_unfolded_equals_oracle_and_whole_zone():
    cfg, ibz, fz, R, nsym, rs, rf = _runners("c6", program="sloss", vertex_map="igk")
    out_s, out_f = rs.full(), rf.full()
    pb = _oracle_problem(ibz, rs, R, nsym, map_mode=1)
    S0 = orc.sloss_s0(pb, cfg.nval, cfg.no, rs.domega, cfg.eta_ev / H.HARTREE, ibz.vcell)
    assert H.rel_max(out_s["S0"], S0) < TOL
    for k in ("S0", "epsinv", "W"):
        assert H.rel_max(out_s[k], out_f[k]) < TOL, k
    rs.close()
    rf.close()


@pytest.mark.parametrize("program", ["sloss", "optabs"])
def test_mixed_unfolded_against_fp64_oracle(program):
    cfg, ibz, fz, R, nsym, ngd = _pair("c6")
    r = host.Runner(ibz, host.RunConfig(program=program, precision="mixed", ngd=ngd), ops=(R, nsym))
    out = r.full()
    pb = _oracle_problem(ibz, r, R, nsym)
    if program == "sloss":
        S0 = orc.sloss_s0(pb, cfg.nval, cfg.no, r.domega, cfg.eta_ev / H.HARTREE, ibz.vcell)
    else:
        S0, _ = orc.optabs_s0(pb, cfg.nval, cfg.no, 1, r.domega, cfg.eta_ev / H.HARTREE, r.rc.lor_cut,
                              r.rc.df_cut, 0.0, 0.0, ibz.vcell, ibz.gcar)
    err = H.rel_max(out["S0"], S0)
    print(f"[symmetry] mixed {program} S0 rel max {err:.2e}", flush=True)
    assert err < TOL_MIXED
    r.close()


@pytest.mark.parametrize("form", ["col", "gather", "staged"])
def test_every_vertex_form_request_serves_the_unfolding(form, monkeypatch):
    """The column form needs translation-only targets: asked for under point
    operations, the library runs a gather form instead; every request gives
    the oracle's vertices."""
    monkeypatch.setenv("RPA_VERTEX_FORM", form)
    cfg, ibz, fz, R, nsym, ngd = _pair("c3")
    for prec, tol in (("fp64", TOL), ("mixed", TOL_MIXED)):
        r = host.Runner(ibz, host.RunConfig(program="sloss", precision=prec, ngd=ngd), ops=(R, nsym))
        pb = _oracle_problem(ibz, r, R, nsym)
        v = orc.vertices(pb, 0, ibz.gcar, 0, cfg.nval, cfg.nval, cfg.nband)
        assert H.rel_max(r.vertices()[0].T, v) < tol, prec
        r.close()


def test_batched_q_loop_with_operations_equals_single_q_runs(monkeypatch):
    """With the vertex form forced for both, each q of a batch with point
    operations gives bitwise what a single-q run gives."""
    monkeypatch.setenv("RPA_VERTEX_FORM", "staged")
    cfg, ibz, fz, R, nsym, ngd = _pair("c6")
    r = host.Runner(ibz, host.RunConfig(program="sloss", ngd=ngd), ops=(R, nsym))
    pts = [r.q_point(iq, nlf) for iq, nlf in ((2, 9), (3, 13))]
    assert all("R1" in p for p in pts)
    ctx = r.ctx
    ctx.batch_set(pts, r.ktot)
    ctx.batch_run()
    for iq, pt in enumerate(pts):
        r.set_q_point(pt)
        ctx.run()
        for name in ("EPSINV", "W"):
            assert np.array_equal(ctx.batch_result(iq, name), ctx.result(name)), (iq, name)
    r.close()


def test_symmetry_state_errors():
    cfg, ibz, fz, R, nsym, ngd = _pair("c3")
    r = host.Runner(ibz, host.RunConfig(program="sloss", ngd=ngd), ops=(R, nsym))
    ctx = r.ctx
    # a 45-degree rotation does not map the hexagonal lattice onto itself
    bad = R.copy()
    c = np.sqrt(0.5)
    bad[1] = np.array([[c, -c, 0.0], [c, c, 0.0], [0.0, 0.0, 1.0]])
    with pytest.raises(abi.RpaError) as e:
        ctx.set_symmetry(bad, 2)
    assert e.value.code == abi.RPA_E_ARG and "lattice" in str(e.value)
    # operation numbers beyond Nsym
    with pytest.raises(abi.RpaError) as e:
        ctx.set_q_sym(r.K1, r.K2, r.iG0, r.R1, np.full_like(r.R2, nsym + 1), r.ktot, r.q)
    assert e.value.code == abi.RPA_E_ARG
    # a new G list clears the operations
    ctx.set_gvectors(ibz.mill_g, ibz.kc)
    with pytest.raises(abi.RpaError) as e:
        ctx.set_q_sym(r.K1, r.K2, r.iG0, r.R1, r.R2, r.ktot, r.q)
    assert e.value.code == abi.RPA_E_STATE
    ctx.set_symmetry(R, nsym)
    ctx.set_local_fields(r.lf_idx, r.parg)
    ctx.set_q_sym(r.K1, r.K2, r.iG0, r.R1, r.R2, r.ktot, r.q)
    ctx.run()
    r.close()
