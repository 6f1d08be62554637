"""Boundary behaviour of the C ABI: singular dielectric matrices, band-block
lists longer than one vertex launch takes, the accumulation's LDS limit at
large Nlf, and the state a new G list leaves behind.

The reference's gjel prints 'singular matrix in gjel' on a zero pivot and goes
on with the division (Chi_RPA/gjel.f:59-62).  Here the run completes, returns
RPA_E_SINGULAR, and its results stay readable; the flag does not outlive the
run.  The condition is forced from the host with non-finite inputs (a NaN
spectral function or a NaN coefficient), which the pivot search treats like a
zero pivot; every kernel still exits normally.
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
TOL_MIXED = 5e-6


def _runner(s, **kw):
    return host.Runner(s, host.RunConfig(**kw))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_singular_eps_from_loaded_s0_is_reported_and_not_sticky(precision):
    s = H.system("tiny")
    r = _runner(s, program="sloss", precision=precision)
    ctx = r.ctx
    good = r.full()
    bad = good["S0"].copy()
    bad[3, 0, 0] = np.nan
    ctx.load_s0(bad)
    ctx.kk(want=False)
    ctx.coulomb()
    with pytest.raises(abi.RpaError) as e:
        ctx.dyson_w()
    assert e.value.code == abi.RPA_E_SINGULAR
    assert not np.isfinite(ctx.result("EPSINV")).all()
    # a finite S0 on the same context: no error, the previous eps^-1 again
    ctx.load_s0(good["S0"])
    ctx.kk(want=False)
    ctx.coulomb()
    epsinv, _, W = ctx.dyson_w()
    assert np.array_equal(epsinv, good["epsinv"])
    assert np.array_equal(W, good["W"])
    r.close()


@pytest.mark.parametrize("graph", [False, True])
def test_singular_eps_in_the_chain_is_reported_and_not_sticky(graph):
    s = H.system("tiny")
    r = _runner(s, program="sloss")
    ctx = r.ctx
    ctx.run(graph=graph)
    ref = ctx.result("EPSINV")
    n = int(s.ngk[0])
    evc = s.evc[0, :, :n].T.copy()
    evc[0, 0] = np.nan
    ctx.upload_basis(1, s.mill_k[0, :n], evc)
    with pytest.raises(abi.RpaError) as e:
        ctx.run(graph=graph)
    assert e.value.code == abi.RPA_E_SINGULAR
    assert "singular" in str(e.value)
    assert not np.isfinite(ctx.result("EPSINV")).all()  # results stay readable
    ctx.upload_basis(1, s.mill_k[0, :n], s.evc[0, :, :n].T)
    ctx.run(graph=graph)
    assert np.array_equal(ctx.result("EPSINV"), ref)
    r.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("temp", [0.5, 0.0])
def test_optabs_forty_bands_all_pairs(precision, temp):
    """40 bands with all pairs (T > 0, or df_cut = 0 at T = 0) make 15 band
    blocks of 8 x 16: more than one vertex launch takes, so the library runs
    them as several launches."""
    s = H.system("tiny", nval=16, ncond=24, ngk_target=120.0)
    r = _runner(s, program="optabs", pol="xx", precision=precision, temp_ev=temp,
                df_cut=(1.0e-3 if temp > 0 else 0.0))
    cfg = s.cfg
    out = r.full()
    pb = orc.Problem(s, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=0, ngd=r.ngd, E_hartree=r.E)
    S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, 1, r.domega, cfg.eta_ev / H.HARTREE, r.rc.lor_cut,
                          r.rc.df_cut, 0.0, temp / H.HARTREE, s.vcell, s.gcar)
    tol = TOL if precision == "fp64" else TOL_MIXED
    assert H.rel_max(out["S0"], S0) < tol
    assert H.rel_max(out["Qeff"], Q) < tol
    pi0 = orc.kk_optabs(S0, r.domega)
    pi11, si, _ = orc.pi_assembly(pi0, Q, r.domega, r.rc.eta_intra_ev / H.HARTREE, cfg.c)
    assert H.rel_max(out["pi11"], pi11) < tol
    r.close()


def _ecut_for(s, lo, hi):
    """Ecut (Hartree) of the 'xyz' rule that selects between lo and hi fields."""
    g = H.synth.gcart(s.mill_g, s.kc) * s.gcar
    e = np.sort(np.unique(np.round(0.5 * np.einsum("ij,ij->i", g, g), 12)))
    for v in e:
        n = int((0.5 * np.einsum("ij,ij->i", g, g) <= v + 1e-12).sum())
        if lo <= n <= hi:
            return float(v + 1e-9), n
    raise AssertionError("no shell count in range")


def test_fp64_optabs_accumulation_beyond_64k_lds():
    """FP64 Optabs stages 4352 + 512 Nlf bytes of LDS: past 64 KB from
    Nlf = 120 (the library raises the kernel's limit at setup)."""
    s = H.system("tiny", lf="xyz", nlf_target=118)
    ecut, n = _ecut_for(s, 120, 128)
    r = host.Runner(s, host.RunConfig(program="optabs", pol="zz"), ecut_lf=ecut)
    assert 120 <= r.nlf <= 128
    cfg = s.cfg
    out = r.full()
    pb = orc.Problem(s, r.K1, r.K2, r.iG0, r.ktot, r.q, r.lf_idx, map_mode=0, ngd=r.ngd, E_hartree=r.E)
    S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, 3, r.domega, cfg.eta_ev / H.HARTREE, r.rc.lor_cut,
                          r.rc.df_cut, 0.0, 0.0, s.vcell, s.gcar)
    assert H.rel_max(out["S0"], S0) < TOL
    r.close()


def test_new_g_list_requires_local_fields_and_q_again():
    s = H.system("tiny")
    r = _runner(s, program="sloss")
    ctx = r.ctx
    ctx.run()
    ref = ctx.result("EPSINV")
    ctx.set_gvectors(s.mill_g, s.kc)
    with pytest.raises(abi.RpaError) as e:
        ctx.run()
    assert e.value.code == abi.RPA_E_STATE and "local_fields" in str(e.value)
    ctx.set_local_fields(r.lf_idx, r.parg)
    with pytest.raises(abi.RpaError) as e:
        ctx.run()
    assert e.value.code == abi.RPA_E_STATE and "set_q" in str(e.value)
    ctx.set_q(r.K1, r.K2, r.iG0, r.ktot, r.q)
    ctx.run()
    assert np.array_equal(ctx.result("EPSINV"), ref)
    r.close()
