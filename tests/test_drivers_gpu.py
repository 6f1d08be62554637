"""The Fortran drivers end to end on the GPU: synthetic QE-format inputs in a
run directory, sloss_hip.x / optabs_hip.x, and their output files against the
oracle on the same inputs.

Tolerances follow the file formats: values written with full precision
(Loss_Qi es24.15, list-directed doubles) to 1e-10 relative; W_Qi (10F15.5)
to its 5 decimals; Corrfun/Qeff (2F15.10) to their 1e-10 absolute
quantisation.  The mixed-precision runs (namelist precision = 'mixed') use
the 4x4-k test system, whose 64 transitions per frequency leave the FP32
rounding of the vertices unaveraged: the loss and W to 3e-6, Pi to 5e-6
(the production-size bounds, eps^-1 and W to 1e-6, are asserted in
test_production_gpu.py).
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
build = importlib.import_module("2dmaterialsopticalprops_amd.build")
synth = H.synth


def _exe(name):
    path = os.path.join(build.BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    return path


def _run(exe, cwd):
    print(f"[drivers] {os.path.basename(exe)} in {cwd}", flush=True)
    r = subprocess.run([exe], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _floats(line):
    """Real numbers of a line (those with a decimal point or an exponent)."""
    pat = r"[-+]?(?:\d+\.\d*|\.\d+)(?:[eEdD][-+]?\d+)?|[-+]?\d+[eEdD][-+]?\d+"
    return [float(x.replace("D", "E").replace("d", "e")) for x in re.findall(pat, line)]


def _table(path):
    rows = [_floats(l) for l in open(path) if l.strip()]
    if len({len(r) for r in rows}) != 1 or len(rows[0]) < 2:
        head = "".join(open(path).readlines()[:4])
        raise AssertionError(f"unexpected layout of {path}:\n{head}")
    return np.array(rows)


def _sloss_oracle(s):
    cfg = s.cfg
    pb = H.setup(s, "sloss")
    dw = H.sloss_grid(cfg)
    S0 = orc.sloss_s0(pb, cfg.nval, cfg.no, dw, cfg.eta_ev / H.HARTREE, s.vcell)
    chi0 = orc.kk_sloss(S0, dw)
    V = orc.genV(1e-4, pb.q[0] * s.gcar, pb.q[1] * s.gcar, cfg.c, pb.parg, pb.glf, s.gcar)
    outs = [orc.dyson_w(chi0[f], V) for f in range(cfg.no - 1)]
    E = np.array([o[0] for o in outs])
    W = np.array([o[2] for o in outs])
    return dw, S0, V, E, W


def test_sloss_driver_outputs(tmp_path):
    s = H.system("tiny")
    cfg = s.cfg
    run = tmp_path / "sloss"
    synth.write_all(str(run), s, "sloss")
    _run(_exe("sloss_hip.x"), run)
    dw, S0, V, E, W = _sloss_oracle(s)
    iq = cfg.iq
    loss = _table(run / f"Loss_Q{iq}")
    np.testing.assert_allclose(loss[:, 0], np.arange(cfg.no - 1) * dw * H.HARTREE, rtol=1e-14)
    assert H.rel_max(loss[:, 1], -E[:, 0, 0].imag) < 1e-10
    re_w = _table(run / "fort.10008")
    im_w = _table(run / "fort.20008")
    assert H.rel_max(re_w[:, 1], W[:, 0, 0].real) < 1e-10
    assert H.rel_max(im_w[:, 1], W[:, 0, 0].imag) < 1e-10
    im22 = _table(run / "fort.22308")
    assert H.rel_max(im22[:, 1], W[:, 1, 1].imag) < 1e-10
    # W_Qi: io = 1, rows of (re, im) pairs over jG, iG outer, 10F15.5
    vals = np.array([v for l in open(run / f"W_Q{iq}").readlines()[1:] for v in _floats(l)])
    want = np.stack([W[0].real, W[0].imag], axis=-1).reshape(-1)
    np.testing.assert_allclose(vals, want, rtol=0, atol=6e-6)
    # Kramers-Kronig check from the oracle's W and genW1W2 (Sloss.f90:489-534)
    SW = S0.copy()
    SW[:cfg.no - 1] = -(1.0 / np.pi) * W.imag
    kks = skk = 0.0
    for io in range(1, cfg.no):
        w1, w2 = orc.w1w2(io, 1, 1, SW, dw)
        if io == 1:
            g0 = w1
        wind = (W[io - 1, 0, 0] - V[0, 0]).real
        fact = 0.5 * dw if io in (1, cfg.no - 1) else dw
        kks += fact * ((w1 - w2) - wind) ** 2
        skk += fact * wind * wind
    txt = open(run / f"Kramers-Kron_Q{iq}").read().splitlines()
    got_kks = _floats(txt[2])[-1]
    got_skk = _floats(txt[3])[-1]
    assert got_kks == pytest.approx(kks, rel=1e-8)
    assert got_skk == pytest.approx(skk, rel=1e-9)
    assert _floats(txt[8])[-1] == pytest.approx(g0, rel=1e-9)
    info = open(run / "Info").read()
    assert re.search(r"Number of local field vectors is\s+%d" % len(V), info)


@pytest.mark.parametrize("temp", [0.0, 0.5])
def test_optabs_driver_outputs(tmp_path, temp):
    s = H.system("tiny")
    cfg = s.cfg
    run = tmp_path / "optabs"
    synth.write_all(str(run), s, "optabs", extra={"pol": "xx", "calc": 3, "T": temp})
    _run(_exe("optabs_hip.x"), run)
    pb = H.setup(s, "optabs")
    dw = H.optabs_grid(cfg)
    S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, 1, dw, cfg.eta_ev / H.HARTREE, 1e-3, 1e-3, 0.0,
                          temp / H.HARTREE, s.vcell, s.gcar)
    nlf = S0.shape[1]
    # Corrfun_xx: per io a header line, then nlf^2 lines (iG outer, jG inner)
    lines = open(run / "Corrfun_xx").read().splitlines()
    assert len(lines) == (2 * cfg.no + 1) * (nlf * nlf + 1)
    got = np.array([_floats(l) for k, l in enumerate(lines) if k % (nlf * nlf + 1)])
    got = (got[:, 0] + 1j * got[:, 1]).reshape(2 * cfg.no + 1, nlf, nlf)
    quant = 0.51e-10  # half the last decimal of F15.10, per component
    assert np.abs((got - S0).view(np.float64)).max() <= quant + 1e-12 * np.abs(S0).max()
    gq = _table(run / "Qeff_xx")
    gq = (gq[:, 0] + 1j * gq[:, 1]).reshape(nlf, nlf)
    assert np.abs((gq - Q).view(np.float64)).max() <= quant + 1e-12 * np.abs(Q).max()
    if temp > 0:
        assert np.abs(Q).max() > 1e-6
    # second part from the files, as the reference does
    Pi0 = orc.kk_optabs(got, dw)
    pi11, si, sa = orc.pi_assembly(Pi0, gq, dw, 0.01 / H.HARTREE, cfg.c)
    pr = _table(run / "Pi_RPA_xx")
    assert H.rel_max(pr[:, 1] + 1j * pr[:, 2], pi11) < 1e-10
    f401 = _table(run / "fort.401")
    f402 = _table(run / "fort.402")
    assert H.rel_max(f401[:, 1], si[1:]) < 1e-10
    if temp > 0:
        assert H.rel_max(f402[:, 1], sa[1:]) < 1e-10
    assert os.path.exists(run / "Info_xx.out")


def test_sloss_driver_mixed_precision(tmp_path):
    """precision = 'mixed' in &system: the loss and W traces within the mixed
    chain's bound of the FP64 oracle."""
    s = H.system("tiny")
    cfg = s.cfg
    run = tmp_path / "sloss_mixed"
    synth.write_all(str(run), s, "sloss", extra={"precision": "mixed"})
    _run(_exe("sloss_hip.x"), run)
    dw, S0, V, E, W = _sloss_oracle(s)
    loss = _table(run / f"Loss_Q{cfg.iq}")
    assert H.rel_max(loss[:, 1], -E[:, 0, 0].imag) < 3e-6
    re_w = _table(run / "fort.10008")
    im_w = _table(run / "fort.20008")
    assert H.rel_max(re_w[:, 1] + 1j * im_w[:, 1], W[:, 0, 0]) < 3e-6


def test_optabs_driver_mixed_precision(tmp_path):
    s = H.system("tiny")
    cfg = s.cfg
    run = tmp_path / "optabs_mixed"
    synth.write_all(str(run), s, "optabs", extra={"pol": "xx", "calc": 3, "precision": "mixed"})
    _run(_exe("optabs_hip.x"), run)
    pb = H.setup(s, "optabs")
    dw = H.optabs_grid(cfg)
    S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, 1, dw, cfg.eta_ev / H.HARTREE, 1e-3, 1e-3, 0.0, 0.0,
                          s.vcell, s.gcar)
    Pi0 = orc.kk_optabs(S0, dw)
    pi11, si, sa = orc.pi_assembly(Pi0, Q, dw, 0.01 / H.HARTREE, cfg.c)
    pr = _table(run / "Pi_RPA_xx")
    assert H.rel_max(pr[:, 1] + 1j * pr[:, 2], pi11) < 5e-6


def test_driver_rejects_unknown_precision(tmp_path):
    s = H.system("tiny")
    run = tmp_path / "sloss_bad"
    synth.write_all(str(run), s, "sloss", extra={"precision": "fp16"})
    r = subprocess.run([_exe("sloss_hip.x")], cwd=run, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "precision must be" in r.stdout + r.stderr


def _as_read(ibz, R):
    """The symmetric system's inputs as the drivers parse them: KC and k to the
    6 decimals of scf.out and the band file, R to the 7 of scf.out."""
    import dataclasses
    six = np.vectorize(lambda x: float(f"{x:.6f}"))
    s2 = dataclasses.replace(ibz, kc=six(ibz.kc), kcart=six(ibz.kcart))
    R2 = np.vectorize(lambda x: float(f"{x:.7f}"))(R)
    return s2, R2


def _sym_oracle(s2, R2, nsym, program, ecut, eps=1e-4):
    """Oracle-side setup with point operations: genFBZ, PointR, findMinQ,
    findKinIBZ / findKQinBZ over the operations, genGlf."""
    import ctypes as C
    cfg = s2.cfg
    g = synth.gcart(s2.mill_g, s2.kc)
    kt = np.zeros((3, 48 * len(s2.kcart)), order="F")
    nt = C.c_int()
    assert orc.lib().orc_gen_fbz(len(s2.kcart), nsym, eps, np.asfortranarray(s2.kcart.T), np.asfortranarray(R2),
                                 C.byref(nt), kt) == 0
    ktot = kt[:, :nt.value].T.copy()
    RI = orc.point_inverse(R2, nsym)
    q = orc.find_min_q(ktot, cfg.iq)
    R1 = np.zeros(len(ktot), np.int32)
    K1 = np.zeros(len(ktot), np.int32)
    for ik in range(len(ktot)):
        R1[ik], K1[ik] = orc.find_kin_ibz(ik + 1, s2.kcart, ktot[ik], nsym, RI, eps)
    iG0, R2k, K2 = orc.find_kq_sym_all(ktot, s2.kcart, g, q, nsym, RI, eps)
    rc, nlf, glf, parg, idx = orc.gen_glf(cfg.lf, ecut, s2.gcar, g, s2.mill_g, 1000)
    assert rc == 0
    pb = orc.Problem(s2, K1, K2, iG0, ktot, q, idx, map_mode=0, ngd=int(s2.ngk.min()),
                     E_hartree=H.energies(s2, program), R=R2, RI=RI, R1=R1, R2=R2k, eps=eps)
    pb.glf, pb.parg = glf, parg
    assert (R1 != 1).any() and (R2k != 1).any()
    return pb


def _symmetric_pair(kind):
    import dataclasses
    cfg = dataclasses.replace(synth.CONFIGS["tiny"], nk=6, lf="xyz", nlf_target=9)
    ibz, fz, R, nsym, kinfo, ngd = synth.make_symmetric_pair(cfg, kind, 90, 60)
    assert ngd <= int(ibz.ngk.min())
    return ibz, R, nsym


def test_sloss_driver_unfolds_point_operations(tmp_path):
    """Nsym > 1 in scf.out: the driver inverts the operations (PointR), finds
    R1/K1 and R2/K2 of every FBZ k and runs the unfolded chain; its loss and
    W against the oracle on the inputs as the driver reads them."""
    ibz, R, nsym = _symmetric_pair("c6")
    cfg = ibz.cfg
    run = tmp_path / "sloss_sym"
    synth.write_all(str(run), ibz, "sloss", ops=(R, nsym))
    out = _run(_exe("sloss_hip.x"), run)
    s2, R2 = _as_read(ibz, R)
    pb = _sym_oracle(s2, R2, nsym, "sloss", synth.local_field_ecut(ibz))
    dw = H.sloss_grid(cfg)
    S0 = orc.sloss_s0(pb, cfg.nval, cfg.no, dw, cfg.eta_ev / H.HARTREE, s2.vcell)
    chi0 = orc.kk_sloss(S0, dw)
    V = orc.genV(1e-4, pb.q[0] * s2.gcar, pb.q[1] * s2.gcar, cfg.c, pb.parg, pb.glf, s2.gcar)
    outs = [orc.dyson_w(chi0[f], V) for f in range(cfg.no - 1)]
    E = np.array([o[0] for o in outs])
    W = np.array([o[2] for o in outs])
    loss = _table(run / f"Loss_Q{cfg.iq}")
    assert H.rel_max(loss[:, 1], -E[:, 0, 0].imag) < 1e-10
    re_w = _table(run / "fort.10008")
    im_w = _table(run / "fort.20008")
    assert H.rel_max(re_w[:, 1] + 1j * im_w[:, 1], W[:, 0, 0]) < 1e-10
    info = open(run / "Info").read()
    assert re.search(r"Number of point symmetry operation is\s+%d" % nsym, info), out


def test_optabs_driver_unfolds_point_operations(tmp_path):
    ibz, R, nsym = _symmetric_pair("c3v")
    cfg = ibz.cfg
    run = tmp_path / "optabs_sym"
    synth.write_all(str(run), ibz, "optabs", extra={"pol": "yy", "calc": 3}, ops=(R, nsym))
    _run(_exe("optabs_hip.x"), run)
    s2, R2 = _as_read(ibz, R)
    pb = _sym_oracle(s2, R2, nsym, "optabs", synth.local_field_ecut(ibz))
    dw = H.optabs_grid(cfg)
    S0, Q = orc.optabs_s0(pb, cfg.nval, cfg.no, 2, dw, cfg.eta_ev / H.HARTREE, 1e-3, 1e-3, 0.0, 0.0,
                          s2.vcell, s2.gcar)
    nlf = S0.shape[1]
    lines = open(run / "Corrfun_yy").read().splitlines()
    got = np.array([_floats(l) for k, l in enumerate(lines) if k % (nlf * nlf + 1)])
    got = (got[:, 0] + 1j * got[:, 1]).reshape(2 * cfg.no + 1, nlf, nlf)
    quant = 0.51e-10  # half the last decimal of F15.10, per component
    assert np.abs((got - S0).view(np.float64)).max() <= quant + 1e-12 * np.abs(S0).max()
    gq = _table(run / "Qeff_yy")
    gq = (gq[:, 0] + 1j * gq[:, 1]).reshape(nlf, nlf)
    assert np.abs((gq - Q).view(np.float64)).max() <= quant + 1e-12 * np.abs(Q).max()
    # the second part from the files, as the reference does
    Pi0 = orc.kk_optabs(got, dw)
    pi11, si, sa = orc.pi_assembly(Pi0, gq, dw, 0.01 / H.HARTREE, cfg.c)
    pr = _table(run / "Pi_RPA_yy")
    assert H.rel_max(pr[:, 1] + 1j * pr[:, 2], pi11) < 1e-10


@pytest.mark.parametrize("program,kind", [("sloss", "c6"), ("optabs", "c3v")])
def test_drivers_take_operations_and_axes_from_the_xml_file(tmp_path, program, kind):
    """&directories xml_file: the operations and reciprocal axes come from a
    data-file-schema.xml (Modules/io_xml.f90) instead of scf.out, whose
    operation and axes lines are removed.  The same system run both ways:
    Sloss's loss and W, Optabs' Corrfun/Pi, equal within the difference of
    the two sources' operations (the xml rebuilds them from integer crystal
    forms and the 6-decimal axes, scf.out has 7-decimal cartesian rows)."""
    ibz, R, nsym = _symmetric_pair(kind)
    cfg = ibz.cfg
    extra = {"pol": "yy", "calc": 3} if program == "optabs" else None
    outs = {}
    for how in ("scf", "xml"):
        run = tmp_path / how
        synth.write_all(str(run), ibz, program, extra=extra, ops=(R, nsym))
        if how == "xml":
            synth.write_data_file_schema(str(run / "data-file-schema.xml"), ibz, ops=(R, nsym))
            cfgtxt = open(run / "config.in").read().replace("&directories", "&directories\n xml_file = "
                                                              "'data-file-schema.xml',", 1)
            open(run / "config.in", "w").write(cfgtxt)
            scf = open(run / "scf.out").read().splitlines()
            keep = [l for l in scf if "reciprocal axes" not in l and " b(" not in l and "Sym. Ops" not in l
                    and "cryst." not in l and "cart." not in l]
            open(run / "scf.out", "w").write("\n".join(keep) + "\n")
        print(f"[drivers] {program} {kind} operations from {how}", flush=True)
        _run(_exe(f"{program}_hip.x"), run)
        if program == "sloss":
            outs[how] = (_table(run / f"Loss_Q{cfg.iq}")[:, 1], _table(run / "fort.10008")[:, 1],
                         _table(run / "fort.20008")[:, 1])
        else:
            outs[how] = (_table(run / "Pi_RPA_yy")[:, 1:3],)
    # the two operation sets differ by ~1e-6 (the 6-decimal axes in KC s KC^-1):
    # the unfolded k list, q and the current vertices R(R1) G follow them
    for a, b in zip(outs["xml"], outs["scf"]):
        assert H.rel_max(a, b) < 2e-5
