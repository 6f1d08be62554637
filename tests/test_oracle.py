"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

Each oracle routine is checked against something that does not restate it:
  - the reference's own routines (tests/golden/ref_golden.npz, written by
    tests/golden/make_golden.py from oracle/_ref/libref.so, which is compiled
    from Modules/response_function_KK.f90, qe_gksort.f90, local_field.f90 and
    Chi_RPA/gjel.f where they lie), and the live library when it is built;
  - closed forms: orthonormality of the Bloch coefficients, the slab-averaged
    2D Coulomb integral, the continuum principal value on a uniform grid;
  - brute force in numpy on tiny inputs (Miller-index dictionaries instead of
    the oracle's index arithmetic);
  - identities: eps^-1 = I + chi V, Hermiticity, LAPACK inverses.
"""

import ctypes as C
import math
import os

import numpy as np
import pytest

import helpers as H
from oracle import oracle as orc

synth = H.synth
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_golden.npz"))
NOS = (5, 6, 9, 17)


# ---------------------------------------------------------------------------
# Kramers-Kronig stencils

def _k1_table(no, dw):
    """Optabs K1 stencil of the oracle as a table (io, w, part_in, part_out)."""
    nw = 2 * no + 1
    t = np.zeros((no - 1, nw, 2, 2))
    for w in range(nw):
        for part in (0, 1):
            s = np.zeros((nw, 1, 1), dtype=np.complex128)
            s[w, 0, 0] = 1.0 if part == 0 else 1.0j
            out = orc.kk_optabs(s, dw)[:, 0, 0]
            t[:, w, part, 0] = out.real
            t[:, w, part, 1] = out.imag
    return t


def _k2_table(no, dw):
    """Sloss K2 stencil of the oracle: (io, j) -> chi0 for a unit S0(j)."""
    t = np.zeros((no - 1, no), dtype=np.complex128)
    for j in range(no):
        s = np.zeros((no, 1, 1))
        s[j, 0, 0] = 1.0
        t[:, j] = orc.kk_sloss(s, dw)[:, 0, 0]
    return t


@pytest.mark.parametrize("no", NOS)
def test_kk_optabs_matches_reference_golden(no):
    dw = float(GOLD["kk_dw"])
    ref = GOLD[f"kk_cmplx_No{no}"]
    got = _k1_table(no, dw)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-14 * scale


@pytest.mark.parametrize("no", NOS)
def test_kk_module_real_matches_reference_golden(no):
    dw = float(GOLD["kk_dw"])
    ref = GOLD[f"kk_real_No{no}"]
    nw = 2 * no + 1
    got = np.zeros_like(ref)
    for w in range(nw):
        s = np.zeros((nw, 1, 1))
        s[w, 0, 0] = 1.0
        for io in range(1, no):
            got[io - 1, w] = orc.kk_module_real(io, 1, 1, s, dw)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


def test_kk_against_live_reference_on_random_spectra():
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref/libref.so not built (needs the reference sources)")
    import ctypes as C
    rng = np.random.default_rng(7)
    no, nlf, dw = 23, 3, 0.0123
    s = np.asfortranarray(rng.standard_normal((2 * no + 1, nlf, nlf))
                          + 1j * rng.standard_normal((2 * no + 1, nlf, nlf)))
    pi0 = orc.kk_optabs(s, dw)
    a, b = C.c_double(), C.c_double()
    for io in (1, 2, 7, no - 2, no - 1):
        for iG, jG in ((1, 1), (2, 3), (3, 1)):
            ref.ref_kk_cmplx(io, no, nlf, iG, jG, (io - 1) * dw, dw, s, C.byref(a), C.byref(b))
            z = pi0[io - 1, iG - 1, jG - 1]
            assert abs(z.real - a.value) <= 1e-13 * max(1.0, abs(a.value))
            assert abs(z.imag - b.value) <= 1e-13 * max(1.0, abs(b.value))


@pytest.mark.parametrize("no", NOS)
def test_kk_sloss_static_and_second_rows_from_reference(no):
    """K2 rows io = 1 and io = 2 equal the module's real variant at the same
    rows, read one frequency index higher (the module indexes S0 on -No:No)."""
    dw = float(GOLD["kk_dw"])
    ref = GOLD[f"kk_real_No{no}"]          # (io, w, re/im) for unit S0m(w - no)
    k2 = _k2_table(no, dw)
    for io in (1, 2):
        for j in range(1, no + 1):          # Sloss S0(j), frequency (j-1) dw
            want = ref[io - 1, no + j, 0] if no + j < 2 * no + 1 else 0.0
            got = k2[io - 1, j - 1].real
            assert abs(got - want) <= 1e-14 * max(1.0, abs(want)), (io, j)


@pytest.mark.parametrize("no", NOS)
def test_kk_sloss_interior_rows(no):
    """For io >= 3: Re chi0 coefficient on S0(j) is the uniform-grid principal
    value 1/(io-j) (0 at j = io) minus the K1 'plus' weight, which the
    reference golden table holds at frequency -(j-1); Im chi0 = -pi S0(io)."""
    dw = float(GOLD["kk_dw"])
    ref = GOLD[f"kk_cmplx_No{no}"]
    k2 = _k2_table(no, dw)
    for io in range(3, no):
        for j in range(1, no + 1):
            plus = ref[io - 1, no - (j - 1), 0, 0]
            if j == 1:
                plus = 0.5 / (io - 1)        # S0(0) carries both halves in K1
            minus = 0.0 if j == io else 1.0 / (io - j)
            assert abs(k2[io - 1, j - 1].real - (minus - plus)) <= 1e-13, (io, j)
            want_im = -np.pi if j == io else 0.0
            assert k2[io - 1, j - 1].imag == pytest.approx(want_im, abs=1e-15)


def test_kk_k1_plus_weight_at_zero_frequency_is_half():
    """K1 puts c+_1 = c-_1 = dw/(2 w_i) on S0(0) for io >= 3 (half weights), so
    the golden coefficient on S0(0) is 1/(io-1) in grid units."""
    no = 17
    ref = GOLD[f"kk_cmplx_No{no}"]
    for io in range(3, no - 1):
        assert ref[io - 1, no, 0, 0] == pytest.approx(1.0 / (io - 1), rel=1e-14)


def test_w1w2_consistent_with_k2():
    rng = np.random.default_rng(3)
    no, nlf, dw = 19, 2, 0.02
    s = np.asfortranarray(rng.standard_normal((no, nlf, nlf)))
    chi0 = orc.kk_sloss(s, dw)
    for io in range(1, no):
        w1, w2 = orc.w1w2(io, 2, 1, s, dw)
        if io == 1:
            assert w1 == pytest.approx(0.5 * chi0[0, 1, 0].real, rel=1e-13)
            assert w2 == -w1
        else:
            assert w1 - w2 == pytest.approx(chi0[io - 1, 1, 0].real, rel=1e-12, abs=1e-12)


def test_kk_continuum_limit_of_a_box_spectrum():
    """Loose check of the discretisation: S0 = 1 on [a, b] gives
    Re chi0(w) -> -(ln|(w-b)/(w-a)| + ln((w+b)/(w+a))) as dw -> 0 (K2 form
    Re chi0 = P int S0(w')[1/(w-w') - 1/(w+w')] dw')."""
    no = 2001
    w = np.arange(no) * (4.0 / (no - 1))
    a, b = 1.0, 2.0
    s = ((w >= a) & (w <= b)).astype(float)
    dw = w[1] - w[0]
    chi0 = orc.kk_sloss(s.reshape(no, 1, 1), dw)[:, 0, 0]
    for x in (0.5, 2.6, 3.1):
        io = int(round(x / dw)) + 1
        wi = (io - 1) * dw
        exact = np.log(abs((wi - a) / (wi - b))) - np.log((wi + b) / (wi + a))
        assert chi0[io - 1].real == pytest.approx(exact, abs=5e-3)


# ---------------------------------------------------------------------------
# gjel, gk_sort, genGlf

@pytest.mark.parametrize("n", (1, 2, 3, 5, 11))
def test_gjel_matches_reference_golden_and_lapack(n):
    a = GOLD[f"gjel_in_{n}"]
    got = orc.gjel(a)
    ref = GOLD[f"gjel_out_{n}"]
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    assert np.abs(got - np.linalg.inv(a)).max() <= 1e-12 * np.abs(ref).max()


def test_gk_sort_matches_reference_golden():
    ks, g, gcutw, igk = GOLD["igk_k"], GOLD["igk_g"], float(GOLD["igk_gcutw"]), GOLD["igk_out"]
    for ik in range(len(ks)):
        rc, lst = orc.gk_sort(ks[ik], g, gcutw, igk.shape[0])
        assert rc == 0
        want = igk[:, ik]
        want = want[want > 0]
        np.testing.assert_array_equal(lst, want)


def test_gk_sort_brute_force_rule():
    """|k+G|^2 <= gcutw, stable by |k+G|^2 with index tie-break (eps 1e-8)."""
    ks, g, gcutw = GOLD["igk_k"], GOLD["igk_g"], float(GOLD["igk_gcutw"])
    for ik in range(len(ks)):
        q2 = ((ks[ik] + g) ** 2).sum(1)
        sel = np.nonzero(q2 <= gcutw)[0]
        # hpsort_eps compares within 1e-8: bucket equal keys, keep index order
        key = np.round(q2[sel] / 1e-8).astype(np.int64)
        order = sel[np.lexsort((sel, key))] + 1
        rc, lst = orc.gk_sort(ks[ik], g, gcutw, 400)
        np.testing.assert_array_equal(lst, order)


@pytest.mark.parametrize("name", ["tiny", "cfg0"])
@pytest.mark.parametrize("lf", ["z", "xyz"])
def test_gen_glf_matches_reference_golden(name, lf):
    s = synth.make_system(synth.CONFIGS[name], with_coeffs=False)
    g = synth.gcart(s.mill_g, s.kc)
    ecut = float(GOLD[f"glf_{name}_{lf}_ecut"])
    rc, n, glf, parg, idx = orc.gen_glf(lf, ecut, s.gcar, g, s.mill_g, 4000)
    assert rc == 0
    assert n == int(GOLD[f"glf_{name}_{lf}_n"])
    np.testing.assert_array_equal(glf, GOLD[f"glf_{name}_{lf}_glf"])
    np.testing.assert_array_equal(parg, GOLD[f"glf_{name}_{lf}_parg"])
    np.testing.assert_array_equal(g[idx - 1], glf)


# ---------------------------------------------------------------------------
# k / k+q mapping

def test_find_kq_brute_force():
    s = H.system("tiny")
    g = synth.gcart(s.mill_g, s.kc)
    for iq in (1, 2, 3):
        q = orc.find_min_q(s.kcart, iq)
        K2, iG0 = orc.find_kq_all(s.kcart, s.kcart, g, q, 1e-4)
        for ik in range(len(s.kcart)):
            kq = s.kcart[ik] + q
            # k+q = G0 + k2 with G0 the first G (in list order) that works
            np.testing.assert_allclose(kq, g[iG0[ik] - 1] + s.kcart[K2[ik] - 1], atol=1e-4)
            d = np.abs(kq[None, None, :] - g[:, None, :] - s.kcart[None, :, :]).max(2)
            hits = np.argwhere(d < 1e-4)
            assert len(hits) >= 1
            assert hits[0, 0] + 1 == iG0[ik]


# ---------------------------------------------------------------------------
# vertices

def _vertices_brute(pb, kind, gcar, n_lo, n_hi, m_lo, m_hi):
    s = pb.sys
    mlf = s.mill_g[pb.lf_idx - 1]
    glf = synth.gcart(mlf, s.kc)
    gpos = {tuple(m): i for i, m in enumerate(s.mill_g)}
    nn, nm = n_hi - n_lo, m_hi - m_lo
    out = np.zeros((len(pb.K1) * nn * nm, pb.nlf), dtype=np.complex128)
    for ik in range(len(pb.K1)):
        K1, K2 = pb.K1[ik] - 1, pb.K2[ik] - 1
        m0 = s.mill_g[pb.iG0[ik] - 1]
        if pb.map_mode == 0:
            l1 = s.mill_g[:pb.ngd]
            look = gpos
            n2 = s.ngk[K2]
        else:
            l1 = s.mill_k[K1, :s.ngk[K1]]
            look = {tuple(m): i for i, m in enumerate(s.mill_k[K2, :s.ngk[K2]])}
            n2 = s.ngk[K2]
        g1 = synth.gcart(l1, s.kc)
        for iG in range(pb.nlf):
            idx = np.array([look.get(tuple(m), -1) for m in (l1 + mlf[iG] + m0)])
            ok = (idx >= 0) & (idx < n2)
            c1 = s.evc[K1, n_lo:n_hi, :len(l1)][:, ok]
            c2 = s.evc[K2, m_lo:m_hi][:, idx[ok]]
            if kind == 0:
                wgt = np.ones(ok.sum())
            else:
                p = kind - 1
                wgt = 0.5 * gcar * (pb.q[p] + 2 * pb.ktot[ik, p] + glf[iG, p] + 2 * g1[ok, p])
            v = np.conj(c1) @ (c2 * wgt).T                  # (nn, nm)
            out[ik * nn * nm:(ik + 1) * nn * nm, iG] = v.ravel()
    return out


@pytest.mark.parametrize("map_mode", [0, 1])
@pytest.mark.parametrize("kind", [0, 1, 3])
def test_vertices_brute_force_tiny(map_mode, kind):
    s = H.system("tiny")
    pb = H.setup(s, "sloss" if kind == 0 else "optabs", map_mode=map_mode)
    nv, nb = s.cfg.nval, s.cfg.nband
    got = orc.vertices(pb, kind, s.gcar, 0, nv, 0, nb)
    want = _vertices_brute(pb, kind, s.gcar, 0, nv, 0, nb)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_vertices_q0_orthonormality_and_current_closed_form():
    """q = 0, G = 0, whole sphere: rho_nm = delta_nm and
    j_nn = Gcar sum_i |C_n(i)|^2 (k + G_i)_pol."""
    s = H.system("tiny")
    pb = H.setup(s, "optabs", map_mode=1, iq=1)
    assert np.all(pb.K2 == pb.K1) and np.allclose(s.mill_g[pb.iG0 - 1], 0)
    g0 = int(np.nonzero((s.mill_g[pb.lf_idx - 1] == 0).all(1))[0][0])
    nb = s.cfg.nband
    rho = orc.vertices(pb, 0, s.gcar, 0, nb, 0, nb)[:, g0].reshape(len(pb.K1), nb, nb)
    eye = np.broadcast_to(np.eye(nb), rho.shape)
    assert np.abs(rho - eye).max() < 1e-13
    for kind in (1, 2, 3):
        j = orc.vertices(pb, kind, s.gcar, 0, nb, 0, nb)[:, g0].reshape(len(pb.K1), nb, nb)
        p = kind - 1
        for ik in range(len(pb.K1)):
            n1 = s.ngk[ik]
            kg = s.kcart[ik, p] + synth.gcart(s.mill_k[ik, :n1], s.kc)[:, p]
            want = s.gcar * (np.abs(s.evc[ik, :, :n1]) ** 2 @ kg)
            np.testing.assert_allclose(np.diag(j[ik]).real, want, rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------------------
# S0 accumulation

def _sloss_s0_brute(pb, nocc, no, dw, eta, vcell):
    s = pb.sys
    nb = s.cfg.nband
    rho = orc.vertices(pb, 0, s.gcar, 0, nocc, nocc, nb)
    nt = len(pb.K1)
    S0 = np.zeros((no, pb.nlf, pb.nlf))
    o = np.arange(no) * dw
    t = 0
    for ik in range(nt):
        for n in range(nocc):
            for m in range(nocc, nb):
                dE = o + pb.E[pb.K1[ik] - 1, n] - pb.E[pb.K2[ik] - 1, m]
                lor = -eta / (dE ** 2 + eta ** 2)
                lor[np.abs(lor) < 1e-3 / eta] = 0.0
                r = rho[t]
                S0 -= (2 * lor)[:, None, None] * np.real(np.outer(r, np.conj(r)))[None] / (np.pi * nt * vcell)
                t += 1
    return S0


def test_sloss_s0_brute_force_and_symmetry():
    s = H.system("tiny")
    pb = H.setup(s, "sloss")
    dw = H.sloss_grid(s.cfg)
    eta = s.cfg.eta_ev / H.HARTREE
    S0 = orc.sloss_s0(pb, s.cfg.nval, s.cfg.no, dw, eta, s.vcell, nthreads=2)
    want = _sloss_s0_brute(pb, s.cfg.nval, s.cfg.no, dw, eta, s.vcell)
    assert np.abs(S0 - want).max() <= 1e-12 * np.abs(want).max()
    assert np.allclose(S0, S0.transpose(0, 2, 1), rtol=0, atol=1e-14 * np.abs(S0).max())
    assert (np.einsum("wii->wi", S0) >= 0).all()


def test_sloss_s0_thread_count_only_reorders_sums():
    s = H.system("tiny")
    pb = H.setup(s, "sloss")
    dw = H.sloss_grid(s.cfg)
    eta = s.cfg.eta_ev / H.HARTREE
    a = orc.sloss_s0(pb, s.cfg.nval, s.cfg.no, dw, eta, s.vcell, nthreads=1)
    b = orc.sloss_s0(pb, s.cfg.nval, s.cfg.no, dw, eta, s.vcell, nthreads=3)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()


def test_optabs_s0_brute_force_hermitian_and_qeff():
    s = H.system("tiny")
    pb = H.setup(s, "optabs")
    cfg = s.cfg
    dw = H.optabs_grid(cfg)
    eta = cfg.eta_ev / H.HARTREE
    T = 0.05 / H.HARTREE
    ef = pb.E[:, cfg.nval - 1].max() - 0.1 / H.HARTREE  # Drude terms near the valence top
    lor_cut, df_cut = 1e-3, 1e-3
    nb, no = cfg.nband, cfg.no
    S0, Q = orc.optabs_s0(pb, cfg.nval, no, 1, dw, eta, lor_cut, df_cut, ef, T, s.vcell, s.gcar,
                          nthreads=2)
    j = orc.vertices(pb, 1, s.gcar, 0, nb, 0, nb)
    nt = len(pb.K1)
    o = np.arange(-no, no + 1) * dw
    want = np.zeros_like(S0)
    wq = np.zeros_like(Q)

    def occ(e):  # Fermi function, exactly 1 / 0 beyond |x| = 20
        x = (e - ef) / T
        if x < -20:
            return 1.0, 0.0
        if x > 20:
            return 0.0, x
        return 1.0 / (np.exp(x) + 1.0), np.exp(x)

    t = 0
    for ik in range(nt):
        for n in range(nb):
            for m in range(nb):
                e1 = pb.E[pb.K1[ik] - 1, n]
                e2 = pb.E[pb.K2[ik] - 1, m]
                f1, x1 = occ(e1)
                f2, _ = occ(e2)
                v = j[t]
                t += 1
                if n == m:
                    if abs(e1 - ef) <= 10 * T:
                        fact = -x1 / ((x1 + 1) ** 2 * T)
                        wq += 2 * fact * np.outer(v, np.conj(v)) / (nt * s.vcell)
                    continue
                if abs(f1 - f2) < df_cut:
                    continue
                dE = o + e1 - e2
                lor = eta / (dE ** 2 + eta ** 2)
                lor[lor < lor_cut / eta] = 0.0
                want -= (2 * (f1 - f2) * lor)[:, None, None] * np.outer(v, np.conj(v))[None] / (np.pi * nt * s.vcell)
    scale = np.abs(want).max()
    assert np.abs(S0 - want).max() <= 1e-12 * scale
    assert np.abs(S0 - np.conj(S0.transpose(0, 2, 1))).max() <= 1e-14 * scale
    assert np.abs(wq).max() > 0
    assert np.abs(Q - wq).max() <= 1e-12 * np.abs(wq).max()


# ---------------------------------------------------------------------------
# Coulomb, Dyson, Pi

def test_genV_is_the_slab_averaged_coulomb_integral():
    """V(Gz, Gz') = (1/c) int int (2 pi/Q) e^{-Q|z-z'|} e^{-i Gz z + i Gz' z'}
    over the slab |z|, |z'| <= c/2 (lf = z local fields, Q = |q|)."""
    from scipy import integrate
    c0, a0 = 20.0, 4.65
    gcar = 2 * np.pi / a0
    ns = [0, 1, -1, 2]
    glf = np.array([[0.0, 0.0, n * a0 / c0] for n in ns])
    parg = np.array([1 if n % 2 == 0 else -1 for n in ns], dtype=np.int32)
    qx, qy = 0.04, 0.03
    V = orc.genV(1e-4, qx, qy, c0, parg, glf, gcar)
    Q = np.hypot(qx, qy)
    for i in range(len(ns)):
        for j in range(len(ns)):
            gi, gj = glf[i, 2] * gcar, glf[j, 2] * gcar
            f = lambda z, zp: (2 * np.pi / Q) * np.exp(-Q * abs(z - zp)) * np.cos(-gi * z + gj * zp)
            r, _ = integrate.dblquad(f, -c0 / 2, c0 / 2, -c0 / 2, c0 / 2, epsabs=1e-9, epsrel=1e-9)
            assert V[i, j] == pytest.approx(r / c0, rel=2e-7, abs=1e-7)


def test_genV_zero_between_different_in_plane_fields():
    c0, a0 = 20.0, 4.65
    gcar = 2 * np.pi / a0
    glf = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 0.0, 0.2325], [0.0, 0.0, 0.2325]])
    parg = np.array([1, 1, -1, -1], dtype=np.int32)
    V = orc.genV(1e-4, 0.01, 0.0, c0, parg, glf, gcar)
    for i in range(4):
        for j in range(4):
            if glf[i, 0] != glf[j, 0]:
                assert V[i, j] == 0.0
    # the in-plane field G = (1,0) sees Q = |q + G Gcar|
    Q = (0.01 + gcar)
    want = 4 * np.pi / Q ** 2 * (1 - (1 - np.exp(-Q * c0)) / (Q * c0))
    assert V[1, 1] == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("nlf", [1, 4, 9])
def test_dyson_w_identities_and_lapack(nlf):
    rng = np.random.default_rng(nlf)
    chi0 = 0.05 * (rng.standard_normal((nlf, nlf)) + 1j * rng.standard_normal((nlf, nlf)))
    V = rng.standard_normal((nlf, nlf))
    V = V + V.T
    e, c, w = orc.dyson_w(chi0, V)
    eps = np.eye(nlf) - chi0 @ V
    ref_inv = np.linalg.inv(eps)
    assert np.abs(e - ref_inv).max() <= 1e-12 * np.abs(ref_inv).max()
    assert np.abs(e - (np.eye(nlf) + c @ V)).max() <= 1e-12 * np.abs(e).max()
    assert np.abs(c - ref_inv @ chi0).max() <= 1e-12 * np.abs(c).max()
    assert np.abs(w - (V + V @ c @ V)).max() <= 1e-12 * np.abs(w).max()
    if nlf == 1:
        assert c[0, 0] == pytest.approx(chi0[0, 0] / (1 - V[0, 0] * chi0[0, 0]), rel=1e-14)


def test_pi_assembly_single_field_closed_form():
    rng = np.random.default_rng(11)
    no, dw, eta_i, c0 = 12, 0.01, 0.004, 20.0
    pi0 = (rng.standard_normal(no - 1) + 1j * rng.standard_normal(no - 1)).reshape(no - 1, 1, 1)
    qeff = np.array([[0.3 - 0.1j]])
    pi11, si, sa = orc.pi_assembly(pi0, qeff, dw, eta_i, c0)
    w = np.arange(no - 1) * dw
    intra = qeff[0, 0] * w / (w + 1j * eta_i)
    inter = pi0[:, 0, 0] - pi0[0, 0, 0].real
    np.testing.assert_allclose(pi11, inter + intra, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(si[1:], (-1j * c0 * inter[1:] / w[1:]).real, rtol=1e-13)
    np.testing.assert_allclose(sa[1:], (-1j * c0 * intra[1:] / w[1:]).real, rtol=1e-13)
    assert si[0] == 0.0 and sa[0] == 0.0


# ---------------------------------------------------------------------------
# findMinQ, genFBZ and the Pi assembly at N_G >= 2: pinned by properties of
# the k-grid, of the point group and of the reference's statement order.

def _mp_grid(n, kc):
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    kcrys = np.stack([ii.ravel() / n, jj.ravel() / n, np.zeros(n * n)], axis=1)
    return np.round(H.synth.gcart(kcrys, kc), 6)


@pytest.mark.parametrize("n", [4, 12, 48, 64])
def test_find_min_q_is_the_shortest_grid_step(n):
    """Optabs.f90:743-773: q = (iq-1) k*, k* the nonzero FBZ vector of least
    norm (below 1 in 2pi/a0 units), the first in list order among equals.  On
    an N x N grid of the hexagonal cell the shortest steps are b1/N, b2/N and
    (b1+b2)/N, all of norm |b|/N."""
    kc = H.synth.reciprocal_axes(4.65, 20.0)
    k = _mp_grid(n, kc)
    norms = np.sqrt(np.einsum("ij,ij->i", k, k))
    b = np.linalg.norm(kc[:, 0])
    steps = np.round(np.stack([kc[:, 0], kc[:, 1], kc[:, 0] + kc[:, 1]]) / n, 6)
    for iq in (1, 2, 5):
        q = orc.find_min_q(k, iq)
        if iq == 1:
            assert not q.any()
            continue
        ks = q / (iq - 1)
        assert abs(np.linalg.norm(ks) - b / n) < 2e-6
        assert any(np.allclose(ks, s_, rtol=0, atol=1e-12) for s_ in steps)
        nz = norms[norms > 0]
        assert np.linalg.norm(ks) <= nz.min() + 1e-15           # no shorter vector exists
        first = int(np.flatnonzero(norms == np.linalg.norm(ks))[0])
        np.testing.assert_array_equal(ks, k[first])              # the first of equal norm


def test_find_min_q_takes_the_first_of_equal_vectors_in_list_order():
    k = np.array([[0, 0, 0], [0.3, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [0.5, 0.5, 0]], dtype=np.float64)
    np.testing.assert_array_equal(orc.find_min_q(k, 3), [0.2, 0, 0])
    np.testing.assert_array_equal(orc.find_min_q(k[[0, 1, 3, 2, 4]], 3), [0, 0.2, 0])
    # vectors of norm 1 or more never qualify (kmin starts at 1.0)
    k2 = np.array([[0, 0, 0], [1.0, 0, 0], [0.9, 0, 0]])
    np.testing.assert_array_equal(orc.find_min_q(k2, 2), [0.9, 0, 0])


def _gen_fbz(ki, Rs, eps=1e-4):
    nki = len(ki)
    R = np.zeros((48, 3, 3), order="F")
    for i, r in enumerate(Rs):
        R[i] = r
    kt = np.zeros((3, 48 * nki), order="F")
    nt = C.c_int()
    kif = np.asfortranarray(np.asarray(ki, dtype=np.float64).T)
    rc = orc.lib().orc_gen_fbz(nki, len(Rs), eps, kif, R, C.byref(nt), kt)
    assert rc == 0
    return kt[:, :nt.value].T.copy()


def test_gen_fbz_identity_keeps_the_list_and_drops_repeats():
    """Optabs.f90:871-933 with Nsym = 1: the IBZ list itself, in order; a
    point within eps of an earlier one appears once."""
    kc = H.synth.reciprocal_axes(4.65, 20.0)
    k = _mp_grid(6, kc)
    np.testing.assert_array_equal(_gen_fbz(k, [np.eye(3)]), k)
    dup = np.vstack([k[:5], k[2:3] + 5e-5, k[5:]])
    np.testing.assert_array_equal(_gen_fbz(dup, [np.eye(3)]), k)


def test_gen_fbz_mirror_unfolds_half_a_grid_to_the_whole_grid():
    """The swap of crystal coordinates (b1 <-> b2) is a mirror of the
    hexagonal lattice.  From the half grid i <= j it must give every point of
    the N x N grid exactly once: the half first, in order, then the mirror
    images of its off-diagonal points in the same order."""
    n = 8
    kc = H.synth.reciprocal_axes(4.65, 20.0)
    P = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1]], dtype=np.float64)
    M = kc @ P @ np.linalg.inv(kc)
    pairs = [(i, j) for i in range(n) for j in range(n) if i <= j]
    half = np.array([kc @ np.array([i / n, j / n, 0.0]) for i, j in pairs])
    fbz = _gen_fbz(half, [np.eye(3), M])
    full = np.array([kc @ np.array([i / n, j / n, 0.0]) for i in range(n) for j in range(n)])
    assert len(fbz) == n * n
    d = np.abs(fbz[:, None, :] - full[None, :, :]).max(axis=2)
    assert (d.min(axis=1) < 1e-12).all() and (d.min(axis=0) < 1e-12).all()
    np.testing.assert_array_equal(fbz[:len(half)], half)
    off = [p for p, (i, j) in enumerate(pairs) if i != j]
    np.testing.assert_allclose(fbz[len(half):], half[off] @ M.T, rtol=0, atol=1e-15)


def test_gen_fbz_orbit_sizes_of_the_sixfold_rotations():
    """Six rotations about z: a generic k has six distinct images, Gamma one."""
    Rs = []
    for t in range(6):
        c, s_ = math.cos(t * math.pi / 3), math.sin(t * math.pi / 3)
        Rs.append(np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1.0]]))
    assert len(_gen_fbz(np.array([[0.13, 0.05, 0.0]]), Rs)) == 6
    assert len(_gen_fbz(np.array([[0.0, 0.0, 0.0]]), Rs)) == 1
    assert len(_gen_fbz(np.array([[0.0, 0.0, 0.0], [0.2, 0.0, 0.0]]), Rs)) == 7


def test_pi_assembly_two_fields_reads_pi11_after_each_update():
    """Optabs.f90:575-603 at Nlf = 2: Pi_inter = Pi_tot(1,1) is read after
    every (iG, jG) update, and Pi_tot(1,1) gets its intraband term at the
    first update, so fort.401 holds Re(-i c0 [Pi0 - Re Pi0(0) + Qeff
    omega/(omega + i eta)]_11 / omega), intraband included (at Nlf = 1 it is
    read before that term).  A pure-Python walk of the reference's loop
    order gives the same."""
    rng = np.random.default_rng(12)
    no, dw, eta_i, c0 = 10, 0.02, 0.005, 15.0
    pi0 = rng.standard_normal((no - 1, 2, 2)) + 1j * rng.standard_normal((no - 1, 2, 2))
    qeff = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    pi11, si, sa = orc.pi_assembly(pi0, qeff, dw, eta_i, c0)
    w = np.arange(no - 1) * dw
    intra = qeff[0, 0] * w / (w + 1j * eta_i)
    tot11 = pi0[:, 0, 0] - pi0[0, 0, 0].real + intra
    np.testing.assert_allclose(pi11, tot11, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(si[1:], c0 * tot11[1:].imag / w[1:], rtol=1e-13)
    np.testing.assert_allclose(sa[1:], c0 * intra[1:].imag / w[1:], rtol=1e-13)
    # the single-field assembly of the same (1,1) data leaves intra out of si
    _, s1, _ = orc.pi_assembly(pi0[:, :1, :1].copy(), qeff[:1, :1].copy(), dw, eta_i, c0)
    np.testing.assert_allclose(s1[1:], c0 * (tot11 - intra)[1:].imag / w[1:], rtol=1e-13)
    # the reference's loop nest, statement by statement
    dia = np.zeros((2, 2), dtype=complex)
    for io in range(1, no):
        oi = (io - 1) * dw
        tot = np.zeros((2, 2), dtype=complex)
        for iG in range(2):
            for jG in range(2):
                if io == 1:
                    dia[iG, jG] = -complex(pi0[0, iG, jG].real, 0.0)
                tot[iG, jG] = pi0[io - 1, iG, jG] + dia[iG, jG]
                p_inter = tot[0, 0]
                tot[iG, jG] = tot[iG, jG] + qeff[iG, jG] * oi / (oi + 1j * eta_i)
        if io > 1:
            assert si[io - 1] == pytest.approx((-1j * c0 * p_inter / oi).real, rel=1e-13)


# ---------------------------------------------------------------------------
# Point-group unfolding (Nsym > 1): the oracle's rotated vertex map and k
# bookkeeping against the same system written out on the whole unfolded zone
# with rotated coefficients, C_Rk(R G) = C_k(G), and E(Rk) = E(k), run with
# Nsym = 1.  The two must agree to rounding: a wrong rotation, inverse,
# operation choice or index map breaks the equality.

def _sym_setup(kind, nk, nlf_target=9, map_mode=0):
    import dataclasses
    cfg = dataclasses.replace(H.synth.CONFIGS["tiny"], nk=nk, lf="xyz", nlf_target=nlf_target)
    ibz, fz, R, nsym, kinfo, ngd = H.synth.make_symmetric_pair(cfg, kind, 90, 60)
    g = H.synth.gcart(ibz.mill_g, ibz.kc)
    RI = orc.point_inverse(R, nsym)
    # genFBZ with the operations
    kif = np.asfortranarray(ibz.kcart.T)
    Rf = np.asfortranarray(R)
    kt = np.zeros((3, 48 * len(ibz.kcart)), order="F")
    nt = C.c_int()
    assert orc.lib().orc_gen_fbz(len(ibz.kcart), nsym, 1e-4, kif, Rf, C.byref(nt), kt) == 0
    ktot = kt[:, :nt.value].T.copy()
    np.testing.assert_allclose(ktot, fz.kcart, rtol=0, atol=1e-12)
    R1 = np.zeros(len(ktot), np.int32)
    K1 = np.zeros(len(ktot), np.int32)
    for ik in range(len(ktot)):
        R1[ik], K1[ik] = orc.find_kin_ibz(ik + 1, ibz.kcart, ktot[ik], nsym, RI, 1e-4)
    assert [(r - 1, k - 1) for r, k in zip(R1, K1)] == kinfo
    q = orc.find_min_q(ktot, 2)
    iG0, R2, K2 = orc.find_kq_sym_all(ktot, ibz.kcart, g, q, nsym, RI, 1e-4)
    ecut = H.synth.local_field_ecut(ibz)
    rc, nlf, glf, parg, idx = orc.gen_glf("xyz", ecut, ibz.gcar, g, ibz.mill_g, 100)
    pb_s = orc.Problem(ibz, K1, K2, iG0, ktot, q, idx, map_mode=map_mode, ngd=ngd,
                       E_hartree=H.energies(ibz, "sloss"), R=R, RI=RI, R1=R1, R2=R2)
    K2f, iG0f = orc.find_kq_all(ktot, ktot, g, q, 1e-4)
    pb_f = orc.Problem(fz, np.arange(1, len(ktot) + 1, dtype=np.int32), K2f, iG0f, ktot, q, idx,
                       map_mode=map_mode, ngd=ngd, E_hartree=H.energies(fz, "sloss"))
    # the unfolded k+q point is the same FBZ point in both descriptions
    assert np.array_equal(iG0, iG0f)
    return cfg, ibz, fz, pb_s, pb_f, nsym


@pytest.mark.parametrize("kind,nk,map_mode", [("c2", 4, 0), ("c3", 6, 0), ("c6", 6, 0), ("c3v", 6, 0),
                                               ("c3v", 6, 1)])
def test_unfolded_vertices_and_s0_equal_the_whole_zone(kind, nk, map_mode):
    cfg, ibz, fz, pb_s, pb_f, nsym = _sym_setup(kind, nk, map_mode=map_mode)
    nv, nb = cfg.nval, cfg.nband
    assert nsym > 1 and (pb_s.R1 != 1).any() and (pb_s.R2 != 1).any()
    v_s = orc.vertices(pb_s, 0, ibz.gcar, 0, nv, nv, nb)
    v_f = orc.vertices(pb_f, 0, fz.gcar, 0, nv, nv, nb)
    assert np.abs(v_s).max() > 0.1
    assert H.rel_max(v_s, v_f) < 1e-12
    dw = H.sloss_grid(cfg)
    S_s = orc.sloss_s0(pb_s, nv, cfg.no, dw, cfg.eta_ev / H.HARTREE, ibz.vcell)
    S_f = orc.sloss_s0(pb_f, nv, cfg.no, dw, cfg.eta_ev / H.HARTREE, fz.vcell)
    assert H.rel_max(S_s, S_f) < 1e-12
    # current vertices: J carries R(R1) G'' of the unfolded plane wave
    for pol in (1, 2, 4):
        a1, a2 = orc.vertices(pb_s, pol, ibz.gcar, 0, nv, nv, nb, both=True)
        b1, b2 = orc.vertices(pb_f, pol, fz.gcar, 0, nv, nv, nb, both=True)
        assert H.rel_max(a1, b1) < 1e-12 and H.rel_max(a2, b2) < 1e-12


def test_point_inverse_matches_lapack():
    R, nsym = H.synth.hex_point_ops("c3v")
    RI = orc.point_inverse(R, nsym)
    for i in range(nsym):
        np.testing.assert_allclose(RI[i], np.linalg.inv(R[i]), rtol=0, atol=1e-15)


# ---------------------------------------------------------------------------
# interpolate.x (Chi_RPA/interpolate.f90:103-186)

def _dense_by_np_interp(pi0, qeff, no_interp, dwt, eta):
    """The dense grid written another way: Pi_tot = Pi0 - Re Pi0(1) on the
    coarse nodes x = 0..No-2, np.interp at x = (io-1) + jo/No_interp, plus
    Qeff oi/(oi + i eta) with oi = counter * domega_tot."""
    no = pi0.shape[0] + 1
    tot = pi0 - pi0[0].real[None]
    x = np.arange(no - 1, dtype=np.float64)
    t = np.array([(io - 1) + jo / no_interp for io in range(1, no - 1) for jo in range(1, no_interp)])
    oi = np.arange(1, len(t) + 1) * dwt
    nlf = pi0.shape[1]
    out = np.zeros((len(t), nlf, nlf), dtype=np.complex128)
    for i in range(nlf):
        for j in range(nlf):
            out[:, i, j] = (np.interp(t, x, tot[:, i, j].real) + 1j * np.interp(t, x, tot[:, i, j].imag)
                            + qeff[i, j] * oi / (oi + 1j * eta))
    return out, qeff[0, 0] * oi / (oi + 1j * eta)


@pytest.mark.parametrize("no,nlf,no_interp", [(9, 3, 10), (17, 2, 2), (6, 1, 5), (40, 4, 7)])
def test_interpolate_is_linear_interpolation_of_the_kk_response(no, nlf, no_interp):
    """orc_interpolate against np.interp on Pi0 - Re Pi0(io=1) plus the
    intraband closed form; a divisor No_interp - 1, a counter starting at 0,
    Pi_dia taken at another io or Qeff(1,1) on every pair fails it."""
    rng = np.random.default_rng(no * 100 + nlf)
    dw, dwt, eta = 0.013, 0.0021, 0.004
    s0 = (rng.standard_normal((2 * no + 1, nlf, nlf)) + 1j * rng.standard_normal((2 * no + 1, nlf, nlf)))
    qeff = rng.standard_normal((nlf, nlf)) + 1j * rng.standard_normal((nlf, nlf))
    pi0 = orc.kk_optabs(s0, dw)
    got, intra = orc.interpolate(pi0, qeff, no_interp, dwt, eta)
    want, want_intra = _dense_by_np_interp(pi0, qeff, no_interp, dwt, eta)
    assert got.shape == ((no_interp - 1) * (no - 2), nlf, nlf)
    scale = np.abs(want).max()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(intra, want_intra, rtol=1e-14, atol=0)


def test_interpolate_one_point_per_interval_is_empty_and_the_static_real_part_vanishes():
    """No_interp = 1 gives No_tot = 0 (the reference's jo loop is empty);
    Re Pi_tot(io = 1) = 0 by Pi_dia, so with Qeff = 0 the first No_interp - 1
    dense points have real part Re Pi_tot(2) * jo / No_interp exactly up to
    rounding."""
    rng = np.random.default_rng(5)
    no, nlf = 8, 2
    s0 = rng.standard_normal((2 * no + 1, nlf, nlf)) + 1j * rng.standard_normal((2 * no + 1, nlf, nlf))
    pi0 = orc.kk_optabs(s0, 0.02)
    got, intra = orc.interpolate(pi0, np.zeros((nlf, nlf)), 1, 0.01, 0.003)
    assert got.shape[0] == 0 and intra.shape[0] == 0
    n = 4
    got, _ = orc.interpolate(pi0, np.zeros((nlf, nlf)), n, 0.01, 0.003)
    re2 = (pi0[1] - pi0[0]).real
    for jo in range(1, n):
        np.testing.assert_allclose(got[jo - 1].real, re2 * jo / n, rtol=1e-14, atol=1e-15)


# ---------------------------------------------------------------------------
# Photon.x (BSE/Photonew.f90)

def test_lu_invert_matches_lapack_and_reports_a_zero_pivot():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 33, 66):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        inv, info = orc.lu_invert(a)
        assert info == 0
        np.testing.assert_allclose(inv, np.linalg.inv(a), rtol=0, atol=1e-12 * np.abs(np.linalg.inv(a)).max())
    a = np.ones((3, 3), dtype=np.complex128)
    _, info = orc.lu_invert(a)
    assert info == 2


def _slab_integrals(beta, gi, gj, c0):
    """(1/c0) int int over the slab of e^{-i Gi z} K(z - z') e^{i Gj z'} for
    K = e^{i beta |z - z'|} and K = sign(z - z') e^{i beta |z - z'|}, the inner
    integral split at z' = z."""
    from scipy import integrate
    h = c0 / 2

    def inner(z, sgn):
        def f(zp, part):
            v = np.exp(1j * beta * abs(z - zp)) * np.exp(1j * gj * zp) * (np.sign(z - zp) if sgn else 1.0)
            return v.real if part == 0 else v.imag
        r = sum(integrate.quad(f, a, b, args=(0,), epsabs=1e-13, epsrel=1e-12)[0] for a, b in ((-h, z), (z, h)))
        i = sum(integrate.quad(f, a, b, args=(1,), epsabs=1e-13, epsrel=1e-12)[0] for a, b in ((-h, z), (z, h)))
        return r + 1j * i

    out = []
    for sgn in (False, True):
        def g(z, part):
            v = np.exp(-1j * gi * z) * inner(z, sgn)
            return v.real if part == 0 else v.imag
        r = integrate.quad(g, -h, h, args=(0,), epsabs=1e-12, epsrel=1e-11, limit=200)[0]
        i = integrate.quad(g, -h, h, args=(1,), epsabs=1e-12, epsrel=1e-11, limit=200)[0]
        out.append((r + 1j * i) / c0)
    return out


def test_photon_d0_is_the_slab_average_of_the_free_propagator():
    """genD0's A + diagonal1 and B + diagonal2 (Photonew.f90:632-643) are the
    slab double integrals of e^{i beta|z-z'|} and sign(z-z') e^{i beta|z-z'|}
    between plane waves of Gz = 2 pi m / c0, parity (-1)^m; Dxx0, Dyy0, Dyz0
    and Dzz0 are them times the propagator's prefactors.  A swapped index, a
    wrong parity or sign fails it."""
    c0, o, eta, q = 7.3, 0.35, 0.01, 0.02
    ms = np.array([0, 1, -1, 2])
    gz = 2 * np.pi * ms / c0
    par = np.where(ms % 2 == 0, 1, -1).astype(np.int32)
    D = orc.photon_d0(q, o, eta, c0, par, gz)
    gam = 1 / 137.0
    oi = o + 1j * eta
    beta = np.sqrt(gam ** 2 * oi ** 2 - q ** 2 + 0j)
    a_d1 = D[:, :, 0] * beta / (2j * np.pi * gam ** 2)
    b_d2 = D[:, :, 3] * oi ** 2 / (-2j * np.pi * q)
    for i, j in [(0, 0), (1, 1), (0, 1), (1, 2), (3, 1), (2, 0)]:
        ia, ib = _slab_integrals(beta, gz[i], gz[j], c0)
        assert abs(a_d1[i, j] - ia) < 1e-8 * max(abs(ia), 1e-3), (i, j, a_d1[i, j], ia)
        assert abs(b_d2[i, j] - ib) < 1e-8 * max(abs(ib), 1e-3), (i, j, b_d2[i, j], ib)
    np.testing.assert_allclose(D[:, :, 1], 2j * np.pi * beta * a_d1 / oi ** 2, rtol=1e-13)
    np.testing.assert_allclose(D[:, :, 2], 2j * np.pi * q ** 2 * a_d1 / (beta * oi ** 2)
                               - np.eye(4) * 4 * np.pi / oi ** 2, rtol=1e-13, atol=1e-13)


def _photon_inputs(n, seed, scale=0.2):
    rng = np.random.default_rng(seed)
    c0 = 9.0
    ms = np.array([0] + [(k + 1) // 2 * (1 if k % 2 else -1) for k in range(1, n)])
    gz = 2 * np.pi * ms / c0
    par = np.where(ms % 2 == 0, 1, -1).astype(np.int32)
    pi0 = scale * (rng.standard_normal((n, n, 5)) + 1j * rng.standard_normal((n, n, 5)))
    return c0, gz, par, np.asfortranarray(pi0)


@pytest.mark.parametrize("n", [1, 3, 6])
def test_photon_point_solves_the_dyson_equation_and_sums_the_slab_integral(n):
    """The screened blocks satisfy Pi = Pi0 + Pi0 D0 Pi (s mode with Dxx0, p
    mode with the 2x2 block propagator), the reference's TS/TP inversion
    written as a numpy solve; the sums equal (1/c0) sum Pi(G,G') int_0^{c0/2}
    e^{-iGz} dz int_0^{c0/2} e^{iG'z'} dz' by quadrature; with Pi0 = 0 the
    absorption and reflection vanish."""
    from scipy import integrate
    c0, gz, par, pi0 = _photon_inputs(n, seed=n)
    q, o, eta, theta = 0.03, 0.4, 0.02, 0.0
    pi, sums, spec, st = orc.photon_point(q, o, eta, c0, theta, par, gz, pi0)
    assert st == 0
    D = orc.photon_d0(q, o, eta, c0, par, gz)
    Dxx, Dyy, Dzz, Dyz = (D[:, :, k] for k in range(4))
    xx0, yy0, yz0, zy0, zz0 = (pi0[:, :, k] for k in range(5))
    np.testing.assert_allclose(pi[:, :, 0], xx0 + xx0 @ Dxx @ pi[:, :, 0], rtol=0, atol=1e-12)
    P0 = np.block([[yy0, yz0], [zy0, zz0]])
    Dp = np.block([[Dyy, Dyz], [Dyz, Dzz]])
    Pp = np.block([[pi[:, :, 1], pi[:, :, 2]], [pi[:, :, 3], pi[:, :, 4]]])
    np.testing.assert_allclose(Pp, P0 + P0 @ Dp @ Pp, rtol=0, atol=1e-12)
    h = c0 / 2

    def cquad(f):
        return (integrate.quad(lambda z: f(z).real, 0, h, epsabs=1e-13)[0]
                + 1j * integrate.quad(lambda z: f(z).imag, 0, h, epsabs=1e-13)[0])

    left = np.array([cquad(lambda z, g=g: np.exp(-1j * g * z)) for g in gz])
    right = np.array([cquad(lambda z, g=g: np.exp(1j * g * z)) for g in gz])
    for m, k in zip(range(5), (0, 1, 4, 2, 3)):  # sums: xx, yy, zz, yz, zy
        want = left @ pi[:, :, k] @ right / c0
        assert abs(sums[m] - want) < 1e-11 * max(1.0, abs(want)), (m, sums[m], want)
    _, _, spec0, _ = orc.photon_point(q, o, eta, c0, theta, par, gz, np.zeros_like(pi0))
    assert np.all(spec0 == 0.0)


# ---------------------------------------------------------------------------
# GW propagators (Modules/propagators.f90)

@pytest.mark.parametrize("no", [5, 6, 9, 17, 40])
def test_gamma_propagator_weights_are_the_pinned_k1_stencil(no):
    """genGammaPropagator's W1 and W2 are the K1 sums (the module's chain of
    plain ifs): a real spectrum S with S(1) = 0 placed on the positive
    branch of an Optabs S0(-no:no) gives Re K1[S0] = W1, on the negative
    branch Re K1[S0] = W2 (orc_kk_optabs is pinned by the reference's
    golden K1 tables); Im Gamma+ = -pi S(io) with the module's pi =
    3.141592654, Gamma- real."""
    rng = np.random.default_rng(no)
    dw = 0.037
    S = rng.standard_normal(no)
    S[0] = 0.0
    D = -3.141592654 * S * 1j + rng.standard_normal(no)  # Im D = -pi S
    gp, gm, _, _ = orc.gamma_propagator(D, dw)
    np.testing.assert_allclose(-(1.0 / 3.141592654) * D.imag, S, rtol=1e-15)
    pos = np.zeros((2 * no + 1, 1, 1), dtype=np.complex128)
    neg = np.zeros_like(pos)
    for jo in range(1, no + 1):
        pos[no + jo - 1, 0, 0] = S[jo - 1]
        neg[no - (jo - 1), 0, 0] = S[jo - 1]
    w1 = orc.kk_optabs(pos, dw)[:, 0, 0].real
    w2 = orc.kk_optabs(neg, dw)[:, 0, 0].real
    scale = np.abs(S).max() * no
    np.testing.assert_allclose(gp.real, w1, rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(gm.real, -w2, rtol=0, atol=1e-13 * scale)
    np.testing.assert_array_equal(gp.imag, -3.141592654 * (-(1.0 / 3.141592654) * D.imag[:no - 1]))
    assert np.all(gm.imag == 0.0)


def test_gamma_propagator_reconstructs_a_causal_response():
    """For a retarded D(w) = 1/(w - w0 + i g) - 1/(w + w0 + i g) on a fine
    grid, Re Gamma+ + Re Gamma- = W1 - W2 reconstructs Re D (the module's own
    Kramers-Kronig check: KKS/SKK small), and Re Gamma+(0) = Re D(0)/2 (the
    static row, W2 = -W1); a sign error in either sum breaks both."""
    no, dw, w0, g = 2001, 0.002, 1.0, 0.05
    w = np.arange(no) * dw
    D = 1.0 / (w - w0 + 1j * g) - 1.0 / (w + w0 + 1j * g)
    gp, gm, kks, skk = orc.gamma_propagator(D, dw)
    assert kks / skk < 1e-3
    rec = gp.real + gm.real
    assert np.abs(rec - D.real[:no - 1]).max() < 0.02 * np.abs(D.real).max()
    assert gp[0].real == pytest.approx(D[0].real / 2, rel=1e-2)


# ---------------------------------------------------------------------------
# ladder (BSE) frequency loop (Chi_ladd_up/BSE.f90:1116-1176)

def test_ladder_is_the_summed_neumann_series():
    """lladd = (I - L0 F/NV)^-1 L0 F L0 is the ladder series
    sum_m (L0 F / NV)^m L0 F L0: chi_ladd from the oracle's gjel inversion
    equals the series summed with numpy for a non-symmetric kernel (a
    transposed kernel or L0 on the wrong side fails it); F = 0 gives no
    ladder term and the bubble sum_k |j|^2 (o/Delta) L0 valley / NV."""
    rng = np.random.default_rng(7)
    nk, no, ntot, vcell, valley = 9, 5, 16, 2.5, 2.0
    omin, dw, eta = 0.1, 0.03, 0.02
    nv = ntot * vcell
    F = 0.05 * nv * (rng.standard_normal((nk, nk)) + 1j * rng.standard_normal((nk, nk)))
    delta = rng.uniform(0.5, 1.5, nk)
    cur = rng.standard_normal((2, nk)) + 1j * rng.standard_normal((2, nk))
    cl, cp = orc.ladder(F, delta, cur, no, omin, dw, eta, ntot, vcell, valley)
    for io in range(no):
        o = omin + io * dw
        L0 = np.diag(1.0 / (o + delta + 1j * eta))
        A = L0 @ F / nv
        term = L0 @ F @ L0
        lad = np.zeros_like(term)
        for _ in range(80):
            lad += term
            term = A @ term
        for pol in range(2):
            want = -valley * cur[pol] @ lad @ cur[pol].conj() / nv ** 2
            assert abs(cl[pol, io] - want) < 1e-12 * abs(want)
            bub = valley * np.sum(np.abs(cur[pol]) ** 2 * (o / delta) / (o + delta + 1j * eta)) / nv
            assert abs(cp[pol, io] - bub) < 1e-13 * abs(bub)
    cl0, cp0 = orc.ladder(np.zeros((nk, nk)), delta, cur, no, omin, dw, eta, ntot, vcell, valley)
    assert np.all(cl0 == 0.0)
    np.testing.assert_allclose(cp0, cp, rtol=1e-15)


def test_fock_kernel_contraction_diagonal_and_transfer_rule():
    """Chi_ladd's Fock kernel (Chi_ladd_up/BSE.f90:305,542,672-676,804-816):
    off the diagonal -M1^H W(iq) M2 with a non-symmetric W (numpy: a swapped
    (G, G') order or the conjugate on the wrong vertex fails it); on the
    diagonal -Vcell Ntot ln(1 + alpha epsq)/alpha, with alpha recovered from
    a W(2,1,1) built from a chosen alpha; pairs past nq, or below qmin, 0."""
    rng = np.random.default_rng(9)
    nk, nlf, nq, ntot, c0, vcell = 6, 3, 5, 100, 20.0, 800.0
    gcar, qmin, pi = 2 * np.pi / 4.65, 0.01, 3.141592654
    kbz = np.zeros((3, nk))
    kbz[0] = np.array([0.0, 0.0017, 0.0040, 0.0061, 0.0125, 0.0003])  # 2pi/a0 units
    wt = rng.standard_normal((nq, nlf, nlf)) + 1j * rng.standard_normal((nq, nlf, nlf))
    alpha0 = 3.7
    wt[1, 0, 0] = c0 * (2 * pi / qmin) / (1 + alpha0 * qmin) + 0.4j
    mnm = rng.standard_normal((2, nlf, nk, nk)) + 1j * rng.standard_normal((2, nlf, nk, nk))
    f = orc.fock_kernel(kbz, wt, mnm, nq, ntot, gcar, qmin, c0, vcell)
    epsq = np.sqrt(4 * pi * c0 / (vcell * ntot))
    diag = -vcell * ntot * np.log(1 + alpha0 * epsq) / alpha0
    for k1 in range(nk):
        for k2 in range(nk):
            if k1 == k2:
                assert f[k1, k2] == pytest.approx(diag, rel=1e-12)
                continue
            q = gcar * abs(kbz[0, k2] - kbz[0, k1])
            iq = int(q / qmin) + 1
            if iq < 2 or iq > nq:
                assert f[k1, k2] == 0
            else:
                want = -(mnm[0, :, k1, k2].conj() @ wt[iq - 1] @ mnm[1, :, k1, k2])
                assert abs(f[k1, k2] - want) < 1e-13 * abs(want)
    assert (np.abs(f) > 0).sum() > nk + 4  # both kinds of pairs occur
