"""Host-side runner for in-memory inputs: the reference drivers' setup steps
(Optabs.f90:160-276, Sloss.f90:182-299) done through the library's own host
entry points, then the device chain.  The Fortran drivers do the same from
files; this runner serves the tests, the benchmark and smoke().
"""

from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

from . import _abi
from ._abi import RPA_OPTABS, RPA_SLOSS

HARTREE = 2.0 * 13.6056923  # Optabs.f90:46, Sloss.f90:71


@dataclasses.dataclass
class RunConfig:
    program: str = "sloss"           # 'sloss' | 'optabs'
    vertex_map: str = "direct"       # 'direct' | 'igk'
    pol: str = "xx"
    iq: Optional[int] = None
    dgw_ev: float = 0.0              # Optabs scissor (Sloss hard-codes 1 eV)
    temp_ev: float = 0.0
    efermi_ev: float = 0.0
    lor_cut: float = 1.0e-3
    df_cut: float = 1.0e-3
    eta_intra_ev: float = 0.01
    omin_ev: float = 0.0
    eps: float = 1.0e-4
    ngd: Optional[int] = None
    device: int = 0
    precision: str = "fp64"          # 'fp64' | 'mixed'



def energies(sys_, rc: RunConfig) -> np.ndarray:
    """E(NkI, Nband) in Hartree after the drivers' scissor (loadkIandE)."""
    E = sys_.e_ev / HARTREE
    nv = sys_.cfg.nval
    if rc.program == "sloss":
        E[:, nv:] = E[:, nv:] + 1.0 / HARTREE          # Sloss.f90:1521
    else:
        E[:, nv:] = E[:, nv:] + rc.dgw_ev / HARTREE    # Optabs.f90:1350
    return E


def local_fields_by_qg(g: np.ndarray, mill_g: np.ndarray, gcar: float, q: np.ndarray, nlf: int):
    """The nlf local fields of smallest |q + G| (ties by G index), for a q loop
    whose local-field count follows a |q+G| cutoff.

    g is the density G list (2pi/a units, G(1) = 0), q in the same units.
    Returns (lf_idx 1-based, parG) with the parity rule of genGlfandParity
    (Sloss.f90:850-854, 865-869: +1 for even Miller z, -1 for odd).
    """
    qg = (np.asarray(q, dtype=np.float64)[None, :] + g) * gcar
    d2 = np.einsum("ij,ij->i", qg, qg)
    order = np.lexsort((np.arange(len(g)), np.round(d2, 12)))[:nlf]
    order = np.sort(order)                   # keep the G-list order, as genGlf does
    mz = np.asarray(mill_g)[order, 2]
    parg = np.where((mz // 2) * 2 == mz, 1, -1).astype(np.int32)
    return (order + 1).astype(np.int32), parg


def _window_count(a, b, io_lo: int, io_hi: int, dw: float, eta: float, cut: float) -> int:
    """Bins io in [io_lo, io_hi] with eta / (dE^2 + eta^2) >= cut for
    dE = io dw + a - b, per pair (a, b).  The passing bins are contiguous
    around the resonance; their range comes from |dE| <= sqrt(eta/cut -
    eta^2), and both ends are settled by the exact test, so the count equals
    a bin-by-bin evaluation."""
    hw = math.sqrt(max(eta / cut - eta * eta, 0.0))

    def ok(io):
        d = (io * dw + a) - b
        return np.abs(-eta / (d * d + eta * eta)) >= cut

    c = (b - a) / dw
    lo = np.ceil(c - hw / dw).astype(np.int64)
    hi = np.floor(c + hw / dw).astype(np.int64)
    lo = np.where(ok(lo - 1), lo - 1, np.where(ok(lo), lo, lo + 1))
    hi = np.where(ok(hi + 1), hi + 1, np.where(ok(hi), hi, hi - 1))
    n = np.minimum(hi, io_hi) - np.maximum(lo, io_lo) + 1
    return int(np.maximum(n, 0).sum())


def _pairs(E, K1, K2, n_rng, m_rng):
    e1 = E[K1 - 1][:, n_rng[0]:n_rng[1]]
    e2 = E[K2 - 1][:, m_rng[0]:m_rng[1]]
    shp = (len(K1), e1.shape[1], e2.shape[1])
    return np.broadcast_to(e1[:, :, None], shp).ravel(), np.broadcast_to(e2[:, None, :], shp).ravel()


def nonzero_weights(E, K1, K2, nv: int, nb: int, no: int, dw: float, eta: float) -> int:
    """Number of (transition, frequency) pairs whose Sloss weight passes the
    cutoff |L| >= 1e-3/eta, with L = -eta/(dE^2 + eta^2) and
    dE = (io-1) dw + E(K1,n) - E(K2,m), n <= Nocc < m (Sloss.f90:407-412)."""
    K1 = np.asarray(K1)
    K2 = np.asarray(K2)
    total = 0
    for k0 in range(0, len(K1), 4096):
        a, b = _pairs(E, K1[k0:k0 + 4096], K2[k0:k0 + 4096], (0, nv), (nv, nb))
        total += _window_count(a, b, 0, no - 1, dw, eta, 1.0e-3 / eta)
    return total


def nonzero_weights_optabs(E, K1, K2, nv: int, nb: int, no: int, dw: float, eta: float, lor_cut: float) -> int:
    """The same for the Optabs weight at T = 0: pairs n != m with |f_n - f_m|
    = 1 (valence to conduction and back), L = eta/(dE^2 + eta^2) >=
    Lor_cut/eta, dE = io dw + E(K1,n) - E(K2,m), io = -no..no
    (Optabs.f90:356-402)."""
    K1 = np.asarray(K1)
    K2 = np.asarray(K2)
    total = 0
    for k0 in range(0, len(K1), 2048):
        k1, k2 = K1[k0:k0 + 2048], K2[k0:k0 + 2048]
        for n_rng, m_rng in (((0, nv), (nv, nb)), ((nv, nb), (0, nv))):
            a, b = _pairs(E, k1, k2, n_rng, m_rng)
            total += _window_count(a, b, -no, no, dw, eta, lor_cut / eta)
    return total


def frequency_step(cfg, rc: RunConfig) -> float:
    if rc.program == "sloss":                          # Sloss.f90:212
        return (cfg.omax_ev / HARTREE - 0.0) / (cfg.no - 1)
    omin = rc.omin_ev / HARTREE                        # Optabs.f90:201-204
    omax = cfg.omax_ev / HARTREE + omin
    return (omax - omin) / (cfg.no - 1)


class Runner:
    """Sets up one synthetic system for one q on the device.

    ops = (R, nsym): point operations R[op, i, j] (cartesian, identity first)
    of an irreducible k list; the FBZ is unfolded from it (genFBZ) and every
    k carries its operations R1, R2 (findKinIBZ / findKQinBZ).  Without ops
    the k list is the whole zone (Nsym = 1)."""

    def __init__(self, sys_, rc: RunConfig, ecut_lf: Optional[float] = None, ops=None):
        from . import synth
        cfg = sys_.cfg
        self.sys, self.rc, self.cfg = sys_, rc, cfg
        prog = RPA_SLOSS if rc.program == "sloss" else RPA_OPTABS
        nki = len(sys_.ngk)
        # FBZ from the IBZ list with the operations (genFBZ), q, k+q map
        if ops is None:
            R = np.zeros((48, 3, 3))
            R[0] = np.eye(3)
            nsym = 1
        else:
            R, nsym = np.asarray(ops[0], dtype=np.float64), int(ops[1])
        self.R, self.nsym = R, nsym
        Rf = np.asfortranarray(R)
        ktot = np.zeros((48 * nki, 3))
        import ctypes as C
        nt = C.c_int()
        ki = np.ascontiguousarray(sys_.kcart, dtype=np.float64)
        rcode = _abi.lib().rpa_gen_fbz(nki, nsym, rc.eps, ki.ctypes.data, Rf.ctypes.data, C.byref(nt),
                                       ktot.ctypes.data)
        if rcode:
            raise _abi.RpaError(rcode, "rpa_gen_fbz")
        self.ktot = ktot[:nt.value].copy()
        iq = cfg.iq if rc.iq is None else rc.iq
        self.q = _abi.find_min_q(self.ktot, iq)
        self.g = _abi.gcart(sys_.mill_g, sys_.kc)
        self.RI = _abi.point_inverse(R, nsym) if nsym > 1 else None
        self.R1 = self.R2 = None
        if nsym > 1:
            self.R1, self.K1, self.R2, self.K2, self.iG0 = _abi.find_kq_sym(self.ktot, ki, self.g, self.q, nsym,
                                                                            self.RI, rc.eps)
        else:
            self.K1, self.K2, self.iG0 = _abi.find_kq(self.ktot, ki, self.g, self.q, rc.eps)
        self.ecut_lf = synth.local_field_ecut(sys_) if ecut_lf is None else ecut_lf
        nlfd = max(cfg.nlf_target, 1) + 10
        code, nlf, glf, parg, idx = _abi.gen_glf(cfg.lf, self.ecut_lf, sys_.gcar, self.g, sys_.mill_g, nlfd)
        if code:
            raise _abi.RpaError(code, "Nlf is bigger than Nlfd")
        self.nlf, self.glf, self.parg, self.lf_idx = nlf, glf, parg, idx
        self.E = energies(sys_, rc)
        self.domega = frequency_step(cfg, rc)
        prec = _abi.RPA_MIXED if rc.precision == "mixed" else _abi.RPA_FP64
        self.ctx = _abi.Context(nki, cfg.nband, sys_.npwx, nlfd, cfg.no, prog, precision=prec,
                                device=rc.device)
        ctx = self.ctx
        ctx.set_gvectors(sys_.mill_g, sys_.kc)
        if nsym > 1:
            ctx.set_symmetry(R, nsym)
        for ik in range(nki):
            n = int(sys_.ngk[ik])
            ctx.upload_basis(ik + 1, sys_.mill_k[ik, :n], sys_.evc[ik, :, :n].T)
        ctx.set_bands(self.E)
        ctx.set_local_fields(idx, parg)
        ngd = int(sys_.ngk.min()) if rc.ngd is None else rc.ngd
        self.ngd = ngd
        eta = cfg.eta_ev / HARTREE
        ctx.set_params(nval=cfg.nval,
                       vertex_map=_abi.RPA_MAP_DIRECT if rc.vertex_map == "direct" else _abi.RPA_MAP_IGK,
                       ngd=ngd, pol=_abi.POL[rc.pol], eps=rc.eps, gcar=sys_.gcar,
                       domega=self.domega, eta=eta, lor_cut=rc.lor_cut, df_cut=rc.df_cut,
                       efermi=rc.efermi_ev / HARTREE, temp=rc.temp_ev / HARTREE,
                       vcell=sys_.vcell, c0=cfg.c, eta_intra=rc.eta_intra_ev / HARTREE)
        self._set_q(self.K1, self.K2, self.iG0, self.R1, self.R2, self.q)

    def _set_q(self, K1, K2, iG0, R1, R2, q):
        if R1 is None:
            self.ctx.set_q(K1, K2, iG0, self.ktot, q)
        else:
            self.ctx.set_q_sym(K1, K2, iG0, R1, R2, self.ktot, q)

    def q_point(self, iq: int, nlf: Optional[int] = None) -> dict:
        """Host setup of one q of the Sloss q loop (findMinQ, findKQinBZ,
        Sloss.f90:292-299): the k+q map and, with nlf given, the nlf local
        fields of smallest |q+G|.  Nothing is sent to the device."""
        q = _abi.find_min_q(self.ktot, iq)
        pt = {"iq": iq, "q": q}
        if self.nsym > 1:
            pt["R1"], pt["K1"], pt["R2"], pt["K2"], pt["iG0"] = _abi.find_kq_sym(
                self.ktot, self.sys.kcart, self.g, q, self.nsym, self.RI, self.rc.eps)
        else:
            pt["K1"], pt["K2"], pt["iG0"] = _abi.find_kq(self.ktot, self.sys.kcart, self.g, q, self.rc.eps)
        if nlf is None:
            pt["lf_idx"], pt["parg"] = self.lf_idx, self.parg
        else:
            pt["lf_idx"], pt["parg"] = local_fields_by_qg(self.g, self.sys.mill_g, self.sys.gcar, q, nlf)
        return pt

    def set_q_point(self, pt: dict) -> None:
        """Make pt (from q_point) the context's current q."""
        if pt["lf_idx"] is not self.lf_idx:
            self.ctx.set_local_fields(pt["lf_idx"], pt["parg"])
            self.lf_idx, self.parg, self.nlf = pt["lf_idx"], pt["parg"], len(pt["lf_idx"])
            self.glf = self.g[self.lf_idx - 1].copy()      # Glf(:, n) = G(:, iG)
        self._set_q(pt["K1"], pt["K2"], pt["iG0"], pt.get("R1"), pt.get("R2"), pt["q"])
        self.q, self.K1, self.K2, self.iG0 = pt["q"], pt["K1"], pt["K2"], pt["iG0"]
        self.R1, self.R2 = pt.get("R1"), pt.get("R2")

    def close(self):
        self.ctx.close()

    # step by step (each copies its result to the host)
    def vertices(self, both=False):
        return self.ctx.get_vertices(both)

    def s0(self):
        self.ctx.vertices()
        return self.ctx.accumulate_s0()

    def chi0(self):
        self.s0()
        return self.ctx.kk()

    def full(self):
        """Every output of the chain for the current q."""
        out = {}
        out["S0"], out["Qeff"] = self.s0()
        out["chi0"] = self.ctx.kk()
        if self.rc.program == "sloss":
            out["V"] = self.ctx.coulomb()
            out["epsinv"], out["chi"], out["W"] = self.ctx.dyson_w()
        out["loss"], out["pi11"], out["sig"] = self.ctx.epilogue()
        return out

    def ntransitions(self) -> int:
        import ctypes as C
        n = C.c_longlong()
        _abi._check(self.ctx.h, _abi.lib().rpa_get_vertices(self.ctx.h, None, None, C.byref(n)))
        return n.value
