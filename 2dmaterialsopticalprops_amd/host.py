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


def frequency_step(cfg, rc: RunConfig) -> float:
    if rc.program == "sloss":                          # Sloss.f90:212
        return (cfg.omax_ev / HARTREE - 0.0) / (cfg.no - 1)
    omin = rc.omin_ev / HARTREE                        # Optabs.f90:201-204
    omax = cfg.omax_ev / HARTREE + omin
    return (omax - omin) / (cfg.no - 1)


class Runner:
    """Sets up one synthetic system for one q on the device."""

    def __init__(self, sys_, rc: RunConfig, ecut_lf: Optional[float] = None):
        from . import synth
        cfg = sys_.cfg
        self.sys, self.rc, self.cfg = sys_, rc, cfg
        prog = RPA_SLOSS if rc.program == "sloss" else RPA_OPTABS
        nki = len(sys_.ngk)
        # FBZ from the IBZ list with the identity (genFBZ), q, k+q map
        R = np.zeros((48, 3, 3), order="F")
        R[0] = np.eye(3)
        ktot = np.zeros((48 * nki, 3))
        import ctypes as C
        nt = C.c_int()
        ki = np.ascontiguousarray(sys_.kcart, dtype=np.float64)
        rcode = _abi.lib().rpa_gen_fbz(nki, 1, rc.eps, ki.ctypes.data, R.ctypes.data, C.byref(nt),
                                       ktot.ctypes.data)
        if rcode:
            raise _abi.RpaError(rcode, "rpa_gen_fbz")
        self.ktot = ktot[:nt.value].copy()
        iq = cfg.iq if rc.iq is None else rc.iq
        self.q = _abi.find_min_q(self.ktot, iq)
        self.g = _abi.gcart(sys_.mill_g, sys_.kc)
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
        ctx.set_q(self.K1, self.K2, self.iG0, self.ktot, self.q)

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
