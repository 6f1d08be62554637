"""ctypes access to the CPU oracle (librpa_oracle.so) and to the reference
modules compiled under oracle/_ref (libref.so, present only where the
reference sources were available at build time).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product never imports this module.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

PI = 4.0 * np.arctan(1.0)
HARTREE = 2.0 * 13.6056923

P_DBL = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
P_CPLX = np.ctypeslib.ndpointer(dtype=np.complex128, flags="F_CONTIGUOUS")
P_INT = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")


class OrcProblem(C.Structure):
    _fields_ = [
        ("nki", C.c_int), ("nband", C.c_int), ("npwx", C.c_int),
        ("ngk", C.c_void_p), ("mill_k", C.c_void_p), ("evc", C.c_void_p),
        ("ngm", C.c_int), ("mill_g", C.c_void_p), ("kc", C.c_void_p), ("E", C.c_void_p),
        ("ntot", C.c_int), ("ktot", C.c_void_p), ("K1", C.c_void_p), ("K2", C.c_void_p),
        ("iG0", C.c_void_p), ("q", C.c_double * 3), ("nlf", C.c_int), ("lf_idx", C.c_void_p),
        ("map_mode", C.c_int), ("ngd", C.c_int), ("eps", C.c_double),
        ("nsym", C.c_int), ("R", C.c_void_p), ("RI", C.c_void_p), ("R1", C.c_void_p), ("R2", C.c_void_p),
    ]


class OrcSlossParams(C.Structure):
    _fields_ = [("nocc", C.c_int), ("no", C.c_int), ("domega", C.c_double),
                ("eta", C.c_double), ("vcell", C.c_double)]


class OrcOptabsParams(C.Structure):
    _fields_ = [("nval", C.c_int), ("no", C.c_int), ("pol", C.c_int),
                ("domega", C.c_double), ("eta_inter", C.c_double), ("lor_cut", C.c_double),
                ("df_cut", C.c_double), ("efermi", C.c_double), ("T", C.c_double),
                ("vcell", C.c_double), ("gcar", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "librpa_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: run `make -C oracle liboracle`")
        _lib = C.CDLL(path)
        L = _lib
        L.orc_sloss_s0.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcSlossParams), P_DBL, C.c_int]
        L.orc_optabs_s0.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcOptabsParams), P_CPLX,
                                    C.c_void_p, C.c_int]
        L.orc_vertices.argtypes = [C.POINTER(OrcProblem), C.c_int, C.c_double, C.c_int, C.c_int,
                                   C.c_int, C.c_int, P_CPLX, C.c_void_p]
        L.orc_kk_optabs.argtypes = [C.c_int, C.c_int, C.c_double, P_CPLX, P_CPLX]
        L.orc_kk_sloss.argtypes = [C.c_int, C.c_int, C.c_double, P_DBL, P_CPLX]
        L.orc_w1w2.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, P_DBL,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_kk_module_real.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                         P_DBL, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_genV.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, P_INT,
                               P_DBL, P_DBL, P_DBL]
        L.orc_gjel.argtypes = [P_CPLX, C.c_int, C.c_int, P_CPLX, C.c_int, C.c_int]
        L.orc_dyson_w.argtypes = [C.c_int, P_CPLX, P_DBL, P_CPLX, P_CPLX, P_CPLX]
        L.orc_pi_assembly.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      P_CPLX, P_CPLX, P_CPLX, P_DBL, P_DBL]
        L.orc_fock_kernel.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, P_DBL, P_CPLX, P_CPLX, P_CPLX]
        L.orc_ladder.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                 C.c_double, P_CPLX, P_DBL, P_CPLX, P_CPLX, P_CPLX]
        L.orc_gamma_propagator.argtypes = [C.c_int, C.c_double, P_CPLX, P_CPLX, P_CPLX,
                                           C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_lu_invert.argtypes = [P_CPLX, C.c_int]
        L.orc_photon_d0.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, P_INT,
                                    P_DBL, P_CPLX]
        L.orc_photon_point.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_int, P_INT, P_DBL, P_CPLX, P_CPLX, P_CPLX, P_DBL]
        L.orc_interpolate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                      P_CPLX, P_CPLX, P_CPLX, P_CPLX]
        L.orc_gen_glf.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_int, P_DBL, P_INT,
                                  C.c_int, C.POINTER(C.c_int), P_DBL, P_INT, P_INT]
        L.orc_occupation.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.orc_gen_fbz.argtypes = [C.c_int, C.c_int, C.c_double, P_DBL, P_DBL,
                                  C.POINTER(C.c_int), P_DBL]
        L.orc_find_min_q.argtypes = [C.c_int, P_DBL, C.c_int, P_DBL]
        L.orc_find_kq.argtypes = [P_DBL, C.c_double, C.c_int, C.c_int, C.c_int, P_DBL, P_DBL,
                                  P_DBL, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_gcart.argtypes = [C.c_int, P_INT, P_DBL, P_DBL]
        L.orc_gk_sort.argtypes = [P_DBL, C.c_int, P_DBL, C.c_double, C.c_int,
                                  C.POINTER(C.c_int), P_INT]
        L.orc_gfast_linear.argtypes = [C.c_int, P_DBL, C.c_int, P_DBL, C.c_int, P_DBL, P_DBL,
                                       C.c_double, P_INT]
        L.orc_gfast_miller.argtypes = [C.c_int, P_INT, C.c_int, P_INT, C.c_int, P_INT, P_INT, P_INT]
        L.orc_point_inverse.argtypes = [C.c_int, P_DBL, P_DBL]
        L.orc_find_kin_ibz.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, P_DBL, P_DBL, P_DBL,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_find_kq_sym.argtypes = [P_DBL, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, P_DBL, P_DBL,
                                      P_DBL, P_DBL, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    return _lib


def ref_lib() -> Optional[C.CDLL]:
    """The reference modules compiled from source (None when not built here)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        _ref = C.CDLL(path)
        R = _ref
        R.ref_kk_cmplx.argtypes = [C.c_int] * 5 + [C.c_double, C.c_double, P_CPLX,
                                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        R.ref_kk_real.argtypes = [C.c_int] * 5 + [C.c_double, C.c_double, P_DBL,
                                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
        R.ref_gjel.argtypes = [P_CPLX, C.c_int, C.c_int, P_CPLX, C.c_int, C.c_int]
        R.ref_init_igk.argtypes = [C.c_int, P_DBL, C.c_int, C.c_int, P_DBL, C.c_double, P_INT]
        R.ref_genglf.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, P_DBL, P_INT, C.c_int,
                                 C.POINTER(C.c_int), P_DBL, P_INT]
    return _ref


def F(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


# ---------------------------------------------------------------------------
# host setup (restated reference subroutines)

def gen_glf(lf, ecut, gcar, g, mill_g, nlfd):
    g = F(np.asarray(g).T, np.float64)            # (3, ng)
    mill = F(np.asarray(mill_g).T, np.int32)
    ng = g.shape[1]
    glf = np.zeros((3, max(nlfd, 1)), order="F")
    parg = np.zeros(max(nlfd, 1), dtype=np.int32, order="F")
    idx = np.zeros(max(nlfd, 1), dtype=np.int32, order="F")
    nlf = C.c_int()
    rc = lib().orc_gen_glf(lf.encode(), ecut, gcar, ng, g, mill, nlfd, C.byref(nlf), glf, parg, idx)
    n = nlf.value
    return rc, n, glf[:, :n].T.copy(), parg[:n].copy(), idx[:n].copy()


def occupation(ei, ef, T):
    e, f = C.c_double(), C.c_double()
    lib().orc_occupation(ei, ef, T, C.byref(e), C.byref(f))
    return e.value, f.value


def find_min_q(ktot, iq):
    k = F(np.asarray(ktot).T, np.float64)
    q = np.zeros(3, order="F")
    lib().orc_find_min_q(k.shape[1], k, iq, q)
    return q


def find_kq_all(ktot, ki, g, q, eps):
    """(K2, iG0) 1-based for every FBZ k, Nsym = 1 (Optabs.f90:308-320)."""
    kt = F(np.asarray(ktot).T, np.float64)
    kI = F(np.asarray(ki).T, np.float64)
    gg = F(np.asarray(g).T, np.float64)
    ntot = kt.shape[1]
    K2 = np.zeros(ntot, dtype=np.int32)
    iG0 = np.zeros(ntot, dtype=np.int32)
    a, b = C.c_int(), C.c_int()
    for ik in range(ntot):
        kq = F(kt[:, ik] + q, np.float64)
        rc = lib().orc_find_kq(kq, eps, kI.shape[1], ntot, gg.shape[1], kt, kI, gg,
                               C.byref(a), C.byref(b))
        if rc != 0:
            raise RuntimeError(f"k+q not found for k={ik + 1} (code {rc})")
        iG0[ik], K2[ik] = a.value, b.value
    return K2, iG0


def point_inverse(R, nsym):
    """PointR's inverses of R(48,3,3) (Chi_RPA/PointR.f90:119-138)."""
    Rf = F(R, np.float64)
    RI = np.zeros((48, 3, 3), order="F")
    lib().orc_point_inverse(nsym, Rf, RI)
    return RI


def find_kin_ibz(ik, ki, k, nsym, RI, eps):
    kI = F(np.asarray(ki).T, np.float64)
    a, b = C.c_int(), C.c_int()
    rc = lib().orc_find_kin_ibz(ik, kI.shape[1], nsym, eps, F(k, np.float64), F(RI, np.float64), kI,
                                C.byref(a), C.byref(b))
    if rc:
        raise RuntimeError(f"k {ik} not in the IBZ")
    return a.value, b.value


def find_kq_sym_all(ktot, ki, g, q, nsym, RI, eps):
    """(iG0, R2, K2) 1-based for every FBZ k (Optabs.f90:822-869)."""
    kt = F(np.asarray(ktot).T, np.float64)
    kI = F(np.asarray(ki).T, np.float64)
    gg = F(np.asarray(g).T, np.float64)
    RIf = F(RI, np.float64)
    ntot = kt.shape[1]
    out = np.zeros((3, ntot), dtype=np.int32)
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    for ik in range(ntot):
        kq = F(kt[:, ik] + q, np.float64)
        rc = lib().orc_find_kq_sym(kq, eps, nsym, kI.shape[1], ntot, gg.shape[1], kt, kI, gg, RIf,
                                   C.byref(a), C.byref(b), C.byref(c))
        if rc:
            raise RuntimeError(f"k+q not found for k={ik + 1} (code {rc})")
        out[:, ik] = a.value, b.value, c.value
    return out[0], out[1], out[2]


def gk_sort(k, g, gcutw, npwx):
    kk = F(k, np.float64)
    gg = F(np.asarray(g).T, np.float64)
    igk = np.zeros(npwx, dtype=np.int32, order="F")
    n = C.c_int()
    rc = lib().orc_gk_sort(kk, gg.shape[1], gg, gcutw, npwx, C.byref(n), igk)
    return rc, igk[:n.value].copy()


def gfast_linear(l1, l2, glf, g0, eps):
    l1 = F(np.asarray(l1).T, np.float64)
    l2 = F(np.asarray(l2).T, np.float64)
    gl = F(np.asarray(glf).T, np.float64)
    n1, n2, nlf = l1.shape[1], l2.shape[1], gl.shape[1]
    out = np.zeros(nlf * n1, dtype=np.int32, order="F")
    lib().orc_gfast_linear(n1, l1, n2, l2, nlf, gl, F(g0, np.float64), eps, out)
    return out.reshape(nlf, n1)


def gfast_miller(m1, m2, mlf, m0):
    m1 = F(np.asarray(m1).T, np.int32)
    m2 = F(np.asarray(m2).T, np.int32)
    ml = F(np.asarray(mlf).T, np.int32)
    n1, n2, nlf = m1.shape[1], m2.shape[1], ml.shape[1]
    out = np.zeros(nlf * n1, dtype=np.int32, order="F")
    lib().orc_gfast_miller(n1, m1, n2, m2, nlf, ml, F(m0, np.int32), out)
    return out.reshape(nlf, n1)


# ---------------------------------------------------------------------------
# problem description shared by the response routines

class Problem:
    """Holds the arrays of one (system, q) and the ctypes view of them."""

    def __init__(self, sys, K1, K2, iG0, ktot, q, lf_idx, map_mode=0, ngd=None, eps=1e-4,
                 E_hartree=None, R=None, RI=None, R1=None, R2=None):
        self.sys = sys
        nki = len(sys.ngk)
        self.ngk = F(sys.ngk, np.int32)
        self.mill_k = np.ascontiguousarray(sys.mill_k, dtype=np.int32)       # (nki,npwx,3) C == (3,npwx,nki) F
        self.evc = np.ascontiguousarray(sys.evc, dtype=np.complex128)         # (nki,nband,npwx) C
        self.mill_g = np.ascontiguousarray(sys.mill_g, dtype=np.int32)
        self.kc = F(sys.kc, np.float64)
        self.E = F(E_hartree, np.float64)                                    # (nki, nband)
        self.ktot = np.ascontiguousarray(ktot, dtype=np.float64)
        self.K1 = np.ascontiguousarray(K1, dtype=np.int32)
        self.K2 = np.ascontiguousarray(K2, dtype=np.int32)
        self.iG0 = np.ascontiguousarray(iG0, dtype=np.int32)
        self.lf_idx = np.ascontiguousarray(lf_idx, dtype=np.int32)
        self.q = np.asarray(q, dtype=np.float64)
        self.map_mode = map_mode
        self.ngd = int(ngd if ngd is not None else sys.ngk.min())
        self.eps = eps
        s = OrcProblem()
        s.nki, s.nband, s.npwx = nki, sys.evc.shape[1], sys.npwx
        s.ngk = self.ngk.ctypes.data
        s.mill_k = self.mill_k.ctypes.data
        s.evc = self.evc.ctypes.data
        s.ngm = len(self.mill_g)
        s.mill_g = self.mill_g.ctypes.data
        s.kc = self.kc.ctypes.data
        s.E = self.E.ctypes.data
        s.ntot = len(self.K1)
        s.ktot = self.ktot.ctypes.data
        s.K1 = self.K1.ctypes.data
        s.K2 = self.K2.ctypes.data
        s.iG0 = self.iG0.ctypes.data
        for i in range(3):
            s.q[i] = self.q[i]
        s.nlf = len(self.lf_idx)
        s.lf_idx = self.lf_idx.ctypes.data
        s.map_mode = map_mode
        s.ngd = self.ngd
        s.eps = eps
        # point-group unfolding: R, RI (48,3,3) cartesian, R1/R2 1-based per FBZ k
        s.nsym = 1
        if R1 is not None:
            self.R = F(R, np.float64)
            self.RI = F(RI, np.float64)
            self.R1 = np.ascontiguousarray(R1, dtype=np.int32)
            self.R2 = np.ascontiguousarray(R2, dtype=np.int32)
            s.nsym = int(max(self.R1.max(), self.R2.max()))
            s.R, s.RI = self.R.ctypes.data, self.RI.ctypes.data
            s.R1, s.R2 = self.R1.ctypes.data, self.R2.ctypes.data
        self.s = s

    @property
    def nlf(self):
        return len(self.lf_idx)


def vertices(pb: Problem, kind, gcar, n_lo, n_hi, m_lo, m_hi, both=False):
    ntrans = len(pb.K1) * (n_hi - n_lo) * (m_hi - m_lo)
    r1 = np.zeros((ntrans, pb.nlf), dtype=np.complex128, order="F")
    r2 = np.zeros((ntrans, pb.nlf), dtype=np.complex128, order="F") if both else None
    rc = lib().orc_vertices(C.byref(pb.s), kind, gcar, n_lo, n_hi, m_lo, m_hi, r1,
                            r2.ctypes.data if both else None)
    if rc:
        raise RuntimeError(f"orc_vertices failed ({rc})")
    return (r1, r2) if both else r1


def sloss_s0(pb: Problem, nocc, no, domega, eta, vcell, nthreads=None):
    nthreads = nthreads or os.cpu_count()
    S0 = np.zeros((no, pb.nlf, pb.nlf), order="F")
    p = OrcSlossParams(nocc, no, domega, eta, vcell)
    rc = lib().orc_sloss_s0(C.byref(pb.s), C.byref(p), S0, nthreads)
    if rc:
        raise RuntimeError(f"orc_sloss_s0 failed ({rc})")
    return S0


def optabs_s0(pb: Problem, nval, no, pol, domega, eta_inter, lor_cut, df_cut, efermi, T, vcell,
              gcar, nthreads=None):
    nthreads = nthreads or os.cpu_count()
    S0 = np.zeros((2 * no + 1, pb.nlf, pb.nlf), dtype=np.complex128, order="F")
    Q = np.zeros((pb.nlf, pb.nlf), dtype=np.complex128, order="F")
    p = OrcOptabsParams(nval, no, pol, domega, eta_inter, lor_cut, df_cut, efermi, T, vcell, gcar)
    rc = lib().orc_optabs_s0(C.byref(pb.s), C.byref(p), S0, Q.ctypes.data, nthreads)
    if rc:
        raise RuntimeError(f"orc_optabs_s0 failed ({rc})")
    return S0, Q


def kk_optabs(S0, domega):
    nw, nlf, _ = S0.shape
    no = (nw - 1) // 2
    out = np.zeros((no - 1, nlf, nlf), dtype=np.complex128, order="F")
    lib().orc_kk_optabs(no, nlf, domega, F(S0, np.complex128), out)
    return out


def kk_sloss(S0, domega):
    no, nlf, _ = S0.shape
    out = np.zeros((no - 1, nlf, nlf), dtype=np.complex128, order="F")
    lib().orc_kk_sloss(no, nlf, domega, F(S0, np.float64), out)
    return out


def w1w2(io, iG, jG, S0, domega):
    no, nlf, _ = S0.shape
    a, b = C.c_double(), C.c_double()
    lib().orc_w1w2(io, no, nlf, iG, jG, domega, F(S0, np.float64), C.byref(a), C.byref(b))
    return a.value, b.value


def kk_module_real(io, iG, jG, S0m, domega):
    nw, nlf, _ = S0m.shape
    no = (nw - 1) // 2
    a, b = C.c_double(), C.c_double()
    lib().orc_kk_module_real(io, no, nlf, iG, jG, domega, F(S0m, np.float64), C.byref(a), C.byref(b))
    return a.value, b.value


def genV(eps, qx, qy, c0, parg, glf, gcar):
    glf = np.asarray(glf, dtype=np.float64)
    nlf = len(glf)
    V = np.zeros((nlf, nlf), order="F")
    lib().orc_genV(eps, qx, qy, c0, nlf, F(parg, np.int32), F(glf.T, np.float64),
                   F((gcar * glf).T, np.float64), V)
    return V


def gjel(a):
    a = F(a, np.complex128).copy(order="F")
    n = a.shape[0]
    b = np.zeros((n, n), dtype=np.complex128, order="F")
    lib().orc_gjel(a, n, n, b, n, n)
    return a


def dyson_w(chi0, V):
    nlf = chi0.shape[0]
    e = np.zeros((nlf, nlf), dtype=np.complex128, order="F")
    c = np.zeros_like(e)
    w = np.zeros_like(e)
    lib().orc_dyson_w(nlf, F(chi0, np.complex128), F(V, np.float64), e, c, w)
    return e, c, w


def pi_assembly(Pi0, Qeff, domega, eta_intra, c0):
    no = Pi0.shape[0] + 1
    nlf = Pi0.shape[1]
    pi11 = np.zeros(no - 1, dtype=np.complex128, order="F")
    si = np.zeros(no - 1, order="F")
    sa = np.zeros(no - 1, order="F")
    lib().orc_pi_assembly(no, nlf, domega, eta_intra, c0, F(Pi0, np.complex128),
                          F(Qeff, np.complex128), pi11, si, sa)
    return pi11, si, sa


def interpolate(Pi0, Qeff, no_interp, domega_tot, eta_intra):
    """interpolate.x after the KK (Chi_RPA/interpolate.f90:103-186):
    Pi_tot_interp(No_tot,Nlf,Nlf) and the intraband column (No_tot) from
    Pi0 = kk_optabs(S0, domega)."""
    no = Pi0.shape[0] + 1
    nlf = Pi0.shape[1]
    no_tot = (no_interp - 1) * (no - 2)
    out = np.zeros((max(no_tot, 1), nlf, nlf), dtype=np.complex128, order="F")
    intra = np.zeros(max(no_tot, 1), dtype=np.complex128, order="F")
    if no_tot > 0:
        lib().orc_interpolate(no, nlf, no_interp, domega_tot, eta_intra, F(Pi0, np.complex128),
                              F(Qeff, np.complex128), out, intra)
    return out[:no_tot], intra[:no_tot]


def lu_invert(a):
    """zgetrf + zgetri as Photon's invert uses them (BSE/Photonew.f90:388-462);
    returns (inverse, info)."""
    a = F(a, np.complex128).copy(order="F")
    info = lib().orc_lu_invert(a, a.shape[0])
    return a, info


def photon_d0(q, o, eta, c0, parg, glfz):
    """genD0 (BSE/Photonew.f90:618-663): (Dxx0, Dyy0, Dzz0, Dyz0) as (N,N,4)."""
    n = len(glfz)
    D = np.zeros((n, n, 4), dtype=np.complex128, order="F")
    lib().orc_photon_d0(q, o, eta, c0, n, F(parg, np.int32), F(glfz, np.float64), D)
    return D


def photon_point(q, o, eta, c0, theta, parg, glfz, pi0):
    """One point of Photon's loops (BSE/Photonew.f90:229-330): pi0 (N,N,5) =
    xx, yy, yz, zy, zz -> (screened (N,N,5), sums (5: xx, yy, zz, yz, zy),
    (A_p, R_p), status)."""
    n = len(glfz)
    pi = np.zeros((n, n, 5), dtype=np.complex128, order="F")
    sums = np.zeros(5, dtype=np.complex128)
    spec = np.zeros(2)
    st = lib().orc_photon_point(q, o, eta, c0, theta, n, F(parg, np.int32), F(glfz, np.float64),
                                F(pi0, np.complex128), pi, sums, spec)
    return pi, sums, spec, st


def gamma_propagator(D, domega):
    """genGammaPropagator (Modules/propagators.f90:11-108) of one series D(No):
    (Gammap (No-1), Gammam (No-1), KKS, SKK)."""
    D = F(D, np.complex128)
    no = len(D)
    gp = np.zeros(no - 1, dtype=np.complex128, order="F")
    gm = np.zeros(no - 1, dtype=np.complex128, order="F")
    a, b = C.c_double(), C.c_double()
    lib().orc_gamma_propagator(no, domega, D, gp, gm, C.byref(a), C.byref(b))
    return gp, gm, a.value, b.value


def ladder(fockk, delta, current, no, omin, domega, eta, ntot, vcell, valley):
    """Chi_ladd's frequency loop (Chi_ladd_up/BSE.f90:1116-1176):
    (chi_ladd, chi_para) (2, no) complex."""
    nk = len(delta)
    cl = np.zeros((2, no), dtype=np.complex128, order="F")
    cp = np.zeros((2, no), dtype=np.complex128, order="F")
    lib().orc_ladder(nk, no, omin, domega, eta, ntot, vcell, valley, F(fockk, np.complex128),
                     F(delta, np.float64), F(current, np.complex128), cl, cp)
    return cl, cp


def fock_kernel(kbz, wt, mnm, nq, ntot, gcar, qmin, c0, vcell):
    """The Fock kernel of Chi_ladd (Chi_ladd_up/BSE.f90:305,542,804-816):
    kbz (3, nk), wt (nq, nlf, nlf), mnm (2, nlf, nk, nk) -> fockk (nk, nk)."""
    kbz = F(kbz, np.float64)
    nk = kbz.shape[1]
    nlf = wt.shape[1]
    out = np.zeros((nk, nk), dtype=np.complex128, order="F")
    lib().orc_fock_kernel(nk, nlf, nq, ntot, gcar, qmin, c0, vcell, kbz, F(wt, np.complex128),
                          F(mnm, np.complex128), out)
    return out
