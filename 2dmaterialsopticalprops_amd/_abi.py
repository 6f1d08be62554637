"""ctypes binding of librpa_hip.so (include/rpa_hip.h).

The library is the product: loading fails loudly when it has not been built;
there is no host fallback for any device entry point.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "librpa_hip.so")

RPA_OK, RPA_E_ARG, RPA_E_STATE, RPA_E_HIP, RPA_E_NOTFOUND, RPA_E_LIMIT, RPA_E_SINGULAR = range(7)
RPA_SLOSS, RPA_OPTABS = 0, 1
RPA_FP64, RPA_MIXED = 0, 1
RPA_MAP_DIRECT, RPA_MAP_IGK = 0, 1
POL = {"xx": 1, "x": 1, "yy": 2, "y": 2, "zz": 3, "z": 3, "yz": 4}

EXPORTS = [
    "rpa_ctx_create", "rpa_ctx_destroy", "rpa_last_error", "rpa_last_run_ms",
    "rpa_set_gvectors", "rpa_upload_basis", "rpa_set_bands", "rpa_set_local_fields",
    "rpa_set_params", "rpa_set_q", "rpa_vertices", "rpa_accumulate_s0", "rpa_kk",
    "rpa_coulomb", "rpa_dyson_w", "rpa_epilogue", "rpa_run", "rpa_run_graph",
    "rpa_get_vertices", "rpa_gen_fbz", "rpa_find_min_q", "rpa_find_kq", "rpa_gen_glf",
    "rpa_gk_sort", "rpa_gcart", "rpa_set_profiling", "rpa_stage_ms", "rpa_get_result",
    "rpa_load_s0", "rpa_w_kk", "rpa_kk_noalloc", "rpa_set_autotune", "rpa_batch_set", "rpa_batch_run",
    "rpa_batch_result", "rpa_set_symmetry", "rpa_set_q_sym", "rpa_batch_set_sym", "rpa_find_kq_sym",
    "rpa_point_inverse", "rpa_interpolate", "rpa_photon", "rpa_gamma_w", "rpa_gamma_propagator",
    "rpa_ladder", "rpa_fock_kernel",
]
# enum rpa_result
RES = {"S0": 0, "QEFF": 1, "CHI0": 2, "V": 3, "EPSINV": 4, "CHI": 5, "W": 6, "LOSS": 7,
       "PI11": 8, "SIG": 9}
STAGES = ["sort", "vertices", "accumulate", "reduce", "kk", "dyson_or_pi"]


class RpaDims(C.Structure):
    _fields_ = [("nki", C.c_int), ("nband", C.c_int), ("npwx", C.c_int), ("nlfd", C.c_int),
                ("no", C.c_int), ("program", C.c_int), ("precision", C.c_int), ("device", C.c_int)]


class RpaParams(C.Structure):
    _fields_ = [("nval", C.c_int), ("vertex_map", C.c_int), ("ngd", C.c_int), ("pol", C.c_int),
                ("eps", C.c_double), ("gcar", C.c_double), ("domega", C.c_double),
                ("eta", C.c_double), ("lor_cut", C.c_double), ("df_cut", C.c_double),
                ("efermi", C.c_double), ("temp", C.c_double), ("vcell", C.c_double),
                ("c0", C.c_double), ("eta_intra", C.c_double)]


class RpaInterpParams(C.Structure):
    _fields_ = [("no", C.c_int), ("nlf", C.c_int), ("no_interp", C.c_int), ("all", C.c_int),
                ("device", C.c_int), ("domega", C.c_double), ("domega_tot", C.c_double),
                ("eta_intra", C.c_double)]


class RpaPhotonParams(C.Structure):
    _fields_ = [("no", C.c_int), ("nlf", C.c_int), ("nline", C.c_int), ("device", C.c_int),
                ("omin", C.c_double), ("domega", C.c_double), ("eta", C.c_double), ("c0", C.c_double)]


class RpaFockParams(C.Structure):
    _fields_ = [("nk", C.c_int), ("nlf", C.c_int), ("nq", C.c_int), ("ntot", C.c_int), ("device", C.c_int),
                ("pad", C.c_int), ("gcar", C.c_double), ("qmin", C.c_double), ("c0", C.c_double),
                ("vcell", C.c_double)]


class RpaLadderParams(C.Structure):
    _fields_ = [("nk", C.c_int), ("no", C.c_int), ("ntot", C.c_int), ("device", C.c_int),
                ("omin", C.c_double), ("domega", C.c_double), ("eta", C.c_double), ("vcell", C.c_double),
                ("valley", C.c_double)]


class RpaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rpa error {code}: {msg}")
        self.code = code


_lib = None

P = C.c_void_p


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.rpa_ctx_create.argtypes = [C.POINTER(RpaDims), C.POINTER(P)]
        L.rpa_ctx_destroy.argtypes = [P]
        L.rpa_last_error.argtypes = [P]
        L.rpa_last_error.restype = C.c_char_p
        L.rpa_last_run_ms.argtypes = [P]
        L.rpa_last_run_ms.restype = C.c_double
        L.rpa_set_gvectors.argtypes = [P, C.c_int, P, P]
        L.rpa_upload_basis.argtypes = [P, C.c_int, C.c_int, P, P]
        L.rpa_set_bands.argtypes = [P, P]
        L.rpa_set_local_fields.argtypes = [P, C.c_int, P, P]
        L.rpa_set_params.argtypes = [P, C.POINTER(RpaParams)]
        L.rpa_set_q.argtypes = [P, C.c_int, P, P, P, P, P]
        L.rpa_vertices.argtypes = [P]
        L.rpa_accumulate_s0.argtypes = [P, P, P]
        L.rpa_kk.argtypes = [P, P]
        L.rpa_coulomb.argtypes = [P, P]
        L.rpa_dyson_w.argtypes = [P, P, P, P]
        L.rpa_epilogue.argtypes = [P, P, P, P]
        L.rpa_run.argtypes = [P]
        L.rpa_run_graph.argtypes = [P]
        L.rpa_get_vertices.argtypes = [P, P, P, C.POINTER(C.c_longlong)]
        L.rpa_gen_fbz.argtypes = [C.c_int, C.c_int, C.c_double, P, P, C.POINTER(C.c_int), P]
        L.rpa_find_min_q.argtypes = [C.c_int, P, C.c_int, P]
        L.rpa_find_kq.argtypes = [C.c_int, P, C.c_int, P, C.c_int, P, P, C.c_double, P, P, P]
        L.rpa_gen_glf.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_int, P, P, C.c_int,
                                  C.POINTER(C.c_int), P, P, P]
        L.rpa_gk_sort.argtypes = [P, C.c_int, P, C.c_double, C.c_int, C.POINTER(C.c_int), P]
        L.rpa_gcart.argtypes = [C.c_int, P, P, P]
        L.rpa_set_profiling.argtypes = [P, C.c_int]
        L.rpa_stage_ms.argtypes = [P, P, C.c_int]
        L.rpa_get_result.argtypes = [P, C.c_int, P]
        L.rpa_load_s0.argtypes = [P, P, P]
        L.rpa_w_kk.argtypes = [P, P, P]
        L.rpa_kk_noalloc.argtypes = [P, C.c_void_p]
        L.rpa_set_autotune.argtypes = [P, C.c_int]
        L.rpa_batch_set.argtypes = [P, C.c_int, P, P, P, C.c_int, P, P, P, P, P]
        L.rpa_batch_run.argtypes = [P]
        L.rpa_batch_result.argtypes = [P, C.c_int, C.c_int, P]
        L.rpa_set_symmetry.argtypes = [P, C.c_int, P]
        L.rpa_set_q_sym.argtypes = [P, C.c_int, P, P, P, P, P, P, P]
        L.rpa_batch_set_sym.argtypes = [P, C.c_int, P, P, P, C.c_int, P, P, P, P, P, P, P]
        L.rpa_find_kq_sym.argtypes = [C.c_int, P, C.c_int, P, C.c_int, P, P, C.c_int, P, C.c_double,
                                      P, P, P, P, P]
        L.rpa_point_inverse.argtypes = [C.c_int, P, P]
        L.rpa_interpolate.argtypes = [C.POINTER(RpaInterpParams), P, P, P, P]
        L.rpa_ladder.argtypes = [C.POINTER(RpaLadderParams), P, P, P, P, P]
        L.rpa_fock_kernel.argtypes = [C.POINTER(RpaFockParams), P, P, P, P]
        L.rpa_gamma_w.argtypes = [P, P, P]
        L.rpa_gamma_propagator.argtypes = [C.c_int, C.c_int, C.c_double, P, P, P, P, P, C.c_int]
        L.rpa_photon.argtypes = [C.POINTER(RpaPhotonParams), P, P, P, P, P, P, P, P, P, P, P, P, P, P]
        _lib = L
    return _lib


def _ptr(a):
    """Address of a numpy array.  The caller must hold a reference to `a`
    until the C call returns (never pass the address of a temporary)."""
    return a.ctypes.data if a is not None else None


def _f(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


# ---------------------------------------------------------------------------
# host setup entry points (no device needed)

def gcart(mill, kc):
    mill = np.ascontiguousarray(mill, dtype=np.int32)
    g = np.zeros((len(mill), 3))
    kcf = _f(kc, np.float64)  # keep the converted array alive across the call
    _check(None, lib().rpa_gcart(len(mill), _ptr(mill), _ptr(kcf), _ptr(g)))
    return g


def find_min_q(ktot, iq):
    k = np.ascontiguousarray(ktot, dtype=np.float64)
    q = np.zeros(3)
    _check(None, lib().rpa_find_min_q(len(k), _ptr(k), iq, _ptr(q)))
    return q


def find_kq(ktot, ki, g, q, eps):
    k = np.ascontiguousarray(ktot, dtype=np.float64)
    kI = np.ascontiguousarray(ki, dtype=np.float64)
    gg = np.ascontiguousarray(g, dtype=np.float64)
    qq = np.ascontiguousarray(q, dtype=np.float64)
    n = len(k)
    K1, K2, iG0 = (np.zeros(n, dtype=np.int32) for _ in range(3))
    _check(None, lib().rpa_find_kq(n, _ptr(k), len(kI), _ptr(kI), len(gg), _ptr(gg), _ptr(qq), eps,
                                   _ptr(K1), _ptr(K2), _ptr(iG0)))
    return K1, K2, iG0


def point_inverse(R, nsym):
    """PointR's inverses RI(48,3,3) of R(48,3,3) (numpy arrays indexed [op, i, j])."""
    Rf = _f(R, np.float64)
    RI = np.zeros((48, 3, 3), order="F")
    _check(None, lib().rpa_point_inverse(nsym, _ptr(Rf), _ptr(RI)))
    return np.ascontiguousarray(RI)


def find_kq_sym(ktot, ki, g, q, nsym, RI, eps):
    """(R1, K1, R2, K2, iG0), 1-based, of every FBZ k with point operations."""
    k = np.ascontiguousarray(ktot, dtype=np.float64)
    kI = np.ascontiguousarray(ki, dtype=np.float64)
    gg = np.ascontiguousarray(g, dtype=np.float64)
    qq = np.ascontiguousarray(q, dtype=np.float64)
    RIf = _f(RI, np.float64)
    n = len(k)
    R1, K1, R2, K2, iG0 = (np.zeros(n, dtype=np.int32) for _ in range(5))
    _check(None, lib().rpa_find_kq_sym(n, _ptr(k), len(kI), _ptr(kI), len(gg), _ptr(gg), _ptr(qq), nsym,
                                       _ptr(RIf), eps, _ptr(R1), _ptr(K1), _ptr(R2), _ptr(K2), _ptr(iG0)))
    return R1, K1, R2, K2, iG0


def gen_glf(lf, ecut, gcar, g, mill, nlfd):
    gg = np.ascontiguousarray(g, dtype=np.float64)
    mm = np.ascontiguousarray(mill, dtype=np.int32)
    glf = np.zeros((max(nlfd, 1), 3))
    parg = np.zeros(max(nlfd, 1), dtype=np.int32)
    idx = np.zeros(max(nlfd, 1), dtype=np.int32)
    n = C.c_int()
    rc = lib().rpa_gen_glf(lf.encode(), ecut, gcar, len(gg), _ptr(gg), _ptr(mm), nlfd, C.byref(n),
                           _ptr(glf), _ptr(parg), _ptr(idx))
    m = min(n.value, nlfd)
    return rc, n.value, glf[:m].copy(), parg[:m].copy(), idx[:m].copy()


def gk_sort(k, g, gcutw, npwx):
    kk = np.ascontiguousarray(k, dtype=np.float64)
    gg = np.ascontiguousarray(g, dtype=np.float64)
    igk = np.zeros(npwx, dtype=np.int32)
    n = C.c_int()
    rc = lib().rpa_gk_sort(_ptr(kk), len(gg), _ptr(gg), gcutw, npwx, C.byref(n), _ptr(igk))
    return rc, igk[:n.value].copy()


def interpolate(S0, qeff, no_interp, domega, domega_tot, eta_intra, all_pairs=True, device=0):
    """interpolate.x on the device (rpa_interpolate): S0(-No:No,Nlf,Nlf) and
    Qeff(Nlf,Nlf) as read from Corrfun_{pol} / Qeff_{pol} ->
    (Pi_tot_interp(No_tot,Nlf,Nlf), or its (1,1) column (No_tot) when
    all_pairs is False, and the intraband column (No_tot))."""
    S0f = _f(S0, np.complex128)
    qf = _f(qeff, np.complex128)
    if S0f.ndim != 3 or S0f.shape[1] != S0f.shape[2] or S0f.shape[0] % 2 != 1:
        raise ValueError("S0 must be (2 No + 1, Nlf, Nlf)")
    nlf = S0f.shape[1]
    if qf.shape != (nlf, nlf):
        raise ValueError("Qeff must be (Nlf, Nlf)")
    no = (S0f.shape[0] - 1) // 2
    no_tot = max(no_interp - 1, 0) * max(no - 2, 0)
    shape = (no_tot, nlf, nlf) if all_pairs else (no_tot,)
    out = np.zeros(shape, dtype=np.complex128, order="F")
    intra = np.zeros(no_tot, dtype=np.complex128)
    p = RpaInterpParams(no, nlf, no_interp, 1 if all_pairs else 0, device, domega, domega_tot, eta_intra)
    rc = lib().rpa_interpolate(C.byref(p), _ptr(S0f), _ptr(qf), _ptr(out), _ptr(intra))
    if rc != RPA_OK:
        raise RpaError(rc, lib().rpa_last_error(None).decode())
    return out, intra


def fock_kernel(kbz, wt, mnm, ntot, gcar, qmin, c0, vcell, device=0):
    """The Fock kernel of Chi_ladd on the device (rpa_fock_kernel): kbz (3, nk),
    wt (nq, nlf, nlf), mnm (2, nlf, nk, nk) -> fockk (nk, nk) complex."""
    k = _f(kbz, np.float64)
    w = _f(wt, np.complex128)
    m = _f(mnm, np.complex128)
    nk = k.shape[1]
    nq, nlf = w.shape[0], w.shape[1]
    if k.shape != (3, nk) or w.shape != (nq, nlf, nlf) or m.shape != (2, nlf, nk, nk):
        raise ValueError("kbz (3, nk), wt (nq, nlf, nlf) and mnm (2, nlf, nk, nk) expected")
    out = np.zeros((nk, nk), dtype=np.complex128, order="F")
    p = RpaFockParams(nk, nlf, nq, ntot, device, 0, gcar, qmin, c0, vcell)
    rc = lib().rpa_fock_kernel(C.byref(p), _ptr(k), _ptr(w), _ptr(m), _ptr(out))
    if rc != RPA_OK:
        raise RpaError(rc, lib().rpa_last_error(None).decode())
    return out


def ladder(fockk, delta, current, no, omin, domega, eta, ntot, vcell, valley, device=0):
    """Chi_ladd's frequency loop on the device (rpa_ladder): fockk (nk, nk),
    delta (nk), current (2, nk) -> chi_ladd, chi_para (2, no) complex."""
    F = _f(fockk, np.complex128)
    d = _f(delta, np.float64)
    j = _f(current, np.complex128)
    nk = F.shape[0]
    if F.shape != (nk, nk) or d.shape != (nk,) or j.shape != (2, nk):
        raise ValueError("fockk (nk, nk), delta (nk) and current (2, nk) expected")
    cl = np.zeros((2, no), dtype=np.complex128, order="F")
    cp = np.zeros((2, no), dtype=np.complex128, order="F")
    p = RpaLadderParams(nk, no, ntot, device, omin, domega, eta, vcell, valley)
    rc = lib().rpa_ladder(C.byref(p), _ptr(F), _ptr(d), _ptr(j), _ptr(cl), _ptr(cp))
    if rc != RPA_OK:
        raise RpaError(rc, lib().rpa_last_error(None).decode())
    return cl, cp


def gamma_propagator(D, domega, device=0):
    """genGammaPropagator on the device for the columns of D (No, nser):
    returns (Gammap, Gammam) (No-1, nser) complex and (KKS, SKK) (nser)."""
    Df = _f(D, np.complex128)
    if Df.ndim == 1:
        Df = Df.reshape(-1, 1, order="F")
    no, nser = Df.shape
    gp = np.zeros((no - 1, nser), dtype=np.complex128, order="F")
    gm = np.zeros_like(gp)
    kks = np.zeros(nser)
    skk = np.zeros(nser)
    rc = lib().rpa_gamma_propagator(no, nser, domega, _ptr(Df), _ptr(gp), _ptr(gm), _ptr(kks), _ptr(skk), device)
    if rc != RPA_OK:
        raise RpaError(rc, lib().rpa_last_error(None).decode())
    return gp, gm, kks, skk


def photon(pixx0, piyy0, pizz0, glfz, parg, lines, omin, domega, eta, c0, piyz0=None, pizy0=None,
           full=False, device=0):
    """Photon.x's per-point work on the device (rpa_photon).  Pi0 arrays are
    (No, Nlf, Nlf) complex as loadPi0 reads them; lines is a list of
    (q, theta, qmode) with qmode 1 for q = gamma o sin(theta).  Returns a dict
    with pi11 (No-1, nline, 3), psum (No-1, nline, 3), spectra (No-1, nline, 2)
    and, with full=True, pi (Nlf, Nlf, 5, No-1, nline)."""
    arrs = [_f(a, np.complex128) if a is not None else None for a in (pixx0, piyy0, pizz0, piyz0, pizy0)]
    no, nlf = arrs[0].shape[0], arrs[0].shape[1]
    for a in arrs:
        if a is not None and a.shape != (no, nlf, nlf):
            raise ValueError("Pi0 arrays must all be (No, Nlf, Nlf)")
    gz = _f(glfz, np.float64)
    pg = _f(parg, np.int32)
    if gz.shape != (nlf,) or pg.shape != (nlf,):
        raise ValueError("glfz and parg must have Nlf entries")
    nline = len(lines)
    ql = np.array([l[0] for l in lines], dtype=np.float64)
    tl = np.array([l[1] for l in lines], dtype=np.float64)
    qm = np.array([l[2] for l in lines], dtype=np.int32)
    nf = no - 1
    pi11 = np.zeros((3, nf, nline), dtype=np.complex128, order="F")
    psum = np.zeros((3, nf, nline), dtype=np.complex128, order="F")
    spec = np.zeros((2, nf, nline), order="F")
    pif = np.zeros((nlf, nlf, 5, nf, nline), dtype=np.complex128, order="F") if full else None
    p = RpaPhotonParams(no, nlf, nline, device, omin, domega, eta, c0)
    rc = lib().rpa_photon(C.byref(p), _ptr(gz), _ptr(pg), *[_ptr(a) for a in arrs], _ptr(ql), _ptr(tl),
                          _ptr(qm), _ptr(pi11), _ptr(psum), _ptr(spec), _ptr(pif))
    if rc != RPA_OK:
        raise RpaError(rc, lib().rpa_last_error(None).decode())
    out = {"pi11": np.moveaxis(pi11, 0, -1), "psum": np.moveaxis(psum, 0, -1), "spectra": np.moveaxis(spec, 0, -1)}
    if full:
        out["pi"] = pif
    return out


def _check(ctx, rc):
    if rc != RPA_OK:
        msg = lib().rpa_last_error(ctx).decode() if ctx else ""
        raise RpaError(rc, msg)


class Context:
    """One device context (rpa_ctx) with numpy-facing methods."""

    def __init__(self, nki, nband, npwx, nlfd, no, program, precision=RPA_FP64, device=0):
        self.dims = RpaDims(nki, nband, npwx, nlfd, no, program, precision, device)
        h = P()
        rc = lib().rpa_ctx_create(C.byref(self.dims), C.byref(h))
        if rc != RPA_OK:
            raise RpaError(rc, "rpa_ctx_create failed")
        self.h = h
        self.nlf = 0
        self.no = no
        self.program = program

    def close(self):
        if self.h:
            lib().rpa_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(self.h, rc)

    def set_gvectors(self, mill, kc):
        mill = np.ascontiguousarray(mill, dtype=np.int32)
        kcf = _f(kc, np.float64)
        self._c(lib().rpa_set_gvectors(self.h, len(mill), _ptr(mill), _ptr(kcf)))

    def upload_basis(self, ik, mill, evc):
        """evc: (ngk, nband) complex128 (the file's band records as columns)."""
        mill = np.ascontiguousarray(mill, dtype=np.int32)
        e = np.asfortranarray(evc, dtype=np.complex128)
        self._c(lib().rpa_upload_basis(self.h, ik, len(mill), _ptr(mill), _ptr(e)))

    def set_bands(self, E):
        Ef = _f(E, np.float64)
        self._c(lib().rpa_set_bands(self.h, _ptr(Ef)))

    def set_local_fields(self, lf_idx, parg):
        li = np.ascontiguousarray(lf_idx, dtype=np.int32)
        pg = np.ascontiguousarray(parg, dtype=np.int32)
        self.nlf = len(li)
        self._c(lib().rpa_set_local_fields(self.h, len(li), _ptr(li), _ptr(pg)))

    def set_params(self, **kw):
        p = RpaParams()
        for k, v in kw.items():
            setattr(p, k, v)
        self.params = p
        self._c(lib().rpa_set_params(self.h, C.byref(p)))

    def set_q(self, K1, K2, iG0, ktot, q):
        K1 = np.ascontiguousarray(K1, dtype=np.int32)
        K2 = np.ascontiguousarray(K2, dtype=np.int32)
        iG0 = np.ascontiguousarray(iG0, dtype=np.int32)
        kt = np.ascontiguousarray(ktot, dtype=np.float64)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        self._c(lib().rpa_set_q(self.h, len(K1), _ptr(K1), _ptr(K2), _ptr(iG0), _ptr(kt), _ptr(qq)))

    def set_symmetry(self, R, nsym):
        Rf = _f(R, np.float64)
        self._c(lib().rpa_set_symmetry(self.h, nsym, _ptr(Rf)))

    def set_q_sym(self, K1, K2, iG0, R1, R2, ktot, q):
        K1, K2, iG0, R1, R2 = (np.ascontiguousarray(x, dtype=np.int32) for x in (K1, K2, iG0, R1, R2))
        kt = np.ascontiguousarray(ktot, dtype=np.float64)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        self._c(lib().rpa_set_q_sym(self.h, len(K1), _ptr(K1), _ptr(K2), _ptr(iG0), _ptr(R1), _ptr(R2), _ptr(kt),
                                    _ptr(qq)))

    # ---- hot path
    def vertices(self):
        self._c(lib().rpa_vertices(self.h))

    def accumulate_s0(self, want=True):
        nlf = self.nlf
        S0 = Q = None
        if want:
            if self.program == RPA_SLOSS:
                S0 = np.zeros((self.no, nlf, nlf), order="F")
            else:
                S0 = np.zeros((2 * self.no + 1, nlf, nlf), dtype=np.complex128, order="F")
                Q = np.zeros((nlf, nlf), dtype=np.complex128, order="F")
        self._c(lib().rpa_accumulate_s0(self.h, _ptr(S0), _ptr(Q)))
        return S0, Q

    def kk(self, want=True):
        out = np.zeros((self.no - 1, self.nlf, self.nlf), dtype=np.complex128, order="F") if want else None
        self._c(lib().rpa_kk(self.h, _ptr(out)))
        return out

    def coulomb(self):
        V = np.zeros((self.nlf, self.nlf), order="F")
        self._c(lib().rpa_coulomb(self.h, _ptr(V)))
        return V

    def dyson_w(self, want=True):
        shp = (self.no - 1, self.nlf, self.nlf)
        outs = [np.zeros(shp, dtype=np.complex128, order="F") for _ in range(3)] if want else [None] * 3
        self._c(lib().rpa_dyson_w(self.h, *[_ptr(o) for o in outs]))
        return outs

    def epilogue(self):
        nf = self.no - 1
        loss = np.zeros(nf)
        pi11 = np.zeros(nf, dtype=np.complex128)
        sig = np.zeros((2, nf))
        self._c(lib().rpa_epilogue(self.h, _ptr(loss), _ptr(pi11), _ptr(sig)))
        return loss, pi11, sig

    def run(self, graph=False):
        self._c((lib().rpa_run_graph if graph else lib().rpa_run)(self.h))
        return lib().rpa_last_run_ms(self.h)

    def set_profiling(self, on=True):
        self._c(lib().rpa_set_profiling(self.h, 1 if on else 0))

    def stage_ms(self):
        ms = np.zeros(6)
        self._c(lib().rpa_stage_ms(self.h, _ptr(ms), 6))
        return dict(zip(STAGES, ms.tolist()))

    def result(self, name):
        """A result of the last run, downloaded without recomputation."""
        nlf, nf = self.nlf, self.no - 1
        shapes = {
            "S0": ((self.no, nlf, nlf), np.float64) if self.program == RPA_SLOSS
            else ((2 * self.no + 1, nlf, nlf), np.complex128),
            "QEFF": ((nlf, nlf), np.complex128), "CHI0": ((nf, nlf, nlf), np.complex128),
            "V": ((nlf, nlf), np.float64), "EPSINV": ((nf, nlf, nlf), np.complex128),
            "CHI": ((nf, nlf, nlf), np.complex128), "W": ((nf, nlf, nlf), np.complex128),
            "LOSS": ((nf,), np.float64), "PI11": ((nf,), np.complex128), "SIG": ((nf, 2), np.float64),
        }
        shp, dt = shapes[name]
        out = np.zeros(shp, dtype=dt, order="F")
        self._c(lib().rpa_get_result(self.h, RES[name], _ptr(out)))
        return out

    def load_s0(self, S0, qeff=None):
        dt = np.float64 if self.program == RPA_SLOSS else np.complex128
        s = np.asfortranarray(S0, dtype=dt)
        q = None if qeff is None else np.asfortranarray(qeff, dtype=np.complex128)
        self._c(lib().rpa_load_s0(self.h, _ptr(s), _ptr(q)))

    def w_kk(self):
        nf = self.no - 1
        w1, w2 = np.zeros(nf), np.zeros(nf)
        self._c(lib().rpa_w_kk(self.h, _ptr(w1), _ptr(w2)))
        return w1, w2

    def gamma_w(self):
        """GW propagators (Gammap, Gammam) (no-1, nlf, nlf) of the resident W."""
        shape = (self.no - 1, self.nlf, self.nlf)
        gp = np.zeros(shape, dtype=np.complex128, order="F")
        gm = np.zeros(shape, dtype=np.complex128, order="F")
        self._c(lib().rpa_gamma_w(self.h, _ptr(gp), _ptr(gm)))
        return gp, gm

    def set_autotune(self, on=True):
        self._c(lib().rpa_set_autotune(self.h, 1 if on else 0))

    def batch_set(self, points, ktot):
        """points: dicts with lf_idx, parg, K1, K2, iG0 and q (host.Runner.q_point)."""
        self.batch_nlf = [len(p["lf_idx"]) for p in points]
        nl = np.ascontiguousarray(self.batch_nlf, dtype=np.int32)
        li = np.ascontiguousarray(np.concatenate([p["lf_idx"] for p in points]), dtype=np.int32)
        pg = np.ascontiguousarray(np.concatenate([p["parg"] for p in points]), dtype=np.int32)
        K1 = np.ascontiguousarray(np.concatenate([p["K1"] for p in points]), dtype=np.int32)
        K2 = np.ascontiguousarray(np.concatenate([p["K2"] for p in points]), dtype=np.int32)
        iG0 = np.ascontiguousarray(np.concatenate([p["iG0"] for p in points]), dtype=np.int32)
        kt = np.ascontiguousarray(ktot, dtype=np.float64)
        qq = np.ascontiguousarray(np.stack([p["q"] for p in points]), dtype=np.float64)
        ntot = len(points[0]["K1"])
        if "R1" in points[0]:
            R1 = np.ascontiguousarray(np.concatenate([p["R1"] for p in points]), dtype=np.int32)
            R2 = np.ascontiguousarray(np.concatenate([p["R2"] for p in points]), dtype=np.int32)
            self._c(lib().rpa_batch_set_sym(self.h, len(points), _ptr(nl), _ptr(li), _ptr(pg), ntot, _ptr(K1),
                                            _ptr(K2), _ptr(iG0), _ptr(R1), _ptr(R2), _ptr(kt), _ptr(qq)))
            return
        self._c(lib().rpa_batch_set(self.h, len(points), _ptr(nl), _ptr(li), _ptr(pg), ntot, _ptr(K1), _ptr(K2),
                                    _ptr(iG0), _ptr(kt), _ptr(qq)))

    def batch_run(self):
        self._c(lib().rpa_batch_run(self.h))
        return lib().rpa_last_run_ms(self.h)

    def batch_result(self, iq, name):
        nlf, nf = self.batch_nlf[iq], self.no - 1
        shapes = {"V": ((nlf, nlf), np.float64), "EPSINV": ((nf, nlf, nlf), np.complex128),
                  "W": ((nf, nlf, nlf), np.complex128), "LOSS": ((nf,), np.float64),
                  "QEFF": ((nlf, nlf), np.complex128), "PI11": ((nf,), np.complex128), "SIG": ((nf, 2), np.float64)}
        shp, dt = shapes[name]
        out = np.zeros(shp, dtype=dt, order="F")
        self._c(lib().rpa_batch_result(self.h, iq, RES[name], _ptr(out)))
        return out

    def get_vertices(self, both=False):
        n = C.c_longlong()
        self._c(lib().rpa_get_vertices(self.h, None, None, C.byref(n)))
        r1 = np.zeros((self.nlf, n.value), dtype=np.complex128, order="F")
        r2 = np.zeros_like(r1) if both else None
        self._c(lib().rpa_get_vertices(self.h, _ptr(r1), _ptr(r2), C.byref(n)))
        return r1, r2
