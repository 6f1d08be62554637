"""Seeded synthetic inputs in Quantum ESPRESSO format for the RPA response path.

The reference reads a QE save directory and these text outputs:

* ``charge-density.dat``  -- Fortran unformatted: (gamma_only, ngm, nspin),
  (b1, b2, b3), mill(3, ngm)                      (Optabs_qe6.f90:1139-1157)
* ``wfc{ik}.dat``         -- (ik, xk, ispin, gamma_only, scalef),
  (ngw, igwx, npol, nbnd), (b1, b2, b3), mill(3, igwx), then one complex128
  record per band                                 (Optabs_qe6.f90:1213-1234)
* ``*.band``               -- bands.x filband: header, per k ``(10x,3f10.6)``
  and energies ``(10f8.3)`` eV                      (Optabs.f90:1333-1339)
* ``scf.out``              -- the lines PointR (Chi_RPA/PointR.f90:41-91) and
  loadKC (Optabs.f90:652-680) scan for.
* ``config.in``            -- the namelists of Optabs.f90:14-142 / Sloss.f90:39-172.

Construction: hexagonal 2D cell (a, c), unshifted N x N x 1 Monkhorst-Pack grid
in crystal coordinates [0, 1), two-band nearest-neighbour tight-binding bands
(t = 2.7 eV, gap 1 eV) replicated with rigid 3 eV shifts, E_F mid-gap, and
coefficients (N(0,1) + i N(0,1)) exp(-|k+G|^2/8) orthonormalised per k by
modified Gram-Schmidt.  Everything is a pure function of the seed.
"""

from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

HARTREE = 2.0 * 13.6056923  # eV, the drivers' value (Optabs.f90:46, Sloss.f90:71)


@dataclasses.dataclass
class SynthConfig:
    """One synthetic system.  Lengths in bohr, energies in eV."""

    name: str
    a: float = 4.65
    c: float = 20.0
    nk: int = 12                 # N x N x 1 grid
    nval: int = 4
    ncond: int = 4
    ngk_target: float = 285.0    # mean plane waves per k (sets the cutoff)
    lf: str = "z"
    nlf_target: int = 11
    no: int = 201                # frequencies
    omax_ev: float = 20.0
    eta_ev: float = 0.05
    iq: int = 2                  # q = (iq-1) k_min (on-grid, Optabs.f90:743-773)
    seed: int = 20261202
    t_hop: float = 2.7
    gap: float = 1.0
    shift: float = 3.0

    @property
    def nband(self) -> int:
        return self.nval + self.ncond


# The configurations the path is exercised on (sizes from the workload table).
CONFIGS = {
    "tiny": SynthConfig("tiny", nk=4, nval=2, ncond=2, ngk_target=90.0, nlf_target=5,
                        no=41, omax_ev=16.0, seed=20261202 + 100),
    "cfg0": SynthConfig("cfg0", nk=12, nval=4, ncond=4, ngk_target=285.0, nlf_target=11,
                        no=201, omax_ev=20.0, seed=20261202 + 0),
    "cfg1": SynthConfig("cfg1", nk=64, nval=8, ncond=12, ngk_target=1500.0, nlf_target=31,
                        no=701, omax_ev=30.0, seed=20261202 + 1),
    "cfg2": SynthConfig("cfg2", nk=64, nval=8, ncond=12, ngk_target=1500.0, nlf_target=51,
                        no=701, omax_ev=30.0, seed=20261202 + 2),
    # batched W(q, omega): 100 q, per-q N_G from 21 to 61 chosen by |q+G|
    # (host.local_fields_by_qg); nlf_target is the largest N_G
    "cfg3": SynthConfig("cfg3", nk=48, nval=8, ncond=8, ngk_target=1000.0, lf="xyz", nlf_target=61,
                        no=700, omax_ev=25.0, eta_ev=0.08, seed=20261202 + 3),
    # spectral function + KK (Optabs route): 2 no + 1 = 2001 bins over 0-40 eV,
    # KK to no - 1 = 999 frequencies; a bilayer cell treated as one 2D cell
    # (a = 4.68 bohr, c = 40 bohr); 24 q, per-q N_G from 61 to 81 by |q+G|
    "cfg4": SynthConfig("cfg4", a=4.68, c=40.0, nk=36, nval=16, ncond=24, ngk_target=3200.0, lf="xyz",
                        nlf_target=81, no=1000, omax_ev=40.0, eta_ev=0.05, seed=20261202 + 4),
}

# q points of the batched workloads: the drivers' q loop (q = (iq-1) k_min,
# Sloss.f90:292, 595-597), iq = 2 .. 101 for batched W and 2 .. 25 for the
# Optabs route, N_G rising linearly over the q points.
BATCHED_Q = {"cfg3": dict(iq=list(range(2, 102)), nlf_min=21, nlf_max=61, program="sloss"),
             "cfg4": dict(iq=list(range(2, 26)), nlf_min=61, nlf_max=81, program="optabs")}


def batched_q_points(name: str, nq: Optional[int] = None):
    """(iq, N_G) pairs of a batched-W workload, the first nq of them (spread
    over the whole N_G range when nq is smaller than the list)."""
    b = BATCHED_Q[name]
    iqs = b["iq"]
    n = len(iqs)
    ngs = [b["nlf_min"] + int(round((b["nlf_max"] - b["nlf_min"]) * j / max(n - 1, 1))) for j in range(n)]
    pts = list(zip(iqs, ngs))
    if nq is None or nq >= n:
        return pts
    pick = sorted({int(round(j * (n - 1) / max(nq - 1, 1))) for j in range(nq)})
    return [pts[j] for j in pick]


def reciprocal_axes(a: float, c: float, decimals: int | None = 6) -> np.ndarray:
    """KC(3,3) in 2pi/a units, rounded to the 6 decimals scf.out carries
    (decimals=None: unrounded, so that the hexagonal rotations map the
    lattice onto itself exactly).

    KC[:, m] is b_m (cartesian), as loadKC stores KC(1:3, j) (Optabs.f90:668).
    """
    b1 = np.array([1.0, -1.0 / math.sqrt(3.0), 0.0])
    b2 = np.array([0.0, 2.0 / math.sqrt(3.0), 0.0])
    b3 = np.array([0.0, 0.0, a / c])
    kc = np.stack([b1, b2, b3], axis=1)
    return kc if decimals is None else np.round(kc, decimals)


def cell_volume(a: float, c: float) -> float:
    return math.sqrt(3.0) / 2.0 * a * a * c


def gcart(mill: np.ndarray, kc: np.ndarray) -> np.ndarray:
    """G(n) = sum_m KC(n,m) mill(m), accumulated left to right (Optabs_qe6.f90:1162-1167)."""
    mill = np.asarray(mill, dtype=np.float64)
    g = np.zeros(mill.shape, dtype=np.float64)
    for b in range(3):
        g = g + kc[:, b][None, :] * mill[:, b:b + 1]
    return g


def density_gvectors(kc: np.ndarray, a: float, gmax_au: float) -> np.ndarray:
    """Miller list with |G| <= gmax (a.u.), sorted by |G|^2 then Miller (G(1) = 0)."""
    gcar = 2.0 * math.pi / a
    bl = np.linalg.norm(kc, axis=0) * gcar
    nmax = [int(math.ceil(gmax_au / bl[i])) + 1 for i in range(3)]
    r = [np.arange(-n, n + 1) for n in nmax]
    mm = np.stack(np.meshgrid(*r, indexing="ij"), axis=-1).reshape(-1, 3)
    g = gcart(mm, kc) * gcar
    g2 = np.einsum("ij,ij->i", g, g)
    keep = g2 <= gmax_au * gmax_au + 1e-12
    mm, g2 = mm[keep], g2[keep]
    order = np.lexsort((mm[:, 2], mm[:, 1], mm[:, 0], np.round(g2, 10)))
    mm = mm[order].astype(np.int32)
    assert (mm[0] == 0).all()
    return mm


def gk_sort(k: np.ndarray, g: np.ndarray, gcutw: float) -> np.ndarray:
    """|k+G|^2 <= gcutw in 2pi/a units, ordered by |k+G|^2 (ties by index).

    Same rule as Modules/qe_gksort.f90:44-94 (gk_sort + hpsort_eps, eps8).
    Returns 0-based indices into g.
    """
    kg = k[None, :] + g
    q = (kg[:, 0] * kg[:, 0] + kg[:, 1] * kg[:, 1]) + kg[:, 2] * kg[:, 2]
    q = np.where(q <= 1e-8, 0.0, q)
    idx = np.nonzero(q <= gcutw)[0]
    qs = q[idx]
    # group values closer than eps8, then order groups by value, members by index
    order = np.lexsort((idx, qs))
    qs_sorted = qs[order]
    keys = np.zeros(len(order), dtype=np.int64)
    if len(order):
        gap = np.diff(qs_sorted) >= 1e-8
        keys[1:] = np.cumsum(gap)
    final = np.lexsort((idx[order], keys))
    return idx[order][final]


def tb_bands(kcrys: np.ndarray, cfg: SynthConfig) -> np.ndarray:
    """Band energies (nk, nband) in eV, ascending per k."""
    ph1 = np.exp(2j * math.pi * kcrys[:, 0])
    ph2 = np.exp(2j * math.pi * kcrys[:, 1])
    f = np.abs(1.0 + ph1 + ph2)
    ep = np.sqrt((cfg.gap / 2.0) ** 2 + (cfg.t_hop * f) ** 2)
    vals = [-ep - cfg.shift * j for j in range(cfg.nval)][::-1]
    conds = [ep + cfg.shift * j for j in range(cfg.ncond)]
    return np.stack(vals + conds, axis=1)


@dataclasses.dataclass
class SynthSystem:
    cfg: SynthConfig
    kc: np.ndarray          # (3,3)
    mill_g: np.ndarray      # (ngm, 3) int32
    kcart: np.ndarray       # (nk, 3) 2pi/a units, 6 decimals
    e_ev: np.ndarray        # (nk, nband) eV, 3 decimals
    ngk: np.ndarray         # (nk,)
    npwx: int
    mill_k: np.ndarray      # (nk, npwx, 3) int32 (zero padded)
    evc: np.ndarray         # (nk, nband, npwx) complex128 (zero padded)
    gcutw: float            # 2pi/a units squared
    ecutwfc_ry: float
    vcell: float

    @property
    def gcar(self) -> float:
        return 2.0 * math.pi / self.cfg.a


def make_system(cfg: SynthConfig, with_coeffs: bool = True) -> SynthSystem:
    a, c = cfg.a, cfg.c
    kc = reciprocal_axes(a, c)
    gcar = 2.0 * math.pi / a
    vol = cell_volume(a, c)
    # plane-wave sphere radius (a.u.) that gives ngk_target on average
    kmax = (cfg.ngk_target * 8.0 * math.pi ** 3 / (vol * 4.0 * math.pi / 3.0)) ** (1.0 / 3.0)
    gcutw = (kmax / gcar) ** 2
    mill_g = density_gvectors(kc, a, 2.0 * kmax + 2.0 * gcar)
    g = gcart(mill_g, kc)

    n = cfg.nk
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    kcrys = np.stack([ii.ravel() / n, jj.ravel() / n, np.zeros(n * n)], axis=1)
    kcart = np.round(gcart(kcrys, kc), 6)
    e_ev = np.round(tb_bands(kcrys, cfg), 3)

    lists = [gk_sort(kcart[ik], g, gcutw) for ik in range(n * n)]
    ngk = np.array([len(l) for l in lists], dtype=np.int32)
    npwx = int(((ngk.max() + 63) // 64) * 64)
    mill_k = np.zeros((n * n, npwx, 3), dtype=np.int32)
    for ik, l in enumerate(lists):
        mill_k[ik, :len(l)] = mill_g[l]
    evc = np.zeros((n * n, cfg.nband, npwx), dtype=np.complex128)
    if with_coeffs:
        for ik in range(n * n):
            rng = np.random.default_rng([cfg.seed, ik])
            npw = ngk[ik]
            kg = (kcart[ik][None, :] + g[lists[ik]]) * gcar
            env = np.exp(-np.einsum("ij,ij->i", kg, kg) / 8.0)
            z = (rng.standard_normal((cfg.nband, npw)) + 1j * rng.standard_normal((cfg.nband, npw)))
            z *= env[None, :]
            for b in range(cfg.nband):  # modified Gram-Schmidt
                v = z[b]
                for p in range(b):
                    v = v - np.vdot(z[p], v) * z[p]
                z[b] = v / math.sqrt(np.vdot(v, v).real)
            evc[ik, :, :npw] = z
    return SynthSystem(cfg=cfg, kc=kc, mill_g=mill_g, kcart=kcart, e_ev=e_ev, ngk=ngk,
                       npwx=npwx, mill_k=mill_k, evc=evc, gcutw=gcutw,
                       ecutwfc_ry=2.0 * 0.5 * kmax * kmax, vcell=vol)


def local_field_ecut(sys: SynthSystem) -> float:
    """Ecut (Hartree) that selects cfg.nlf_target local fields with cfg.lf."""
    g = gcart(sys.mill_g, sys.kc)
    gcar = sys.gcar
    if sys.cfg.lf == "z":
        sel = (g[:, 0] == 0.0) & (g[:, 1] == 0.0)
        eref = gcar * gcar * (g[sel, 2] * g[sel, 2]) / 2.0
    else:
        eref = gcar * gcar * ((g[:, 0] * g[:, 0] + g[:, 1] * g[:, 1]) + g[:, 2] * g[:, 2]) / 2.0
    vals = np.unique(eref)
    counts = np.array([(eref <= v).sum() for v in vals])
    j = int(np.argmin(np.abs(counts - sys.cfg.nlf_target)))
    hi = vals[j + 1] if j + 1 < len(vals) else vals[j] * 2
    return float(0.5 * (vals[j] + hi))


# --------------------------------------------------------------------------
# Writers (byte-exact Fortran sequential records, fixed-format text)

def _rec(f, payload: bytes) -> None:
    n = np.int32(len(payload)).tobytes()
    f.write(n)
    f.write(payload)
    f.write(n)


def write_charge_density(path: str, sys: SynthSystem) -> None:
    b = (sys.kc * sys.gcar).T.astype(np.float64)  # b1, b2, b3 in a.u.
    with open(path, "wb") as f:
        _rec(f, np.int32(0).tobytes() + np.int32(len(sys.mill_g)).tobytes() + np.int32(1).tobytes())
        _rec(f, b.tobytes())
        _rec(f, np.ascontiguousarray(sys.mill_g, dtype=np.int32).tobytes())


def write_wfc(path: str, sys: SynthSystem, ik: int) -> None:
    npw = int(sys.ngk[ik])
    b = (sys.kc * sys.gcar).T.astype(np.float64)
    xk = (sys.kcart[ik] * sys.gcar).astype(np.float64)
    with open(path, "wb") as f:
        _rec(f, np.int32(ik + 1).tobytes() + xk.tobytes() + np.int32(1).tobytes()
             + np.int32(0).tobytes() + np.float64(1.0).tobytes())
        _rec(f, np.array([npw, npw, 1, sys.cfg.nband], dtype=np.int32).tobytes())
        _rec(f, b.tobytes())
        _rec(f, np.ascontiguousarray(sys.mill_k[ik, :npw], dtype=np.int32).tobytes())
        for bnd in range(sys.cfg.nband):
            _rec(f, np.ascontiguousarray(sys.evc[ik, bnd, :npw]).tobytes())


def write_band_file(path: str, sys: SynthSystem) -> None:
    nk, nb = sys.e_ev.shape
    with open(path, "w") as f:
        f.write(f" &plot nbnd={nb:4d}, nks={nk:6d} /\n")
        for ik in range(nk):
            k = sys.kcart[ik]
            f.write(" " * 10 + "".join(f"{x:10.6f}" for x in k) + "\n")
            e = sys.e_ev[ik]
            for s in range(0, nb, 10):
                f.write("".join(f"{x:8.3f}" for x in e[s:s + 10]) + "\n")


def write_scf_out(path: str, sys: SynthSystem, ops=None) -> None:
    """scf.out with the reciprocal axes (6 decimals) and the point operations:
    the identity, or ops = (R, nsym) with R[op] cartesian, each written as
    pw.x does (the crystal form, then the cartesian rows to 7 decimals, the
    values from column 20 as PointR reads them)."""
    kc = sys.kc
    lines = [
        "",
        "     Program PWSCF (synthetic output for the RPA drop-in)",
        "",
        f"     lattice parameter (alat)  = {sys.cfg.a:12.4f}  a.u.",
        "",
        "     reciprocal axes: (cart. coord. in units 2 pi/alat)",
    ]
    for j in range(3):
        lines.append(" " * 15 + f"b({j + 1}) = (" + "".join(f"{kc[i, j]:10.6f}" for i in range(3)) + " )  ")
    lines += [
        "",
        "     atomic species   valence    mass     pseudopotential",
        "        X1            14.00    95.95000     X1( 1.00)",
        "        X2             6.00    32.06000     X2( 1.00)",
        "",
    ]
    if ops is None:
        R, nsym = np.eye(3)[None], 1
    else:
        R, nsym = np.asarray(ops[0]), int(ops[1])
    kci = np.linalg.inv(kc)
    lines += [f"     {nsym} Sym. Ops. (no inversion) found", "",
              "                                    s                        frac. trans.", ""]
    for op in range(nsym):
        cr = np.rint(kci @ R[op] @ kc).astype(int).T  # crystal form (rows of the direct-lattice operation)
        lines += [f"      isym = {op + 1:2d}     {'identity' if op == 0 else 'operation'}", ""]
        for i in range(3):
            head = f" cryst.   s({op + 1:2d}) = (" if i == 0 else " " * 18 + "("
            lines.append(head + "".join(f"{cr[i, j]:6d}     " for j in range(3)) + " )")
        lines.append("")
        for i in range(3):
            head = f" cart.    s({op + 1:2d}) = (" if i == 0 else " " * 18 + "("
            lines.append(head + "".join(f"{R[op][i, j]:11.7f}" for j in range(3)) + " )")
        lines.append("")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def xml_crystal_ops(kc: np.ndarray, R: np.ndarray, nsym: int) -> np.ndarray:
    """Integer crystal forms s = KC^-1 R KC of cartesian operations: s acts on
    reciprocal crystal components, the way pw.x rotates k-points."""
    kci = np.linalg.inv(kc)
    return np.stack([np.rint(kci @ R[op] @ kc) for op in range(nsym)])


def write_data_file_schema(path: str, sys: SynthSystem, ops=None, noinv: bool = True,
                           force_symmorphic: bool = True, efermi_ha: float = 0.0) -> None:
    """A pw.x data-file-schema.xml with the elements io_xml reads: the input
    cell and Monkhorst-Pack grid, symmetry flags and smearing, the output
    symmetries (each <rotation> as three lines, the columns of s), the
    reciprocal lattice (the scf.out axes sys.kc), and per IBZ k-point its
    cartesian k (2pi/alat), npw, eigenvalues (Hartree) and occupations, five
    values per line."""
    cfg = sys.cfg
    kc = sys.kc
    at = np.linalg.inv(kc).T * cfg.a  # direct lattice vectors (columns), bohr
    if ops is None:
        R, nsym = np.eye(3)[None], 1
    else:
        R, nsym = np.asarray(ops[0]), int(ops[1])
    S = xml_crystal_ops(kc, R, nsym)
    nk, nb = sys.e_ev.shape
    e_ha = sys.e_ev / HARTREE
    occ = (e_ha < efermi_ha).astype(float)

    def vec(v):
        return " ".join(repr(float(x)) for x in v)

    def block(vals):
        out = []
        for i in range(0, len(vals), 5):
            out.append("  " + " ".join(f"{x:.15e}" for x in vals[i:i + 5]))
        return out

    L = ['<?xml version="1.0" encoding="UTF-8"?>',
         '<qes:espresso Units="Hartree atomic units">',
         '  <input>',
         f'    <atomic_structure nat="2" alat="{cfg.a!r}" bravais_index="4">',
         '      <cell>']
    L += [f'        <a{i + 1}>{vec(at[:, i])}</a{i + 1}>' for i in range(3)]
    L += ['      </cell>',
          '    </atomic_structure>',
          '    <basis>',
          '      <ecutwfc>1.0e1</ecutwfc>',
          '      <ecutrho>4.0e1</ecutrho>',
          '    </basis>',
          '    <bands>',
          '      <smearing degauss="0.0">gaussian</smearing>',
          '      <occupations>fixed</occupations>',
          '    </bands>',
          '    <k_points_IBZ>',
          f'      <monkhorst_pack nk1="{cfg.nk}" nk2="{cfg.nk}" nk3="1" k1="0" k2="0" k3="0">Monkhorst-Pack'
          '</monkhorst_pack>',
          '    </k_points_IBZ>',
          '    <symmetry_flags>',
          '      <nosym>false</nosym>',
          f'      <noinv>{"true" if noinv else "false"}</noinv>',
          '      <no_t_rev>false</no_t_rev>',
          f'      <force_symmorphic>{"true" if force_symmorphic else "false"}</force_symmorphic>',
          '    </symmetry_flags>',
          '  </input>',
          '  <output>',
          '    <symmetries>',
          f'      <nsym>{nsym}</nsym>',
          f'      <nrot>{nsym}</nrot>',
          '      <space_group>0</space_group>']
    for op in range(nsym):
        L += ['      <symmetry>',
              '        <info name="crystal_symmetry">crystal_symmetry</info>',
              '        <rotation rank="2" dims="3 3" order="F">']
        L += ["  " + " ".join(f"{S[op][i, c]:.15e}" for i in range(3)) for c in range(3)]
        L += ['        </rotation>',
              '      </symmetry>']
    L += ['    </symmetries>',
          '    <basis_set>',
          '      <ecutwfc>1.0e1</ecutwfc>',
          '      <ecutrho>4.0e1</ecutrho>',
          '      <reciprocal_lattice>']
    L += [f'        <b{j + 1}>{vec(kc[:, j])}</b{j + 1}>' for j in range(3)]
    L += ['      </reciprocal_lattice>',
          '    </basis_set>',
          '    <band_structure>',
          f'      <nbnd>{nb}</nbnd>',
          f'      <nelec>{2.0 * cfg.nval!r}</nelec>',
          f'      <fermi_energy>{efermi_ha!r}</fermi_energy>',
          '      <starting_k_points>',
          f'        <monkhorst_pack nk1="{cfg.nk}" nk2="{cfg.nk}" nk3="1" k1="0" k2="0" k3="0">Monkhorst-Pack'
          '</monkhorst_pack>',
          '      </starting_k_points>',
          f'      <nks>{nk}</nks>']
    for ik in range(nk):
        L += ['      <ks_energies>',
              f'        <k_point weight="{2.0 / nk!r}">{vec(sys.kcart[ik])}</k_point>',
              f'        <npw>{int(sys.ngk[ik])}</npw>',
              f'        <eigenvalues size="{nb}">']
        L += block(e_ha[ik])
        L += ['        </eigenvalues>',
              f'        <occupations size="{nb}">']
        L += block(occ[ik])
        L += ['        </occupations>',
              '      </ks_energies>']
    L += ['    </band_structure>',
          '  </output>',
          '</qes:espresso>']
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_save(dirpath: str, sys: SynthSystem) -> None:
    os.makedirs(dirpath, exist_ok=True)
    write_charge_density(os.path.join(dirpath, "charge-density.dat"), sys)
    for ik in range(len(sys.ngk)):
        write_wfc(os.path.join(dirpath, f"wfc{ik + 1}.dat"), sys, ik)


def write_config(path: str, sys: SynthSystem, program: str, savedir: str,
                 band_file: str = "synthetic.band", scf_file: str = "scf.out",
                 extra: Optional[dict] = None) -> None:
    """config.in for the Optabs-style (program='optabs') or Sloss-style driver."""
    cfg = sys.cfg
    nk = len(sys.ngk)
    ecut_lf = local_field_ecut(sys)
    ngd = int(sys.ngk.min())
    x = dict(extra or {})
    if program == "optabs":
        text = f"""&directories
 rundir = './'
 savedir = '{savedir}'
 scf_file = '{scf_file}'
 band_file = '{band_file}'
/
&system
 lf = '{cfg.lf}'
 pol = '{x.get("pol", "xx")}'
 calc = {x.get("calc", 3)}
 jump = 1
 omin = {x.get("omin", 0.0):.17e}
 omax = {cfg.omax_ev:.17e}
 qmin = {cfg.iq}
 qmax = {cfg.iq}
 eta_intra = {x.get("eta_intra", 0.01):.17e}
 eta_inter = {cfg.eta_ev:.17e}
 Lor_cut = {x.get("lor_cut", 1.0e-3):.17e}
 df_cut = {x.get("df_cut", 1.0e-3):.17e}
 vertex_map = '{x.get("vertex_map", "direct")}'
 precision = '{x.get("precision", "fp64")}'
/
&config
 NG = {len(sys.mill_g)}
 NGd = {ngd}
 NkI = {nk}
 Nband = {cfg.nband}
 Nval = {cfg.nval}
 NelQE = {2 * cfg.nval}
 no = {cfg.no}
 Nlfd = {max(cfg.nlf_target, 1) + 10}
/
&parameters
 Efermi = 0.0
 dGW = {x.get("dgw", 0.0):.17e}
 eps = 1.0e-4
 T = {x.get("T", 0.0):.17e}
 Ecut = {ecut_lf:.17e}
 a0 = {cfg.a:.17e}
 c0 = {cfg.c:.17e}
 Vcell = {sys.vcell:.17e}
/
&parallel
 Nthreads = 1
/
"""
    else:
        text = f"""&directories
 rundir = './'
 savedir = '{savedir}'
 scf_file = '{scf_file}'
 band_file = '{band_file}'
/
&system
 lf = '{cfg.lf}'
 jump = 1
 omin = 0.0
 omax = {cfg.omax_ev / HARTREE:.17e}
 qmin = {cfg.iq}
 qmax = {cfg.iq}
 vertex_map = '{x.get("vertex_map", "direct")}'
 precision = '{x.get("precision", "fp64")}'
/
&config
 NkI = {nk}
 Nband = {cfg.nband}
 Nocc = {cfg.nval}
 NelQE = {2 * cfg.nval}
 NGd = {ngd}
 NG = {len(sys.mill_g)}
 no = {cfg.no}
 nq = 1
 Nlfd = {max(cfg.nlf_target, 1) + 10}
/
&parameters
 Efermi = 0.0
 a0 = {cfg.a:.17e}
 c0 = {cfg.c:.17e}
 eps = 1.0e-4
 T = 0.0
 eta = {cfg.eta_ev:.17e}
 Ecut = {ecut_lf:.17e}
 Vcell = {sys.vcell:.17e}
/
&parallel
 Nthreads = 1
/
"""
    with open(path, "w") as f:
        f.write(text)


def write_all(rundir: str, sys: SynthSystem, program: str, extra: Optional[dict] = None, ops=None) -> None:
    """Complete run directory: save/, scf.out, synthetic.band, config.in
    (ops = (R, nsym): the point operations of an irreducible k list)."""
    os.makedirs(rundir, exist_ok=True)
    write_save(os.path.join(rundir, "save"), sys)
    write_scf_out(os.path.join(rundir, "scf.out"), sys, ops)
    write_band_file(os.path.join(rundir, "synthetic.band"), sys)
    write_config(os.path.join(rundir, "config.in"), sys, program, "./save/", extra=extra)


# --------------------------------------------------------------------------
# Point-group test systems (Nsym > 1)

def hex_point_ops(kind: str) -> np.ndarray:
    """Cartesian operations of the hexagonal lattice, R(48,3,3) with the
    identity first: 'c2' (identity, 180-degree rotation about z), 'c3', 'c6'
    (rotations about z), 'c3v' (c3 and the mirrors y -> -y and its images)."""
    def rot(t):
        c, s = math.cos(t), math.sin(t)
        return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    my = np.diag([1.0, -1.0, 1.0])
    if kind == "c2":
        ops = [rot(0.0), rot(math.pi)]
    elif kind == "c3":
        ops = [rot(2.0 * math.pi * j / 3.0) for j in range(3)]
    elif kind == "c6":
        ops = [rot(math.pi * j / 3.0) for j in range(6)]
    elif kind == "c3v":
        r = [rot(2.0 * math.pi * j / 3.0) for j in range(3)]
        ops = r + [x @ my for x in r]
    else:
        raise ValueError(kind)
    R = np.zeros((48, 3, 3))
    for i, o in enumerate(ops):
        R[i] = np.where(np.abs(o) < 1e-15, 0.0, o)
    return R, len(ops)


def make_symmetric_pair(cfg: SynthConfig, kind: str, npw_target: int, ngd_target: int, seed_shift: int = 0):
    """Two descriptions of one physical system for testing the unfolding:

    ibz: the irreducible k-points of an N x N grid under the point group
         `kind` (the first point of each orbit in grid order), seeded
         orthonormal coefficients over the first npw global G, energies from
         the tight-binding bands of those points;
    fz:  every point of R(i) kI(j) (the unfolded zone, in the order genFBZ
         gives it) as a plain Nsym = 1 list, with coefficients C_Rk(R G) =
         C_k(G) and energies E(Rk) = E(k).

    npw and NGd are rounded up to counts whose G windows every operation maps
    onto themselves, so rotations keep both coefficient windows.  Returns (ibz, fz, R, nsym, kinfo) with
    kinfo[j] = (op, IBZ point) of fz point j, both 0-based."""
    R, nsym = hex_point_ops(kind)
    a, c = cfg.a, cfg.c
    kc = reciprocal_axes(a, c, decimals=None)
    gcar = 2.0 * math.pi / a
    vol = cell_volume(a, c)
    kmax = (npw_target * 8.0 * math.pi ** 3 / (vol * 4.0 * math.pi / 3.0)) ** (1.0 / 3.0)
    mill_g = density_gvectors(kc, a, 2.0 * kmax + 2.0 * gcar)
    g = gcart(mill_g, kc)
    g2 = np.einsum("ij,ij->i", g, g)
    index = {tuple(m): i for i, m in enumerate(mill_g.tolist())}
    kcinv = np.linalg.inv(kc)

    def rot_index(op, i):  # global index of R(op) G(i), len(g) when outside the list
        m = np.rint(kcinv @ (R[op] @ g[i])).astype(int)
        return index.get(tuple(m.tolist()), len(g))

    def closed(n):  # smallest count >= n whose window every operation maps onto itself
        while any(rot_index(op, i) >= n for op in range(nsym) for i in range(n)):
            n += 1
        return n
    npw, ngd = closed(npw_target), closed(ngd_target)

    n = cfg.nk
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    kcrys = np.stack([ii.ravel() / n, jj.ravel() / n, np.zeros(n * n)], axis=1)
    kcart = np.round(gcart(kcrys, kc), 6)
    covered = np.zeros(n * n, bool)
    reps = []
    for p in range(n * n):
        if covered[p]:
            continue
        reps.append(p)
        for op in range(nsym):
            kr = kcinv @ (R[op] @ kcart[p])
            q = np.rint(np.mod(kr, 1.0) * n).astype(int) % n
            covered[q[0] * n + q[1]] = True
    ki = kcart[reps]
    e_ibz = np.round(tb_bands(kcrys[reps], cfg), 3)
    nb = cfg.nband
    evc_i = np.zeros((len(reps), nb, npw), dtype=np.complex128)
    for j in range(len(reps)):
        rng = np.random.default_rng([cfg.seed + seed_shift, j])
        kg = (ki[j][None, :] + g[:npw]) * gcar
        env = np.exp(-np.einsum("ij,ij->i", kg, kg) / 8.0)
        z = (rng.standard_normal((nb, npw)) + 1j * rng.standard_normal((nb, npw))) * env[None, :]
        for b in range(nb):
            v = z[b]
            for p_ in range(b):
                v = v - np.vdot(z[p_], v) * z[p_]
            z[b] = v / math.sqrt(np.vdot(v, v).real)
        evc_i[j] = z
    # unfolded zone in genFBZ order (operation outer, IBZ point inner; a
    # point within eps of an earlier one is dropped)
    kt = []
    for op in range(nsym):
        for j in range(len(reps)):
            v = R[op] @ ki[j]
            if not any(np.all(np.abs(v - w) <= 1e-4) for w in kt):
                kt.append(v)
    kt = np.array(kt)
    # each point's (operation, IBZ point) by the reference's rule: the first
    # operation, in order, whose inverse takes it to an IBZ point
    # (findKinIBZ / findKQinBZ, Optabs.f90:776-869)
    kinfo = []
    for v in kt:
        hit = None
        for op in range(nsym):
            w = np.linalg.inv(R[op]) @ v
            for j in range(len(reps)):
                if np.all(np.abs(w - ki[j]) <= 1e-4):
                    hit = (op, j)
                    break
            if hit:
                break
        kinfo.append(hit)
    evc_f = np.zeros((len(kt), nb, npw), dtype=np.complex128)
    for t, (op, j) in enumerate(kinfo):
        perm = np.array([rot_index(op, i) for i in range(npw)])
        evc_f[t][:, perm] = evc_i[j]
    e_fz = e_ibz[[j for _, j in kinfo]]

    def system(kcart_, e_, evc_):
        nk = len(kcart_)
        npwx = int((npw + 63) // 64 * 64)
        ev = np.zeros((nk, nb, npwx), dtype=np.complex128)
        ev[:, :, :npw] = evc_
        mk = np.zeros((nk, npwx, 3), dtype=np.int32)
        mk[:, :npw] = mill_g[:npw]
        return SynthSystem(cfg=dataclasses.replace(cfg, name=cfg.name + "-" + kind), kc=kc, mill_g=mill_g,
                           kcart=np.asarray(kcart_), e_ev=e_, ngk=np.full(nk, npw, dtype=np.int32), npwx=npwx,
                           mill_k=mk, evc=ev, gcutw=float(g2[npw - 1]), ecutwfc_ry=0.0, vcell=vol)
    return system(ki, e_ibz, evc_i), system(kt, e_fz, evc_f), R, nsym, kinfo, ngd
