/*
 * CPU oracle: FP64 restatement of the RPA response path of the reference
 * programs Optabs.x (Chi_RPA/Optabs.f90) and Sloss.x (W/Sloss.f90) and of the
 * reference modules (the Modules directory).  Every function cites the reference
 * lines it follows.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * or as the timed CPU baseline.  The product library never uses it.
 *
 * Restrictions (documented, not silent):
 *   - symmetry operations are the identity (Nsym = 1): R = RI = I, R1 = R2 = 1;
 *   - the G-vector lookup of the vertex map uses exact Miller arithmetic with a
 *     hash (orc_gfast_miller); orc_gfast_linear restates the reference's
 *     linear search with the eps tolerance and the tests check they agree.
 */
#include "rpa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double ORC_PI = 3.14159265358979323846; /* 4*atan(1), Optabs.f90:44 */

/* ------------------------------------------------------------------------ */
/* a4: local fields.  Optabs.f90:1016-1059 (genGlf), Sloss.f90:824-881
 * (genGlfandParity), Modules/local_field.f90:11-94.  parity from the Miller z
 * index as Sloss.f90:1349 stores it.  Returns -1 when Nlf > Nlfd (the
 * reference stops, Optabs.f90:1054-1057).                                    */
int orc_gen_glf(const char *lf, double ecut, double gcar, int ng, const double *g,
                const int *mill_g, int nlfd, int *nlf, double *glf, int *parg,
                int *idx) {
  int n = 0;
  int is_z = (strcmp(lf, "z") == 0);
  int is_xyz = (strcmp(lf, "xyz") == 0);
  if (!is_z && !is_xyz) return -2;
  for (int iG = 0; iG < ng; ++iG) {
    const double *G = g + 3 * iG;
    double eref;
    int take = 0;
    if (is_z) {
      if (G[0] == 0.0 && G[1] == 0.0) {
        eref = gcar * gcar * (G[2] * G[2]) / 2.0;
        take = (eref <= ecut);
      }
    } else {
      eref = gcar * gcar * ((G[0] * G[0] + G[1] * G[1]) + G[2] * G[2]) / 2.0;
      take = (eref <= ecut);
    }
    if (!take) continue;
    if (n < nlfd) {
      if (glf) {
        glf[3 * n + 0] = is_z ? 0.0 : G[0];
        glf[3 * n + 1] = is_z ? 0.0 : G[1];
        glf[3 * n + 2] = G[2];
      }
      if (parg) {
        int pz = mill_g ? mill_g[3 * iG + 2] : 0;
        parg[n] = ((pz / 2) * 2 == pz) ? 1 : -1;
      }
      if (idx) idx[n] = iG + 1;
    }
    ++n;
  }
  *nlf = n;
  return (n > nlfd) ? -1 : 0;
}

/* a5: Fermi-Dirac occupation, Optabs.f90:935-951 (genOccupation). */
void orc_occupation(double ei, double ef, double T, double *expo, double *f) {
  double x = (ei - ef) / T;
  if (x < -20) {
    *expo = 0.0;
    *f = 1.0;
  } else if (x > 20) {
    *expo = x;
    *f = 0.0;
  } else {
    *expo = exp(x);
    *f = 1.0 / (*expo + 1.0);
  }
}

/* a3: FBZ generation, Optabs.f90:871-933 (genFBZ; Sloss.f90:697-759 the same).  R is the
 * reference's R(48,3,3).  Writing fbz_check.dat is left to the caller.       */
int orc_gen_fbz(int nki, int nsym, double eps, const double *ki, const double *R,
                int *ntot, double *ktot) {
  int nk = 48 * nki;
  double *k = (double *)malloc(sizeof(double) * 3 * (size_t)nk);
  if (!k) return -1;
  int jk = 0, nt = 0;
  for (int i = 0; i < nsym; ++i) {
    for (int ik = 0; ik < nki; ++ik) {
      int it = 1;
      for (int n = 0; n < 3; ++n) {
        double s = 0.0;
        for (int m = 0; m < 3; ++m) s = s + R[i + 48 * (n + 3 * m)] * ki[3 * ik + m];
        k[3 * jk + n] = s;
      }
      for (int lk = 0; lk < jk; ++lk)
        if (fabs(k[3 * jk] - k[3 * lk]) <= eps && fabs(k[3 * jk + 1] - k[3 * lk + 1]) <= eps &&
            fabs(k[3 * jk + 2] - k[3 * lk + 2]) <= eps)
          it = 2;
      if (it == 1) {
        memcpy(ktot + 3 * nt, k + 3 * jk, 3 * sizeof(double));
        ++nt;
      }
      ++jk;
    }
  }
  free(k);
  *ntot = nt;
  return 0;
}

/* a3: q = (iq-1) * k_min, Optabs.f90:743-773 (findMinQ). */
void orc_find_min_q(int ntot, const double *ktot, int iq, double q[3]) {
  double kmin = 1.0;
  int ikmin = 0;
  for (int i = 0; i < ntot; ++i) {
    const double *k = ktot + 3 * i;
    double kref = sqrt((k[0] * k[0] + k[1] * k[1]) + k[2] * k[2]);
    if (kref == 0.0) continue;
    if (kref < kmin) {
      kmin = kref;
      ikmin = i;
    }
  }
  for (int c = 0; c < 3; ++c) q[c] = (double)(iq - 1) * ktot[3 * ikmin + c];
}

/* a3: k+q = G0 + K2, Optabs.f90:820-869 / Sloss.f90:646-695 (findKQinBZ)
 * with Nsym = 1.  Returns 1 (not in FBZ) or 2 (not in IBZ) as the reference's
 * two stop paths, 0 on success.  iG0, K2 are 1-based.                        */
int orc_find_kq(const double kq[3], double eps, int nki, int ntot, int ng,
                const double *ktot, const double *ki, const double *g, int *iG0,
                int *K2) {
  int it = 1;
  for (int iG = 0; iG < ng; ++iG) {
    for (int jk = 0; jk < ntot; ++jk) {
      if (fabs(kq[0] - g[3 * iG] - ktot[3 * jk]) <= eps &&
          fabs(kq[1] - g[3 * iG + 1] - ktot[3 * jk + 1]) <= eps &&
          fabs(kq[2] - g[3 * iG + 2] - ktot[3 * jk + 2]) <= eps) {
        it = 2;
        *iG0 = iG + 1;
        /* identity symmetry: K11 = sum(RI(1,1,1:3)*ktot(1:3,jk)) */
        double K11 = (1.0 * ktot[3 * jk] + 0.0 * ktot[3 * jk + 1]) + 0.0 * ktot[3 * jk + 2];
        double K22 = (0.0 * ktot[3 * jk] + 1.0 * ktot[3 * jk + 1]) + 0.0 * ktot[3 * jk + 2];
        double K33 = (0.0 * ktot[3 * jk] + 0.0 * ktot[3 * jk + 1]) + 1.0 * ktot[3 * jk + 2];
        for (int j = 0; j < nki; ++j) {
          if (fabs(K11 - ki[3 * j]) <= eps && fabs(K22 - ki[3 * j + 1]) <= eps &&
              fabs(K33 - ki[3 * j + 2]) <= eps) {
            *K2 = j + 1;
            return 0;
          }
        }
      }
    }
  }
  return it == 1 ? 1 : 2;
}

/* a3 with symmetry.  R and RI are (48,3,3) Fortran arrays: R(i,n,m) is
 * R[i + 48*(n + 3*m)].  PointR inverts each R(i,:,:) with gjel on a complex
 * 3x3 copy and keeps the real part (Chi_RPA/PointR.f90:119-138).          */
void orc_point_inverse(int nsym, const double *R, double *RI) {
  for (int i = 0; i < nsym; ++i) {
    orc_z T[9], B[9];
    for (int n = 0; n < 3; ++n)
      for (int m = 0; m < 3; ++m) {
        T[n + 3 * m] = R[i + 48 * (n + 3 * m)] + 0.0 * I;
        B[n + 3 * m] = (n == m) ? 1.0 : 0.0;
      }
    orc_gjel(T, 3, 3, B, 3, 3);
    for (int n = 0; n < 3; ++n)
      for (int m = 0; m < 3; ++m) RI[i + 48 * (n + 3 * m)] = creal(T[n + 3 * m]);
  }
}

static inline double orc_rot(const double *M, int i, int n, const double v[3]) {
  /* M(i,n,1)*v1 + M(i,n,2)*v2 + M(i,n,3)*v3, left to right as the reference */
  return (M[i + 48 * (n + 0)] * v[0] + M[i + 48 * (n + 3)] * v[1]) + M[i + 48 * (n + 6)] * v[2];
}

/* findKinIBZ, Optabs.f90:776-819: the first ik <= NkI are the IBZ list
 * itself (R1 = 1); the others are searched over operations 2..Nsym.        */
int orc_find_kin_ibz(int ik, int nki, int nsym, double eps, const double k[3],
                     const double *RI, const double *ki, int *R1, int *K1) {
  if (ik <= nki) {
    *R1 = 1;
    *K1 = ik;
    return 0;
  }
  for (int i = 1; i < nsym; ++i) {
    double v0 = orc_rot(RI, i, 0, k), v1 = orc_rot(RI, i, 1, k), v2 = orc_rot(RI, i, 2, k);
    for (int j = 0; j < nki; ++j)
      if (fabs(v0 - ki[3 * j]) <= eps && fabs(v1 - ki[3 * j + 1]) <= eps && fabs(v2 - ki[3 * j + 2]) <= eps) {
        *R1 = i + 1;
        *K1 = j + 1;
        return 0;
      }
  }
  return 1;
}

/* findKQinBZ, Optabs.f90:822-869: the first (iG, jk) with k+q - G(iG) =
 * ktot(jk), then the first operation i and IBZ point j with RI(i) ktot(jk)
 * = kI(j); the search over (iG, jk) goes on when none matches.            */
int orc_find_kq_sym(const double kq[3], double eps, int nsym, int nki, int ntot, int ng,
                    const double *ktot, const double *ki, const double *g, const double *RI,
                    int *iG0, int *R2, int *K2) {
  int it = 1;
  for (int iG = 0; iG < ng; ++iG)
    for (int jk = 0; jk < ntot; ++jk) {
      if (fabs(kq[0] - g[3 * iG] - ktot[3 * jk]) <= eps && fabs(kq[1] - g[3 * iG + 1] - ktot[3 * jk + 1]) <= eps &&
          fabs(kq[2] - g[3 * iG + 2] - ktot[3 * jk + 2]) <= eps) {
        it = 2;
        *iG0 = iG + 1;
        const double *kt = ktot + 3 * jk;
        for (int i = 0; i < nsym; ++i) {
          double v0 = orc_rot(RI, i, 0, kt), v1 = orc_rot(RI, i, 1, kt), v2 = orc_rot(RI, i, 2, kt);
          for (int j = 0; j < nki; ++j)
            if (fabs(v0 - ki[3 * j]) <= eps && fabs(v1 - ki[3 * j + 1]) <= eps &&
                fabs(v2 - ki[3 * j + 2]) <= eps) {
              *R2 = i + 1;
              *K2 = j + 1;
              return 0;
            }
        }
      }
    }
  return it == 1 ? 1 : 2;
}

/* G(n) = sum_m KC(n,m) * mill(m), Optabs_qe6.f90:1162-1167 (loadG_QE6). */
void orc_gcart(int n, const int *mill, const double *kc, double *g) {
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      double s = 0.0;
      for (int b = 0; b < 3; ++b) s = s + kc[a + 3 * b] * (double)mill[3 * i + b];
      g[3 * i + a] = s;
    }
}

/* a2: |k+G|^2 <= gcutw, sorted by |k+G| with index tie-break within eps8.
 * Modules/qe_gksort.f90:44-94 (gk_sort) and 96-203 (hpsort_eps).  igk is
 * 1-based.  Returns -1 on npwx overflow, -2 when the G list ends before the
 * |G| > |k| + sqrt(ecut) exit (both are error stops in the reference).      */
typedef struct {
  double q;
  int ind;
} orc_gk_entry;
static int orc_gk_cmp(const void *a, const void *b) {
  const orc_gk_entry *x = (const orc_gk_entry *)a, *y = (const orc_gk_entry *)b;
  if (fabs(x->q - y->q) >= 1.0e-8) return (x->q < y->q) ? -1 : 1;
  return (x->ind < y->ind) ? -1 : (x->ind > y->ind);
}
int orc_gk_sort(const double k[3], int ngm, const double *g, double gcutw, int npwx,
                int *ngk, int *igk) {
  const double eps8 = 1.0e-8;
  double kk = sqrt((k[0] * k[0] + k[1] * k[1]) + k[2] * k[2]);
  double q2x = (kk + sqrt(gcutw)) * (kk + sqrt(gcutw));
  orc_gk_entry *e = (orc_gk_entry *)malloc(sizeof(orc_gk_entry) * (size_t)(npwx + 1));
  int n = 0, ng;
  for (ng = 0; ng < ngm; ++ng) {
    const double *G = g + 3 * ng;
    double a0 = k[0] + G[0], a1 = k[1] + G[1], a2 = k[2] + G[2];
    double q = (a0 * a0 + a1 * a1) + a2 * a2;
    if (q <= eps8) q = 0.0;
    if (q <= gcutw) {
      if (n >= npwx) {
        free(e);
        return -1;
      }
      e[n].q = q;
      e[n].ind = ng + 1;
      ++n;
    } else if ((G[0] * G[0] + G[1] * G[1]) + G[2] * G[2] > q2x + eps8) {
      break;
    }
  }
  if (ng >= ngm) {
    free(e);
    return -2;
  }
  qsort(e, (size_t)n, sizeof(orc_gk_entry), orc_gk_cmp);
  for (int i = 0; i < n; ++i) igk[i] = e[i].ind;
  *ngk = n;
  free(e);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* a6: vertex index map with the reference's linear search and tolerance,
 * Sloss.f90:907-946 / Optabs.f90:1399-1449 / vertices.f90:120-168, R = RI = I.
 * gfast[iG*n1 + i] is the 0-based position in list 2, or -1.                */
void orc_gfast_linear(int n1, const double *l1, int n2, const double *l2, int nlf,
                      const double *glf, const double g0[3], double eps, int *gfast) {
  for (int iG = 0; iG < nlf; ++iG)
    for (int i = 0; i < n1; ++i) {
      double t0 = (l1[3 * i] + glf[3 * iG]) + g0[0];
      double t1 = (l1[3 * i + 1] + glf[3 * iG + 1]) + g0[1];
      double t2 = (l1[3 * i + 2] + glf[3 * iG + 2]) + g0[2];
      int hit = -1;
      for (int j = 0; j < n2; ++j)
        if (fabs(l2[3 * j] - t0) < eps && fabs(l2[3 * j + 1] - t1) < eps &&
            fabs(l2[3 * j + 2] - t2) < eps) {
          hit = j;
          break;
        }
      gfast[iG * n1 + i] = hit;
    }
}

/* Same map by exact Miller arithmetic (first match kept), O(n1 + n2). */
static inline unsigned orc_hkey(int a, int b, int c) {
  unsigned h = (unsigned)(a * 73856093) ^ (unsigned)(b * 19349663) ^ (unsigned)(c * 83492791);
  return h * 2654435761u;
}
int orc_gfast_miller(int n1, const int *m1, int n2, const int *m2, int nlf,
                     const int *mlf, const int m0[3], int *gfast) {
  unsigned cap = 1;
  while (cap < 2u * (unsigned)n2 + 2u) cap <<= 1;
  int *slot = (int *)malloc(sizeof(int) * cap);
  if (!slot) return -1;
  for (unsigned s = 0; s < cap; ++s) slot[s] = -1;
  for (int j = 0; j < n2; ++j) {
    unsigned h = orc_hkey(m2[3 * j], m2[3 * j + 1], m2[3 * j + 2]) & (cap - 1);
    for (;;) {
      int o = slot[h];
      if (o < 0) {
        slot[h] = j;
        break;
      }
      if (m2[3 * o] == m2[3 * j] && m2[3 * o + 1] == m2[3 * j + 1] &&
          m2[3 * o + 2] == m2[3 * j + 2])
        break; /* keep first occurrence */
      h = (h + 1) & (cap - 1);
    }
  }
  for (int iG = 0; iG < nlf; ++iG)
    for (int i = 0; i < n1; ++i) {
      int a = m1[3 * i] + mlf[3 * iG] + m0[0];
      int b = m1[3 * i + 1] + mlf[3 * iG + 1] + m0[1];
      int c = m1[3 * i + 2] + mlf[3 * iG + 2] + m0[2];
      unsigned h = orc_hkey(a, b, c) & (cap - 1);
      int hit = -1;
      for (;;) {
        int o = slot[h];
        if (o < 0) break;
        if (m2[3 * o] == a && m2[3 * o + 1] == b && m2[3 * o + 2] == c) {
          hit = o;
          break;
        }
        h = (h + 1) & (cap - 1);
      }
      gfast[iG * n1 + i] = hit;
    }
  free(slot);
  return 0;
}

/* Per-k lists of the two coefficient sets (a6 inputs).  Direct mode:
 * coefficient i <-> global G(i) for i < NGd (C1) and i < ngk(K2) (C2),
 * Sloss.f90:907-946.  IGK mode: coefficient i <-> mill_k(:,i,K),
 * vertices.f90:139,162.                                                     */
typedef struct {
  int n1, n2;
  const int *m1, *m2;   /* Miller lists */
  const double *c1base; /* evc(:,:,K1) */
  const double *c2base; /* evc(:,:,K2) */
  int K1, K2;           /* 0-based */
  int m0[3];
} orc_kpair;

static int orc_make_pair(const orc_problem *p, int ik, orc_kpair *kp) {
  int K1 = p->K1[ik] - 1, K2 = p->K2[ik] - 1, ig0 = p->iG0[ik] - 1;
  kp->K1 = K1;
  kp->K2 = K2;
  for (int c = 0; c < 3; ++c) kp->m0[c] = p->mill_g[3 * ig0 + c];
  size_t kstride = (size_t)2 * p->npwx * p->nband;
  kp->c1base = p->evc + kstride * K1;
  kp->c2base = p->evc + kstride * K2;
  if (p->map_mode == ORC_MAP_DIRECT) {
    if (p->ngd > p->ngk[K1] || p->ngd > p->ngk[K2]) return -3; /* Optabs.f90:349-352,377-380 */
    kp->n1 = p->ngd;
    kp->n2 = p->ngk[K2];
    if (kp->n2 > p->ngm) return -4;
    kp->m1 = p->mill_g;
    kp->m2 = p->mill_g;
  } else {
    kp->n1 = p->ngk[K1];
    kp->n2 = p->ngk[K2];
    kp->m1 = p->mill_k + (size_t)3 * p->npwx * K1;
    kp->m2 = p->mill_k + (size_t)3 * p->npwx * K2;
  }
  return 0;
}

/* a6 with symmetry (Optabs.f90:1405-1449, Sloss.f90:917-946; the igk map of
 * Modules/vertices.f90:130-160, 343-370 the same on the k lists): plane wave i
 * of C1 carries G'' = R(R1) G(i) (cartesian, in the reference's order), the
 * target is RI(R2) (G'' + Glf(G) + G(iG0)), matched against the G list of
 * C2 within eps per component.  The match takes the lattice vector nearest
 * to the target (its Miller indices rounded through KC^-1) and keeps it when
 * every component lies within eps: the reference's linear search finds the
 * same vector whenever eps is below half a lattice spacing.  l1 receives
 * R(R1) G(i) for the current vertices.                                       */
static int orc_gfast_sym(const orc_problem *p, int ik, const orc_kpair *kp, int nlf, const int *mlf,
                         int *gfast, double *l1) {
  const int r1 = p->R1[ik] - 1, r2 = p->R2[ik] - 1;
  const double *kc = p->kc;
  /* KC^-1 by cofactors */
  double inv[9];
  {
    const double a = kc[0], b = kc[3], c = kc[6], d = kc[1], e = kc[4], f = kc[7], g = kc[2], h = kc[5], k = kc[8];
    const double det = a * (e * k - f * h) - b * (d * k - f * g) + c * (d * h - e * g);
    inv[0] = (e * k - f * h) / det;
    inv[3] = (c * h - b * k) / det;
    inv[6] = (b * f - c * e) / det;
    inv[1] = (f * g - d * k) / det;
    inv[4] = (a * k - c * g) / det;
    inv[7] = (c * d - a * f) / det;
    inv[2] = (d * h - e * g) / det;
    inv[5] = (b * g - a * h) / det;
    inv[8] = (a * e - b * d) / det;
  }
  const int n1 = kp->n1, n2 = kp->n2;
  unsigned cap = 1;
  while (cap < 2u * (unsigned)n2 + 2u) cap <<= 1;
  int *slot = (int *)malloc(sizeof(int) * cap);
  double *glf = (double *)malloc(sizeof(double) * 3 * nlf);
  if (!slot || !glf) {
    free(slot);
    free(glf);
    return -1;
  }
  for (unsigned t = 0; t < cap; ++t) slot[t] = -1;
  const int *m2 = kp->m2;
  for (int j = 0; j < n2; ++j) {
    unsigned hh = orc_hkey(m2[3 * j], m2[3 * j + 1], m2[3 * j + 2]) & (cap - 1);
    while (slot[hh] >= 0) hh = (hh + 1) & (cap - 1);
    slot[hh] = j;
  }
  orc_gcart(nlf, mlf, kc, glf);
  double g0[3];
  orc_gcart(1, kp->m0, kc, g0);
  for (int i = 0; i < n1; ++i) {
    double gi[3];
    orc_gcart(1, kp->m1 + 3 * i, kc, gi);
    const double k11 = orc_rot(p->R, r1, 0, gi), k22 = orc_rot(p->R, r1, 1, gi), k33 = orc_rot(p->R, r1, 2, gi);
    l1[3 * i] = k11;
    l1[3 * i + 1] = k22;
    l1[3 * i + 2] = k33;
    for (int iG = 0; iG < nlf; ++iG) {
      const double t[3] = {k11 + glf[3 * iG] + g0[0], k22 + glf[3 * iG + 1] + g0[1], k33 + glf[3 * iG + 2] + g0[2]};
      double w[3];
      for (int n = 0; n < 3; ++n) w[n] = orc_rot(p->RI, r2, n, t);
      int mm[3];
      for (int n = 0; n < 3; ++n) mm[n] = (int)lround(inv[n] * w[0] + inv[n + 3] * w[1] + inv[n + 6] * w[2]);
      unsigned hh = orc_hkey(mm[0], mm[1], mm[2]) & (cap - 1);
      int hit = -1;
      while (slot[hh] >= 0) {
        const int o = slot[hh];
        if (m2[3 * o] == mm[0] && m2[3 * o + 1] == mm[1] && m2[3 * o + 2] == mm[2]) {
          hit = o;
          break;
        }
        hh = (hh + 1) & (cap - 1);
      }
      if (hit >= 0) {
        double gj[3];
        orc_gcart(1, m2 + 3 * hit, kc, gj);
        if (!(fabs(gj[0] - w[0]) < p->eps && fabs(gj[1] - w[1]) < p->eps && fabs(gj[2] - w[2]) < p->eps)) hit = -1;
      }
      gfast[iG * n1 + i] = hit;
    }
  }
  free(slot);
  free(glf);
  return 0;
}

/* The index map of a k pair and, for the current vertices, the cartesian G
 * of C1's plane waves (rotated by R(R1) under symmetry).                    */
static int orc_pair_maps(const orc_problem *p, int ik, const orc_kpair *kp, int nlf, const int *mlf, int *gfast,
                         double *l1) {
  if (p->R1 && (p->R1[ik] != 1 || p->R2[ik] != 1))  /* both maps: kp holds the mode's lists */
    return orc_gfast_sym(p, ik, kp, nlf, mlf, gfast, l1);
  orc_gfast_miller(kp->n1, kp->m1, kp->n2, kp->m2, nlf, mlf, kp->m0, gfast);
  if (l1) orc_gcart(kp->n1, kp->m1, p->kc, l1);
  return 0;
}

static void orc_lf_miller(const orc_problem *p, int *mlf) {
  for (int iG = 0; iG < p->nlf; ++iG)
    for (int c = 0; c < 3; ++c) mlf[3 * iG + c] = p->mill_g[3 * (p->lf_idx[iG] - 1) + c];
}

/* a7: charge vertex rho(iG) = sum_i conj(C1(i)) C2(gfast), Sloss.f90:905-951. */
static void orc_charge_vertex(int nlf, int n1, const int *gfast, const double *c1,
                              const double *c2, double *rr, double *ri) {
  for (int iG = 0; iG < nlf; ++iG) {
    double sr = 0.0, si = 0.0;
    const int *gf = gfast + (size_t)iG * n1;
    for (int i = 0; i < n1; ++i) {
      int j = gf[i];
      if (j < 0) continue;
      double ar = c1[2 * i], ai = -c1[2 * i + 1];
      double br = c2[2 * j], bi = c2[2 * j + 1];
      sr = sr + (ar * br - ai * bi);
      si = si + (ar * bi + ai * br);
    }
    rr[iG] = sr;
    ri[iG] = si;
  }
}

/* a8: current vertex, Optabs.f90:1399-1469 (direct) and vertices.f90:343-390
 * (_v2): j(iG) += ((0.5*conj(C1)) * J) * C2 with
 * J = (q + 2k + Glf(iG) + 2 R G(iG1))_pol * Gcar.                           */
static void orc_current_vertex(int pol, int nlf, int n1, const int *gfast,
                               const double *l1, const double *glf, const double q[3],
                               const double kv[3], double gcar, const double *c1,
                               const double *c2, double *rr, double *ri) {
  for (int iG = 0; iG < nlf; ++iG) {
    double sr = 0.0, si = 0.0;
    const int *gf = gfast + (size_t)iG * n1;
    for (int i = 0; i < n1; ++i) {
      int j = gf[i];
      if (j < 0) continue;
      double J = (((q[pol] + 2.0 * kv[pol]) + glf[3 * iG + pol]) + 2.0 * l1[3 * i + pol]) * gcar;
      double ar = 0.5 * c1[2 * i], ai = 0.5 * (-c1[2 * i + 1]);
      ar = ar * J;
      ai = ai * J;
      double br = c2[2 * j], bi = c2[2 * j + 1];
      sr = sr + (ar * br - ai * bi);
      si = si + (ar * bi + ai * br);
    }
    rr[iG] = sr;
    ri[iG] = si;
  }
}

int orc_vertices(const orc_problem *p, int kind, double gcar, int n_lo, int n_hi,
                 int m_lo, int m_hi, orc_z *rho1, orc_z *rho2) {
  int nlf = p->nlf;
  int nn = n_hi - n_lo, nm = m_hi - m_lo;
  size_t ntrans = (size_t)p->ntot * nn * nm;
  double *glf = (double *)malloc(sizeof(double) * 3 * nlf);
  int *mlf = (int *)malloc(sizeof(int) * 3 * nlf);
  orc_lf_miller(p, mlf);
  orc_gcart(nlf, mlf, p->kc, glf);
  int err = 0;
#pragma omp parallel
  {
    int *gfast = (int *)malloc(sizeof(int) * (size_t)nlf * (p->ngd > p->npwx ? p->ngd : p->npwx));
    double *l1 = (double *)malloc(sizeof(double) * 3 * (size_t)(p->ngd > p->npwx ? p->ngd : p->npwx));
    double *vr = (double *)malloc(sizeof(double) * nlf * 2);
    double *vi = (double *)malloc(sizeof(double) * nlf * 2);
#pragma omp for schedule(static)
    for (int ik = 0; ik < p->ntot; ++ik) {
      orc_kpair kp;
      if (orc_make_pair(p, ik, &kp)) {
        err = -3;
        continue;
      }
      if (orc_pair_maps(p, ik, &kp, nlf, mlf, gfast, l1)) {
        err = -5;
        continue;
      }
      for (int n = n_lo; n < n_hi; ++n) {
        const double *c1 = kp.c1base + (size_t)2 * p->npwx * n;
        for (int m = m_lo; m < m_hi; ++m) {
          const double *c2 = kp.c2base + (size_t)2 * p->npwx * m;
          size_t t = ((size_t)ik * nn + (n - n_lo)) * nm + (m - m_lo);
          if (kind == 0) {
            orc_charge_vertex(nlf, kp.n1, gfast, c1, c2, vr, vi);
            for (int iG = 0; iG < nlf; ++iG) rho1[(size_t)iG * ntrans + t] = vr[iG] + I * vi[iG];
          } else {
            const double *kv = p->ktot + 3 * ik;
            int pol1 = (kind == 4) ? 1 : kind - 1;
            orc_current_vertex(pol1, nlf, kp.n1, gfast, l1, glf, p->q, kv, gcar, c1, c2, vr, vi);
            for (int iG = 0; iG < nlf; ++iG) rho1[(size_t)iG * ntrans + t] = vr[iG] + I * vi[iG];
            if (rho2) {
              if (kind == 4)
                orc_current_vertex(2, nlf, kp.n1, gfast, l1, glf, p->q, kv, gcar, c1, c2, vr, vi);
              for (int iG = 0; iG < nlf; ++iG) rho2[(size_t)iG * ntrans + t] = vr[iG] + I * vi[iG];
            }
          }
        }
      }
    }
    free(gfast);
    free(l1);
    free(vr);
    free(vi);
  }
  free(glf);
  free(mlf);
  return err;
}

/* ------------------------------------------------------------------------ */
/* a9/a10 Sloss: real S0(no,Nlf,Nlf), Sloss.f90:355-430.
 *   dE = o + E(K1,n) - E(K2,m);  Lor = -eta/(dE^2+eta^2);
 *   if |Lor| >= 1e-3/eta:  S0 -= Re[2 Lor rho(iG) conj(rho(jG))] / (pi Ntot Vcell)
 * The per-k partial sums are added in k order within each thread's static
 * block, then the thread blocks in thread order (deterministic for a fixed
 * thread count).                                                           */
int orc_sloss_s0(const orc_problem *p, const orc_sloss_params *sp, double *S0,
                 int nthreads) {
  const int nlf = p->nlf, no = sp->no, nband = p->nband, nocc = sp->nocc;
  const size_t s0n = (size_t)no * nlf * nlf;
  const double den = (ORC_PI * (double)p->ntot) * sp->vcell;
  const double cut = 1.0e-3 / sp->eta;
  int *mlf = (int *)malloc(sizeof(int) * 3 * nlf);
  orc_lf_miller(p, mlf);
  if (nthreads <= 0) nthreads = 1;
  double *acc = (double *)calloc((size_t)nthreads * s0n, sizeof(double));
  int err = 0;
#pragma omp parallel num_threads(nthreads)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *mine = acc + (size_t)tid * s0n;
    /* internal layout [io][iG][jG] for locality; transposed at the end */
    double *part = (double *)malloc(sizeof(double) * s0n);
    int nmax = p->ngd > p->npwx ? p->ngd : p->npwx;
    int *gfast = (int *)malloc(sizeof(int) * (size_t)nlf * nmax);
    double *l1s = (double *)malloc(sizeof(double) * 3 * (size_t)nmax);
    double *rr = (double *)malloc(sizeof(double) * nlf);
    double *ri = (double *)malloc(sizeof(double) * nlf);
#pragma omp for schedule(static)
    for (int ik = 0; ik < p->ntot; ++ik) {
      orc_kpair kp;
      if (orc_make_pair(p, ik, &kp)) {
        err = -3;
        continue;
      }
      if (orc_pair_maps(p, ik, &kp, nlf, mlf, gfast, l1s)) {
        err = -5;
        continue;
      }
      memset(part, 0, sizeof(double) * s0n);
      for (int n = 0; n < nocc; ++n) {
        const double *c1 = kp.c1base + (size_t)2 * p->npwx * n;
        const double e1 = p->E[kp.K1 + p->nki * n];
        for (int m = nocc; m < nband; ++m) {
          const double *c2 = kp.c2base + (size_t)2 * p->npwx * m;
          const double e2 = p->E[kp.K2 + p->nki * m];
          orc_charge_vertex(nlf, kp.n1, gfast, c1, c2, rr, ri);
          for (int io = 0; io < no; ++io) {
            double o = (double)io * sp->domega;
            double dE = (o + e1) - e2;
            double lor = -sp->eta / (dE * dE + sp->eta * sp->eta);
            if (!(fabs(lor) >= cut)) continue;
            double t = 2.0 * lor;
            double *row = part + (size_t)io * nlf * nlf;
            for (int iG = 0; iG < nlf; ++iG) {
              double ar = t * rr[iG], ai = t * ri[iG];
              double *dst = row + (size_t)iG * nlf;
              for (int jG = 0; jG < nlf; ++jG) dst[jG] -= (ar * rr[jG] + ai * ri[jG]) / den;
            }
          }
        }
      }
      for (size_t x = 0; x < s0n; ++x) mine[x] += part[x];
    }
    free(part);
    free(gfast);
    free(l1s);
    free(rr);
    free(ri);
  }
  for (int io = 0; io < no; ++io)
    for (int iG = 0; iG < nlf; ++iG)
      for (int jG = 0; jG < nlf; ++jG) {
        double s = 0.0;
        size_t x = ((size_t)io * nlf + iG) * nlf + jG;
        for (int t = 0; t < nthreads; ++t) s += acc[(size_t)t * s0n + x];
        S0[io + (size_t)no * (iG + (size_t)nlf * jG)] = s;
      }
  free(acc);
  free(mlf);
  return err;
}

/* a9/a10 Optabs: complex S0(-no:no,Nlf,Nlf) and Qeff(Nlf,Nlf),
 * Optabs.f90:327-461:
 *   pair kept if |f_n - f_m| >= df_cut or n == m (360);
 *   n /= m: dE = o + E(K1,n) - E(K2,m), Lor = eta/(dE^2+eta^2),
 *           if |Lor| >= Lor_cut/eta: S0 -= 2 df Lor j1(iG) conj(j2(jG))/(pi Ntot V);
 *   n == m and |E-Ef| <= 10 T: Qeff += 2 fact j1 conj(j2)/(Ntot V),
 *           fact = -expo/((expo+1)^2 T) (430-442).                          */
int orc_optabs_s0(const orc_problem *p, const orc_optabs_params *op, orc_z *S0,
                  orc_z *Qeff, int nthreads) {
  const int nlf = p->nlf, no = op->no, nband = p->nband, nw = 2 * no + 1;
  const size_t s0n = (size_t)nw * nlf * nlf;
  const double den = (ORC_PI * (double)p->ntot) * op->vcell;
  const double dq = (double)p->ntot * op->vcell;
  const double cut = op->lor_cut / op->eta_inter;
  int *mlf = (int *)malloc(sizeof(int) * 3 * nlf);
  double *glf = (double *)malloc(sizeof(double) * 3 * nlf);
  orc_lf_miller(p, mlf);
  orc_gcart(nlf, mlf, p->kc, glf);
  if (nthreads <= 0) nthreads = 1;
  double *acc = (double *)calloc((size_t)nthreads * 2 * s0n, sizeof(double));
  double *qacc = (double *)calloc((size_t)nthreads * 2 * nlf * nlf, sizeof(double));
  int pol1 = (op->pol == 4) ? 1 : op->pol - 1;
  int pol2 = (op->pol == 4) ? 2 : op->pol - 1;
  int err = 0;
#pragma omp parallel num_threads(nthreads)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *mine = acc + (size_t)tid * 2 * s0n;
    double *qmine = qacc + (size_t)tid * 2 * nlf * nlf;
    double *part = (double *)malloc(sizeof(double) * 2 * s0n);
    double *qpart = (double *)malloc(sizeof(double) * 2 * nlf * nlf);
    int nmax = p->ngd > p->npwx ? p->ngd : p->npwx;
    int *gfast = (int *)malloc(sizeof(int) * (size_t)nlf * nmax);
    double *l1 = (double *)malloc(sizeof(double) * 3 * (size_t)nmax);
    double *ar = (double *)malloc(sizeof(double) * nlf), *ai = (double *)malloc(sizeof(double) * nlf);
    double *br = (double *)malloc(sizeof(double) * nlf), *bi = (double *)malloc(sizeof(double) * nlf);
#pragma omp for schedule(static)
    for (int ik = 0; ik < p->ntot; ++ik) {
      orc_kpair kp;
      if (orc_make_pair(p, ik, &kp)) {
        err = -3;
        continue;
      }
      if (orc_pair_maps(p, ik, &kp, nlf, mlf, gfast, l1)) {
        err = -5;
        continue;
      }
      const double *kv = p->ktot + 3 * ik;
      memset(part, 0, sizeof(double) * 2 * s0n);
      memset(qpart, 0, sizeof(double) * 2 * nlf * nlf);
      for (int n = 0; n < nband; ++n) {
        double expo1, f1, expo2, f2;
        const double e1 = p->E[kp.K1 + p->nki * n];
        orc_occupation(e1, op->efermi, op->T, &expo1, &f1);
        const double *c1 = kp.c1base + (size_t)2 * p->npwx * n;
        for (int m = 0; m < nband; ++m) {
          const double e2 = p->E[kp.K2 + p->nki * m];
          orc_occupation(e2, op->efermi, op->T, &expo2, &f2);
          double df = f1 - f2;
          if (!((fabs(df) >= op->df_cut) || (n == m))) continue;
          int intra = (n == m) && (fabs(e1 - op->efermi) <= 10.0 * op->T);
          if (n == m && !intra) continue; /* vertex computed but unused (430) */
          const double *c2 = kp.c2base + (size_t)2 * p->npwx * m;
          orc_current_vertex(pol1, nlf, kp.n1, gfast, l1, glf, p->q, kv, op->gcar, c1, c2, ar, ai);
          if (pol2 != pol1)
            orc_current_vertex(pol2, nlf, kp.n1, gfast, l1, glf, p->q, kv, op->gcar, c1, c2, br, bi);
          else {
            memcpy(br, ar, sizeof(double) * nlf);
            memcpy(bi, ai, sizeof(double) * nlf);
          }
          if (n != m) {
            for (int io = -no; io <= no; ++io) {
              double o = (double)io * op->domega;
              double dE = (o + e1) - e2;
              double lor = op->eta_inter / (dE * dE + op->eta_inter * op->eta_inter);
              if (!(fabs(lor) >= cut)) continue;
              double t = (2.0 * df) * lor;
              double *row = part + (size_t)2 * (io + no) * nlf * nlf;
              for (int iG = 0; iG < nlf; ++iG) {
                double xr = t * ar[iG], xi = t * ai[iG];
                double *dst = row + (size_t)2 * iG * nlf;
                for (int jG = 0; jG < nlf; ++jG) {
                  /* (xr + i xi) * (br - i bi) */
                  double pr = xr * br[jG] + xi * bi[jG];
                  double pi_ = xi * br[jG] - xr * bi[jG];
                  dst[2 * jG] -= pr / den;
                  dst[2 * jG + 1] -= pi_ / den;
                }
              }
            }
          } else {
            double fact = expo1 / ((expo1 + 1.0) * (expo1 + 1.0));
            fact = -fact / op->T;
            double t = 2.0 * fact;
            for (int iG = 0; iG < nlf; ++iG)
              for (int jG = 0; jG < nlf; ++jG) {
                double xr = t * ar[iG], xi = t * ai[iG];
                double pr = xr * br[jG] + xi * bi[jG];
                double pi_ = xi * br[jG] - xr * bi[jG];
                qpart[2 * (iG * nlf + jG)] += pr / dq;
                qpart[2 * (iG * nlf + jG) + 1] += pi_ / dq;
              }
          }
        }
      }
      for (size_t x = 0; x < 2 * s0n; ++x) mine[x] += part[x];
      for (int x = 0; x < 2 * nlf * nlf; ++x) qmine[x] += qpart[x];
    }
    free(part);
    free(qpart);
    free(gfast);
    free(l1);
    free(ar);
    free(ai);
    free(br);
    free(bi);
  }
  for (int w = 0; w < nw; ++w)
    for (int iG = 0; iG < nlf; ++iG)
      for (int jG = 0; jG < nlf; ++jG) {
        double sr = 0.0, si = 0.0;
        size_t x = 2 * (((size_t)w * nlf + iG) * nlf + jG);
        for (int t = 0; t < nthreads; ++t) {
          sr += acc[(size_t)t * 2 * s0n + x];
          si += acc[(size_t)t * 2 * s0n + x + 1];
        }
        S0[w + (size_t)nw * (iG + (size_t)nlf * jG)] = sr + I * si;
      }
  if (Qeff)
    for (int iG = 0; iG < nlf; ++iG)
      for (int jG = 0; jG < nlf; ++jG) {
        double sr = 0.0, si = 0.0;
        for (int t = 0; t < nthreads; ++t) {
          sr += qacc[(size_t)t * 2 * nlf * nlf + 2 * (iG * nlf + jG)];
          si += qacc[(size_t)t * 2 * nlf * nlf + 2 * (iG * nlf + jG) + 1];
        }
        Qeff[iG + (size_t)nlf * jG] = sr + I * si;
      }
  free(acc);
  free(qacc);
  free(mlf);
  free(glf);
  return err;
}

/* ------------------------------------------------------------------------ */
/* a12 K1: Optabs genReChi0/genImChi0, Optabs.f90:1112-1318 (== the complex
 * variant of Modules/response_function_KK.f90:74-178).  S0(-no:no,Nlf,Nlf),
 * Pi0(no-1,Nlf,Nlf) for io = 1..no-1.                                       */
static double orc_k1_part(int io, int no, double oi, double domega, const orc_z *s,
                          int want_re) {
  /* s points at S0(0,iG,jG) (frequency index 0) */
#define S0P(j) (want_re ? creal(s[(j)]) : cimag(s[(j)]))
  double acc = 0.0, fact = 0.0, oj;
  if (io == 1) {
    for (int jo = 2; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      fact = domega / oj;
      if (jo == 2)
        fact = 3.0 / 2.0;
      else if (jo == no)
        fact = 0.5 * domega / oj;
      acc = acc + fact * (S0P(-jo + 1) - S0P(jo - 1));
    }
  } else if (io == 2) {
    for (int jo = 1; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      if (jo != io) fact = domega / (oi - oj);
      if (jo == 1)
        fact = 1.0;
      else if (jo == 2)
        fact = 0.0;
      else if (jo == 3)
        fact = -3.0 / 2.0;
      else if (jo == no)
        fact = 0.5 * domega / (oi - oj);
      acc = acc + fact * S0P(jo - 1);
      fact = domega / (oi + oj);
      if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
      acc = acc + fact * S0P(-jo + 1);
    }
  } else if (io == no - 1) {
    for (int jo = 1; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      if (jo != io) fact = domega / (oi - oj);
      if (jo == 1)
        fact = 0.5 * domega / (oi - oj);
      else if (jo == no - 2)
        fact = 3.0 / 2.0;
      else if (jo == no - 1)
        fact = 0.0;
      else if (jo == no)
        fact = -1.0;
      acc = acc + fact * S0P(jo - 1);
      fact = domega / (oi + oj);
      if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
      acc = acc + fact * S0P(-jo + 1);
    }
  } else {
    for (int jo = 1; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      if (jo != io) fact = domega / (oi - oj);
      if (jo == 1)
        fact = 0.5 * domega / (oi - oj);
      else if (jo == io - 1)
        fact = 3.0 / 2.0;
      else if (jo == io)
        fact = 0.0;
      else if (jo == io + 1)
        fact = -3.0 / 2.0;
      else if (jo == no)
        fact = 0.5 * domega / (oi - oj);
      acc = acc + fact * S0P(jo - 1);
      fact = domega / (oi + oj);
      if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
      acc = acc + fact * S0P(-jo + 1);
    }
  }
#undef S0P
  return acc;
}

void orc_kk_optabs(int no, int nlf, double domega, const orc_z *S0, orc_z *Pi0) {
  const int nw = 2 * no + 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int jG = 0; jG < nlf; ++jG)
    for (int iG = 0; iG < nlf; ++iG) {
      const orc_z *s = S0 + (size_t)nw * (iG + (size_t)nlf * jG) + no; /* frequency 0 */
      for (int io = 1; io <= no - 1; ++io) {
        double oi = (double)(io - 1) * domega;
        double re = orc_k1_part(io, no, oi, domega, s, 1);
        re = re + ORC_PI * cimag(s[io - 1]);
        double im = orc_k1_part(io, no, oi, domega, s, 0);
        im = im - ORC_PI * creal(s[io - 1]);
        Pi0[(io - 1) + (size_t)(no - 1) * (iG + (size_t)nlf * jG)] = re + I * im;
      }
    }
}

/* a14 K2: Sloss genChi0, Sloss.f90:954-1059 (if / else-if chain: the
 * overrides only act at jo == io).  S0(no,Nlf,Nlf) real, chi0(no-1,Nlf,Nlf). */
static double orc_k2_re(int io, int no, double oi, double domega, const double *s) {
  double re = 0.0, fact = 0.0, oj;
  if (io == 1) {
    for (int jo = 2; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      fact = domega / oj;
      if (jo == 2)
        fact = 3.0 / 2.0;
      else if (jo == no)
        fact = 0.5 * domega / oj;
      re = re + fact * s[jo - 1];
    }
    re = -2.0 * re;
    return re;
  }
  for (int jo = 1; jo <= no; ++jo) {
    oj = (double)(jo - 1) * domega;
    if (jo != io) {
      fact = domega / (oi - oj);
    } else if (io == 2) {
      if (jo == 1)
        fact = 1.0;
      else if (jo == 2)
        fact = 0.0;
      else if (jo == 3)
        fact = -3.0 / 2.0;
      else if (jo == no)
        fact = 0.5 * domega / (oi - oj);
    } else if (io == no - 1) {
      if (jo == 1)
        fact = 0.5 * domega / (oi - oj);
      else if (jo == no - 2)
        fact = 3.0 / 2.0;
      else if (jo == no - 1)
        fact = 0.0;
      else if (jo == no)
        fact = -1.0;
    } else {
      if (jo == 1)
        fact = 0.5 * domega / (oi - oj);
      else if (jo == io - 1)
        fact = 3.0 / 2.0;
      else if (jo == io)
        fact = 0.0;
      else if (jo == io + 1)
        fact = -3.0 / 2.0;
      else if (jo == no)
        fact = 0.5 * domega / (oi - oj);
    }
    re = re + fact * s[jo - 1];
    fact = domega / (oi + oj);
    if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
    re = re - fact * s[jo - 1];
  }
  return re;
}

void orc_kk_sloss(int no, int nlf, double domega, const double *S0, orc_z *chi0) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int jG = 0; jG < nlf; ++jG)
    for (int iG = 0; iG < nlf; ++iG) {
      const double *s = S0 + (size_t)no * (iG + (size_t)nlf * jG);
      for (int io = 1; io <= no - 1; ++io) {
        double oi = (double)(io - 1) * domega;
        double re = orc_k2_re(io, no, oi, domega, s);
        double im = -ORC_PI * s[io - 1];
        chi0[(io - 1) + (size_t)(no - 1) * (iG + (size_t)nlf * jG)] = re + I * im;
      }
    }
}

/* a20: genW1W2, Sloss.f90:1062-1171 (same if / else-if chain). */
void orc_w1w2(int io, int no, int nlf, int iG, int jG, double domega, const double *S0,
              double *W1, double *W2) {
  const double *s = S0 + (size_t)no * ((iG - 1) + (size_t)nlf * (jG - 1));
  double w1 = 0.0, w2 = 0.0, fact = 0.0, oj;
  double oi = (double)(io - 1) * domega;
  if (io == 1) {
    for (int jo = 2; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      fact = domega / oj;
      if (jo == 2)
        fact = 3.0 / 2.0;
      else if (jo == no)
        fact = 0.5 * domega / oj;
      w1 = w1 - fact * s[jo - 1];
    }
    w2 = -w1;
  } else {
    for (int jo = 1; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      if (jo != io) {
        fact = domega / (oi - oj);
      } else if (io == 2) {
        if (jo == 1)
          fact = 1.0;
        else if (jo == 2)
          fact = 0.0;
        else if (jo == 3)
          fact = -3.0 / 2.0;
        else if (jo == no)
          fact = 0.5 * domega / (oi - oj);
      } else if (io == no - 1) {
        if (jo == 1)
          fact = 0.5 * domega / (oi - oj);
        else if (jo == no - 2)
          fact = 3.0 / 2.0;
        else if (jo == no - 1)
          fact = 0.0;
        else if (jo == no)
          fact = -1.0;
      } else {
        if (jo == 1)
          fact = 0.5 * domega / (oi - oj);
        else if (jo == io - 1)
          fact = 3.0 / 2.0;
        else if (jo == io)
          fact = 0.0;
        else if (jo == io + 1)
          fact = -3.0 / 2.0;
        else if (jo == no)
          fact = 0.5 * domega / (oi - oj);
      }
      w1 = w1 + fact * s[jo - 1];
      fact = domega / (oi + oj);
      if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
      w2 = w2 + fact * s[jo - 1];
    }
  }
  *W1 = w1;
  *W2 = w2;
}

/* K3: real-S0 variant of Modules/response_function_KK.f90:181-273.  S0m is
 * the module's S0(-No:No,Nlf,Nlf) array (real).                             */
void orc_kk_module_real(int io, int no, int nlf, int iG, int jG, double domega,
                        const double *S0m, double *re_out, double *im_out) {
  const int nw = 2 * no + 1;
  const double *s = S0m + (size_t)nw * ((iG - 1) + (size_t)nlf * (jG - 1)) + no; /* index 0 */
  double re = 0.0, fact = 0.0, oj;
  double oi = (double)(io - 1) * domega;
  if (io == 1) {
    for (int jo = 2; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      fact = domega / oj;
      if (jo == 2) fact = 3.0 / 2.0;
      if (jo == no) fact = 0.5 * domega / oj;
      re = re + fact * s[jo];
    }
    re = -2.0 * re;
  } else if (io == 2) {
    for (int jo = 1; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      if (jo != io)
        fact = domega / (oi - oj);
      else if (jo == 1)
        fact = 1.0;
      else if (jo == 2)
        fact = 0.0;
      else if (jo == 3)
        fact = -3.0 / 2.0;
      else if (jo == no)
        fact = 0.5 * domega / (oi - oj);
      re = re + fact * s[jo];
      fact = domega / (oi + oj);
      if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
      re = re - fact * s[jo];
    }
  } else if (io == no - 1) {
    for (int jo = 1; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      if (jo != io) fact = domega / (oi - oj);
      if (jo == 1) fact = 0.5 * domega / (oi - oj);
      if (jo == no - 2) fact = 3.0 / 2.0;
      if (jo == no - 1) fact = 0.0;
      if (jo == no) fact = -1.0;
      re = re + fact * s[jo];
      fact = domega / (oi + oj);
      if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
      re = re - fact * s[jo];
    }
  } else {
    for (int jo = 1; jo <= no; ++jo) {
      oj = (double)(jo - 1) * domega;
      if (jo != io) fact = domega / (oi - oj);
      if (jo == 1) fact = 0.5 * domega / (oi - oj);
      if (jo == io - 1) fact = 3.0 / 2.0;
      if (jo == io) fact = 0.0;
      if (jo == io + 1) fact = -3.0 / 2.0;
      if (jo == no) fact = 0.5 * domega / (oi - oj);
      re = re + fact * s[jo];
      fact = domega / (oi + oj);
      if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
      re = re - fact * s[jo];
    }
  }
  *re_out = re;
  *im_out = -ORC_PI * s[io];
}

/* ------------------------------------------------------------------------ */
/* a15: slab Coulomb V(G,G'), Sloss.f90:1173-1208.  qx, qy and GlfV are
 * cartesian (a.u.); glf in 2pi/a0 units (compared exactly as the reference). */
void orc_genV(double eps, double qx, double qy, double c0, int nlf, const int *parg,
              const double *glf, const double *glfv, double *V) {
  for (int iG = 0; iG < nlf; ++iG) {
    double a = qx + glfv[3 * iG], b = qy + glfv[3 * iG + 1];
    double gabs = sqrt(a * a + b * b);
    if (gabs == 0.0) gabs = eps;
    for (int jG = 0; jG < nlf; ++jG) {
      double v = 0.0;
      if (glf[3 * jG] == glf[3 * iG] && glf[3 * jG + 1] == glf[3 * iG + 1]) {
        double gz_i = glfv[3 * iG + 2], gz_j = glfv[3 * jG + 2];
        v = 4.0 * ORC_PI * (1.0 - exp(-gabs * c0)) / (gabs * c0);
        v = v * (gabs * gabs - gz_i * gz_j);
        v = v / (gabs * gabs + gz_i * gz_i);
        v = v / (gabs * gabs + gz_j * gz_j);
        v = -(double)parg[iG] * (double)parg[jG] * v;
        if (glf[3 * jG + 2] == glf[3 * iG + 2]) v = 4.0 * ORC_PI / (gabs * gabs + gz_i * gz_i) + v;
      }
      V[iG + (size_t)nlf * jG] = v;
    }
  }
}

/* a17: Gauss-Jordan inversion with full pivoting, in place (A <- A^-1, B
 * transformed alongside), Chi_RPA/gjel.f:1-91.  Column-major a(np,np),
 * b(np,mp).  A singular pivot is reported by the reference and then divided
 * through (gjel.f:60-63); the repeated-pivot case stops (gjel.f:40-43),
 * here the routine returns early leaving A partially reduced.              */
void orc_gjel(orc_z *a, int n, int np, orc_z *b, int m, int mp) {
  (void)mp;
  int *indxc = (int *)malloc(sizeof(int) * n), *indxr = (int *)malloc(sizeof(int) * n);
  int *ipiv = (int *)calloc((size_t)n, sizeof(int));
  int irow = 0, icol = 0;
#define A(i, j) a[(i) + (size_t)np * (j)]
#define B(i, j) b[(i) + (size_t)np * (j)]
  for (int i = 0; i < n; ++i) {
    double big = 0.0;
    for (int j = 0; j < n; ++j) {
      if (ipiv[j] != 1) {
        for (int k = 0; k < n; ++k) {
          if (ipiv[k] == 0) {
            if (cabs(A(j, k)) >= big) {
              big = cabs(A(j, k));
              irow = j;
              icol = k;
            }
          } else if (ipiv[k] > 1) {
            goto done; /* 'singular matrix in gjel' + stop */
          }
        }
      }
    }
    ipiv[icol] += 1;
    if (irow != icol) {
      for (int l = 0; l < n; ++l) {
        orc_z d = A(irow, l);
        A(irow, l) = A(icol, l);
        A(icol, l) = d;
      }
      for (int l = 0; l < m; ++l) {
        orc_z d = B(irow, l);
        B(irow, l) = B(icol, l);
        B(icol, l) = d;
      }
    }
    indxr[i] = irow;
    indxc[i] = icol;
    orc_z pivinv = 1.0 / A(icol, icol);
    A(icol, icol) = 1.0;
    for (int l = 0; l < n; ++l) A(icol, l) = A(icol, l) * pivinv;
    for (int l = 0; l < m; ++l) B(icol, l) = B(icol, l) * pivinv;
    for (int ll = 0; ll < n; ++ll) {
      if (ll != icol) {
        orc_z d = A(ll, icol);
        A(ll, icol) = 0.0;
        for (int l = 0; l < n; ++l) A(ll, l) = A(ll, l) - A(icol, l) * d;
        for (int l = 0; l < m; ++l) B(ll, l) = B(ll, l) - B(icol, l) * d;
      }
    }
  }
  for (int l = n - 1; l >= 0; --l) {
    if (indxr[l] != indxc[l]) {
      for (int k = 0; k < n; ++k) {
        orc_z d = A(k, indxr[l]);
        A(k, indxr[l]) = A(k, indxc[l]);
        A(k, indxc[l]) = d;
      }
    }
  }
done:
#undef A
#undef B
  free(indxc);
  free(indxr);
  free(ipiv);
}

/* a16-a19: eps = I - chi0 V (Sloss.f90:1210-1234), eps <- eps^-1 by gjel
 * (Sloss.f90:462), chi = eps^-1 chi0 (1236-1253), W = V + V chi V
 * (1255-1277).  All (Nlf,Nlf) column-major.                                 */
void orc_dyson_w(int nlf, const orc_z *chi0, const double *V, orc_z *epsinv,
                 orc_z *chi, orc_z *W) {
  size_t nn = (size_t)nlf * nlf;
  orc_z *e = (orc_z *)malloc(sizeof(orc_z) * nn);
  orc_z *bm = (orc_z *)calloc(nn, sizeof(orc_z));
  orc_z *c = (orc_z *)malloc(sizeof(orc_z) * nn);
  for (int iG = 0; iG < nlf; ++iG)
    for (int jG = 0; jG < nlf; ++jG) {
      orc_z d = (iG == jG) ? 1.0 : 0.0;
      for (int k = 0; k < nlf; ++k) d = d - chi0[iG + (size_t)nlf * k] * V[k + (size_t)nlf * jG];
      e[iG + (size_t)nlf * jG] = d;
    }
  orc_gjel(e, nlf, nlf, bm, nlf, nlf);
  for (int iG = 0; iG < nlf; ++iG)
    for (int jG = 0; jG < nlf; ++jG) {
      orc_z s = 0.0;
      for (int k = 0; k < nlf; ++k) s = s + e[iG + (size_t)nlf * k] * chi0[k + (size_t)nlf * jG];
      c[iG + (size_t)nlf * jG] = s;
    }
  if (W)
    for (int iG = 0; iG < nlf; ++iG)
      for (int jG = 0; jG < nlf; ++jG) {
        orc_z s = 0.0;
        for (int k1 = 0; k1 < nlf; ++k1)
          for (int k2 = 0; k2 < nlf; ++k2)
            s = s + V[iG + (size_t)nlf * k1] * c[k1 + (size_t)nlf * k2] * V[k2 + (size_t)nlf * jG];
        W[iG + (size_t)nlf * jG] = V[iG + (size_t)nlf * jG] + s;
      }
  if (epsinv) memcpy(epsinv, e, sizeof(orc_z) * nn);
  if (chi) memcpy(chi, c, sizeof(orc_z) * nn);
  free(e);
  free(bm);
  free(c);
}

/* f4: the Fock kernel of Chi_ladd, Chi_ladd_up/BSE.f90 in its statement
 * order: epsq (:305), alpha from wt(2,1,1) (:542), per pair (ik1, ik2) the
 * transfer q and iq (:672-676, pairs past nq skipped and left 0 here), the
 * contraction and the Q = 0 diagonal (:804-816).  wt(nq,Nlf,Nlf),
 * mnm(2,Nlf,nk,nk), fockk(nk,nk), all column-major.                       */
void orc_fock_kernel(int nk, int nlf, int nq, int ntot, double gcar, double qmin, double c0,
                     double vcell, const double *kbz, const orc_z *wt, const orc_z *mnm,
                     orc_z *fockk) {
  const double pi = 3.141592654;
  double epsq = sqrt(4.0 * pi * c0 / (vcell * ntot));
  double vq = (2.0 * pi) / qmin;
  double alpha = (c0 * vq / creal(wt[1]) - 1.0) / qmin;
  for (int ik1 = 0; ik1 < nk; ++ik1)
    for (int ik2 = 0; ik2 < nk; ++ik2) {
      size_t f = (size_t)ik1 + (size_t)nk * ik2;
      fockk[f] = 0.0;
      double qx = kbz[3 * ik2] - kbz[3 * ik1], qy = kbz[3 * ik2 + 1] - kbz[3 * ik1 + 1];
      double q = gcar * sqrt(qx * qx + qy * qy);
      int iq = (int)(q / qmin) + 1;
      if (iq > nq) continue;
      if (ik1 != ik2) {
        if (iq < 2) continue; /* wt(1,:,:) is never read by the reference */
        const orc_z *m = mnm + 2 * (size_t)nlf * f;
        orc_z s = 0.0;
        for (int ig = 0; ig < nlf; ++ig)
          for (int jg = 0; jg < nlf; ++jg) {
            orc_z w = wt[(size_t)(iq - 1) + (size_t)nq * (ig + (size_t)nlf * jg)];
            s = s - w * conj(m[2 * ig]) * m[2 * jg + 1];
          }
        fockk[f] = s;
      } else {
        fockk[f] = -vcell * ntot * log(1.0 + alpha * epsq) / alpha;
      }
    }
}

/* f4: the ladder frequency loop of Chi_ladd, Chi_ladd_up/BSE.f90:1116-1176,
 * in its statement order (fock and xi0, gjel, lladd = fock^-1 xi0, the two
 * closures); the reference's COMPLEX*8 arrays are FP64 here.              */
void orc_ladder(int nk, int no, double omin, double domega, double eta, int ntot, double vcell,
                double valley, const orc_z *fockk, const double *delta, const orc_z *current,
                orc_z *chi_ladd, orc_z *chi_para) {
  const size_t nn = (size_t)nk * nk;
  orc_z *fock = (orc_z *)malloc(sizeof(orc_z) * nn);
  orc_z *xi0 = (orc_z *)malloc(sizeof(orc_z) * nn);
  orc_z *lladd = (orc_z *)malloc(sizeof(orc_z) * nn);
  orc_z *unit = (orc_z *)malloc(sizeof(orc_z) * nn);
#define M_(A, i, j) (A)[(size_t)(i) + (size_t)nk * (j)]
  for (int io = 1; io <= no; ++io) {
    double o = omin + (io - 1) * domega;
    for (int i = 0; i < nk; ++i)
      for (int j = 0; j < nk; ++j) M_(unit, i, j) = (i == j) ? 1.0 : 0.0;
    for (int i = 0; i < nk; ++i) {
      orc_z l0i = 1.0 / (o + delta[i] + I * eta);
      for (int j = 0; j < nk; ++j) {
        orc_z l0j = 1.0 / (o + delta[j] + I * eta);
        M_(fock, i, j) = M_(unit, i, j) - l0i * M_(fockk, i, j) / (ntot * vcell);
        M_(xi0, i, j) = l0i * M_(fockk, i, j) * l0j;
      }
    }
    orc_gjel(fock, nk, nk, unit, nk, nk);
    for (int i = 0; i < nk; ++i)
      for (int j = 0; j < nk; ++j) {
        orc_z s = 0.0;
        for (int k = 0; k < nk; ++k) s += M_(fock, i, k) * M_(xi0, k, j);
        M_(lladd, i, j) = s;
      }
    for (int pol = 0; pol < 2; ++pol) {
      orc_z cl = 0.0, cp = 0.0;
      for (int i = 0; i < nk; ++i) {
        orc_z l0i = (o / delta[i]) * (1.0 / (o + delta[i] + I * eta));
        orc_z j1 = current[pol + 2 * (size_t)i];
        for (int j = 0; j < nk; ++j) {
          orc_z j2 = conj(current[pol + 2 * (size_t)j]);
          cl = cl - valley * j1 * M_(lladd, i, j) * j2 / (ntot * vcell * ntot * vcell);
        }
        cp = cp + valley * j1 * l0i * conj(j1) / (ntot * vcell);
      }
      chi_ladd[pol + 2 * (size_t)(io - 1)] = cl;
      chi_para[pol + 2 * (size_t)(io - 1)] = cp;
    }
  }
#undef M_
  free(fock);
  free(xi0);
  free(lladd);
  free(unit);
}

/* f4: genGammaPropagator, Modules/propagators.f90:11-108, its statement
 * order: the weights as the module's chain of plain ifs sets them (later
 * ifs override earlier ones), pi = 3.141592654 as the module declares it. */
void orc_gamma_propagator(int no, double domega, const orc_z *D, orc_z *gp, orc_z *gm,
                          double *kks, double *skk) {
  const double pi = 3.141592654;
  double *S = (double *)malloc(sizeof(double) * (size_t)no);
  for (int j = 0; j < no; ++j) S[j] = -(1.0 / pi) * cimag(D[j]);
  double KKS = 0.0, SKK = 0.0, fact = 0.0;
  for (int io = 1; io <= no - 1; ++io) {
    double oi = (io - 1) * domega, W1 = 0.0, W2 = 0.0;
    if (io == 1) {
      for (int jo = 2; jo <= no; ++jo) {
        double oj = (jo - 1) * domega;
        fact = domega / oj;
        if (jo == 2) fact = 3.0 / 2.0;
        if (jo == no) fact = 0.5 * domega / oj;
        W1 = W1 - fact * S[jo - 1];
      }
      W2 = -W1;
    } else if (io == 2) {
      for (int jo = 1; jo <= no; ++jo) {
        double oj = (jo - 1) * domega;
        if (jo != io) fact = domega / (oi - oj);
        if (jo == 1) fact = 1.0;
        if (jo == 2) fact = 0.0;
        if (jo == 3) fact = -3.0 / 2.0;
        if (jo == no) fact = 0.5 * domega / (oi - oj);
        W1 = W1 + fact * S[jo - 1];
        fact = domega / (oi + oj);
        if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
        W2 = W2 + fact * S[jo - 1];
      }
    } else if (io == no - 1) {
      for (int jo = 1; jo <= no; ++jo) {
        double oj = (jo - 1) * domega;
        if (jo != io) fact = domega / (oi - oj);
        if (jo == 1) fact = 0.5 * domega / (oi - oj);
        if (jo == no - 2) fact = 3.0 / 2.0;
        if (jo == no - 1) fact = 0.0;
        if (jo == no) fact = -1.0;
        W1 = W1 + fact * S[jo - 1];
        fact = domega / (oi + oj);
        if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
        W2 = W2 + fact * S[jo - 1];
      }
    } else {
      for (int jo = 1; jo <= no; ++jo) {
        double oj = (jo - 1) * domega;
        if (jo != io) fact = domega / (oi - oj);
        if (jo == 1) fact = 0.5 * domega / (oi - oj);
        if (jo == io - 1) fact = 3.0 / 2.0;
        if (jo == io) fact = 0.0;
        if (jo == io + 1) fact = -3.0 / 2.0;
        if (jo == no) fact = 0.5 * domega / (oi - oj);
        W1 = W1 + fact * S[jo - 1];
        fact = domega / (oi + oj);
        if (jo == 1 || jo == no) fact = 0.5 * domega / (oi + oj);
        W2 = W2 + fact * S[jo - 1];
      }
    }
    gp[io - 1] = W1 + I * (-pi * S[io - 1]);
    gm[io - 1] = -W2 + I * 0.0;
    double wind = creal(D[io - 1]);
    double windkk = creal(gp[io - 1]) + creal(gm[io - 1]);
    fact = domega;
    if (io == 1 || io == no - 1) fact = 0.5 * domega;
    KKS = KKS + fact * (windkk - wind) * (windkk - wind);
    SKK = SKK + fact * wind * wind;
  }
  *kks = KKS;
  *skk = SKK;
  free(S);
}

/* f3: Photon.x.  zgetrf (partial pivoting on |Re| + |Im|, the first
 * maximum) then the inverse column by column from the LU factors: the same
 * matrix zgetri returns, in another operation order.                       */
int orc_lu_invert(orc_z *a, int n) {
#define A_(i, j) a[(size_t)(i) + (size_t)n * (j)]
  int *ip = (int *)malloc(sizeof(int) * (size_t)n);
  int info = 0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = -1.0;
    for (int i = k; i < n; ++i) {
      double m = fabs(creal(A_(i, k))) + fabs(cimag(A_(i, k)));
      if (m > best) {
        best = m;
        p = i;
      }
    }
    ip[k] = p;
    if (best == 0.0) {
      if (!info) info = k + 1;
      continue;
    }
    if (p != k)
      for (int j = 0; j < n; ++j) {
        orc_z t = A_(k, j);
        A_(k, j) = A_(p, j);
        A_(p, j) = t;
      }
    for (int i = k + 1; i < n; ++i) {
      A_(i, k) = A_(i, k) / A_(k, k);
      for (int j = k + 1; j < n; ++j) A_(i, j) = A_(i, j) - A_(i, k) * A_(k, j);
    }
  }
  if (!info) {
    orc_z *lu = (orc_z *)malloc(sizeof(orc_z) * (size_t)n * n);
    memcpy(lu, a, sizeof(orc_z) * (size_t)n * n);
    for (int c = 0; c < n; ++c) {
      orc_z *x = a + (size_t)n * c; /* column c of the inverse: solve A x = e_c */
      for (int i = 0; i < n; ++i) x[i] = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < n; ++k)
        if (ip[k] != k) {
          orc_z t = x[k];
          x[k] = x[ip[k]];
          x[ip[k]] = t;
        }
      for (int i = 0; i < n; ++i)
        for (int k = 0; k < i; ++k) x[i] = x[i] - lu[i + (size_t)n * k] * x[k];
      for (int i = n - 1; i >= 0; --i) {
        for (int k = i + 1; k < n; ++k) x[i] = x[i] - lu[i + (size_t)n * k] * x[k];
        x[i] = x[i] / lu[i + (size_t)n * i];
      }
    }
    free(lu);
  }
  free(ip);
  return info;
#undef A_
}

static const double ORC_GAMMA = 1.0 / 137.0; /* Photonew.f90:38 */

/* genD0, Photonew.f90:618-663, in its statement order */
void orc_photon_d0(double q, double o, double eta, double c0, int nlf, const int *parg,
                   const double *glfz, orc_z *D) {
  const size_t nn = (size_t)nlf * nlf;
  orc_z oi = o + I * eta;
  orc_z beta = csqrt(ORC_GAMMA * ORC_GAMMA * oi * oi - q * q + 0.0 * I);
  for (int i = 0; i < nlf; ++i) {
    orc_z aii = parg[i] * (1.0 - cexp(I * beta * c0)) / (beta * beta - glfz[i] * glfz[i]);
    for (int j = 0; j < nlf; ++j) {
      orc_z ajj = aii * parg[j] / (beta * beta - glfz[j] * glfz[j]);
      orc_z aij = 2.0 * (beta * beta + glfz[i] * glfz[j]) * ajj / c0;
      orc_z bij = 2.0 * beta * (glfz[i] + glfz[j]) * ajj / c0;
      orc_z d1 = 0.0, d2 = 0.0;
      if (i == j) {
        d1 = 2.0 * I * beta / (beta * beta - glfz[i] * glfz[i]);
        d2 = 2.0 * I * glfz[i] / (beta * beta - glfz[i] * glfz[i]);
      }
      size_t x = (size_t)i + (size_t)nlf * j;
      D[x] = 2.0 * ORC_PI * I * ORC_GAMMA * ORC_GAMMA * (aij + d1) / beta;
      D[nn + x] = 2.0 * ORC_PI * I * beta * (aij + d1) / (oi * oi);
      D[3 * nn + x] = -2.0 * ORC_PI * I * q * (bij + d2) / (oi * oi);
      D[2 * nn + x] = 2.0 * ORC_PI * I * q * q * (aij + d1) / (beta * oi * oi);
    }
    D[2 * nn + (size_t)i * (nlf + 1)] -= 4.0 * ORC_PI / (oi * oi);
  }
}

int orc_photon_point(double q, double o, double eta, double c0, double theta, int nlf,
                     const int *parg, const double *glfz, const orc_z *pi0, orc_z *pi,
                     orc_z *sums, double *spec) {
  const int N = nlf, M = 2 * nlf;
  const size_t nn = (size_t)N * N;
  orc_z *D = (orc_z *)malloc(sizeof(orc_z) * 4 * nn);
  orc_z *TS = (orc_z *)malloc(sizeof(orc_z) * nn);
  orc_z *TP = (orc_z *)malloc(sizeof(orc_z) * 4 * nn);
  orc_photon_d0(q, o, eta, c0, N, parg, glfz, D);
  const orc_z *Dxx = D, *Dyy = D + nn, *Dzz = D + 2 * nn, *Dyz = D + 3 * nn, *Dzy = Dyz;
  const orc_z *Pxx = pi0, *Pyy = pi0 + nn, *Pyz = pi0 + 2 * nn, *Pzy = pi0 + 3 * nn, *Pzz = pi0 + 4 * nn;
#define E_(A, i, j) (A)[(size_t)(i) + (size_t)N * (j)]
  /* genTS (890-917) */
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      orc_z s = 0.0;
      for (int k = 0; k < N; ++k) s += E_(Pxx, i, k) * E_(Dxx, k, j);
      E_(TS, i, j) = (i == j ? 1.0 : 0.0) - s;
    }
  int st_s = orc_lu_invert(TS, N);
  /* genTP (919-966) */
  for (int I2 = 0; I2 < M; ++I2)
    for (int J2 = 0; J2 < M; ++J2) {
      orc_z s = 0.0;
      int i = I2 % N, j = J2 % N;
      for (int k = 0; k < N; ++k) {
        if (I2 < N && J2 < N)
          s += E_(Pyy, i, k) * E_(Dyy, k, j) + E_(Pyz, i, k) * E_(Dzy, k, j);
        else if (I2 < N)
          s += E_(Pyy, i, k) * E_(Dyz, k, j) + E_(Pyz, i, k) * E_(Dzz, k, j);
        else if (J2 < N)
          s += E_(Pzy, i, k) * E_(Dyy, k, j) + E_(Pzz, i, k) * E_(Dzy, k, j);
        else
          s += E_(Pzy, i, k) * E_(Dyz, k, j) + E_(Pzz, i, k) * E_(Dzz, k, j);
      }
      TP[(size_t)I2 + (size_t)M * J2] = (I2 == J2 ? 1.0 : 0.0) - s;
    }
  int st_p = orc_lu_invert(TP, M);
#define T_(i, j) TP[(size_t)(i) + (size_t)M * (j)]
  /* genScreenedPi (744-787) */
  orc_z *Xx = pi, *Yy = pi + nn, *Yz = pi + 2 * nn, *Zy = pi + 3 * nn, *Zz = pi + 4 * nn;
  for (size_t x = 0; x < 5 * nn; ++x) pi[x] = 0.0;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int k = 0; k < N; ++k) {
        E_(Xx, i, j) += E_(TS, i, k) * E_(Pxx, k, j);
        E_(Yy, i, j) += T_(i, k) * E_(Pyy, k, j) + T_(i, k + N) * E_(Pzy, k, j);
        E_(Yz, i, j) += T_(i, k) * E_(Pyz, k, j) + T_(i, k + N) * E_(Pzz, k, j);
        E_(Zy, i, j) += T_(i + N, k) * E_(Pyy, k, j) + T_(i + N, k + N) * E_(Pzy, k, j);
        E_(Zz, i, j) += T_(i + N, k) * E_(Pyz, k, j) + T_(i + N, k + N) * E_(Pzz, k, j);
      }
  /* sumIntScreenedPi (826-888): order xx, yy, zz, yz, zy */
  const orc_z *mats[5] = {Xx, Yy, Zz, Yz, Zy};
  for (int m = 0; m < 5; ++m) sums[m] = 0.0;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      orc_z w;
      if (i == 0 && j != 0)
        w = (-I * (-1.0 + cexp(I * (glfz[j] * c0 / 2)))) * c0 / (2 * glfz[j]);
      else if (i != 0 && j == 0)
        w = (-I * (1.0 - cexp(-I * (glfz[i] * c0 / 2)))) * c0 / (2 * glfz[i]);
      else if (i == 0 && j == 0)
        w = c0 * c0 / 4;
      else
        w = -(-1.0 + cexp(I * (glfz[i] * c0 / 2))) * (-1.0 + cexp(I * (glfz[j] * c0 / 2))) /
            (glfz[i] * glfz[j] * cexp(I * (glfz[i] * c0 / 2)));
      /* the last branch subtracts P (.)(.)/(.) in the reference: w carries the sign */
      for (int m = 0; m < 5; ++m) sums[m] += E_(mats[m], i, j) * w;
    }
  for (int m = 0; m < 5; ++m) sums[m] = 1 / c0 * sums[m];
  /* genSpectra (968-1064), p mode */
  orc_z oi = o + I * eta;
  orc_z beta = csqrt(ORC_GAMMA * ORC_GAMMA * oi * oi - q * q + 0.0 * I);
  orc_z dyy_r = 0.0, dzz_r = 0.0, dyy_a = 0.0, dzz_a = 0.0, dyz_a = 0.0;
  for (int i = 0; i < N; ++i) {
    double ff1 = creal((2.0 * parg[i] / sqrt(c0)) * csin(beta * c0 / 2.0) / (beta + glfz[i]));
    for (int j = 0; j < N; ++j) {
      double ff2 = creal((2.0 * parg[j] / sqrt(c0)) * csin(beta * c0 / 2.0) / (beta + glfz[j]));
      double gf2 = creal((2.0 * parg[j] / sqrt(c0)) * csin(beta * c0 / 2.0) / (beta - glfz[j]));
      dyy_r += ff1 * E_(Yy, i, j) * gf2;
      dzz_r += ff1 * E_(Zz, i, j) * gf2;
      dyy_a += ff1 * E_(Yy, i, j) * ff2;
      dzz_a += ff1 * E_(Zz, i, j) * ff2;
      dyz_a += ff1 * E_(Yz, i, j) * ff2;
    }
  }
  orc_z dyy0 = 2 * ORC_PI * ORC_GAMMA * I / oi;
  orc_z ref_p = dyy0 * (dyy_r * cos(theta) - dzz_r * sin(theta) * sin(theta) / cos(theta));
  orc_z abso_p = cos(theta) * dyy_a * cos(theta) + sin(theta) * dzz_a * sin(theta) -
                 2.0 * cos(theta) * dyz_a * sin(theta);
  spec[0] = (4.0 * ORC_PI * ORC_GAMMA / o) * cimag(abso_p);
  spec[1] = creal(ref_p * conj(ref_p));
#undef T_
#undef E_
  free(D);
  free(TS);
  free(TP);
  return st_s ? st_s : (st_p ? -st_p : 0);
}

/* f2: interpolate.x, Chi_RPA/interpolate.f90:103-186.  Pi0(no-1,Nlf,Nlf) is
 * (ReChi0, ImChi0) of every io = 1..No-1: interpolate.f90's contained
 * genReChi0/genImChi0 (:242-464) are Optabs' K1, i.e. orc_kk_optabs.
 * Pi_tot(io) = Pi0(io) + Pi_dia with Pi_dia = -Re Pi0(1) (:116-123); then
 * for io = 1..No-2, jo = 1..No_interp-1 (:166-181):
 *   counter += 1, oi = counter domega_tot,
 *   pi_interp(counter) = Pi_tot(io) + (Pi_tot(io+1) - Pi_tot(io))/No_interp jo
 *                        + Qeff oi/(oi + i eta_intra)
 * pi_interp(No_tot,Nlf,Nlf); pi_intra(No_tot) = Qeff(1,1) oi/(oi + i eta).  */
void orc_interpolate(int no, int nlf, int no_interp, double domega_tot, double eta_intra,
                     const orc_z *Pi0, const orc_z *Qeff, orc_z *pi_interp, orc_z *pi_intra) {
  const size_t nn = (size_t)nlf * nlf, nf = (size_t)(no - 1);
  const size_t no_tot = (size_t)(no_interp - 1) * (size_t)(no - 2);
  orc_z *tot = (orc_z *)malloc(sizeof(orc_z) * nf * nn);
  for (size_t x = 0; x < nn; ++x) {
    orc_z dia = -(creal(Pi0[nf * x]) + I * 0.0);
    for (size_t io = 0; io < nf; ++io) tot[io + nf * x] = Pi0[io + nf * x] + dia;
  }
  size_t counter = 0;
  for (int io = 1; io <= no - 2; ++io)
    for (int jo = 1; jo <= no_interp - 1; ++jo) {
      counter += 1;
      double oi = (double)counter * domega_tot;
      orc_z den = oi + I * eta_intra;
      for (size_t x = 0; x < nn; ++x) {
        orc_z a = tot[(io - 1) + nf * x], b = tot[io + nf * x];
        orc_z tmp = (b - a) / (double)no_interp;
        orc_z v = a + tmp * (double)jo;
        pi_interp[(counter - 1) + no_tot * x] = v + Qeff[x] * oi / den;
      }
      if (pi_intra) pi_intra[counter - 1] = Qeff[0] * oi / den;
    }
  free(tot);
}

/* a13: Pi assembly, Optabs.f90:575-611.  Pi0(no-1,Nlf,Nlf) from orc_kk_optabs;
 * pi11[io-1] = Pi_tot(1,1); sig_inter/sig_intra = Re(-i c0 Pi/oi) (io > 1,
 * entry 0 left 0).  Pi_inter is read back after each (iG,jG) update exactly as
 * the reference does (Optabs.f90:594).                                     */
void orc_pi_assembly(int no, int nlf, double domega, double eta_intra, double c0,
                     const orc_z *Pi0, const orc_z *Qeff, orc_z *pi11,
                     double *sig_inter, double *sig_intra) {
  size_t nn = (size_t)nlf * nlf;
  orc_z *dia = (orc_z *)calloc(nn, sizeof(orc_z));
  orc_z *tot = (orc_z *)calloc(nn, sizeof(orc_z));
  for (int io = 1; io <= no - 1; ++io) {
    double oi = (double)(io - 1) * domega;
    orc_z pinter = 0.0, pintra = 0.0;
    for (int iG = 0; iG < nlf; ++iG)
      for (int jG = 0; jG < nlf; ++jG) {
        orc_z p0 = Pi0[(io - 1) + (size_t)(no - 1) * (iG + (size_t)nlf * jG)];
        size_t x = iG + (size_t)nlf * jG;
        if (io == 1) dia[x] = -(creal(p0) + I * 0.0);
        tot[x] = p0;
        tot[x] = tot[x] + dia[x];
        pinter = tot[0];
        orc_z den = oi + I * eta_intra;
        pintra = Qeff[0] * oi / den;
        tot[x] = tot[x] + Qeff[x] * oi / den;
      }
    pi11[io - 1] = tot[0];
    if (io > 1) {
      sig_inter[io - 1] = creal(-I * c0 * pinter / oi);
      sig_intra[io - 1] = creal(-I * c0 * pintra / oi);
    } else {
      sig_inter[0] = 0.0;
      sig_intra[0] = 0.0;
    }
  }
  free(dia);
  free(tot);
}
