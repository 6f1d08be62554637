// Host-side setup of the RPA path: the reference's setup subroutines that
// run before its k-loop (FBZ, q, k+q map, local fields, k-sphere sort).
// Integer / tolerance logic on small arrays; no device work.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/rpa_hip.h"

namespace {

// G(n) = sum_m KC(n,m) mill(m) accumulated left to right (Optabs_qe6.f90:1162-1167)
inline void kc_times(const double *kc, const int *m, double *g) {
  for (int a = 0; a < 3; ++a) {
    double s = 0.0;
    for (int b = 0; b < 3; ++b) s = s + kc[a + 3 * b] * (double)m[b];
    g[a] = s;
  }
}

inline bool near3(const double *x, const double *y, double eps) {
  return std::fabs(x[0] - y[0]) <= eps && std::fabs(x[1] - y[1]) <= eps && std::fabs(x[2] - y[2]) <= eps;
}

}  // namespace

extern "C" {

int rpa_gcart(int n, const int *mill, const double *kc, double *g) {
  if (n < 0 || (n && (!mill || !kc || !g))) return RPA_E_ARG;
  for (int i = 0; i < n; ++i) kc_times(kc, mill + 3 * i, g + 3 * i);
  return RPA_OK;
}

// genFBZ (Optabs.f90:871-933, Sloss.f90:697-759): apply every point operation to every IBZ k and
// keep the distinct results in order of appearance.
int rpa_gen_fbz(int nki, int nsym, double eps, const double *ki, const double *R, int *ntot,
                double *ktot) {
  if (nki <= 0 || nsym <= 0 || nsym > 48 || !ki || !R || !ntot || !ktot) return RPA_E_ARG;
  std::vector<double> all;
  all.reserve(3 * (size_t)nki * nsym);
  int nt = 0;
  for (int s = 0; s < nsym; ++s)
    for (int k = 0; k < nki; ++k) {
      double v[3];
      for (int n = 0; n < 3; ++n) {
        double acc = 0.0;
        for (int m = 0; m < 3; ++m) acc = acc + R[s + 48 * (n + 3 * m)] * ki[3 * k + m];
        v[n] = acc;
      }
      bool seen = false;
      for (size_t j = 0; j < all.size() / 3 && !seen; ++j) seen = near3(v, &all[3 * j], eps);
      all.insert(all.end(), v, v + 3);
      if (!seen) {
        std::memcpy(ktot + 3 * (size_t)nt, v, sizeof v);
        ++nt;
      }
    }
  *ntot = nt;
  return RPA_OK;
}

// findMinQ (Optabs.f90:743-773): the shortest nonzero FBZ vector below 1.0,
// first one wins; q = (iq-1) k_min.
int rpa_find_min_q(int ntot, const double *ktot, int iq, double *q) {
  if (ntot <= 0 || !ktot || !q) return RPA_E_ARG;
  double kmin = 1.0;
  int imin = -1;
  for (int i = 0; i < ntot; ++i) {
    const double *k = ktot + 3 * i;
    double kr = std::sqrt((k[0] * k[0] + k[1] * k[1]) + k[2] * k[2]);
    if (kr == 0.0) continue;
    if (kr < kmin) {
      kmin = kr;
      imin = i;
    }
  }
  if (imin < 0) return RPA_E_NOTFOUND;
  for (int a = 0; a < 3; ++a) q[a] = (double)(iq - 1) * ktot[3 * imin + a];
  return RPA_OK;
}

// findKinIBZ + findKQinBZ for every FBZ k, identity symmetry
// (Optabs.f90:776-869): K1 = k; the first (G0, k') in G-major order with
// k + q - G0 = k' within eps, then k' located in the IBZ list.
int rpa_find_kq(int ntot, const double *ktot, int nki, const double *ki, int ng, const double *g,
                const double *q, double eps, int *K1, int *K2, int *iG0) {
  if (ntot <= 0 || nki <= 0 || ng <= 0 || !ktot || !ki || !g || !q || !K1 || !K2 || !iG0)
    return RPA_E_ARG;
  if (ntot > nki) return RPA_E_ARG;  // Nsym = 1: the FBZ is the IBZ list
  for (int k = 0; k < ntot; ++k) {
    K1[k] = k + 1;  // findKinIBZ: ik <= NkI gives R1 = 1, K1 = ik
    const double kq[3] = {ktot[3 * k] + q[0], ktot[3 * k + 1] + q[1], ktot[3 * k + 2] + q[2]};
    int found = 0;
    for (int ig = 0; ig < ng && !found; ++ig) {
      const double *G = g + 3 * ig;
      for (int jk = 0; jk < ntot; ++jk) {
        const double *kk = ktot + 3 * jk;
        if (std::fabs(kq[0] - G[0] - kk[0]) <= eps && std::fabs(kq[1] - G[1] - kk[1]) <= eps &&
            std::fabs(kq[2] - G[2] - kk[2]) <= eps) {
          iG0[k] = ig + 1;
          found = 1;
          for (int j = 0; j < nki; ++j)
            if (near3(kk, ki + 3 * j, eps)) {
              K2[k] = j + 1;
              found = 2;
              break;
            }
          break;
        }
      }
    }
    if (found != 2) return RPA_E_NOTFOUND;  // 'Can not find wave vector K+Q'
  }
  return RPA_OK;
}

// PointR (Chi_RPA/PointR.f90:119-138): RI(i) = R(i)^-1 for every operation
// by Gauss-Jordan elimination with full pivoting in gjel's order
// (Chi_RPA/gjel.f: the pivot is the last entry of largest modulus in a
// row-by-row scan of the rows and columns not yet used; rows are swapped to
// put it on the diagonal, columns restored at the end).  PointR runs it on a
// complex copy whose imaginary parts stay zero, so real arithmetic gives the
// same values.  A zero pivot (gjel's 'singular matrix') is an error here.
int rpa_point_inverse(int nsym, const double *R, double *RI) {
  if (nsym <= 0 || nsym > 48 || !R || !RI) return RPA_E_ARG;
  for (int op = 0; op < nsym; ++op) {
    double a[3][3];
    for (int n = 0; n < 3; ++n)
      for (int m = 0; m < 3; ++m) a[n][m] = R[op + 48 * (n + 3 * m)];
    int used[3] = {0, 0, 0}, prow[3], pcol[3];
    for (int step = 0; step < 3; ++step) {
      double big = 0.0;
      int ir = -1, ic = -1;
      for (int j = 0; j < 3; ++j) {
        if (used[j]) continue;
        for (int k = 0; k < 3; ++k)
          if (!used[k] && std::fabs(a[j][k]) >= big) {
            big = std::fabs(a[j][k]);
            ir = j;
            ic = k;
          }
      }
      if (ic < 0 || a[ir][ic] == 0.0) return RPA_E_SINGULAR;
      used[ic] = 1;
      if (ir != ic)
        for (int l = 0; l < 3; ++l) std::swap(a[ir][l], a[ic][l]);
      prow[step] = ir;
      pcol[step] = ic;
      const double inv = 1.0 / a[ic][ic];
      a[ic][ic] = 1.0;
      for (int l = 0; l < 3; ++l) a[ic][l] *= inv;
      for (int r = 0; r < 3; ++r) {
        if (r == ic) continue;
        const double f = a[r][ic];
        a[r][ic] = 0.0;
        for (int l = 0; l < 3; ++l) a[r][l] -= a[ic][l] * f;
      }
    }
    for (int step = 2; step >= 0; --step)
      if (prow[step] != pcol[step])
        for (int k = 0; k < 3; ++k) std::swap(a[k][prow[step]], a[k][pcol[step]]);
    for (int n = 0; n < 3; ++n)
      for (int m = 0; m < 3; ++m) RI[op + 48 * (n + 3 * m)] = a[n][m];
  }
  return RPA_OK;
}

namespace {
// RI(op) v accumulated left to right, as the reference's sums
inline void rot3(const double *RI, int op, const double *v, double *w) {
  for (int n = 0; n < 3; ++n)
    w[n] = (RI[op + 48 * n] * v[0] + RI[op + 48 * (n + 3)] * v[1]) + RI[op + 48 * (n + 6)] * v[2];
}
}  // namespace

// findKinIBZ + findKQinBZ with the point operations, for every FBZ k
// (Optabs.f90:306-320, 776-869):
//   findKinIBZ: k <= NkI is the IBZ point itself (R1 = 1); otherwise the first
//     operation i = 2..Nsym and IBZ point j with RI(i) k = kI(j) within eps;
//   findKQinBZ: the first (G0, k') in G-major order with k + q - G0 = k'
//     within eps for which some operation i = 1..Nsym and IBZ point j give
//     RI(i) k' = kI(j) (the first such i, then j).
// RPA_E_NOTFOUND at the reference's stops ('Can not find wave vector ...').
int rpa_find_kq_sym(int ntot, const double *ktot, int nki, const double *ki, int ng, const double *g,
                    const double *q, int nsym, const double *RI, double eps, int *R1, int *K1, int *R2,
                    int *K2, int *iG0) {
  if (ntot <= 0 || nki <= 0 || ng <= 0 || nsym <= 0 || nsym > 48 || !ktot || !ki || !g || !q || !RI ||
      !R1 || !K1 || !R2 || !K2 || !iG0)
    return RPA_E_ARG;
  auto in_ibz = [&](const double *v) -> int {
    for (int j = 0; j < nki; ++j)
      if (near3(v, ki + 3 * j, eps)) return j;
    return -1;
  };
  for (int k = 0; k < ntot; ++k) {
    const double *kv = ktot + 3 * k;
    if (k < nki) {
      R1[k] = 1;
      K1[k] = k + 1;
    } else {
      int hit = -1;
      for (int op = 1; op < nsym && hit < 0; ++op) {
        double w[3];
        rot3(RI, op, kv, w);
        const int j = in_ibz(w);
        if (j >= 0) {
          hit = op;
          R1[k] = op + 1;
          K1[k] = j + 1;
        }
      }
      if (hit < 0) return RPA_E_NOTFOUND;  // 'Can not find wave vector K= .. in I.B.Z.'
    }
    const double kq[3] = {kv[0] + q[0], kv[1] + q[1], kv[2] + q[2]};
    int state = 1;  // 1: no k' in the FBZ, 2: k' but not in the IBZ, 3: found
    for (int ig = 0; ig < ng && state < 3; ++ig) {
      const double *G = g + 3 * ig;
      for (int jk = 0; jk < ntot && state < 3; ++jk) {
        const double *kk = ktot + 3 * jk;
        if (!(std::fabs(kq[0] - G[0] - kk[0]) <= eps && std::fabs(kq[1] - G[1] - kk[1]) <= eps &&
              std::fabs(kq[2] - G[2] - kk[2]) <= eps))
          continue;
        state = 2;
        iG0[k] = ig + 1;
        for (int op = 0; op < nsym; ++op) {
          double w[3];
          rot3(RI, op, kk, w);
          const int j = in_ibz(w);
          if (j >= 0) {
            R2[k] = op + 1;
            K2[k] = j + 1;
            state = 3;
            break;
          }
        }
      }
    }
    if (state != 3) return RPA_E_NOTFOUND;  // 'Can not find wave vector K+Q'
  }
  return RPA_OK;
}

// genGlf / genGlfandParity (Optabs.f90:1016-1059, Sloss.f90:824-881):
// 'z' keeps G with Gx = Gy = 0 and Gcar^2 Gz^2 / 2 <= Ecut, 'xyz' keeps
// Gcar^2 |G|^2 / 2 <= Ecut; parity +1 for even Miller z.
int rpa_gen_glf(const char *lf, double ecut, double gcar, int ng, const double *g, const int *mill,
                int nlfd, int *nlf, double *glf, int *parg, int *lf_idx) {
  if (!lf || ng <= 0 || !g || !nlf) return RPA_E_ARG;
  const std::string mode(lf);
  if (mode != "z" && mode != "xyz") return RPA_E_ARG;
  int n = 0;
  for (int i = 0; i < ng; ++i) {
    const double *G = g + 3 * i;
    bool take;
    if (mode == "z") {
      take = (G[0] == 0.0 && G[1] == 0.0) && (gcar * gcar * (G[2] * G[2]) / 2.0 <= ecut);
    } else {
      take = gcar * gcar * ((G[0] * G[0] + G[1] * G[1]) + G[2] * G[2]) / 2.0 <= ecut;
    }
    if (!take) continue;
    if (n < nlfd) {
      if (glf) {
        glf[3 * n] = mode == "z" ? 0.0 : G[0];
        glf[3 * n + 1] = mode == "z" ? 0.0 : G[1];
        glf[3 * n + 2] = G[2];
      }
      if (parg) {
        int z = mill ? mill[3 * i + 2] : 0;
        parg[n] = ((z / 2) * 2 == z) ? 1 : -1;
      }
      if (lf_idx) lf_idx[n] = i + 1;
    }
    ++n;
  }
  *nlf = n;
  return n > nlfd ? RPA_E_LIMIT : RPA_OK;
}

// qe_gksort.init_igk for one k (Modules/qe_gksort.f90:44-94): the G with
// |k+G|^2 <= gcutw in list order, ordered by |k+G|^2 where values closer than
// 1e-8 count as equal and keep list order.  The scan stops at the first G
// with |G|^2 > (|k| + sqrt(gcutw))^2 + 1e-8; a list that ends first is an
// error, as is overflowing npwx.
int rpa_gk_sort(const double *k, int ngm, const double *g, double gcutw, int npwx, int *ngk,
                int *igk) {
  if (!k || ngm <= 0 || !g || npwx <= 0 || !ngk || !igk) return RPA_E_ARG;
  const double eps8 = 1.0e-8;
  const double kk = std::sqrt((k[0] * k[0] + k[1] * k[1]) + k[2] * k[2]);
  const double q2x = (kk + std::sqrt(gcutw)) * (kk + std::sqrt(gcutw));
  std::vector<std::pair<double, int>> v;
  int ig = 0;
  for (; ig < ngm; ++ig) {
    const double *G = g + 3 * ig;
    const double a = k[0] + G[0], b = k[1] + G[1], c = k[2] + G[2];
    double q = (a * a + b * b) + c * c;
    if (q <= eps8) q = 0.0;
    if (q <= gcutw) {
      if ((int)v.size() >= npwx) return RPA_E_LIMIT;
      v.emplace_back(q, ig + 1);
    } else if ((G[0] * G[0] + G[1] * G[1]) + G[2] * G[2] > q2x + eps8) {
      break;
    }
  }
  if (ig >= ngm) return RPA_E_LIMIT;
  // order by value; then within runs of values closer than eps8, by index
  std::stable_sort(v.begin(), v.end(), [](const std::pair<double, int> &x,
                                          const std::pair<double, int> &y) { return x.first < y.first; });
  size_t s = 0;
  while (s < v.size()) {
    size_t e = s + 1;
    while (e < v.size() && v[e].first - v[e - 1].first < eps8) ++e;
    std::sort(v.begin() + s, v.begin() + e,
              [](const std::pair<double, int> &x, const std::pair<double, int> &y) { return x.second < y.second; });
    s = e;
  }
  for (size_t i = 0; i < v.size(); ++i) igk[i] = v[i].second;
  *ngk = (int)v.size();
  return RPA_OK;
}

}  // extern "C"
