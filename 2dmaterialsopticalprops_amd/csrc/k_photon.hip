// Photon.x (BSE/Photonew.f90) on the device: the photon Dyson equation of
// a slab for every (spectrum line, frequency) point.
//
// Per point (o = omin + (io-1) domega, oi = o + i eta, q fixed or
// gamma o sin(theta), beta = sqrt(gamma^2 oi^2 - q^2); Photonew.f90:229-249):
//   D0      free photon propagator blocks Dxx, Dyy, Dzz, Dyz = Dzy  (genD0, 618-663)
//   TS      = I - Pixx0 Dxx0                                       (genTS, 890-917)
//   TP      = I - [Piyy0 Piyz0; Pizy0 Pizz0] [Dyy0 Dyz0; Dzy0 Dzz0] (genTP, 919-966)
//   Pixx    = TS^-1 Pixx0,  [Piyy Piyz; Pizy Pizz] = TP^-1 [Piyy0 Piyz0; Pizy0 Pizz0]
//                                                                  (invert 388-462, genScreenedPi 744-787)
//   sums    slab integrals of the screened Pi over 0..c0/2           (sumIntScreenedPi, 826-888)
//   A_p, R_p absorption and reflection of the p mode                (genSpectra, 968-1064)
//
// D0 is rank two plus a diagonal in (G, G'): with u = par/(beta^2 - Gz^2)
// and e = 1 - exp(i beta c0), A(G,G') = (2e/c0)(beta^2 + Gz Gz') u u' and
// B(G,G') = (2e beta/c0)(Gz + Gz') u u'.  So Pi0 D0 needs only the two
// vectors Pi0 u and Pi0 (Gz u) per Pi0 block: the T matrices are built in
// O(N^2) per point instead of the reference's O(N^3) products.  The inverses
// are Gauss-Jordan with row pivoting in LDS (in a global scratch slice past
// 44 local fields), one workgroup per point, and the screened blocks follow
// as T^-1 Pi0 products.  Sums and spectra are reduced from the values each
// thread computes.
#include <string>

#include "dev_linalg.h"
#include "rpa_internal.h"

namespace rpa {

namespace {

constexpr int kPhMaxNlf = 64;     // local fields per point (TP is 2N x 2N)
constexpr int kPhLdsNlf = 44;     // TP in LDS up to here (64 N^2 bytes)
constexpr double kGamma = 1.0 / 137.0;  // Photonew.f90:38

// dynamic LDS of k_photon: T (LDS form) + 19 N double2 + 3 N double + 2 N int
inline size_t photon_lds_bytes(int N, bool global) {
  const size_t M = 2 * (size_t)N;
  return (global ? 0 : sizeof(double2) * M * M) + sizeof(double2) * 19 * N + sizeof(double) * 3 * N +
         sizeof(int) * 2 * N;
}

}  // namespace

struct PhotonArgs {
  int N, nf, p0;
  double omin, dw, eta, c0;
  const double *gz;     // [N] Glf(3,:) cartesian
  const int *par;       // [N] parity +-1
  const double2 *P0;    // [io][5][N][N]: xx, yy, yz, zy, zz (row-major iG, jG)
  const double *qline, *thline;
  const int *qmode;     // per line: 0 q fixed, 1 q = gamma o sin(theta)
  double2 *W;           // optional [point][5][N][N] screened Pi (chunk-relative)
  double2 *T;           // GLOBAL form: [point][2N][2N] (chunk-relative)
  double2 *pi11;        // [point][3] Pixx(1,1), Piyy(1,1), Pizz(1,1)
  double2 *psum;        // [point][3] sPixx, sPiyy, sPizz
  double *spec;         // [point][2] A_p, R_p
  int *status;
};

template <bool GLOBAL>
__global__ __launch_bounds__(256) void k_photon(PhotonArgs a) {
  // dynamic LDS: [2N][2N] T (LDS form only), then the per-field tables
  // (photon_lds_bytes), so that two workgroups share a CU at N_G = 33
  extern __shared__ double2 lds_dyn[];
  __shared__ double red_v[4];
  __shared__ int red_i[4];
  __shared__ double red[5 * 20];
  __shared__ double2 s_11[3];
  const int N = a.N, M = 2 * N, tid = threadIdx.x;
  const int lp = blockIdx.x, p = a.p0 + lp;
  const int line = p / a.nf, io = p - line * a.nf + 1;
  double2 *sa = GLOBAL ? a.T + (size_t)lp * M * M : lds_dyn;
  double2 *tab = lds_dyn + (GLOBAL ? 0 : (size_t)M * M);
  // [0,N) s_vs; [N,15N) the tables of the T builds, then (solve form) the
  // eight p-mode right-hand sides [8][2N] in [N,17N); [15N,17N) the two
  // s-mode right-hand sides; [17N,19N) fcol
  double2 *s_vs = tab, *s_aii = tab + N, *s_pd = tab + 2 * N, *s_d1 = tab + 3 * N, *s_d2 = tab + 4 * N;
  double2 *s_u = tab + 5 * N, *s_v = tab + 10 * N;  // [5][N]
  double2 *ys = tab + 15 * N, *yp = tab + N;
  double2 *fcol = tab + 17 * N;                      // [2N]
  double *s_gz = (double *)(tab + 19 * N), *s_ff = s_gz + N, *s_gf = s_gz + 2 * N;
  int *piv = (int *)(s_gz + 3 * N);                  // [2N]

  // ---- the point: frequency, q, beta (Photonew.f90:229-249) -------------
  const double o = a.omin + (double)(io - 1) * a.dw;
  const double theta = a.thline[line];
  const double q = a.qmode[line] ? kGamma * o * sin(theta) : a.qline[line];
  const double2 oi = make_double2(o, a.eta);
  const double2 oi2 = c_mul(oi, oi);
  const double2 beta = c_sqrt(make_double2(kGamma * kGamma * oi2.x - q * q, kGamma * kGamma * oi2.y));
  const double2 bb = c_mul(beta, beta);
  const double c0 = a.c0;
  const double2 e1 = c_sub(make_double2(1.0, 0.0), c_exp(make_double2(-beta.y * c0, beta.x * c0)));
  // block factors: Dxx = kx (A + d1), Dyy = ky (A + d1), Dyz = Dzy = kz (B + d2),
  // Dzz = kw (A + d1) - delta 4 pi / oi^2  (genD0, 652-660)
  const double2 two_pi_i = make_double2(0.0, 2.0 * M_PI);
  const double2 kx = c_div(c_scale(two_pi_i, kGamma * kGamma), beta);
  const double2 ky = c_div(c_mul(two_pi_i, beta), oi2);
  const double2 kz = c_div(c_scale(two_pi_i, -q), oi2);
  const double2 kw = c_div(c_scale(two_pi_i, q * q), c_mul(beta, oi2));
  const double2 fpi = c_div(make_double2(4.0 * M_PI, 0.0), oi2);

  // ---- per local field -----------------------------------------------------
  for (int i = tid; i < N; i += 256) {
    const double gz = a.gz[i], pr = (double)a.par[i];
    const double2 den = make_double2(bb.x - gz * gz, bb.y);
    s_gz[i] = gz;
    s_aii[i] = c_div(c_scale(e1, pr), den);                      // Aii (634-635)
    s_pd[i] = c_div(make_double2(pr, 0.0), den);                 // par_j / (beta^2 - Gz_j^2)
    s_d1[i] = c_div(c_mul(make_double2(0.0, 2.0), beta), den);   // diagonal1 (640)
    s_d2[i] = c_div(make_double2(0.0, 2.0 * gz), den);           // diagonal2 (641-642)
    // the slab integral vs(i) = int_0^{c0/2} e^{i Gz z} dz (c0/2 at the first
    // field), once per field: the sums below read it per element
    const double2 eg = c_expi(gz * c0 / 2.0);
    s_vs[i] = i == 0 ? make_double2(c0 / 2.0, 0.0) : make_double2(eg.y / gz, -(eg.x - 1.0) / gz);
    // form factors of genSpectra (1009-1013), real parts as the real Ff/Gf take them
    const double2 sn = c_scale(c_sin(c_scale(beta, c0 / 2.0)), 2.0 * pr / sqrt(c0));
    s_ff[i] = c_div(sn, make_double2(beta.x + gz, beta.y)).x;
    s_gf[i] = c_div(sn, make_double2(beta.x - gz, beta.y)).x;
  }
  __syncthreads();
  // u_c = P0_c aii, v_c = P0_c (Gz aii) for the five Pi0 blocks
  const double2 *P0 = a.P0 + (size_t)(io - 1) * 5 * N * N;
  for (int x = tid; x < 5 * N; x += 256) {
    const int c = x / N, i = x - c * N;
    const double2 *row = P0 + ((size_t)c * N + i) * N;
    double2 u = make_double2(0.0, 0.0), v = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k) {
      const double2 t = c_mul(row[k], s_aii[k]);
      u = c_add(u, t);
      v = c_add(v, c_scale(t, s_gz[k]));
    }
    s_u[c * N + i] = u;
    s_v[c * N + i] = v;
  }
  __syncthreads();
  const double tc = 2.0 / c0;
  const double2 tbc = c_scale(beta, 2.0 / c0);
  // sum_k P0_c(i,k) (A + d1)(k,j) and sum_k P0_c(i,k) (B + d2)(k,j)
  // (Aii carries the factor 1 - exp(i beta c0), so u and v do too)
  auto SA = [&](int c, int i, int j, double2 pij) {
    const double2 w = c_add(c_mul(bb, s_u[c * N + i]), c_scale(s_v[c * N + i], s_gz[j]));
    return c_add(c_mul(c_scale(s_pd[j], tc), w), c_mul(pij, s_d1[j]));
  };
  auto SB = [&](int c, int i, int j, double2 pij) {
    const double2 w = c_add(s_v[c * N + i], c_scale(s_u[c * N + i], s_gz[j]));
    return c_add(c_mul(c_mul(s_pd[j], tbc), w), c_mul(pij, s_d2[j]));
  };
  auto P = [&](int c, int i, int j) { return P0[((size_t)c * N + i) * N + j]; };
  // sumIntScreenedPi weight of (iG, jG) (Photonew.f90:846-879), times 1/c0 at
  // the end: it factors as w(i,j) = us(i) vs(j) with vs(j) = int_0^{c0/2}
  // e^{i Gz_j z} dz and us(i) = conj(vs(i)) (c0/2 at the first field), the
  // reference's four cases (0,0), (0,j), (i,0), (i,j) included
  auto vslab = [&](int j) { return s_vs[j]; };
  auto uslab = [&](int i) { return c_conj(s_vs[i]); };
  auto wsum = [&](int i, int j) { return c_mul(c_conj(s_vs[i]), s_vs[j]); };
  double acc[20];
#pragma unroll
  for (int e = 0; e < 20; ++e) acc[e] = 0.0;

  // ---- s mode: TS = I - Pixx0 Dxx0, Pixx = TS^-1 Pixx0 ----------------------
  for (int x = tid; x < N * N; x += 256) {
    const int i = x / N, j = x - i * N;
    const double2 t = c_mul(kx, SA(0, i, j, P(0, i, j)));
    sa[x] = make_double2((i == j ? 1.0 : 0.0) - t.x, -t.y);
  }
  __syncthreads();
  double2 *Wp = a.W ? a.W + (size_t)lp * 5 * N * N : nullptr;
  if (!Wp) {
    // solve form: every output is x^T T^-1 Pi0 y, so TS z = Pi0 y for y = e_1
    // (Pixx(1,1) = z_1) and y = vs (sPixx = us^T z / c0), vs and us the slab
    // integrals of the plane waves (the rank-one form of the sum's weights)
    for (int x = tid; x < 2 * N; x += 256) {
      const int r = x / N, i = x - r * N;
      double2 b = make_double2(0.0, 0.0);
      if (r == 0) {
        b = P(0, i, 0);
      } else {
        for (int j = 0; j < N; ++j) b = c_add(b, c_mul(P(0, i, j), vslab(j)));
      }
      ys[x] = b;
    }
    __syncthreads();
    gauss_solve<GLOBAL>(sa, N, ys, 2, fcol, red_v, red_i, a.status);
    if (tid == 0) s_11[0] = ys[0];
    for (int i = tid; i < N; i += 256) {
      const double2 t = c_mul(uslab(i), ys[N + i]);
      acc[0] += t.x;
      acc[1] += t.y;
    }
  } else {
    gj_invert<GLOBAL>(sa, N, fcol, piv, red_v, red_i, a.status);
    for (int x = tid; x < N * N; x += 256) {
      const int i = x / N, j = x - i * N;
      double2 s = make_double2(0.0, 0.0);
      for (int k = 0; k < N; ++k) s = c_add(s, c_mul(sa[i * N + k], P(0, k, j)));
      Wp[x] = s;
      if (x == 0) s_11[0] = s;
      const double2 ws = c_mul(s, wsum(i, j));
      acc[0] += ws.x;
      acc[1] += ws.y;
    }
  }
  __syncthreads();

  // ---- p mode: TP = I - [Pyy Pyz; Pzy Pzz] [Dyy Dyz; Dzy Dzz] ---------------
  for (int x = tid; x < M * M; x += 256) {
    const int I = x / M, J = x - I * M;
    const int i = I % N, j = J % N;
    double2 t;
    if (I < N && J < N) {  // Pyy Dyy + Pyz Dzy
      t = c_add(c_mul(ky, SA(1, i, j, P(1, i, j))), c_mul(kz, SB(2, i, j, P(2, i, j))));
    } else if (I < N) {  // Pyy Dyz + Pyz Dzz
      const double2 pyz = P(2, i, j);
      t = c_add(c_mul(kz, SB(1, i, j, P(1, i, j))), c_sub(c_mul(kw, SA(2, i, j, pyz)), c_mul(pyz, fpi)));
    } else if (J < N) {  // Pzy Dyy + Pzz Dzy
      t = c_add(c_mul(ky, SA(3, i, j, P(3, i, j))), c_mul(kz, SB(4, i, j, P(4, i, j))));
    } else {  // Pzy Dyz + Pzz Dzz
      const double2 pzz = P(4, i, j);
      t = c_add(c_mul(kz, SB(3, i, j, P(3, i, j))), c_sub(c_mul(kw, SA(4, i, j, pzz)), c_mul(pzz, fpi)));
    }
    sa[x] = make_double2((I == J ? 1.0 : 0.0) - t.x, -t.y);
  }
  __syncthreads();
  if (!Wp) {
    // solve form: TP z = B0 y with B0 = [Piyy0 Piyz0; Pizy0 Pizz0] restricted to
    // its y or z columns (blk) and y one of e_1, vs, Gf, Ff (vec): r = 2 vec + blk
    for (int x = tid; x < 8 * M; x += 256) {
      const int r = x / M, K = x - r * M;
      const int blk = r & 1, vec = r >> 1;
      const int c = (K < N ? 1 : 3) + blk, kk = K < N ? K : K - N;
      double2 b = make_double2(0.0, 0.0);
      if (vec == 0) {
        b = P(c, kk, 0);
      } else {
        for (int j = 0; j < N; ++j) {
          const double2 pij = P(c, kk, j);
          b = c_add(b, vec == 1 ? c_mul(pij, vslab(j)) : c_scale(pij, vec == 2 ? s_gf[j] : s_ff[j]));
        }
      }
      yp[x] = b;
    }
    __syncthreads();
    gauss_solve<GLOBAL>(sa, M, yp, 8, fcol, red_v, red_i, a.status);
    if (tid == 0) {
      s_11[1] = yp[0];      // Piyy(1,1)
      s_11[2] = yp[M + N];  // Pizz(1,1)
    }
    for (int i = tid; i < N; i += 256) {
      const double2 us = uslab(i);
      const double fi = s_ff[i];
      const double2 syy = c_mul(us, yp[2 * M + i]), szz = c_mul(us, yp[3 * M + N + i]);
      acc[2] += syy.x, acc[3] += syy.y;
      acc[8] += szz.x, acc[9] += szz.y;
      acc[10] += fi * yp[4 * M + i].x, acc[11] += fi * yp[4 * M + i].y;          // Dyy_r
      acc[12] += fi * yp[5 * M + N + i].x, acc[13] += fi * yp[5 * M + N + i].y;  // Dzz_r
      acc[14] += fi * yp[6 * M + i].x, acc[15] += fi * yp[6 * M + i].y;          // Dyy_a
      acc[16] += fi * yp[7 * M + N + i].x, acc[17] += fi * yp[7 * M + N + i].y;  // Dzz_a
      acc[18] += fi * yp[7 * M + i].x, acc[19] += fi * yp[7 * M + i].y;          // Dyz_a
    }
  } else {
    gj_invert<GLOBAL>(sa, M, fcol, piv, red_v, red_i, a.status);
    // [Piyy Piyz; Pizy Pizz] = TP^-1 [Piyy0 Piyz0; Pizy0 Pizz0]
    for (int x = tid; x < M * M; x += 256) {
      const int I = x / M, J = x - I * M;
      const int j = J % N, cj = J < N ? 0 : 1;  // column block: y or z
      double2 s = make_double2(0.0, 0.0);
      for (int k = 0; k < N; ++k) s = c_add(s, c_mul(sa[I * M + k], P(1 + cj, k, j)));       // rows of Piy.0
      for (int k = 0; k < N; ++k) s = c_add(s, c_mul(sa[I * M + N + k], P(3 + cj, k, j)));   // rows of Piz.0
      const int i = I % N;
      const int c = (I < N ? 1 : 3) + cj;  // 1 yy, 2 yz, 3 zy, 4 zz
      Wp[((size_t)c * N + i) * N + j] = s;
      if (i == 0 && j == 0 && (c == 1 || c == 4)) s_11[c == 1 ? 1 : 2] = s;
      const double2 ws = c_mul(s, wsum(i, j));
      // genSpectra sums (1016-1030): reflection Ff Pi Gf, absorption Ff Pi Ff
      const double fi = s_ff[i], gj = s_gf[j], fj = s_ff[j];
      if (c == 1) {
        acc[2] += ws.x, acc[3] += ws.y;
        acc[10] += fi * s.x * gj, acc[11] += fi * s.y * gj;    // Dyy_r
        acc[14] += fi * s.x * fj, acc[15] += fi * s.y * fj;    // Dyy_a
      } else if (c == 2) {
        acc[4] += ws.x, acc[5] += ws.y;
        acc[18] += fi * s.x * fj, acc[19] += fi * s.y * fj;    // Dyz_a
      } else if (c == 3) {
        acc[6] += ws.x, acc[7] += ws.y;
      } else {
        acc[8] += ws.x, acc[9] += ws.y;
        acc[12] += fi * s.x * gj, acc[13] += fi * s.y * gj;    // Dzz_r
        acc[16] += fi * s.x * fj, acc[17] += fi * s.y * fj;    // Dzz_a
      }
    }
  }
  __syncthreads();
  block_sum<20>(acc, red);
  if (tid == 0) {
    const double *r = red + 80;
    const double ic0 = 1.0 / c0;
    double2 *p11 = a.pi11 + (size_t)p * 3;
    double2 *ps = a.psum + (size_t)p * 3;
    p11[0] = s_11[0];
    p11[1] = s_11[1];
    p11[2] = s_11[2];
    ps[0] = make_double2(r[0] * ic0, r[1] * ic0);
    ps[1] = make_double2(r[2] * ic0, r[3] * ic0);
    ps[2] = make_double2(r[8] * ic0, r[9] * ic0);
    // p-mode reflection and absorption (1040-1059)
    const double2 dyy0 = c_div(make_double2(0.0, 2.0 * M_PI * kGamma), oi);
    const double ct = cos(theta), st = sin(theta);
    const double2 dyy_r = make_double2(r[10], r[11]), dzz_r = make_double2(r[12], r[13]);
    const double2 dyy_a = make_double2(r[14], r[15]), dzz_a = make_double2(r[16], r[17]);
    const double2 dyz_a = make_double2(r[18], r[19]);
    const double2 ref_p = c_mul(dyy0, c_sub(c_scale(dyy_r, ct), c_scale(dzz_r, st * st / ct)));
    const double2 abso_p =
        c_sub(c_add(c_scale(dyy_a, ct * ct), c_scale(dzz_a, st * st)), c_scale(dyz_a, 2.0 * ct * st));
    a.spec[2 * (size_t)p] = (4.0 * M_PI * kGamma / o) * abso_p.y;
    a.spec[2 * (size_t)p + 1] = c_mul(ref_p, c_conj(ref_p)).x;
  }
}

// One Pi0 component as the reference holds it, Pi0(No, Nlf, Nlf) column-major
// (frequency fastest), into the blocks P0[io][c][iG][jG] of io = 1..No-1:
// reads follow the frequencies, writes are scattered.
__global__ void k_photon_pack(int c, int no, int nf, int N, const double2 *__restrict__ raw,
                              double2 *__restrict__ P0) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long NN = (long long)N * N;
  if (t >= (long long)nf * NN) return;
  const int f = (int)(t % nf);
  const long long pr = t / nf;  // iG + N jG
  const int i = (int)(pr % N), j = (int)(pr / N);
  P0[(((size_t)f * 5 + c) * N + i) * N + j] = raw[(size_t)f + (size_t)no * pr];
}

namespace {

struct PhotonCall {
  hipStream_t s = nullptr;
  double2 *raw = nullptr;
  double2 *P0 = nullptr, *W = nullptr, *T = nullptr, *pi11 = nullptr, *psum = nullptr;
  double *gz = nullptr, *ql = nullptr, *tl = nullptr, *spec = nullptr;
  int *par = nullptr, *qm = nullptr, *status = nullptr;
  ~PhotonCall() {
    if (s) (void)hipStreamSynchronize(s);
    for (void *p : {(void *)raw, (void *)P0, (void *)W, (void *)T, (void *)pi11, (void *)psum, (void *)gz, (void *)ql, (void *)tl,
                    (void *)spec, (void *)par, (void *)qm, (void *)status})
      if (p) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
};

int perr(int code, const std::string &msg) {
  set_free_error("rpa_photon: " + msg);
  return code;
}
int phip(hipError_t e, const char *what) { return perr(RPA_E_HIP, std::string(what) + ": " + hipGetErrorString(e)); }

template <class T>
hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
}

}  // namespace

}  // namespace rpa

using namespace rpa;

extern "C" int rpa_photon(const rpa_photon_params *p, const double *glfz, const int *parg, const double *pixx0,
                          const double *piyy0, const double *pizz0, const double *piyz0, const double *pizy0,
                          const double *q_line, const double *theta_line, const int *qmode_line, double *pi11,
                          double *psum, double *spectra, double *pi_full) {
  if (!p || !glfz || !parg || !pixx0 || !piyy0 || !pizz0 || !q_line || !theta_line || !qmode_line || !pi11 ||
      !psum || !spectra)
    return perr(RPA_E_ARG, "null argument");
  const int no = p->no, N = p->nlf, nline = p->nline;
  if (no < 2) return perr(RPA_E_ARG, "No must be at least 2");
  if (N < 1) return perr(RPA_E_ARG, "Nlf must be positive");
  if (N > kPhMaxNlf) return perr(RPA_E_LIMIT, "Nlf above " + std::to_string(kPhMaxNlf));
  if (nline < 1) return perr(RPA_E_ARG, "no spectrum line");
  if (!(p->domega > 0.0) || !(p->c0 > 0.0) || !(p->eta >= 0.0))
    return perr(RPA_E_ARG, "domega and c0 must be positive, eta not negative");
  for (int l = 0; l < nline; ++l)
    if (qmode_line[l] != 0 && qmode_line[l] != 1) return perr(RPA_E_ARG, "qmode must be 0 or 1");
  const int nf = no - 1;
  const long long npts = (long long)nf * nline;
  if (npts > 0x7fffffffLL) return perr(RPA_E_LIMIT, "too many points");
  const size_t NN = (size_t)N * N, M = 2 * (size_t)N;
  const bool global = N > kPhLdsNlf;
  const int chunk = (int)std::min<long long>(npts, global ? 2048 : 8192);

  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return phip(e, "hipSetDevice");
  PhotonCall call;
  if ((e = hipStreamCreateWithFlags(&call.s, hipStreamNonBlocking)) != hipSuccess)
    return phip(e, "hipStreamCreateWithFlags");
  if ((e = dalloc(&call.P0, (size_t)nf * 5 * NN)) != hipSuccess || (e = dalloc(&call.raw, (size_t)no * NN)) != hipSuccess ||
      (e = dalloc(&call.gz, N)) != hipSuccess ||
      (e = dalloc(&call.par, N)) != hipSuccess || (e = dalloc(&call.ql, nline)) != hipSuccess ||
      (e = dalloc(&call.tl, nline)) != hipSuccess || (e = dalloc(&call.qm, nline)) != hipSuccess ||
      (e = dalloc(&call.pi11, (size_t)npts * 3)) != hipSuccess ||
      (e = dalloc(&call.psum, (size_t)npts * 3)) != hipSuccess ||
      (e = dalloc(&call.spec, (size_t)npts * 2)) != hipSuccess || (e = dalloc(&call.status, 1)) != hipSuccess ||
      (pi_full && (e = dalloc(&call.W, (size_t)chunk * 5 * NN)) != hipSuccess) ||
      (global && (e = dalloc(&call.T, (size_t)chunk * M * M)) != hipSuccess))
    return phip(e, "hipMalloc");
  struct {
    void *d;
    const void *s;
    size_t b;
  } up[] = {{call.gz, glfz, sizeof(double) * N},
            {call.par, parg, sizeof(int) * N},             {call.ql, q_line, sizeof(double) * nline},
            {call.tl, theta_line, sizeof(double) * nline}, {call.qm, qmode_line, sizeof(int) * nline}};
  for (auto &u : up)
    if ((e = hipMemcpyAsync(u.d, u.s, u.b, hipMemcpyHostToDevice, call.s)) != hipSuccess)
      return phip(e, "hipMemcpyAsync");
  // Pi0 blocks [io][5][N][N] from the reference's Pi0(No, Nlf, Nlf), one
  // component at a time through the staging buffer; Piyz0 and Pizy0 are zero
  // when not given (loadPi0, Photonew.f90:693-694)
  if ((e = hipMemsetAsync(call.P0, 0, sizeof(double2) * (size_t)nf * 5 * NN, call.s)) != hipSuccess)
    return phip(e, "hipMemsetAsync");
  const double *src[5] = {pixx0, piyy0, piyz0, pizy0, pizz0};
  for (int c = 0; c < 5; ++c) {
    if (!src[c]) continue;
    if ((e = hipMemcpyAsync(call.raw, src[c], sizeof(double2) * (size_t)no * NN, hipMemcpyHostToDevice, call.s)) !=
        hipSuccess)
      return phip(e, "hipMemcpyAsync");
    const long long n = (long long)nf * NN;
    hipLaunchKernelGGL(k_photon_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, call.s, c, no, nf, N,
                       (const double2 *)call.raw, call.P0);
    if ((e = hipGetLastError()) != hipSuccess) return phip(e, "k_photon_pack launch");
  }
  if ((e = hipMemsetAsync(call.status, 0, sizeof(int), call.s)) != hipSuccess) return phip(e, "hipMemsetAsync");

  PhotonArgs a{};
  a.N = N;
  a.nf = nf;
  a.omin = p->omin;
  a.dw = p->domega;
  a.eta = p->eta;
  a.c0 = p->c0;
  a.gz = call.gz;
  a.par = call.par;
  a.P0 = call.P0;
  a.qline = call.ql;
  a.thline = call.tl;
  a.qmode = call.qm;
  a.W = call.W;
  a.T = call.T;
  a.pi11 = call.pi11;
  a.psum = call.psum;
  a.spec = call.spec;
  a.status = call.status;
  const size_t lds = photon_lds_bytes(N, global);
  if ((e = hipFuncSetAttribute(global ? (const void *)k_photon<true> : (const void *)k_photon<false>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)) != hipSuccess)
    return phip(e, "hipFuncSetAttribute");
  std::vector<double> hw(pi_full ? (size_t)chunk * 5 * NN * 2 : 0);
  for (long long p0 = 0; p0 < npts; p0 += chunk) {
    const int n = (int)std::min<long long>(chunk, npts - p0);
    a.p0 = (int)p0;
    if (global)
      hipLaunchKernelGGL(k_photon<true>, dim3(n), dim3(256), lds, call.s, a);
    else
      hipLaunchKernelGGL(k_photon<false>, dim3(n), dim3(256), lds, call.s, a);
    if ((e = hipGetLastError()) != hipSuccess) return phip(e, "k_photon launch");
    if (pi_full) {
      // [point][5][N][N] row-major -> (Nlf, Nlf, 5, No-1, nline) column-major
      if ((e = hipMemcpyAsync(hw.data(), call.W, sizeof(double) * 2 * (size_t)n * 5 * NN, hipMemcpyDeviceToHost,
                              call.s)) != hipSuccess ||
          (e = hipStreamSynchronize(call.s)) != hipSuccess)
        return phip(e, "copy of the screened Pi");
      for (int l = 0; l < n; ++l)
        for (int c = 0; c < 5; ++c)
          for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
              const size_t s = 2 * ((((size_t)l * 5 + c) * N + i) * N + j);
              const size_t d = 2 * ((size_t)i + N * (j + (size_t)N * (c + 5 * (size_t)(p0 + l))));
              pi_full[d] = hw[s];
              pi_full[d + 1] = hw[s + 1];
            }
    }
  }
  int status = 0;
  if ((e = hipMemcpyAsync(pi11, call.pi11, sizeof(double2) * npts * 3, hipMemcpyDeviceToHost, call.s)) != hipSuccess ||
      (e = hipMemcpyAsync(psum, call.psum, sizeof(double2) * npts * 3, hipMemcpyDeviceToHost, call.s)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(spectra, call.spec, sizeof(double) * npts * 2, hipMemcpyDeviceToHost, call.s)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(&status, call.status, sizeof(int), hipMemcpyDeviceToHost, call.s)) != hipSuccess ||
      (e = hipStreamSynchronize(call.s)) != hipSuccess)
    return phip(e, "copy of the results");
  if (status) return perr(RPA_E_SINGULAR, "singular TS or TP matrix at some point (results there are not finite)");
  return RPA_OK;
}
