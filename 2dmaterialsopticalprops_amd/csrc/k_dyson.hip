// Coulomb matrix, Dyson step, screened interaction and scalar epilogues.
//
//  k_coulomb   V(G,G') of the slab (Sloss.f90:1173-1208 genV); V is
//              independent of frequency, so it is built once per q instead of
//              once per frequency as the reference does (Sloss.f90:456).
//  k_dyson     one workgroup per frequency io = 1..no-1 (k_dyson_col for
//              N_G <= 64, the general form k_dyson up to the LDS limit):
//                eps = I - chi0 V            (genDielectricEpsilon, 1210-1234)
//                eps <- eps^-1               (gjel, Sloss.f90:462; here
//                                             Gauss-Jordan with row pivoting,
//                                             in place in LDS)
//                chi = eps^-1 chi0           (genChi, 1236-1253)
//                W   = V + V chi V            (genWT, 1255-1277) evaluated
//                                             as V eps^-1 (chi V = eps^-1 - I),
//                                             one O(N^3) product
//                loss = -Im eps^-1(1,1)
//  k_pi        Optabs Pi assembly (Optabs.f90:575-611): Pi_tot(1,1) and the
//              conductivities of fort.401 / fort.402.
#include <cstdlib>
#include <string>

#include "rpa_internal.h"

namespace rpa {

__global__ void k_coulomb(int nlf, const double *glf, const int *parg, double qx, double qy,
                          double gcar, double c0, double eps, double *V) {
#pragma clang fp contract(off)
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nlf * nlf) return;
  const int iG = idx / nlf, jG = idx - iG * nlf;
  const double *gi = glf + 3 * iG, *gj = glf + 3 * jG;
  const double a = qx + gcar * gi[0], b = qy + gcar * gi[1];
  double gabs = sqrt(a * a + b * b);
  if (gabs == 0.0) gabs = eps;
  double v = 0.0;
  if (gj[0] == gi[0] && gj[1] == gi[1]) {
    const double gzi = gcar * gi[2], gzj = gcar * gj[2];
    v = 4.0 * M_PI * (1.0 - exp(-gabs * c0)) / (gabs * c0);
    v = v * (gabs * gabs - gzi * gzj);
    v = v / (gabs * gabs + gzi * gzi);
    v = v / (gabs * gabs + gzj * gzj);
    v = -(double)parg[iG] * (double)parg[jG] * v;
    if (gj[2] == gi[2]) v = 4.0 * M_PI / (gabs * gabs + gzi * gzi) + v;
  }
  V[idx] = v;
}

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cdiv1(double2 a) {  // 1 / a (scaled)
  double s = fabs(a.x) > fabs(a.y) ? fabs(a.x) : fabs(a.y);
  double ar = a.x / s, ai = a.y / s;
  double d = ar * ar + ai * ai;
  return make_double2(ar / (d * s), -ai / (d * s));
}

struct DysonArgs {
  int nlf, nfreq;
  const double2 *chi0;  // [nfreq][nlf][nlf]
  const double *V;      // [nlf][nlf]
  double2 *epsinv, *chi, *W;
  double *loss;
  int *status;
};

// General form (N_G > 64): the working matrix eps -> eps^-1 lives in LDS when
// its 16 N_G^2 bytes fit (N_G <= 96), else in place in the eps^-1 output of
// the frequency (GLOBAL: every access goes through the L2; the barriers order
// the workgroup's global accesses as they order its LDS ones).  chi0 and V
// are read from the L2.
template <bool GLOBAL>
__global__ __launch_bounds__(256) void k_dyson(DysonArgs a) {
  extern __shared__ double2 lds_eps[];  // [nlf][nlf] eps, then eps^-1 (LDS form)
  __shared__ double2 fcol[kMaxNlf];
  __shared__ int piv[kMaxNlf];
  __shared__ double red_v[4];
  __shared__ int red_i[4];
  const int N = a.nlf, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double *V = a.V;
  const int f = blockIdx.x;
  const double2 *X0 = a.chi0 + (size_t)f * N * N;
  double2 *Einv = a.epsinv + (size_t)f * N * N;
  double2 *sa = GLOBAL ? Einv : lds_eps;
  // eps = I - chi0 V
  for (int x = tid; x < N * N; x += 256) {
    int i = x / N, j = x - i * N;
    double2 s = make_double2(i == j ? 1.0 : 0.0, 0.0);
    for (int k = 0; k < N; ++k) {
      double2 c = X0[i * N + k];
      double v = V[k * N + j];
      s.x -= c.x * v;
      s.y -= c.y * v;
    }
    sa[x] = s;
  }
  __syncthreads();
  // in-place Gauss-Jordan inversion with row pivoting
  for (int k = 0; k < N; ++k) {
    // pivot: largest |Re| + |Im| in column k at or below row k, lowest row on ties
    double best = -1.0;
    int bi = N;
    for (int i = k + tid; i < N; i += 256) {
      double2 z = sa[i * N + k];
      double m = fabs(z.x) + fabs(z.y);
      if (m > best) {
        best = m;
        bi = i;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off);
      const int oi = __shfl_xor(bi, off);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if (lane == 0) {
      red_v[wave] = best;
      red_i[wave] = bi;
    }
    __syncthreads();
    double pv = red_v[0];
    int p = red_i[0];
    for (int w = 1; w < 4; ++w)
      if (red_v[w] > pv || (red_v[w] == pv && red_i[w] < p)) {
        pv = red_v[w];
        p = red_i[w];
      }
    if (!(pv > 0.0)) {  // singular: flag it and go on without a swap
      if (tid == 0) atomicOr(a.status, 1);
      p = k;
    }
    if (tid == 0) piv[k] = p;
    if (p != k)
      for (int j = tid; j < N; j += 256) {
        double2 t = sa[k * N + j];
        sa[k * N + j] = sa[p * N + j];
        sa[p * N + j] = t;
      }
    __syncthreads();
    const double2 inv = cdiv1(sa[k * N + k]);
    __syncthreads();
    for (int j = tid; j < N; j += 256) sa[k * N + j] = (j == k) ? inv : cmul(sa[k * N + j], inv);
    for (int i = tid; i < N; i += 256) fcol[i] = sa[i * N + k];
    __syncthreads();
    for (int i = tid; i < N; i += 256)
      if (i != k) sa[i * N + k] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int x = tid; x < N * N; x += 256) {
      int i = x / N, j = x - i * N;
      if (i != k) {
        double2 fk = fcol[i];
        double2 r = sa[k * N + j];
        double2 v = sa[x];
        v.x -= fk.x * r.x - fk.y * r.y;
        v.y -= fk.x * r.y + fk.y * r.x;
        sa[x] = v;
      }
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, last first
  for (int k = N - 1; k >= 0; --k) {
    const int p = piv[k];
    if (p != k)
      for (int i = tid; i < N; i += 256) {
        double2 t = sa[i * N + k];
        sa[i * N + k] = sa[i * N + p];
        sa[i * N + p] = t;
      }
    __syncthreads();
  }
  double2 *Ch = a.chi + (size_t)f * N * N;
  double2 *Wo = a.W + (size_t)f * N * N;
  if (tid == 0) a.loss[f] = -sa[0].y;
  for (int x = tid; x < N * N; x += 256) {
    int i = x / N, j = x - i * N;
    if (!GLOBAL) Einv[x] = sa[x];
    // chi = eps^-1 chi0
    double2 s = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k) {
      double2 e = sa[i * N + k], c = X0[k * N + j];
      s.x += e.x * c.x - e.y * c.y;
      s.y += e.x * c.y + e.y * c.x;
    }
    Ch[x] = s;
    // W = V + V (chi V) = V eps^-1, since chi V = eps^-1 chi0 V = eps^-1 (I - eps)
    double2 w = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k) {
      double v = V[i * N + k];
      double2 e = sa[k * N + j];
      w.x += v * e.x;
      w.y += v * e.y;
    }
    Wo[x] = w;
  }
}

// The same solve for N_G <= 64 with lane = column j and rows i = wave + 4u:
// no index divisions, and each O(N^3) product reads one operand per k and
// takes the other as a broadcast.  Only eps (then eps^-1) lives in LDS
// (16 N_G^2 bytes), so three frequencies share a CU; chi0 and V are read
// from L2, the broadcast operand through the scalar cache (row addresses are
// wave-uniform).  Rows past N_G read row N_G-1 and their results are dropped,
// so the row loops have no branch and their loads are issued together.
// Gauss-Jordan step k: wave 0 finds the pivot, reads the pivot row, row p and
// column k and writes the scaled pivot row, the swapped row p and the
// multipliers; after one barrier all waves eliminate.
constexpr int kDyRows = 16;  // rows per wave (N_G <= 64)

__global__ __launch_bounds__(256, 3) void k_dyson_col(DysonArgs a) {
  extern __shared__ double2 sa[];  // eps -> eps^-1 [N][N]
  __shared__ double2 fcol[64];
  __shared__ int piv[64];
  const int N = a.nlf, tid = threadIdx.x, j = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int f = blockIdx.x;
  const double2 *X0 = a.chi0 + (size_t)f * N * N;
  const double *V = a.V;
  const bool jok = j < N;
  const int jc = jok ? j : N - 1;  // lanes past N work on a valid column, results dropped
  auto ic = [&](int u) { return min(wave + 4 * u, N - 1); };

  // eps = I - chi0 V.  chi0 of this frequency is first copied to LDS in
  // bursts of four loads per thread, then read as wave-uniform broadcasts;
  // V is read four k ahead.  (Read row by row from L2 inside the k loop, the
  // loads of each k waited a full latency: 70 of the kernel's 290 us.)
  {
    const int nn = N * N;
    for (int x0 = 0; x0 < nn; x0 += 1024) {
      double2 t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) t[u] = X0[min(x0 + tid + 256 * u, nn - 1)];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (x0 + tid + 256 * u < nn) sa[x0 + tid + 256 * u] = t[u];
    }
  }
  __syncthreads();
  // The three O(N^3) products run on the FP64 matrix cores: wave w takes the
  // 16 x 16 tiles w, w + 4, ... of the (N padded to 16) result, K in steps of
  // 4, each operand one value per lane (A[row lane & 15][k lane >> 4],
  // B[k lane >> 4][column lane & 15]), read from clamped addresses and zeroed
  // past N; D[4 e + (lane >> 4)][lane & 15] in register e.  (As one FMA per
  // LDS broadcast on the vector units the products took 92 of the kernel's
  // 204 us per frequency at N_G = 51.)
  const int nT = (N + 15) >> 4, ntile = nT * nT, nk4 = (N + 3) >> 2;
  const int lr = j & 15, lk = j >> 4;
  // eps = I - chi0 V: A = Re, Im chi0 (LDS), B = V (L2)
  rpa_d4 er[4], ei[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    er[t] = rpa_d4{0.0, 0.0, 0.0, 0.0};
    ei[t] = rpa_d4{0.0, 0.0, 0.0, 0.0};
    const int tile = wave + 4 * t;
    if (tile < ntile) {  // uniform
      const int row = 16 * (tile / nT) + lr, col = 16 * (tile % nT) + lr;
      const int rc = min(row, N - 1), cc = min(col, N - 1);
#pragma unroll 4
      for (int kk = 0; kk < nk4; ++kk) {
        const int k = 4 * kk + lk, kc = min(k, N - 1);
        const double2 x = sa[rc * N + kc];
        const double v = V[kc * N + cc];
        const bool ok = row < N && k < N;  // B past N meets only dropped columns
        const double ax = ok ? x.x : 0.0, ay = ok ? x.y : 0.0;
        er[t] = mfma_f64_16x16x4(ax, v, er[t]);
        ei[t] = mfma_f64_16x16x4(ay, v, ei[t]);
      }
    }
  }
  __syncthreads();  // every wave has read chi0
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int tile = wave + 4 * t;
    if (tile < ntile) {
      const int col = 16 * (tile % nT) + lr;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = 16 * (tile / nT) + 4 * e + lk;
        if (i < N && col < N) sa[i * N + col] = make_double2((i == col ? 1.0 : 0.0) - er[t][e], -ei[t][e]);
      }
    }
  }
  __syncthreads();

  // in-place Gauss-Jordan inversion with row pivoting.  Wave 0 alone finds
  // the pivot, reads the pivot row, row p and column k, and writes the
  // scaled pivot row, the swapped row, the multipliers and the zeroed column
  // k: its LDS operations run in program order, so no barrier separates
  // those reads from those writes.  All waves then eliminate.
  for (int k = 0; k < N; ++k) {
    if (wave == 0) {
      // pivot: largest |Re| + |Im| in column k at or below row k, lowest row on ties
      // (the maximum by a butterfly, then the lowest lane holding it by a ballot:
      // lane j holds row j)
      const double2 z = sa[jc * N + k];  // column k of row j before the swap
      const double mine = (j >= k && jok) ? fabs(z.x) + fabs(z.y) : -1.0;
      double best = mine;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) best = fmax(best, __shfl_xor(best, off));
      int p = __ffsll((unsigned long long)__ballot(mine == best)) - 1;
      if (!(best > 0.0)) {  // singular: flag it and go on without a swap
        if (j == 0) atomicOr(a.status, 1);
        p = k;
      }
      if (j == 0) piv[k] = p;
      const double2 rk = sa[k * N + jc];
      const double2 rp = sa[p * N + jc];
      const double2 inv = cdiv1(sa[p * N + k]);  // pivot value: column k of the new row k
      const double2 okk = sa[k * N + k];         // column k of the old row k (the new row p)
      if (jok) {
        sa[k * N + j] = (j == k) ? inv : cmul(rp, inv);
        if (p != k && j != k) sa[p * N + j] = rk;
        if (j != k) {
          // per component: a select of whole double2 values goes through scratch memory
          const bool sw = j == p;
          fcol[j] = make_double2(sw ? okk.x : z.x, sw ? okk.y : z.y);
          sa[j * N + k] = make_double2(0.0, 0.0);
        }
      }
    }
    __syncthreads();
    {
      const double2 r = sa[k * N + jc];
      // two batches of 8 rows keep the live registers low
#pragma unroll
      for (int h = 0; h < kDyRows; h += 8) {
        double2 fk[8], v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          fk[u] = fcol[ic(h + u)];
          v[u] = sa[ic(h + u) * N + jc];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          v[u].x -= fk[u].x * r.x - fk[u].y * r.y;
          v[u].y -= fk[u].x * r.y + fk[u].y * r.x;
        }
        if (jok) {
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int i = wave + 4 * (h + u);
            if (i < N && i != k) sa[i * N + j] = v[u];
          }
        }
      }
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, last first: column j of
  // the result is column c(j) of the eliminated matrix, c = t_(N-1) .. t_0 (j)
  int cj = jc;
  for (int k = 0; k < N; ++k) {
    const int pk = piv[k];
    cj = (cj == k) ? pk : (cj == pk) ? k : cj;
  }
  // (the values move as native vectors: an array of HIP's double2 struct,
  // only copied, went to scratch memory)
  typedef double dv2 __attribute__((ext_vector_type(2)));
  dv2 pv[kDyRows];
#pragma unroll
  for (int u = 0; u < kDyRows; ++u) pv[u] = *(const dv2 *)&sa[ic(u) * N + cj];
  __syncthreads();
  double2 *Einv = a.epsinv + (size_t)f * N * N;
  double2 *Ch = a.chi + (size_t)f * N * N;
  double2 *Wo = a.W + (size_t)f * N * N;
  if (jok) {
#pragma unroll
    for (int u = 0; u < kDyRows; ++u) {
      const int i = wave + 4 * u;
      if (i < N) {
        *(dv2 *)&sa[i * N + j] = pv[u];
        *(dv2 *)&Einv[i * N + j] = pv[u];
      }
    }
  }
  __syncthreads();
  if (tid == 0) a.loss[f] = -sa[0].y;
  // chi = eps^-1 chi0: A = eps^-1 (LDS), B = chi0 (L2); then W = V eps^-1
  // (= V + V chi V): A = V (L2), B = eps^-1 (LDS)
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int tile = wave + 4 * t;
    if (tile < ntile) {  // uniform
      const int row = 16 * (tile / nT) + lr, col = 16 * (tile % nT) + lr;
      const int rc = min(row, N - 1), cc = min(col, N - 1);
      rpa_d4 cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
      rpa_d4 wr = {0.0, 0.0, 0.0, 0.0}, wi = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
      for (int kk = 0; kk < nk4; ++kk) {
        const int k = 4 * kk + lk, kc = min(k, N - 1);
        const bool oka = row < N && k < N, okb = k < N && col < N;
        const double2 e = sa[rc * N + kc];
        const double2 c = X0[kc * N + cc];
        const double v = V[rc * N + kc];
        const double2 eb = sa[kc * N + cc];
        const double ex = oka ? e.x : 0.0, ey = oka ? e.y : 0.0;
        const double cx = okb ? c.x : 0.0, cy = okb ? c.y : 0.0;
        const double va = oka ? v : 0.0;
        const double bx = okb ? eb.x : 0.0, by = okb ? eb.y : 0.0;
        cr = mfma_f64_16x16x4(ex, cx, cr);
        cr = mfma_f64_16x16x4(-ey, cy, cr);
        ci = mfma_f64_16x16x4(ex, cy, ci);
        ci = mfma_f64_16x16x4(ey, cx, ci);
        wr = mfma_f64_16x16x4(va, bx, wr);
        wi = mfma_f64_16x16x4(va, by, wi);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = 16 * (tile / nT) + 4 * q + lk;
        if (i < N && col < N) {
          Ch[i * N + col] = make_double2(cr[q], ci[q]);
          Wo[i * N + col] = make_double2(wr[q], wi[q]);
        }
      }
    }
  }
}

// Mixed-precision Dyson step (N_G <= 64): eps = I - chi0 V in FP64, its
// inverse X by Gauss-Jordan with row pivoting on the FP32 copy of eps, then
// one FP64 refinement step
//     R = I - eps X   (FP64 residual; eps staged from the chi output buffer)
//     X1 = X + X R    (the correction X R in FP32: R is of the order of the
//                      FP32 rounding, so its FP32 product errs by ~1e-14)
// and, when max |R| exceeds kRefineAgain, a second one on X1 (the error
// after a step is about |R|^2, so 1e-8 at the threshold).  chi = X1 chi0, W = V X1 and the loss follow as
// in k_dyson_col.  Layout as k_dyson_col: lane = column j, rows wave + 4u.
// LDS: chi0 of the frequency (FP64) while eps is formed, then X and R
// (FP32), then X1 (FP64); eps itself waits in the chi output of the
// frequency, which chi overwrites last.
constexpr double kRefineAgain = 1.0e-4;
constexpr int kMxKB = 8;  // eps columns staged per pass of the residual product

__global__ __launch_bounds__(256, 3) void k_dyson_mx(DysonArgs a) {
  extern __shared__ double2 sa[];  // chi0 [N][N] FP64 | X [N][N], R [N][N] FP32 | X1 [N][N] FP64
  __shared__ float2 fcol[64];
  __shared__ int piv[64];
  __shared__ double2 seps[64][kMxKB];  // eps(:, k0 .. k0+KB-1) of the residual product
  __shared__ float redmax[4];
  float2 *sx = (float2 *)sa;
  const int N = a.nlf, tid = threadIdx.x, j = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int f = blockIdx.x;
  const double2 *X0 = a.chi0 + (size_t)f * N * N;
  const double *V = a.V;
  double2 *Ch = a.chi + (size_t)f * N * N;  // eps until chi is written
  const bool jok = j < N;
  const int jc = jok ? j : N - 1;
  double2 acc[kDyRows];
  auto ic = [&](int u) { return min(wave + 4 * u, N - 1); };

  // eps = I - chi0 V (chi0 staged to LDS in bursts, V four k ahead), as k_dyson_col
  {
    const int nn = N * N;
    for (int x0 = 0; x0 < nn; x0 += 1024) {
      double2 t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) t[u] = X0[min(x0 + tid + 256 * u, nn - 1)];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (x0 + tid + 256 * u < nn) sa[x0 + tid + 256 * u] = t[u];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kDyRows; ++u) acc[u] = make_double2(0.0, 0.0);
  for (int k0 = 0; k0 < N; k0 += 4) {
    double v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = V[min(k0 + q, N - 1) * N + jc];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (k0 + q < N) {  // uniform
#pragma unroll
        for (int u = 0; u < kDyRows; ++u) {
          const double2 c = sa[ic(u) * N + k0 + q];
          acc[u].x += c.x * v[q];
          acc[u].y += c.y * v[q];
        }
      }
    }
  }
  __syncthreads();  // every wave has read chi0
  if (jok) {
#pragma unroll
    for (int u = 0; u < kDyRows; ++u) {
      const int i = wave + 4 * u;
      if (i < N) {
        const double2 e = make_double2((i == j ? 1.0 : 0.0) - acc[u].x, -acc[u].y);
        Ch[i * N + j] = e;
        sx[i * N + j] = make_float2((float)e.x, (float)e.y);
      }
    }
  }
  __syncthreads();

  // FP32 Gauss-Jordan inversion with row pivoting, as in k_dyson_col
  for (int k = 0; k < N; ++k) {
    if (wave == 0) {
      const float2 z = sx[jc * N + k];
      const float mine = (j >= k && jok) ? fabsf(z.x) + fabsf(z.y) : -1.0f;
      float best = mine;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) best = fmaxf(best, __shfl_xor(best, off));
      int p = __ffsll((unsigned long long)__ballot(mine == best)) - 1;
      if (!(best > 0.0f)) {  // singular: flag it and go on without a swap
        if (j == 0) atomicOr(a.status, 1);
        p = k;
      }
      if (j == 0) piv[k] = p;
      const float2 rk = sx[k * N + jc];
      const float2 rp = sx[p * N + jc];
      const float2 pz = sx[p * N + k];
      const float2 okk = sx[k * N + k];
      // 1 / pz, scaled
      const float sc = fmaxf(fabsf(pz.x), fabsf(pz.y));
      const float pr = pz.x / sc, pi = pz.y / sc, pd = pr * pr + pi * pi;
      const float2 inv = make_float2(pr / (pd * sc), -pi / (pd * sc));
      if (jok) {
        sx[k * N + j] = (j == k) ? inv : make_float2(rp.x * inv.x - rp.y * inv.y, rp.x * inv.y + rp.y * inv.x);
        if (p != k && j != k) sx[p * N + j] = rk;
        if (j != k) {
          const bool sw = j == p;
          fcol[j] = make_float2(sw ? okk.x : z.x, sw ? okk.y : z.y);
          sx[j * N + k] = make_float2(0.0f, 0.0f);
        }
      }
    }
    __syncthreads();
    {
      const float2 r = sx[k * N + jc];
#pragma unroll
      for (int h = 0; h < kDyRows; h += 8) {
        float2 fk[8], v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          fk[u] = fcol[ic(h + u)];
          v[u] = sx[ic(h + u) * N + jc];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          v[u].x -= fk[u].x * r.x - fk[u].y * r.y;
          v[u].y -= fk[u].x * r.y + fk[u].y * r.x;
        }
        if (jok) {
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int i = wave + 4 * (h + u);
            if (i < N && i != k) sx[i * N + j] = v[u];
          }
        }
      }
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges
  int cj = jc;
  for (int k = 0; k < N; ++k) {
    const int pk = piv[k];
    cj = (cj == k) ? pk : (cj == pk) ? k : cj;
  }
  // X (FP64) waits in the eps^-1 output of the frequency, each lane holding
  // its own entries there (rows wave + 4u, column j), between the phases
  double2 *Einv = a.epsinv + (size_t)f * N * N;
  {
    float2 t[kDyRows];
#pragma unroll
    for (int u = 0; u < kDyRows; ++u) t[u] = sx[ic(u) * N + cj];
#pragma unroll
    for (int u = 0; u < kDyRows; ++u) {
      const int i = wave + 4 * u;
      if (jok && i < N) Einv[i * N + j] = make_double2((double)t[u].x, (double)t[u].y);
    }
  }
  __syncthreads();  // every wave has read the Gauss-Jordan result

  // refinement steps: R = I - eps X with X read from LDS in FP64, then
  // X += X R with X and R in FP32 (LDS).  A second step runs when
  // max |R| > kRefineAgain (uniform across the workgroup).
  double2 *sxd = sa;            // X in FP64 [N][N] during the residual
  float2 *sr = sx + N * N;      // R in FP32 [N][N] during the correction
  for (int step = 0; step < 2; ++step) {
#pragma unroll
    for (int u = 0; u < kDyRows; ++u) {
      const int i = wave + 4 * u;
      if (jok && i < N) sxd[i * N + j] = Einv[i * N + j];
    }
#pragma unroll
    for (int u = 0; u < kDyRows; ++u) acc[u] = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < N; k0 += kMxKB) {
      // eps(:, k0 .. k0+KB-1) from the chi buffer into LDS
      for (int x = tid; x < N * kMxKB; x += 256) {
        const int i = x / kMxKB, kk = x - i * kMxKB;
        seps[i][kk] = (k0 + kk < N) ? Ch[i * N + k0 + kk] : make_double2(0.0, 0.0);
      }
      __syncthreads();  // (first pass: also X in LDS)
#pragma unroll 2
      for (int kk = 0; kk < kMxKB; ++kk) {
        if (k0 + kk < N) {  // uniform
          const double2 xv = sxd[(k0 + kk) * N + jc];
#pragma unroll
          for (int u = 0; u < kDyRows; ++u) {
            const double2 e = seps[ic(u)][kk];
            acc[u].x += e.x * xv.x - e.y * xv.y;
            acc[u].y += e.x * xv.y + e.y * xv.x;
          }
        }
      }
      __syncthreads();  // seps and X read before the next pass writes them
    }
    // R = I - eps X and X, both in FP32, to LDS; the largest |R| component
    float rmax = 0.0f;
#pragma unroll
    for (int u = 0; u < kDyRows; ++u) {
      const int i = wave + 4 * u;
      if (jok && i < N) {
        const double rr = (i == j ? 1.0 : 0.0) - acc[u].x, ri = -acc[u].y;
        rmax = fmaxf(rmax, fmaxf(fabsf((float)rr), fabsf((float)ri)));
        sr[i * N + j] = make_float2((float)rr, (float)ri);
        const double2 xo = Einv[i * N + j];
        sx[i * N + j] = make_float2((float)xo.x, (float)xo.y);
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) rmax = fmaxf(rmax, __shfl_xor(rmax, off));
    if (j == 0) redmax[wave] = rmax;
    __syncthreads();
    // X += X R (FP32 products, added in FP64), in two halves of the rows
#pragma unroll
    for (int h = 0; h < kDyRows; h += 8) {
      float2 cr[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) cr[u] = make_float2(0.0f, 0.0f);
      for (int k = 0; k < N; ++k) {
        const float2 rv = sr[k * N + jc];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const float2 xv = sx[ic(h + u) * N + k];
          cr[u].x += xv.x * rv.x - xv.y * rv.y;
          cr[u].y += xv.x * rv.y + xv.y * rv.x;
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = wave + 4 * (h + u);
        if (jok && i < N) {
          double2 xo = Einv[i * N + j];
          xo.x += (double)cr[u].x;
          xo.y += (double)cr[u].y;
          Einv[i * N + j] = xo;
        }
      }
    }
    const float rall = fmaxf(fmaxf(redmax[0], redmax[1]), fmaxf(redmax[2], redmax[3]));
    __syncthreads();  // X, R and redmax read by every wave
    if (!(rall > (float)kRefineAgain)) break;  // uniform
  }

  // X1 (FP64) to LDS for chi and W
#pragma unroll
  for (int u = 0; u < kDyRows; ++u) {
    const int i = wave + 4 * u;
    if (jok && i < N) sa[i * N + j] = Einv[i * N + j];
  }
  __syncthreads();
  if (tid == 0) a.loss[f] = -sa[0].y;
  // chi = X1 chi0 (chi0 rows read four k ahead)
#pragma unroll
  for (int u = 0; u < kDyRows; ++u) acc[u] = make_double2(0.0, 0.0);
  for (int k0 = 0; k0 < N; k0 += 4) {
    double2 c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) c[q] = X0[min(k0 + q, N - 1) * N + jc];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (k0 + q < N) {  // uniform
#pragma unroll
        for (int u = 0; u < kDyRows; ++u) {
          const double2 e = sa[ic(u) * N + k0 + q];
          acc[u].x += e.x * c[q].x - e.y * c[q].y;
          acc[u].y += e.x * c[q].y + e.y * c[q].x;
        }
      }
    }
  }
  __syncthreads();  // eps in the chi buffer is read by nobody any more
  if (jok) {
#pragma unroll
    for (int u = 0; u < kDyRows; ++u) {
      const int i = wave + 4 * u;
      if (i < N) Ch[i * N + j] = acc[u];
    }
  }
  // W = V X1
  double2 *Wo = a.W + (size_t)f * N * N;
#pragma unroll
  for (int u = 0; u < kDyRows; ++u) acc[u] = make_double2(0.0, 0.0);
  for (int k0 = 0; k0 < N; k0 += 2) {
    double vv[2][kDyRows];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int u = 0; u < kDyRows; ++u) vv[q][u] = V[ic(u) * N + min(k0 + q, N - 1)];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (k0 + q < N) {  // uniform
        const double2 e = sa[(k0 + q) * N + jc];
#pragma unroll
        for (int u = 0; u < kDyRows; ++u) {
          acc[u].x += vv[q][u] * e.x;
          acc[u].y += vv[q][u] * e.y;
        }
      }
    }
  }
  if (jok) {
#pragma unroll
    for (int u = 0; u < kDyRows; ++u) {
      const int i = wave + 4 * u;
      if (i < N) Wo[i * N + j] = acc[u];
    }
  }
}

// Optabs Pi assembly per io (Optabs.f90:575-611).  Pi_inter is read after each
// (iG,jG) update in the reference, so for Nlf > 1 it holds Pi_tot(1,1)
// including the intraband term (Optabs.f90:594).
__global__ void k_pi(int nfreq, int nlf, double dw, double eta_intra, double c0,
                     const double2 *pi0, const double *qeff, double *pi11, double *sig) {
  const int iop = blockIdx.x * blockDim.x + threadIdx.x;
  if (iop >= nfreq) return;
  const int io = iop + 1;
  const double oi = (double)(io - 1) * dw;
  const double2 p = pi0[(size_t)iop * nlf * nlf];
  const double dia = -pi0[0].x;
  const double tr = p.x + dia, ti = p.y + 0.0;
  // Qeff(1,1) * oi / (oi + i eta_intra)
  const double qr = qeff[0] * oi, qi = qeff[1] * oi;
  const double dr = oi, di = eta_intra;
  const double den = dr * dr + di * di;
  const double ir = (qr * dr + qi * di) / den, ii = (qi * dr - qr * di) / den;
  const double totr = tr + ir, toti = ti + ii;
  pi11[2 * iop] = totr;
  pi11[2 * iop + 1] = toti;
  double inter_r = (nlf > 1) ? totr : tr, inter_i = (nlf > 1) ? toti : ti;
  if (io > 1) {
    // Re(-i c0 P / oi) = c0 Im(P) / oi
    sig[2 * iop] = c0 * inter_i / oi;
    sig[2 * iop + 1] = c0 * ii / oi;
  } else {
    sig[2 * iop] = 0.0;
    sig[2 * iop + 1] = 0.0;
  }
  (void)inter_r;
}

int launch_coulomb(rpa_ctx *c) {
  const int nn = c->nlf * c->nlf;
  hipLaunchKernelGGL(k_coulomb, dim3((nn + 255) / 256), dim3(256), 0, c->stream, c->nlf,
                     (const double *)c->d_glf, (const int *)c->d_parg, c->p.gcar * c->q[0],
                     c->p.gcar * c->q[1], c->p.gcar, c->p.c0, c->p.eps, c->d_V);
  RPA_HIP(c, hipGetLastError());
  return 0;
}

// Which Dyson kernel serves the current N_G:
//   N_G <= 64   k_dyson_col (FP64 Gauss-Jordan), or k_dyson_mx (FP32
//               Gauss-Jordan + FP64 refinement) on the mixed path with
//               RPA_DYSON=mixed
//   N_G <= 100  k_dyson<false>: eps in LDS (16 N_G^2 bytes)
//   N_G <= 128  k_dyson<true>:  eps in place in the eps^-1 output (L2)
enum DysonForm { kDyCol = 0, kDyMx = 1, kDyLds = 2, kDyGlobal = 3 };
constexpr size_t kLdsMax = 160 * 1024;
// static LDS next to the dynamic part: k_dyson (fcol, piv, reductions),
// k_dyson_mx (fcol, piv, staged eps columns, reductions)
constexpr size_t kDysonStatic = (sizeof(double2) + sizeof(int)) * kMaxNlf + 64;
constexpr size_t kDysonMxStatic = (sizeof(float2) + sizeof(int)) * 64 + sizeof(double2) * 64 * kMxKB + 64;

// The mixed path keeps the FP64 Gauss-Jordan by default: at N_G = 51 the
// FP32 Gauss-Jordan with its FP64 refinement took 0.32 ms for the 700
// frequencies against 0.24 ms with the FP64 form's products on the vector
// units (the refinement adds an FP64 product as large as the elimination
// saves), 0.19 ms with them on the matrix cores.  RPA_DYSON=mixed selects it.
static bool mixed_dyson(const rpa_ctx *c) {
  if (c->d.precision != RPA_MIXED) return false;
  const char *env = std::getenv("RPA_DYSON");
  return env && std::string(env) == "mixed";
}

static int dyson_form(const rpa_ctx *c) {
  const int n = c->nlf;
  if (n <= 64) return mixed_dyson(c) ? kDyMx : kDyCol;
  if (sizeof(double2) * n * n + kDysonStatic <= kLdsMax) return kDyLds;
  return kDyGlobal;
}

static size_t dyson_lds(int form, int nlf) {
  return form == kDyGlobal ? 0 : sizeof(double2) * (size_t)nlf * nlf;
}

static const void *dyson_kernel(int form) {
  switch (form) {
    case kDyCol: return (const void *)k_dyson_col;
    case kDyMx: return (const void *)k_dyson_mx;
    case kDyLds: return (const void *)k_dyson<false>;
    default: return (const void *)k_dyson<true>;
  }
}

// Outside any stream capture (prepare): allow the kernel the dynamic LDS of
// the current N_G when that exceeds the default 64 KB.
int configure_dyson(rpa_ctx *c) {
  const int form = dyson_form(c);
  const size_t lds = dyson_lds(form, c->nlf);
  const size_t st = form == kDyMx ? kDysonMxStatic : kDysonStatic;
  if (lds + st <= 64 * 1024 || lds <= c->dyson_lds_attr[form]) return 0;
  RPA_HIP(c, hipFuncSetAttribute(dyson_kernel(form), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  c->dyson_lds_attr[form] = lds;
  return 0;
}

// Clears the singular-matrix flag ahead of each Dyson launch.  A kernel, not
// a memset, so a captured chain holds no memset node.
__global__ void k_clear_status(int *status) {
  if (threadIdx.x == 0) status[0] = 0;
}

int launch_dyson(rpa_ctx *c) {
  DysonArgs a{};
  a.nlf = c->nlf;
  a.nfreq = c->d.no - 1;
  a.chi0 = (const double2 *)c->d_chi0;
  a.V = c->d_V;
  a.epsinv = (double2 *)c->d_epsinv;
  a.chi = (double2 *)c->d_chi;
  a.W = (double2 *)c->d_W;
  a.loss = c->d_loss;
  a.status = c->d_status;
  if (c->nlf > kMaxNlf) return fail(c, RPA_E_LIMIT, "Nlf too large for the Dyson step");
  const int form = dyson_form(c);
  const size_t lds = dyson_lds(form, c->nlf);
  const size_t st = form == kDyMx ? kDysonMxStatic : kDysonStatic;
  if (lds + st > 64 * 1024 && lds > c->dyson_lds_attr[form])
    return fail(c, RPA_E_STATE, "Dyson LDS limit not configured");
  hipLaunchKernelGGL(k_clear_status, dim3(1), dim3(64), 0, c->stream, c->d_status);
  switch (form) {
    case kDyCol: hipLaunchKernelGGL(k_dyson_col, dim3(a.nfreq), dim3(256), lds, c->stream, a); break;
    case kDyMx: hipLaunchKernelGGL(k_dyson_mx, dim3(a.nfreq), dim3(256), lds, c->stream, a); break;
    case kDyLds: hipLaunchKernelGGL(k_dyson<false>, dim3(a.nfreq), dim3(256), lds, c->stream, a); break;
    default: hipLaunchKernelGGL(k_dyson<true>, dim3(a.nfreq), dim3(256), 0, c->stream, a); break;
  }
  RPA_HIP(c, hipGetLastError());
  return 0;
}

int launch_epilogue(rpa_ctx *c) {
  if (c->d.program != RPA_OPTABS) return 0;
  const int nf = c->d.no - 1;
  hipLaunchKernelGGL(k_pi, dim3((nf + 255) / 256), dim3(256), 0, c->stream, nf, c->nlf,
                     c->p.domega, c->p.eta_intra, c->p.c0, (const double2 *)c->d_chi0,
                     (const double *)c->d_qeff, c->d_pi11, c->d_sig);
  RPA_HIP(c, hipGetLastError());
  return 0;
}

}  // namespace rpa
