// The ladder (BSE) current-current response of Chi_ladd (Chi_ladd_up/BSE.f90,
// its frequency loop 1116-1176) on the device: for every frequency
// o = omin + (io-1) domega, io = 1..no, with L0(k) = 1/(o + Delta(k) + i eta)
// and NV = Ntot Vcell,
//   fock   = I - L0 F / NV                   (F the Fock kernel FockK)
//   xi0    = L0 F L0
//   lladd  = fock^-1 xi0                      (gjel inversion, 1137-1152)
//   chi_ladd(pol) = -valley sum_{k,k'} j(k) lladd(k,k') conj(j(k')) / NV^2
//   chi_para(pol) =  valley sum_k j(k) (o/Delta(k)) L0(k) conj(j(k)) / NV
// for the two current components j = current(pol, :).  The closing vector
// y = xi0 conj(j) = L0 F (L0 conj(j)) is formed first (O(n^2)), then
// fock z = y is solved by Gaussian elimination with row pivoting (in LDS up
// to 88 k, past that in a per-frequency global slice), and chi_ladd = -c
// j^T z: neither the inverse nor the reference's n^3 product lladd is formed.
// One workgroup per frequency.
#include <cmath>
#include <string>

#include "dev_linalg.h"
#include "rpa_internal.h"

namespace rpa {

namespace {
constexpr int kLdMaxK = 1024;  // k-points of the kernel (fock is n x n)
constexpr int kLdLdsK = 88;    // fock in LDS up to here (16 n^2 bytes)
}  // namespace

struct LadderArgs {
  int n, no, f0;
  double omin, dw, eta, nv, valley;
  const double2 *F;      // [k][k'] Fock kernel (row-major)
  const double *delta;   // [n]
  const double2 *cur;    // [2][n] current(pol, k)
  double2 *T;            // GLOBAL form: [frequency in chunk][n][n]
  double2 *chi_ladd;     // [io][2]
  double2 *chi_para;     // [io][2]
  int *status;
};

// dynamic LDS: fock (LDS form), then L0 [n], y [2][n], fcol [n] (double2), piv [n] (int)
inline size_t ladder_lds_bytes(int n, bool global) {
  return (global ? 0 : sizeof(double2) * (size_t)n * n) + sizeof(double2) * 4 * (size_t)n + sizeof(int) * n;
}

template <bool GLOBAL>
__global__ __launch_bounds__(256) void k_ladder(LadderArgs a) {
  extern __shared__ double2 lds_dyn[];
  __shared__ double red_v[4];
  __shared__ int red_i[4];
  __shared__ double red[5 * 8];
  const int n = a.n, tid = threadIdx.x;
  const int lf = blockIdx.x, io = a.f0 + lf + 1;
  double2 *sa = GLOBAL ? a.T + (size_t)lf * n * n : lds_dyn;
  double2 *tab = lds_dyn + (GLOBAL ? 0 : (size_t)n * n);
  double2 *l0 = tab, *y = tab + n, *fcol = tab + 3 * n;  // y: [2][n]
  int *piv = (int *)(tab + 4 * n);
  const double o = a.omin + (double)(io - 1) * a.dw;
  for (int k = tid; k < n; k += 256) l0[k] = c_div(make_double2(1.0, 0.0), make_double2(o + a.delta[k], a.eta));
  __syncthreads();
  // y_pol = L0 F (L0 conj(j_pol)) and fock = I - L0 F / NV
  const double inv_nv = 1.0 / a.nv;
  for (int k = tid; k < n; k += 256) {
    double2 s0 = make_double2(0.0, 0.0), s1 = make_double2(0.0, 0.0);
    const double2 *row = a.F + (size_t)k * n;
    for (int k2 = 0; k2 < n; ++k2) {
      const double2 f = row[k2];
      s0 = c_add(s0, c_mul(f, c_mul(l0[k2], c_conj(a.cur[k2]))));
      s1 = c_add(s1, c_mul(f, c_mul(l0[k2], c_conj(a.cur[n + k2]))));
    }
    y[k] = c_mul(l0[k], s0);
    y[n + k] = c_mul(l0[k], s1);
  }
  for (int x = tid; x < n * n; x += 256) {
    const int k = x / n, k2 = x - k * n;
    const double2 t = c_scale(c_mul(l0[k], a.F[x]), inv_nv);
    sa[x] = make_double2((k == k2 ? 1.0 : 0.0) - t.x, -t.y);
  }
  __syncthreads();
  (void)piv;
  gauss_solve<GLOBAL>(sa, n, y, 2, fcol, red_v, red_i, a.status);
  // chi_ladd(pol) = -valley j^T fock^-1 y / NV^2 (z = fock^-1 y now in y),
  // chi_para from the diagonal
  double acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) acc[e] = 0.0;
  for (int k = tid; k < n; k += 256) {
    const double2 z0 = y[k], z1 = y[n + k];
    const double2 j0 = a.cur[k], j1 = a.cur[n + k];
    const double2 t0 = c_mul(j0, z0), t1 = c_mul(j1, z1);
    acc[0] += t0.x, acc[1] += t0.y, acc[2] += t1.x, acc[3] += t1.y;
    const double2 lp = c_scale(l0[k], o / a.delta[k]);  // (o / Delta) L0
    const double2 p0 = c_mul(c_mul(j0, lp), c_conj(j0)), p1 = c_mul(c_mul(j1, lp), c_conj(j1));
    acc[4] += p0.x, acc[5] += p0.y, acc[6] += p1.x, acc[7] += p1.y;
  }
  block_sum<8>(acc, red);
  if (tid == 0) {
    const double *r = red + 32;
    const double cl = -a.valley * inv_nv * inv_nv, cp = a.valley * inv_nv;
    double2 *ol = a.chi_ladd + 2 * (size_t)(io - 1), *op = a.chi_para + 2 * (size_t)(io - 1);
    ol[0] = make_double2(cl * r[0], cl * r[1]);
    ol[1] = make_double2(cl * r[2], cl * r[3]);
    op[0] = make_double2(cp * r[4], cp * r[5]);
    op[1] = make_double2(cp * r[6], cp * r[7]);
  }
}

// Fock kernel of Chi_ladd (Chi_ladd_up/BSE.f90:804-816): one wave per pair
// (k, k') of the kernel's k-points, fockk(k,k') = -sum_{G,G'} W(iq,G,G')
// conj(M1(G)) M2(G') with iq = int(Gcar |k' - k|_xy / qmin) + 1, and on the
// diagonal the Q = 0 term -Vcell Ntot ln(1 + alpha epsq) / alpha.  Lane j
// takes columns G' = j, j + 64, ...: t = sum_G conj(M1(G)) W(iq,G,G'), then
// the wave sums t M2(G').  Pairs whose iq falls outside 2..nq get 0 (the
// reference reads W only for those and leaves such entries unset).
__global__ __launch_bounds__(256) void k_fock(int nk, int nlf, int nq, double gcar, double qmin, double diag,
                                              const double *__restrict__ kbz, const double2 *__restrict__ W,
                                              const double2 *__restrict__ M, double2 *__restrict__ fockk) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const long long pair = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (pair >= (long long)nk * nk) return;
  const int k1 = (int)(pair % nk), k2 = (int)(pair / nk);
  double2 f = make_double2(0.0, 0.0);
  if (k1 == k2) {
    f = make_double2(diag, 0.0);
  } else {
    const double qx = kbz[3 * k2] - kbz[3 * k1], qy = kbz[3 * k2 + 1] - kbz[3 * k1 + 1];
    const double q = gcar * sqrt(qx * qx + qy * qy);
    const int iq = (int)(q / qmin) + 1;
    if (iq >= 2 && iq <= nq) {
      const double2 *m = M + 2 * (size_t)nlf * pair;  // [ni][G] of the pair
      const double2 *w = W + (size_t)(iq - 1) * nlf * nlf;  // [G][G']
      double2 acc = make_double2(0.0, 0.0);
      for (int j = lane; j < nlf; j += 64) {
        double2 t = make_double2(0.0, 0.0);
        for (int g = 0; g < nlf; ++g) t = c_add(t, c_mul(c_conj(m[2 * g]), w[(size_t)g * nlf + j]));
        acc = c_add(acc, c_mul(t, m[2 * j + 1]));
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        acc.x += __shfl_xor(acc.x, off);
        acc.y += __shfl_xor(acc.y, off);
      }
      f = make_double2(-acc.x, -acc.y);
    }
  }
  if (lane == 0) fockk[pair] = f;
}

}  // namespace rpa

using namespace rpa;

extern "C" int rpa_fock_kernel(const rpa_fock_params *p, const double *kbz, const double *wt, const double *mnm,
                               double *fockk) {
  auto err = [](int code, const std::string &m) {
    set_free_error("rpa_fock_kernel: " + m);
    return code;
  };
  if (!p || !kbz || !wt || !mnm || !fockk) return err(RPA_E_ARG, "null argument");
  const int nk = p->nk, N = p->nlf, nq = p->nq;
  if (nk < 1 || N < 1 || nq < 2) return err(RPA_E_ARG, "nkbz and Nlf must be positive and nq at least 2");
  if (N > kMaxNlf) return err(RPA_E_LIMIT, "Nlf above " + std::to_string(kMaxNlf));
  if (!(p->qmin > 0.0) || !(p->c0 > 0.0) || !(p->vcell > 0.0) || p->ntot < 1)
    return err(RPA_E_ARG, "qmin, c0, Vcell and Ntot must be positive");
  const size_t NN = (size_t)N * N, npair = (size_t)nk * nk;
  // alpha from W(q = qmin) at G = G' = 0 (BSE.f90:542), epsq (:305), the
  // diagonal Q = 0 term (:816)
  const double pi = 3.141592654;  // BSE.f90's pi
  const double vq = (2.0 * pi) / p->qmin;
  const double w11 = wt[2 * (size_t)1];  // Re wt(2,1,1): (nq, Nlf, Nlf) column-major, iq = 2 first
  const double alpha = (p->c0 * vq / w11 - 1.0) / p->qmin;
  const double epsq = std::sqrt(4.0 * pi * p->c0 / (p->vcell * p->ntot));
  const double diag = -p->vcell * p->ntot * std::log(1.0 + alpha * epsq) / alpha;
  // W(nq, Nlf, Nlf) -> [iq][G][G']
  std::vector<double> hw(2 * (size_t)nq * NN);
  for (int q = 0; q < nq; ++q)
    for (int g = 0; g < N; ++g)
      for (int h = 0; h < N; ++h)
        for (int r = 0; r < 2; ++r)
          hw[2 * (((size_t)q * N + g) * N + h) + r] = wt[2 * ((size_t)q + (size_t)nq * (g + (size_t)N * h)) + r];
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return err(RPA_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  hipStream_t st = nullptr;
  double *dk = nullptr, *dw = nullptr, *dm = nullptr, *df = nullptr;
  auto cleanup = [&]() {
    if (st) (void)hipStreamSynchronize(st);
    for (double *q : {dk, dw, dm, df})
      if (q) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
  };
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc((void **)&dk, sizeof(double) * 3 * nk)) != hipSuccess ||
      (e = hipMalloc((void **)&dw, sizeof(double) * hw.size())) != hipSuccess ||
      (e = hipMalloc((void **)&dm, sizeof(double) * 4 * N * npair)) != hipSuccess ||
      (e = hipMalloc((void **)&df, sizeof(double) * 2 * npair)) != hipSuccess ||
      (e = hipMemcpyAsync(dk, kbz, sizeof(double) * 3 * nk, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(dw, hw.data(), sizeof(double) * hw.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(dm, mnm, sizeof(double) * 4 * N * npair, hipMemcpyHostToDevice, st)) != hipSuccess) {
    cleanup();
    return err(RPA_E_HIP, hipGetErrorString(e));
  }
  hipLaunchKernelGGL(k_fock, dim3((unsigned)((npair + 3) / 4)), dim3(256), 0, st, nk, N, nq, p->gcar, p->qmin, diag,
                     (const double *)dk, (const double2 *)dw, (const double2 *)dm, (double2 *)df);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(fockk, df, sizeof(double) * 2 * npair, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess) {
    cleanup();
    return err(RPA_E_HIP, hipGetErrorString(e));
  }
  cleanup();
  return RPA_OK;
}

extern "C" int rpa_ladder(const rpa_ladder_params *p, const double *fockk, const double *delta,
                          const double *current, double *chi_ladd, double *chi_para) {
  auto err = [](int code, const std::string &m) {
    set_free_error("rpa_ladder: " + m);
    return code;
  };
  if (!p || !fockk || !delta || !current || !chi_ladd || !chi_para) return err(RPA_E_ARG, "null argument");
  const int n = p->nk, no = p->no;
  if (n < 1 || no < 1) return err(RPA_E_ARG, "nkbz and no must be positive");
  if (n > kLdMaxK) return err(RPA_E_LIMIT, "more than " + std::to_string(kLdMaxK) + " k-points in the kernel");
  if (!(p->ntot > 0) || !(p->vcell > 0.0)) return err(RPA_E_ARG, "Ntot and Vcell must be positive");
  for (int k = 0; k < n; ++k)
    if (delta[k] == 0.0) return err(RPA_E_ARG, "a zero Delta(k) (chi_para divides by it)");
  const bool global = n > kLdLdsK;
  const size_t nn = (size_t)n * n;
  // GLOBAL form: at most 1 GiB of per-frequency fock slices per launch
  const int chunk = global ? (int)std::min<size_t>(no, std::max<size_t>(1, ((size_t)1 << 30) / (16 * nn))) : no;
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return err(RPA_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  // FockK(nkbz, nkbz) column-major -> [k][k'] rows; current(2, nkbz) -> [pol][k]
  std::vector<double> hF(2 * nn), hc(4 * (size_t)n);
  for (int k = 0; k < n; ++k)
    for (int k2 = 0; k2 < n; ++k2)
      for (int r = 0; r < 2; ++r) hF[2 * ((size_t)k * n + k2) + r] = fockk[2 * ((size_t)k + (size_t)n * k2) + r];
  for (int pol = 0; pol < 2; ++pol)
    for (int k = 0; k < n; ++k)
      for (int r = 0; r < 2; ++r) hc[2 * ((size_t)pol * n + k) + r] = current[2 * ((size_t)pol + 2 * (size_t)k) + r];
  hipStream_t st = nullptr;
  double2 *dF = nullptr, *dc = nullptr, *dT = nullptr, *dl = nullptr, *dp = nullptr;
  double *dd = nullptr;
  int *ds = nullptr;
  auto cleanup = [&]() {
    if (st) (void)hipStreamSynchronize(st);
    for (void *q : {(void *)dF, (void *)dc, (void *)dT, (void *)dl, (void *)dp, (void *)dd, (void *)ds})
      if (q) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
  };
  const size_t lds = ladder_lds_bytes(n, global);
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc((void **)&dF, sizeof(double2) * nn)) != hipSuccess ||
      (e = hipMalloc((void **)&dc, sizeof(double2) * 2 * n)) != hipSuccess ||
      (e = hipMalloc((void **)&dd, sizeof(double) * n)) != hipSuccess ||
      (e = hipMalloc((void **)&dl, sizeof(double2) * 2 * (size_t)no)) != hipSuccess ||
      (e = hipMalloc((void **)&dp, sizeof(double2) * 2 * (size_t)no)) != hipSuccess ||
      (e = hipMalloc((void **)&ds, sizeof(int))) != hipSuccess ||
      (global && (e = hipMalloc((void **)&dT, sizeof(double2) * nn * chunk)) != hipSuccess) ||
      (e = hipMemcpyAsync(dF, hF.data(), sizeof(double) * hF.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(dc, hc.data(), sizeof(double) * hc.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(dd, delta, sizeof(double) * n, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemsetAsync(ds, 0, sizeof(int), st)) != hipSuccess ||
      (e = hipFuncSetAttribute(global ? (const void *)k_ladder<true> : (const void *)k_ladder<false>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)) != hipSuccess) {
    cleanup();
    return err(RPA_E_HIP, hipGetErrorString(e));
  }
  LadderArgs a{};
  a.n = n;
  a.no = no;
  a.omin = p->omin;
  a.dw = p->domega;
  a.eta = p->eta;
  a.nv = (double)p->ntot * p->vcell;
  a.valley = p->valley;
  a.F = dF;
  a.delta = dd;
  a.cur = dc;
  a.T = dT;
  a.chi_ladd = dl;
  a.chi_para = dp;
  a.status = ds;
  for (int f0 = 0; f0 < no; f0 += chunk) {
    a.f0 = f0;
    const int nf = std::min(chunk, no - f0);
    if (global)
      hipLaunchKernelGGL(k_ladder<true>, dim3(nf), dim3(256), lds, st, a);
    else
      hipLaunchKernelGGL(k_ladder<false>, dim3(nf), dim3(256), lds, st, a);
    if ((e = hipGetLastError()) != hipSuccess) {
      cleanup();
      return err(RPA_E_HIP, std::string("k_ladder launch: ") + hipGetErrorString(e));
    }
  }
  int status = 0;
  if ((e = hipMemcpyAsync(chi_ladd, dl, sizeof(double2) * 2 * (size_t)no, hipMemcpyDeviceToHost, st)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(chi_para, dp, sizeof(double2) * 2 * (size_t)no, hipMemcpyDeviceToHost, st)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(&status, ds, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess) {
    cleanup();
    return err(RPA_E_HIP, hipGetErrorString(e));
  }
  cleanup();
  if (status) return err(RPA_E_SINGULAR, "singular I - L0 F / NV at some frequency (results there are not finite)");
  return RPA_OK;
}
