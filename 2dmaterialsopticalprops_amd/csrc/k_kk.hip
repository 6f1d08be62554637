// Kramers-Kronig transform of S0 as a GEMM with the reference's discrete
// stencil:  X[io'][col] = sum_w K[io'][w] S0[w][col],  io = io'+1 = 1..no-1.
//
//  K2 (Sloss genChi0, Sloss.f90:954-1059): if / else-if chain, so the
//      corrections act only at jo == io (coefficient 0); io = 1 is the static
//      limit -2 sum f_j S0(j) with f_2 = 3/2 and a half weight at jo = no.
//  K1 (Optabs genReChi0/genImChi0, Optabs.f90:1112-1318; equal to the complex
//      variant of Modules/response_function_KK.f90:74-178): the same c+ and c-
//      weights with the +-3/2, 1, -1 and half-weight corrections, applied to
//      S0(jo-1) and S0(-jo+1) of the (-no:no) grid; the +-pi S0(io-1) terms
//      are added in the epilogue.
// Each stencil entry is evaluated independently with the reference's chain,
// so the matrix is exactly the coefficients the reference multiplies with.
#include <mutex>
#include <string>

#include "rpa_internal.h"

namespace rpa {

// c- weight of the Optabs stencil for output io at input jo (io >= 2)
__device__ inline double k1_cminus(int io, int jo, int no, double oi, double oj, double dw) {
#pragma clang fp contract(off)
  double fact = 0.0;
  if (jo != io) fact = dw / (oi - oj);
  if (io == 2) {
    if (jo == 1) fact = 1.0;
    else if (jo == 2) fact = 0.0;
    else if (jo == 3) fact = -3.0 / 2.0;
    else if (jo == no) fact = 0.5 * dw / (oi - oj);
  } else if (io == no - 1) {
    if (jo == 1) fact = 0.5 * dw / (oi - oj);
    else if (jo == no - 2) fact = 3.0 / 2.0;
    else if (jo == no - 1) fact = 0.0;
    else if (jo == no) fact = -1.0;
  } else {
    if (jo == 1) fact = 0.5 * dw / (oi - oj);
    else if (jo == io - 1) fact = 3.0 / 2.0;
    else if (jo == io) fact = 0.0;
    else if (jo == io + 1) fact = -3.0 / 2.0;
    else if (jo == no) fact = 0.5 * dw / (oi - oj);
  }
  return fact;
}

// c- weight of the Sloss stencil (io >= 2): plain 1/(oi-oj), zero at jo == io
__device__ inline double k2_cminus(int io, int jo, double oi, double oj, double dw) {
#pragma clang fp contract(off)
  return (jo != io) ? dw / (oi - oj) : 0.0;
}

__device__ inline double c_plus(int jo, int no, double oi, double oj, double dw) {
#pragma clang fp contract(off)
  double fact = dw / (oi + oj);
  if (jo == 1 || jo == no) fact = 0.5 * dw / (oi + oj);
  return fact;
}

// static-limit weight f_j (io = 1), both stencils
__device__ inline double f_static(int jo, int no, double oj, double dw) {
#pragma clang fp contract(off)
  double fact = dw / oj;
  if (jo == 2) fact = 3.0 / 2.0;
  else if (jo == no) fact = 0.5 * dw / oj;
  return fact;
}

__global__ void k_kstencil(int program, int no, int nw, double dw, double *K) {
#pragma clang fp contract(off)
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)(no - 1) * nw) return;
  const int iop = (int)(idx / nw), w = (int)(idx - (long long)iop * nw);
  const int io = iop + 1;
  const double oi = (double)(io - 1) * dw;
  double coef = 0.0;
  if (program == RPA_SLOSS) {
    const int jo = w + 1;  // S0(jo), jo = 1..no
    const double oj = (double)(jo - 1) * dw;
    if (io == 1) {
      coef = (jo >= 2) ? -2.0 * f_static(jo, no, oj, dw) : 0.0;
    } else {
      coef = k2_cminus(io, jo, oi, oj, dw) - c_plus(jo, no, oi, oj, dw);
    }
  } else {
    const int f = w - no;  // frequency index of S0(f), f = -no..no
    if (io == 1) {
      if (f > 0 && f + 1 <= no) {
        int jo = f + 1;
        coef = -f_static(jo, no, (double)(jo - 1) * dw, dw);
      } else if (f < 0 && -f + 1 <= no) {
        int jo = -f + 1;
        coef = f_static(jo, no, (double)(jo - 1) * dw, dw);
      }
    } else {
      if (f >= 0 && f + 1 <= no) {
        int jo = f + 1;
        coef += k1_cminus(io, jo, no, oi, (double)(jo - 1) * dw, dw);
      }
      if (f <= 0 && -f + 1 <= no) {
        int jo = -f + 1;
        coef += c_plus(jo, no, oi, (double)(jo - 1) * dw, dw);
      }
    }
  }
  K[idx] = coef;
}

// C[M][N] = A[M][K] B[K][N], row-major, FP64 matrix cores, 64x64 tiles,
// K in chunks of 64.  The next chunk's operands are loaded into registers
// (from clamped addresses, out-of-range values zeroed when stored to LDS)
// while the current chunk runs on the matrix cores.
constexpr int kGemmKC = 64;
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, const double *__restrict__ A,
                                                   const double *__restrict__ B, double *__restrict__ C) {
  constexpr int KC = kGemmKC, U = 64 * KC / 256;
  __shared__ double sA[64][KC + 1];
  __shared__ double sB[KC][64 + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, q4 = lane >> 4;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  rpa_d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = rpa_d4{0.0, 0.0, 0.0, 0.0};
  // thread roles: A element (x >> 6 row, x & 63 k), B element (x >> 6 k, x & 63 column), x = tid + 256 u
  double ra[U], rb[U];
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int x = tid + 256 * u, hi = x >> 6, lo = x & 63;
      ra[u] = A[(size_t)min(m0 + hi, M - 1) * K + min(k0 + lo, K - 1)];
      rb[u] = B[(size_t)min(k0 + hi, K - 1) * N + min(n0 + lo, N - 1)];
    }
  };
  auto store = [&](int k0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int x = tid + 256 * u, hi = x >> 6, lo = x & 63;
      sA[hi][lo] = (m0 + hi < M && k0 + lo < K) ? ra[u] : 0.0;
      sB[hi][lo] = (k0 + hi < K && n0 + lo < N) ? rb[u] : 0.0;
    }
  };
  load(0);
  for (int k0 = 0; k0 < K; k0 += KC) {
    __syncthreads();  // the previous chunk's reads are done
    store(k0);
    __syncthreads();
    load(min(k0 + KC, K - 1));  // past the end: a clamped reload, never stored
#pragma unroll
    for (int kk = 0; kk < KC / 4; ++kk) {
      const int k = kk * 4 + q4;
      double av[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i] = sA[wm + 16 * i + r][k];
#pragma unroll
      for (int j = 0; j < 2; ++j) bv[j] = sB[k][wn + 16 * j + r];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma_f64_16x16x4(av[i], bv[j], acc[i][j]);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int gm = m0 + wm + 16 * i + 4 * e + q4, gn = n0 + wn + 16 * j + r;
        if (gm < M && gn < N) C[(size_t)gm * N + gn] = acc[i][j][e];
      }
}

// chi0[io'][iG][jG] (complex) from X = K S0 and S0:
//   Sloss:  (X, -pi S0(io))                          Sloss.f90:1053-1054
//   Optabs: (X_re + pi Im S0(io-1), X_im - pi Re S0(io-1))  Optabs.f90:1209,1316
__global__ void k_chi0_epilogue(int program, int no, int nlf, int ncol_pad, const double *X,
                                const double *s0, double *chi0) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nn = (long long)nlf * nlf;
  if (idx >= (long long)(no - 1) * nn) return;
  const int iop = (int)(idx / nn);
  const int pr = (int)(idx - (long long)iop * nn);
  double re, im;
  if (program == RPA_SLOSS) {
    re = X[(size_t)iop * ncol_pad + pr];
    im = -M_PI * s0[(size_t)iop * ncol_pad + pr];
  } else {
    const size_t w = (size_t)iop + no;  // frequency index io-1 = iop
    re = X[(size_t)iop * ncol_pad + 2 * pr] + M_PI * s0[w * ncol_pad + 2 * pr + 1];
    im = X[(size_t)iop * ncol_pad + 2 * pr + 1] - M_PI * s0[w * ncol_pad + 2 * pr];
  }
  chi0[2 * idx] = re;
  chi0[2 * idx + 1] = im;
}

int launch_kstencil(rpa_ctx *c) {
  const int no = c->d.no, nw = c->nw;
  const long long n = (long long)(no - 1) * nw;
  hipLaunchKernelGGL(k_kstencil, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     c->d.program, no, nw, c->p.domega, c->d_kmat);
  RPA_HIP(c, hipGetLastError());
  c->kmat_no = no;
  c->kmat_dw = c->p.domega;
  return 0;
}

int launch_kk(rpa_ctx *c) {
  const int no = c->d.no, nw = c->nw;
  dim3 grid((c->ncol_pad + 63) / 64, (no - 1 + 63) / 64);
  hipLaunchKernelGGL(k_gemm_f64, grid, dim3(256), 0, c->stream, no - 1, c->ncol_pad, nw,
                     (const double *)c->d_kmat, (const double *)c->d_s0, c->d_x);
  const long long n = (long long)(no - 1) * c->nlf * c->nlf;
  hipLaunchKernelGGL(k_chi0_epilogue, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     c->d.program, no, c->nlf, c->ncol_pad, (const double *)c->d_x,
                     (const double *)c->d_s0, c->d_chi0);
  RPA_HIP(c, hipGetLastError());
  return 0;
}

hipError_t kk_optabs_columns(hipStream_t s, int no, int nlf, int ncol, double dw, const double *s0, double *kmat,
                             double *x, double *chi0) {
  const int nw = 2 * no + 1;
  const long long nk = (long long)(no - 1) * nw;
  hipLaunchKernelGGL(k_kstencil, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, s, (int)RPA_OPTABS, no, nw, dw,
                     kmat);
  dim3 grid((ncol + 63) / 64, (no - 1 + 63) / 64);
  hipLaunchKernelGGL(k_gemm_f64, grid, dim3(256), 0, s, no - 1, ncol, nw, (const double *)kmat, s0, x);
  const long long n = (long long)(no - 1) * nlf * nlf;
  hipLaunchKernelGGL(k_chi0_epilogue, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int)RPA_OPTABS, no, nlf,
                     ncol, (const double *)x, s0, chi0);
  return hipGetLastError();
}

// ---- Kramers-Kronig without allocation (rpa_kk_noalloc) ------------------
// The stencil K and the product X = K S0 of this call live in a file-scope
// workspace that every call on a device shares, instead of the context's
// buffers; the stencil is rebuilt on each call, since another context (or
// another shape) may have used the workspace since.  Calls that overlap on
// one device race on it.
constexpr size_t kNoAllocDoubles = (size_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[kMaxDevices];

// Once per device, with that device current: the device address of the
// workspace (the array's name is only a host-side stand-in for it).
int noalloc_init_device(int device) {
  if (device < 0 || device >= kMaxDevices) return RPA_E_LIMIT;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return 0;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return RPA_E_HIP;
  g_scratch[device] = (double *)ptr;
  return 0;
}

int launch_kk_noalloc(rpa_ctx *c, double *chi0_dev) {
  const int dev = c->d.device;
  double *ws = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ws = (dev >= 0 && dev < kMaxDevices) ? g_scratch[dev] : nullptr;
  }
  if (!ws) return fail(c, RPA_E_STATE, "rpa_kk_noalloc: no workspace address for this device");
  const int no = c->d.no, nw = c->nw;
  const size_t nk = (size_t)(no - 1) * nw, nx = (size_t)(no - 1) * c->ncol_pad;
  if (nk + nx > kNoAllocDoubles)
    return fail(c, RPA_E_LIMIT, "rpa_kk_noalloc: stencil and product need " + std::to_string((nk + nx) * 8 >> 20) +
                                    " MiB, more than the 64 MiB workspace");
  double *kmat = ws, *x = ws + (nk + 31) / 32 * 32;
  if ((size_t)(x - ws) + nx > kNoAllocDoubles)
    return fail(c, RPA_E_LIMIT, "rpa_kk_noalloc: stencil and product exceed the 64 MiB workspace");
  hipLaunchKernelGGL(k_kstencil, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, c->stream, c->d.program, no, nw,
                     c->p.domega, kmat);
  dim3 grid((c->ncol_pad + 63) / 64, (no - 1 + 63) / 64);
  hipLaunchKernelGGL(k_gemm_f64, grid, dim3(256), 0, c->stream, no - 1, c->ncol_pad, nw, (const double *)kmat,
                     (const double *)c->d_s0, x);
  const long long n = (long long)(no - 1) * c->nlf * c->nlf;
  hipLaunchKernelGGL(k_chi0_epilogue, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->d.program, no,
                     c->nlf, c->ncol_pad, (const double *)x, (const double *)c->d_s0, chi0_dev);
  RPA_HIP(c, hipGetLastError());
  RPA_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

// ---- GW propagators Gamma+- (the Kramers-Kronig of a real spectrum S) ----
// Stencil rows io' = io-1 = 0..no-2 give W1(io), rows no-1+io' give W2(io):
//   io = 1:  W1 = -sum_{jo>=2} f_jo S(jo), W2 = -W1 (static limit)
//   io >= 2: W1 = sum_jo c-(io,jo) S(jo), W2 = sum_jo c+(io,jo) S(jo)
// with c- of the Sloss stencil K2 (genW1W2, Sloss.f90:1062-1171) or of the
// Optabs stencil K1 (genGammaPropagator, Modules/propagators.f90:35-92,
// whose chain of ifs gives K1's weights for No >= 4).
__global__ void k_gamma_stencil(int kind, int no, double dw, double *K) {
#pragma clang fp contract(off)
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 2LL * (no - 1) * no) return;
  const int r = (int)(idx / no), jo = (int)(idx - (long long)r * no) + 1;
  const bool w2 = r >= no - 1;
  const int io = (w2 ? r - (no - 1) : r) + 1;
  const double oi = (double)(io - 1) * dw, oj = (double)(jo - 1) * dw;
  double c;
  if (io == 1) {
    c = (jo >= 2) ? -f_static(jo, no, oj, dw) : 0.0;
    if (w2) c = -c;
  } else if (!w2) {
    c = (kind == RPA_KK_SLOSS) ? k2_cminus(io, jo, oi, oj, dw) : k1_cminus(io, jo, no, oi, oj, dw);
  } else {
    c = c_plus(jo, no, oi, oj, dw);
  }
  K[idx] = c;
}

// X = [W1; W2] (rows io', series columns) -> Gamma+ = (W1, -pi S(io)),
// Gamma- = (-W2, 0) as [io'][series] complex
__global__ void k_gamma_epilogue(int no, int nser, double pi_s, const double *X, const double *S, double *gp,
                                 double *gm) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)(no - 1) * nser) return;
  const long long io0 = idx / nser, ser = idx - io0 * nser;
  const double w1 = X[idx], w2 = X[(size_t)(no - 1) * nser + idx];
  gp[2 * idx] = w1;
  gp[2 * idx + 1] = -pi_s * S[idx];
  gm[2 * idx] = -w2;
  gm[2 * idx + 1] = 0.0;
  (void)io0;
  (void)ser;
}

// genGammaPropagator's Kramers-Kronig check per series (propagators.f90:97-105):
// KKS = sum_io f (W1 - W2 - Re D(io))^2, SKK = sum_io f Re D(io)^2, f = domega
// (half at io = 1 and No-1), io = 1..No-1 in order
__global__ void k_gamma_check(int no, int nser, double dw, const double *X, const double *dre, double *kks,
                              double *skk) {
#pragma clang fp contract(off)
  const int ser = blockIdx.x * blockDim.x + threadIdx.x;
  if (ser >= nser) return;
  double a = 0.0, b = 0.0;
  for (int io = 1; io <= no - 1; ++io) {
    const size_t x = (size_t)(io - 1) * nser + ser;
    const double wind = dre[x];
    const double windkk = X[x] + (-X[(size_t)(no - 1) * nser + x]);
    double f = dw;
    if (io == 1 || io == no - 1) f = 0.5 * dw;
    a = a + f * (windkk - wind) * (windkk - wind);
    b = b + f * wind * wind;
  }
  kks[ser] = a;
  skk[ser] = b;
}

// D(No, nser) column-major (frequency fastest) -> S[jo][series] = -(1/pi) Im D
// and R[jo][series] = Re D
__global__ void k_gamma_in(int no, int nser, double pi_s, const double2 *__restrict__ D, double *__restrict__ S,
                           double *__restrict__ R) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)no * nser) return;
  const int jo = (int)(t % no);
  const long long s = t / no;
  const double2 d = D[t];
  S[(size_t)jo * nser + s] = -((1.0 / pi_s) * d.y);
  R[(size_t)jo * nser + s] = d.x;
}

// [io'][series] complex -> (No-1, nser) column-major
__global__ void k_gamma_out(int nf, int nser, const double2 *__restrict__ in, double2 *__restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nf * nser) return;
  const int i = (int)(t % nf);
  const long long s = t / nf;
  out[t] = in[(size_t)i * nser + s];
}

// S(jo, pair) of W for the GW propagators (Sloss.f90:489): -(1/pi) Im W for
// jo = 1..no-1; S(no) keeps chi0's spectral value S0(no), which the
// reference does not overwrite
__global__ void k_gamma_s_of_w(int no, int npair, int ncol_pad, const double *W, const double *s0, double *S) {
#pragma clang fp contract(off)
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)no * npair) return;
  const int jo = (int)(idx / npair), pr = (int)(idx - (long long)jo * npair);
  const double pi = 3.14159265358979323846;
  S[idx] = (jo < no - 1) ? -((1.0 / pi) * W[2 * ((size_t)jo * npair + pr) + 1]) : s0[(size_t)(no - 1) * ncol_pad + pr];
}

// stencil, product and epilogue of the Gamma propagators for nser series
// S[no][nser] on a stream; ws holds 2(no-1)no + 2(no-1)nser doubles
static hipError_t gamma_series(hipStream_t st, int kind, int no, int nser, double dw, double pi_s, const double *S,
                               double *ws, double *gp, double *gm, double **x_out) {
  double *K = ws, *X = ws + 2 * (size_t)(no - 1) * no;
  const long long nk = 2LL * (no - 1) * no;
  hipLaunchKernelGGL(k_gamma_stencil, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, st, kind, no, dw, K);
  dim3 grid((nser + 63) / 64, (2 * (no - 1) + 63) / 64);
  hipLaunchKernelGGL(k_gemm_f64, grid, dim3(256), 0, st, 2 * (no - 1), nser, no, (const double *)K, S, X);
  const long long n = (long long)(no - 1) * nser;
  hipLaunchKernelGGL(k_gamma_epilogue, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, no, nser, pi_s,
                     (const double *)X, S, gp, gm);
  if (x_out) *x_out = X;
  return hipGetLastError();
}

size_t gamma_ws_doubles(int no, int nser) { return 2 * (size_t)(no - 1) * no + 2 * (size_t)(no - 1) * nser; }

int launch_gamma_w(rpa_ctx *c, double *ws, double *S, double *gp, double *gm) {
  const int no = c->d.no, np = c->nlf * c->nlf;
  const long long n = (long long)no * np;
  hipLaunchKernelGGL(k_gamma_s_of_w, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, no, np,
                     c->ncol_pad, (const double *)c->d_W, (const double *)c->d_s0, S);
  RPA_HIP(c, gamma_series(c->stream, RPA_KK_SLOSS, no, np, c->p.domega, 3.14159265358979323846, S, ws, gp, gm,
                          nullptr));
  return 0;
}

// Kramers-Kronig self-check of the screened interaction (Sloss.f90:489-534,
// genW1W2 1062-1171) at (G, G') = (1, 1): the spectral function
// S_W(io) = -(1/pi) Im W(io,1,1) for io = 1..no-1 (S0(no,1,1) keeps its
// chi0 value, the reference overwrites only 1..no-1), then per io
//   io = 1:  W1 = -sum_{jo>=2} f_jo S_W(jo),  W2 = -W1
//   io >= 2: W1 = sum_jo c-(io,jo) S_W(jo),  W2 = sum_jo c+(io,jo) S_W(jo)
// with the K2 weights.  One workgroup; S_W staged in LDS.
__global__ void k_w_kk(int no, double dw, int nlf, const double *W, const double *s0_last,
                       double *w12) {
#pragma clang fp contract(off)
  extern __shared__ double sw[];
  const double pi = 3.14159265358979323846;
  const size_t stride = 2 * (size_t)nlf * nlf;  // complex [f][i][j]
  for (int j = threadIdx.x; j < no; j += blockDim.x)
    sw[j] = (j < no - 1) ? -((1.0 / pi) * W[(size_t)j * stride + 1]) : s0_last[0];
  __syncthreads();
  for (int io = 1 + threadIdx.x; io <= no - 1; io += blockDim.x) {
    const double oi = (double)(io - 1) * dw;
    double w1 = 0.0, w2 = 0.0;
    if (io == 1) {
      for (int jo = 2; jo <= no; ++jo) w1 = w1 - f_static(jo, no, (double)(jo - 1) * dw, dw) * sw[jo - 1];
      w2 = -w1;
    } else {
      for (int jo = 1; jo <= no; ++jo) {
        const double oj = (double)(jo - 1) * dw;
        w1 = w1 + k2_cminus(io, jo, oi, oj, dw) * sw[jo - 1];
        w2 = w2 + c_plus(jo, no, oi, oj, dw) * sw[jo - 1];
      }
    }
    w12[io - 1] = w1;
    w12[no - 1 + io - 1] = w2;
  }
}

int launch_w_kk(rpa_ctx *c) {
  const int no = c->d.no;
  const size_t lds = sizeof(double) * (size_t)no;
  if (lds > 64 * 1024) return fail(c, RPA_E_LIMIT, "too many frequencies for the W Kramers-Kronig check");
  hipLaunchKernelGGL(k_w_kk, dim3(1), dim3(256), lds, c->stream, no, c->p.domega, c->nlf,
                     (const double *)c->d_W, (const double *)c->d_s0 + (size_t)(no - 1) * c->ncol_pad,
                     c->d_wkk);
  RPA_HIP(c, hipGetLastError());
  return 0;
}

}  // namespace rpa

using namespace rpa;

extern "C" int rpa_gamma_propagator(int no, int nser, double domega, const double *D, double *gammap, double *gammam,
                                    double *kks, double *skk, int device) {
  auto err = [](int code, const std::string &m) {
    set_free_error("rpa_gamma_propagator: " + m);
    return code;
  };
  if (!D || !gammap || !gammam || !kks || !skk) return err(RPA_E_ARG, "null argument");
  if (no < 4 || nser < 1) return err(RPA_E_ARG, "No must be at least 4 and the series at least one");
  if (!(domega > 0.0)) return err(RPA_E_ARG, "domega must be positive");
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return err(RPA_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  const double pi_s = 3.141592654;  // propagators.f90's pi
  const size_t nS = (size_t)no * nser, nG = (size_t)(no - 1) * nser;
  hipStream_t st = nullptr;
  double *S = nullptr, *R = nullptr, *ws = nullptr, *gp = nullptr, *gm = nullptr, *a = nullptr, *b = nullptr,
         *dD = nullptr, *go = nullptr;
  auto cleanup = [&]() {
    if (st) (void)hipStreamSynchronize(st);
    for (double *p : {S, R, ws, gp, gm, a, b, dD, go})
      if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  };
  double *X = nullptr;
  const long long nin = (long long)nS, nout = (long long)nG;
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc((void **)&dD, sizeof(double) * 2 * nS)) != hipSuccess ||
      (e = hipMalloc((void **)&S, sizeof(double) * nS)) != hipSuccess ||
      (e = hipMalloc((void **)&R, sizeof(double) * nS)) != hipSuccess ||
      (e = hipMalloc((void **)&ws, sizeof(double) * gamma_ws_doubles(no, nser))) != hipSuccess ||
      (e = hipMalloc((void **)&gp, sizeof(double) * 2 * nG)) != hipSuccess ||
      (e = hipMalloc((void **)&gm, sizeof(double) * 2 * nG)) != hipSuccess ||
      (e = hipMalloc((void **)&go, sizeof(double) * 2 * nG)) != hipSuccess ||
      (e = hipMalloc((void **)&a, sizeof(double) * nser)) != hipSuccess ||
      (e = hipMalloc((void **)&b, sizeof(double) * nser)) != hipSuccess ||
      (e = hipMemcpyAsync(dD, D, sizeof(double) * 2 * nS, hipMemcpyHostToDevice, st)) != hipSuccess) {
    cleanup();
    return err(RPA_E_HIP, hipGetErrorString(e));
  }
  hipLaunchKernelGGL(k_gamma_in, dim3((unsigned)((nin + 255) / 256)), dim3(256), 0, st, no, nser, pi_s,
                     (const double2 *)dD, S, R);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = gamma_series(st, RPA_KK_OPTABS, no, nser, domega, pi_s, S, ws, gp, gm, &X)) != hipSuccess) {
    cleanup();
    return err(RPA_E_HIP, hipGetErrorString(e));
  }
  hipLaunchKernelGGL(k_gamma_check, dim3((unsigned)((nser + 255) / 256)), dim3(256), 0, st, no, nser, domega,
                     (const double *)X, (const double *)R, a, b);
  const dim3 go_grid((unsigned)((nout + 255) / 256));
  hipLaunchKernelGGL(k_gamma_out, go_grid, dim3(256), 0, st, no - 1, nser, (const double2 *)gp, (double2 *)go);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(gammap, go, sizeof(double) * 2 * nG, hipMemcpyDeviceToHost, st)) != hipSuccess) {
    cleanup();
    return err(RPA_E_HIP, hipGetErrorString(e));
  }
  // Gamma+ in [io'][series] has been read by the first transposition (stream
  // order), so its buffer takes the transposed Gamma-
  hipLaunchKernelGGL(k_gamma_out, go_grid, dim3(256), 0, st, no - 1, nser, (const double2 *)gm, (double2 *)gp);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(gammam, gp, sizeof(double) * 2 * nG, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipMemcpyAsync(kks, a, sizeof(double) * nser, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipMemcpyAsync(skk, b, sizeof(double) * nser, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess) {
    cleanup();
    return err(RPA_E_HIP, hipGetErrorString(e));
  }
  cleanup();
  return RPA_OK;
}
