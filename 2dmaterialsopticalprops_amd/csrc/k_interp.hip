// interpolate.x (Chi_RPA/interpolate.f90) on the device: the current-current
// response Pi of Corrfun_{pol} / Qeff_{pol} on a dense frequency grid.
//
//   Pi_tot(io) = K1[S0](io) + Pi_dia,  Pi_dia = -Re K1[S0](1)   io = 1..No-1
//        (genReChi0/genImChi0 are Optabs' K1 stencil, interpolate.f90:242-464;
//         interpolate.f90:105-135)
//   dense point c = (io-1)(No_interp-1) + jo, io = 1..No-2, jo = 1..No_interp-1,
//   oi = c * domega_tot:
//   Pi_interp(c) = Pi_tot(io) + (Pi_tot(io+1) - Pi_tot(io))/No_interp * jo
//                  + Qeff oi / (oi + i eta_intra)          interpolate.f90:165-180
//
// The KK part is the Optabs GEMM of k_kk.hip over the (iG, jG) columns the
// call asks for (every pair, or only (1,1)); k_pi_dense then forms a tile of
// 64 dense points x 32 pairs in LDS from the coarse Pi rows (pair-contiguous
// reads) and writes it point-contiguous, which is the reference's
// Pi_tot_interp(No_tot, Nlf, Nlf) layout.  The kernel is bound by its HBM
// writes: 16 bytes per (point, pair) against 32 bytes read per
// (No_interp-1) points.
#include <string>

#include "rpa_internal.h"

namespace rpa {

constexpr int kIpC = 64;  // dense points per tile
constexpr int kIpP = 32;  // (iG, jG) pairs per tile

// chi0[io-1][pair] (complex, pair-contiguous) -> out[pair][c] (complex,
// point-contiguous), pair = iG + Nlf jG; intra[c] = Qeff(1,1) oi/(oi + i eta)
__global__ __launch_bounds__(256) void k_pi_dense(int no_interp, long long no_tot, int npair, double dwt, double eta,
                                                  const double *__restrict__ chi0, const double *__restrict__ qeff,
                                                  double *__restrict__ out, double *__restrict__ intra) {
#pragma clang fp contract(off)
  __shared__ double tr[kIpC][kIpP + 1], ti[kIpC][kIpP + 1];
  const long long c0 = (long long)blockIdx.x * kIpC;
  const int p0 = blockIdx.y * kIpP, tid = threadIdx.x;
  const int ni1 = no_interp - 1;
  const double rn = (double)no_interp;
#pragma unroll
  for (int u = 0; u < kIpC * kIpP / 256; ++u) {
    const int cl = tid / kIpP + (256 / kIpP) * u, pl = tid % kIpP;
    const long long c = c0 + cl;  // 0-based dense point, counter - 1
    const int pr = p0 + pl;
    if (c < no_tot && pr < npair) {
      const long long io = c / ni1 + 1;  // 1..No-2
      const int jo = (int)(c - (io - 1) * ni1) + 1;
      const double dia = -chi0[2 * (size_t)pr];  // -Re Pi0(1), interpolate.f90:118-120
      const size_t a = 2 * ((size_t)(io - 1) * npair + pr), b = a + 2 * (size_t)npair;
      const double ar = chi0[a] + dia, ai = chi0[a + 1];
      const double br = chi0[b] + dia, bi = chi0[b + 1];
      const double dr = (br - ar) / rn, di = (bi - ai) / rn;
      double vr = ar + dr * (double)jo, vi = ai + di * (double)jo;
      // + Qeff oi / (oi + i eta), interpolate.f90:176
      const double oi = (double)(c + 1) * dwt;
      const double qr = qeff[2 * pr] * oi, qi = qeff[2 * pr + 1] * oi;
      const double den = oi * oi + eta * eta;
      vr += (qr * oi + qi * eta) / den;
      vi += (qi * oi - qr * eta) / den;
      tr[cl][pl] = vr;
      ti[cl][pl] = vi;
      if (pr == 0 && intra) {
        intra[2 * c] = (qr * oi + qi * eta) / den;
        intra[2 * c + 1] = (qi * oi - qr * eta) / den;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kIpC * kIpP / 256; ++u) {
    const int cl = tid % kIpC, pl = tid / kIpC + (256 / kIpC) * u;
    const long long c = c0 + cl;
    const int pr = p0 + pl;
    if (c < no_tot && pr < npair) {
      const size_t o = 2 * ((size_t)pr * no_tot + c);
      out[o] = tr[cl][pl];
      out[o + 1] = ti[cl][pl];
    }
  }
}

// S0(-No:No, pairs) as the caller holds it (frequency fastest) -> the KK's
// frequency rows s0[w][pair] (interleaved complex)
__global__ void k_interp_pack(long long nw, int npair, const double2 *__restrict__ raw, double2 *__restrict__ s0) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nw * npair) return;
  const long long w = t % nw, pr = t / nw;
  s0[w * npair + pr] = raw[t];
}

namespace {

// device buffers and the stream of one call, released on every return path
struct InterpCall {
  hipStream_t s = nullptr;
  double *s0 = nullptr, *kmat = nullptr, *x = nullptr, *chi0 = nullptr, *q = nullptr, *out = nullptr,
         *intra = nullptr, *raw = nullptr;
  ~InterpCall() {
    if (s) (void)hipStreamSynchronize(s);
    for (double *p : {s0, kmat, x, chi0, q, out, intra, raw})
      if (p) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
};

int err(int code, const std::string &msg) {
  set_free_error("rpa_interpolate: " + msg);
  return code;
}

int hip_err(hipError_t e, const char *what) {
  return err(RPA_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

}  // namespace rpa

using namespace rpa;

extern "C" int rpa_interpolate(const rpa_interp_params *p, const double *S0, const double *qeff, double *pi_dense,
                               double *pi_intra) {
  if (!p || !S0 || !qeff || !pi_dense) return err(RPA_E_ARG, "null argument");
  const int no = p->no, nlf = p->nlf, nint = p->no_interp;
  if (no < 4) return err(RPA_E_ARG, "No must be at least 4");
  if (nlf < 1) return err(RPA_E_ARG, "Nlf must be positive");
  if (nlf > 4096) return err(RPA_E_LIMIT, "Nlf above 4096");
  if (nint < 1) return err(RPA_E_ARG, "No_interp must be positive");
  if (!(p->domega > 0.0) || !(p->domega_tot > 0.0) || !(p->eta_intra >= 0.0))
    return err(RPA_E_ARG, "domega and domega_tot must be positive, eta_intra not negative");
  const long long no_tot = (long long)(nint - 1) * (no - 2);  // interpolate.f90:54
  if (no_tot == 0) return RPA_OK;
  const int npair = p->all ? nlf * nlf : 1;
  const int ncol = 2 * npair;
  const size_t nw = 2 * (size_t)no + 1, nf = (size_t)(no - 1);
  if ((no_tot + kIpC - 1) / kIpC > 0x7fffffffLL) return err(RPA_E_LIMIT, "too many dense points");

  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return hip_err(e, "hipSetDevice");
  InterpCall call;
  if ((e = hipStreamCreateWithFlags(&call.s, hipStreamNonBlocking)) != hipSuccess)
    return hip_err(e, "hipStreamCreateWithFlags");
  const size_t n_out = 2 * (size_t)no_tot * npair;
  struct {
    double **ptr;
    size_t n;
  } bufs[] = {{&call.s0, nw * ncol},  {&call.kmat, nf * nw},   {&call.x, nf * ncol},       {&call.chi0, nf * ncol},
              {&call.q, (size_t)ncol}, {&call.out, n_out},     {&call.intra, 2 * (size_t)no_tot},
              {&call.raw, nw * ncol}};
  for (auto &b : bufs)
    if ((e = hipMalloc((void **)b.ptr, sizeof(double) * b.n)) != hipSuccess) return hip_err(e, "hipMalloc");

  // S0(-No:No, Nlf, Nlf) and Qeff(Nlf, Nlf), column-major complex: column
  // pair = iG + Nlf jG of the device rows is the reference's own order (the
  // first npair columns: all of them, or (1,1) alone)
  if ((e = hipMemcpyAsync(call.raw, S0, sizeof(double) * nw * ncol, hipMemcpyHostToDevice, call.s)) != hipSuccess)
    return hip_err(e, "hipMemcpyAsync");
  {
    const long long n = (long long)nw * npair;
    hipLaunchKernelGGL(k_interp_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, call.s, (long long)nw, npair,
                       (const double2 *)call.raw, (double2 *)call.s0);
    if ((e = hipGetLastError()) != hipSuccess) return hip_err(e, "k_interp_pack launch");
  }
  if ((e = hipMemcpyAsync(call.q, qeff, sizeof(double) * ncol, hipMemcpyHostToDevice, call.s)) != hipSuccess)
    return hip_err(e, "hipMemcpyAsync");
  if ((e = kk_optabs_columns(call.s, no, p->all ? nlf : 1, ncol, p->domega, call.s0, call.kmat, call.x, call.chi0)) !=
      hipSuccess)
    return hip_err(e, "Kramers-Kronig launch");
  dim3 grid((unsigned)((no_tot + kIpC - 1) / kIpC), (unsigned)((npair + kIpP - 1) / kIpP));
  hipLaunchKernelGGL(k_pi_dense, grid, dim3(256), 0, call.s, nint, no_tot, npair, p->domega_tot, p->eta_intra,
                     (const double *)call.chi0, (const double *)call.q, call.out, call.intra);
  if ((e = hipGetLastError()) != hipSuccess) return hip_err(e, "k_pi_dense launch");
  if ((e = hipMemcpyAsync(pi_dense, call.out, sizeof(double) * n_out, hipMemcpyDeviceToHost, call.s)) != hipSuccess)
    return hip_err(e, "hipMemcpyAsync");
  if (pi_intra &&
      (e = hipMemcpyAsync(pi_intra, call.intra, sizeof(double) * 2 * no_tot, hipMemcpyDeviceToHost, call.s)) !=
          hipSuccess)
    return hip_err(e, "hipMemcpyAsync");
  if ((e = hipStreamSynchronize(call.s)) != hipSuccess) return hip_err(e, "hipStreamSynchronize");
  return RPA_OK;
}
