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
// fock z = y is solved with row pivoting, and chi_ladd = -c j^T z: neither
// the inverse nor the reference's n^3 product lladd is formed.
//
// Up to 88 k-points one workgroup per frequency eliminates in LDS.  Past
// that the solve is blocked over all frequencies of a chunk at once, each
// fock in its own slice [n][ld] with y in columns n, n+1:
//   k_lb_form    fock and y per frequency;
//   k_lb_panel   per panel of NB columns, one workgroup per frequency: the
//                panel's active rows in registers (one per thread), NB
//                pivot steps (|Re| + |Im|, lowest row on ties) that choose
//                rows without moving them, the multipliers written in place,
//                the active-row list compacted, and U12 = L11^-1 A12 on the
//                pivot rows for every column right of the panel;
//   k_lb_update  the trailing rows lose L21 U12, a batched complex GEMM on
//                the FP64 matrix cores (64 x 64 tiles, frequency in z);
//   k_lb_back    back substitution over the pivot order in blocks of NB,
//                right-looking (a block's triangle by one wave, then every
//                earlier pivot row loses the block's columns), then
//                chi_ladd and chi_para.
// The trailing block is read and written once per panel, n^3 / (3 NB)
// complex values per frequency, where the unblocked solve streamed it once
// per pivot through one compute unit.  The frequencies of a chunk run as
// two independent parts on two streams, so one part's panels (a serial
// pivot chain) overlap the other's trailing updates.
#include <algorithm>
#include <climits>
#include <cmath>
#include <string>

#include "dev_linalg.h"
#include "rpa_internal.h"

namespace rpa {

namespace {
constexpr int kLdMaxK = 3072;  // k-points of the kernel (fock is n x n); the reference's nkd is 2700
constexpr int kLdLdsK = 88;    // fock in LDS up to here (16 n^2 bytes)
constexpr int kLbParts = 2;    // blocked form: frequency parts on their own streams (4 measured slower)
}  // namespace

struct LadderArgs {
  int n, no, f0;
  double omin, dw, eta, nv, valley;
  const double2 *F;      // [k][k'] Fock kernel (row-major)
  const double *delta;   // [n]
  const double2 *cur;    // [2][n] current(pol, k)
  double2 *chi_ladd;     // [io][2]
  double2 *chi_para;     // [io][2]
  int *status;
};

// dynamic LDS: fock, then L0 [n], y [2][n], fcol [n] (double2)
inline size_t ladder_lds_bytes(int n) { return sizeof(double2) * ((size_t)n * n + 4 * (size_t)n); }

// chi_ladd(pol) = -valley j^T z / NV^2 and chi_para(pol) from the diagonal
// (z = fock^-1 y in z[2][n]), written by thread 0 of the workgroup
__device__ void ladder_close(const double2 *cur, const double *delta, const double2 *z, int n, double o, double eta,
                             double nv, double valley, int io, double2 *chi_ladd, double2 *chi_para, double *red) {
  const int tid = threadIdx.x;
  double acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) acc[e] = 0.0;
  for (int k = tid; k < n; k += 256) {
    const double2 z0 = z[k], z1 = z[n + k];
    const double2 j0 = cur[k], j1 = cur[n + k];
    const double2 t0 = c_mul(j0, z0), t1 = c_mul(j1, z1);
    acc[0] += t0.x, acc[1] += t0.y, acc[2] += t1.x, acc[3] += t1.y;
    const double2 l0 = c_div(make_double2(1.0, 0.0), make_double2(o + delta[k], eta));
    const double2 lp = c_scale(l0, o / delta[k]);  // (o / Delta) L0
    const double2 p0 = c_mul(c_mul(j0, lp), c_conj(j0)), p1 = c_mul(c_mul(j1, lp), c_conj(j1));
    acc[4] += p0.x, acc[5] += p0.y, acc[6] += p1.x, acc[7] += p1.y;
  }
  block_sum<8>(acc, red);
  if (tid == 0) {
    const double *r = red + 32;
    const double inv_nv = 1.0 / nv;
    const double cl = -valley * inv_nv * inv_nv, cp = valley * inv_nv;
    double2 *ol = chi_ladd + 2 * (size_t)(io - 1), *op = chi_para + 2 * (size_t)(io - 1);
    ol[0] = make_double2(cl * r[0], cl * r[1]);
    ol[1] = make_double2(cl * r[2], cl * r[3]);
    op[0] = make_double2(cp * r[4], cp * r[5]);
    op[1] = make_double2(cp * r[6], cp * r[7]);
  }
}

__global__ __launch_bounds__(256) void k_ladder(LadderArgs a) {
  extern __shared__ double2 lds_dyn[];
  __shared__ double red_v[4];
  __shared__ int red_i[4];
  __shared__ double red[5 * 8];
  const int n = a.n, tid = threadIdx.x;
  const int lf = blockIdx.x, io = a.f0 + lf + 1;
  double2 *sa = lds_dyn;
  double2 *tab = lds_dyn + (size_t)n * n;
  double2 *l0 = tab, *y = tab + n, *fcol = tab + 3 * n;  // y: [2][n]
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
  gauss_solve<false>(sa, n, y, 2, fcol, red_v, red_i, a.status);
  ladder_close(a.cur, a.delta, y, n, o, a.eta, a.nv, a.valley, io, a.chi_ladd, a.chi_para, red);
}

// ---------------------------------------------------------------------------
// blocked form (more than 88 k-points)

struct LbArgs {
  int n, ld, f0;
  double omin, dw, eta, nv, valley;
  const double2 *F;      // [k][k'] Fock kernel (row-major)
  const double *delta;   // [n]
  const double2 *cur;    // [2][n]
  double2 *T;            // [frequency in chunk][n][ld]: fock | y0 | y1 | 0
  int *act;              // [frequency][n] rows not yet chosen as pivots, in row order
  int *piv;              // [frequency][n] pivot row of each column
  double2 *chi_ladd;     // [io][2]
  double2 *chi_para;     // [io][2]
  int *status;
};

// fock = I - L0 F / NV and y = L0 F (L0 conj(j)): rows k0 .. k0 + 15 of one
// frequency per workgroup (blockIdx.y the frequency), one wave per row
constexpr int kLbFormRows = 16;
__global__ __launch_bounds__(256) void k_lb_form(LbArgs a) {
  extern __shared__ double2 lds_dyn[];  // L0 conj(j) [2][n]
  const int n = a.n, ld = a.ld, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lf = blockIdx.y, io = a.f0 + lf + 1, k0 = blockIdx.x * kLbFormRows;
  double2 *T = a.T + (size_t)lf * n * ld;
  double2 *lc = lds_dyn;
  const double o = a.omin + (double)(io - 1) * a.dw;
  for (int k = tid; k < n; k += 256) {
    const double2 l = c_div(make_double2(1.0, 0.0), make_double2(o + a.delta[k], a.eta));
    lc[k] = c_mul(l, c_conj(a.cur[k]));
    lc[n + k] = c_mul(l, c_conj(a.cur[n + k]));
  }
  if (blockIdx.x == 0)
    for (int k = tid; k < n; k += 256) a.act[(size_t)lf * n + k] = k;
  __syncthreads();
  const double inv_nv = 1.0 / a.nv;
  for (int k = k0 + wave; k < min(k0 + kLbFormRows, n); k += 4) {
    const double2 *row = a.F + (size_t)k * n;
    double2 *out = T + (size_t)k * ld;
    const double2 lk = c_div(make_double2(1.0, 0.0), make_double2(o + a.delta[k], a.eta));
    double2 s0 = make_double2(0.0, 0.0), s1 = make_double2(0.0, 0.0);
    for (int k2 = lane; k2 < n; k2 += 64) {
      const double2 f = row[k2];
      s0 = c_add(s0, c_mul(f, lc[k2]));
      s1 = c_add(s1, c_mul(f, lc[n + k2]));
      const double2 t = c_scale(c_mul(lk, f), inv_nv);
      out[k2] = make_double2((k == k2 ? 1.0 : 0.0) - t.x, -t.y);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      s0.x += __shfl_xor(s0.x, off);
      s0.y += __shfl_xor(s0.y, off);
      s1.x += __shfl_xor(s1.x, off);
      s1.y += __shfl_xor(s1.y, off);
    }
    if (lane == 0) {
      out[n] = c_mul(lk, s0);
      out[n + 1] = c_mul(lk, s1);
    }
    if (lane >= 2 && n + lane < ld) out[n + lane] = make_double2(0.0, 0.0);
  }
}

// One panel p0 .. p0 + w - 1 (w = min(NB, n - p0)) of every frequency: NT
// threads, thread t holds active rows act[t + NT i], i < R (n - p0 <= R NT).
// mode 0: factor and U12; 1: factor only; 2: U12 only (L11 read back from
// the pivot rows, after their trailing columns were brought up to date)
template <int NB, int NT, int R>
__global__ __launch_bounds__(NT) void k_lb_panel(LbArgs a, int p0, int mode) {
  constexpr int NW = NT / 64;
  __shared__ double2 s_prow[NB];  // the pivot row of the step
  __shared__ double2 s_inv;       // 1 / pivot
  __shared__ double s_rv[NW];     // each wave's best row: |Re| + |Im| and row index
  __shared__ int s_ri[NW];
  __shared__ double2 s_l11[NB][NB];
  __shared__ int s_piv[NB];
  __shared__ int s_cnt[R][NW];
  const int n = a.n, ld = a.ld, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lf = blockIdx.x;
  const int m = n - p0, w = min(NB, m);
  double2 *T = a.T + (size_t)lf * n * ld;
  int *act = a.act + (size_t)lf * n;
  if (mode != 2) {
    bool mine[R], live[R];
    int row[R];
    double2 v[R][NB];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      mine[i] = tid + NT * i < m;
      row[i] = mine[i] ? act[tid + NT * i] : 0;
      live[i] = mine[i];
      const double2 *src = T + (size_t)row[i] * ld + p0;
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        const double2 x = src[min(c, w - 1)];
        v[i][c] = (mine[i] && c < w) ? x : make_double2(0.0, 0.0);
      }
    }
    // Per pivot step: the waves' best rows meet in LDS (barrier), the pivot
    // row's owner publishes its values and 1 / pivot (barrier), the active
    // rows eliminate.  Fully unrolled so that v[][] stays in registers.
#pragma clang loop unroll(full)
    for (int j = 0; j < NB; ++j) {
      if (j < w) {  // uniform over the workgroup
        double best = -1.0;
        int bi = INT_MAX;
#pragma unroll
        for (int i = 0; i < R; ++i)
          if (live[i]) {
            double mm = fabs(v[i][j].x) + fabs(v[i][j].y);
            mm = (mm >= 0.0) ? mm : 0.0;  // a NaN entry competes as zero
            if (mm > best || (mm == best && row[i] < bi)) {
              best = mm;
              bi = row[i];
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
          s_rv[wave] = best;
          s_ri[wave] = bi;
        }
        __syncthreads();
        double pv = s_rv[0];
        int p = s_ri[0];
#pragma unroll
        for (int q = 1; q < NW; ++q)
          if (s_rv[q] > pv || (s_rv[q] == pv && s_ri[q] < p)) {
            pv = s_rv[q];
            p = s_ri[q];
          }
#pragma unroll
        for (int i = 0; i < R; ++i)
          if (live[i] && row[i] == p) {
#pragma unroll
            for (int c = j; c < NB; ++c) s_prow[c] = v[i][c];
#pragma unroll
            for (int c = 0; c < j; ++c) s_l11[j][c] = v[i][c];
            s_inv = c_div(make_double2(1.0, 0.0), v[i][j]);
            s_piv[j] = p;
            live[i] = false;
          }
        if (tid == 0 && !(pv > 0.0)) atomicOr(a.status, 1);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < R; ++i)
          if (live[i]) {
            const double2 l = c_mul(v[i][j], s_inv);
            v[i][j] = l;
#pragma unroll
            for (int c = j + 1; c < NB; ++c) v[i][c] = c_sub(v[i][c], c_mul(l, s_prow[c]));
          }
      }
    }
    // the panel in place: pivot rows hold their U11 row (and L11 multipliers
    // left of it), the other rows their L21 multipliers
#pragma unroll
    for (int i = 0; i < R; ++i)
      if (mine[i]) {
        double2 *dst = T + (size_t)row[i] * ld + p0;
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c < w) dst[c] = v[i][c];
      }
    // rows still active, in row-list order (slot i of every thread, then i + 1)
    unsigned long long bal[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      bal[i] = __ballot(live[i]);
      if (lane == 0) s_cnt[i][wave] = __popcll(bal[i]);
    }
    __syncthreads();  // also: every thread has read its act entries and s_l11 / s_piv are complete
    int base = 0;
#pragma unroll
    for (int i = 0; i < R; ++i) {
      int off = base + __popcll(bal[i] & ((1ull << lane) - 1ull));
      for (int q = 0; q < NW; ++q) {
        if (q < wave) off += s_cnt[i][q];
        base += s_cnt[i][q];
      }
      if (live[i]) act[off] = row[i];
    }
    if (tid < w) a.piv[(size_t)lf * n + p0 + tid] = s_piv[tid];
  } else {
    for (int x = tid; x < NB * NB; x += NT) {
      const int jj = x / NB, i = x - jj * NB;
      const double2 t = T[(size_t)a.piv[(size_t)lf * n + p0 + min(jj, w - 1)] * ld + p0 + i];
      s_l11[jj][i] = (jj < w && i < jj) ? t : make_double2(0.0, 0.0);
    }
    if (tid < NB) s_piv[tid] = a.piv[(size_t)lf * n + p0 + min(tid, w - 1)];
    __syncthreads();
  }
  if (mode == 1) return;
  // U12 = L11^-1 A12 on the pivot rows, a column per thread and pass (one
  // pass when R = 1: the host launches NT >= ld - NB threads there, and a
  // loop would let the compiler hold all of L11 in registers)
  auto solve_col = [&](int c) {
    double2 x[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) x[j] = (j < w) ? T[(size_t)s_piv[j] * ld + c] : make_double2(0.0, 0.0);
#pragma unroll
    for (int j = 1; j < NB; ++j) {
      if (j < w) {
#pragma unroll
        for (int i = 0; i < j; ++i) x[j] = c_sub(x[j], c_mul(s_l11[j][i], x[i]));
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < w) T[(size_t)s_piv[j] * ld + c] = x[j];
  };
  if constexpr (R == 1) {
    const int c = p0 + w + tid;
    if (c < ld) solve_col(c);
  } else {
    for (int c = p0 + w + tid; c < ld; c += NT) solve_col(c);
  }
}

// Update with the multipliers of columns p0 .. p0 + w - 1: for m rows r (the
// active list, or the pivot rows of columns rows_at ..) and columns c in
// cbeg .. cend - 1,
//   T[r][c] -= sum_{j<w} T[r][p0 + j] T[piv(p0 + j)][c]
// 64 x 64 complex tiles, frequency in blockIdx.z; each wave 32 x 32 as
// 2 x 2 blocks of 16 x 16, four real products per complex one.  The L and
// U operands pass through LDS in steps of 16 (34 KB, four workgroups per
// compute unit); the tile of T itself goes straight to the accumulators.
template <int NB>
__global__ __launch_bounds__(256) void k_lb_update(LbArgs a, int p0, int w, int rows_at, int m, int cbeg, int cend) {
  constexpr int KH = NB < 16 ? NB : 16;
  __shared__ double2 sL[64][KH + 1];
  __shared__ double2 sU[KH][65];
  __shared__ int s_row[64];
  __shared__ int s_pr[NB];
  const int n = a.n, ld = a.ld, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lf = blockIdx.z;
  const int c0 = cbeg + blockIdx.x * 64, r0 = blockIdx.y * 64;
  double2 *T = a.T + (size_t)lf * n * ld;
  // rows: the active list, or (rows_at >= 0) the pivot rows of columns rows_at ..
  const int *act = rows_at < 0 ? a.act + (size_t)lf * n : a.piv + (size_t)lf * n + rows_at;
  const int *piv = a.piv + (size_t)lf * n + p0;
  const int r = lane & 15, q4 = lane >> 4, wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  if (tid < 64) s_row[tid] = act[min(r0 + tid, m - 1)];
  if (tid < NB) s_pr[tid] = piv[min(tid, w - 1)];
  __syncthreads();
  rpa_d4 ar[2][2], ai[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int rr = wm + 16 * i + 4 * e + q4, cc = min(c0 + wn + 16 * j + r, cend - 1);
        const double2 cv = T[(size_t)s_row[rr] * ld + cc];
        ar[i][j][e] = cv.x;
        ai[i][j][e] = cv.y;
      }
  // the next step's L and U values are loaded into registers before this
  // step's products, so their loads are in flight while the products run
  constexpr int NU = 64 * KH / 256;
  double2 rl[NU], ru[NU];
  auto load_step = [&](int h) {
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int x = tid + 256 * u;
      const int rr = x / KH, k = x - rr * KH, kg = h * KH + k;
      rl[u] = T[(size_t)s_row[rr] * ld + p0 + min(kg, w - 1)];
      const int k2 = x >> 6, cc = x & 63, kg2 = h * KH + k2;
      ru[u] = T[(size_t)s_pr[kg2] * ld + min(c0 + cc, cend - 1)];
    }
  };
  load_step(0);
#pragma unroll
  for (int h = 0; h < NB / KH; ++h) {
    if (h > 0) __syncthreads();  // the previous step's operand reads are done
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int x = tid + 256 * u;
      const int rr = x / KH, k = x - rr * KH, kg = h * KH + k;
      sL[rr][k] = (kg < w) ? rl[u] : make_double2(0.0, 0.0);
      const int k2 = x >> 6, cc = x & 63, kg2 = h * KH + k2;
      sU[k2][cc] = (kg2 < w) ? ru[u] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (h + 1 < NB / KH) load_step(h + 1);
#pragma unroll
    for (int kk = 0; kk < KH / 4; ++kk) {
      const int k = 4 * kk + q4;
      double2 lv[2], uv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) lv[i] = sL[wm + 16 * i + r][k];
#pragma unroll
      for (int j = 0; j < 2; ++j) uv[j] = sU[k][wn + 16 * j + r];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          ar[i][j] = mfma_f64_16x16x4(-lv[i].x, uv[j].x, ar[i][j]);  // -= Lr Ur
          ar[i][j] = mfma_f64_16x16x4(lv[i].y, uv[j].y, ar[i][j]);   // += Li Ui
          ai[i][j] = mfma_f64_16x16x4(-lv[i].x, uv[j].y, ai[i][j]);  // -= Lr Ui
          ai[i][j] = mfma_f64_16x16x4(-lv[i].y, uv[j].x, ai[i][j]);  // -= Li Ur
        }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int rr = wm + 16 * i + 4 * e + q4, cc = c0 + wn + 16 * j + r;
        if (r0 + rr < m && cc < cend) T[(size_t)s_row[rr] * ld + cc] = make_double2(ar[i][j][e], ai[i][j][e]);
      }
}

// Back substitution z_j = (y(P_j) - sum_{c>j} U(P_j, c) z_c) / U(P_j, j) in
// blocks of NB pivots from the last, right-looking: the right-hand sides of
// all pivots stay in LDS, each replaced by its z once solved; once a block's
// z is solved (its triangle by one wave with shuffles), every earlier pivot
// row subtracts that block's columns (a thread per row, its NB values
// independent loads).  Then chi_ladd and chi_para.
template <int NB>
__global__ __launch_bounds__(256) void k_lb_back(LbArgs a) {
  // dynamic LDS: rhs [2][n] (position j holds z_j once its block is
  // solved: unknown j belongs to pivot step j), pivot rows [n] (int)
  extern __shared__ double2 lds_dyn[];
  __shared__ double2 s_u[NB][NB + 1];
  __shared__ double red[5 * 8];
  const int n = a.n, ld = a.ld, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lf = blockIdx.x, io = a.f0 + lf + 1;
  const double2 *T = a.T + (size_t)lf * n * ld;
  double2 *rhs = lds_dyn, *z = lds_dyn;
  int *sp = (int *)(lds_dyn + 2 * (size_t)n);
  for (int k = tid; k < n; k += 256) {
    const int pr = a.piv[(size_t)lf * n + k];
    sp[k] = pr;
    rhs[k] = T[(size_t)pr * ld + n];
    rhs[n + k] = T[(size_t)pr * ld + n + 1];
  }
  __syncthreads();
  for (int jb = ((n - 1) / NB) * NB; jb >= 0; jb -= NB) {
    const int je = min(jb + NB, n), bw = je - jb;
    for (int x = tid; x < bw * bw; x += 256) {
      const int jj = x / bw, c = x - jj * bw;
      s_u[jj][c] = T[(size_t)sp[jb + jj] * ld + jb + c];
    }
    __syncthreads();
    if (wave == 0) {  // lane = 32 rhs + row of the block
      const int rr = lane >> 5, i = lane & 31;
      double2 sv = (i < bw) ? rhs[(size_t)rr * n + jb + i] : make_double2(0.0, 0.0), zi = make_double2(0.0, 0.0);
      for (int j = bw - 1; j >= 0; --j) {
        if (i == j) zi = c_div(sv, s_u[i][i]);
        const double2 zj = make_double2(__shfl(zi.x, 32 * rr + j), __shfl(zi.y, 32 * rr + j));
        if (i < j) sv = c_sub(sv, c_mul(s_u[i][j], zj));
      }
      if (i < bw) z[(size_t)rr * n + jb + i] = zi;
    }
    __syncthreads();
    for (int j = tid; j < jb; j += 256) {
      const double2 *urow = T + (size_t)sp[j] * ld + jb;
      double2 u[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) u[c] = urow[min(c, bw - 1)];
      double2 s0 = make_double2(0.0, 0.0), s1 = make_double2(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < bw) {
          s0 = c_add(s0, c_mul(u[c], z[jb + c]));
          s1 = c_add(s1, c_mul(u[c], z[n + jb + c]));
        }
      rhs[j] = c_sub(rhs[j], s0);
      rhs[n + j] = c_sub(rhs[n + j], s1);
    }
    __syncthreads();
  }
  const double o = a.omin + (double)(io - 1) * a.dw;
  ladder_close(a.cur, a.delta, z, n, o, a.eta, a.nv, a.valley, io, a.chi_ladd, a.chi_para, red);
}

template <int NB, int NT, int R>
hipError_t ladder_blocked(hipStream_t st, LbArgs a, int nf) {
  const int n = a.n, ld = a.ld;
  // R active rows per thread; one U12 column per thread when R = 1
  if (n > R * NT || (R == 1 && ld - NB > NT) || n <= NB) return hipErrorInvalidValue;
  const size_t form_lds = sizeof(double2) * 2 * n;                    // 96 KB at 3072 k-points
  const size_t back_lds = sizeof(double2) * 2 * n + sizeof(int) * n;  // 108 KB at 3072 k-points
  hipError_t e = hipFuncSetAttribute((const void *)k_lb_form, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)form_lds);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)k_lb_back<NB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)back_lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_lb_form, dim3((n + kLbFormRows - 1) / kLbFormRows, nf), dim3(256), form_lds, st, a);
  auto update = [&](auto kern, int p0, int w, int rows_at, int m, int cbeg, int cend) {
    if (m > 0 && cend > cbeg)
      hipLaunchKernelGGL(kern, dim3((cend - cbeg + 63) / 64, (m + 63) / 64, nf), dim3(256), 0, st, a, p0, w, rows_at, m,
                         cbeg, cend);
  };
  // Panels in pairs: panel p with its U12; panel p+1's columns brought up to
  // date (narrow update); panel p+1 factored; its pivot rows brought up to
  // date on the columns right of it; its U12; then one update of the trailing
  // block with the 2 NB multiplier columns of both, which is read and
  // written once per pair instead of once per panel.
  for (int p0 = 0; p0 < n;) {
    const int w1 = std::min(NB, n - p0), q0 = p0 + w1;
    hipLaunchKernelGGL((k_lb_panel<NB, NT, R>), dim3(nf), dim3(NT), 0, st, a, p0, 0);
    if (q0 >= n) break;
    const int w2 = std::min(NB, n - q0), e0 = q0 + w2;
    update(k_lb_update<NB>, p0, w1, -1, n - q0, q0, e0);
    hipLaunchKernelGGL((k_lb_panel<NB, NT, R>), dim3(nf), dim3(NT), 0, st, a, q0, 1);
    update(k_lb_update<NB>, p0, w1, q0, w2, e0, ld);
    hipLaunchKernelGGL((k_lb_panel<NB, NT, R>), dim3(nf), dim3(NT), 0, st, a, q0, 2);
    update(k_lb_update<2 * NB>, p0, w1 + w2, -1, n - e0, e0, ld);
    p0 = e0;
  }
  hipLaunchKernelGGL(k_lb_back<NB>, dim3(nf), dim3(256), back_lds, st, a);
  return hipGetLastError();
}

// Fock kernel of Chi_ladd (Chi_ladd_up/BSE.f90:804-816): one wave per pair
// (k, k') of the kernel's k-points, fockk(k,k') = -sum_{G,G'} W(iq,G,G')
// conj(M1(G)) M2(G') with iq = int(Gcar |k' - k|_xy / qmin) + 1, and on the
// diagonal the Q = 0 term -Vcell Ntot ln(1 + alpha epsq) / alpha.  Up to 64
// local fields the lanes walk the W block contiguously over (G, G') and add
// conj(M1(G)) W M2(G') (M1, M2 one value per lane, shared by shuffles);
// beyond, lane j takes columns G' = j, j + 64, ...: t = sum_G conj(M1(G))
// W(iq,G,G'), then t M2(G').  The wave sums the lanes.  Pairs whose iq
// falls outside 2..nq get 0 (the reference reads W only for those and leaves
// such entries unset).
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
      if (nlf <= 64) {
        // one load per lane for M1 and M2 (lane = G), the W block read
        // contiguously over (G, G') with M1(G), M2(G') taken from their lanes
        // by shuffles: 2 + nlf^2/64 load instructions per pair instead of
        // 2 nlf, the address unit being what bounds this kernel
        const double2 m1 = lane < nlf ? m[2 * lane] : make_double2(0.0, 0.0);
        const double2 m2 = lane < nlf ? m[2 * lane + 1] : make_double2(0.0, 0.0);
        const int nn = nlf * nlf;
        for (int e0 = 0; e0 < nn; e0 += 64) {  // same trip count on every lane (shuffles)
          const int e = e0 + lane, ec = min(e, nn - 1), g = ec / nlf, j = ec - g * nlf;
          const double2 wv = w[ec];
          const double2 a1 = make_double2(__shfl(m1.x, g), __shfl(m1.y, g));
          const double2 a2 = make_double2(__shfl(m2.x, j), __shfl(m2.y, j));
          if (e < nn) acc = c_add(acc, c_mul(c_mul(c_conj(a1), wv), a2));
        }
      } else {
        for (int j = lane; j < nlf; j += 64) {
          double2 t = make_double2(0.0, 0.0);
          for (int g = 0; g < nlf; ++g) t = c_add(t, c_mul(c_conj(m[2 * g]), w[(size_t)g * nlf + j]));
          acc = c_add(acc, c_mul(t, m[2 * j + 1]));
        }
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
  const bool blocked = n > kLdLdsK;
  const size_t nn = (size_t)n * n;
  // blocked form: fock | y per frequency in rows of ld (a multiple of 4)
  // complex values, at most 4 GiB of slices per chunk of frequencies
  const int ld = (n + 2 + 3) & ~3;
  const size_t slice = (size_t)n * ld;
  // The blocked solve runs as kLbParts independent parts of each chunk's
  // frequencies on as many streams: some parts' panels (a serial pivot
  // chain, one workgroup per compute unit) overlap the others' trailing
  // updates (HBM-bound).  Each part keeps its own region of the slices.
  const int chunk = blocked ? (int)std::min<size_t>(
                                  std::min<size_t>(no, 65535),
                                  std::max<size_t>(1, ((size_t)4 << 30) / (sizeof(double2) * slice)))
                            : no;
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
  const int part = (chunk + kLbParts - 1) / kLbParts;  // region per stream, in slices
  hipStream_t st = nullptr, sq[kLbParts] = {};  // sq[0] is st
  hipEvent_t ev_in = nullptr, ev_out[kLbParts] = {};
  double2 *dF = nullptr, *dc = nullptr, *dT = nullptr, *dl = nullptr, *dp = nullptr;
  double *dd = nullptr;
  int *ds = nullptr, *dact = nullptr, *dpiv = nullptr;
  auto cleanup = [&]() {
    if (st) (void)hipStreamSynchronize(st);
    for (int i = 1; i < kLbParts; ++i)
      if (sq[i]) (void)hipStreamSynchronize(sq[i]);
    for (void *q : {(void *)dF, (void *)dc, (void *)dT, (void *)dl, (void *)dp, (void *)dd, (void *)ds, (void *)dact,
                    (void *)dpiv})
      if (q) (void)hipFree(q);
    if (ev_in) (void)hipEventDestroy(ev_in);
    for (int i = 0; i < kLbParts; ++i)
      if (ev_out[i]) (void)hipEventDestroy(ev_out[i]);
    if (st) (void)hipStreamDestroy(st);
    for (int i = 1; i < kLbParts; ++i)
      if (sq[i]) (void)hipStreamDestroy(sq[i]);
  };
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess ||
      (blocked && (e = hipEventCreateWithFlags(&ev_in, hipEventDisableTiming)) != hipSuccess) ||
      (e = hipMalloc((void **)&dF, sizeof(double2) * nn)) != hipSuccess ||
      (e = hipMalloc((void **)&dc, sizeof(double2) * 2 * n)) != hipSuccess ||
      (e = hipMalloc((void **)&dd, sizeof(double) * n)) != hipSuccess ||
      (e = hipMalloc((void **)&dl, sizeof(double2) * 2 * (size_t)no)) != hipSuccess ||
      (e = hipMalloc((void **)&dp, sizeof(double2) * 2 * (size_t)no)) != hipSuccess ||
      (e = hipMalloc((void **)&ds, sizeof(int))) != hipSuccess ||
      (blocked && (e = hipMalloc((void **)&dT, sizeof(double2) * slice * kLbParts * part)) != hipSuccess) ||
      (blocked && (e = hipMalloc((void **)&dact, sizeof(int) * (size_t)n * kLbParts * part)) != hipSuccess) ||
      (blocked && (e = hipMalloc((void **)&dpiv, sizeof(int) * (size_t)n * kLbParts * part)) != hipSuccess) ||
      (e = hipMemcpyAsync(dF, hF.data(), sizeof(double) * hF.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(dc, hc.data(), sizeof(double) * hc.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(dd, delta, sizeof(double) * n, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemsetAsync(ds, 0, sizeof(int), st)) != hipSuccess ||
      (!blocked && (e = hipFuncSetAttribute((const void *)k_ladder, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)ladder_lds_bytes(n))) != hipSuccess)) {
    cleanup();
    return err(RPA_E_HIP, hipGetErrorString(e));
  }
  // the other streams start after the inputs are up
  if (blocked) {
    sq[0] = st;
    e = hipEventRecord(ev_in, st);
    for (int i = 1; i < kLbParts && e == hipSuccess; ++i)
      if ((e = hipStreamCreateWithFlags(&sq[i], hipStreamNonBlocking)) == hipSuccess &&
          (e = hipEventCreateWithFlags(&ev_out[i], hipEventDisableTiming)) == hipSuccess)
        e = hipStreamWaitEvent(sq[i], ev_in, 0);
    if (e != hipSuccess) {
      cleanup();
      return err(RPA_E_HIP, hipGetErrorString(e));
    }
  }
  const double nv = (double)p->ntot * p->vcell;
  for (int f0 = 0; f0 < no; f0 += chunk) {
    const int nf = std::min(chunk, no - f0);
    if (blocked) {
      LbArgs b{};
      b.n = n;
      b.ld = ld;
      b.f0 = f0;
      b.omin = p->omin;
      b.dw = p->domega;
      b.eta = p->eta;
      b.nv = nv;
      b.valley = p->valley;
      b.F = dF;
      b.delta = dd;
      b.cur = dc;
      b.T = dT;
      b.act = dact;
      b.piv = dpiv;
      b.chi_ladd = dl;
      b.chi_para = dp;
      b.status = ds;
      // 32-column panels with 512 threads; past 512 k-points 16 columns with
      // 1024 threads (one active row per thread); past 1024 8 columns with
      // 512 threads of four or six active rows (256 registers a lane there).  Part i (frequencies
      // f0 + i q .. f0 + (i + 1) q - 1, q = ceil(nf / kLbParts) <= part) on
      // stream i in region [i part, (i + 1) part).
      const int q = (nf + kLbParts - 1) / kLbParts;
      for (int i = 0; i < kLbParts && e == hipSuccess; ++i) {
        const int lo = i * q, cnt = std::min(nf, lo + q) - lo;
        if (cnt <= 0) break;
        LbArgs bi = b;
        bi.f0 = f0 + lo;
        bi.T = dT + slice * part * i;
        bi.act = dact + (size_t)n * part * i;
        bi.piv = dpiv + (size_t)n * part * i;
        e = n <= 512    ? ladder_blocked<32, 512, 1>(sq[i], bi, cnt)
            : n <= 1024 ? ladder_blocked<16, 1024, 1>(sq[i], bi, cnt)
            : n <= 2048 ? ladder_blocked<8, 512, 4>(sq[i], bi, cnt)
                        : ladder_blocked<8, 512, 6>(sq[i], bi, cnt);
      }
    } else {
      LadderArgs a{};
      a.n = n;
      a.no = no;
      a.f0 = f0;
      a.omin = p->omin;
      a.dw = p->domega;
      a.eta = p->eta;
      a.nv = nv;
      a.valley = p->valley;
      a.F = dF;
      a.delta = dd;
      a.cur = dc;
      a.chi_ladd = dl;
      a.chi_para = dp;
      a.status = ds;
      hipLaunchKernelGGL(k_ladder, dim3(nf), dim3(256), ladder_lds_bytes(n), st, a);
      e = hipGetLastError();
    }
    if (e != hipSuccess) {
      cleanup();
      return err(RPA_E_HIP, std::string("ladder launch: ") + hipGetErrorString(e));
    }
  }
  if (blocked) {
    for (int i = 1; i < kLbParts && e == hipSuccess; ++i)
      if ((e = hipEventRecord(ev_out[i], sq[i])) == hipSuccess) e = hipStreamWaitEvent(st, ev_out[i], 0);
    if (e != hipSuccess) {
      cleanup();
      return err(RPA_E_HIP, hipGetErrorString(e));
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
