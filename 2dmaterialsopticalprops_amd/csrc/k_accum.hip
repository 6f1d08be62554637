// Spectral-function accumulation S0(w,G,G') = sum_t W(w,t) rho_t(G) conj(rho'_t(G'))
// (Sloss.f90:407-430 real part; Optabs.f90:386-461 complex) as a banded
// grouped GEMM on the FP64 matrix cores.
//
//  1. k_transitions   per transition: energies, occupation filter, and the
//                     first frequency tile ("home") its Lorentzian window
//                     reaches (-1 if none; intraband Qeff pairs get the
//                     pseudo tile ntiles).
//  2. k_tile_hist / k_tile_scan / k_tile_scatter
//                     stable counting sort by home tile -> perm[t] (natural ->
//                     sorted position).  The vertex kernel writes its rows at
//                     perm[t], so no data moves for the sort.
//  3. k_worklist      per tile b the candidates are the contiguous sorted
//                     range of homes [b-span, b]; cut into chunks of kChunkT.
//  4. k_weight_table  W of every chunk (32 frequencies x its transitions) from
//                     the energies with the reference's exact arithmetic and
//                     cutoff test, once per run.
//     k_accum         persistent kernel over (chunk, column block): a 32 x 256
//                     tile of S0 += W (32 x K) P (K x 256), W copied from the
//                     table, P from the vertices staged in LDS.  Partials per
//                     chunk.
//  5. k_reduce        fixed-order Neumaier sum of the chunk partials of each
//                     tile -> S0 (deterministic, FP64).
//  6. k_qeff          Optabs intraband Qeff (Optabs.f90:430-442).
//
// Mixed precision replaces step 4 by k_accum_mx: the FP64 weights times the
// FP32 vertices, split into BF16 hi + lo parts, on the BF16 matrix cores as
// hi*hi + hi*lo + lo*hi with FP32 accumulation over a chunk of kChunkTMx
// transitions; step 5 still sums the chunks in FP64 with compensation.
#include "rpa_internal.h"

#include <algorithm>
#include <cmath>

namespace rpa {

constexpr int kKS = 16;          // transitions per LDS stage
constexpr int kMaxTiles = 1024;

struct TransArgs {
  const int *k1, *k2;
  const double *E;
  int nki, nband;
  PairTable pairs;
  long long ntrans;
  int program;
  double domega, hw;
  int shift;  // frequency index offset (Optabs: no)
  int nw, ntiles;
  double efermi, temp, df_cut;
  int *key;
};

__device__ inline void trans_decode(const PairTable &pt, long long t, int &k, int &n, int &m) {
  k = (int)(t / pt.tpk);
  int r = (int)(t - (long long)k * pt.tpk);
  int pb = 0;
#pragma unroll
  for (int b = 1; b < kMaxPairBlocks; ++b)
    if (b < pt.nblocks && r >= pt.off[b]) pb = b;
  const PairBlock &B = pt.b[pb];
  int loc = r - pt.off[pb];
  int nm = B.m_hi - B.m_lo;
  n = B.n_lo + loc / nm;
  m = B.m_lo + loc % nm;
}

// Optabs genOccupation (Optabs.f90:935-951)
__device__ inline void occupation(double ei, double ef, double T, double &expo, double &f) {
#pragma clang fp contract(off)
  double x = (ei - ef) / T;
  if (x < -20.0) {
    expo = 0.0;
    f = 1.0;
  } else if (x > 20.0) {
    expo = x;
    f = 0.0;
  } else {
    expo = exp(x);
    f = 1.0 / (expo + 1.0);
  }
}

__global__ void k_transitions(TransArgs a) {
#pragma clang fp contract(off)
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntrans) return;
  int k, n, m;
  trans_decode(a.pairs, t, k, n, m);
  const int K1 = a.k1[k] - 1, K2 = a.k2[k] - 1;
  const double e1 = a.E[K1 + (size_t)a.nki * n];
  const double e2 = a.E[K2 + (size_t)a.nki * m];
  bool inter = true, intra = false;
  if (a.program == RPA_OPTABS) {
    double x1, f1, x2, f2;
    occupation(e1, a.efermi, a.temp, x1, f1);
    occupation(e2, a.efermi, a.temp, x2, f2);
    double df = f1 - f2;
    inter = (n != m) && (fabs(df) >= a.df_cut);
    intra = (n == m) && (fabs(e1 - a.efermi) <= 10.0 * a.temp);
  }
  int key = -1;
  if (inter) {
    int wlo, whi;
    if (isinf(a.hw)) {
      wlo = 0;
      whi = a.nw - 1;
    } else {
      double res = e2 - e1;
      double lo = ceil((res - a.hw) / a.domega) - 2.0 + a.shift;
      double hi = floor((res + a.hw) / a.domega) + 2.0 + a.shift;
      lo = fmax(lo, 0.0);
      hi = fmin(hi, (double)(a.nw - 1));
      wlo = (int)lo;
      whi = (int)hi;
      if (!(lo <= hi)) wlo = 1, whi = 0;
    }
    if (wlo <= whi) key = wlo / kTileW;
  } else if (intra) {
    key = a.ntiles;
  }
  a.key[t] = key;
}

__global__ __launch_bounds__(256) void k_tile_hist(const int *key, long long ntrans, int nkeys,
                                                    int nblk, int *bcnt) {
  __shared__ int h[kMaxTiles + 1];
  for (int i = threadIdx.x; i < nkeys; i += blockDim.x) h[i] = 0;
  __syncthreads();
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntrans) {
    int k = key[t];
    if (k >= 0) atomicAdd(&h[k], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nkeys; i += blockDim.x) bcnt[(size_t)i * nblk + blockIdx.x] = h[i];
}

// Exclusive scan of bcnt (key-major) in place by one workgroup; toff[key] =
// first sorted position of key, toff[nkeys] = total.  Tiles of 4096
// consecutive counts, four per thread (coalesced loads), a shuffle scan of
// the thread totals per wave and the 16 wave totals through LDS; the running
// total of earlier tiles is carried.  The next tile's counts are loaded
// while the current one is scanned.  (One count per thread: 0.040 ms for the
// 35k counts of a 393k-transition run; four per thread, each tile's loads
// waited on before its scan: 0.025 ms.)
__global__ __launch_bounds__(1024) void k_tile_scan(int *bcnt, long long n, int nkeys, int nblk,
                                                     int *toff) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int carry = 0;  // total of the earlier tiles
  int vn[4];
  auto load = [&](long long base) {  // clamped: no conditional load
#pragma unroll
    for (int q = 0; q < 4; ++q) vn[q] = bcnt[min(base + 4 * tid + q, n - 1)];
  };
  load(0);
  for (long long base = 0; base < n; base += 4096) {
    const long long i0 = base + 4 * tid;
    int v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = (i0 + q < n) ? vn[q] : 0;
    load(base + 4096);  // past the end: clamped loads, never used
    const int ts = (v[0] + v[1]) + (v[2] + v[3]);
    int x = ts;  // inclusive scan of the thread totals within the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int wpre = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      const int t = wsum[w];
      wpre += (w < wv) ? t : 0;
      total += t;
    }
    int e = carry + wpre + x - ts;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long i = i0 + q;
      if (i < n) {
        bcnt[i] = e;
        if (i % nblk == 0) toff[i / nblk] = e;  // first block of key i / nblk
      }
      e += v[q];
    }
    carry += total;
    __syncthreads();  // wsum is rewritten by the next tile
  }
  if (tid == 0) toff[nkeys] = carry;
}

struct ScatterArgs {
  const int *key;
  const int *bcnt;
  long long ntrans;
  int nblk;
  const int *k1, *k2;
  const double *E;
  int nki;
  PairTable pairs;
  int program;
  int ntiles;
  double efermi, temp;
  int *perm, *nat;
  double *te;
};

__global__ __launch_bounds__(256) void k_tile_scatter(ScatterArgs a) {
#pragma clang fp contract(off)
  __shared__ int keys[256];
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int key = (t < a.ntrans) ? a.key[t] : -1;
  keys[threadIdx.x] = key;
  __syncthreads();
  if (t >= a.ntrans) return;
  if (key < 0) {
    a.perm[t] = -1;
    return;
  }
  int rank = 0;
  for (int j = 0; j < (int)threadIdx.x; ++j) rank += (keys[j] == key);
  int pos = a.bcnt[(size_t)key * a.nblk + blockIdx.x] + rank;
  a.perm[t] = pos;
  a.nat[pos] = (int)t;
  int k, n, m;
  trans_decode(a.pairs, t, k, n, m);
  const int K1 = a.k1[k] - 1, K2 = a.k2[k] - 1;
  const double e1 = a.E[K1 + (size_t)a.nki * n];
  const double e2 = a.E[K2 + (size_t)a.nki * m];
  double z = 1.0;
  if (a.program == RPA_OPTABS) {
    double x1, f1, x2, f2;
    occupation(e1, a.efermi, a.temp, x1, f1);
    occupation(e2, a.efermi, a.temp, x2, f2);
    if (key == a.ntiles) {
      // intraband: fact = -expo/((expo+1)^2 T)  (Optabs.f90:432-433)
      double fact = x1 / ((x1 + 1.0) * (x1 + 1.0));
      z = -fact / a.temp;
    } else {
      z = f1 - f2;
    }
  }
  a.te[3 * (size_t)pos + 0] = e1;
  a.te[3 * (size_t)pos + 1] = e2;
  a.te[3 * (size_t)pos + 2] = z;
}

__global__ __launch_bounds__(1024) void k_worklist(const int *toff, int ntiles, int span, int kc,
                                                    int *items, int *first_item, int *nitems) {
  __shared__ int cnt[kMaxTiles + 1];
  for (int b = threadIdx.x; b < ntiles; b += blockDim.x) {
    int lo = toff[b - span > 0 ? b - span : 0], hi = toff[b + 1];
    cnt[b] = hi > lo ? (hi - lo + kc - 1) / kc : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int b = 0; b < ntiles; ++b) {
      int v = cnt[b];
      cnt[b] = run;
      run += v;
    }
    cnt[ntiles] = run;
    nitems[0] = run;
  }
  __syncthreads();
  for (int b = threadIdx.x; b <= ntiles; b += blockDim.x) first_item[b] = cnt[b];
  for (int b = threadIdx.x; b < ntiles; b += blockDim.x) {
    int lo = toff[b - span > 0 ? b - span : 0], hi = toff[b + 1];
    int it = cnt[b];
    for (int s = lo; s < hi; s += kc, ++it) {
      items[3 * it + 0] = b;
      items[3 * it + 1] = s;
      items[3 * it + 2] = s + kc < hi ? s + kc : hi;
    }
  }
}

struct AccArgs {
  const double *te;
  const double *wt;  // weight table [item][32][kChunkT]
  const double2 *rho1, *rho2;
  const int *items;
  const int *nitems;
  int npb;  // column blocks of 256
  int nlf, nw, ncol, ncol_pad;
  int program;
  int shift;
  double domega, eta, cut, den;
  double *part;
};

// Weight of transition (e1, e2, z) at frequency row w, with the reference's
// exact arithmetic and cutoff test (Sloss.f90:408-411, Optabs.f90:391-402).
__device__ inline double weight(int program, int w, int shift, double domega, double eta,
                                double cut, double den, double e1, double e2, double z) {
#pragma clang fp contract(off)
  double o = (double)(w - shift) * domega;
  double dE = (o + e1) - e2;
  if (program == RPA_SLOSS) {
    double lor = -eta / (dE * dE + eta * eta);
    if (!(fabs(lor) >= cut)) return 0.0;
    return -(2.0 * lor) / den;
  }
  double lor = eta / (dE * dE + eta * eta);
  if (!(fabs(lor) >= cut)) return 0.0;
  return -((2.0 * z) * lor) / den;
}

// W of chunk `item` for its 32 frequency rows: wt[(item*32 + row)*CHUNK +
// t - t_lo], zero for frequencies past nw and for t in [t_hi, t_lo +
// round16(t_hi - t_lo)).  Computed once per run so the workgroups that share
// a chunk copy it instead of each evaluating the Lorentzians.  Workgroup =
// (chunk, 256 transitions); a thread reads its transition's energies once
// and writes its column of all 32 rows (one row per workgroup re-read them
// 32 times: 0.042 ms for the 12.6M weights of a 393k-transition run).
template <class T, int CHUNK>
__global__ __launch_bounds__(256) void k_weight_table(const double *te, const int *items, const int *nitems,
                                                      int program, int shift, double domega, double eta,
                                                      double cut, double den, int nw, T *wt) {
  constexpr int NX = CHUNK / 256;
  const int item = blockIdx.x / NX, x = (blockIdx.x - item * NX) * 256 + threadIdx.x;
  if (item >= nitems[0]) return;
  const int b = items[3 * item], t_lo = items[3 * item + 1], t_hi = items[3 * item + 2];
  const int len = (t_hi - t_lo + 15) / 16 * 16;
  if (x >= len) return;
  const int t = t_lo + x;
  const bool tok = t < t_hi;
  const double *e = te + 3 * (size_t)(tok ? t : t_lo);
  const double e1 = e[0], e2 = e[1], z = e[2];
  T *dst = wt + (size_t)item * kTileW * CHUNK + x;
#pragma unroll 4
  for (int row = 0; row < kTileW; ++row) {
    const int w = b * kTileW + row;
    double v = 0.0;
    if (tok && w < nw) v = weight(program, w, shift, domega, eta, cut, den, e1, e2, z);
    dst[(size_t)row * CHUNK] = (T)v;
  }
}

// RU: vertex values staged per thread and stage (16 nlf <= 256 RU); the next
// stage's vertices and weights are loaded into registers during the products.
template <bool CPLX, int RU>
__global__ __launch_bounds__(256) void k_accum(AccArgs a) {
  extern __shared__ double smem[];
  double *sW = smem;                                   // [32][kKS+1]
  double2 *sR1 = (double2 *)(smem + kTileW * (kKS + 1));  // [kKS][nlf]
  double2 *sR2 = sR1 + kKS * a.nlf;                    // [kKS][nlf] (CPLX)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, q4 = lane >> 4;
  // XCD-aware: workgroup b runs on XCD b % 8 and takes the items with
  // item % 8 == b % 8, so every column block of one chunk reads the chunk's
  // vertices from the same L2 (the grid is a multiple of 8)
  const int xcd = blockIdx.x & 7, nsl = gridDim.x >> 3;
  const int nix = (a.nitems[0] - xcd + 7) / 8;
  const int total = nix * a.npb;

  for (int wk = blockIdx.x >> 3; wk < total; wk += nsl) {
    const int il = wk / a.npb, pbk = wk - il * a.npb;
    const int item = 8 * il + xcd;
    const int t_lo = a.items[3 * item + 1], t_hi = a.items[3 * item + 2];
    const double *wsrc = a.wt + (size_t)item * kTileW * kChunkT;

    // columns owned by this lane in the four 16-wide blocks of its wave
    int cg_i[4], cg_j[4], cg_p[4];
    bool cg_ok[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      int col = pbk * kColsWG + wave * 64 + cb * 16 + r;
      int pair = CPLX ? (col >> 1) : col;
      cg_p[cb] = CPLX ? (col & 1) : 0;
      cg_i[cb] = pair / a.nlf;
      cg_j[cb] = pair - cg_i[cb] * a.nlf;
      cg_ok[cb] = col < a.ncol;
      if (!cg_ok[cb]) cg_i[cb] = cg_j[cb] = 0;
    }
    rpa_d4 acc[2][4];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = rpa_d4{0.0, 0.0, 0.0, 0.0};

    double rw[2];
    double2 rx[RU], ry[RU];
    const int nst = kKS * a.nlf;
    auto load = [&](int t0) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int x = tid + 256 * u, row = x / kKS, kt = x - row * kKS;
        rw[u] = wsrc[(size_t)row * kChunkT + (t0 - t_lo) + kt];
      }
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        const int x = tid + 256 * u;
        const bool ok = x < nst && t0 + x / a.nlf < t_hi;
        rx[u] = ok ? a.rho1[(size_t)t0 * a.nlf + x] : make_double2(0.0, 0.0);
        if (CPLX) ry[u] = ok ? a.rho2[(size_t)t0 * a.nlf + x] : make_double2(0.0, 0.0);
      }
    };
    auto store = [&]() {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int x = tid + 256 * u, row = x / kKS, kt = x - row * kKS;
        sW[row * (kKS + 1) + kt] = rw[u];
      }
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        const int x = tid + 256 * u;
        if (x < nst) {
          sR1[x] = rx[u];
          if (CPLX) sR2[x] = ry[u];
        }
      }
    };
    load(t_lo);
    for (int t0 = t_lo; t0 < t_hi; t0 += kKS) {
      __syncthreads();  // the previous stage's reads are done
      store();
      __syncthreads();
      if (t0 + kKS < t_hi) load(t0 + kKS);
#pragma unroll
      for (int st = 0; st < kKS / 4; ++st) {
        const int kt = st * 4 + q4;
        double av[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) av[rb] = sW[(rb * 16 + r) * (kKS + 1) + kt];
        const double2 *r1 = sR1 + kt * a.nlf;
        const double2 *r2 = (CPLX ? sR2 : sR1) + kt * a.nlf;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
          double2 x = r1[cg_i[cb]];
          double2 y = r2[cg_j[cb]];
          double bv;
          if (!CPLX) {
            bv = x.x * y.x + x.y * y.y;  // Re(x conj y)
          } else {
            bv = cg_p[cb] == 0 ? (x.x * y.x + x.y * y.y) : (x.y * y.x - x.x * y.y);
          }
          bv = cg_ok[cb] ? bv : 0.0;
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) acc[rb][cb] = mfma_f64_16x16x4(av[rb], bv, acc[rb][cb]);
        }
      }
    }
    double *dst = a.part + ((size_t)item * a.npb + pbk) * (kTileW * kColsWG);
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          int row = rb * 16 + 4 * e + q4;
          int col = wave * 64 + cb * 16 + r;
          dst[row * kColsWG + col] = acc[rb][cb][e];
        }
  }
}


// ---------------------------------------------------------------------------
// Mixed precision: S0 tile (32 frequencies x all G') of one local field G per
// GEMM,  S0(w, G, G') += sum_(t,c) A_G[w][(t,c)] B[(t,c)][G'],
//   A_G = W(w,t) x_t(G)_c       (x = rho1, c = Re/Im; W from FP64)
//   B   = y_t(G')_c             (Re S0: y = rho2;  Im S0 (Optabs):
//                                B' = (-Im y, Re y) for c = (Re, Im))
// with K = 2 x kMxKS = 32 per MFMA, each operand split as hi + lo BF16.
// Two passes before the products, so that the workgroups sharing a chunk do
// not each redo them: k_weight_table tabulates W of every chunk as FP32, and
// k_mx_split writes the BF16 hi/lo parts of y transposed (local field major,
// Re/Im of one transition in one 32-bit word) so a stage of B is a row copy.
constexpr int kMxKS = 16;   // transitions per stage (K = 32)
constexpr int kMxRow = 40;  // bf16 per staged B row (32 + pad, 80 bytes)

__device__ inline void split_bf16(float v, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

__device__ inline uint32_t bf16_pair(__bf16 re, __bf16 im) {
  return (uint32_t)__builtin_bit_cast(uint16_t, re) | ((uint32_t)__builtin_bit_cast(uint16_t, im) << 16);
}

// y[t][j] (complex FP32) -> rbt[hl][j][t] = (Re, Im) of the hi (hl = 0) or
// lo (hl = 1) BF16 part, through an LDS tile of 64 transitions.
__global__ __launch_bounds__(256) void k_mx_split(const float2 *y, long long nt, int nlf, uint32_t *rbt) {
  extern __shared__ float2 ytile[];  // [64][nlf]
  const long long t0 = (long long)blockIdx.x * 64;
  const int nrow = (int)min(64LL, nt - t0);
  const float2 *src = y + t0 * nlf;
  // eight loads per thread in flight, from clamped addresses (a load per
  // iteration followed by its LDS store waited a full latency each time)
  const int n = nrow * nlf;
  for (int x0 = 0; x0 < n; x0 += 256 * 8) {
    float2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[min(x0 + (int)threadIdx.x + 256 * u, n - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (x0 + (int)threadIdx.x + 256 * u < n) ytile[x0 + threadIdx.x + 256 * u] = v[u];
  }
  __syncthreads();
  for (int x = threadIdx.x; x < 64 * nlf; x += 256) {
    const int j = x >> 6, tt = x & 63;
    if (tt >= nrow) continue;
    const float2 v = ytile[tt * nlf + j];
    __bf16 rh, rl, ih, il;
    split_bf16(v.x, rh, rl);
    split_bf16(v.y, ih, il);
    const size_t o = (size_t)j * nt + t0 + tt;
    rbt[o] = bf16_pair(rh, ih);
    rbt[(size_t)nlf * nt + o] = bf16_pair(rl, il);
  }
}

struct AccMxArgs {
  const float *wmx;
  const float2 *rho1;
  const uint32_t *rbt;  // [2][nlf][nt]
  long long nt;
  const int *items;
  const int *nitems;
  int nlf, nw, ncol, ncol_pad, npb, nib;
  float *part;  // FP32 chunk partials (the FP32 accumulators as they are)
};

// IPW local fields G per wave (4*IPW per workgroup), NCB blocks of 16 G'.
// The next stage's W, x and B words are loaded into registers while the
// current stage runs on the matrix cores.
template <bool CPLX, int NCB, int IPW>
__global__ __launch_bounds__(256, (NCB <= 4) ? 2 : 1) void k_accum_mx(AccMxArgs a) {
  constexpr int NB = CPLX ? 2 : 1;  // B operands (Re, Im)
  constexpr int NJ = 16 * NCB;
  constexpr int NI = 4 * IPW;       // local fields of the workgroup
  constexpr int NBV = 2 * NCB;      // staged B words per thread: 2 x NJ x 16 / 256
  __shared__ float sW[kTileW][kMxKS];
  __shared__ float2 sX[kMxKS][NI];
  __shared__ __align__(16) __bf16 sB[NB][2][NJ][kMxRow];  // [re/im][hi/lo][G'][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, kg = lane >> 4;
  // XCD-aware as in k_accum: all local-field blocks of one chunk on one XCD
  const int xcd = blockIdx.x & 7, nsl = gridDim.x >> 3;
  const int nix = (a.nitems[0] - xcd + 7) / 8;
  const int total = nix * a.nib;

  for (int wk = blockIdx.x >> 3; wk < total; wk += nsl) {
    const int il = wk / a.nib, ib = wk - il * a.nib;
    const int item = 8 * il + xcd;
    const int t_lo = a.items[3 * item + 1], t_hi = a.items[3 * item + 2];
    const int ig0 = ib * NI;                 // first local field of the workgroup
    const int i0 = wave * IPW;               // first local field of the wave (local)
    const float *wsrc = a.wmx + (size_t)item * kTileW * kChunkTMx;

    // staging in registers: W (2 per thread), x (tid < 16 NI), B words
    float rw[2];
    float2 rx;
    uint32_t rb[NBV];
    auto load = [&](int t0) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = tid + 256 * u, row = e >> 4, kt = e & 15;
        rw[u] = wsrc[(size_t)row * kChunkTMx + (t0 - t_lo) + kt];
      }
      if (tid < kMxKS * NI) {
        const int kt = tid / NI, ii = tid - kt * NI;
        const int t = t0 + kt, g = ig0 + ii;
        rx = (t < t_hi && g < a.nlf) ? a.rho1[(size_t)t * a.nlf + g] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < NBV; ++u) {
        const int e = tid + 256 * u, kt = e & 15, jh = e >> 4;
        const int hl = jh / NJ, j = jh - hl * NJ;
        const int t = t0 + kt;
        rb[u] = (t < t_hi && j < a.nlf) ? a.rbt[((size_t)hl * a.nlf + j) * a.nt + t] : 0u;
      }
    };
    auto store = [&]() {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = tid + 256 * u;
        sW[e >> 4][e & 15] = rw[u];
      }
      if (tid < kMxKS * NI) sX[tid / NI][tid % NI] = rx;
#pragma unroll
      for (int u = 0; u < NBV; ++u) {
        const int e = tid + 256 * u, kt = e & 15, jh = e >> 4;
        const int hl = jh / NJ, j = jh - hl * NJ;
        *(uint32_t *)&sB[0][hl][j][2 * kt] = rb[u];
        if (CPLX)  // (-Im, Re): swap the halves, flip the sign of the new low half
          *(uint32_t *)&sB[NB - 1][hl][j][2 * kt] = ((rb[u] & 0xffffu) << 16) | ((rb[u] >> 16) ^ 0x8000u);
      }
    };

    rpa_f4 acc[NB][IPW][2][NCB];
#pragma unroll
    for (int s = 0; s < NB; ++s)
#pragma unroll
      for (int q = 0; q < IPW; ++q)
#pragma unroll
        for (int rb2 = 0; rb2 < 2; ++rb2)
#pragma unroll
          for (int cb = 0; cb < NCB; ++cb) acc[s][q][rb2][cb] = rpa_f4{0.f, 0.f, 0.f, 0.f};

    load(t_lo);
    for (int t0 = t_lo; t0 < t_hi; t0 += kMxKS) {
      __syncthreads();  // the previous stage's reads are done
      store();
      __syncthreads();
      if (t0 + kMxKS < t_hi) load(t0 + kMxKS);
      // B fragments: column G' = 16 cb + r, K slice 8 kg .. 8 kg + 7
      rpa_bf8 bh[NB][NCB], bl[NB][NCB];
#pragma unroll
      for (int s = 0; s < NB; ++s)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
          bh[s][cb] = *(const rpa_bf8 *)&sB[s][0][16 * cb + r][8 * kg];
          bl[s][cb] = *(const rpa_bf8 *)&sB[s][1][16 * cb + r][8 * kg];
        }
      if constexpr (!CPLX) {
        // real S0: the A operands of all (local field, row block) pairs
        // first, then the products, one run of independent matrix-core work
        // per stage (3% faster at N_G = 51; the complex kernel's second B
        // operand leaves no registers for it)
        rpa_bf8 ahs[IPW][2], als[IPW][2];
#pragma unroll
        for (int q = 0; q < IPW; ++q) {
          float2 xv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) xv[u] = sX[4 * kg + u][i0 + q];
#pragma unroll
          for (int rb2 = 0; rb2 < 2; ++rb2) {
            const float *wrow = &sW[16 * rb2 + r][4 * kg];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
              const float w = wrow[v >> 1];
              const float xc = (v & 1) ? xv[v >> 1].y : xv[v >> 1].x;
              __bf16 h, l;
              split_bf16(w * xc, h, l);
              ahs[q][rb2][v] = h;
              als[q][rb2][v] = l;
            }
          }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < IPW; ++q) {
          if (ig0 + i0 + q >= a.nlf) continue;  // uniform over the wave
#pragma unroll
          for (int rb2 = 0; rb2 < 2; ++rb2) {
#pragma unroll
            for (int s = 0; s < NB; ++s)
#pragma unroll
              for (int cb = 0; cb < NCB; ++cb) {
                rpa_f4 c = acc[s][q][rb2][cb];
                c = mfma_bf16_16x16x32(ahs[q][rb2], bh[s][cb], c);
                c = mfma_bf16_16x16x32(ahs[q][rb2], bl[s][cb], c);
                c = mfma_bf16_16x16x32(als[q][rb2], bh[s][cb], c);
                acc[s][q][rb2][cb] = c;
              }
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < IPW; ++q) {
          if (ig0 + i0 + q >= a.nlf) continue;  // uniform over the wave
          float2 xv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) xv[u] = sX[4 * kg + u][i0 + q];
#pragma unroll
          for (int rb2 = 0; rb2 < 2; ++rb2) {
            const float *wrow = &sW[16 * rb2 + r][4 * kg];
            rpa_bf8 ah, al;
#pragma unroll
            for (int v = 0; v < 8; ++v) {
              const float w = wrow[v >> 1];
              const float xc = (v & 1) ? xv[v >> 1].y : xv[v >> 1].x;
              __bf16 h, l;
              split_bf16(w * xc, h, l);
              ah[v] = h;
              al[v] = l;
            }
#pragma unroll
            for (int s = 0; s < NB; ++s)
#pragma unroll
              for (int cb = 0; cb < NCB; ++cb) {
                rpa_f4 c = acc[s][q][rb2][cb];
                c = mfma_bf16_16x16x32(ah, bh[s][cb], c);
                c = mfma_bf16_16x16x32(ah, bl[s][cb], c);
                c = mfma_bf16_16x16x32(al, bh[s][cb], c);
                acc[s][q][rb2][cb] = c;
              }
          }
        }
      }
    }
    // partials in the layout of k_accum, part[item][col/256][row][col%256],
    // in FP32: the reduction reads half the bytes and sums the same values
    float *dst = a.part + (size_t)item * a.npb * (kTileW * kColsWG);
#pragma unroll
    for (int q = 0; q < IPW; ++q) {
      const int i = ig0 + i0 + q;
      if (i >= a.nlf) continue;
#pragma unroll
      for (int rb2 = 0; rb2 < 2; ++rb2)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
          const int j = 16 * cb + r;
          if (j >= a.nlf) continue;
#pragma unroll
          for (int s = 0; s < NB; ++s) {
            const int col = CPLX ? 2 * (i * a.nlf + j) + s : i * a.nlf + j;
            float *d = dst + (size_t)(col / kColsWG) * (kTileW * kColsWG) + col % kColsWG;
#pragma unroll
            for (int e = 0; e < 4; ++e) d[(16 * rb2 + 4 * kg + e) * kColsWG] = acc[s][q][rb2][cb][e];
          }
        }
    }
    if (ib == 0) {  // padded columns stay zero
      for (int x = tid; x < kTileW * (a.ncol_pad - a.ncol); x += 256) {
        const int row = x / (a.ncol_pad - a.ncol), col = a.ncol + x % (a.ncol_pad - a.ncol);
        dst[(size_t)(col / kColsWG) * (kTileW * kColsWG) + row * kColsWG + col % kColsWG] = 0.f;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Mixed precision, home panels.  A transition's Lorentzian window starts in
// its home tile b and spans at most 2 hw / domega + 5 rows, so it lies in
// the row panel [32 b, 32 b + 16 NRB).  One work item is a chunk of up to
// kChunkHp transitions of one home tile with that whole panel, so every
// transition enters exactly one item and its products
//   P_t(G, G') = x_t(G) conj y_t(G')     (Re for Sloss; Re and Im for Optabs)
// are formed once:  S0(w, col) += sum_t W(w, t) P_t(col),  A = W (panel rows
// x transitions), B = P (transitions x columns), K = transitions, on the BF16
// matrix cores as hi*hi + hi*lo + lo*hi with FP32 accumulation over the chunk.
// Columns come in blocks of 16 with G fixed and G' = G'0 .. G'0 + 15.  The
// real S0 of Sloss is symmetric, so only G' >= G is formed (blocks G'0 = G +
// 16 j, G-major); the complex S0 of Optabs takes every G' (G'0 = 16 j).
constexpr int kHpRowA = 40;     // bf16 per staged weight row (32 + pad: conflict-free 16-byte reads)
constexpr int kHpRowX = 34;     // float2 per staged vertex row (32 + pad)

// (G, G'0) of column block bl (hp_nblocks gives the count)
__device__ inline void hp_block(int bl, int nlf, bool cplx, int &G, int &G0) {
  if (cplx) {
    const int ncb = (nlf + 15) / 16;
    G = bl / ncb;
    G0 = 16 * (bl - G * ncb);
    return;
  }
  int g = 0, rem = bl;
  while (g < nlf) {
    const int nb = (nlf - g + 15) / 16;
    if (rem < nb) break;
    rem -= nb;
    ++g;
  }
  G = g;
  G0 = g + 16 * rem;
}

// W of every item and stage as BF16 hi and lo parts:
// tab[(((item * nst + s) * 2 + hl) * rows + row) * 32 + tt] = part hl of
// W(32 b + row, t_lo + 32 s + tt), zero past t_hi and past nw.  The weight is
// the reference's FP64 arithmetic and cutoff test, rounded to FP32 and split.
// (Formed inside the panel kernel instead, from the energies, the weights
// made that kernel VALU-bound: 0.26 vs 0.23 ms with this table at N_G = 51.)
__global__ __launch_bounds__(256) void k_wsplit(const double *te, const int *items, const int *nitems, int nst,
                                                int rows, int program, int shift, double domega, double eta,
                                                double cut, double den, int nw, uint16_t *tab) {
  const int item = blockIdx.x / nst, s = blockIdx.x - item * nst;
  if (item >= nitems[0]) return;
  const int b = items[3 * item], t_lo = items[3 * item + 1], t_hi = items[3 * item + 2];
  const int t0 = t_lo + s * kHpKS;
  if (t0 >= t_hi) return;
  const int tt = threadIdx.x & 31, rg = threadIdx.x >> 5;
  const int t = t0 + tt;
  const bool tok = t < t_hi;
  const double *e = te + 3 * (size_t)(tok ? t : t_lo);
  const double e1 = e[0], e2 = e[1], z = e[2];
  uint16_t *hi = tab + ((size_t)item * nst + s) * 2 * rows * kHpKS;
  uint16_t *lo = hi + (size_t)rows * kHpKS;
  for (int row = rg; row < rows; row += 8) {
    const int w = b * kTileW + row;
    double v = 0.0;
    if (tok && w < nw) v = weight(program, w, shift, domega, eta, cut, den, e1, e2, z);
    __bf16 h, l;
    split_bf16((float)v, h, l);
    hi[row * kHpKS + tt] = __builtin_bit_cast(uint16_t, h);
    lo[row * kHpKS + tt] = __builtin_bit_cast(uint16_t, l);
  }
}

struct AccHpArgs {
  const uint16_t *tab;  // weights, k_wsplit
  const float2 *x, *y;  // vertices at sorted positions, [t][nlf]; y == x unless two sets
  const int *items, *nitems;
  int nlf, nbl, ngroup, nst;
  float *part;  // [item][nbl][16 NRB rows][16][NS]
};

// One work unit = (item, group of WPG NCW column blocks): wave w takes
// blocks (WPG group + w) NCW .. + NCW - 1.  The next stage's weights and
// vertices are loaded into registers during the products, every load from a
// clamped in-bounds address with its validity applied when the values are
// stored (a conditional load joined with a zero fill makes the compiler wait
// for each load before the next issues).  XU: staged vertex values per
// thread and set (32 nlf <= 64 WPG XU).
template <bool CPLX, bool TWO, int NRB, int NCW, int XU, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_accum_hp(AccHpArgs a) {
  constexpr int NS = CPLX ? 2 : 1;
  constexpr int ROWS = 16 * NRB;
  constexpr int NT = 64 * WPG;
  constexpr int NAU = (ROWS * 8 + NT - 1) / NT;  // 16-byte weight units per thread
  extern __shared__ __align__(16) unsigned char hp_raw[];
  uint16_t *sA = (uint16_t *)hp_raw;                                        // [2][ROWS][kHpRowA]
  float2 *sX = (float2 *)(hp_raw + sizeof(uint16_t) * 2 * ROWS * kHpRowA);  // [nlf + 1][kHpRowX]
  float2 *sY = TWO ? sX + (size_t)(a.nlf + 1) * kHpRowX : sX;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, kg = lane >> 4;
  const int nlf = a.nlf;
  // row nlf of the staged vertices stays zero: columns G' >= nlf read it
  for (int x = tid; x < kHpRowX; x += NT) {
    sX[(size_t)nlf * kHpRowX + x] = make_float2(0.f, 0.f);
    if (TWO) sY[(size_t)nlf * kHpRowX + x] = make_float2(0.f, 0.f);
  }
  // staging assignment, fixed for the kernel: weight unit q = tid + NT u ->
  // LDS offset; vertex value e = tid + NT u of a stage -> (tt, G) -> offset
  int ao[NAU];
#pragma unroll
  for (int u = 0; u < NAU; ++u) {
    const int q = min(tid + NT * u, ROWS * 8 - 1);
    const int hl = q / (ROWS * 4), row = (q >> 2) % ROWS, part = q & 3;
    ao[u] = (hl * ROWS + row) * kHpRowA + 8 * part;
  }
  const int nxe = kHpKS * nlf;  // vertex values per stage and set
  int xo[XU], xt[XU];
#pragma unroll
  for (int u = 0; u < XU; ++u) {
    const int e = tid + NT * u;
    const int tt = e / nlf, g = e - tt * nlf;
    xt[u] = e < nxe ? tt : kHpKS;  // kHpKS: not a value of the stage
    xo[u] = e < nxe ? g * kHpRowX + tt : 0;
  }
  const int xcd = blockIdx.x & 7, nsl = gridDim.x >> 3;
  const int nix = (a.nitems[0] - xcd + 7) / 8;
  const int total = nix * a.ngroup;

  for (int wk = blockIdx.x >> 3; wk < total; wk += nsl) {
    const int il = wk / a.ngroup, grp = wk - il * a.ngroup;
    const int item = 8 * il + xcd;
    const int t_lo = a.items[3 * item + 1], t_hi = a.items[3 * item + 2];
    const int nstg = (t_hi - t_lo + kHpKS - 1) / kHpKS;
    int Gb[NCW], Pb[NCW];
    bool okb[NCW];
#pragma unroll
    for (int cb = 0; cb < NCW; ++cb) {
      const int bl = (WPG * grp + wave) * NCW + cb;
      okb[cb] = bl < a.nbl;
      hp_block(okb[cb] ? bl : 0, nlf, CPLX, Gb[cb], Pb[cb]);
    }
    // 16-byte units as a native vector type: an array of HIP's uint4 struct
    // here was kept in scratch memory
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 *tsrc = (const u32x4 *)(a.tab + (size_t)item * a.nst * 2 * ROWS * kHpKS);

    rpa_f4 acc[NRB][NCW][NS];
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
      for (int cb = 0; cb < NCW; ++cb)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[rb][cb][s] = rpa_f4{0.f, 0.f, 0.f, 0.f};

    u32x4 ra[NAU];
    float2 rx[XU], ry[XU];
    int nvl = 0;  // valid transitions of the stage in registers
    auto load = [&](int st) {
      const u32x4 *src = tsrc + (size_t)st * (ROWS * 8);
#pragma unroll
      for (int u = 0; u < NAU; ++u) ra[u] = src[min(tid + NT * u, ROWS * 8 - 1)];
      const int t0 = t_lo + st * kHpKS;
      nvl = min(kHpKS, t_hi - t0);
      const int last = nvl * nlf - 1;  // the stage's last valid value
      const float2 *xs = a.x + (size_t)t0 * nlf;
      const float2 *ys = a.y + (size_t)t0 * nlf;
#pragma unroll
      for (int u = 0; u < XU; ++u) {
        const int e = min(tid + NT * u, last);
        rx[u] = xs[e];
        if (TWO) ry[u] = ys[e];
      }
    };
    auto store = [&]() {
#pragma unroll
      for (int u = 0; u < NAU; ++u)
        if (tid + NT * u < ROWS * 8) *(u32x4 *)(sA + ao[u]) = ra[u];
#pragma unroll
      for (int u = 0; u < XU; ++u) {
        if (xt[u] < kHpKS) {
          const bool ok = xt[u] < nvl;
          sX[xo[u]] = ok ? rx[u] : make_float2(0.f, 0.f);
          if (TWO) sY[xo[u]] = ok ? ry[u] : make_float2(0.f, 0.f);
        }
      }
    };

    load(0);
    for (int st = 0; st < nstg; ++st) {
      __syncthreads();  // the previous stage's reads are done
      store();
      __syncthreads();
      if (st + 1 < nstg) load(st + 1);
      // B fragments of every block: P = x conj y for the lane's column and
      // the transitions 8 kg .. 8 kg + 7, split into BF16 hi and lo
      rpa_bf8 bh[NCW], bl[NCW], ch[NCW], cl[NCW];
#pragma unroll
      for (int cb = 0; cb < NCW; ++cb) {
        const int gp = Pb[cb] + r;
        const float2 *xrow = sX + (size_t)Gb[cb] * kHpRowX + 8 * kg;
        const float2 *yrow = sY + (size_t)(gp < nlf ? gp : nlf) * kHpRowX + 8 * kg;
#pragma unroll
        for (int v = 0; v < 8; v += 2) {
          const float4 p = *(const float4 *)(xrow + v), q = *(const float4 *)(yrow + v);
          const float xr[2] = {p.x, p.z}, xi[2] = {p.y, p.w}, yr[2] = {q.x, q.z}, yi[2] = {q.y, q.w};
#pragma unroll
          for (int h2 = 0; h2 < 2; ++h2) {
            __bf16 h, l;
            split_bf16(xr[h2] * yr[h2] + xi[h2] * yi[h2], h, l);  // Re(x conj y)
            bh[cb][v + h2] = h;
            bl[cb][v + h2] = l;
            if (CPLX) {
              split_bf16(xi[h2] * yr[h2] - xr[h2] * yi[h2], h, l);  // Im(x conj y)
              ch[cb][v + h2] = h;
              cl[cb][v + h2] = l;
            }
          }
        }
      }
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb) {
        const rpa_bf8 ah = *(const rpa_bf8 *)(sA + (16 * rb + r) * kHpRowA + 8 * kg);
        const rpa_bf8 al = *(const rpa_bf8 *)(sA + (ROWS + 16 * rb + r) * kHpRowA + 8 * kg);
#pragma unroll
        for (int cb = 0; cb < NCW; ++cb) {
          rpa_f4 c = acc[rb][cb][0];
          c = mfma_bf16_16x16x32(ah, bh[cb], c);
          c = mfma_bf16_16x16x32(ah, bl[cb], c);
          c = mfma_bf16_16x16x32(al, bh[cb], c);
          acc[rb][cb][0] = c;
          if (CPLX) {
            rpa_f4 d = acc[rb][cb][NS - 1];
            d = mfma_bf16_16x16x32(ah, ch[cb], d);
            d = mfma_bf16_16x16x32(ah, cl[cb], d);
            d = mfma_bf16_16x16x32(al, ch[cb], d);
            acc[rb][cb][NS - 1] = d;
          }
        }
      }
    }
    float *dst = a.part + (size_t)item * a.nbl * ROWS * 16 * NS;
#pragma unroll
    for (int cb = 0; cb < NCW; ++cb) {
      if (!okb[cb]) continue;
      const int bl = (WPG * grp + wave) * NCW + cb;
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 16 * rb + 4 * kg + e;
#pragma unroll
          for (int s = 0; s < NS; ++s) dst[(((size_t)bl * ROWS + row) * 16 + r) * NS + s] = acc[rb][cb][s][e];
        }
    }
  }
}

// S0[w][col] from the panels: the items of every home tile b whose panel
// [32 b, 32 b + rows) holds row w, in tile then item order, summed in FP64
// (Neumaier).  Thread = (w, column block, column r, part s); the symmetric
// real S0 writes (G, G') and (G', G).  Columns past ncol are set to zero.
__global__ void k_reduce_hp(const float *part, const int *first_item, int rows, int nlf, int nbl, int cplx,
                            int nw, int ncol, int ncol_pad, double *s0) {
  const int NS = cplx ? 2 : 1;
  const long long per_w = (long long)nbl * 16 * NS;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long npad = (long long)nw * (ncol_pad - ncol);
  if (idx >= (long long)nw * per_w) {
    const long long z = idx - (long long)nw * per_w;
    if (z < npad) {
      const int w = (int)(z / (ncol_pad - ncol)), c = ncol + (int)(z - (long long)w * (ncol_pad - ncol));
      s0[(size_t)w * ncol_pad + c] = 0.0;
    }
    return;
  }
  const int w = (int)(idx / per_w);
  const int rest = (int)(idx - (long long)w * per_w);
  const int bl = rest / (16 * NS), rr = (rest / NS) % 16, s = rest % NS;
  int G, G0;
  hp_block(bl, nlf, cplx != 0, G, G0);
  const int gp = G0 + rr;
  if (gp >= nlf) return;
  const int b_hi = w / kTileW;
  const int b_lo = w >= rows ? (w - rows) / kTileW + 1 : 0;
  double sum = 0.0, comp = 0.0;
  for (int b = b_lo; b <= b_hi; ++b) {
    const int row = w - b * kTileW;
    for (int it = first_item[b]; it < first_item[b + 1]; ++it) {
      const double v = (double)part[((((size_t)it * nbl + bl) * rows + row) * 16 + rr) * NS + s];
      const double tt = sum + v;
      if (fabs(sum) >= fabs(v))
        comp += (sum - tt) + v;
      else
        comp += (v - tt) + sum;
      sum = tt;
    }
  }
  const double val = sum + comp;
  double *row_out = s0 + (size_t)w * ncol_pad;
  if (cplx) {
    row_out[2 * ((size_t)G * nlf + gp) + s] = val;
  } else {
    row_out[(size_t)G * nlf + gp] = val;
    if (gp != G) row_out[(size_t)gp * nlf + G] = val;
  }
}

// S0[w][col] = sum over the chunks of tile w/32 of their partials (Neumaier,
// in FP64; the partials are FP64 from k_accum, FP32 from k_accum_mx).
template <class T>
__global__ void k_reduce(const T *part, const int *first_item, int npb, int nw, int ncol_pad, double *s0) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)nw * ncol_pad) return;
  const int w = (int)(idx / ncol_pad), col = (int)(idx - (long long)w * ncol_pad);
  const int b = w / kTileW, row = w - b * kTileW;
  const int pbk = col / kColsWG, cc = col - pbk * kColsWG;
  const int i0 = first_item[b], i1 = first_item[b + 1];
  double s = 0.0, comp = 0.0;
  for (int it = i0; it < i1; ++it) {
    const double v = (double)part[((size_t)it * npb + pbk) * (kTileW * kColsWG) + row * kColsWG + cc];
    double tt = s + v;
    if (fabs(s) >= fabs(v))
      comp += (s - tt) + v;
    else
      comp += (v - tt) + s;
    s = tt;
  }
  s0[idx] = s + comp;
}

// Qeff(iG,jG) = sum_intra 2 fact j1(iG) conj(j2(jG)) / (Ntot Vcell)
template <class C2>
__global__ void k_qeff(const double *te, const C2 *rho1, const C2 *rho2, const int *toff,
                       int ntiles, int nlf, double dq, double *qeff) {
#pragma clang fp contract(off)
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nlf * nlf) return;
  int iG = idx / nlf, jG = idx - iG * nlf;
  double sr = 0.0, si = 0.0;
  for (int t = toff[ntiles]; t < toff[ntiles + 1]; ++t) {
    double f = 2.0 * te[3 * (size_t)t + 2];
    const C2 x = rho1[(size_t)t * nlf + iG], y = rho2[(size_t)t * nlf + jG];
    const double xr = f * (double)x.x, xi = f * (double)x.y;
    const double yr = y.x, yi = y.y;
    sr += (xr * yr + xi * yi) / dq;
    si += (xi * yr - xr * yi) / dq;
  }
  qeff[2 * idx] = sr;
  qeff[2 * idx + 1] = si;
}

// ---------------------------------------------------------------------------

int launch_transitions(rpa_ctx *c) {
  const long long nt = c->ntrans;
  const int nblk = (int)((nt + 255) / 256);
  const int nkeys = c->ntiles + 1;
  if (nkeys > kMaxTiles) return fail(c, RPA_E_LIMIT, "too many frequency tiles");
  if (nt > (long long)INT32_MAX) return fail(c, RPA_E_LIMIT, "too many transitions");
  TransArgs ta{};
  ta.k1 = c->d_k1;
  ta.k2 = c->d_k2;
  ta.E = c->d_E;
  ta.nki = c->d.nki;
  ta.nband = c->d.nband;
  ta.pairs = c->pairs;
  ta.ntrans = nt;
  ta.program = c->d.program;
  ta.domega = c->p.domega;
  ta.shift = (c->d.program == RPA_OPTABS) ? c->d.no : 0;
  ta.nw = c->nw;
  ta.ntiles = c->ntiles;
  ta.efermi = c->p.efermi;
  ta.temp = c->p.temp;
  ta.df_cut = c->p.df_cut;
  ta.key = c->d_key;
  // window half-width: |Lor| >= cut  <=>  dE^2 <= eta/cut - eta^2
  double h2 = c->p.eta / c->wcut - c->p.eta * c->p.eta;
  ta.hw = (c->wcut > 0.0) ? std::sqrt(std::max(h2, 0.0)) : INFINITY;
  if (nt == 0) return 0;
  hipLaunchKernelGGL(k_transitions, dim3(nblk), dim3(256), 0, c->stream, ta);
  hipLaunchKernelGGL(k_tile_hist, dim3(nblk), dim3(256), 0, c->stream, (const int *)c->d_key, nt,
                     nkeys, nblk, c->d_bcnt);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, c->stream, c->d_bcnt,
                     (long long)nkeys * nblk, nkeys, nblk, c->d_toff);
  ScatterArgs sa{};
  sa.key = c->d_key;
  sa.bcnt = c->d_bcnt;
  sa.ntrans = nt;
  sa.nblk = nblk;
  sa.k1 = c->d_k1;
  sa.k2 = c->d_k2;
  sa.E = c->d_E;
  sa.nki = c->d.nki;
  sa.pairs = c->pairs;
  sa.program = c->d.program;
  sa.ntiles = c->ntiles;
  sa.efermi = c->p.efermi;
  sa.temp = c->p.temp;
  sa.perm = c->d_perm;
  sa.nat = c->d_nat;
  sa.te = c->d_te;
  hipLaunchKernelGGL(k_tile_scatter, dim3(nblk), dim3(256), 0, c->stream, sa);
  hipLaunchKernelGGL(k_worklist, dim3(1), dim3(1024), 0, c->stream, (const int *)c->d_toff,
                     c->ntiles, c->wl_span, c->wl_chunk,
                     c->d_items, c->d_first_item, c->d_nitems);
  RPA_HIP(c, hipGetLastError());
  return 0;
}

template <bool CPLX, int NCB, int IPW>
static void launch_mx(rpa_ctx *c, const AccMxArgs &a, int grid) {
  hipLaunchKernelGGL((k_accum_mx<CPLX, NCB, IPW>), dim3(grid), dim3(256), 0, c->stream, a);
}

template <bool CPLX>
static int launch_mx_ncb(rpa_ctx *c, AccMxArgs &a, long long items) {
  constexpr int IPW = CPLX ? 1 : 2;
  const int ncb = (c->nlf + 15) / 16;
  a.nib = (c->nlf + 4 * IPW - 1) / (4 * IPW);
  const int grid = (int)std::min<long long>(std::max<long long>((items * a.nib + 7) / 8 * 8, 8), 8192);
  switch (ncb) {
    case 1: launch_mx<CPLX, 1, IPW>(c, a, grid); break;
    case 2: launch_mx<CPLX, 2, IPW>(c, a, grid); break;
    case 3: launch_mx<CPLX, 3, IPW>(c, a, grid); break;
    case 4: launch_mx<CPLX, 4, IPW>(c, a, grid); break;
    case 5: launch_mx<CPLX, 5, IPW>(c, a, grid); break;
    case 6: launch_mx<CPLX, 6, IPW>(c, a, grid); break;
    default: return fail(c, RPA_E_LIMIT, "mixed precision accumulation supports Nlf <= 96");
  }
  return 0;
}

// Home-panel launch: NRB row blocks, XU staged vertex values per thread.
// Column blocks per wave: as many as fit 256 registers without spills
// (measured at N_G = 51, 112 panel rows: 2 blocks 0.25 ms, 3 blocks 0.205
// ms; workgroups of 4 waves, two per CU, were no faster than one of 8); the
// complex S0 keeps two accumulator sets per block.
constexpr int kHpWaves = 8;
constexpr int hp_ncw(bool cplx, int nrb, int xu) {
  return cplx ? ((nrb <= 4 && xu == 4) ? 2 : 1) : ((nrb >= 7 && xu == 8) ? 2 : 3);
}
template <bool CPLX, bool TWO, int NRB>
static const void *hp_kernel(int nlf) {
  if (nlf <= 64) return (const void *)k_accum_hp<CPLX, TWO, NRB, hp_ncw(CPLX, NRB, 4), 4, kHpWaves>;
  return (const void *)k_accum_hp<CPLX, TWO, NRB, hp_ncw(CPLX, NRB, 8), 8, kHpWaves>;
}
template <bool CPLX, bool TWO>
static const void *hp_kernel_nrb(int nrb, int nlf) {
  switch (nrb) {
    case 4: return hp_kernel<CPLX, TWO, 4>(nlf);
    case 5: return hp_kernel<CPLX, TWO, 5>(nlf);
    case 6: return hp_kernel<CPLX, TWO, 6>(nlf);
    case 7: return hp_kernel<CPLX, TWO, 7>(nlf);
    default: return hp_kernel<CPLX, TWO, 8>(nlf);
  }
}
static const void *hp_kernel_of(const rpa_ctx *c) {
  const bool cplx = (c->d.program == RPA_OPTABS);
  const bool two = cplx && c->d_rho2 != nullptr;
  if (!cplx) return hp_kernel_nrb<false, false>(c->hp_nrb, c->nlf);
  return two ? hp_kernel_nrb<true, true>(c->hp_nrb, c->nlf) : hp_kernel_nrb<true, false>(c->hp_nrb, c->nlf);
}
static size_t hp_lds(const rpa_ctx *c) {
  const bool two = (c->d.program == RPA_OPTABS) && c->d_rho2 != nullptr;
  return sizeof(uint16_t) * 2 * 16 * (size_t)c->hp_nrb * kHpRowA +
         sizeof(float2) * (size_t)(c->nlf + 1) * kHpRowX * (two ? 2 : 1);
}

static int launch_hp(rpa_ctx *c, int shift, double den) {
  const bool cplx = (c->d.program == RPA_OPTABS);
  const int rows = 16 * c->hp_nrb, nst = kChunkHp / kHpKS;
  hipLaunchKernelGGL(k_wsplit, dim3((unsigned)(c->max_items * nst)), dim3(256), 0, c->stream,
                     (const double *)c->d_te, (const int *)c->d_items, (const int *)c->d_nitems, nst, rows,
                     c->d.program, shift, c->p.domega, c->p.eta, c->wcut, den, c->nw, c->d_wtab);
  AccHpArgs a{};
  a.tab = c->d_wtab;
  a.x = (const float2 *)c->d_rho1;
  a.y = (const float2 *)(c->d_rho2 ? c->d_rho2 : c->d_rho1);
  a.items = c->d_items;
  a.nitems = c->d_nitems;
  a.nlf = c->nlf;
  a.nbl = hp_nblocks(c->nlf, cplx);
  const int ncw = hp_ncw(cplx, c->hp_nrb, c->nlf <= 64 ? 4 : 8);
  a.ngroup = (a.nbl + kHpWaves * ncw - 1) / (kHpWaves * ncw);
  a.nst = nst;
  a.part = (float *)c->d_part;
  const size_t lds = hp_lds(c);
  if (lds > 64 * 1024 && lds > c->accum_lds_attr)
    return fail(c, RPA_E_STATE, "accumulation LDS limit not configured");
  const long long units = (long long)c->max_items * a.ngroup;
  const int grid = (int)std::min<long long>(std::max<long long>((units + 7) / 8 * 8, 8), 8192);
  const void *fn = hp_kernel_of(c);
  void *args[] = {&a};
  RPA_HIP(c, hipLaunchKernel(fn, dim3(grid), dim3(64 * kHpWaves), args, lds, c->stream));
  if (c->profiling) RPA_HIP(c, hipEventRecord(c->prof_ev[3], c->stream));
  const long long n = (long long)c->nw * a.nbl * 16 * (cplx ? 2 : 1) + (long long)c->nw * (c->ncol_pad - c->ncol);
  hipLaunchKernelGGL(k_reduce_hp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     (const float *)c->d_part, (const int *)c->d_first_item, rows, c->nlf, a.nbl, cplx ? 1 : 0,
                     c->nw, c->ncol, c->ncol_pad, c->d_s0);
  return 0;
}

// FP64 k_accum stages kKS vertices of every local field (both vertex sets
// for Optabs) next to the weight rows: 4352 + 512 Nlf bytes for Optabs, past
// the default 64 KB from Nlf = 120.  Outside any stream capture (prepare):
// raise the instance's dynamic-LDS limit to what the current Nlf needs.
static size_t accum_lds(const rpa_ctx *c) {
  const bool cplx = (c->d.program == RPA_OPTABS);
  return sizeof(double) * kTileW * (kKS + 1) + sizeof(double2) * kKS * c->nlf * (cplx ? 2 : 1);
}

int configure_accumulate(rpa_ctx *c) {
  if (c->hp_nrb) {  // the instance depends on the q's Nlf: set its limit every time
    const size_t lds = hp_lds(c);
    if (lds <= 64 * 1024) return 0;
    if (lds > 160 * 1024) return fail(c, RPA_E_LIMIT, "Nlf too large for the accumulation staging");
    RPA_HIP(c, hipFuncSetAttribute(hp_kernel_of(c), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    c->accum_lds_attr = std::max(c->accum_lds_attr, lds);
    return 0;
  }
  if (c->d.precision == RPA_MIXED) return 0;
  const size_t lds = accum_lds(c);
  if (lds <= 64 * 1024 || lds <= c->accum_lds_attr) return 0;
  if (lds > 160 * 1024) return fail(c, RPA_E_LIMIT, "Nlf too large for the accumulation staging");
  const int ru = (kKS * c->nlf + 255) / 256;
  const bool cplx = (c->d.program == RPA_OPTABS);
  const void *fn = cplx ? (ru <= 2 ? (const void *)k_accum<true, 2> : ru <= 4 ? (const void *)k_accum<true, 4>
                                                                              : (const void *)k_accum<true, 8>)
                        : (ru <= 2 ? (const void *)k_accum<false, 2> : ru <= 4 ? (const void *)k_accum<false, 4>
                                                                               : (const void *)k_accum<false, 8>);
  RPA_HIP(c, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  c->accum_lds_attr = lds;
  return 0;
}

int launch_accumulate(rpa_ctx *c) {
  const bool cplx = (c->d.program == RPA_OPTABS);
  const bool mixed = (c->d.precision == RPA_MIXED);
  const int npb = c->ncol_pad / kColsWG;
  const double den = (M_PI * (double)c->ntot) * c->p.vcell;
  const int shift = cplx ? c->d.no : 0;
  if (mixed && c->hp_nrb) {
    if (int e = launch_hp(c, shift, den)) return e;
  } else if (mixed) {
    if (c->nlf > 96) return fail(c, RPA_E_LIMIT, "mixed precision accumulation supports Nlf <= 96");
    const long long nt = std::max<long long>(c->ntrans, 1);
    const float2 *y = (const float2 *)(c->d_rho2 ? c->d_rho2 : c->d_rho1);
    hipLaunchKernelGGL((k_weight_table<float, kChunkTMx>), dim3((unsigned)(c->max_items * (kChunkTMx / 256))), dim3(256), 0,
                       c->stream, (const double *)c->d_te, (const int *)c->d_items, (const int *)c->d_nitems,
                       c->d.program, shift, c->p.domega, c->p.eta, c->wcut, den, c->nw, c->d_wmx);
    hipLaunchKernelGGL(k_mx_split, dim3((unsigned)((nt + 63) / 64)), dim3(256), 64 * c->nlf * sizeof(float2),
                       c->stream, y, nt, c->nlf, c->d_rbt);
    AccMxArgs a{};
    a.wmx = c->d_wmx;
    a.rho1 = (const float2 *)c->d_rho1;
    a.rbt = c->d_rbt;
    a.nt = nt;
    a.items = c->d_items;
    a.nitems = c->d_nitems;
    a.nlf = c->nlf;
    a.nw = c->nw;
    a.ncol = c->ncol;
    a.ncol_pad = c->ncol_pad;
    a.npb = npb;
    a.part = (float *)c->d_part;
    const int e = cplx ? launch_mx_ncb<true>(c, a, (long long)c->max_items)
                       : launch_mx_ncb<false>(c, a, (long long)c->max_items);
    if (e) return e;
  } else {
    hipLaunchKernelGGL((k_weight_table<double, kChunkT>), dim3((unsigned)(c->max_items * (kChunkT / 256))), dim3(256), 0,
                       c->stream, (const double *)c->d_te, (const int *)c->d_items, (const int *)c->d_nitems,
                       c->d.program, shift, c->p.domega, c->p.eta, c->wcut, den, c->nw, c->d_wt64);
    AccArgs a{};
    a.te = c->d_te;
    a.wt = c->d_wt64;
    a.rho1 = (const double2 *)c->d_rho1;
    a.rho2 = (const double2 *)(c->d_rho2 ? c->d_rho2 : c->d_rho1);
    a.items = c->d_items;
    a.nitems = c->d_nitems;
    a.npb = npb;
    a.nlf = c->nlf;
    a.nw = c->nw;
    a.ncol = c->ncol;
    a.ncol_pad = c->ncol_pad;
    a.program = c->d.program;
    a.shift = shift;
    a.domega = c->p.domega;
    a.eta = c->p.eta;
    a.cut = c->wcut;
    a.den = den;
    a.part = c->d_part;
    const size_t lds = accum_lds(c);
    if (lds > 64 * 1024 && lds > c->accum_lds_attr)
      return fail(c, RPA_E_STATE, "accumulation LDS limit not configured");
    long long work = (long long)c->max_items * a.npb;
    int grid = (int)std::min<long long>(std::max<long long>((work + 7) / 8 * 8, 8), 4096);
    const int ru = (kKS * c->nlf + 255) / 256;
    if (ru > 8) return fail(c, RPA_E_LIMIT, "Nlf too large for the accumulation staging");
    if (cplx) {
      if (ru <= 2)
        hipLaunchKernelGGL((k_accum<true, 2>), dim3(grid), dim3(256), lds, c->stream, a);
      else if (ru <= 4)
        hipLaunchKernelGGL((k_accum<true, 4>), dim3(grid), dim3(256), lds, c->stream, a);
      else
        hipLaunchKernelGGL((k_accum<true, 8>), dim3(grid), dim3(256), lds, c->stream, a);
    } else {
      if (ru <= 2)
        hipLaunchKernelGGL((k_accum<false, 2>), dim3(grid), dim3(256), lds, c->stream, a);
      else if (ru <= 4)
        hipLaunchKernelGGL((k_accum<false, 4>), dim3(grid), dim3(256), lds, c->stream, a);
      else
        hipLaunchKernelGGL((k_accum<false, 8>), dim3(grid), dim3(256), lds, c->stream, a);
    }
  }
  if (!c->hp_nrb && c->profiling) RPA_HIP(c, hipEventRecord(c->prof_ev[3], c->stream));
  const long long n = (long long)c->nw * c->ncol_pad;
  if (mixed && c->hp_nrb) {
    // reduced in launch_hp
  } else if (mixed)
    hipLaunchKernelGGL(k_reduce<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       (const float *)c->d_part, (const int *)c->d_first_item, npb, c->nw, c->ncol_pad, c->d_s0);
  else
    hipLaunchKernelGGL(k_reduce<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       (const double *)c->d_part, (const int *)c->d_first_item, npb, c->nw, c->ncol_pad, c->d_s0);
  if (cplx) {
    const int nn = c->nlf * c->nlf;
    const dim3 g((nn + 255) / 256);
    if (mixed)
      hipLaunchKernelGGL(k_qeff<float2>, g, dim3(256), 0, c->stream, (const double *)c->d_te,
                         (const float2 *)c->d_rho1, (const float2 *)(c->d_rho2 ? c->d_rho2 : c->d_rho1),
                         (const int *)c->d_toff, c->ntiles, c->nlf, (double)c->ntot * c->p.vcell,
                         c->d_qeff);
    else
      hipLaunchKernelGGL(k_qeff<double2>, g, dim3(256), 0, c->stream, (const double *)c->d_te,
                         (const double2 *)c->d_rho1, (const double2 *)(c->d_rho2 ? c->d_rho2 : c->d_rho1),
                         (const int *)c->d_toff, c->ntiles, c->nlf, (double)c->ntot * c->p.vcell,
                         c->d_qeff);
  }
  RPA_HIP(c, hipGetLastError());
  return 0;
}

}  // namespace rpa
