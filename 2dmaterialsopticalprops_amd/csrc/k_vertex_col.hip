// Charge vertices of the direct map by column correlation (Sloss).
//
//   rho_nm(k,q,G) = sum_{i < NGd} conj(C1_n(i)) C2_m(map(G,i))   (Sloss.f90:905-951)
//
// In the direct map plane wave i of C1 is the global G(i), and map(G,i) is the
// global position of G(i) + Glf(G) + G0, used when it lies below ngk(K2)
// (Sloss.f90:928-944).  Group the Miller triples of G(0..NGd-1) into columns
// of equal (mx, my).  For a local field G = (lx, ly, lz) the target of
// (mx, my, z) is (mx + lx + ox, my + ly + oy, z + lz + oz): the local fields of
// one in-plane class (lx, ly) differ only by a shift lz along the column.  So
// the k+q rows of a column -- the target column, z from zlo + lzmin to
// zhi + lzmax -- are staged in LDS once and serve every local field of the
// class at an offset of lz - lzmin rows.  The sum becomes a correlation along
// z: each staged k+q value enters up to (fields of the class) products, and
// L2 delivers a k+q row once per class instead of once per local field.
//
// Matrix cores: rows are (8 bands n) x (Re, Im) of the real embedding of
// conj(C1), K runs over positions (plane waves), columns are (4 shifts) x
// (4 bands m); Re and Im of C2 are two MFMAs.  One accumulator per tile
// (group of 4 shifts, group of 4 bands m); the four waves split the tiles and
// each walks every position of the chunk, loading the position's A value and
// B row offset once and reusing them for all of its tiles.  The summation
// runs over positions in column order instead of the reference's i order.
// Mixed charge vertices run the same tiles on the BF16 matrix cores as hi/lo
// split products (ColBf3, k_vertex_col_split); FP64, and FP32 current
// vertices, on the FP64 / FP32 16x16x4 MFMAs.
//
// Current vertices (Optabs, Optabs.f90:1412-1466) use the same correlation:
// with 0.5 J(G,i) = alpha(G) + beta(i), alpha = 0.5 Gcar (q + 2k + Glf(G))_pol
// and beta = Gcar G(i)_pol,
//   j(G) = alpha(G) rho(G) + sum_i conj(C1(i)) beta(i) C2(map(G,i)),
// so each staged k+q value also feeds a second (pol yz: third) accumulator
// whose A operand is scaled by beta of the position; alpha is applied when
// the results are written.
//
// Workgroup = (k, in-plane class, band block); workgroups of one k on one XCD.
#include <algorithm>
#include <array>
#include <climits>
#include <cstdlib>
#include <numeric>
#include <string>

#include "rpa_internal.h"

namespace rpa {

constexpr int kColInfo = 2 + kColMaxShift / 4;  // per class: rowoff, nsg, sg list
constexpr size_t kColLdsBudget = 80 * 1024; // default: two workgroups per CU
// positions per chunk and staged k+q rows per thread (rows <= 64 NRU)
constexpr int kKcF32 = 128, kKcF64 = 64;
constexpr int kNruF32 = 8, kNruF64 = 4;

struct ColArgs {
  const void *box[2];  // basis arrays, rows of bw[g] bands (column-box or global order)
  int bw[2], blo1, nbg;  // band widths, first band of array 1, arrays in use
  int lim_ngk;           // 1: rows past ngk(K2) read as absent (global order); 0: stored as zero
  const int *ngk, *k1, *k2, *g0idx;
  const int *slot;   // [npos] C1 slot (global index) of position p
  const int *prow;   // [ncls][npos] B row of p relative to its chunk
  const int *chunk;  // [nchunk + 1] first position of each chunk
  const int *crow;   // [ncls][nchunk + 1] first B row of each chunk in the class
  const int *bmap;   // [ng0][nrows] global index of each B row's target, -1 none
  const int *shift;  // [ncls][kColMaxShift] local field of each shift slot, -1 none
  const int *cls;    // [ncls][kColInfo]
  int npos, nchunk, ncls, nrows;
  int nband, npwx, nlf, ntot;
  int blocks;        // ncls * nvb
  int ms;            // complex values per staged k+q row (bank-conflict-free stride)
  int nvb;
  VBlock vb[kVertexMaxBlocks];
  PairTable pairs;
  const int *perm;
  void *rho, *rho2;
  // current vertices: beta of every position for pol1, pol2; alpha inputs
  const double *beta;  // [2][npos]
  const double *ktot, *glf;
  double q[3];
  double gcar;
  int pol1, pol2;
  // LDS DMA form: the k+q box array as (hi, lo) BF16 words, its values and
  // row stride in values
  const uint32_t *boxw;
  long long boxw_nv;
  int bww;
};

struct ColF64 {
  using real = double;
  using cplx = double2;
  using acc = rpa_d4;
  static __device__ inline acc mma(real a, real b, acc c) { return mfma_f64_16x16x4(a, b, c); }
  static __device__ inline int arow_band(int r) { return r & 7; }
  static __device__ inline bool arow_im(int r) { return r >= 8; }
  static __device__ inline int out_band(int q4, int p) { return 4 * p + q4; }
  static __device__ inline int out_re(int p) { return p; }
  static __device__ inline int out_im(int p) { return p + 2; }
  static __device__ inline cplx make(real x, real y) { return make_double2(x, y); }
  static __device__ inline acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
  static constexpr bool split = false;
};

struct ColF32 {
  using real = float;
  using cplx = float2;
  using acc = rpa_f4;
  static __device__ inline acc mma(real a, real b, acc c) { return mfma_f32_16x16x4(a, b, c); }
  static __device__ inline int arow_band(int r) { return r >> 1; }
  static __device__ inline bool arow_im(int r) { return (r & 1) != 0; }
  static __device__ inline int out_band(int q4, int p) { return 2 * q4 + p; }
  static __device__ inline int out_re(int p) { return 2 * p; }
  static __device__ inline int out_im(int p) { return 2 * p + 1; }
  static __device__ inline cplx make(real x, real y) { return make_float2(x, y); }
  static __device__ inline acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
  static constexpr bool split = false;
};

// FP32 vertices on the BF16 matrix cores (charge vertices): every staged
// value is split as hi = bf16(x), lo = bf16(x - hi) when it is written to
// LDS, and a product is hi.hi + hi.lo + lo.hi with FP32 accumulation
// (relative error of one product below 2^-16; the dropped lo.lo term is
// below 2^-18).  Rows and result layout as ColF32; K of one MFMA is 8
// positions x (Re, Im) x (hi, lo): K index 8 kg + 4 j + 2 h + c <-> position
// 8 t + 2 kg + j, part h, component c.  The hi and lo words of a staged k+q
// value lie side by side, so a lane's B operand per tile is two 8-byte LDS
// reads; A holds (hi, hi) for the hi.hi + hi.lo products and (lo, 0) for
// lo.hi, so a 16-position step costs 4 MFMAs where an operand layout with
// K = 16 positions x (Re, Im) costs 3, but that one needs 8 register moves
// per tile to gather hi and lo words into separate operands (measured at
// N_G = 51: 1.04 ms with this layout and batched reads, 1.16 ms with that
// one).
struct ColBf3 : ColF32 {
  static constexpr bool split = true;
};

typedef uint32_t rpa_u4 __attribute__((ext_vector_type(4)));
typedef int rpa_i4 __attribute__((ext_vector_type(4)));

// hi and lo BF16 parts of (Re, Im), Re in the low half of each word
__device__ inline void split_pair(float2 z, uint32_t &h, uint32_t &l) {
  const __bf16 hx = (__bf16)z.x, hy = (__bf16)z.y;
  const __bf16 lx = (__bf16)(z.x - (float)hx), ly = (__bf16)(z.y - (float)hy);
  h = (uint32_t)__builtin_bit_cast(uint16_t, hx) | ((uint32_t)__builtin_bit_cast(uint16_t, hy) << 16);
  l = (uint32_t)__builtin_bit_cast(uint16_t, lx) | ((uint32_t)__builtin_bit_cast(uint16_t, ly) << 16);
}

// Staging of one chunk in registers (loaded one chunk ahead, stored to LDS
// after the barrier).  Thread t: C1 rows of positions t/8 + 32u, band t%8;
// k+q rows t/4 + 64u, bands m t%4 + 4e (16 consecutive lanes store to 16
// distinct bank pairs: 2% faster than bands 4(t%4)..+3 at N_G = 51, half
// the LDS bank conflicts); B row of position t (t < KC).
// Every load is issued unconditionally from a clamped, in-bounds address and
// its presence kept in a mask; the zero fill of absent values happens in the
// LDS store.  (A conditional load joined with a zero fill makes the compiler
// wait for each load before the next one issues.)
template <class P, int KC, int NRU>
struct ColStage {
  static_assert(4 * NRU <= 32 && KC / 32 <= 32, "the 32-bit presence masks hold 4 NRU and KC / 32 bits");
  typename P::cplx a[KC / 32];
  typename P::cplx b[NRU][4];
  int prow;
  typename P::real beta[2];
  unsigned amask;  // bit u: C1 value u present
  unsigned bmask;  // bit 4u + e: k+q value (u, e) present
  int np, nr;      // positions and B rows of the chunk
};

// Products of one staged chunk in the split form (mixed charge vertices):
// 8 positions per block, K = (position, Re/Im, hi/lo): a lane's B operand is
// the (hi, lo) word pairs of its 2 positions, two 8-byte LDS reads straight
// into the operand registers.  A1 = (hi, hi) of C1 against (hi, lo) gives
// hi.hi + hi.lo, A2 = (lo, 0) gives lo.hi.  Software pipelined: after a
// tile's two products its registers take the next block's B values of that
// tile, so the reads are in flight during the other tiles' products (the
// group barriers keep the compiler from sinking each read to its product,
// which left one LDS latency per tile exposed); rows and A parts are read two
// blocks ahead.  Positions past the chunk (to the next multiple of 8 <= KC)
// have zero A parts and B row 0; reads past the last block use clamped
// positions and are never consumed.  The A row of an Im row is (-Im, Re) of
// C1: the halves of each word swapped and the sign of the new low half
// flipped.  sAw: [hi, lo][8 bands n][KC + 4] words; sP: [KC] B row of each
// position; sBw: [rows][ms][hi, lo] words.
// After block t < nissue the caller's issue(t) runs (the next chunk's
// requests, spread over the products); the rest after the last block.
struct NoIssue {
  __device__ void operator()(int) const {}
};

template <int TPW, int KC, class Issue = NoIssue>
__device__ __forceinline__ void split_products(rpa_f4 (&acc)[TPW], const uint32_t *sAw, const int *sP,
                                               const uint32_t *sBw, int nblk, int ms, int lane_off,
                                               const int (&toff)[TPW], int band_r, bool im_row, int kg,
                                               Issue &&issue = Issue{}, int nissue = 0) {
  constexpr int KCA = KC + 4;
  const uint32_t *aph = sAw + band_r * KCA, *apl = aph + 8 * KCA;
  auto rot = [&](uint32_t w) { return im_row ? (__builtin_amdgcn_alignbit(w, w, 16) ^ 0x8000u) : w; };
  auto qpos = [&](int blk) { return min(8 * blk, KC - 8) + 2 * kg; };
  // block 0: rows, A parts and B values; block 1: rows and A parts
  uint2 ah = *(const uint2 *)&aph[qpos(0)];
  uint2 al = *(const uint2 *)&apl[qpos(0)];
  rpa_u4 b[TPW];
  {
    const int2 pr = *(const int2 *)&sP[qpos(0)];
    const int rb0 = pr.x * ms + lane_off, rb1 = pr.y * ms + lane_off;
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
      const uint2 v0 = *(const uint2 *)&sBw[2 * (rb0 + toff[tt])];
      const uint2 v1 = *(const uint2 *)&sBw[2 * (rb1 + toff[tt])];
      b[tt][0] = v0.x;
      b[tt][1] = v0.y;
      b[tt][2] = v1.x;
      b[tt][3] = v1.y;
    }
  }
  int2 prn = *(const int2 *)&sP[qpos(1)];
  uint2 ahn = *(const uint2 *)&aph[qpos(1)];
  uint2 aln = *(const uint2 *)&apl[qpos(1)];
  // all of them landed (lgkmcnt(0)): the loop's own waits then cover one
  // pass of reads.  Entered with these in flight, in whatever order the
  // compiler issued them, the loop began with a full LDS wait on every pass.
  __builtin_amdgcn_s_waitcnt(0xc07f);
  for (int t = 0; t < nblk; ++t) {
    rpa_u4 a1, a2;
    {
      const uint32_t hx = rot(ah.x), hy = rot(ah.y), lx = rot(al.x), ly = rot(al.y);
      a1[0] = hx; a1[1] = hx; a1[2] = hy; a1[3] = hy;
      a2[0] = lx; a2[1] = 0u; a2[2] = ly; a2[3] = 0u;
    }
    const rpa_bf8 a_1 = __builtin_bit_cast(rpa_bf8, a1), a_2 = __builtin_bit_cast(rpa_bf8, a2);
    const int rb0 = prn.x * ms + lane_off, rb1 = prn.y * ms + lane_off;
    ah = ahn;
    al = aln;
    prn = *(const int2 *)&sP[qpos(t + 2)];
    ahn = *(const uint2 *)&aph[qpos(t + 2)];
    aln = *(const uint2 *)&apl[qpos(t + 2)];
    // each tile's products, then the next block's B values of that tile
    // into the same registers: the reads are in flight during the
    // remaining tiles' products
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
      const rpa_bf8 bb = __builtin_bit_cast(rpa_bf8, b[tt]);
      acc[tt] = mfma_bf16_16x16x32(a_2, bb, acc[tt]);
      acc[tt] = mfma_bf16_16x16x32(a_1, bb, acc[tt]);
      const uint2 v0 = *(const uint2 *)&sBw[2 * (rb0 + toff[tt])];
      const uint2 v1 = *(const uint2 *)&sBw[2 * (rb1 + toff[tt])];
      b[tt][0] = v0.x;
      b[tt][1] = v0.y;
      b[tt][2] = v1.x;
      b[tt][3] = v1.y;
      __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
    }
    if (t < nissue) issue(t);
  }
  for (int t = nblk; t < nissue; ++t) issue(t);
}

// Writes the tiles of one workgroup: tile (sg, mg), lane column r -> shift
// 4 sg + (r & 3), band m 4 mg + (r >> 2), rows as the precision's layout;
// current vertices add alpha(G) rho.
template <class P, int KIND, int TPW, int NV>
__device__ __forceinline__ void col_write(const ColArgs &a, const typename P::acc (&acc)[NV][TPW], int k, int cl,
                                          const VBlock &vb, int wave, int r, int kg, const int *info, int nt,
                                          int nmg) {
  using real = typename P::real;
  using cplx = typename P::cplx;
  // The output slots (perm lookups) of all tiles are loaded first, from
  // clamped addresses, so that the lookups are in flight together.
  const PairBlock &pbk = a.pairs.b[vb.pb];
  const int nm = pbk.m_hi - pbk.m_lo;
  cplx *rho = (cplx *)a.rho, *rho2 = (cplx *)a.rho2;
  const int *shift = a.shift + (size_t)cl * kColMaxShift;
  const double *kv = a.ktot + 3 * (size_t)k;
  // (slot indices fit 31 bits: perm holds them as int, and the vertex buffer
  // of 2^31 transitions would exceed device memory)
  int out_g[TPW];
  int out_pos[TPW][2];
  int sgv[TPW];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const int T = wave + 4 * tt;
    sgv[tt] = info[2 + (T < nt ? T : 0) / nmg];
  }
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) out_g[tt] = shift[4 * sgv[tt] + (r & 3)];
  long long idx[TPW][2];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const int T = wave + 4 * tt;
    const int mo = 4 * ((T < nt ? T : 0) % nmg) + (r >> 2);
    const bool ok = T < nt && out_g[tt] >= 0 && mo < vb.ncols;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int onr = P::out_band(kg, p);
      const bool v = ok && onr < vb.nrows;
      idx[tt][p] = v ? (long long)k * a.pairs.tpk + a.pairs.off[vb.pb] + (long long)onr * nm + mo +
                           (long long)(vb.n0 - pbk.n_lo) * nm + (vb.m0 - pbk.m_lo)
                     : -1;
    }
  }
  if (a.perm) {
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int pv = a.perm[idx[tt][p] >= 0 ? idx[tt][p] : 0];
        out_pos[tt][p] = idx[tt][p] >= 0 ? pv : -1;
      }
  } else {
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
      for (int p = 0; p < 2; ++p) out_pos[tt][p] = (int)idx[tt][p];
  }
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const int iG = out_g[tt];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const long long pos = out_pos[tt][p];
      if (pos < 0) continue;
      const real xr = acc[0][tt][P::out_re(p)], xi = acc[0][tt][P::out_im(p)];
      if (KIND == 0) {
        rho[pos * a.nlf + iG] = P::make(xr, xi);
      } else {
        // j = alpha(G) rho + rho_beta, alpha = 0.5 Gcar (q + 2k + Glf(G))_pol
        const real al1 = (real)(0.5 * (((a.q[a.pol1] + 2.0 * kv[a.pol1]) + a.glf[3 * iG + a.pol1]) * a.gcar));
        const cplx j1 = P::make(al1 * xr + acc[1][tt][P::out_re(p)], al1 * xi + acc[1][tt][P::out_im(p)]);
        rho[pos * a.nlf + iG] = j1;
        if (KIND == 2) {
          const real al2 = (real)(0.5 * (((a.q[a.pol2] + 2.0 * kv[a.pol2]) + a.glf[3 * iG + a.pol2]) * a.gcar));
          rho2[pos * a.nlf + iG] = P::make(al2 * xr + acc[NV - 1][tt][P::out_re(p)],
                                           al2 * xi + acc[NV - 1][tt][P::out_im(p)]);
        } else if (rho2) {
          rho2[pos * a.nlf + iG] = j1;
        }
      }
    }
  }
}

// PIPE: the next chunk is loaded into registers while the current one runs
// (fewer waves, hidden latency) instead of loaded between the barriers.
// KIND: 0 charge, 1 current (one polarisation), 2 current (pol yz: two).
template <class P, int KIND, int TPW, int KC, int NRU, bool PIPE>
__device__ __forceinline__ void col_body(const ColArgs &a) {
  using real = typename P::real;
  using cplx = typename P::cplx;
  using acc_t = typename P::acc;
  constexpr int NV = 1 + (KIND >= 1) + (KIND == 2);  // A variants: C1, beta1 C1, beta2 C1
  // steps per pass of the product loop: two in FP32 (FP64 would pass 256
  // registers and halve the waves per SIMD)
  constexpr int SU = (sizeof(real) == 4) ? 2 : 1;
  constexpr bool SPLIT = P::split;
  static_assert(!SPLIT || KIND == 0, "the split form runs charge vertices");
  constexpr int KCA = KC + 4;  // split form: A row stride (words), 16-byte aligned rows
  extern __shared__ __align__(16) unsigned char col_raw[];
  cplx *sA = (cplx *)col_raw;         // [KC][8]   C1 of position p, band n
  uint32_t *sAw = (uint32_t *)col_raw;  // split: [hi, lo][8 bands n][KCA] BF16 pairs of C1
  int *sP = SPLIT ? (int *)(sAw + 2 * 8 * KCA) : (int *)(sA + KC * 8);  // [KC] B row of position p
  real *sBeta = (real *)(sP + KC);    // [2][KC]   beta of position p (current vertices)
  cplx *sB = (cplx *)(sBeta + 2 * KC);  // [rows][ms] C2 of the targets, bands m
  uint32_t *sBw = (uint32_t *)(sP + KC);  // split: [rows][ms][hi, lo] BF16 pairs of C2

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int L = blockIdx.x;
  const int xcd = L & 7;
  const int slot = L >> 3;
  const int k = xcd + 8 * (slot / a.blocks);
  if (k >= a.ntot) return;  // whole workgroup leaves together
  const int rest = slot % a.blocks;
  const int cl = rest % a.ncls;
  const VBlock vb = a.vb[rest / a.ncls];

  const int K1 = a.k1[k] - 1, K2 = a.k2[k] - 1;
  const int lim2 = a.lim_ngk ? a.ngk[K2] : 0x7fffffff;
  const int g1 = (a.nbg == 2 && vb.n0 >= a.blo1) ? 1 : 0;
  const int g2 = (a.nbg == 2 && vb.m0 >= a.blo1) ? 1 : 0;
  const int s1 = a.bw[g1], s2 = a.bw[g2];  // row strides of the C1 and C2 arrays
  const cplx *c1 = (const cplx *)a.box[g1] + (size_t)K1 * a.npwx * s1 + (vb.n0 - (g1 ? a.blo1 : 0));
  const cplx *c2 = (const cplx *)a.box[g2] + (size_t)K2 * a.npwx * s2 + (vb.m0 - (g2 ? a.blo1 : 0));
  const int *info = a.cls + (size_t)cl * kColInfo;
  const int *bmap = a.bmap + (size_t)a.g0idx[k] * a.nrows + info[0];
  const int *prow = a.prow + (size_t)cl * a.npos;
  const int *crow = a.crow + (size_t)cl * (a.nchunk + 1);
  const int nmg = (vb.ncols + 3) >> 2;
  const int nt = info[1] * nmg;  // tiles of this class and block

  // ---- staging roles
  const int sa_p = tid >> 3, sa_n = tid & 7;       // C1: position sa_p + 32u, band sa_n
  const int sb_r = tid >> 2, sb_m = tid & 3;  // k+q: row sb_r + 64u, bands sb_m + 4e
  ColStage<P, KC, NRU> st;
  auto load = [&](int ch) {
    const int p0 = a.chunk[ch], np = a.chunk[ch + 1] - p0;
    const int r0 = crow[ch], nr = crow[ch + 1] - r0;
    st.np = np;
    st.nr = nr;
    const bool nok = sa_n < vb.nrows;
    const int nn = nok ? sa_n : 0;
    int slot_u[KC / 32];
#pragma unroll
    for (int u = 0; u < KC / 32; ++u) slot_u[u] = a.slot[p0 + min(sa_p + 32 * u, np - 1)];
    const int tp = min(tid, np - 1);
    st.prow = prow[p0 + tp];
    if (KIND >= 1) {
#pragma unroll
      for (int v = 0; v < NV - 1; ++v) st.beta[v] = (real)a.beta[(size_t)v * a.npos + p0 + tp];
    }
    int jj[NRU];
#pragma unroll
    for (int u = 0; u < NRU; ++u) jj[u] = bmap[r0 + min(sb_r + 64 * u, nr - 1)];
    unsigned am = 0;
#pragma unroll
    for (int u = 0; u < KC / 32; ++u) {
      st.a[u] = c1[(size_t)slot_u[u] * s1 + nn];
      am |= (nok && sa_p + 32 * u < np) ? (1u << u) : 0u;
    }
    st.amask = am;
    // row groups u with no row in the chunk (uniform test) issue no loads;
    // their registers are masked, never written with a fill
    unsigned bm = 0;
#pragma unroll
    for (int u = 0; u < NRU; ++u) {
      {
        const bool ok = sb_r + 64 * u < nr && jj[u] >= 0 && jj[u] < lim2;
        const int row = ok ? jj[u] : 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const bool eok = sb_m + 4 * e < vb.ncols;
          st.b[u][e] = c2[(size_t)row * s2 + (eok ? sb_m + 4 * e : 0)];
          bm |= (ok && eok) ? (1u << (4 * u + e)) : 0u;
        }
      }
    }
    st.bmask = bm;
  };
  // zero fill per component (a select of whole vector values keeps the
  // staging struct in scratch memory for double2)
  auto keep = [](bool ok, cplx z) { return P::make(ok ? z.x : (real)0, ok ? z.y : (real)0); };
  auto store = [&]() {
    if constexpr (SPLIT) {
#pragma unroll
      for (int u = 0; u < KC / 32; ++u) {
        uint32_t h, l;
        split_pair(keep((st.amask >> u) & 1u, st.a[u]), h, l);
        sAw[sa_n * KCA + sa_p + 32 * u] = h;
        sAw[(8 + sa_n) * KCA + sa_p + 32 * u] = l;
      }
      if (tid < KC) sP[tid] = tid < st.np ? st.prow : 0;
#pragma unroll
      for (int u = 0; u < NRU; ++u) {
        const int rr = sb_r + 64 * u;
        if (rr < st.nr) {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (sb_m + 4 * e < a.ms) {
              uint32_t h, l;
              split_pair(keep((st.bmask >> (4 * u + e)) & 1u, st.b[u][e]), h, l);
              *(uint2 *)&sBw[2 * (rr * a.ms + sb_m + 4 * e)] = make_uint2(h, l);
            }
        }
      }
      return;
    }
#pragma unroll
    for (int u = 0; u < KC / 32; ++u) sA[(sa_p + 32 * u) * 8 + sa_n] = keep((st.amask >> u) & 1u, st.a[u]);
    const bool pok = tid < st.np;
    if (tid < KC) sP[tid] = pok ? st.prow : 0;
    if (KIND >= 1 && tid < KC) {
#pragma unroll
      for (int v = 0; v < NV - 1; ++v) sBeta[v * KC + tid] = pok ? st.beta[v] : (real)0;
    }
#pragma unroll
    for (int u = 0; u < NRU; ++u) {
      const int rr = sb_r + 64 * u;
      if (rr < st.nr) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (sb_m + 4 * e < a.ms) sB[rr * a.ms + sb_m + 4 * e] = keep((st.bmask >> (4 * u + e)) & 1u, st.b[u][e]);
      }
    }
  };

  // ---- matrix-core roles
  const int r = lane & 15, kg = lane >> 4;
  const int band_r = P::arow_band(r);
  const bool im_row = P::arow_im(r);
  // Row stride ms: the 7 rows x 4 bands one step reads fall in distinct banks
  const int lane_off = (r & 3) * a.ms + (r >> 2);  // shift r&3, band m r>>2
  int toff[TPW];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const int T = wave + 4 * tt;
    toff[tt] = (T < nt) ? 4 * info[2 + T / nmg] * a.ms + 4 * (T % nmg) : 0;
  }
  acc_t acc[NV][TPW];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) acc[v][tt] = P::zero();

  if (PIPE) load(0);
  for (int ch = 0; ch < a.nchunk; ++ch) {
    if (!PIPE) load(ch);
    __syncthreads();  // the previous chunk's reads are done
    store();
    __syncthreads();
    // in flight during this chunk's products (after the last chunk a reload
    // of it, never stored: no branch around the loads)
    if (PIPE) load(min(ch + 1, a.nchunk - 1));
    const int nsteps = (a.chunk[ch + 1] - a.chunk[ch] + 3) >> 2;
    // A variants of the position's C1 value: real embedding row, scaled by beta
    auto avals = [&](cplx z, const real *bt, real *are, real *aim) {
      are[0] = im_row ? -z.y : z.x;
      aim[0] = im_row ? z.x : z.y;
#pragma unroll
      for (int v = 1; v < NV; ++v) {
        are[v] = bt[v - 1] * are[0];
        aim[v] = bt[v - 1] * aim[0];
      }
    };
    // every wave runs TPW tiles; tiles past nt read row group 0 and their
    // results are dropped (no branch in the loop, so the LDS reads of all
    // tiles are issued ahead of the products)
    if constexpr (SPLIT) {
      split_products<TPW, KC>(acc[0], sAw, sP, sBw, (a.chunk[ch + 1] - a.chunk[ch] + 7) >> 3, a.ms, lane_off, toff,
                              band_r, im_row, kg);
    } else if constexpr (SU == 2) {
      // two steps per pass: the LDS reads of both steps' A values and B rows
      // are issued together, so one latency chain feeds 2 x TPW x NV x 2
      // products.  An odd last step reads positions past the chunk, whose A
      // values are zero (at most position 4 * 31 + 3 < KC).
      for (int t = 0; t < nsteps; t += 2) {
        int prw[2];
        cplx za[2];
        real bt[2][2] = {{(real)0, (real)0}, {(real)0, (real)0}};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int p = 4 * (t + h) + kg;
          prw[h] = sP[p];
          za[h] = sA[p * 8 + band_r];
#pragma unroll
          for (int v = 0; v < NV - 1; ++v) bt[h][v] = sBeta[v * KC + p];
        }
        cplx zb[2][TPW];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const cplx *brow = sB + prw[h] * a.ms + lane_off;
#pragma unroll
          for (int tt = 0; tt < TPW; ++tt) zb[h][tt] = brow[toff[tt]];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          real are[NV], aim[NV];
          avals(za[h], bt[h], are, aim);
#pragma unroll
          for (int tt = 0; tt < TPW; ++tt) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
              acc[v][tt] = P::mma(are[v], zb[h][tt].x, acc[v][tt]);
              acc[v][tt] = P::mma(aim[v], zb[h][tt].y, acc[v][tt]);
            }
          }
        }
      }
    } else {
      // one step per pass; the position's A value, beta and B row are read
      // one step ahead
      int prow_c = sP[kg];
      cplx za_c = sA[kg * 8 + band_r];
      real bt_c[2] = {(real)0, (real)0};
#pragma unroll
      for (int v = 0; v < NV - 1; ++v) bt_c[v] = sBeta[v * KC + kg];
      for (int t = 0; t < nsteps; ++t) {
        const int pn = min(4 * (t + 1) + kg, KC - 1);
        const int prow_n = sP[pn];
        const cplx za_n = sA[pn * 8 + band_r];
        real bt_n[2] = {(real)0, (real)0};
#pragma unroll
        for (int v = 0; v < NV - 1; ++v) bt_n[v] = sBeta[v * KC + pn];
        const cplx *brow = sB + prow_c * a.ms + lane_off;
        real are[NV], aim[NV];
        avals(za_c, bt_c, are, aim);
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt) {
          const cplx zb = brow[toff[tt]];
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            acc[v][tt] = P::mma(are[v], zb.x, acc[v][tt]);
            acc[v][tt] = P::mma(aim[v], zb.y, acc[v][tt]);
          }
        }
        prow_c = prow_n;
        za_c = za_n;
#pragma unroll
        for (int v = 0; v < NV - 1; ++v) bt_c[v] = bt_n[v];
      }
    }
  }

  col_write<P, KIND, TPW, NV>(a, acc, k, cl, vb, wave, r, kg, info, nt, nmg);
}

template <class P, int KIND, int TPW, int KC, int NRU, bool PIPE>
__global__ __launch_bounds__(256) void k_vertex_col(ColArgs a) {
  col_body<P, KIND, TPW, KC, NRU, PIPE>(a);
}

// The split form (mixed charge vertices) with two waves per SIMD held: the
// batched B values of all tiles would otherwise take the register count past
// 256 and one wave per SIMD.
template <int TPW, int KC, int NRU, bool PIPE>
__global__ __launch_bounds__(256, 2) void k_vertex_col_split(ColArgs a) {
  col_body<ColBf3, 0, TPW, KC, NRU, PIPE>(a);
}

// The split form with the k+q rows staged by LDS DMA (buffer loads to LDS,
// 16 bytes a lane).  The k+q side reads a copy of its box array already
// split into (hi, lo) BF16 words, so a staged row passes through no
// registers and no conversion.  Each chunk's rows are requested one chunk
// ahead into the other of two LDS buffers and land while the current chunk's
// products run; the C1 values and position rows of the next chunk come one
// chunk ahead in registers, and the slot and row indices they are addressed
// by two chunks ahead, so no load of an iteration waits on another.
//   * Buffer loads, not global ones: the compiler takes a global load to LDS
//     for a flat access that may also count on the LDS counter, and while
//     one is in flight it waits for every LDS read before any product (one
//     full LDS drain per block of 8 positions).
//   * The chunk bounds sit in registers, entry lane + 64 q of register q,
//     read with readlane: an LDS or global read of them in the loop waited
//     behind the other workgroup's product reads or the loads in flight.
//   * The two buffers are distinct LDS variables and each half of the
//     unrolled loop names its own, so the compiler's wait for LDS DMA stores
//     stays off the other buffer's reads.
// Unit u (16 bytes: values 2 part, 2 part + 1 of a row) of a chunk is row
// u / (ms / 2), part u % (ms / 2): a buffer holds [rows][ms][hi, lo] words as
// the register-staged form does, and the 64 lanes of one request fill 64
// consecutive units.  Absent rows read past the buffer resource's range,
// which loads zeros; values of a row past the block's bands come from the
// next bands or row of the copy and meet only output columns that are
// dropped.  Products and results as the split form, so the results are
// bitwise those of k_vertex_col_split with the same chunks.
constexpr int kColDmaUnits = 8 * 256;  // 16-byte units per LDS buffer (32 KB)
constexpr int kColDmaChunks = 255;     // chunks of one plan (4 bound registers)

template <int TPW>
__global__ __launch_bounds__(256, 2) void k_vertex_col_dma(ColArgs a) {
  constexpr int KC = kKcF32, KCA = KC + 4, NA = KC / 32, NBJ = kColDmaUnits / 256;
  __shared__ __align__(16) uint32_t sAw[2 * 8 * KCA];
  __shared__ __align__(16) int sP[KC];
  __shared__ __align__(16) uint32_t sB0[4 * kColDmaUnits];
  __shared__ __align__(16) uint32_t sB1[4 * kColDmaUnits];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int L = blockIdx.x;
  const int xcd = L & 7;
  const int slot = L >> 3;
  const int k = xcd + 8 * (slot / a.blocks);
  if (k >= a.ntot) return;  // whole workgroup leaves together
  const int rest = slot % a.blocks;
  const int cl = rest % a.ncls;
  const VBlock vb = a.vb[rest / a.ncls];

  const int K1 = a.k1[k] - 1, K2 = a.k2[k] - 1;
  const int g1 = (a.nbg == 2 && vb.n0 >= a.blo1) ? 1 : 0;
  const int s1 = a.bw[g1];
  const float2 *c1 = (const float2 *)a.box[g1] + (size_t)K1 * a.npwx * s1 + (vb.n0 - (g1 ? a.blo1 : 0));
  // k+q rows: a buffer resource from this k's first row at the block's first
  // band, its range the rest of the copy (at most 2^31 - 1 bytes; the rows
  // read lie within the first npwx rows and 20 values past them)
  const long long c2off = (long long)K2 * a.npwx * a.bww + (vb.m0 - (a.nbg == 2 ? a.blo1 : 0));
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(a.boxw + 2 * c2off), (short)0, (int)min(8 * (a.boxw_nv - c2off), 0x7fffffffLL), 0x00020000);
  const int *info = a.cls + (size_t)cl * kColInfo;
  const int *bmap = a.bmap + (size_t)a.g0idx[k] * a.nrows + info[0];
  const int *prow = a.prow + (size_t)cl * a.npos;
  const int *crow = a.crow + (size_t)cl * (a.nchunk + 1);
  const int nmg = (vb.ncols + 3) >> 2;
  const int nt = info[1] * nmg;
  const int hu = a.ms >> 1;  // 16-byte units per staged row

  // chunk bounds: entry i of chunk / crow in lane i & 63 of register i >> 6
  int chv[4], crv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = min(lane + 64 * q, a.nchunk);
    chv[q] = a.chunk[i];
    crv[q] = crow[i];
  }
  auto bound = [&](const int (&v)[4], int i) {
    const int q = i >> 6;
    const int x = q == 0 ? v[0] : q == 1 ? v[1] : q == 2 ? v[2] : v[3];
    return __builtin_amdgcn_readlane(x, i & 63);
  };

  // ---- staging roles: C1 positions sa_p + 32 u, band sa_n; DMA unit tid + 256 j
  const int sa_p = tid >> 3, sa_n = tid & 7;
  const bool nok = sa_n < vb.nrows;
  const int nn = nok ? sa_n : 0;
  int urow[NBJ];
#pragma unroll
  for (int j = 0; j < NBJ; ++j) urow[j] = (tid + 256 * j) / hu;
  int sl[NA];          // C1 slots of the next chunk whose values are loaded
  int bm[NBJ];         // box rows of the next chunk whose rows are requested
  float2 av[NA];       // C1 values of the chunk stored next
  unsigned am = 0;     // bit u: av[u] present
  int pr = 0, npc = 0; // row of position tid, positions of that chunk
  auto load_idx = [&](int ch) {
    const int p0 = bound(chv, ch), np = bound(chv, ch + 1) - p0;
    const int r0 = bound(crv, ch), nr = bound(crv, ch + 1) - r0;
#pragma unroll
    for (int u = 0; u < NA; ++u) sl[u] = a.slot[p0 + min(sa_p + 32 * u, np - 1)];
#pragma unroll
    for (int j = 0; j < NBJ; ++j) bm[j] = bmap[r0 + min(urow[j], nr - 1)];
  };
  auto load_a = [&](int ch) {
    const int p0 = bound(chv, ch), np = bound(chv, ch + 1) - p0;
    unsigned m = 0;
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      av[u] = c1[(size_t)sl[u] * s1 + nn];
      m |= (nok && sa_p + 32 * u < np) ? (1u << u) : 0u;
    }
    am = m;
    pr = prow[p0 + min(tid, np - 1)];
    npc = np;
  };
  // requests j of this wave with (j >> 1) == i, i.e. two of them (of a
  // chunk of nu units); requests of whole waves past the units are skipped,
  // lanes of the last request past them fill units of the buffer that no
  // read reaches
  auto dma = [&](int i, int nu, uint32_t *buf) {
#pragma unroll
    for (int j = 0; j < NBJ; ++j) {
      const int ub = 256 * j + 64 * wave;
      if ((j >> 1) == i && ub < nu) {
        const bool ok = ub + lane < nu && bm[j] >= 0;
        const int part = tid + 256 * j - urow[j] * hu;
        const int off = ok ? 8 * (bm[j] * a.bww + 2 * part) : (int)0x80000000;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (void __attribute__((address_space(3))) *)(buf + 4 * ub), 16,
                                                 off, 0, 0, 0);
      }
    }
  };
  auto units = [&](int ch) { return (bound(crv, ch + 1) - bound(crv, ch)) * hu; };
  auto store_a = [&]() {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      uint32_t h, l;
      const bool ok = (am >> u) & 1u;
      split_pair(make_float2(ok ? av[u].x : 0.f, ok ? av[u].y : 0.f), h, l);
      sAw[sa_n * KCA + sa_p + 32 * u] = h;
      sAw[(8 + sa_n) * KCA + sa_p + 32 * u] = l;
    }
    if (tid < KC) sP[tid] = tid < npc ? pr : 0;
  };

  // ---- matrix-core roles (as the split form)
  const int r = lane & 15, kg = lane >> 4;
  const int band_r = r >> 1;
  const bool im_row = (r & 1) != 0;
  const int lane_off = (r & 3) * a.ms + (r >> 2);
  int toff[TPW];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const int T = wave + 4 * tt;
    toff[tt] = (T < nt) ? 4 * info[2 + T / nmg] * a.ms + 4 * (T % nmg) : 0;
  }
  rpa_f4 acc[1][TPW];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) acc[0][tt] = rpa_f4{0.f, 0.f, 0.f, 0.f};

  // one chunk: its C1 values into LDS, then the products from buffer cur
  // with the next chunk's requests and loads spread over their first blocks
  // (two requests a block, then the C1 loads, then the indices): issued
  // together they stalled every wave of the CU on the memory pipeline
  // (about 4700 cycles a chunk at N_G = 51)
  auto step = [&](int ch, const uint32_t *cur, uint32_t *nxt) {
    __builtin_amdgcn_s_waitcnt(0);  // this chunk's rows and values have landed
    __syncthreads();                // ... in every wave; the last products are done
    store_a();
    __syncthreads();
    const int np = npc;
    const bool more = ch + 1 < a.nchunk, more2 = ch + 2 < a.nchunk;
    const int nu = more ? units(ch + 1) : 0;
    auto issue = [&](int i) {
      if (i < NBJ / 2) {
        dma(i, nu, nxt);
      } else if (i == NBJ / 2) {
        if (more) load_a(ch + 1);
      } else if (more2) {
        load_idx(ch + 2);
      }
    };
    split_products<TPW, KC>(acc[0], sAw, sP, cur, (np + 7) >> 3, a.ms, lane_off, toff, band_r, im_row, kg, issue,
                            NBJ / 2 + 2);
  };
  load_idx(0);
  for (int i = 0; i < NBJ / 2; ++i) dma(i, units(0), sB0);
  load_a(0);
  if (a.nchunk > 1) load_idx(1);
  for (int ch = 0; ch < a.nchunk; ch += 2) {
    step(ch, sB0, sB1);
    if (ch + 1 < a.nchunk) step(ch + 1, sB1, sB0);
  }
  col_write<ColF32, 0, TPW, 1>(a, acc, k, cl, vb, wave, r, kg, info, nt, nmg);
}

template <class P, int KIND, int KC, int NRU, bool PIPE>
static const void *col_kernel(int tpw) {
  switch (tpw) {
    case 1: return (const void *)k_vertex_col<P, KIND, 1, KC, NRU, PIPE>;
    case 2: return (const void *)k_vertex_col<P, KIND, 2, KC, NRU, PIPE>;
    case 4: return (const void *)k_vertex_col<P, KIND, 4, KC, NRU, PIPE>;
    case 6: return (const void *)k_vertex_col<P, KIND, 6, KC, NRU, PIPE>;
    case 8: return (const void *)k_vertex_col<P, KIND, 8, KC, NRU, PIPE>;
    case 10: return (const void *)k_vertex_col<P, KIND, 10, KC, NRU, PIPE>;
    case 13: return (const void *)k_vertex_col<P, KIND, 13, KC, NRU, PIPE>;
    default: return (const void *)k_vertex_col<P, KIND, 16, KC, NRU, PIPE>;
  }
}

template <class P, int KIND, int KC, int NRU, bool PIPE>
static void launch_col_p(rpa_ctx *c, const ColArgs &a, int tpw, dim3 grid) {
  const size_t lds = c->col.lds;
  switch (tpw) {
    case 1: hipLaunchKernelGGL((k_vertex_col<P, KIND, 1, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 2: hipLaunchKernelGGL((k_vertex_col<P, KIND, 2, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 4: hipLaunchKernelGGL((k_vertex_col<P, KIND, 4, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 6: hipLaunchKernelGGL((k_vertex_col<P, KIND, 6, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 8: hipLaunchKernelGGL((k_vertex_col<P, KIND, 8, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 10: hipLaunchKernelGGL((k_vertex_col<P, KIND, 10, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 13: hipLaunchKernelGGL((k_vertex_col<P, KIND, 13, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    default: hipLaunchKernelGGL((k_vertex_col<P, KIND, 16, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
  }
}

template <int KC, int NRU, bool PIPE>
static const void *split_kernel(int tpw) {
  switch (tpw) {
    case 1: return (const void *)k_vertex_col_split<1, KC, NRU, PIPE>;
    case 2: return (const void *)k_vertex_col_split<2, KC, NRU, PIPE>;
    case 4: return (const void *)k_vertex_col_split<4, KC, NRU, PIPE>;
    case 6: return (const void *)k_vertex_col_split<6, KC, NRU, PIPE>;
    case 8: return (const void *)k_vertex_col_split<8, KC, NRU, PIPE>;
    case 10: return (const void *)k_vertex_col_split<10, KC, NRU, PIPE>;
    case 13: return (const void *)k_vertex_col_split<13, KC, NRU, PIPE>;
    default: return (const void *)k_vertex_col_split<16, KC, NRU, PIPE>;
  }
}

template <int KC, int NRU, bool PIPE>
static void launch_split_p(rpa_ctx *c, const ColArgs &a, int tpw, dim3 grid) {
  const size_t lds = c->col.lds;
  switch (tpw) {
    case 1: hipLaunchKernelGGL((k_vertex_col_split<1, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 2: hipLaunchKernelGGL((k_vertex_col_split<2, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 4: hipLaunchKernelGGL((k_vertex_col_split<4, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 6: hipLaunchKernelGGL((k_vertex_col_split<6, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 8: hipLaunchKernelGGL((k_vertex_col_split<8, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 10: hipLaunchKernelGGL((k_vertex_col_split<10, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    case 13: hipLaunchKernelGGL((k_vertex_col_split<13, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
    default: hipLaunchKernelGGL((k_vertex_col_split<16, KC, NRU, PIPE>), grid, dim3(256), lds, c->stream, a); break;
  }
}

static void launch_dma(rpa_ctx *c, const ColArgs &a, int tpw, dim3 grid) {
  switch (tpw) {
    case 1: hipLaunchKernelGGL((k_vertex_col_dma<1>), grid, dim3(256), 0, c->stream, a); break;
    case 2: hipLaunchKernelGGL((k_vertex_col_dma<2>), grid, dim3(256), 0, c->stream, a); break;
    case 4: hipLaunchKernelGGL((k_vertex_col_dma<4>), grid, dim3(256), 0, c->stream, a); break;
    case 6: hipLaunchKernelGGL((k_vertex_col_dma<6>), grid, dim3(256), 0, c->stream, a); break;
    case 8: hipLaunchKernelGGL((k_vertex_col_dma<8>), grid, dim3(256), 0, c->stream, a); break;
    case 10: hipLaunchKernelGGL((k_vertex_col_dma<10>), grid, dim3(256), 0, c->stream, a); break;
    case 13: hipLaunchKernelGGL((k_vertex_col_dma<13>), grid, dim3(256), 0, c->stream, a); break;
    default: hipLaunchKernelGGL((k_vertex_col_dma<16>), grid, dim3(256), 0, c->stream, a); break;
  }
}

// Kernel of a (precision, kind, staging) choice; current vertices always
// load the next chunk during the products.
template <class P, int KC, int NRU>
static const void *col_kernel_of(int kind, bool pipe, int tpw) {
  if (kind == 1) return col_kernel<P, 1, KC, NRU, true>(tpw);
  if (kind == 2) return col_kernel<P, 2, KC, NRU, true>(tpw);
  return pipe ? col_kernel<P, 0, KC, NRU, true>(tpw) : col_kernel<P, 0, KC, NRU, false>(tpw);
}

template <class P, int KC, int NRU>
static void launch_col_of(rpa_ctx *c, const ColArgs &a, int kind, bool pipe, int tpw, dim3 grid) {
  if (kind == 1)
    launch_col_p<P, 1, KC, NRU, true>(c, a, tpw, grid);
  else if (kind == 2)
    launch_col_p<P, 2, KC, NRU, true>(c, a, tpw, grid);
  else if (pipe)
    launch_col_p<P, 0, KC, NRU, true>(c, a, tpw, grid);
  else
    launch_col_p<P, 0, KC, NRU, false>(c, a, tpw, grid);
}

// Mixed charge vertices run the split BF16 form unless RPA_VERTEX_COL_SPLIT=0
// asks for the FP32 matrix cores.
static bool col_split(bool mixed, int kind) {
  if (!mixed || kind != 0) return false;
  const char *env = std::getenv("RPA_VERTEX_COL_SPLIT");
  return !(env && env[0] == '0');
}

// 1 when RPA_VERTEX_FORM=col (or the autotuner) asks for this form
// regardless of shape
static int vertex_form_forced(const rpa_ctx *c) {
  const char *env = std::getenv("RPA_VERTEX_FORM");
  return (env ? std::string(env) == "col" : c->vform == "col") ? 1 : 0;
}

static int tpw_of(int tiles) {
  const int w = (tiles + 3) / 4;
  for (int t : {1, 2, 4, 6, 8, 10, 13, 16})
    if (w <= t) return t;
  return 0;
}


void free_col_plan(rpa_ctx *c) {
  ColPlan &P = c->col;
  for (void *b : {(void *)P.d_slot, (void *)P.d_prow, (void *)P.d_chunk, (void *)P.d_crow, (void *)P.d_bmap,
                  (void *)P.d_shift, (void *)P.d_cls, (void *)P.d_beta})
    if (b) (void)hipFree(b);
  P = ColPlan{};
}

// Column-box basis (see rpa_ctx::d_box): row R of k holds the coefficients of
// global G index gidx[R] when that index is below ngk(k), else zero.
template <class T>
__global__ void k_build_box_k(const T *evc, const int *ngk, const int *gidx, int npwx, int nband, int nrow, int nbg,
                              int blo1, T *box0, T *box1) {
  const int k = blockIdx.y;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nrow * nband) return;
  const int R = idx / nband, b = idx - R * nband;
  const int g = gidx[R];
  T v;
  if (g >= 0 && g < ngk[k]) {
    v = evc[((size_t)k * npwx + g) * nband + b];
  } else {
    v.x = 0;
    v.y = 0;
  }
  if (nbg == 2 && b >= blo1)
    box1[((size_t)k * npwx + R) * (nband - blo1) + (b - blo1)] = v;
  else
    box0[((size_t)k * npwx + R) * (nbg == 2 ? blo1 : nband) + b] = v;
}

// Builds the column-box basis when it is missing or its band split no longer
// matches the band blocks; leaves box_ready false (the plan then indexes the
// global order) when the columns' z ranges need more rows than npwx.
static int ensure_box(rpa_ctx *c) {
  const int nb = c->d.nband, nv = c->p.nval;
  bool two = nv > 0 && nv < nb;
  for (int i = 0; i < c->pairs.nblocks && two; ++i) {
    const PairBlock &b = c->pairs.b[i];
    const bool n_ok = b.n_hi <= nv || b.n_lo >= nv, m_ok = b.m_hi <= nv || b.m_lo >= nv;
    two = n_ok && m_ok;
  }
  const int nbg = two ? 2 : 1, blo1 = two ? nv : nb;
  if (c->box_ready && c->box_nbg == nbg && c->box_blo1 == blo1) return 0;
  c->box_ready = false;
  c->boxw_ready = false;
  int npw = 0;
  for (int k = 0; k < c->d.nki; ++k) npw = std::max(npw, c->ngk[k]);
  npw = std::min(npw, c->ngm);
  const int *mg = c->mill_g.data();
  std::vector<std::array<int, 4>> cols;  // (mx, my, zmin, zmax)
  {
    std::vector<int> ord(npw);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int x, int y) {
      for (int d = 0; d < 3; ++d)
        if (mg[3 * x + d] != mg[3 * y + d]) return mg[3 * x + d] < mg[3 * y + d];
      return false;
    });
    for (int i : ord) {
      const int *m = mg + 3 * i;
      if (cols.empty() || cols.back()[0] != m[0] || cols.back()[1] != m[1])
        cols.push_back({m[0], m[1], m[2], m[2]});
      else
        cols.back()[3] = m[2];
    }
  }
  std::vector<int> base(cols.size() + 1, 0);
  for (size_t u = 0; u < cols.size(); ++u) base[u + 1] = base[u] + cols[u][3] - cols[u][2] + 1;
  const int nrow = base.back();
  if (nrow > c->d.npwx) return 0;  // holes in the columns: keep the global order
  std::vector<int> gidx(nrow, -1);
  c->h_rbox.assign((size_t)c->gbox.size(), -1);
  for (int i = 0; i < npw; ++i) {
    const int *m = mg + 3 * i;
    const auto it = std::lower_bound(cols.begin(), cols.end(), std::array<int, 4>{m[0], m[1], INT_MIN, INT_MIN});
    const int u = (int)(it - cols.begin());
    const int R = base[u] + m[2] - cols[u][2];
    gidx[R] = i;
    c->h_rbox[c->gbox.index(m[0], m[1], m[2])] = R;
  }
  const bool mixed = (c->d.precision == RPA_MIXED);
  const size_t csz = mixed ? sizeof(float2) : sizeof(double2);
  const size_t rows = (size_t)c->d.nki * c->d.npwx;
  int e = 0;
  unsigned char *b0 = (unsigned char *)c->d_box[0], *b1 = (unsigned char *)c->d_box[1];
  if (!e) e = ensure(c, b0, c->cap_box[0], rows * (nbg == 2 ? blo1 : nb) * csz);
  if (!e && nbg == 2) e = ensure(c, b1, c->cap_box[1], rows * (nb - blo1) * csz);
  c->d_box[0] = b0;
  c->d_box[1] = b1;
  if (!e) e = ensure(c, c->d_box_gidx, c->cap_box_gidx, (size_t)nrow);
  if (e) return e;
  const Upload up[] = {{c->d_box_gidx, gidx.data(), sizeof(int) * gidx.size()}};
  if (int e2 = upload(c, up, 1)) return e2;
  const long long n = (long long)nrow * nb;
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)c->d.nki);  // grid.y: k
  if (mixed)
    hipLaunchKernelGGL(k_build_box_k<float2>, grid, dim3(256), 0, c->stream, (const float2 *)c->d_evc32, c->d_ngk,
                       c->d_box_gidx, c->d.npwx, nb, nrow, nbg, blo1, (float2 *)c->d_box[0], (float2 *)c->d_box[1]);
  else
    hipLaunchKernelGGL(k_build_box_k<double2>, grid, dim3(256), 0, c->stream, (const double2 *)c->d_evc, c->d_ngk,
                       c->d_box_gidx, c->d.npwx, nb, nrow, nbg, blo1, (double2 *)c->d_box[0], (double2 *)c->d_box[1]);
  RPA_HIP(c, hipGetLastError());
  c->box_nbg = nbg;
  c->box_blo1 = blo1;
  c->box_rows = nrow;
  c->box_ready = true;
  return 0;
}

// Split copy of a box array of bw bands: value j of row (k, R) is the (hi,
// lo) word pair 2 (row bww + j), + 1; zero for j >= bw and for rows past the
// box (R >= nrow); nv values in all, the last ntail of them a zero tail.
__global__ void k_build_boxw(const float2 *box, int npwx, int nrow, int bw, int bww, long long nrows_all,
                             long long nv, uint32_t *out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  float2 v = make_float2(0.f, 0.f);
  if (i < nrows_all * bww) {
    const long long row = i / bww;
    const int j = (int)(i - row * bww);
    if (j < bw && (int)(row % npwx) < nrow) v = box[row * bw + j];
  }
  uint32_t h, l;
  split_pair(v, h, l);
  *(uint2 *)&out[2 * i] = make_uint2(h, l);
}

// Builds the split copy of the k+q side's box array (array 1 when the bands
// are split, else array 0) once per box.  The tail of 32 zero values covers
// reads of up to 20 values past the last row.
static int ensure_boxw(rpa_ctx *c) {
  if (c->boxw_ready) return 0;
  const int g = c->box_nbg == 2 ? 1 : 0;
  const int bw = g ? c->d.nband - c->box_blo1 : c->d.nband;
  const int bww = (bw + 1) & ~1;
  const long long rows = (long long)c->d.nki * c->d.npwx;
  const long long nv = rows * bww + 32;
  if (int e = ensure(c, c->d_boxw, c->cap_boxw, 2 * (size_t)nv)) return e;
  hipLaunchKernelGGL(k_build_boxw, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, c->stream,
                     (const float2 *)c->d_box[g], c->d.npwx, c->box_rows, bw, bww, rows, nv, c->d_boxw);
  RPA_HIP(c, hipGetLastError());
  c->boxw_bww = bww;
  c->boxw_nv = nv;
  c->boxw_ready = true;
  return 0;
}

// The plan of the current q: positions in column order, chunks that fit the
// LDS budget, in-plane classes of the local fields with their shift slots,
// and for every distinct G0 the global index of each staged k+q row.
// Leaves col.ok false (the other vertex forms run) where it does not apply.
int build_col_plan(rpa_ctx *c) {
  ColPlan &P = c->col;
  P.ok = false;
  // the column correlation needs targets G(i) + Glf + G0: no point operations
  if (c->p.vertex_map != RPA_MAP_DIRECT || c->qsym) return 0;
  const bool mixed = (c->d.precision == RPA_MIXED);
  const int kc = mixed ? kKcF32 : kKcF64;
  const size_t csz = mixed ? sizeof(float2) : sizeof(double2);
  const size_t rsz = mixed ? sizeof(float) : sizeof(double);
  int kind = 0, pol1 = 0, pol2 = 0;
  if (c->d.program == RPA_OPTABS) {
    kind = (c->p.pol == RPA_POL_YZ) ? 2 : 1;
    pol1 = (kind == 2) ? 1 : c->p.pol - 1;
    pol2 = (kind == 2) ? 2 : pol1;
  }
  const int ngd = c->p.ngd;
  const int *mg = c->mill_g.data();

  // positions: G(i), i < NGd, ordered by (mx, my, mz, i); segments: runs of
  // one column, at most kc long
  std::vector<int> ord(ngd);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
    for (int d = 0; d < 3; ++d)
      if (mg[3 * x + d] != mg[3 * y + d]) return mg[3 * x + d] < mg[3 * y + d];
    return false;
  });
  struct Seg {
    int mx, my, zlo, zhi, p0, p1;
  };
  std::vector<Seg> segs;
  for (int p = 0; p < ngd; ++p) {
    const int *m = mg + 3 * ord[p];
    if (segs.empty() || segs.back().mx != m[0] || segs.back().my != m[1] || segs.back().p1 - segs.back().p0 >= kc)
      segs.push_back(Seg{m[0], m[1], m[2], m[2], p, p + 1});
    else {
      segs.back().zhi = m[2];
      segs.back().p1 = p + 1;
    }
  }

  // in-plane classes of the local fields
  struct Cls {
    int lx, ly, lzmin, lzmax;
    std::vector<int> fields;
  };
  std::vector<Cls> cls;
  for (int g = 0; g < c->nlf; ++g) {
    const int *m = mg + 3 * (c->lf_idx[g] - 1);
    size_t u = 0;
    while (u < cls.size() && (cls[u].lx != m[0] || cls[u].ly != m[1])) ++u;
    if (u == cls.size()) cls.push_back(Cls{m[0], m[1], m[2], m[2], {}});
    cls[u].lzmin = std::min(cls[u].lzmin, m[2]);
    cls[u].lzmax = std::max(cls[u].lzmax, m[2]);
    cls[u].fields.push_back(g);
  }
  const int ncls = (int)cls.size();
  // many classes of few fields (|q+G| sets in all directions) stage a k+q
  // row set per class for little reuse; there the gather forms are faster
  // (measured at 4-9 fields per class)
  if (c->nlf < 12 * ncls && vertex_form_forced(c) != 1) return 0;
  // current vertices run two (yz: three) accumulator sets where the staged
  // form applies J(G, i) to each term and runs one (two): with local fields
  // in several classes the staged form measured faster (2.2x in FP64, 1.8x
  // in FP32 at N_G = 61-81 in 4-7 classes), with one class the column form
  // (FP64 14.2 vs 16.1 ms at N_G = 31)
  if (kind >= 1 && ncls > 1 && vertex_form_forced(c) != 1) return 0;
  std::vector<int> shift((size_t)ncls * kColMaxShift, -1), info((size_t)ncls * kColInfo, 0);
  int spanmax = 0, padmax = 0, tilemax = 0;
  for (int u = 0; u < ncls; ++u) {
    const int span = cls[u].lzmax - cls[u].lzmin;
    const int nslot = 4 * ((span + 4) / 4);
    if (nslot > kColMaxShift) return 0;
    for (int g : cls[u].fields) {
      const int s = mg[3 * (c->lf_idx[g] - 1) + 2] - cls[u].lzmin;
      if (shift[(size_t)u * kColMaxShift + s] >= 0) return 0;  // repeated local field
      shift[(size_t)u * kColMaxShift + s] = g;
    }
    int nsg = 0;
    for (int sg = 0; sg < nslot / 4; ++sg) {
      bool used = false;
      for (int s = 4 * sg; s < 4 * sg + 4; ++s) used = used || shift[(size_t)u * kColMaxShift + s] >= 0;
      if (used) info[(size_t)u * kColInfo + 2 + nsg++] = sg;
    }
    info[(size_t)u * kColInfo + 1] = nsg;
    spanmax = std::max(spanmax, span);
    padmax = std::max(padmax, nslot - 1 - span);
    tilemax = std::max(tilemax, nsg);
  }
  // groups of 4 bands m of the widest band block (blocks of at most 16 m)
  int nmg = 1;
  for (int b = 0; b < c->pairs.nblocks; ++b)
    nmg = std::max(nmg, (std::min(16, c->pairs.b[b].m_hi - c->pairs.b[b].m_lo) + 3) / 4);
  // every wave runs TPW tiles, so TPW follows the tiles the shape needs
  const int tpw = tpw_of(tilemax * nmg);
  if (tpw == 0) return 0;
  if (!mixed && kind == 2 && tpw >= 13) return 0;  // three FP64 accumulator sets would spill
  // k+q row stride: padded so that 7 consecutive rows of 4 bands map to
  // distinct LDS banks
  const int ms = (nmg == 1) ? 4 : (nmg <= 3) ? 12 : 20;
  // loading the next chunk during the products measured faster in both
  // precisions at N_G = 51 (FP32 3.2 vs 4.6 ms, FP64 5.3 vs 5.4 ms); the
  // split form reads every tile's B values before its products, and the
  // staging registers on top spill (1.28 vs 1.04 ms loading between the
  // barriers)
  const bool split = col_split(mixed, kind);
  bool pipe = !split;
  if (const char *env = std::getenv("RPA_VERTEX_COL_PIPE")) pipe = env[0] == '1';

  // chunks: consecutive segments, at most kc positions and rlim rows
  const size_t fixed = (size_t)kc * (8 * csz + sizeof(int) + 2 * rsz);
  size_t budget = kColLdsBudget;
  if (const char *env = std::getenv("RPA_VERTEX_COL_LDS")) budget = std::max<size_t>(fixed + 4096, 1024 * (size_t)std::atoi(env));
  budget = std::min<size_t>(budget, 160 * 1024);
  const int rmax = std::min((int)((budget - fixed) / (ms * csz)), 64 * (mixed ? kNruF32 : kNruF64));
  // the column-box basis, where it can be built (RPA_VERTEX_BOX=0: global order)
  const char *env_box = std::getenv("RPA_VERTEX_BOX");
  bool box = !(env_box && env_box[0] == '0');
  if (box) {
    if (int e2 = ensure_box(c)) return e2;
    box = c->box_ready;
  }
  // the split form stages its k+q rows by LDS DMA from the split box copy
  // (RPA_VERTEX_COL_DMA=0: through registers); its chunks fit one DMA buffer
  // whichever way they are staged, so both stagings sum alike
  const char *env_dma = std::getenv("RPA_VERTEX_COL_DMA");
  bool dma = split && box && !(env_dma && env_dma[0] == '0');
  const int rlim = split ? std::min(rmax, kColDmaUnits / (ms / 2)) : rmax;
  std::vector<int> chunk{0};
  std::vector<std::pair<int, int>> chunk_segs;  // [first seg, end seg)
  {
    int s0 = 0, npos = 0, rows = padmax;
    for (int s = 0; s < (int)segs.size(); ++s) {
      const int cnt = segs[s].p1 - segs[s].p0;
      const int ext = segs[s].zhi - segs[s].zlo + 1 + spanmax;
      if (ext + padmax > rlim) return 0;
      if (s > s0 && (npos + cnt > kc || rows + ext > rlim)) {
        chunk_segs.push_back({s0, s});
        chunk.push_back(segs[s].p0);
        s0 = s;
        npos = 0;
        rows = padmax;
      }
      npos += cnt;
      rows += ext;
    }
    chunk_segs.push_back({s0, (int)segs.size()});
    chunk.push_back(ngd);
  }
  dma = dma && (int)chunk_segs.size() <= kColDmaChunks;
  if (dma)
    if (int e2 = ensure_boxw(c)) return e2;
  const int nchunk = (int)chunk_segs.size();

  // per class: B rows of each chunk, chunk-relative row of each position
  std::vector<int> prow((size_t)ncls * ngd), crow((size_t)ncls * (nchunk + 1));
  std::vector<int> rowoff(ncls);
  int nrows = 0, rows_used = 0;
  for (int u = 0; u < ncls; ++u) {
    const int span = cls[u].lzmax - cls[u].lzmin;
    const int pad = 4 * ((span + 4) / 4) - 1 - span;
    rowoff[u] = nrows;
    info[(size_t)u * kColInfo] = nrows;
    int r = 0;
    for (int ch = 0; ch < nchunk; ++ch) {
      crow[(size_t)u * (nchunk + 1) + ch] = r;
      int rs = 0;
      for (int s = chunk_segs[ch].first; s < chunk_segs[ch].second; ++s) {
        for (int p = segs[s].p0; p < segs[s].p1; ++p)
          prow[(size_t)u * ngd + p] = rs + (mg[3 * ord[p] + 2] - segs[s].zlo);
        rs += segs[s].zhi - segs[s].zlo + 1 + span;
      }
      rs += pad;
      rows_used = std::max(rows_used, rs);
      r += rs;
    }
    crow[(size_t)u * (nchunk + 1) + nchunk] = r;
    nrows += r;
  }
  // target row of every B row (global index, or box row), per distinct G0
  const int ng0 = (int)c->g0_list.size();
  std::vector<int> bmap((size_t)ng0 * nrows, -1);
  for (int z = 0; z < ng0; ++z) {
    const int *o = mg + 3 * (c->g0_list[z] - 1);
    for (int u = 0; u < ncls; ++u) {
      const int span = cls[u].lzmax - cls[u].lzmin;
      for (int ch = 0; ch < nchunk; ++ch) {
        int row = rowoff[u] + crow[(size_t)u * (nchunk + 1) + ch];
        for (int s = chunk_segs[ch].first; s < chunk_segs[ch].second; ++s) {
          const int ext = segs[s].zhi - segs[s].zlo + 1 + span;
          for (int e = 0; e < ext; ++e, ++row) {
            const int bi = c->gbox.index(segs[s].mx + cls[u].lx + o[0], segs[s].my + cls[u].ly + o[1],
                                         segs[s].zlo + cls[u].lzmin + e + o[2]);
            bmap[(size_t)z * nrows + row] = bi >= 0 ? (box ? c->h_rbox[bi] : c->h_gbox[bi]) : -1;
          }
        }
      }
    }
  }
  std::vector<int> slot(ngd);
  for (int p = 0; p < ngd; ++p) slot[p] = box ? c->h_rbox[c->gbox.index(mg[3 * ord[p]], mg[3 * ord[p] + 1],
                                                                         mg[3 * ord[p] + 2])]
                                              : ord[p];
  // beta(i) = Gcar G(i)_pol of every position (current vertices), with the
  // cartesian component summed as the gather forms do
  std::vector<double> beta(2 * (size_t)ngd, 0.0);
  if (kind >= 1)
    for (int v = 0; v < 2; ++v) {
      const int pol = v ? pol2 : pol1;
      for (int p = 0; p < ngd; ++p) {
        const int *m = mg + 3 * ord[p];
        const double g = ((0.0 + c->kc[pol + 0] * (double)m[0]) + c->kc[pol + 3] * (double)m[1]) +
                         c->kc[pol + 6] * (double)m[2];
        beta[(size_t)v * ngd + p] = g * c->p.gcar;
      }
    }

  int e = ensure(c, P.d_slot, P.cap_slot, (size_t)ngd);
  if (!e) e = ensure(c, P.d_prow, P.cap_prow, prow.size());
  if (!e) e = ensure(c, P.d_chunk, P.cap_chunk, chunk.size());
  if (!e) e = ensure(c, P.d_crow, P.cap_crow, crow.size());
  if (!e) e = ensure(c, P.d_bmap, P.cap_bmap, std::max<size_t>(bmap.size(), 1));
  if (!e) e = ensure(c, P.d_shift, P.cap_shift, shift.size());
  if (!e) e = ensure(c, P.d_cls, P.cap_cls, info.size());
  if (!e) e = ensure(c, P.d_beta, P.cap_beta, beta.size());
  if (e) return e;
  const Upload up[] = {{P.d_slot, slot.data(), sizeof(int) * slot.size()},
                       {P.d_prow, prow.data(), sizeof(int) * prow.size()},
                       {P.d_chunk, chunk.data(), sizeof(int) * chunk.size()},
                       {P.d_crow, crow.data(), sizeof(int) * crow.size()},
                       {P.d_bmap, bmap.data(), sizeof(int) * bmap.size()},
                       {P.d_shift, shift.data(), sizeof(int) * shift.size()},
                       {P.d_cls, info.data(), sizeof(int) * info.size()},
                       {P.d_beta, beta.data(), sizeof(double) * beta.size()}};
  if (int e2 = upload(c, up, 8)) return e2;
  P.npos = ngd;
  P.nchunk = nchunk;
  P.ncls = ncls;
  P.nrows = nrows;
  P.kc = kc;
  P.tpw = tpw;
  P.ms = ms;
  P.pipe = pipe || kind >= 1;
  P.kind = kind;
  P.pol1 = pol1;
  P.pol2 = pol2;
  P.lds = fixed + (size_t)rows_used * ms * csz;
  // the split form keeps hi and lo words of each staged value (the bytes of
  // one complex FP32 value) and no beta rows
  P.split = split;
  P.box = box;
  P.dma = dma;
  if (P.lds > 64 * 1024) {
    const void *fn = P.split ? (pipe ? split_kernel<kKcF32, kNruF32, true>(tpw)
                                     : split_kernel<kKcF32, kNruF32, false>(tpw))
                     : mixed ? col_kernel_of<ColF32, kKcF32, kNruF32>(kind, pipe, tpw)
                             : col_kernel_of<ColF64, kKcF64, kNruF64>(kind, pipe, tpw);
    RPA_HIP(c, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P.lds));
  }
  P.ok = true;
  return 0;
}

int launch_vertices_col(rpa_ctx *c, const VBlock *vb, int nvb, void *rho, void *rho2, const int *perm) {
  const ColPlan &P = c->col;
  const bool mixed = (c->d.precision == RPA_MIXED);
  ColArgs a{};
  if (P.box) {
    a.box[0] = c->d_box[0];
    a.box[1] = c->d_box[1] ? c->d_box[1] : c->d_box[0];
    a.nbg = c->box_nbg;
    a.blo1 = c->box_blo1;
    a.bw[0] = c->box_nbg == 2 ? c->box_blo1 : c->d.nband;
    a.bw[1] = c->d.nband - c->box_blo1;
    a.lim_ngk = 0;
  } else {
    a.box[0] = a.box[1] = mixed ? (const void *)c->d_evc32 : (const void *)c->d_evc;
    a.nbg = 1;
    a.bw[0] = a.bw[1] = c->d.nband;
    a.lim_ngk = 1;
  }
  a.ngk = c->d_ngk;
  a.k1 = c->d_k1;
  a.k2 = c->d_k2;
  a.g0idx = c->d_g0idx;
  a.slot = P.d_slot;
  a.prow = P.d_prow;
  a.chunk = P.d_chunk;
  a.crow = P.d_crow;
  a.bmap = P.d_bmap;
  a.shift = P.d_shift;
  a.cls = P.d_cls;
  a.npos = P.npos;
  a.nchunk = P.nchunk;
  a.ncls = P.ncls;
  a.nrows = P.nrows;
  a.nband = c->d.nband;
  a.npwx = c->d.npwx;
  a.nlf = c->nlf;
  a.ntot = c->ntot;
  a.nvb = nvb;
  for (int b = 0; b < nvb; ++b) a.vb[b] = vb[b];
  a.blocks = P.ncls * nvb;
  a.pairs = c->pairs;
  a.perm = perm;
  a.rho = rho;
  a.rho2 = rho2;
  a.beta = P.d_beta;
  a.ktot = c->d_ktot;
  a.glf = c->d_glf;
  for (int i = 0; i < 3; ++i) a.q[i] = c->q[i];
  a.gcar = c->p.gcar;
  a.pol1 = P.pol1;
  a.pol2 = P.pol2;
  const long long kslots = (c->ntot + 7) / 8 * 8;
  const dim3 grid((unsigned)(kslots * a.blocks));
  a.ms = P.ms;
  // the DMA form takes the k+q rows of every block from the split copy at
  // 16-byte aligned values
  bool dma = P.dma && c->boxw_ready;
  for (int b = 0; b < nvb && dma; ++b) {
    const int m_off = vb[b].m0 - (c->box_nbg == 2 ? c->box_blo1 : 0);
    dma = m_off >= 0 && (m_off & 1) == 0;
  }
  if (dma) {
    a.boxw = c->d_boxw;
    a.boxw_nv = c->boxw_nv;
    a.bww = c->boxw_bww;
    launch_dma(c, a, P.tpw, grid);
  } else if (P.split) {
    if (P.pipe)
      launch_split_p<kKcF32, kNruF32, true>(c, a, P.tpw, grid);
    else
      launch_split_p<kKcF32, kNruF32, false>(c, a, P.tpw, grid);
  } else if (mixed)
    launch_col_of<ColF32, kKcF32, kNruF32>(c, a, P.kind, P.pipe, P.tpw, grid);
  else
    launch_col_of<ColF64, kKcF64, kNruF64>(c, a, P.kind, P.pipe, P.tpw, grid);
  RPA_HIP(c, hipGetLastError());
  return 0;
}

}  // namespace rpa
