// Plane-wave vertices for every transition (k, n, m) of the current q.
//
//   charge  rho_nm(k,q,G) = sum_i conj(C1_n(i)) C2_m(map(G,i))
//           (Sloss.f90:883-952 genMnmK1K2; vertices.f90:88-189 _v2)
//   current j_nm(k,q,G)   = sum_i 0.5 conj(C1_n(i)) J(G,i) C2_m(map(G,i)),
//           J = Gcar (q + 2k + Glf(G) + 2 G(i))_pol
//           (Optabs.f90:1354-1472 genStrujniVrhovi; vertices.f90:293-394)
//
// map(G,i) is the reference's Gfast: the position of G(i) + Glf(G) + G0 in the
// k+q coefficient list (Sloss.f90:907-946), found with exact Miller arithmetic
// in a dense Miller box instead of a linear search (R = RI = I).  In the direct
// map the target is a global G index that depends on k only through G0, so
// k_direct_map tabulates it once per distinct G0 of the q; each k then only
// checks it against ngk(K2).
//
// For one (k, G) the sums over all band pairs form a complex product
// [conj(C1)]_(n x i) [C2 o map]_(i x m).  It runs on the FP64 matrix cores as
// a real 16x16x4 MFMA: rows are (n, re/im) -- rows 0..7 give Re for 8 bands n,
// rows 8..15 give Im -- the K index is (i, re/im), columns are 16 bands m:
//   A[n][(i,0)] =  C1r   A[n][(i,1)] = C1i   A[8+n][(i,0)] = -C1i  A[8+n][(i,1)] = C1r
//   B[(i,0)][m] =  C2r(map(i))*J           B[(i,1)][m] = C2i(map(i))*J
// In mixed precision the same contraction runs in complex FP32 on the FP32
// matrix cores (P = PrecF32); only the row order of A and of the result
// differs, because the FP32 instruction returns rows 4*(l/16)+e.
// A workgroup owns 4 G of one k (one wave each).  It walks the plane waves in
// chunks: the C1 rows and the gathered C2 rows of chunk c+1 are loaded into
// registers while chunk c runs on the matrix cores from LDS.  Workgroups of
// one k are placed on one XCD so the k+q rows they gather stay in its L2.
#include <algorithm>
#include <cstdlib>
#include <string>

#include "rpa_internal.h"

namespace rpa {

constexpr int kPfC1 = 4;   // prefetched C1 elements per thread
constexpr int kPfB = 8;    // prefetched gathered elements per thread

struct VertexArgs {
  const void *evc;  // double2 (FP64) or float2 (mixed)
  const int *mill_k;
  const int *mill_g;
  const int16_t *kbox;
  const int *gbox;
  const int *dmap;   // direct map [ng0][nlf][ngd] global G index or -1
  const int *g0idx;  // per k: index of its map combination (R1, R2, G0)
  const int *csym;   // per combination: A (9) then B (9), row-major
  const double *crkc;  // per combination: R(R1) KC, column-major
  int sym;           // some combination rotates: targets A m + B (Glf + G0)
  Box kb, gb;
  const int *ngk;
  const int *k1, *k2, *m0;
  const double *ktot;
  const int *lf_key;
  const double *glf;
  double q[3];
  double kc[9];
  double gcar;
  int nki, npwx, nband, nlf, ntot, ngd, map_mode;
  int pol1, pol2;  // 0..2 cartesian component (current vertices)
  PairTable pairs;
  int nvb;
  VBlock vb[kVertexMaxBlocks];
  int groups;      // G groups of 4 per k
  int ch;          // plane waves per LDS chunk
  int na0, nna;    // staged C1 bands [na0, na0+nna)
  int mb0, nmb;    // staged C2 bands [mb0, mb0+nmb)
  int kslots;      // ntot rounded up to a multiple of 8
  const int *perm;
  void *rho1, *rho2;
};

// Precision traits of the vertex contraction.  arow_*: which band and part
// row r of the A operand carries; out_*: for result pair p of lane group q4,
// the band and the result elements holding (Re, Im).
struct PrecF64 {
  using real = double;
  using cplx = double2;
  using acc = rpa_d4;
  static __device__ inline acc mma(real a, real b, acc c) { return mfma_f64_16x16x4(a, b, c); }
  static __device__ inline int arow_band(int r) { return r & 7; }
  static __device__ inline bool arow_im(int r) { return r >= 8; }
  // staged row of band offset n (rows of all blocks: Re then Im halves of 8)
  static __device__ inline int arow_index(int n, bool im) { return 2 * (n & ~7) + (im ? 8 : 0) + (n & 7); }
  static __device__ inline int out_band(int q4, int p) { return 4 * p + q4; }
  static __device__ inline int out_re(int p) { return p; }
  static __device__ inline int out_im(int p) { return p + 2; }
  static __device__ inline cplx make(real x, real y) { return make_double2(x, y); }
  static __device__ inline acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
  // two consecutive complex values from 16-byte aligned LDS
  static __device__ inline void load2c(const cplx *p, cplx v[2]) { v[0] = p[0], v[1] = p[1]; }
  // four consecutive reals from 16-byte aligned LDS
  static __device__ inline void load4(const real *p, real v[4]) {
    const double2 x = *(const double2 *)p, y = *(const double2 *)(p + 2);
    v[0] = x.x, v[1] = x.y, v[2] = y.x, v[3] = y.y;
  }
};

struct PrecF32 {
  using real = float;
  using cplx = float2;
  using acc = rpa_f4;
  static __device__ inline acc mma(real a, real b, acc c) { return mfma_f32_16x16x4(a, b, c); }
  static __device__ inline int arow_band(int r) { return r >> 1; }
  static __device__ inline bool arow_im(int r) { return (r & 1) != 0; }
  static __device__ inline int arow_index(int n, bool im) { return 2 * n + (im ? 1 : 0); }
  static __device__ inline int out_band(int q4, int p) { return 2 * q4 + p; }
  static __device__ inline int out_re(int p) { return 2 * p; }
  static __device__ inline int out_im(int p) { return 2 * p + 1; }
  static __device__ inline cplx make(real x, real y) { return make_float2(x, y); }
  static __device__ inline acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
  static __device__ inline void load2c(const cplx *p, cplx v[2]) {
    const float4 x = *(const float4 *)p;
    v[0] = make_float2(x.x, x.y), v[1] = make_float2(x.z, x.w);
  }
  static __device__ inline void load4(const real *p, real v[4]) {
    const float4 x = *(const float4 *)p;
    v[0] = x.x, v[1] = x.y, v[2] = x.z, v[3] = x.w;
  }
};

// Point operation of k in the igk map: plane wave s of k has the target
// Miller index A s + h in the k+q list, h = B (Glf + G0) (A = B = I without
// operations: s + Glf + G0).  The branch is uniform per k.
struct IgkOp {
  int A[9];
  int h0, h1, h2;
  bool on;
  __device__ inline void init(const VertexArgs &a, int k, int l0, int l1, int l2, int d0, int d1, int d2) {
    on = a.sym != 0;
    h0 = l0 + d0, h1 = l1 + d1, h2 = l2 + d2;
    if (on) {
      const int *cs = a.csym + 18 * a.g0idx[k];
#pragma unroll
      for (int x = 0; x < 9; ++x) A[x] = cs[x];
      const int u0 = h0, u1 = h1, u2 = h2;
      h0 = cs[9] * u0 + cs[10] * u1 + cs[11] * u2;
      h1 = cs[12] * u0 + cs[13] * u1 + cs[14] * u2;
      h2 = cs[15] * u0 + cs[16] * u1 + cs[17] * u2;
    }
  }
  __device__ inline int index(const Box &kb, int s0, int s1, int s2) const {
    if (!on) return kb.index(s0 + h0, s1 + h1, s2 + h2);
    return kb.index((A[0] * s0 + A[1] * s1 + A[2] * s2) + h0, (A[3] * s0 + A[4] * s1 + A[5] * s2) + h1,
                    (A[6] * s0 + A[7] * s1 + A[8] * s2) + h2);
  }
};

// Row pol of the matrix that takes a plane wave's Miller index to the
// G'' of the current vertex J: KC, or R(R1) KC with point operations.
__device__ inline void j_row(const VertexArgs &a, int k, int pol, double r[3]) {
  if (a.sym) {
    const double *m = a.crkc + 9 * a.g0idx[k];
    r[0] = m[pol], r[1] = m[pol + 3], r[2] = m[pol + 6];
  } else {
    r[0] = a.kc[pol], r[1] = a.kc[pol + 3], r[2] = a.kc[pol + 6];
  }
}

// Target global index of A G(i) + B (Glf(g) + G0) for every map combination
// of the q, in Miller indices (A = B = I without point operations).
__global__ void k_direct_map(const int *mill_g, const int *gbox, Box gb, const int *lf_key,
                             const int *g0keys, const int *csym, int ng0, int nlf, int ngd, int *dmap) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)ng0 * nlf * ngd;
  if (idx >= n) return;
  const int i = (int)(idx % ngd);
  const int g = (int)((idx / ngd) % nlf);
  const int z = (int)(idx / ((long long)ngd * nlf));
  int s0, s1, s2, l0, l1, l2, o0, o1, o2;
  mill_unpack(mill_g[i], s0, s1, s2);
  mill_unpack(lf_key[g], l0, l1, l2);
  mill_unpack(g0keys[z], o0, o1, o2);
  const int *A = csym + 18 * z, *B = A + 9;
  const int u0 = l0 + o0, u1 = l1 + o1, u2 = l2 + o2;
  const int t0 = (A[0] * s0 + A[1] * s1 + A[2] * s2) + (B[0] * u0 + B[1] * u1 + B[2] * u2);
  const int t1 = (A[3] * s0 + A[4] * s1 + A[5] * s2) + (B[3] * u0 + B[4] * u1 + B[5] * u2);
  const int t2 = (A[6] * s0 + A[7] * s1 + A[8] * s2) + (B[6] * u0 + B[7] * u1 + B[8] * u2);
  const int bi = gb.index(t0, t1, t2);
  dmap[idx] = bi >= 0 ? gbox[bi] : -1;
}

// KIND: 0 charge, 1 current (one polarisation), 2 current (pol yz: two)
// MAXB: band blocks compiled in; DUAL: two accumulator chains (even/odd steps)
// CH: plane waves per chunk.  Thread t stages the gathered row (w, c) =
// divmod(t / TPR, CH) for its G and columns m = t % TPR + TPR u, and the C1
// row c = t / TPA, bands n = t % TPA + TPA u; the assignment is fixed for the
// whole walk, so a chunk costs one map lookup per thread and no division.
//
// The operands are staged in LDS already in matrix-core form, transposed so
// that consecutive K steps of one lane are contiguous:
//   sA[ii][part][row][s]      A value of plane wave c = 2s + ii (the sign and
//                             Re/Im choice of the complex embedding applied)
//   sB[w][ii][part][m][s]     Re or Im of the gathered C2 (times J for
//                             current vertices)
// so a lane loads four K steps of A and of B with two 16-byte reads.  Row
// 2*nna (and column nmb) stay zero and serve the lanes outside a band block.
template <class P, int KIND, int MAXB, bool DUAL, int CH, bool PAIR>
__global__ __launch_bounds__(256) void k_vertex(VertexArgs a) {
  using real = typename P::real;
  using cplx = typename P::cplx;
  using acc_t = typename P::acc;
  // PAIR (every band block at most 8 bands m wide): a wave computes two
  // local fields, its 16 staged B columns holding the first field's bands m
  // in 0-7 and the second's in 8-15, so no matrix-core column stays empty
  constexpr int GPW = PAIR ? 2 : 1;              // local fields per wave
  constexpr int TPR = 256 / (4 * GPW * CH);      // threads per gathered row
  constexpr int TPA = 256 / CH;            // threads per C1 row
  constexpr int S = CH / 2;                // K steps per chunk
  constexpr int SP = S + 16 / sizeof(real);  // padded step stride (16-byte multiple)
  constexpr int NBS = (KIND == 2) ? 2 : 1;   // B sets (J1, J2)
  extern __shared__ __align__(16) unsigned char vsm_raw[];
  const int nna = a.nna, nmb = a.nmb;
  const int zcol = PAIR ? 16 : nmb;  // zero column
  const int arows = 2 * ((nna + 7) / 8) * 8 + 1, bcols = zcol + 1;  // + zero row / column
  const int zrow = arows - 1;
  real *sA = (real *)vsm_raw;                      // [2][2][arows][SP]
  real *sB = sA + 4 * arows * SP;                  // [NBS][4][2][2][bcols][SP]
  auto A_at = [&](int ii_, int part_, int row) { return sA + ((ii_ * 2 + part_) * arows + row) * SP; };
  auto B_at = [&](int set, int w, int ii_, int part_, int m) {
    return sB + ((((set * 4 + w) * 2 + ii_) * 2 + part_) * bcols + m) * SP;
  };

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // XCD-aware placement: linear id L runs on XCD L % 8; all G groups of one
  // k get ids with the same residue.
  const int L = blockIdx.x;
  const int xcd = L & 7;
  const int slot = L >> 3;
  const int k = xcd + 8 * (slot / a.groups);
  const int grp = slot % a.groups;
  if (k >= a.ntot) return;  // whole workgroup leaves together
  const int iG0 = (grp * 4 + wave) * GPW;  // first local field of the wave
  const bool active = iG0 < a.nlf;

  const int K1 = a.k1[k] - 1, K2 = a.k2[k] - 1;
  int d0, d1, d2;
  mill_unpack(a.m0[k], d0, d1, d2);

  const bool direct = (a.map_mode == RPA_MAP_DIRECT);
  const int n1 = direct ? a.ngd : a.ngk[K1];
  const int lim2 = a.ngk[K2];
  const int *src = direct ? a.mill_g : a.mill_k + (size_t)K1 * a.npwx;
  const int16_t *box2 = a.kbox + (size_t)K2 * a.kb.size();
  const cplx *evc = (const cplx *)a.evc;
  const cplx *c1 = evc + (size_t)K1 * a.npwx * a.nband + a.na0;
  const cplx *c2 = evc + (size_t)K2 * a.npwx * a.nband + a.mb0;
  const double *kv = a.ktot + 3 * k;

  // zero rows / columns for masked lanes (never overwritten by staging)
  for (int x = tid; x < 4 * SP; x += 256) A_at(x / SP / 2, (x / SP) & 1, zrow)[x % SP] = 0;
  for (int x = tid; x < NBS * 16 * SP; x += 256) {
    const int q = x / SP;  // (set, w, ii, part)
    B_at(q >> 4, (q >> 2) & 3, (q >> 1) & 1, q & 1, zcol)[x % SP] = 0;
  }

  // this thread's staging assignment
  const int brow = tid / TPR, bsub = tid % TPR;  // gathered row (field, c)
  const int bf = brow / CH, bc = brow % CH;      // field of the workgroup, plane wave
  const int bw = bf / GPW, bh = bf % GPW;        // its wave and column half
  const int bg = grp * 4 * GPW + bf;
  const bool bg_ok = bg < a.nlf;
  const int *dmap = direct ? a.dmap + ((size_t)a.g0idx[k] * a.nlf + (bg_ok ? bg : 0)) * a.ngd : nullptr;
  int lf0 = 0, lf1 = 0, lf2 = 0;
  if (!direct && bg_ok) mill_unpack(a.lf_key[bg], lf0, lf1, lf2);
  IgkOp iop;
  iop.init(a, k, lf0, lf1, lf2, d0, d1, d2);
  double kr1[3] = {0.0, 0.0, 0.0}, kr2[3] = {0.0, 0.0, 0.0};
  if (KIND >= 1) j_row(a, k, a.pol1, kr1);
  if (KIND == 2) j_row(a, k, a.pol2, kr2);
  const int arow = tid / TPA, asub = tid % TPA;  // C1 row c, first band

  const int r = lane & 15, q4 = lane >> 4;
  const int ii = q4 >> 1, part = q4 & 1;
  constexpr bool MK = MAXB >= 4;

  cplx pc1[kPfC1], pb[kPfB];
  unsigned pm_a = 0, pm_b = 0;  // presence of each prefetched value
  real pj1 = 1, pj2 = 1;
  // J without the plane-wave term: Gcar-free (q + 2k + Glf(G))_pol of this thread's G
  double jb1 = 0.0, jb2 = 0.0;
  if (KIND >= 1 && bg_ok) {
    jb1 = (a.q[a.pol1] + 2.0 * kv[a.pol1]) + a.glf[3 * bg + a.pol1];
    if (KIND == 2) jb2 = (a.q[a.pol2] + 2.0 * kv[a.pol2]) + a.glf[3 * bg + a.pol2];
  }

  // Loads of one chunk into registers: the C1 rows, the gathered k+q rows
  // (row j of this thread's G, -1 = absent) and, for current vertices, J
  // from the Miller key mkey of the plane wave.
  // MK (many band blocks): every load is issued from a clamped in-bounds
  // address and its presence kept in a mask, applied when the values are
  // staged -- a conditional load joined with a zero fill makes the compiler
  // wait for each load before the next one issues.  With few blocks the
  // products per chunk are few and the masked form's extra loads cost more
  // (measured: MAXB 8 7% faster masked, MAXB 1 14% slower), so those use
  // conditional loads.
  auto prefetch = [&](int i0, int j, int mkey) {
    {
      const int i = i0 + arow;
      const bool iok = i < n1;
      unsigned m = 0;
      if constexpr (MK) {
        const cplx *row = c1 + (size_t)(iok ? i : 0) * a.nband;
#pragma unroll
        for (int u = 0; u < kPfC1; ++u) {
          const int n = asub + TPA * u;
          const bool ok = iok && n < nna;
          pc1[u] = row[ok ? n : 0];
          m |= ok ? (1u << u) : 0u;
        }
      } else {
        const cplx *row = c1 + (size_t)i * a.nband;
#pragma unroll
        for (int u = 0; u < kPfC1; ++u) {
          const int n = asub + TPA * u;
          pc1[u] = (iok && n < nna) ? row[n] : P::make(0, 0);
        }
        m = ~0u;
      }
      pm_a = m;
    }
    unsigned m = 0;
    if constexpr (MK) {
      const cplx *row = c2 + (size_t)(j >= 0 ? j : 0) * a.nband;
#pragma unroll
      for (int u = 0; u < kPfB; ++u) {
        const int mb = bsub + TPR * u;
        const bool ok = j >= 0 && mb < nmb;
        pb[u] = row[ok ? mb : 0];
        m |= ok ? (1u << u) : 0u;
      }
    } else {
      const cplx *row = c2 + (size_t)(j >= 0 ? j : 0) * a.nband;
#pragma unroll
      for (int u = 0; u < kPfB; ++u) {
        const int mb = bsub + TPR * u;
        pb[u] = (j >= 0 && mb < nmb) ? row[mb] : P::make(0, 0);
      }
      m = ~0u;
    }
    pm_b = m;
    if (KIND >= 1) {
      pj1 = pj2 = 0;
      if (bg_ok && i0 + bc < n1) {
        int s0, s1, s2;
        mill_unpack(mkey, s0, s1, s2);
        // J = Gcar (q + 2k + Glf(G) + 2 G(i))_pol with G(i) = KC mill (left to
        // right; R(R1) G(i) with point operations)
        const double g1 = ((0.0 + kr1[0] * (double)s0) + kr1[1] * (double)s1) + kr1[2] * (double)s2;
        pj1 = (real)(0.5 * ((jb1 + 2.0 * g1) * a.gcar));
        if (KIND == 2) {
          const double g2 = ((0.0 + kr2[0] * (double)s0) + kr2[1] * (double)s1) + kr2[2] * (double)s2;
          pj2 = (real)(0.5 * ((jb2 + 2.0 * g2) * a.gcar));
        }
      }
    }
  };

  // Gather-index pipeline: the row gathered for chunk c+1 is known before
  // its data loads issue, so they never wait on an index load.
  //   direct: raw map value of chunk c+1 in jn (loaded while chunk c-1 ran)
  //   igk:    Miller key of chunk c+2 in kn, its box lookup -> jn one chunk
  //           later (two dependent loads, each hidden behind a chunk)
  // Index loads read clamped addresses; whether plane wave i exists is known
  // from i alone and applied when the value is used.
  auto key_ok = [&](int i) -> bool { return bg_ok && i < n1; };
  auto load_key = [&](int i) -> int { return direct ? dmap[min(i, n1 - 1)] : src[min(i, n1 - 1)]; };
  // box lookup of key (valid when ok): the raw box entry, found = present
  auto lookup_igk = [&](int key, bool ok, bool &found) -> int {
    int s0, s1, s2;
    mill_unpack(key, s0, s1, s2);
    const int bi = iop.index(a.kb, s0, s1, s2);
    found = ok && bi >= 0;
    return (int)box2[found ? bi : 0];
  };
  auto row_of = [&](int jraw) -> int { return (jraw >= 0 && jraw < lim2) ? jraw : -1; };
  auto load_mkey = [&](int i) -> int { return KIND >= 1 ? src[min(i, n1 - 1)] : 0; };

  // registers -> LDS in matrix-core form
  auto stage = [&]() {
    {
      const int ai = arow & 1, as = arow >> 1;
#pragma unroll
      for (int u = 0; u < kPfC1; ++u) {
        const int n = asub + TPA * u;
        if (n < nna) {
          const bool ok = (pm_a >> u) & 1u;
          const cplx z = P::make(ok ? pc1[u].x : (real)0, ok ? pc1[u].y : (real)0);
          const int rre = P::arow_index(n, false), rim = P::arow_index(n, true);
          // Re rows: (C1r, C1i); Im rows: (-C1i, C1r) for part (0, 1)
          A_at(ai, 0, rre)[as] = z.x;
          A_at(ai, 1, rre)[as] = z.y;
          A_at(ai, 0, rim)[as] = -z.y;
          A_at(ai, 1, rim)[as] = z.x;
        }
      }
    }
    const int bi_ = bc & 1, bs = bc >> 1;
#pragma unroll
    for (int u = 0; u < kPfB; ++u) {
      const int m = bsub + TPR * u;
      if (m < nmb) {
        const bool ok = (pm_b >> u) & 1u;
        const cplx z = P::make(ok ? pb[u].x : (real)0, ok ? pb[u].y : (real)0);
        const int mcol = 8 * bh + m;
        B_at(0, bw, bi_, 0, mcol)[bs] = z.x * pj1;
        B_at(0, bw, bi_, 1, mcol)[bs] = z.y * pj1;
        if (KIND == 2) {
          B_at(NBS - 1, bw, bi_, 0, mcol)[bs] = z.x * pj2;
          B_at(NBS - 1, bw, bi_, 1, mcol)[bs] = z.y * pj2;
        }
      }
    }
  };

  acc_t acc1[MAXB], acc2[MAXB], bcc1[MAXB], bcc2[MAXB];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    acc1[b] = acc2[b] = bcc1[b] = bcc2[b] = P::zero();
  }
  // per band block: this lane's A row and B column (zero row/column if masked)
  const real *arowp[MAXB];
  const real *bcolp[MAXB], *bcolp2[MAXB];
  {
    const int nr = P::arow_band(r);
    const bool im_row = P::arow_im(r);
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const VBlock vb = a.vb[b < a.nvb ? b : 0];
      const int arix = (nr < vb.nrows) ? P::arow_index(vb.n0 - a.na0 + nr, im_row) : zrow;
      const int hc = PAIR ? (r >> 3) : 0, mc = PAIR ? (r & 7) : r;
      const int bcix = (mc < vb.ncols) ? 8 * hc + vb.m0 - a.mb0 + mc : zcol;
      arowp[b] = A_at(ii, part, arix);
      bcolp[b] = B_at(0, wave, ii, part, bcix);
      bcolp2[b] = B_at(NBS - 1, wave, ii, part, bcix);
    }
  }

  // jn: raw index of chunk c+1 (valid when jn_ok); kn: igk key of chunk c+2
  // (valid when kn_ok).  MK: only loads write jn and kn, so no zero fill has
  // to wait for a load in flight; else absent keys are kNone.
  constexpr int kNone = -2147483647 - 1;
  auto load_key0 = [&](int i) -> int { return key_ok(i) ? load_key(i) : kNone; };
  auto lookup0 = [&](int key) -> int {
    bool f;
    const int v = lookup_igk(key, key != kNone, f);
    return f ? v : -1;
  };
  int jn, kn = 0, mkn;
  bool jn_ok = true, kn_ok = false;
  if constexpr (MK) {
    const bool ok0 = key_ok(bc);
    const int key0 = load_key(bc);
    int j0;
    if (direct) {
      j0 = ok0 ? key0 : -1;
    } else {
      bool f;
      const int v = lookup_igk(key0, ok0, f);
      j0 = f ? v : -1;
    }
    prefetch(0, row_of(j0), load_mkey(bc));
    const bool ok1 = key_ok(CH + bc);
    const int key1 = load_key(CH + bc);
    if (direct) {
      jn = key1;
      jn_ok = ok1;
    } else {
      jn = lookup_igk(key1, ok1, jn_ok);
      kn_ok = key_ok(2 * CH + bc);
      kn = load_key(2 * CH + bc);
    }
    mkn = load_mkey(CH + bc);
  } else {
    const int key0 = load_key0(bc);
    const int j0 = row_of(direct ? key0 : lookup0(key0));
    prefetch(0, j0, key_ok(bc) ? load_mkey(bc) : 0);
    const int key1 = load_key0(CH + bc);
    jn = direct ? key1 : lookup0(key1);
    mkn = key_ok(CH + bc) ? load_mkey(CH + bc) : 0;
    if (!direct) kn = load_key0(2 * CH + bc);
  }
  for (int i0 = 0; i0 < n1; i0 += CH) {
    __syncthreads();  // the previous chunk's LDS reads are done
    stage();
    __syncthreads();
    if constexpr (MK) {
      prefetch(i0 + CH, jn_ok ? row_of(jn) : -1, mkn);  // past the end: clamped and masked
      const int i2 = i0 + 2 * CH + bc;
      if (direct) {
        jn_ok = key_ok(i2);
        jn = load_key(i2);
      } else {
        jn = lookup_igk(kn, kn_ok, jn_ok);
        kn_ok = key_ok(i2 + CH);
        kn = load_key(i2 + CH);
      }
      mkn = load_mkey(i2);
    } else {
      if (i0 + CH < n1) prefetch(i0 + CH, row_of(jn), mkn);
      const int i2 = i0 + 2 * CH + bc;
      if (direct) {
        jn = load_key0(i2);
      } else {
        jn = lookup0(kn);
        kn = load_key0(i2 + CH);
      }
      mkn = key_ok(i2) ? load_mkey(i2) : 0;
    }
    if (active) {
      // four K steps per pass: two 16-byte reads per operand
#pragma unroll
      for (int s4 = 0; s4 < S; s4 += 4) {
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
          if (b < a.nvb) {
            real av[4], bv[4], bw2[4];
            P::load4(arowp[b] + s4, av);
            P::load4(bcolp[b] + s4, bv);
            if (KIND == 2) P::load4(bcolp2[b] + s4, bw2);
#pragma unroll
            for (int x = 0; x < 4; ++x) {
              acc_t *x1 = (DUAL && (x & 1)) ? bcc1 : acc1;
              acc_t *x2 = (DUAL && (x & 1)) ? bcc2 : acc2;
              x1[b] = P::mma(av[x], bv[x], x1[b]);
              if (KIND == 2) x2[b] = P::mma(av[x], bw2[x], x2[b]);
            }
          }
        }
      }
    }
  }
  if (!active) return;
  if (DUAL) {
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      acc1[b] += bcc1[b];
      acc2[b] += bcc2[b];
    }
  }

  // the two (Re, Im) result pairs of this lane: band out_band(q4, p), column r
  // (PAIR: field iG0 + r / 8, band m r % 8)
  const int hc = PAIR ? (r >> 3) : 0, mc = PAIR ? (r & 7) : r;
  const int iG = iG0 + hc;
  if (iG >= a.nlf) return;
  cplx *rho1 = (cplx *)a.rho1, *rho2 = (cplx *)a.rho2;
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    if (b < a.nvb) {
      const VBlock vb = a.vb[b];
      const PairBlock &pbk = a.pairs.b[vb.pb];
      const int nm = pbk.m_hi - pbk.m_lo;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int onr = P::out_band(q4, p);
        if (onr < vb.nrows && mc < vb.ncols) {
          const int n = vb.n0 + onr, m = vb.m0 + mc;
          const long long t = (long long)k * a.pairs.tpk + a.pairs.off[vb.pb] +
                              (long long)(n - pbk.n_lo) * nm + (m - pbk.m_lo);
          const long long pos = a.perm ? (long long)a.perm[t] : t;
          if (pos >= 0) {
            const cplx v1 = P::make(acc1[b][P::out_re(p)], acc1[b][P::out_im(p)]);
            rho1[pos * a.nlf + iG] = v1;
            if (KIND == 2) rho2[pos * a.nlf + iG] = P::make(acc2[b][P::out_re(p)], acc2[b][P::out_im(p)]);
            else if (KIND == 1 && rho2) rho2[pos * a.nlf + iG] = v1;
          }
        }
      }
    }
  }
}

// Gather-direct form for up to two band blocks.  The K index of one MFMA is
// four plane waves c = 4t + kg (kg = lane / 16) of one part: Re(conj(C1) C2)
// = C1r C2r + C1i C2i and Im = C1r C2i - C1i C2r are two MFMAs sharing the
// loaded values.  Each lane loads its B value (the C2 row of its plane wave,
// column m = lane % 16) straight from L2 -- 32 lanes read one 128-byte row
// segment -- one chunk ahead in registers, so B never passes through LDS.
// The C1 chunk, shared by the four waves, is staged in LDS ([kg][n][t],
// double-buffered, one barrier per chunk).  Row indices of the gathered
// rows are resolved by lanes 0..31 (plane wave i0 + lane) and broadcast
// with a lane shuffle.
constexpr int kGdCH = 32;           // plane waves per chunk
constexpr int kGdNT = kGdCH / 4;    // K groups of four per chunk

template <class P, int KIND, int MAXB, bool DUAL, bool PAIR>
__global__ __launch_bounds__(256) void k_vertex_gd(VertexArgs a) {
  using real = typename P::real;
  using cplx = typename P::cplx;
  using acc_t = typename P::acc;
  constexpr int CH = kGdCH, NT = kGdNT;
  constexpr int UA = 2 * MAXB;  // C1 elements per staging thread: nna <= 8 UA
  extern __shared__ __align__(16) unsigned char vsm_raw[];
  const int nna = a.nna;
  const int nap = nna + 1;  // + zero row
  cplx *sA = (cplx *)vsm_raw;  // [2][4][nap][NT]
  auto A_at = [&](int buf, int q, int n) { return sA + ((buf * 4 + q) * nap + n) * NT; };

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int L = blockIdx.x;
  const int xcd = L & 7;
  const int slot = L >> 3;
  const int k = xcd + 8 * (slot / a.groups);
  const int grp = slot % a.groups;
  if (k >= a.ntot) return;  // whole workgroup leaves together
  // PAIR (every band block at most 8 bands m wide): a wave computes two
  // local fields, matrix-core columns 0-7 the first and 8-15 the second, so
  // the columns an 8-wide block leaves empty carry the second field; its
  // rows are resolved by lanes 32-63
  constexpr int GPW = PAIR ? 2 : 1;
  const int iG0 = (grp * 4 + wave) * GPW;  // first local field of the wave
  const bool active = iG0 < a.nlf;
  const int hr = PAIR ? (lane >> 5) : 0;   // field of this resolution lane
  const int lr = PAIR ? (lane & 31) : lane;  // its plane wave in the chunk
  const int iGr = iG0 + hr;
  const bool res_ok = active && iGr < a.nlf && (PAIR || lane < CH);

  const int K1 = a.k1[k] - 1, K2 = a.k2[k] - 1;
  int d0, d1, d2;
  mill_unpack(a.m0[k], d0, d1, d2);
  const bool direct = (a.map_mode == RPA_MAP_DIRECT);
  const int n1 = direct ? a.ngd : a.ngk[K1];
  const int lim2 = a.ngk[K2];
  const int *src = direct ? a.mill_g : a.mill_k + (size_t)K1 * a.npwx;
  const int16_t *box2 = a.kbox + (size_t)K2 * a.kb.size();
  const cplx *evc = (const cplx *)a.evc;
  const cplx *c1 = evc + (size_t)K1 * a.npwx * a.nband + a.na0;
  const cplx *c2 = evc + (size_t)K2 * a.npwx * a.nband + a.mb0;
  const double *kv = a.ktot + 3 * k;
  const int *dmap = direct ? a.dmap + ((size_t)a.g0idx[k] * a.nlf + min(iGr, a.nlf - 1)) * a.ngd : nullptr;
  int lf0 = 0, lf1 = 0, lf2 = 0;
  if (!direct && iGr < a.nlf) mill_unpack(a.lf_key[iGr], lf0, lf1, lf2);
  IgkOp iop;
  iop.init(a, k, lf0, lf1, lf2, d0, d1, d2);
  double kr1[3] = {0.0, 0.0, 0.0}, kr2[3] = {0.0, 0.0, 0.0};
  if (KIND >= 1) j_row(a, k, a.pol1, kr1);
  if (KIND == 2) j_row(a, k, a.pol2, kr2);

  const int r = lane & 15, kg = lane >> 4;
  const int hc = PAIR ? (r >> 3) : 0;  // field of this matrix-core column
  const int mc = PAIR ? (r & 7) : r;   // its band m in the block
  // zero rows of both buffers
  for (int x = tid; x < 2 * 4 * NT; x += 256) A_at(x / (4 * NT), (x / NT) & 3, nna)[x % NT] = P::make(0, 0);

  // ---- per-lane operand positions of each band block
  int arow[MAXB], bcol[MAXB];
  {
    const int nr = P::arow_band(r);
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const VBlock vb = a.vb[b < a.nvb ? b : 0];
      arow[b] = (nr < vb.nrows) ? vb.n0 - a.na0 + nr : nna;
      bcol[b] = (mc < vb.ncols) ? vb.m0 - a.mb0 + mc : -1;
    }
  }
  const bool im_row = P::arow_im(r);

  // ---- C1 staging: thread -> plane wave ca = tid / 8, bands (tid % 8) + 8u
  const int ca = tid >> 3, na = tid & 7;
  cplx pa[UA];
  auto load_a = [&](int i0) {
    const int i = i0 + ca;
    const cplx *row = c1 + (size_t)i * a.nband;
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      const int n = na + 8 * u;
      if (8 * u < nna) pa[u] = (i < n1 && n < nna) ? row[n] : P::make(0, 0);
    }
  };
  auto store_a = [&](int buf) {
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      const int n = na + 8 * u;
      if (8 * u < nna && n < nna) A_at(buf, ca & 3, n)[ca >> 2] = pa[u];
    }
  };

  // ---- gathered rows: lane l < 32 resolves plane wave i0 + l of its G
  //      (PAIR: lanes 32-63 those of the second field)
  constexpr int kNone = -2147483647 - 1;
  auto load_key = [&](int i) -> int {
    return (res_ok && i < n1) ? (direct ? dmap[i] : src[i]) : kNone;
  };
  auto lookup_igk = [&](int key) -> int {
    if (key == kNone) return -1;
    int s0, s1, s2;
    mill_unpack(key, s0, s1, s2);
    const int bi = iop.index(a.kb, s0, s1, s2);
    return bi >= 0 ? (int)box2[bi] : -1;
  };
  auto row_of = [&](int jraw) -> int { return (jraw >= 0 && jraw < lim2) ? jraw : -1; };
  auto load_mkey = [&](int i) -> int { return (KIND >= 1 && res_ok && i < n1) ? src[i] : 0; };
  // J of plane wave i for current vertices (pol p), lanes < 32
  auto jweight = [&](int mkey, int i, int pol, const double kr[3]) -> real {
    if (!(res_ok && i < n1)) return 0;
    int s0, s1, s2;
    mill_unpack(mkey, s0, s1, s2);
    const double jb = (a.q[pol] + 2.0 * kv[pol]) + a.glf[3 * iGr + pol];
    const double g = ((0.0 + kr[0] * (double)s0) + kr[1] * (double)s1) + kr[2] * (double)s2;
    return (real)(0.5 * ((jb + 2.0 * g) * a.gcar));
  };

  // B registers of one chunk: value of plane wave 4t + kg, column bcol[b]
  cplx bz[NT][MAXB];
  real jw1 = 1, jw2 = 1;  // J of plane wave i0 + lane (lanes < 32) for the chunk in bz
  auto load_b = [&](int jrow) {  // jrow: resolved row of plane wave i0 + lane
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int j = __shfl(jrow, 32 * hc + 4 * t + kg);
#pragma unroll
      for (int b = 0; b < MAXB; ++b)
        bz[t][b] = (b < a.nvb && j >= 0 && bcol[b] >= 0) ? c2[(size_t)j * a.nband + bcol[b]] : P::make(0, 0);
    }
  };

  acc_t acc1[MAXB], acc2[MAXB], bcc1[MAXB], bcc2[MAXB];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) acc1[b] = acc2[b] = bcc1[b] = bcc2[b] = P::zero();

  // ---- prologue: A(0) in LDS, A(1) in registers, B(0) in registers,
  //      rows of chunk 1 resolved (jn), keys of chunk 2 (kn, igk)
  int jn, kn = kNone, mkn = 0;
  real jwn1 = 1, jwn2 = 1;
  load_a(0);
  store_a(0);
  {
    const int key0 = load_key(lr);
    const int j0 = row_of(direct ? key0 : lookup_igk(key0));
    if (KIND >= 1) {
      const int mk0 = load_mkey(lr);
      jw1 = jweight(mk0, lr, a.pol1, kr1);
      if (KIND == 2) jw2 = jweight(mk0, lr, a.pol2, kr2);
    }
    if (active) load_b(j0);
    const int key1 = load_key(CH + lr);
    jn = direct ? key1 : lookup_igk(key1);
    if (KIND >= 1) mkn = load_mkey(CH + lr);
    if (!direct) kn = load_key(2 * CH + lr);
  }
  if (CH < n1) load_a(CH);

  int buf = 0;
  for (int i0 = 0; i0 < n1; i0 += CH, buf ^= 1) {
    __syncthreads();  // A(i0) visible; reads of the other buffer finished
    if (i0 + CH < n1) {
      store_a(buf ^ 1);
      if (i0 + 2 * CH < n1) load_a(i0 + 2 * CH);
    }
    // B values and J of this chunk; then start the next chunk's loads
    cplx bc[NT][MAXB];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int b = 0; b < MAXB; ++b) bc[t][b] = bz[t][b];
    const real w1 = jw1, w2 = jw2;
    if (active && i0 + CH < n1) {
      load_b(row_of(jn));
      if (KIND >= 1) {
        jw1 = jweight(mkn, i0 + CH + lr, a.pol1, kr1);
        if (KIND == 2) jw2 = jweight(mkn, i0 + CH + lr, a.pol2, kr2);
      }
      const int i2 = i0 + 2 * CH + lr;
      if (direct) {
        jn = load_key(i2);
      } else {
        jn = lookup_igk(kn);
        kn = load_key(i2 + CH);
      }
      if (KIND >= 1) mkn = load_mkey(i2);
    }
    if (active) {
#pragma unroll
      for (int t = 0; t < NT; t += 2) {
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
          if (b < a.nvb) {
            // C1 of plane waves 4t + kg and 4(t+1) + kg, band arow[b]
            cplx za[2];
            P::load2c(A_at(buf, kg, arow[b]) + t, za);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const cplx z1 = za[h];
              const real are = im_row ? -z1.y : z1.x;
              const real aim = im_row ? z1.x : z1.y;
              const cplx z2 = bc[t + h][b];
              acc_t *x1 = (DUAL && h) ? bcc1 : acc1;
              acc_t *x2 = (DUAL && h) ? bcc2 : acc2;
              if (KIND >= 1) {
                const real wa = __shfl(w1, 32 * hc + 4 * (t + h) + kg);
                x1[b] = P::mma(are, z2.x * wa, x1[b]);
                x1[b] = P::mma(aim, z2.y * wa, x1[b]);
                if (KIND == 2) {
                  const real wb = __shfl(w2, 32 * hc + 4 * (t + h) + kg);
                  x2[b] = P::mma(are, z2.x * wb, x2[b]);
                  x2[b] = P::mma(aim, z2.y * wb, x2[b]);
                }
              } else {
                x1[b] = P::mma(are, z2.x, x1[b]);
                x1[b] = P::mma(aim, z2.y, x1[b]);
              }
            }
          }
        }
      }
    }
  }
  if (!active) return;
  if (DUAL) {
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      acc1[b] += bcc1[b];
      acc2[b] += bcc2[b];
    }
  }
  cplx *rho1 = (cplx *)a.rho1, *rho2 = (cplx *)a.rho2;
  const int iG = iG0 + hc;
  if (iG >= a.nlf) return;
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    if (b < a.nvb) {
      const VBlock vb = a.vb[b];
      const PairBlock &pbk = a.pairs.b[vb.pb];
      const int nm = pbk.m_hi - pbk.m_lo;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int onr = P::out_band(kg, p);
        if (onr < vb.nrows && mc < vb.ncols) {
          const int n = vb.n0 + onr, m = vb.m0 + mc;
          const long long tt = (long long)k * a.pairs.tpk + a.pairs.off[vb.pb] +
                               (long long)(n - pbk.n_lo) * nm + (m - pbk.m_lo);
          const long long pos = a.perm ? (long long)a.perm[tt] : tt;
          if (pos >= 0) {
            const cplx v1 = P::make(acc1[b][P::out_re(p)], acc1[b][P::out_im(p)]);
            rho1[pos * a.nlf + iG] = v1;
            if (KIND == 2) rho2[pos * a.nlf + iG] = P::make(acc2[b][P::out_re(p)], acc2[b][P::out_im(p)]);
            else if (KIND == 1 && rho2) rho2[pos * a.nlf + iG] = v1;
          }
        }
      }
    }
  }
}

// LDS-resident form for charge vertices of one band-pair block with at most
// 8 bands n (every Sloss case).  A workgroup owns one k and a group of four bands m: the k+q
// coefficients of those bands, all ngk(K2) rows, live in LDS for the whole
// walk (plus a zero row for absent targets), so every local field gathers
// from LDS and L2 supplies each k+q row once per band group.  MFMA columns
// are (4 bands m) x (4 local fields G); rows are (8 bands n, Re/Im); the K
// index is four plane waves of one part (Re and Im of the product are two
// MFMAs).  The gathered row of every (G, plane wave) of a chunk is resolved
// one chunk ahead and staged as int16 in LDS, as is the C1 chunk
// (double-buffered, one barrier per chunk).  The four waves split K: wave w
// takes the K groups t = w and w + 4 of each chunk for all local fields, and
// the partial sums are added through LDS at the end.
constexpr int kLdsCH = 32;             // plane waves per chunk
constexpr int kLdsNT = kLdsCH / 4;     // K groups per chunk
constexpr int kLdsMaxGQ = 24;          // local-field quads (Nlf <= 96)

constexpr int kLdsRowStride = 5;       // complex values per staged k+q row (4 + pad)

struct LdsLayout {
  int rows2;     // staged k+q rows + 1 (zero row)
  int ngp;       // local fields padded to 4
  size_t bytes;  // dynamic LDS
};

// local-field quads compiled for a given Nlf (extra fields gather the zero row)
__host__ __device__ inline int lds_ngq(int nlf) {
  const int q = (nlf + 3) / 4;
  return q <= 2 ? 2 : q <= 4 ? 4 : q <= 8 ? 8 : q <= 13 ? 13 : q <= 16 ? 16 : q <= 20 ? 20 : 24;
}

template <class P>
__host__ __device__ inline LdsLayout lds_layout(int npwx, int nlf) {
  LdsLayout L;
  L.rows2 = npwx + 1;
  L.ngp = 4 * lds_ngq(nlf);
  const size_t c2 = sizeof(typename P::cplx) * kLdsRowStride * (size_t)L.rows2;
  const size_t aa = sizeof(typename P::cplx) * 2 * 4 * 9 * kLdsNT;  // [2][4][8+1][NT]
  const size_t jj = sizeof(int16_t) * 2 * (size_t)L.ngp * kLdsCH;
  const size_t red = sizeof(typename P::real) * (size_t)L.ngp * 256;  // final wave sum
  const size_t st = (c2 + aa + jj + 15) / 16 * 16;
  L.bytes = st > red ? st : red;
  return L;
}

template <class P, int NGQ>
__global__ __launch_bounds__(256) void k_vertex_lds(VertexArgs a) {
  using real = typename P::real;
  using cplx = typename P::cplx;
  using acc_t = typename P::acc;
  constexpr int CH = kLdsCH, NT = kLdsNT;
  extern __shared__ __align__(16) unsigned char vsm_raw[];
  const LdsLayout lay = lds_layout<P>(a.npwx, a.nlf);
  constexpr int RS = kLdsRowStride;
  cplx *sC2 = (cplx *)vsm_raw;                         // [rows2][RS], 4 used
  cplx *sA = sC2 + RS * (size_t)lay.rows2;             // [2][4][9][NT]
  int16_t *sJ = (int16_t *)(sA + 2 * 4 * 9 * NT);      // [2][ngp][CH]
  auto A_at = [&](int buf, int q, int n) { return sA + ((buf * 4 + q) * 9 + n) * NT; };
  auto J_at = [&](int buf, int g) { return sJ + ((size_t)buf * lay.ngp + g) * CH; };

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nmg = (a.pairs.b[0].m_hi - a.pairs.b[0].m_lo + 3) / 4;  // band groups of m
  // XCD-aware placement: linear id L runs on XCD L % 8; the band groups of
  // one k share the XCD (and its copy of the k / k+q rows in L2)
  const int Lid = blockIdx.x;
  const int xcd = Lid & 7;
  const int slot = Lid >> 3;
  const int k = xcd + 8 * (slot / nmg);
  const int mg = slot % nmg;
  if (k >= a.ntot) return;  // whole workgroup leaves together

  const PairBlock &pbk = a.pairs.b[0];  // the single band-pair block
  const int K1 = a.k1[k] - 1, K2 = a.k2[k] - 1;
  int d0, d1, d2;
  mill_unpack(a.m0[k], d0, d1, d2);
  const bool direct = (a.map_mode == RPA_MAP_DIRECT);
  const int n1 = direct ? a.ngd : a.ngk[K1];
  const int lim2 = a.ngk[K2];
  const int zrow = lim2;  // zero row of the staged k+q block
  const int *src = direct ? a.mill_g : a.mill_k + (size_t)K1 * a.npwx;
  const int16_t *box2 = a.kbox + (size_t)K2 * a.kb.size();
  const cplx *evc = (const cplx *)a.evc;
  const cplx *c1 = evc + (size_t)K1 * a.npwx * a.nband + pbk.n_lo;
  const int mlo = pbk.m_lo + 4 * mg;              // first band m of the group
  const int nmv = min(4, pbk.m_hi - mlo);         // valid bands in the group
  const cplx *c2 = evc + (size_t)K2 * a.npwx * a.nband + mlo;
  const int nlf = a.nlf, nrows = pbk.n_hi - pbk.n_lo;

  // ---- the k+q block of the band group, rows 0..lim2-1, and the zero row
  for (int x = tid; x < 4 * (lim2 + 1); x += 256) {
    const int j = x >> 2, m = x & 3;
    sC2[RS * j + m] = (j < lim2 && m < nmv) ? c2[(size_t)j * a.nband + m] : P::make(0, 0);
  }
  for (int x = tid; x < 2 * 4 * NT; x += 256) A_at(x / (4 * NT), (x / NT) & 3, 8)[x % NT] = P::make(0, 0);

  // ---- staging assignments: C1 (plane wave ca, band na) and gather rows
  const int ca = tid >> 3, na = tid & 7;
  constexpr int NJ = (kLdsMaxGQ * 4 * CH + 255) / 256;  // rows resolved per thread (max)
  const int njt = (lay.ngp * CH + 255) / 256;           // used
  cplx pa;
  int pj[NJ];
  auto load_a = [&](int i0) {
    const int i = i0 + ca;
    pa = (i < n1 && na < nrows) ? c1[(size_t)i * a.nband + na] : P::make(0, 0);
  };
  // gather row of entry x = g * CH + c of the chunk at i0: (G g, plane wave i0 + c)
  auto raw_key = [&](int x, int i0) -> int {
    const int g = x / CH, i = i0 + (x % CH);
    if (g >= nlf || i >= n1) return -1;
    return direct ? a.dmap[((size_t)a.g0idx[k] * nlf + g) * a.ngd + i] : src[i];
  };
  auto resolve = [&](int x, int v) -> int {
    if (v < 0 && direct) return zrow;
    const int g = x / CH, i = x % CH;
    (void)i;
    if (g >= nlf) return zrow;
    if (direct) return (v >= 0 && v < lim2) ? v : zrow;
    if (v == -1) return zrow;  // past the end of the list
    int s0, s1, s2, l0, l1, l2;
    mill_unpack(v, s0, s1, s2);
    mill_unpack(a.lf_key[g], l0, l1, l2);
    IgkOp op;
    op.init(a, k, l0, l1, l2, d0, d1, d2);
    const int bi = op.index(a.kb, s0, s1, s2);
    const int j = bi >= 0 ? (int)box2[bi] : -1;
    return j >= 0 ? j : zrow;
  };
  auto load_j = [&](int i0) {
#pragma unroll
    for (int u = 0; u < NJ; ++u) {
      const int x = tid + 256 * u;
      if (u < njt) pj[u] = (x < lay.ngp * CH) ? raw_key(x, i0) : -1;
    }
  };
  auto store = [&](int buf) {
    if (na < 8) A_at(buf, ca & 3, na)[ca >> 2] = pa;
#pragma unroll
    for (int u = 0; u < NJ; ++u) {
      const int x = tid + 256 * u;
      if (u < njt && x < lay.ngp * CH) J_at(buf, x / CH)[x % CH] = (int16_t)resolve(x, pj[u]);
    }
  };

  const int r = lane & 15, kg = lane >> 4;
  const int nr = P::arow_band(r);
  const bool im_row = P::arow_im(r);
  const int arow = nr < nrows ? nr : 8;
  const int ml = r & 3, gl = r >> 2;  // column = (band m ml, local field 4 gq + gl)

  acc_t acc1[NGQ];
#pragma unroll
  for (int q = 0; q < NGQ; ++q) acc1[q] = P::zero();
  constexpr int ngq = NGQ;  // == lay.ngp / 4

  load_a(0);
  load_j(0);
  store(0);
  if (CH < n1) {
    load_a(CH);
    load_j(CH);
  }
  int buf = 0;
  for (int i0 = 0; i0 < n1; i0 += CH, buf ^= 1) {
    __syncthreads();  // chunk i0 staged; the other buffer is free
    if (i0 + CH < n1) {
      store(buf ^ 1);
      if (i0 + 2 * CH < n1) {
        load_a(i0 + 2 * CH);
        load_j(i0 + 2 * CH);
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = wave + 4 * h;  // this wave's K groups of the chunk
      const cplx za = A_at(buf, kg, arow)[t];
      const real are = im_row ? -za.y : za.x;
      const real aim = im_row ? za.x : za.y;
      const int16_t *jrow = J_at(buf, 0) + (size_t)gl * CH + 4 * t + kg;  // K slot of this lane
      // all row indices, then all gathered values, then the MFMAs
      int jq[NGQ];
#pragma unroll
      for (int q = 0; q < NGQ; ++q) jq[q] = jrow[(size_t)(4 * q) * CH];
      cplx bq[NGQ];
#pragma unroll
      for (int q = 0; q < NGQ; ++q) bq[q] = sC2[RS * jq[q] + ml];
#pragma unroll
      for (int q = 0; q < NGQ; ++q) {
        acc1[q] = P::mma(are, bq[q].x, acc1[q]);
        acc1[q] = P::mma(aim, bq[q].y, acc1[q]);
      }
    }
  }

  // ---- add the four waves' partial sums through LDS (the k+q block is done)
  __syncthreads();
  real *red = (real *)vsm_raw;  // [wave][q][4 elems][64 lanes]
#pragma unroll
  for (int q = 0; q < NGQ; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) red[((size_t)(wave * ngq + q) * 4 + e) * 64 + lane] = acc1[q][e];
  __syncthreads();
  cplx *rho1 = (cplx *)a.rho1;
  const int nm = pbk.m_hi - pbk.m_lo;
  // wave w finishes the quads q = w, w + 4, ...
  for (int q = wave; q < ngq; q += 4) {
    real v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      v[e] = 0;
#pragma unroll
      for (int w = 0; w < 4; ++w) v[e] += red[((size_t)(w * ngq + q) * 4 + e) * 64 + lane];
    }
    const int g = 4 * q + gl, m = mlo + ml;
    if (g < nlf && ml < nmv) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int onr = P::out_band(kg, p);
        if (onr < nrows) {
          const int n = pbk.n_lo + onr;
          const long long tt = (long long)k * a.pairs.tpk + a.pairs.off[0] +
                               (long long)(n - pbk.n_lo) * nm + (m - pbk.m_lo);
          const long long pos = a.perm ? (long long)a.perm[tt] : tt;
          if (pos >= 0) rho1[pos * nlf + g] = P::make(v[P::out_re(p)], v[P::out_im(p)]);
        }
      }
    }
  }
}

// kbox[k][box(mill)] = position of mill in k's list.  Miller triples of one
// list are distinct, so each slot is written at most once.
__global__ void k_build_kbox(const int *mill_k, const int *ngk, int npwx, Box kb, int16_t *kbox) {
  const int k = blockIdx.y;
  const int n = ngk[k];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int a0, a1, a2;
    mill_unpack(mill_k[(size_t)k * npwx + i], a0, a1, a2);
    const int bi = kb.index(a0, a1, a2);
    if (bi >= 0) kbox[(size_t)k * kb.size() + bi] = (int16_t)i;
  }
}

int launch_build_kbox(rpa_ctx *c) {
  const size_t n = (size_t)c->d.nki * c->kbox.size();
  if (c->d_kbox) (void)hipFree(c->d_kbox);
  c->d_kbox = nullptr;
  RPA_HIP(c, hipMalloc((void **)&c->d_kbox, sizeof(int16_t) * (n ? n : 1)));
  RPA_HIP(c, hipMemsetAsync(c->d_kbox, 0xff, sizeof(int16_t) * n, c->stream));
  dim3 grid(8, c->d.nki);
  hipLaunchKernelGGL(k_build_kbox, grid, dim3(256), 0, c->stream, (const int *)c->d_mill_k,
                     (const int *)c->d_ngk, c->d.npwx, c->kbox, c->d_kbox);
  RPA_HIP(c, hipGetLastError());
  return 0;
}

int launch_direct_map(rpa_ctx *c) {
  if (c->p.vertex_map != RPA_MAP_DIRECT) return 0;
  const long long n = (long long)c->ng0 * c->nlf * c->p.ngd;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_direct_map, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     (const int *)c->d_mill_g, (const int *)c->d_gbox, c->gbox,
                     (const int *)c->d_lf_key, (const int *)c->d_g0keys, (const int *)c->d_csym, c->ng0, c->nlf,
                     c->p.ngd, c->d_dmap);
  RPA_HIP(c, hipGetLastError());
  return 0;
}

template <class P, int KIND, int MAXB, bool DUAL, bool PAIR = false>
static void launch_one(rpa_ctx *c, const VertexArgs &a, dim3 grid, size_t lds) {
  if (a.ch == 32)
    hipLaunchKernelGGL((k_vertex<P, KIND, MAXB, DUAL, 32, PAIR>), grid, dim3(256), lds, c->stream, a);
  else if (a.ch == 16)
    hipLaunchKernelGGL((k_vertex<P, KIND, MAXB, DUAL, 16, PAIR>), grid, dim3(256), lds, c->stream, a);
  else
    hipLaunchKernelGGL((k_vertex<P, KIND, MAXB, DUAL, 8, PAIR>), grid, dim3(256), lds, c->stream, a);
}

// Whether the staged form runs two local fields per wave: charge vertices
// whose band blocks are all at most 8 bands m wide
static bool staged_pair(const VertexArgs &a, int kind) {
  if (kind != 0 || a.nvb > 2) return false;
  for (int b = 0; b < a.nvb; ++b)
    if (a.vb[b].ncols > 8) return false;
  return true;
}

template <class P, int KIND>
static void launch_kind(rpa_ctx *c, const VertexArgs &a, dim3 grid, size_t lds) {
  if constexpr (KIND == 0) {
    if (staged_pair(a, KIND)) {
      VertexArgs pa = a;
      pa.groups = (a.nlf + 7) / 8;
      const dim3 pg((unsigned)((long long)a.kslots * pa.groups));
      if (a.nvb <= 1)
        launch_one<P, KIND, 1, true, true>(c, pa, pg, lds);
      else
        launch_one<P, KIND, 2, true, true>(c, pa, pg, lds);
      return;
    }
  }
  if (a.nvb <= 1)
    launch_one<P, KIND, 1, true>(c, a, grid, lds);
  else if (a.nvb <= 2)
    launch_one<P, KIND, 2, true>(c, a, grid, lds);
  else if (a.nvb <= 4)
    launch_one<P, KIND, 4, false>(c, a, grid, lds);
  else
    launch_one<P, KIND, 8, false>(c, a, grid, lds);
}

// ---- LDS-resident charge-vertex kernel: selection and launch
template <class P>
static const void *lds_kernel(int ngq) {
  switch (ngq) {
    case 2: return (const void *)k_vertex_lds<P, 2>;
    case 4: return (const void *)k_vertex_lds<P, 4>;
    case 8: return (const void *)k_vertex_lds<P, 8>;
    case 13: return (const void *)k_vertex_lds<P, 13>;
    case 16: return (const void *)k_vertex_lds<P, 16>;
    case 20: return (const void *)k_vertex_lds<P, 20>;
    default: return (const void *)k_vertex_lds<P, 24>;
  }
}

template <class P>
static void launch_lds_p(rpa_ctx *c, const VertexArgs &a, dim3 grid, size_t lds, int ngq) {
  switch (ngq) {
    case 2: hipLaunchKernelGGL((k_vertex_lds<P, 2>), grid, dim3(256), lds, c->stream, a); break;
    case 4: hipLaunchKernelGGL((k_vertex_lds<P, 4>), grid, dim3(256), lds, c->stream, a); break;
    case 8: hipLaunchKernelGGL((k_vertex_lds<P, 8>), grid, dim3(256), lds, c->stream, a); break;
    case 13: hipLaunchKernelGGL((k_vertex_lds<P, 13>), grid, dim3(256), lds, c->stream, a); break;
    case 16: hipLaunchKernelGGL((k_vertex_lds<P, 16>), grid, dim3(256), lds, c->stream, a); break;
    case 20: hipLaunchKernelGGL((k_vertex_lds<P, 20>), grid, dim3(256), lds, c->stream, a); break;
    default: hipLaunchKernelGGL((k_vertex_lds<P, 24>), grid, dim3(256), lds, c->stream, a); break;
  }
}

// Whether the LDS-resident kernel serves this context, and its LDS bytes.
// Vertex form asked for by RPA_VERTEX_FORM ("col", "gather", "staged",
// "lds"), else the one the autotuner chose for the current q; empty: chosen
// by shape.
static std::string vertex_form(const rpa_ctx *c) {
  const char *env = std::getenv("RPA_VERTEX_FORM");
  return env ? std::string(env) : c->vform;
}

static bool lds_form(const rpa_ctx *c, size_t &bytes) {
  if (c->d.program != RPA_SLOSS || c->pairs.nblocks != 1) return false;
  const PairBlock &pb = c->pairs.b[0];
  if (pb.n_hi - pb.n_lo > 8 || c->nlf > 4 * kLdsMaxGQ) return false;
  // on request only: measured at N_G = 51 slower than the other forms
  if (vertex_form(c) != "lds") return false;
  const bool mixed = (c->d.precision == RPA_MIXED);
  const LdsLayout L = mixed ? lds_layout<PrecF32>(c->d.npwx, c->nlf) : lds_layout<PrecF64>(c->d.npwx, c->nlf);
  bytes = L.bytes;
  return bytes <= 160 * 1024;
}

// Outside any stream capture: allow the LDS-resident kernel its dynamic LDS.
int configure_vertex(rpa_ctx *c) {
  const std::string form = vertex_form(c);
  if (form.empty() || form == "col") {
    if (int e = build_col_plan(c)) return e;
  } else {
    c->col.ok = false;
  }
  size_t bytes = 0;
  if (!lds_form(c, bytes) || bytes <= 64 * 1024) return 0;
  const int ngq = lds_ngq(c->nlf);
  const void *fn = (c->d.precision == RPA_MIXED) ? lds_kernel<PrecF32>(ngq) : lds_kernel<PrecF64>(ngq);
  RPA_HIP(c, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return 0;
}

template <class P, int KIND>
static void launch_gd(rpa_ctx *c, VertexArgs a, size_t lds) {
  bool pair = true;  // every band block at most 8 bands m wide
  for (int b = 0; b < a.nvb; ++b) pair = pair && a.vb[b].ncols <= 8;
  a.groups = (a.nlf + (pair ? 7 : 3)) / (pair ? 8 : 4);
  const dim3 grid((unsigned)((long long)a.kslots * a.groups));
  if (pair) {
    if (a.nvb <= 1)
      hipLaunchKernelGGL((k_vertex_gd<P, KIND, 1, true, true>), grid, dim3(256), lds, c->stream, a);
    else
      hipLaunchKernelGGL((k_vertex_gd<P, KIND, 2, true, true>), grid, dim3(256), lds, c->stream, a);
  } else if (a.nvb <= 1) {
    hipLaunchKernelGGL((k_vertex_gd<P, KIND, 1, true, false>), grid, dim3(256), lds, c->stream, a);
  } else {
    hipLaunchKernelGGL((k_vertex_gd<P, KIND, 2, true, false>), grid, dim3(256), lds, c->stream, a);
  }
}

template <class P>
static void launch_prec_gd(rpa_ctx *c, int kind, const VertexArgs &a, size_t lds) {
  if (kind == 0)
    launch_gd<P, 0>(c, a, lds);
  else if (kind == 1)
    launch_gd<P, 1>(c, a, lds);
  else
    launch_gd<P, 2>(c, a, lds);
}

template <class P>
static void launch_prec(rpa_ctx *c, int kind, const VertexArgs &a, dim3 grid, size_t lds) {
  if (kind == 0)
    launch_kind<P, 0>(c, a, grid, lds);
  else if (kind == 1)
    launch_kind<P, 1>(c, a, grid, lds);
  else
    launch_kind<P, 2>(c, a, grid, lds);
}

static int launch_block_group(rpa_ctx *c, VertexArgs a, int kind, bool whole);

int launch_vertices(rpa_ctx *c, void *rho1, void *rho2, const int *perm) {
  VertexArgs a{};
  const bool mixed = (c->d.precision == RPA_MIXED);
  a.evc = mixed ? (const void *)c->d_evc32 : (const void *)c->d_evc;
  a.mill_k = c->d_mill_k;
  a.mill_g = c->d_mill_g;
  a.kbox = c->d_kbox;
  a.gbox = c->d_gbox;
  a.dmap = c->d_dmap;
  a.g0idx = c->d_g0idx;
  a.csym = c->d_csym;
  a.crkc = c->d_crkc;
  a.sym = c->qsym ? 1 : 0;
  a.kb = c->kbox;
  a.gb = c->gbox;
  a.ngk = c->d_ngk;
  a.k1 = c->d_k1;
  a.k2 = c->d_k2;
  a.m0 = c->d_m0;
  a.ktot = c->d_ktot;
  a.lf_key = c->d_lf_key;
  a.glf = c->d_glf;
  for (int i = 0; i < 3; ++i) a.q[i] = c->q[i];
  for (int i = 0; i < 9; ++i) a.kc[i] = c->kc[i];
  a.gcar = c->p.gcar;
  a.nki = c->d.nki;
  a.npwx = c->d.npwx;
  a.nband = c->d.nband;
  a.nlf = c->nlf;
  a.ntot = c->ntot;
  a.ngd = c->p.ngd;
  a.map_mode = c->p.vertex_map;
  int kind = 0;
  if (c->d.program == RPA_OPTABS) {
    if (c->p.pol == RPA_POL_YZ) {
      kind = 2;
      a.pol1 = 1;
      a.pol2 = 2;
    } else {
      kind = 1;
      a.pol1 = a.pol2 = c->p.pol - 1;
    }
  }
  a.pairs = c->pairs;
  a.perm = perm;
  a.rho1 = rho1;
  a.rho2 = rho2;
  a.kslots = (c->ntot + 7) / 8 * 8;
  a.groups = (c->nlf + 3) / 4;
  if ((long long)a.kslots * a.groups <= 0) return 0;
  // band blocks of 8 bands n by 16 bands m over every pair block; the kernels
  // take at most kVertexMaxBlocks of them per launch, so longer lists (many
  // bands with all pairs: finite T or metals in Optabs) run as several
  // launches of consecutive blocks of one pair block each
  std::vector<VBlock> all;
  for (int pbi = 0; pbi < c->pairs.nblocks; ++pbi) {
    const PairBlock &pb = c->pairs.b[pbi];
    for (int n0 = pb.n_lo; n0 < pb.n_hi; n0 += 8)
      for (int m0 = pb.m_lo; m0 < pb.m_hi; m0 += 16)
        all.push_back(VBlock{n0, std::min(8, pb.n_hi - n0), m0, std::min(16, pb.m_hi - m0), pbi});
  }
  const std::string form = vertex_form(c);
  const bool col = c->col.ok && (form.empty() || form == "col");
  if (!col)
    if (int e = launch_direct_map(c)) return e;
  const bool whole = all.size() <= (size_t)kVertexMaxBlocks;
  size_t b0 = 0;
  while (b0 < all.size()) {
    size_t b1 = b0;
    while (b1 < all.size() && b1 - b0 < (size_t)kVertexMaxBlocks && (whole || all[b1].pb == all[b0].pb)) ++b1;
    VertexArgs ga = a;
    ga.nvb = (int)(b1 - b0);
    for (int b = 0; b < ga.nvb; ++b) ga.vb[b] = all[b0 + b];
    if (col) {
      if (int e = launch_vertices_col(c, ga.vb, ga.nvb, rho1, rho2, perm)) return e;
    } else if (int e = launch_block_group(c, ga, kind, whole)) {
      return e;
    }
    b0 = b1;
  }
  return 0;
}

// One launch (or one per pair block) of the gather forms over the band
// blocks a.vb[0 .. a.nvb); `whole` when they are every block of the q.
static int launch_block_group(rpa_ctx *c, VertexArgs a, int kind, bool whole) {
  const bool mixed = (c->d.precision == RPA_MIXED);
  const int nvb = a.nvb;
  int nlo = c->d.nband, nhi = 0, mlo = c->d.nband, mhi = 0;
  for (int b = 0; b < nvb; ++b) {
    nlo = std::min(nlo, a.vb[b].n0);
    nhi = std::max(nhi, a.vb[b].n0 + a.vb[b].nrows);
    mlo = std::min(mlo, a.vb[b].m0);
    mhi = std::max(mhi, a.vb[b].m0 + a.vb[b].ncols);
  }
  a.na0 = nlo;
  a.nna = std::max(nhi - nlo, 1);
  a.mb0 = mlo;
  a.nmb = std::max(mhi - mlo, 1);
  // chunk of 32, 16 or 8 plane waves: the staged bands of one row fit the
  // prefetch registers (256/CH threads per C1 row, 64/CH per gathered row)
  // and the staging fits in 64 KB of LDS
  const size_t rsz = mixed ? sizeof(float) : sizeof(double);
  const int nbs = (kind == 2) ? 2 : 1;
  auto lds_of = [&](int ch) {
    const size_t sp = (size_t)ch / 2 + 16 / rsz;
    const size_t arows = 2 * (((size_t)a.nna + 7) / 8) * 8 + 1;
    const size_t bcols = staged_pair(a, kind) ? 17 : (size_t)a.nmb + 1;  // paired fields: 16 columns
    return rsz * sp * (4 * arows + (size_t)nbs * 16 * bcols);
  };
  auto fits = [&](int ch) {
    return a.nna <= (256 / ch) * kPfC1 && a.nmb <= (64 / ch) * kPfB && lds_of(ch) <= 64 * 1024;
  };
  int ch = 32;
  const char *env_ch = std::getenv("RPA_VERTEX_CH");
  const int vch = env_ch ? std::atoi(env_ch) : c->vch;
  if (vch > 0) ch = vch >= 32 ? 32 : vch >= 16 ? 16 : 8;
  while (ch > 8 && !fits(ch)) ch /= 2;
  if (!fits(ch)) return fail(c, RPA_E_LIMIT, "band count too large for the vertex staging");
  a.ch = ch;
  const size_t lds = lds_of(ch);
  const long long nblocks = (long long)a.kslots * a.groups;
  const std::string form = vertex_form(c);
  size_t lds_bytes = 0;
  if (whole && lds_form(c, lds_bytes)) {
    // one workgroup per (k, group of four bands m), all local fields
    const int nmg = (c->pairs.b[0].m_hi - c->pairs.b[0].m_lo + 3) / 4;
    const dim3 g((unsigned)((long long)a.kslots * nmg));
    const int ngq = lds_ngq(c->nlf);
    if (mixed)
      launch_lds_p<PrecF32>(c, a, g, lds_bytes, ngq);
    else
      launch_lds_p<PrecF64>(c, a, g, lds_bytes, ngq);
    RPA_HIP(c, hipGetLastError());
    return 0;
  }
  dim3 grid((unsigned)nblocks);
  // gather-direct kernel (FP32, at most two band blocks spanning <= 16 bands
  // of n): measured faster than the staged form in FP32 and slower in FP64,
  // where its larger register set lowers the occupancy
  const bool gd = mixed && a.nvb <= 2 && a.nna <= 16 && (form.empty() || form == "gather");
  if (gd) {
    const size_t lds_gd = 2 * 4 * (size_t)(a.nna + 1) * kGdNT * (mixed ? sizeof(float2) : sizeof(double2));
    a.ch = kGdCH;
    launch_prec_gd<PrecF32>(c, kind, a, lds_gd);  // gd is an FP32 form
  } else if (mixed && c->pairs.nblocks > 1) {
    // FP32 staged form over several pair blocks (Optabs: both band
    // directions): one launch per pair block, each staging only its own band
    // ranges, so the chunk can be larger (on the 36x36 k-grid with 16 + 24
    // bands, 16 and 32 plane waves instead of 8 for the 40 bands of both
    // directions together: vertices 205 -> 158 ms per 6 q; in FP64 the LDS
    // allows 16 either way, and the two launches measured 3% slower)
    int b0 = 0;
    while (b0 < nvb) {
      int b1 = b0;
      while (b1 < nvb && a.vb[b1].pb == a.vb[b0].pb) ++b1;
      VertexArgs sa = a;
      sa.nvb = b1 - b0;
      for (int b = 0; b < sa.nvb; ++b) sa.vb[b] = a.vb[b0 + b];
      int snlo = c->d.nband, snhi = 0, smlo = c->d.nband, smhi = 0;
      for (int b = 0; b < sa.nvb; ++b) {
        snlo = std::min(snlo, sa.vb[b].n0);
        snhi = std::max(snhi, sa.vb[b].n0 + sa.vb[b].nrows);
        smlo = std::min(smlo, sa.vb[b].m0);
        smhi = std::max(smhi, sa.vb[b].m0 + sa.vb[b].ncols);
      }
      sa.na0 = snlo;
      sa.nna = std::max(snhi - snlo, 1);
      sa.mb0 = smlo;
      sa.nmb = std::max(smhi - smlo, 1);
      auto slds = [&](int chs) {
        const size_t sp = (size_t)chs / 2 + 16 / rsz;
        const size_t arows = 2 * (((size_t)sa.nna + 7) / 8) * 8 + 1;
        return rsz * sp * (4 * arows + (size_t)nbs * 16 * (sa.nmb + 1));
      };
      auto sfits = [&](int chs) {
        return sa.nna <= (256 / chs) * kPfC1 && sa.nmb <= (64 / chs) * kPfB && slds(chs) <= 64 * 1024;
      };
      int sch = vch > 0 ? ch : 32;  // RPA_VERTEX_CH (or the autotuner) caps the chunk
      while (sch > 8 && !sfits(sch)) sch /= 2;
      if (!sfits(sch)) return fail(c, RPA_E_LIMIT, "band count too large for the vertex staging");
      sa.ch = sch;
      launch_prec<PrecF32>(c, kind, sa, grid, slds(sch));
      RPA_HIP(c, hipGetLastError());
      b0 = b1;
    }
  } else if (mixed) {
    launch_prec<PrecF32>(c, kind, a, grid, lds);
  } else {
    launch_prec<PrecF64>(c, kind, a, grid, lds);
  }
  RPA_HIP(c, hipGetLastError());
  return 0;
}

}  // namespace rpa

namespace rpa {

// Autotuner of the vertex kernel: on the first q of each shape it times every
// form that serves the shape -- the column correlation, the FP32
// gather-direct form (mixed precision), and the staged form with its largest
// chunk and with chunks of 16 plane waves -- on the q's own transitions and
// keeps the fastest for the life of the context.  In mixed precision the
// column form (BF16 split products there) is a candidate only where the
// shape rule would take it.  The shape key is (local
// fields, in-plane classes of them, band-pair blocks, program, precision,
// vertex map); the choice replaces the fixed shape rules of the forms.
int autotune_vertex(rpa_ctx *c) {
  std::vector<std::pair<int, int>> cls;
  for (int g = 0; g < c->nlf; ++g) {
    const int *m = &c->mill_g[3 * (size_t)(c->lf_idx[g] - 1)];
    if (std::find(cls.begin(), cls.end(), std::make_pair(m[0], m[1])) == cls.end()) cls.push_back({m[0], m[1]});
  }
  std::string key = std::to_string(c->nlf) + "/" + std::to_string(cls.size()) + "/" + std::to_string(c->pairs.tpk) +
                    "/" + std::to_string(c->pairs.nblocks) + "/" + std::to_string(c->d.program) + "/" +
                    std::to_string(c->d.precision) + "/" + std::to_string(c->p.vertex_map) + "/" +
                    std::to_string(c->p.pol) + "/" + std::to_string((int)c->qsym);
  for (const auto &kv : c->vform_cache)
    if (kv.first == key) {
      c->vform = kv.second.first;
      c->vch = kv.second.second;
      return configure_vertex(c);
    }
  const bool mixed = (c->d.precision == RPA_MIXED);
  std::vector<std::pair<std::string, int>> cand = {{"staged", 0}, {"staged", 16}};
  // the column form of mixed charge vertices runs BF16 hi/lo split products
  // (each within 2^-16 of FP32): a candidate only where the shape rule takes
  // it too (at least 12 local fields per in-plane class), so the tuner never
  // trades accuracy for speed; in FP64 every form gives the same accuracy
  if (!mixed || c->nlf >= 12 * (int)cls.size()) cand.insert(cand.begin(), {"col", 0});
  if (mixed && c->d.program == RPA_SLOSS) cand.insert(cand.begin() + 1, {"gather", 0});
  if (int e = launch_transitions(c)) return e;
  float best = 1e30f;
  std::pair<std::string, int> choice{"", 0};
  for (const auto &cd : cand) {
    c->vform = cd.first;
    c->vch = cd.second;
    if (int e = configure_vertex(c)) return e;
    if (cd.first == "col" && !c->col.ok) continue;
    // one launch to warm up, two timed
    if (launch_vertices(c, c->d_rho1, c->d_rho2, c->d_perm) != 0) continue;  // the form does not take this shape
    RPA_HIP(c, hipEventRecord(c->ev0, c->stream));
    int e = launch_vertices(c, c->d_rho1, c->d_rho2, c->d_perm);
    if (!e) e = launch_vertices(c, c->d_rho1, c->d_rho2, c->d_perm);
    RPA_HIP(c, hipEventRecord(c->ev1, c->stream));
    RPA_HIP(c, hipStreamSynchronize(c->stream));
    if (e) continue;
    float ms = 0.f;
    RPA_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (ms < best) {
      best = ms;
      choice = cd;
    }
  }
  c->err.clear();
  if (choice.first.empty()) return fail(c, RPA_E_LIMIT, "no vertex kernel form takes this shape");
  c->vform_cache.push_back({key, choice});
  c->vform = choice.first;
  c->vch = choice.second;
  return configure_vertex(c);
}

}  // namespace rpa
