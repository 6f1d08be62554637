// Plane-wave vertices for every transition (k, n, m) of the current q.
//
//   charge  rho_nm(k,q,G) = sum_i conj(C1_n(i)) C2_m(map(G,i))
//           (Sloss.f90:883-952 genMnmK1K2; vertices.f90:88-189 _v2)
//   current j_nm(k,q,G)   = sum_i 0.5 conj(C1_n(i)) J(G,i) C2_m(map(G,i)),
//           J = Gcar (q + 2k + Glf(G) + 2 G(i))_pol
//           (Optabs.f90:1354-1472 genStrujniVrhovi; vertices.f90:293-394)
//
// map(G,i) is the reference's Gfast: the position of G(i) + Glf(G) + G0 in the
// k+q coefficient list (Sloss.f90:907-946), found here with exact Miller
// arithmetic in a dense Miller box instead of a linear search (R = RI = I).
//
// For one (k, G) the sums over all band pairs form a complex product
// [conj(C1)]_(n x i) [C2 o map]_(i x m).  It runs on the FP64 matrix cores as
// a real 16x16x4 MFMA: rows are (n, re/im) -- rows 0..7 give Re for 8 bands n,
// rows 8..15 give Im -- the K index is (i, re/im), columns are 16 bands m:
//   A[n][(i,0)] =  C1r   A[n][(i,1)] = C1i   A[8+n][(i,0)] = -C1i  A[8+n][(i,1)] = C1r
//   B[(i,0)][m] =  C2r(map(i))*J           B[(i,1)][m] = C2i(map(i))*J
// One wave owns one G; a workgroup owns 4 G of one k.  Workgroups of one k are
// placed on one XCD so the k+q coefficient rows they gather stay in its L2.
#include "rpa_internal.h"

namespace rpa {

constexpr int kVertexMaxBlocks = 8;

struct VBlock {
  int n0, nrows;  // 8 bands n per block
  int m0, ncols;  // 16 bands m per block
  int pb;         // pair block index
};

struct VertexArgs {
  const double2 *evc;
  const int *mill_k;
  const int *mill_g;
  const int16_t *kbox;
  const int *gbox;
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
  int kslots;      // ntot rounded up to a multiple of 8
  const int *perm;
  double2 *rho1, *rho2;
};

// KIND: 0 charge, 1 current (one polarisation), 2 current (pol yz: two)
template <int KIND>
__global__ __launch_bounds__(256) void k_vertex_f64(VertexArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // XCD-aware placement: linear id L runs on XCD L % 8; all G groups of one
  // k get ids with the same residue.
  const int L = blockIdx.x;
  const int xcd = L & 7;
  const int slot = L >> 3;
  const int k = xcd + 8 * (slot / a.groups);
  const int grp = slot % a.groups;
  if (k >= a.ntot) return;
  const int iG = grp * 4 + wave;
  if (iG >= a.nlf) return;  // whole wave leaves; no block barrier below

  const int K1 = a.k1[k] - 1, K2 = a.k2[k] - 1;
  int d0, d1, d2, l0, l1, l2;
  mill_unpack(a.m0[k], d0, d1, d2);
  mill_unpack(a.lf_key[iG], l0, l1, l2);
  d0 += l0;
  d1 += l1;
  d2 += l2;

  const bool direct = (a.map_mode == RPA_MAP_DIRECT);
  const int n1 = direct ? a.ngd : a.ngk[K1];
  const int lim2 = a.ngk[K2];
  const int *src = direct ? a.mill_g : a.mill_k + (size_t)K1 * a.npwx;
  const int16_t *box2 = a.kbox + (size_t)K2 * a.kb.size();
  const double2 *c1 = a.evc + (size_t)K1 * a.npwx * a.nband;
  const double2 *c2 = a.evc + (size_t)K2 * a.npwx * a.nband;

  const int r = lane & 15, q4 = lane >> 4;
  const int ii = q4 >> 1, part = q4 & 1;

  double jbase1 = 0.0, jbase2 = 0.0;
  if (KIND >= 1) {
    const double *kv = a.ktot + 3 * k;
    jbase1 = (a.q[a.pol1] + 2.0 * kv[a.pol1]) + a.glf[3 * iG + a.pol1];
    jbase2 = (a.q[a.pol2] + 2.0 * kv[a.pol2]) + a.glf[3 * iG + a.pol2];
  }

  rpa_d4 acc1[kVertexMaxBlocks], acc2[kVertexMaxBlocks];
#pragma unroll
  for (int b = 0; b < kVertexMaxBlocks; ++b) {
    acc1[b] = rpa_d4{0.0, 0.0, 0.0, 0.0};
    acc2[b] = rpa_d4{0.0, 0.0, 0.0, 0.0};
  }

  for (int i0 = 0; i0 < n1; i0 += 2) {
    const int i = i0 + ii;
    int j = -1;
    double w1 = 1.0, w2 = 1.0;
    if (i < n1) {
      int s0, s1, s2;
      mill_unpack(src[i], s0, s1, s2);
      if (direct) {
        int bi = a.gb.index(s0 + d0, s1 + d1, s2 + d2);
        int pos = bi >= 0 ? a.gbox[bi] : -1;
        j = (pos >= 0 && pos < lim2) ? pos : -1;
      } else {
        int bi = a.kb.index(s0 + d0, s1 + d1, s2 + d2);
        j = bi >= 0 ? (int)box2[bi] : -1;
      }
      if (KIND >= 1) {
        // G(i) cartesian = KC * mill (left to right), then J * 0.5
        double g1 = ((0.0 + a.kc[a.pol1 + 0] * (double)s0) + a.kc[a.pol1 + 3] * (double)s1) +
                    a.kc[a.pol1 + 6] * (double)s2;
        w1 = 0.5 * ((jbase1 + 2.0 * g1) * a.gcar);
        if (KIND == 2) {
          double g2 = ((0.0 + a.kc[a.pol2 + 0] * (double)s0) + a.kc[a.pol2 + 3] * (double)s1) +
                      a.kc[a.pol2 + 6] * (double)s2;
          w2 = 0.5 * ((jbase2 + 2.0 * g2) * a.gcar);
        }
      }
    }
    const double *c1row = (const double *)(c1 + (size_t)(i < n1 ? i : 0) * a.nband);
    const double *c2row = (const double *)(c2 + (size_t)(j >= 0 ? j : 0) * a.nband);
#pragma unroll
    for (int b = 0; b < kVertexMaxBlocks; ++b) {
      if (b < a.nvb) {
        const VBlock vb = a.vb[b];
        const int nr = r & 7;
        double av = 0.0;
        if (i < n1 && nr < vb.nrows) {
          const double *z = c1row + 2 * (vb.n0 + nr);
          // rows 0..7: (C1r, C1i); rows 8..15: (-C1i, C1r)
          av = (r < 8) ? z[part] : (part == 0 ? -z[1] : z[0]);
        }
        double bv = 0.0;
        if (j >= 0 && r < vb.ncols) bv = c2row[2 * (vb.m0 + r) + part];
        acc1[b] = mfma_f64_16x16x4(av, bv * w1, acc1[b]);
        if (KIND == 2) acc2[b] = mfma_f64_16x16x4(av, bv * w2, acc2[b]);
      }
    }
  }

  // result element e of lane (q4, r) is row 4e + q4: e = 0,1 give Re of band
  // n = 4e + q4, e = 2,3 the matching Im, all in the same lane
#pragma unroll
  for (int b = 0; b < kVertexMaxBlocks; ++b) {
    if (b < a.nvb) {
      const VBlock vb = a.vb[b];
      const PairBlock &pb = a.pairs.b[vb.pb];
      const int nm = pb.m_hi - pb.m_lo;
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int nr = 4 * e + q4;
        if (nr < vb.nrows && r < vb.ncols) {
          const int n = vb.n0 + nr, m = vb.m0 + r;
          const long long t = (long long)k * a.pairs.tpk + a.pairs.off[vb.pb] +
                              (long long)(n - pb.n_lo) * nm + (m - pb.m_lo);
          const long long pos = a.perm ? (long long)a.perm[t] : t;
          if (pos >= 0) {
            a.rho1[pos * a.nlf + iG] = make_double2(acc1[b][e], acc1[b][e + 2]);
            if (KIND == 2) a.rho2[pos * a.nlf + iG] = make_double2(acc2[b][e], acc2[b][e + 2]);
            else if (KIND == 1 && a.rho2) a.rho2[pos * a.nlf + iG] = make_double2(acc1[b][e], acc1[b][e + 2]);
          }
        }
      }
    }
  }
}

// kbox[k][box(mill)] = position of mill in k's list (first occurrence wins as
// the reference's search stops at the first match).
__global__ void k_build_kbox(const int *mill_k, const int *ngk, int npwx, Box kb, int16_t *kbox) {
  const int k = blockIdx.y;
  const int n = ngk[k];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int a0, a1, a2;
    mill_unpack(mill_k[(size_t)k * npwx + i], a0, a1, a2);
    int bi = kb.index(a0, a1, a2);
    if (bi >= 0) {
      // atomicMin on the 32-bit word holding the 16-bit slot is avoided by
      // construction: Miller triples of one list are distinct.
      kbox[(size_t)k * kb.size() + bi] = (int16_t)i;
    }
  }
}

int launch_build_kbox(rpa_ctx *c) {
  const size_t n = (size_t)c->d.nki * c->kbox.size();
  int16_t *dummy = nullptr;
  size_t cap = 0;
  (void)dummy;
  (void)cap;
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

int launch_vertices(rpa_ctx *c, void *rho1, void *rho2, const int *perm) {
  VertexArgs a{};
  a.evc = (const double2 *)c->d_evc;
  a.mill_k = c->d_mill_k;
  a.mill_g = c->d_mill_g;
  a.kbox = c->d_kbox;
  a.gbox = c->d_gbox;
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
  int nvb = 0;
  for (int pbi = 0; pbi < c->pairs.nblocks; ++pbi) {
    const PairBlock &pb = c->pairs.b[pbi];
    for (int n0 = pb.n_lo; n0 < pb.n_hi; n0 += 8)
      for (int m0 = pb.m_lo; m0 < pb.m_hi; m0 += 16) {
        if (nvb >= kVertexMaxBlocks)
          return fail(c, RPA_E_LIMIT, "too many band blocks for the vertex kernel");
        a.vb[nvb++] = VBlock{n0, std::min(8, pb.n_hi - n0), m0, std::min(16, pb.m_hi - m0), pbi};
      }
  }
  a.nvb = nvb;
  a.groups = (c->nlf + 3) / 4;
  a.kslots = (c->ntot + 7) / 8 * 8;
  a.perm = perm;
  a.rho1 = (double2 *)rho1;
  a.rho2 = (double2 *)rho2;
  const long long nblocks = (long long)a.kslots * a.groups;
  if (nblocks <= 0) return 0;
  dim3 grid((unsigned)nblocks);
  if (kind == 0)
    hipLaunchKernelGGL(k_vertex_f64<0>, grid, dim3(256), 0, c->stream, a);
  else if (kind == 1)
    hipLaunchKernelGGL(k_vertex_f64<1>, grid, dim3(256), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_vertex_f64<2>, grid, dim3(256), 0, c->stream, a);
  RPA_HIP(c, hipGetLastError());
  return 0;
}

}  // namespace rpa
