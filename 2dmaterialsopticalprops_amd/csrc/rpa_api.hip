// C ABI of librpa_hip.so: context, inputs, orchestration and copy-outs.
// Kernels live in k_vertex.hip, k_accum.hip, k_kk.hip and k_dyson.hip.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_map>

#include "rpa_internal.h"

namespace rpa {

// the calls that take no context (rpa_interpolate) keep their message per thread
static thread_local std::string g_free_err = "null context";

void set_free_error(const std::string &msg) { g_free_err = msg; }

int fail(rpa_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(rpa_ctx *c, hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (c) c->err = m;
  return RPA_E_HIP;
}

static void invalidate_graph(rpa_ctx *c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  c->graph_exec = nullptr;
  c->graph_valid = false;
}

// A change of the inputs every q shares (G list, basis, bands, parameters,
// profiling): the batch graph and every q's preparation are stale.
static void invalidate_batch(rpa_ctx *c) {
  if (c->batch_exec) (void)hipGraphExecDestroy(c->batch_exec);
  c->batch_exec = nullptr;
  if (c->slots)
    for (rpa_qslot &s : *c->slots) s.prepared = false;
}

// The device buffers of the current q's fields (a slot swapped in).
void free_q_buffers(rpa_ctx *c) {
  void *bufs[] = {c->d_lf_key, c->d_glf, c->d_parg, c->d_k1, c->d_k2, c->d_m0, c->d_ktot, c->d_g0keys,
                  c->d_g0idx, c->d_dmap, c->d_V, c->d_epsinv, c->d_W, c->d_loss, c->d_pi11, c->d_sig,
                  c->d_qeff, c->d_status, c->d_csym, c->d_crkc};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  free_col_plan(c);
  rpa_qslot empty;
  swap_qslot(c, empty);  // every q field back to its initial value
}

static void free_slots(rpa_ctx *c) {
  if (c->batch_exec) (void)hipGraphExecDestroy(c->batch_exec);
  c->batch_exec = nullptr;
  if (!c->slots) return;
  for (rpa_qslot &s : *c->slots) {
    swap_qslot(c, s);
    free_q_buffers(c);
    swap_qslot(c, s);
  }
  delete c->slots;
  c->slots = nullptr;
  if (c->h_bstatus) (void)hipHostFree(c->h_bstatus);
  c->h_bstatus = nullptr;
}

// Which band pairs can carry weight: Sloss n <= Nocc < m (Sloss.f90:355,378);
// Optabs every (n,m) with |f_n - f_m| >= df_cut or n == m (Optabs.f90:360).
// When T = 0 and the first Nval bands lie below E_F at every k, those pairs
// are exactly (occupied, empty) and (empty, occupied); otherwise all pairs
// are enumerated and the device filter decides.
static void build_pairs(rpa_ctx *c) {
  PairTable &pt = c->pairs;
  const int nb = c->d.nband, nv = c->p.nval;
  pt.nblocks = 0;
  if (c->d.program == RPA_SLOSS) {
    pt.b[pt.nblocks++] = PairBlock{0, nv, nv, nb};
  } else {
    bool clean = (c->p.temp == 0.0) && c->p.df_cut > 0.0 && c->p.df_cut <= 1.0;
    for (int k = 0; clean && k < c->d.nki; ++k)
      for (int n = 0; n < nb; ++n) {
        double e = c->E[k + (size_t)c->d.nki * n];
        if ((n < nv && !(e < c->p.efermi)) || (n >= nv && !(e > c->p.efermi))) {
          clean = false;
          break;
        }
      }
    if (clean) {
      pt.b[pt.nblocks++] = PairBlock{0, nv, nv, nb};
      pt.b[pt.nblocks++] = PairBlock{nv, nb, 0, nv};
    } else {
      pt.b[pt.nblocks++] = PairBlock{0, nb, 0, nb};
    }
  }
  int off = 0;
  for (int i = 0; i < pt.nblocks; ++i) {
    pt.off[i] = off;
    off += (pt.b[i].n_hi - pt.b[i].n_lo) * (pt.b[i].m_hi - pt.b[i].m_lo);
  }
  pt.tpk = off;
}

// An environment switch set to "0".
static bool env_off(const char *name) {
  const char *v = std::getenv(name);
  return v && v[0] == '0';
}

int prepare(rpa_ctx *c) {
  if (c->prepared) return 0;
  if (c->ngm == 0) return fail(c, RPA_E_STATE, "rpa_set_gvectors not called");
  for (int k = 0; k < c->d.nki; ++k)
    if (!c->uploaded[k]) return fail(c, RPA_E_STATE, "basis of k-point " + std::to_string(k + 1) + " not uploaded");
  if (c->E.empty()) return fail(c, RPA_E_STATE, "rpa_set_bands not called");
  if (c->nlf == 0) return fail(c, RPA_E_STATE, "rpa_set_local_fields not called");
  if (!c->have_params) return fail(c, RPA_E_STATE, "rpa_set_params not called");
  if (!c->have_q) return fail(c, RPA_E_STATE, "rpa_set_q not called");
  const rpa_params &p = c->p;
  if (p.vertex_map == RPA_MAP_DIRECT) {
    for (int k = 0; k < c->d.nki; ++k)
      if (p.ngd > c->ngk[k])  // Optabs.f90:349-352, 377-380
        return fail(c, RPA_E_ARG, "NGd is bigger than the plane waves of k-point " + std::to_string(k + 1));
    if (p.ngd > c->ngm) return fail(c, RPA_E_ARG, "NGd is bigger than NG");
  }
  invalidate_graph(c);
  if (p.vertex_map == RPA_MAP_IGK && !c->kbox_ready) {
    for (int a = 0; a < 3; ++a) {
      c->kbox.lo[a] = c->mill_min[a];
      c->kbox.n[a] = c->mill_max[a] - c->mill_min[a] + 1;
    }
    if ((long long)c->kbox.size() * c->d.nki > (1LL << 31))
      return fail(c, RPA_E_LIMIT, "Miller box too large");
    if (int e = launch_build_kbox(c)) return e;
    c->kbox_ready = true;
  }
  build_pairs(c);
  const int no = c->d.no;
  c->nw = (c->d.program == RPA_OPTABS) ? 2 * no + 1 : no;
  c->ntiles = (c->nw + kTileW - 1) / kTileW;
  c->wcut = (c->d.program == RPA_OPTABS) ? p.lor_cut / p.eta : 1.0e-3 / p.eta;
  double hw = INFINITY;
  if (c->wcut > 0.0) hw = std::sqrt(std::max(p.eta / c->wcut - p.eta * p.eta, 0.0));
  if (std::isinf(hw)) {
    c->span = c->ntiles;
  } else {
    double bins = 2.0 * hw / p.domega + 6.0;
    c->span = std::min(c->ntiles, (int)std::ceil(bins / kTileW));
  }
  c->ntrans = (long long)c->ntot * c->pairs.tpk;
  const size_t nt = (size_t)std::max<long long>(c->ntrans, 1);
  const size_t nblk = (nt + 255) / 256;
  // mixed S0 in home panels when every window fits a panel of at most
  // kHpMaxNrb row blocks: a window starts in its home tile (offset <= 31)
  // and spans at most floor(2 hw / domega) + 5 rows (+1 for the rounding
  // of its two ends)
  c->hp_nrb = 0;
  if (c->d.precision == RPA_MIXED && !std::isinf(hw) && !env_off("RPA_ACCUM_HP")) {
    const double rows = (kTileW - 1) + std::floor(2.0 * hw / p.domega) + 6.0;
    const int nrb = (int)std::ceil(rows / 16.0);
    if (nrb <= kHpMaxNrb) c->hp_nrb = std::max(nrb, 4);
  }
  c->wl_span = c->hp_nrb ? 0 : c->span;
  c->wl_chunk = c->hp_nrb ? kChunkHp : (c->d.precision == RPA_MIXED) ? kChunkTMx : kChunkT;
  const size_t chunk = (size_t)c->wl_chunk;
  c->max_items = (size_t)(c->wl_span + 1) * ((nt + chunk - 1) / chunk) + c->ntiles + 1;
  const bool two = (c->d.program == RPA_OPTABS && p.pol == RPA_POL_YZ);
  c->ncol = (c->d.program == RPA_OPTABS ? 2 : 1) * c->nlf * c->nlf;
  c->ncol_pad = (c->ncol + kColsWG - 1) / kColsWG * kColsWG;
  const size_t npb = (size_t)c->ncol_pad / kColsWG;
  const size_t nf = (size_t)(no - 1);
  const size_t nn = (size_t)c->nlf * c->nlf;
  int e = 0;
  size_t cap_perm = c->cap_key, cap_nat = c->cap_key;
  if (!e) e = ensure(c, c->d_key, c->cap_key, nt);
  if (!e) e = ensure(c, c->d_perm, cap_perm, nt);
  if (!e) e = ensure(c, c->d_nat, cap_nat, nt);
  if (!e) e = ensure(c, c->d_te, c->cap_te, 3 * nt);
  if (!e) e = ensure(c, c->d_bcnt, c->cap_bcnt, (size_t)(c->ntiles + 1) * nblk);
  if (!e) e = ensure(c, c->d_toff, c->cap_toff, (size_t)c->ntiles + 2);
  if (!e) e = ensure(c, c->d_items, c->cap_items, 3 * c->max_items);
  if (!e) e = ensure(c, c->d_first_item, c->cap_first, (size_t)c->ntiles + 1);
  if (!e) e = ensure(c, c->d_nitems, c->cap_nitems, 2);
  // vertices: complex FP64, or complex FP32 in mixed precision
  const size_t rho_n = nt * c->nlf * (c->d.precision == RPA_MIXED ? 1 : 2);
  if (!e) e = ensure(c, (float2 *&)c->d_rho1, c->cap_rho, rho_n);
  if (!e && two) e = ensure(c, (float2 *&)c->d_rho2, c->cap_rho2, rho_n);
  if (!e && !two && c->d_rho2) {
    (void)hipFree(c->d_rho2);
    c->d_rho2 = nullptr;
    c->cap_rho2 = 0;
  }
  if (!e) e = ensure(c, c->d_s0, c->cap_s0, (size_t)c->nw * c->ncol_pad);
  if (c->hp_nrb) {
    const bool cplx = (c->d.program == RPA_OPTABS);
    const size_t rows = 16 * (size_t)c->hp_nrb;
    const size_t nbl = (size_t)hp_nblocks(c->nlf, cplx);
    if (!e) e = ensure(c, c->d_part, c->cap_part, c->max_items * nbl * rows * 16 * (cplx ? 2 : 1) / 2 + 1);
    if (!e) e = ensure(c, c->d_wtab, c->cap_wtab, c->max_items * (kChunkHp / kHpKS) * 2 * rows * kHpKS);
  } else {
    if (!e) e = ensure(c, c->d_part, c->cap_part, c->max_items * npb * kTileW * kColsWG);
    if (!e && c->d.precision == RPA_MIXED) {
      e = ensure(c, c->d_wmx, c->cap_wmx, c->max_items * kTileW * kChunkTMx);
      if (!e) e = ensure(c, c->d_rbt, c->cap_rbt, 2 * (size_t)c->nlf * nt);
    }
  }
  if (!e && c->d.precision != RPA_MIXED) e = ensure(c, c->d_wt64, c->cap_wt64, c->max_items * kTileW * kChunkT);
  if (!e) e = ensure(c, c->d_kmat, c->cap_kmat, nf * c->nw);
  if (!e) e = ensure(c, c->d_x, c->cap_x, nf * c->ncol_pad);
  if (!e) e = ensure(c, c->d_chi0, c->cap_chi0, nf * nn * 2);
  if (!e) e = ensure(c, c->d_epsinv, c->cap_epsinv, nf * nn * 2);
  if (!e) e = ensure(c, c->d_chi, c->cap_chi, nf * nn * 2);
  if (!e) e = ensure(c, c->d_W, c->cap_W, nf * nn * 2);
  if (!e) e = ensure(c, c->d_V, c->cap_V, nn);
  if (!e) e = ensure(c, c->d_loss, c->cap_loss, nf);
  if (!e) e = ensure(c, c->d_pi11, c->cap_pi11, 2 * nf);
  if (!e) e = ensure(c, c->d_sig, c->cap_sig, 2 * nf);
  if (!e) e = ensure(c, c->d_qeff, c->cap_qeff, 2 * nn);
  if (!e) e = ensure(c, c->d_wkk, c->cap_wkk, 2 * nf);
  if (!e) e = ensure(c, c->d_status, c->cap_status, 1);
  if (!e && p.vertex_map == RPA_MAP_DIRECT)
    e = ensure(c, c->d_dmap, c->cap_dmap, (size_t)c->ng0 * c->nlf * p.ngd);
  if (e) return e;
  RPA_HIP(c, hipMemsetAsync(c->d_qeff, 0, sizeof(double) * 2 * nn, c->stream));
  if (int e2 = launch_kstencil(c)) return e2;
  if (c->d.program == RPA_SLOSS)
    if (int e2 = configure_dyson(c)) return e2;
  if (c->autotune && !std::getenv("RPA_VERTEX_FORM")) {
    if (int e2 = autotune_vertex(c)) return e2;
  } else if (int e2 = configure_vertex(c)) {
    return e2;
  }
  if (int e2 = configure_accumulate(c)) return e2;
  RPA_HIP(c, hipStreamSynchronize(c->stream));
  c->prepared = true;
  return 0;
}

// The per-q chain after prepare(): bookkeeping, vertices, S0, KK, Dyson/W or Pi.
// With profiling on, events mark the stage boundaries (prof_ev[3] is
// recorded inside launch_accumulate between the matrix-core kernel and the
// reduction).
int chain(rpa_ctx *c) {
  const bool pr = c->profiling;
  if (pr) RPA_HIP(c, hipEventRecord(c->prof_ev[0], c->stream));
  if (int e = launch_transitions(c)) return e;
  if (pr) RPA_HIP(c, hipEventRecord(c->prof_ev[1], c->stream));
  if (int e = launch_vertices(c, c->d_rho1, c->d_rho2, c->d_perm)) return e;
  if (pr) RPA_HIP(c, hipEventRecord(c->prof_ev[2], c->stream));
  if (int e = launch_accumulate(c)) return e;
  if (pr) RPA_HIP(c, hipEventRecord(c->prof_ev[4], c->stream));
  if (int e = launch_kk(c)) return e;
  if (pr) RPA_HIP(c, hipEventRecord(c->prof_ev[5], c->stream));
  if (c->d.program == RPA_SLOSS) {
    if (int e = launch_coulomb(c)) return e;
    if (int e = launch_dyson(c)) return e;
  } else {
    if (int e = launch_epilogue(c)) return e;
  }
  if (pr) RPA_HIP(c, hipEventRecord(c->prof_ev[6], c->stream));
  return 0;
}

static int collect_stage_times(rpa_ctx *c) {
  for (int s = 0; s < 6; ++s) {
    float ms = 0.f;
    RPA_HIP(c, hipEventElapsedTime(&ms, c->prof_ev[s], c->prof_ev[s + 1]));
    c->stage_ms[s] = ms;
  }
  return 0;
}

static int sync(rpa_ctx *c) {
  RPA_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

// The Dyson kernels raise d_status when a pivot is zero (the reference's
// gjel prints 'singular matrix in gjel', Chi_RPA/gjel.f:59-61, and goes on
// with the division).  The flag is cleared by a kernel ahead of every Dyson
// launch; after a run it is copied to pinned memory on the stream and read
// once the stream has drained.  The results stay on the device, so a caller
// can still fetch them after RPA_E_SINGULAR.
static int queue_status(rpa_ctx *c) {
  if (c->d.program != RPA_SLOSS) return 0;
  RPA_HIP(c, hipMemcpyAsync(c->h_status, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  return 0;
}

static int check_status(rpa_ctx *c) {
  if (c->d.program == RPA_SLOSS && c->h_status[0] != 0)
    return fail(c, RPA_E_SINGULAR, "singular matrix in the Dyson step (eps = I - chi0 V has a zero pivot)");
  return 0;
}

template <class T>
static int d2h(rpa_ctx *c, T *dst, const T *src, size_t n) {
  RPA_HIP(c, hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
  RPA_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

// [f][i][j] complex (device) -> (f, iG, jG) column-major complex (reference)
static int copy_freq_mats(rpa_ctx *c, const double *dev, double *out) {
  const size_t nf = (size_t)(c->d.no - 1), N = (size_t)c->nlf;
  std::vector<double> h(nf * N * N * 2);
  if (int e = d2h(c, h.data(), dev, h.size())) return e;
  for (size_t f = 0; f < nf; ++f)
    for (size_t i = 0; i < N; ++i)
      for (size_t j = 0; j < N; ++j) {
        size_t src = ((f * N + i) * N + j) * 2, dst = (f + nf * (i + N * j)) * 2;
        out[dst] = h[src];
        out[dst + 1] = h[src + 1];
      }
  return 0;
}

// S0 in the reference layout: Sloss S0(no,Nlf,Nlf) real, Optabs
// S0(-no:no,Nlf,Nlf) complex; device [w][pair] with pair = i*Nlf + j
static int get_s0(rpa_ctx *c, double *out) {
  const int nw = c->nw, N = c->nlf;
  std::vector<double> h((size_t)nw * c->ncol_pad);
  if (int e = d2h(c, h.data(), (const double *)c->d_s0, h.size())) return e;
  for (int w = 0; w < nw; ++w)
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        const size_t pr = (size_t)i * N + j;
        const size_t dst = (size_t)w + (size_t)nw * (i + (size_t)N * j);
        if (c->d.program == RPA_SLOSS) {
          out[dst] = h[(size_t)w * c->ncol_pad + pr];
        } else {
          out[2 * dst] = h[(size_t)w * c->ncol_pad + 2 * pr];
          out[2 * dst + 1] = h[(size_t)w * c->ncol_pad + 2 * pr + 1];
        }
      }
  return 0;
}

// square (Nlf,Nlf) matrix, device row-major [i][j] -> column-major
static int get_square(rpa_ctx *c, const double *dev, int ncomp, double *out) {
  const size_t N = (size_t)c->nlf;
  std::vector<double> h(N * N * ncomp);
  if (int e = d2h(c, h.data(), dev, h.size())) return e;
  for (size_t i = 0; i < N; ++i)
    for (size_t j = 0; j < N; ++j)
      for (int q = 0; q < ncomp; ++q) out[(i + N * j) * ncomp + q] = h[(i * N + j) * ncomp + q];
  return 0;
}

// sig(no-1, 2): device interleaves (inter, intra) per frequency
static int get_sig(rpa_ctx *c, double *sig) {
  const size_t nf = (size_t)(c->d.no - 1);
  std::vector<double> h(2 * nf);
  if (int e = d2h(c, h.data(), (const double *)c->d_sig, 2 * nf)) return e;
  for (size_t f = 0; f < nf; ++f) {
    sig[f] = h[2 * f];
    sig[nf + f] = h[2 * f + 1];
  }
  return 0;
}

int upload(rpa_ctx *c, const Upload *parts, int n) {
  size_t total = 0;
  for (int i = 0; i < n; ++i) total += (parts[i].bytes + 255) / 256 * 256;
  // the previous batch's copies read the staging buffer until they complete
  RPA_HIP(c, hipStreamSynchronize(c->stream));
  if (total > c->cap_stage) {
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    c->h_stage = nullptr;
    c->cap_stage = 0;
    const size_t cap = std::max<size_t>(total, 1 << 20);
    RPA_HIP(c, hipHostMalloc((void **)&c->h_stage, cap, hipHostMallocDefault));
    c->cap_stage = cap;
  }
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    if (parts[i].bytes == 0) continue;
    std::memcpy(c->h_stage + off, parts[i].src, parts[i].bytes);
    RPA_HIP(c, hipMemcpyAsync(parts[i].dst, c->h_stage + off, parts[i].bytes, hipMemcpyHostToDevice, c->stream));
    off += (parts[i].bytes + 255) / 256 * 256;
  }
  return 0;
}

}  // namespace rpa

using namespace rpa;

extern "C" {

int rpa_ctx_create(const rpa_dims *dims, rpa_ctx **out) {
  if (!dims || !out) return RPA_E_ARG;
  *out = nullptr;
  const rpa_dims &d = *dims;
  if (d.nki <= 0 || d.nband <= 0 || d.npwx <= 0 || d.nlfd <= 0 || d.no < 4) return RPA_E_ARG;
  if (d.program != RPA_SLOSS && d.program != RPA_OPTABS) return RPA_E_ARG;
  if (d.precision != RPA_FP64 && d.precision != RPA_MIXED) return RPA_E_ARG;
  if (d.npwx > 32767) return RPA_E_LIMIT;
  if (d.nlfd > kMaxNlf) return RPA_E_LIMIT;
  rpa_ctx *c = new rpa_ctx();
  c->d = d;
  hipError_t e = hipSetDevice(d.device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  const size_t nel = (size_t)d.nki * d.npwx;
  if (e == hipSuccess) e = hipMalloc(&c->d_evc, sizeof(double2) * nel * d.nband);
  if (e == hipSuccess) e = hipMalloc((void **)&c->d_mill_k, sizeof(int) * nel);
  if (e == hipSuccess) e = hipMemset(c->d_mill_k, 0, sizeof(int) * nel);
  if (e == hipSuccess) e = hipMemset(c->d_evc, 0, sizeof(double2) * nel * d.nband);
  if (e == hipSuccess && d.precision == RPA_MIXED) {
    e = hipMalloc((void **)&c->d_evc32, sizeof(float2) * nel * d.nband);
    if (e == hipSuccess) e = hipMemset(c->d_evc32, 0, sizeof(float2) * nel * d.nband);
  }
  if (e == hipSuccess && noalloc_init_device(d.device) != 0) e = hipErrorInvalidDevice;
  if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_status, sizeof(int), hipHostMallocDefault);
  if (e == hipSuccess) c->h_status[0] = 0;
  if (e == hipSuccess) e = hipMalloc((void **)&c->d_ngk, sizeof(int) * d.nki);
  if (e == hipSuccess) e = hipMalloc((void **)&c->d_E, sizeof(double) * d.nki * d.nband);
  if (e != hipSuccess) {
    rpa_ctx_destroy(c);
    return RPA_E_HIP;
  }
  c->ngk.assign(d.nki, 0);
  c->uploaded.assign(d.nki, 0);
  *out = c;
  return RPA_OK;
}

int rpa_ctx_destroy(rpa_ctx *c) {
  if (!c) return RPA_OK;
  (void)hipSetDevice(c->d.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  invalidate_graph(c);
  free_slots(c);
  void *bufs[] = {c->d_mill_g, c->d_gbox, c->d_evc, c->d_mill_k, c->d_kbox, c->d_ngk, c->d_E,
                  c->d_lf_key, c->d_glf, c->d_parg, c->d_k1, c->d_k2, c->d_m0, c->d_ktot,
                  c->d_key, c->d_perm, c->d_bcnt, c->d_toff, c->d_nat, c->d_te, c->d_items,
                  c->d_nitems, c->d_first_item, c->d_intra, c->d_nintra, c->d_rho1, c->d_rho2,
                  c->d_s0, c->d_part, c->d_qeff, c->d_kmat, c->d_x, c->d_chi0, c->d_V,
                  c->d_epsinv, c->d_chi, c->d_W, c->d_loss, c->d_pi11, c->d_sig,
                  c->d_status, c->d_g0keys, c->d_g0idx, c->d_dmap, c->d_wkk, c->d_evc32,
                  c->d_wmx, c->d_rbt, c->d_wt64, c->d_box[0], c->d_box[1], c->d_box_gidx,
                  c->d_csym, c->d_crkc, c->d_wtab, c->d_boxw};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  free_col_plan(c);
  for (hipEvent_t ev : c->prof_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->h_status) (void)hipHostFree(c->h_status);
  delete c;
  return RPA_OK;
}

const char *rpa_last_error(const rpa_ctx *c) { return c ? c->err.c_str() : g_free_err.c_str(); }

double rpa_last_run_ms(const rpa_ctx *c) { return c ? c->last_ms : 0.0; }

int rpa_set_gvectors(rpa_ctx *c, int ngm, const int *mill, const double *kc) {
  if (!c || ngm <= 0 || !mill || !kc) return fail(c, RPA_E_ARG, "rpa_set_gvectors: bad arguments");
  (void)hipSetDevice(c->d.device);
  if (mill[0] != 0 || mill[1] != 0 || mill[2] != 0)  // Optabs_qe6.f90:1153-1157
    return fail(c, RPA_E_ARG, "G vectors input is wrong: G(1) is not (0,0,0)");
  c->ngm = ngm;
  c->mill_g.assign(mill, mill + 3 * (size_t)ngm);
  std::memcpy(c->kc, kc, sizeof(double) * 9);
  int lo[3] = {1 << 20, 1 << 20, 1 << 20}, hi[3] = {-(1 << 20), -(1 << 20), -(1 << 20)};
  std::vector<int> keys(ngm);
  for (int i = 0; i < ngm; ++i) {
    for (int a = 0; a < 3; ++a) {
      int v = mill[3 * i + a];
      if (v < -kMillBias || v >= kMillBias) return fail(c, RPA_E_LIMIT, "Miller index out of range");
      lo[a] = std::min(lo[a], v);
      hi[a] = std::max(hi[a], v);
    }
    keys[i] = mill_pack(mill[3 * i], mill[3 * i + 1], mill[3 * i + 2]);
  }
  for (int a = 0; a < 3; ++a) {
    c->gbox.lo[a] = lo[a];
    c->gbox.n[a] = hi[a] - lo[a] + 1;
  }
  std::vector<int> box((size_t)c->gbox.size(), -1);
  for (int i = 0; i < ngm; ++i) {
    int bi = c->gbox.index(mill[3 * i], mill[3 * i + 1], mill[3 * i + 2]);
    if (box[bi] < 0) box[bi] = i;
  }
  c->h_gbox = box;
  if (c->d_mill_g) (void)hipFree(c->d_mill_g);
  if (c->d_gbox) (void)hipFree(c->d_gbox);
  c->d_mill_g = c->d_gbox = nullptr;
  RPA_HIP(c, hipMalloc((void **)&c->d_mill_g, sizeof(int) * ngm));
  RPA_HIP(c, hipMalloc((void **)&c->d_gbox, sizeof(int) * box.size()));
  RPA_HIP(c, hipMemcpy(c->d_mill_g, keys.data(), sizeof(int) * ngm, hipMemcpyHostToDevice));
  RPA_HIP(c, hipMemcpy(c->d_gbox, box.data(), sizeof(int) * box.size(), hipMemcpyHostToDevice));
  // local fields and the k+q map index the G list: both must be set again
  // for the new list before the next run
  c->nlf = 0;
  c->lf_idx.clear();
  c->parg.clear();
  c->have_q = false;
  c->ng0 = 0;
  c->g0_list.clear();
  c->box_ready = false;
  // the point operations' Miller forms depend on KC
  c->nsym = 1;
  c->op_m.clear();
  c->op_minv.clear();
  c->op_rkc.clear();
  c->prepared = false;
  invalidate_graph(c);
  invalidate_batch(c);
  return RPA_OK;
}

int rpa_upload_basis(rpa_ctx *c, int ik, int ngk, const int *mill, const double *evc) {
  if (!c || !mill || !evc) return fail(c, RPA_E_ARG, "rpa_upload_basis: null argument");
  if (ik < 1 || ik > c->d.nki) return fail(c, RPA_E_ARG, "rpa_upload_basis: ik out of range");
  if (ngk < 1 || ngk > c->d.npwx) return fail(c, RPA_E_ARG, "rpa_upload_basis: ngk exceeds npwx");
  (void)hipSetDevice(c->d.device);
  const int k = ik - 1, nb = c->d.nband;
  std::vector<int> keys(c->d.npwx, 0);
  for (int i = 0; i < ngk; ++i) {
    for (int a = 0; a < 3; ++a) {
      int v = mill[3 * i + a];
      if (v < -kMillBias || v >= kMillBias) return fail(c, RPA_E_LIMIT, "Miller index out of range");
      c->mill_min[a] = std::min(c->mill_min[a], v);
      c->mill_max[a] = std::max(c->mill_max[a], v);
    }
    keys[i] = mill_pack(mill[3 * i], mill[3 * i + 1], mill[3 * i + 2]);
  }
  // evc(ngk, nband) column-major -> rows of all bands per plane wave
  std::vector<double> rows((size_t)c->d.npwx * nb * 2, 0.0);
  for (int n = 0; n < nb; ++n)
    for (int i = 0; i < ngk; ++i) {
      rows[((size_t)i * nb + n) * 2] = evc[((size_t)i + (size_t)ngk * n) * 2];
      rows[((size_t)i * nb + n) * 2 + 1] = evc[((size_t)i + (size_t)ngk * n) * 2 + 1];
    }
  RPA_HIP(c, hipMemcpy((double2 *)c->d_evc + (size_t)k * c->d.npwx * nb, rows.data(),
                       sizeof(double) * rows.size(), hipMemcpyHostToDevice));
  if (c->d.precision == RPA_MIXED) {
    std::vector<float> rows32(rows.size());
    for (size_t x = 0; x < rows.size(); ++x) rows32[x] = (float)rows[x];
    RPA_HIP(c, hipMemcpy(c->d_evc32 + (size_t)k * c->d.npwx * nb, rows32.data(),
                         sizeof(float) * rows32.size(), hipMemcpyHostToDevice));
  }
  RPA_HIP(c, hipMemcpy(c->d_mill_k + (size_t)k * c->d.npwx, keys.data(), sizeof(int) * c->d.npwx,
                       hipMemcpyHostToDevice));
  RPA_HIP(c, hipMemcpy(c->d_ngk + k, &ngk, sizeof(int), hipMemcpyHostToDevice));
  c->ngk[k] = ngk;
  c->uploaded[k] = 1;
  c->kbox_ready = false;
  c->box_ready = false;
  c->prepared = false;
  invalidate_graph(c);
  invalidate_batch(c);
  return RPA_OK;
}

int rpa_set_bands(rpa_ctx *c, const double *E) {
  if (!c || !E) return fail(c, RPA_E_ARG, "rpa_set_bands: null argument");
  (void)hipSetDevice(c->d.device);
  const size_t n = (size_t)c->d.nki * c->d.nband;
  c->E.assign(E, E + n);
  RPA_HIP(c, hipMemcpy(c->d_E, E, sizeof(double) * n, hipMemcpyHostToDevice));
  c->prepared = false;
  invalidate_graph(c);
  invalidate_batch(c);
  return RPA_OK;
}

int rpa_set_local_fields(rpa_ctx *c, int nlf, const int *lf_idx, const int *parg) {
  if (!c || !lf_idx || !parg || nlf <= 0) return fail(c, RPA_E_ARG, "rpa_set_local_fields: bad arguments");
  if (nlf > c->d.nlfd) return fail(c, RPA_E_LIMIT, "Nlf is bigger than Nlfd");  // Optabs.f90:1054-1057
  if (c->ngm == 0) return fail(c, RPA_E_STATE, "rpa_set_gvectors must precede rpa_set_local_fields");
  (void)hipSetDevice(c->d.device);
  std::vector<int> keys(nlf);
  std::vector<double> glf(3 * (size_t)nlf);
  for (int i = 0; i < nlf; ++i) {
    int g = lf_idx[i] - 1;
    if (g < 0 || g >= c->ngm) return fail(c, RPA_E_ARG, "local-field index outside the G list");
    if (parg[i] != 1 && parg[i] != -1) return fail(c, RPA_E_ARG, "parity must be +1 or -1");
    const int *m = &c->mill_g[3 * (size_t)g];
    keys[i] = mill_pack(m[0], m[1], m[2]);
    for (int a = 0; a < 3; ++a) {
      double s = 0.0;
      for (int b = 0; b < 3; ++b) s = s + c->kc[a + 3 * b] * (double)m[b];
      glf[3 * i + a] = s;
    }
  }
  c->nlf = nlf;
  c->lf_idx.assign(lf_idx, lf_idx + nlf);
  c->parg.assign(parg, parg + nlf);
  int e = ensure(c, c->d_lf_key, c->cap_lf_key, (size_t)nlf);
  if (!e) e = ensure(c, c->d_glf, c->cap_glf, 3 * (size_t)nlf);
  if (!e) e = ensure(c, c->d_parg, c->cap_parg, (size_t)nlf);
  if (e) return e;
  const Upload up[] = {{c->d_lf_key, keys.data(), sizeof(int) * nlf},
                       {c->d_glf, glf.data(), sizeof(double) * 3 * nlf},
                       {c->d_parg, parg, sizeof(int) * nlf}};
  if (int e2 = upload(c, up, 3)) return e2;
  c->prepared = false;
  invalidate_graph(c);
  return RPA_OK;
}

int rpa_set_params(rpa_ctx *c, const rpa_params *p) {
  if (!c || !p) return fail(c, RPA_E_ARG, "rpa_set_params: null argument");
  if (p->nval < 0 || p->nval > c->d.nband) return fail(c, RPA_E_ARG, "nval outside 0..Nband");
  if (p->vertex_map != RPA_MAP_DIRECT && p->vertex_map != RPA_MAP_IGK)
    return fail(c, RPA_E_ARG, "unknown vertex map");
  if (p->vertex_map == RPA_MAP_DIRECT && p->ngd <= 0) return fail(c, RPA_E_ARG, "NGd must be positive");
  if (!(p->eta > 0.0) || !(p->domega > 0.0) || !(p->vcell > 0.0))
    return fail(c, RPA_E_ARG, "eta, domega and Vcell must be positive");
  if (c->d.program == RPA_OPTABS && (p->pol < RPA_POL_X || p->pol > RPA_POL_YZ))
    return fail(c, RPA_E_ARG, "unknown polarisation");
  if (c->d.program == RPA_SLOSS && (p->nval == 0 || p->nval == c->d.nband))
    return fail(c, RPA_E_ARG, "Sloss needs occupied and empty bands");
  c->p = *p;
  c->have_params = true;
  c->prepared = false;
  invalidate_graph(c);
  invalidate_batch(c);
  return RPA_OK;
}

// Point operations R(48,3,3) (Chi_RPA/PointR.f90:41-138).  Each must map the
// reciprocal lattice onto itself: its Miller form KC^-1 R KC, and that of its
// inverse RI (rpa_point_inverse, PointR's gjel), are integer matrices within
// 1e-3.  The reference finds RI(R2) (R(R1) G + Glf + G0) in the G list by a
// tolerance search; for a lattice operation that is the vector of integer
// Miller index A m + B (Glf + G0) used here, A = M(RI(R2)) M(R(R1)),
// B = M(RI(R2)).
int rpa_set_symmetry(rpa_ctx *c, int nsym, const double *R) {
  if (!c || nsym < 1 || nsym > 48 || !R) return fail(c, RPA_E_ARG, "rpa_set_symmetry: bad arguments");
  if (c->ngm == 0) return fail(c, RPA_E_STATE, "rpa_set_gvectors must precede rpa_set_symmetry");
  std::vector<double> RI(48 * 9, 0.0);
  if (rpa_point_inverse(nsym, R, RI.data()) != RPA_OK)
    return fail(c, RPA_E_ARG, "a point operation is singular (gjel)");
  // KC^-1 by cofactors, column-major like KC
  const double *K = c->kc;
  auto at = [&](int i, int j) { return K[i + 3 * j]; };
  const double det = at(0, 0) * (at(1, 1) * at(2, 2) - at(1, 2) * at(2, 1)) -
                     at(0, 1) * (at(1, 0) * at(2, 2) - at(1, 2) * at(2, 0)) +
                     at(0, 2) * (at(1, 0) * at(2, 1) - at(1, 1) * at(2, 0));
  if (!(std::fabs(det) > 0.0)) return fail(c, RPA_E_ARG, "rpa_set_symmetry: KC is singular");
  double Ki[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
      Ki[i + 3 * j] = (at(i1, j1) * at(i2, j2) - at(i1, j2) * at(i2, j1)) / det;
    }
  // Miller form (row-major) of the cartesian operation Q(op) (R or RI)
  auto miller = [&](const double *Q, int op, int *out) -> bool {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double t = 0.0;
        for (int l = 0; l < 3; ++l)
          for (int n = 0; n < 3; ++n) t += Ki[i + 3 * l] * Q[op + 48 * (l + 3 * n)] * K[n + 3 * j];
        const double r = std::nearbyint(t);
        if (std::fabs(t - r) > 1e-3) return false;
        out[3 * i + j] = (int)r;
      }
    return true;
  };
  std::vector<int> m(9 * (size_t)nsym), mi(9 * (size_t)nsym);
  std::vector<double> rkc(9 * (size_t)nsym);
  for (int op = 0; op < nsym; ++op) {
    if (!miller(R, op, &m[9 * op]) || !miller(RI.data(), op, &mi[9 * op]))
      return fail(c, RPA_E_ARG,
                  "point operation " + std::to_string(op + 1) + " does not map the reciprocal lattice onto itself");
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        int t = 0;
        for (int l = 0; l < 3; ++l) t += m[9 * op + 3 * i + l] * mi[9 * op + 3 * l + j];
        if (t != (i == j ? 1 : 0))
          return fail(c, RPA_E_ARG, "point operation " + std::to_string(op + 1) + ": RI is not its inverse");
      }
    // R(op) KC, column-major, sums left to right
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        rkc[9 * op + i + 3 * j] = (R[op + 48 * i] * K[3 * j] + R[op + 48 * (i + 3)] * K[1 + 3 * j]) +
                                  R[op + 48 * (i + 6)] * K[2 + 3 * j];
  }
  c->nsym = nsym;
  c->op_m = m;
  c->op_minv = mi;
  c->op_rkc = rkc;
  // the k+q map refers to operations by number: set it again
  c->have_q = false;
  c->prepared = false;
  invalidate_graph(c);
  invalidate_batch(c);
  return RPA_OK;
}

// The k+q map of a q: K1, K2, iG0 and, with point operations, R1 and R2 of
// every FBZ k (1-based; R1 = R2 = null: the identity).
static int set_q_impl(rpa_ctx *c, int ntot, const int *K1, const int *K2, const int *iG0, const int *R1,
                      const int *R2, const double *ktot, const double *q) {
  if (!c || ntot <= 0 || !K1 || !K2 || !iG0 || !ktot || !q || (!R1) != (!R2))
    return fail(c, RPA_E_ARG, "rpa_set_q: bad arguments");
  if (c->ngm == 0) return fail(c, RPA_E_STATE, "rpa_set_gvectors must precede rpa_set_q");
  (void)hipSetDevice(c->d.device);
  const bool ops = R1 != nullptr;
  if (ops && c->op_m.empty()) return fail(c, RPA_E_STATE, "rpa_set_symmetry must precede rpa_set_q_sym");
  auto is_identity = [&](int op) {  // 0-based
    if (!ops) return true;
    for (int x = 0; x < 9; ++x) {
      const int e = (x % 4 == 0) ? 1 : 0;
      if (c->op_m[9 * op + x] != e || c->op_minv[9 * op + x] != e || c->op_rkc[9 * op + x] != c->kc[x]) return false;
    }
    return true;
  };
  std::vector<int> m0(ntot), g0idx(ntot), g0keys, csym;
  std::vector<double> crkc;
  std::vector<int> g0seen;  // iG0 of each combination
  std::unordered_map<long long, int> combo;
  bool qsym = false;
  for (int k = 0; k < ntot; ++k) {
    if (K1[k] < 1 || K1[k] > c->d.nki || K2[k] < 1 || K2[k] > c->d.nki)
      return fail(c, RPA_E_ARG, "K1/K2 outside 1..NkI");
    if (iG0[k] < 1 || iG0[k] > c->ngm) return fail(c, RPA_E_ARG, "iG0 outside 1..NG");
    const int r1 = ops ? R1[k] - 1 : 0, r2 = ops ? R2[k] - 1 : 0;
    if (r1 < 0 || r1 >= c->nsym || r2 < 0 || r2 >= c->nsym) return fail(c, RPA_E_ARG, "R1/R2 outside 1..Nsym");
    const int *m = &c->mill_g[3 * (size_t)(iG0[k] - 1)];
    m0[k] = mill_pack(m[0], m[1], m[2]);
    const long long key = ((long long)r1 * 48 + r2) * (long long)c->ngm + (iG0[k] - 1);
    auto it = combo.find(key);
    int z;
    if (it == combo.end()) {
      z = (int)g0seen.size();
      combo.emplace(key, z);
      g0seen.push_back(iG0[k]);
      g0keys.push_back(m0[k]);
      int A[9], B[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          B[3 * i + j] = ops ? c->op_minv[9 * r2 + 3 * i + j] : (i == j);
          int t = 0;
          for (int l = 0; l < 3; ++l)
            t += (ops ? c->op_minv[9 * r2 + 3 * i + l] : (i == l)) * (ops ? c->op_m[9 * r1 + 3 * l + j] : (l == j));
          A[3 * i + j] = t;
        }
      csym.insert(csym.end(), A, A + 9);
      csym.insert(csym.end(), B, B + 9);
      for (int x = 0; x < 9; ++x) crkc.push_back(ops ? c->op_rkc[9 * r1 + x] : c->kc[x]);
      qsym = qsym || !is_identity(r1) || !is_identity(r2);
    } else {
      z = it->second;
    }
    g0idx[k] = z;
  }
  c->ntot = ntot;
  c->K1.assign(K1, K1 + ntot);
  c->K2.assign(K2, K2 + ntot);
  c->iG0.assign(iG0, iG0 + ntot);
  if (ops) {
    c->kR1.assign(R1, R1 + ntot);
    c->kR2.assign(R2, R2 + ntot);
  } else {
    c->kR1.clear();
    c->kR2.clear();
  }
  c->qsym = qsym;
  c->ktot.assign(ktot, ktot + 3 * (size_t)ntot);
  for (int a = 0; a < 3; ++a) c->q[a] = q[a];
  c->ng0 = (int)g0keys.size();
  c->g0_list = g0seen;
  const size_t nt = (size_t)ntot;
  int e = ensure(c, c->d_k1, c->cap_k1, nt);
  if (!e) e = ensure(c, c->d_k2, c->cap_k2, nt);
  if (!e) e = ensure(c, c->d_m0, c->cap_m0, nt);
  if (!e) e = ensure(c, c->d_ktot, c->cap_ktot, 3 * nt);
  if (!e) e = ensure(c, c->d_g0keys, c->cap_g0keys, (size_t)c->ng0);
  if (!e) e = ensure(c, c->d_g0idx, c->cap_g0idx, nt);
  if (!e) e = ensure(c, c->d_csym, c->cap_csym, 18 * (size_t)c->ng0);
  if (!e) e = ensure(c, c->d_crkc, c->cap_crkc, 9 * (size_t)c->ng0);
  if (e) return e;
  const Upload up[] = {{c->d_k1, K1, sizeof(int) * nt},
                       {c->d_k2, K2, sizeof(int) * nt},
                       {c->d_m0, m0.data(), sizeof(int) * nt},
                       {c->d_ktot, ktot, sizeof(double) * 3 * nt},
                       {c->d_g0keys, g0keys.data(), sizeof(int) * c->ng0},
                       {c->d_g0idx, g0idx.data(), sizeof(int) * nt},
                       {c->d_csym, csym.data(), sizeof(int) * csym.size()},
                       {c->d_crkc, crkc.data(), sizeof(double) * crkc.size()}};
  if (int e2 = upload(c, up, 8)) return e2;
  c->have_q = true;
  c->prepared = false;
  invalidate_graph(c);
  return RPA_OK;
}

int rpa_set_q(rpa_ctx *c, int ntot, const int *K1, const int *K2, const int *iG0, const double *ktot,
              const double *q) {
  return set_q_impl(c, ntot, K1, K2, iG0, nullptr, nullptr, ktot, q);
}

int rpa_set_q_sym(rpa_ctx *c, int ntot, const int *K1, const int *K2, const int *iG0, const int *R1,
                  const int *R2, const double *ktot, const double *q) {
  if (!R1 || !R2) return fail(c, RPA_E_ARG, "rpa_set_q_sym: R1 and R2 are required");
  return set_q_impl(c, ntot, K1, K2, iG0, R1, R2, ktot, q);
}

int rpa_vertices(rpa_ctx *c) {
  if (!c) return RPA_E_ARG;
  (void)hipSetDevice(c->d.device);
  if (int e = prepare(c)) return e;
  if (int e = launch_transitions(c)) return e;
  if (int e = launch_vertices(c, c->d_rho1, c->d_rho2, c->d_perm)) return e;
  return sync(c);
}

int rpa_accumulate_s0(rpa_ctx *c, double *S0_out, double *qeff_out) {
  if (!c) return RPA_E_ARG;
  if (!c->d_rho1 || c->ntrans == 0) return fail(c, RPA_E_STATE, "rpa_vertices must precede rpa_accumulate_s0");
  (void)hipSetDevice(c->d.device);
  if (int e = launch_accumulate(c)) return e;
  if (int e = sync(c)) return e;
  if (S0_out)
    if (int e = get_s0(c, S0_out)) return e;
  if (qeff_out && c->d.program == RPA_OPTABS)
    if (int e = get_square(c, (const double *)c->d_qeff, 2, qeff_out)) return e;
  return RPA_OK;
}

int rpa_kk(rpa_ctx *c, double *chi0_out) {
  if (!c) return RPA_E_ARG;
  if (!c->d_s0) return fail(c, RPA_E_STATE, "rpa_accumulate_s0 must precede rpa_kk");
  (void)hipSetDevice(c->d.device);
  if (int e = launch_kk(c)) return e;
  if (int e = sync(c)) return e;
  if (chi0_out) return copy_freq_mats(c, c->d_chi0, chi0_out);
  return RPA_OK;
}

int rpa_kk_noalloc(rpa_ctx *c, double *chi0_dev) {
  if (!c || !chi0_dev) return fail(c, RPA_E_ARG, "rpa_kk_noalloc: null argument");
  if (!c->d_s0) return fail(c, RPA_E_STATE, "rpa_accumulate_s0 (or rpa_run, rpa_load_s0) must precede rpa_kk_noalloc");
  (void)hipSetDevice(c->d.device);
  return launch_kk_noalloc(c, chi0_dev);
}

int rpa_coulomb(rpa_ctx *c, double *V_out) {
  if (!c) return RPA_E_ARG;
  if (!c->d_V) return fail(c, RPA_E_STATE, "rpa_vertices must precede rpa_coulomb");
  (void)hipSetDevice(c->d.device);
  if (int e = launch_coulomb(c)) return e;
  if (V_out) return get_square(c, (const double *)c->d_V, 1, V_out);
  return sync(c);
}

int rpa_dyson_w(rpa_ctx *c, double *epsinv_out, double *chi_out, double *W_out) {
  if (!c) return RPA_E_ARG;
  if (c->d.program != RPA_SLOSS) return fail(c, RPA_E_STATE, "the Dyson step belongs to the Sloss path");
  if (!c->d_chi0) return fail(c, RPA_E_STATE, "rpa_kk must precede rpa_dyson_w");
  (void)hipSetDevice(c->d.device);
  if (int e = launch_coulomb(c)) return e;
  if (int e = launch_dyson(c)) return e;
  if (int e = queue_status(c)) return e;
  if (int e = sync(c)) return e;
  if (epsinv_out)
    if (int e = copy_freq_mats(c, c->d_epsinv, epsinv_out)) return e;
  if (chi_out)
    if (int e = copy_freq_mats(c, c->d_chi, chi_out)) return e;
  if (W_out)
    if (int e = copy_freq_mats(c, c->d_W, W_out)) return e;
  return check_status(c);
}

int rpa_epilogue(rpa_ctx *c, double *loss, double *pi11, double *sig) {
  if (!c) return RPA_E_ARG;
  if (!c->d_chi0) return fail(c, RPA_E_STATE, "rpa_kk must precede rpa_epilogue");
  (void)hipSetDevice(c->d.device);
  const size_t nf = (size_t)(c->d.no - 1);
  if (c->d.program == RPA_OPTABS) {
    if (int e = launch_epilogue(c)) return e;
    if (pi11)
      if (int e = d2h(c, pi11, (const double *)c->d_pi11, 2 * nf)) return e;
    if (sig)
      if (int e = get_sig(c, sig)) return e;
  } else if (loss) {
    if (int e = d2h(c, loss, (const double *)c->d_loss, nf)) return e;
  }
  return sync(c);
}

int rpa_run(rpa_ctx *c) {
  if (!c) return RPA_E_ARG;
  (void)hipSetDevice(c->d.device);
  if (int e = prepare(c)) return e;
  RPA_HIP(c, hipEventRecord(c->ev0, c->stream));
  if (int e = chain(c)) return e;
  RPA_HIP(c, hipEventRecord(c->ev1, c->stream));
  if (int e = queue_status(c)) return e;
  if (int e = sync(c)) return e;
  float ms = 0.f;
  RPA_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  if (c->profiling)
    if (int e = collect_stage_times(c)) return e;
  return check_status(c);
}

int rpa_load_s0(rpa_ctx *c, const double *S0, const double *qeff) {
  if (!c || !S0) return RPA_E_ARG;
  (void)hipSetDevice(c->d.device);
  if (int e = prepare(c)) return e;
  const int nw = c->nw, N = c->nlf;
  std::vector<double> h((size_t)nw * c->ncol_pad, 0.0);
  for (int w = 0; w < nw; ++w)
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        const size_t pr = (size_t)i * N + j;
        const size_t src = (size_t)w + (size_t)nw * (i + (size_t)N * j);
        if (c->d.program == RPA_SLOSS) {
          h[(size_t)w * c->ncol_pad + pr] = S0[src];
        } else {
          h[(size_t)w * c->ncol_pad + 2 * pr] = S0[2 * src];
          h[(size_t)w * c->ncol_pad + 2 * pr + 1] = S0[2 * src + 1];
        }
      }
  RPA_HIP(c, hipMemcpyAsync(c->d_s0, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
  if (c->d.program == RPA_OPTABS) {
    std::vector<double> q(2 * (size_t)N * N, 0.0);
    if (qeff)
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
          for (int r = 0; r < 2; ++r) q[2 * ((size_t)i * N + j) + r] = qeff[2 * (i + (size_t)N * j) + r];
    RPA_HIP(c, hipMemcpyAsync(c->d_qeff, q.data(), sizeof(double) * q.size(), hipMemcpyHostToDevice, c->stream));
  }
  if (int e = sync(c)) return e;
  c->prepared = false;  // the next full run starts from a cleared Qeff again
  return RPA_OK;
}

int rpa_get_result(rpa_ctx *c, int which, double *out) {
  if (!c || !out) return RPA_E_ARG;
  if (!c->d_s0) return fail(c, RPA_E_STATE, "nothing computed yet");
  (void)hipSetDevice(c->d.device);
  if (int e = sync(c)) return e;
  const size_t nf = (size_t)(c->d.no - 1);
  const bool sloss = c->d.program == RPA_SLOSS;
  switch (which) {
    case RPA_RES_S0: return get_s0(c, out);
    case RPA_RES_QEFF:
      if (sloss) break;
      return get_square(c, (const double *)c->d_qeff, 2, out);
    case RPA_RES_CHI0: return copy_freq_mats(c, c->d_chi0, out);
    case RPA_RES_V:
      if (!sloss) break;
      return get_square(c, (const double *)c->d_V, 1, out);
    case RPA_RES_EPSINV:
      if (!sloss) break;
      return copy_freq_mats(c, c->d_epsinv, out);
    case RPA_RES_CHI:
      if (!sloss) break;
      return copy_freq_mats(c, c->d_chi, out);
    case RPA_RES_W:
      if (!sloss) break;
      return copy_freq_mats(c, c->d_W, out);
    case RPA_RES_LOSS:
      if (!sloss) break;
      return d2h(c, out, (const double *)c->d_loss, nf);
    case RPA_RES_PI11:
      if (sloss) break;
      return d2h(c, out, (const double *)c->d_pi11, 2 * nf);
    case RPA_RES_SIG:
      if (sloss) break;
      return get_sig(c, out);
    default: return fail(c, RPA_E_ARG, "unknown result selector");
  }
  return fail(c, RPA_E_STATE, "result does not belong to this program");
}

int rpa_w_kk(rpa_ctx *c, double *w1, double *w2) {
  if (!c || !w1 || !w2) return RPA_E_ARG;
  if (c->d.program != RPA_SLOSS) return fail(c, RPA_E_STATE, "W belongs to the Sloss path");
  if (!c->d_W) return fail(c, RPA_E_STATE, "rpa_run or rpa_dyson_w must precede rpa_w_kk");
  (void)hipSetDevice(c->d.device);
  if (int e = launch_w_kk(c)) return e;
  const size_t nf = (size_t)(c->d.no - 1);
  if (int e = d2h(c, w1, (const double *)c->d_wkk, nf)) return e;
  return d2h(c, w2, (const double *)c->d_wkk + nf, nf);
}

int rpa_gamma_w(rpa_ctx *c, double *gammap, double *gammam) {
  if (!c || !gammap || !gammam) return fail(c, RPA_E_ARG, "rpa_gamma_w: null argument");
  if (c->d.program != RPA_SLOSS) return fail(c, RPA_E_STATE, "W belongs to the Sloss path");
  if (!c->d_W) return fail(c, RPA_E_STATE, "rpa_run or rpa_dyson_w must precede rpa_gamma_w");
  (void)hipSetDevice(c->d.device);
  const int no = c->d.no, np = c->nlf * c->nlf;
  const size_t nG = (size_t)(no - 1) * np;
  double *ws = nullptr, *S = nullptr, *gp = nullptr, *gm = nullptr;
  auto release = [&]() {
    (void)hipStreamSynchronize(c->stream);
    for (double *p : {ws, S, gp, gm})
      if (p) (void)hipFree(p);
  };
  hipError_t e;
  if ((e = hipMalloc((void **)&ws, sizeof(double) * gamma_ws_doubles(no, np))) != hipSuccess ||
      (e = hipMalloc((void **)&S, sizeof(double) * (size_t)no * np)) != hipSuccess ||
      (e = hipMalloc((void **)&gp, sizeof(double) * 2 * nG)) != hipSuccess ||
      (e = hipMalloc((void **)&gm, sizeof(double) * 2 * nG)) != hipSuccess) {
    release();
    return hip_fail(c, e, "hipMalloc");
  }
  int r = launch_gamma_w(c, ws, S, gp, gm);
  if (!r) r = copy_freq_mats(c, gp, gammap);
  if (!r) r = copy_freq_mats(c, gm, gammam);
  release();
  return r;
}

int rpa_set_profiling(rpa_ctx *c, int on) {
  if (!c) return RPA_E_ARG;
  (void)hipSetDevice(c->d.device);
  if (on && !c->prof_ev[0])
    for (int i = 0; i < 8; ++i) RPA_HIP(c, hipEventCreate(&c->prof_ev[i]));
  c->profiling = on != 0;
  invalidate_graph(c);  // a captured graph would hold the event nodes
  invalidate_batch(c);
  return RPA_OK;
}

int rpa_stage_ms(const rpa_ctx *c, double *ms, int n) {
  if (!c || !ms || n < 0) return RPA_E_ARG;
  for (int i = 0; i < n && i < 6; ++i) ms[i] = c->stage_ms[i];
  return RPA_OK;
}

int rpa_run_graph(rpa_ctx *c) {
  if (!c) return RPA_E_ARG;
  (void)hipSetDevice(c->d.device);
  if (int e = prepare(c)) return e;
  if (!c->graph_exec) {
    hipGraph_t g = nullptr;
    RPA_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int e = chain(c);
    hipError_t he = hipStreamEndCapture(c->stream, &g);
    if (e) {
      if (g) (void)hipGraphDestroy(g);
      return e;
    }
    if (he != hipSuccess) return hip_fail(c, he, "hipStreamEndCapture");
    he = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (he != hipSuccess) return hip_fail(c, he, "hipGraphInstantiate");
    c->graph_valid = true;
  }
  RPA_HIP(c, hipEventRecord(c->ev0, c->stream));
  RPA_HIP(c, hipGraphLaunch(c->graph_exec, c->stream));
  RPA_HIP(c, hipEventRecord(c->ev1, c->stream));
  if (int e = queue_status(c)) return e;
  if (int e = sync(c)) return e;
  float ms = 0.f;
  RPA_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  return check_status(c);
}

int rpa_set_autotune(rpa_ctx *c, int on) {
  if (!c) return RPA_E_ARG;
  c->autotune = on != 0;
  c->prepared = false;
  invalidate_graph(c);
  invalidate_batch(c);
  return RPA_OK;
}

static int batch_set_impl(rpa_ctx *c, int nq, const int *nlf, const int *lf_idx, const int *parg, int ntot,
                          const int *K1, const int *K2, const int *iG0, const int *R1, const int *R2,
                          const double *ktot, const double *q) {
  if (!c || nq <= 0 || !nlf || !lf_idx || !parg || ntot <= 0 || !K1 || !K2 || !iG0 || !ktot || !q)
    return fail(c, RPA_E_ARG, "rpa_batch_set: bad arguments");
  (void)hipSetDevice(c->d.device);
  free_slots(c);
  c->slots = new std::vector<rpa_qslot>(nq);
  if (hipHostMalloc((void **)&c->h_bstatus, sizeof(int) * nq, hipHostMallocDefault) != hipSuccess) {
    free_slots(c);
    return fail(c, RPA_E_HIP, "rpa_batch_set: pinned status buffer");
  }
  size_t off = 0;
  for (int iq = 0; iq < nq; ++iq) {
    rpa_qslot &sl = (*c->slots)[iq];
    swap_qslot(c, sl);
    int e = rpa_set_local_fields(c, nlf[iq], lf_idx + off, parg + off);
    const size_t o = (size_t)iq * ntot;
    if (!e)
      e = set_q_impl(c, ntot, K1 + o, K2 + o, iG0 + o, R1 ? R1 + o : nullptr, R2 ? R2 + o : nullptr, ktot,
                     q + 3 * (size_t)iq);
    swap_qslot(c, sl);
    if (e) {
      const std::string m = c->err;
      free_slots(c);
      return fail(c, e, "rpa_batch_set: q " + std::to_string(iq + 1) + ": " + m);
    }
    off += (size_t)nlf[iq];
  }
  return RPA_OK;
}

int rpa_batch_set(rpa_ctx *c, int nq, const int *nlf, const int *lf_idx, const int *parg, int ntot, const int *K1,
                  const int *K2, const int *iG0, const double *ktot, const double *q) {
  return batch_set_impl(c, nq, nlf, lf_idx, parg, ntot, K1, K2, iG0, nullptr, nullptr, ktot, q);
}

int rpa_batch_set_sym(rpa_ctx *c, int nq, const int *nlf, const int *lf_idx, const int *parg, int ntot,
                      const int *K1, const int *K2, const int *iG0, const int *R1, const int *R2,
                      const double *ktot, const double *q) {
  if (!R1 || !R2) return fail(c, RPA_E_ARG, "rpa_batch_set_sym: R1 and R2 are required");
  return batch_set_impl(c, nq, nlf, lf_idx, parg, ntot, K1, K2, iG0, R1, R2, ktot, q);
}

// Prepares every q of the batch (the autotuner picks each shape's vertex
// form), then captures the chains of all q, one after the other on the
// context's stream, in one graph.  Shared work buffers are sized for the
// largest q before the capture starts, so no launch in the graph sees a
// buffer move.
static int batch_prepare(rpa_ctx *c) {
  std::vector<rpa_qslot> &sl = *c->slots;
  const bool tune = c->autotune;
  for (size_t iq = 0; iq < sl.size(); ++iq) {
    swap_qslot(c, sl[iq]);
    c->autotune = true;
    int e = prepare(c);
    c->autotune = tune;
    swap_qslot(c, sl[iq]);
    if (e) return fail(c, e, "q " + std::to_string(iq + 1) + " of the batch: " + c->err);
  }
  return 0;
}

// RPA_BATCH_GRAPH=0 runs the prepared q eagerly, one chain after the other,
// instead of replaying the captured graph (same kernels, same results).
static bool batch_graph() {
  const char *env = std::getenv("RPA_BATCH_GRAPH");
  return !(env && env[0] == '0');
}

static int batch_capture(rpa_ctx *c) {
  std::vector<rpa_qslot> &sl = *c->slots;
  if (int e = batch_prepare(c)) return e;
  invalidate_graph(c);
  hipGraph_t g = nullptr;
  RPA_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int e = 0;
  for (size_t iq = 0; iq < sl.size() && !e; ++iq) {
    swap_qslot(c, sl[iq]);
    e = chain(c);
    swap_qslot(c, sl[iq]);
  }
  hipError_t he = hipStreamEndCapture(c->stream, &g);
  if (e) {
    if (g) (void)hipGraphDestroy(g);
    return e;
  }
  if (he != hipSuccess) return hip_fail(c, he, "hipStreamEndCapture");
  he = hipGraphInstantiate(&c->batch_exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (he != hipSuccess) return hip_fail(c, he, "hipGraphInstantiate");
  return 0;
}

int rpa_batch_run(rpa_ctx *c) {
  if (!c) return RPA_E_ARG;
  if (!c->slots) return fail(c, RPA_E_STATE, "rpa_batch_set must precede rpa_batch_run");
  if (c->profiling) return fail(c, RPA_E_STATE, "rpa_batch_run runs without stage profiling");
  (void)hipSetDevice(c->d.device);
  std::vector<rpa_qslot> &sl = *c->slots;
  const bool graph = batch_graph();
  if (graph && !c->batch_exec)
    if (int e = batch_capture(c)) return e;
  if (!graph)
    for (rpa_qslot &q : sl)
      if (!q.prepared)
        if (int e = batch_prepare(c)) return e;
  RPA_HIP(c, hipEventRecord(c->ev0, c->stream));
  if (graph) {
    RPA_HIP(c, hipGraphLaunch(c->batch_exec, c->stream));
  } else {
    for (size_t iq = 0; iq < sl.size(); ++iq) {
      swap_qslot(c, sl[iq]);
      const int e = chain(c);
      swap_qslot(c, sl[iq]);
      if (e) return e;
    }
  }
  RPA_HIP(c, hipEventRecord(c->ev1, c->stream));
  if (c->d.program == RPA_SLOSS)
    for (size_t iq = 0; iq < sl.size(); ++iq)
      RPA_HIP(c, hipMemcpyAsync(c->h_bstatus + iq, sl[iq].d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if (int e = sync(c)) return e;
  float ms = 0.f;
  RPA_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  if (c->d.program == RPA_SLOSS)
    for (size_t iq = 0; iq < sl.size(); ++iq)
      if (c->h_bstatus[iq])
        return fail(c, RPA_E_SINGULAR,
                    "singular matrix in the Dyson step of q " + std::to_string(iq + 1) + " of the batch");
  return RPA_OK;
}

int rpa_batch_result(rpa_ctx *c, int iq, int which, double *out) {
  if (!c || !out) return RPA_E_ARG;
  if (!c->slots) return fail(c, RPA_E_STATE, "rpa_batch_set and rpa_batch_run must precede rpa_batch_result");
  if (iq < 0 || iq >= (int)c->slots->size()) return fail(c, RPA_E_ARG, "rpa_batch_result: q index outside the batch");
  const bool sloss = c->d.program == RPA_SLOSS;
  const bool ok = sloss ? (which == RPA_RES_V || which == RPA_RES_EPSINV || which == RPA_RES_W || which == RPA_RES_LOSS)
                        : (which == RPA_RES_QEFF || which == RPA_RES_PI11 || which == RPA_RES_SIG);
  if (!ok) return fail(c, RPA_E_ARG, "rpa_batch_result: not a per-q result of this program");
  (void)hipSetDevice(c->d.device);
  rpa_qslot &sl = (*c->slots)[iq];
  swap_qslot(c, sl);
  const size_t nf = (size_t)(c->d.no - 1);
  int e = 0;
  switch (which) {
    case RPA_RES_V: e = get_square(c, (const double *)c->d_V, 1, out); break;
    case RPA_RES_EPSINV: e = copy_freq_mats(c, c->d_epsinv, out); break;
    case RPA_RES_W: e = copy_freq_mats(c, c->d_W, out); break;
    case RPA_RES_LOSS: e = d2h(c, out, (const double *)c->d_loss, nf); break;
    case RPA_RES_QEFF: e = get_square(c, (const double *)c->d_qeff, 2, out); break;
    case RPA_RES_PI11: e = d2h(c, out, (const double *)c->d_pi11, 2 * nf); break;
    default: e = get_sig(c, out); break;
  }
  swap_qslot(c, sl);
  return e;
}

int rpa_get_vertices(rpa_ctx *c, double *rho1, double *rho2, long long *ntrans_out) {
  if (!c) return RPA_E_ARG;
  (void)hipSetDevice(c->d.device);
  if (int e = prepare(c)) return e;
  const size_t n = (size_t)c->ntrans * c->nlf;
  if (ntrans_out) *ntrans_out = c->ntrans;
  if (!rho1) return RPA_OK;
  double2 *t1 = nullptr, *t2 = nullptr;
  RPA_HIP(c, hipMalloc((void **)&t1, sizeof(double2) * std::max<size_t>(n, 1)));
  hipError_t he = hipSuccess;
  if (rho2) he = hipMalloc((void **)&t2, sizeof(double2) * std::max<size_t>(n, 1));
  int e = (he == hipSuccess) ? launch_vertices(c, t1, t2, nullptr) : hip_fail(c, he, "hipMalloc");
  if (c->d.precision == RPA_MIXED) {  // complex FP32 results, widened here
    std::vector<float> h(2 * n);
    for (int s = 0; s < 2 && !e; ++s) {
      double *dst = s ? rho2 : rho1;
      if (!dst) continue;
      e = d2h(c, h.data(), (const float *)(s ? t2 : t1), 2 * n);
      for (size_t x = 0; x < 2 * n && !e; ++x) dst[x] = h[x];
    }
  } else {
    if (!e) e = d2h(c, rho1, (const double *)t1, 2 * n);
    if (!e && rho2) e = d2h(c, rho2, (const double *)t2, 2 * n);
  }
  (void)hipFree(t1);
  if (t2) (void)hipFree(t2);
  return e;
}

}  // extern "C"
