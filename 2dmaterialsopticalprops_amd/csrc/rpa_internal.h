// Internal declarations of librpa_hip.so (MI355X / gfx950).
//
// Device data layout (HBM):
//   basis    evc[nki][npwx][nband]  complex  (G-major, bands fastest: one
//            plane wave of all bands is one contiguous 16*nband-byte row, so
//            the gathered vertex operand C2[map(i)][m] is a coalesced row read)
//   mill     mill_k[nki][npwx]      packed Miller key (10 bits per component)
//   kbox     kbox[nki][box]         int16 position of each Miller triple in
//            that k's coefficient list (-1 = absent); gbox[box] int32 position
//            in the global G list (direct map)
//   vertices rho[pos][nlf]          complex, pos = transition position sorted
//            by the first frequency tile its Lorentzian window reaches
//   S0       s0[nw][ncol]           real (Sloss) or interleaved complex
//            (Optabs) columns col = iG*NGP + jG (x2 for complex)
//   chi0     chi0[no-1][nlf][nlf]   complex (row-major iG, jG)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rpa_hip.h"

namespace rpa {

constexpr int kMillBias = 512;
__host__ __device__ inline int mill_pack(int a, int b, int c) {
  return ((a + kMillBias) << 20) | ((b + kMillBias) << 10) | (c + kMillBias);
}
__host__ __device__ inline void mill_unpack(int key, int &a, int &b, int &c) {
  a = ((key >> 20) & 1023) - kMillBias;
  b = ((key >> 10) & 1023) - kMillBias;
  c = (key & 1023) - kMillBias;
}

struct Box {
  int lo[3];
  int n[3];
  __host__ __device__ int size() const { return n[0] * n[1] * n[2]; }
  // flattened index of a Miller triple, -1 if outside
  __host__ __device__ int index(int a, int b, int c) const {
    a -= lo[0];
    b -= lo[1];
    c -= lo[2];
    if ((unsigned)a >= (unsigned)n[0] || (unsigned)b >= (unsigned)n[1] || (unsigned)c >= (unsigned)n[2])
      return -1;
    return (a * n[1] + b) * n[2] + c;
  }
};

// A rectangular block of band pairs (n in [n_lo,n_hi), m in [m_lo,m_hi)).
struct PairBlock {
  int n_lo, n_hi, m_lo, m_hi;
};

constexpr int kMaxPairBlocks = 4;
constexpr int kTileW = 32;     // frequency rows per accumulation tile
constexpr int kColsWG = 256;   // S0 columns per accumulation work item
constexpr int kChunkT = 4096;    // transitions per accumulation work item (FP64)
constexpr int kChunkTMx = 1024;  // the same for the mixed path (FP32 partial sums)
constexpr int kMaxNlf = 128;   // local fields supported by the kernels
constexpr int kChunkHp = 1024;  // mixed S0 in home panels: transitions per item
constexpr int kHpKS = 32;       // the same: transitions per stage (K of one MFMA)
constexpr int kHpMaxNrb = 8;    // the same: panel row blocks of 16 compiled in

// Column blocks of 16 of the home-panel S0 form: G fixed, G' = G'0 .. G'0 +
// 15; the symmetric real S0 takes G'0 = G + 16 j (G-major), the complex S0
// G'0 = 16 j.  Count of all blocks:
inline int hp_nblocks(int nlf, bool cplx) {
  if (cplx) return nlf * ((nlf + 15) / 16);
  int n = 0;
  for (int g = 0; g < nlf; ++g) n += (nlf - g + 15) / 16;
  return n;
}

struct PairTable {
  int nblocks;
  int tpk;  // transitions per k
  int off[kMaxPairBlocks];
  PairBlock b[kMaxPairBlocks];
};

// One vertex work block: up to 8 bands n by up to 16 bands m of a pair block.
constexpr int kVertexMaxBlocks = 8;
struct VBlock {
  int n0, nrows;  // 8 bands n per block
  int m0, ncols;  // 16 bands m per block
  int pb;         // pair block index
};

// Column-correlation plan of the charge vertices in the direct map (built
// per q on the host, see k_vertex_col.hip).
constexpr int kColMaxShift = 128;  // shift slots of one in-plane class
struct ColPlan {
  bool ok = false;
  int npos = 0, nchunk = 0, ncls = 0, nrows = 0;  // nrows: B rows of all classes (one G0)
  int kc = 0, tpw = 0, ms = 0, kind = 0, pol1 = 0, pol2 = 0;
  bool pipe = false;
  bool split = false;  // mixed charge vertices: BF16 hi/lo split products
  bool box = false;    // slots and staged rows index the column-box basis
  bool dma = false;    // split form with the k+q rows staged by LDS DMA
  size_t lds = 0;
  int *d_slot = nullptr, *d_prow = nullptr, *d_chunk = nullptr, *d_crow = nullptr;
  int *d_bmap = nullptr, *d_shift = nullptr, *d_cls = nullptr;
  double *d_beta = nullptr;  // current vertices: Gcar G(i)_pol per position, [2][npos]
  size_t cap_slot = 0, cap_prow = 0, cap_chunk = 0, cap_crow = 0, cap_bmap = 0, cap_shift = 0, cap_cls = 0;
  size_t cap_beta = 0;
};

}  // namespace rpa

struct rpa_ctx {
  rpa_dims d{};
  rpa_params p{};
  bool have_params = false;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // pinned staging of the per-q uploads (q point, local fields, column plan)
  unsigned char *h_stage = nullptr;
  size_t cap_stage = 0;
  double last_ms = 0.0;

  // G list
  int ngm = 0;
  std::vector<int> mill_g;  // 3*ngm
  double kc[9] = {0};
  int *d_mill_g = nullptr;  // packed keys [ngm]
  int *d_gbox = nullptr;    // global box -> index
  rpa::Box gbox{};
  std::vector<int> h_gbox;  // host copy of the global box
  // point operations (rpa_set_symmetry), per operation: the Miller forms of
  // R and of its inverse RI (row-major, m' = M m) and R KC (column-major)
  int nsym = 1;
  std::vector<int> op_m, op_minv;
  std::vector<double> op_rkc;

  // basis
  void *d_evc = nullptr;
  float2 *d_evc32 = nullptr;  // mixed precision: FP32 copy of the basis
  size_t dyson_lds_attr[4] = {0, 0, 0, 0};  // dynamic LDS each Dyson kernel form is allowed
  size_t accum_lds_attr = 0;          // dynamic LDS the FP64 k_accum instance is allowed
  int *d_mill_k = nullptr;
  int16_t *d_kbox = nullptr;
  int *d_ngk = nullptr;
  std::vector<int> ngk;
  std::vector<char> uploaded;
  int mill_min[3] = {1 << 20, 1 << 20, 1 << 20}, mill_max[3] = {-(1 << 20), -(1 << 20), -(1 << 20)};
  rpa::Box kbox{};
  bool kbox_ready = false;

  // bands
  std::vector<double> E;  // E(nki, nband)
  double *d_E = nullptr;

  // local fields
  int nlf = 0;
  int ngp = 0;  // nlf padded to a multiple of 8
  std::vector<int> lf_idx, parg;
  int *d_lf_key = nullptr;
  double *d_glf = nullptr;  // cartesian (2pi/a0) 3*nlf
  int *d_parg = nullptr;
  size_t cap_lf_key = 0, cap_glf = 0, cap_parg = 0;

  // q
  int ntot = 0;
  double q[3] = {0, 0, 0};
  std::vector<int> K1, K2, iG0;
  std::vector<double> ktot;
  int *d_k1 = nullptr, *d_k2 = nullptr, *d_m0 = nullptr;
  double *d_ktot = nullptr;
  size_t cap_k1 = 0, cap_k2 = 0, cap_m0 = 0, cap_ktot = 0, cap_g0keys = 0, cap_g0idx = 0;
  bool have_q = false;
  // map combinations of the q: the distinct (R1, R2, iG0) of its k, in order
  // of appearance (with the identity only, the distinct G0): keys, per-k
  // index, direct map table, and per combination the integer operation of
  // the target, A = M(RI(R2)) M(R(R1)) and B = M(RI(R2)) (target Miller
  // A m + B (Glf + G0)), and R(R1) KC for the current vertices
  int ng0 = 0;
  std::vector<int> g0_list;  // iG0 (1-based) of each combination
  std::vector<int> kR1, kR2;  // per k, 1-based operations (empty: identity)
  int *d_g0keys = nullptr, *d_g0idx = nullptr, *d_dmap = nullptr;
  int *d_csym = nullptr;     // [ng0][18]: A then B, row-major
  double *d_crkc = nullptr;  // [ng0][9]
  size_t cap_csym = 0, cap_crkc = 0;
  bool qsym = false;         // some combination rotates (A or B not the identity)
  rpa::ColPlan col;          // column-correlation vertex plan
  // column-box copy of the basis for the column vertex form: per k the
  // plane waves of global index below the largest ngk, in (mx, my, z) order
  // (a column of the Miller lattice is a contiguous run of rows), zero where
  // the index is past that k's ngk; bands split at Nval into two arrays when
  // every band block lies on one side, so C1 and C2 reads fetch only their
  // own bands
  void *d_box[2] = {nullptr, nullptr};
  size_t cap_box[2] = {0, 0};
  int *d_box_gidx = nullptr;
  size_t cap_box_gidx = 0;
  std::vector<int> h_rbox;  // global Miller box index -> box row (-1: none)
  int box_nbg = 0, box_blo1 = 0, box_rows = 0;
  bool box_ready = false;
  // mixed precision: the k+q side's box array (array 1 when the bands are
  // split, else array 0) as BF16 (hi, lo) word pairs per value, rows of
  // boxw_bww values (even), a zero tail past the last row; read by LDS DMA
  uint32_t *d_boxw = nullptr;
  size_t cap_boxw = 0;
  int boxw_bww = 0;
  long long boxw_nv = 0;  // values, tail included
  bool boxw_ready = false;
  // vertex form of the current q ("" = shape rules; "col", "gather",
  // "staged") and staged chunk (0 = largest that fits), set by the autotuner
  std::string vform;
  int vch = 0;
  bool autotune = false;
  std::vector<std::pair<std::string, std::pair<std::string, int>>> vform_cache;  // shape key -> choice
  size_t cap_dmap = 0;
  double *d_wkk = nullptr;  // W1(no-1), W2(no-1) of the W Kramers-Kronig check
  size_t cap_wkk = 0;

  // transitions
  rpa::PairTable pairs{};
  long long ntrans = 0;
  int nw = 0;       // frequency rows of S0
  int ntiles = 0;   // frequency tiles
  int span = 0;     // extra tiles a Lorentzian window can reach
  double wcut = 0;  // |Lor| threshold
  int nthread_blocks = 0;
  int *d_key = nullptr;      // home tile per natural transition (-1 none)
  int *d_perm = nullptr;     // natural -> sorted position (-1 none)
  int *d_bcnt = nullptr;     // per-block per-tile counts / bases
  int *d_toff = nullptr;     // tile offsets [ntiles+1]
  int *d_nat = nullptr;      // sorted -> natural
  double *d_te = nullptr;    // sorted (E1, E2, df) triples
  int wl_span = 0, wl_chunk = 0;  // worklist: tiles behind a tile's home, transitions per item
  int hp_nrb = 0;            // mixed S0 in home panels of 16 hp_nrb rows (0: the tile form)
  uint16_t *d_wtab = nullptr;  // home panels: BF16 hi/lo weights [item][stage][2][rows][32]
  size_t cap_wtab = 0;
  int *d_items = nullptr;    // accumulation work items
  int *d_nitems = nullptr;   // [0] = item count, [1] = total transitions kept
  int *d_first_item = nullptr;
  int *d_intra = nullptr;    // intraband natural transitions [count]
  int *d_nintra = nullptr;
  size_t max_items = 0;
  size_t cap_trans = 0;

  // vertices (sorted order), complex
  void *d_rho1 = nullptr, *d_rho2 = nullptr;
  size_t cap_rho = 0;

  // S0
  int ncol = 0, ncol_pad = 0;
  double *d_s0 = nullptr;
  double *d_part = nullptr;
  float *d_wmx = nullptr;      // mixed: FP32 weights of every chunk [item][32][kChunkTMx]
  uint32_t *d_rbt = nullptr;   // mixed: BF16 hi/lo of the vertices, [2][nlf][ntrans]
  double *d_wt64 = nullptr;    // FP64: weights of every chunk [item][32][kChunkT]
  size_t cap_wmx = 0, cap_rbt = 0, cap_wt64 = 0;
  double *d_qeff = nullptr;
  size_t cap_part = 0;

  // KK, Dyson
  double *d_kmat = nullptr;  // stencil (no-1) x nw
  int kmat_no = 0;
  double kmat_dw = 0;
  double *d_x = nullptr;     // K * S0 (no-1) x ncol_pad
  double *d_chi0 = nullptr;  // complex (no-1) nlf nlf
  double *d_V = nullptr;     // nlf x nlf
  double *d_epsinv = nullptr, *d_chi = nullptr, *d_W = nullptr;
  double *d_loss = nullptr, *d_pi11 = nullptr, *d_sig = nullptr;
  int *d_status = nullptr;
  int *h_status = nullptr;  // pinned copy of d_status after a run

  // batched q loop (rpa_batch_*): per-q state, exchanged with the fields
  // above while a q is prepared, captured or read; one graph for all q
  std::vector<struct rpa_qslot> *slots = nullptr;
  hipGraphExec_t batch_exec = nullptr;
  int *h_bstatus = nullptr;  // pinned: each q's singular flag after a batch run

  // graph
  hipGraphExec_t graph_exec = nullptr;
  bool graph_valid = false;
  bool prepared = false;
  // stage profiling (rpa_set_profiling)
  bool profiling = false;
  hipEvent_t prof_ev[8] = {nullptr};
  double stage_ms[6] = {0};
  // capacities (elements) of the buffers sized by prepare()
  size_t cap_key = 0, cap_bcnt = 0, cap_toff = 0, cap_items = 0, cap_first = 0, cap_nitems = 0;
  size_t cap_te = 0, cap_s0 = 0, cap_kmat = 0, cap_x = 0, cap_chi0 = 0, cap_epsinv = 0;
  size_t cap_chi = 0, cap_W = 0, cap_V = 0, cap_loss = 0, cap_pi11 = 0, cap_sig = 0;
  size_t cap_qeff = 0, cap_status = 0, cap_rho2 = 0;
};

// The q-dependent fields of a context: local fields, k+q map, vertex plans,
// the shapes derived from them and the per-q results.
#define RPA_Q_FIELDS(X)                                                                                      \
  X(nlf) X(ngp) X(lf_idx) X(parg) X(d_lf_key) X(d_glf) X(d_parg) X(cap_lf_key) X(cap_glf) X(cap_parg) X(ntot)   \
  X(K1) X(K2) X(iG0) X(ktot) X(d_k1) X(d_k2) X(d_m0) X(d_ktot) X(cap_k1) X(cap_k2) X(cap_m0) X(cap_ktot)          \
  X(cap_g0keys) X(cap_g0idx) X(have_q) X(ng0) X(g0_list) X(d_g0keys) X(d_g0idx) X(d_dmap) X(col) X(cap_dmap)       \
  X(kR1) X(kR2) X(d_csym) X(d_crkc) X(cap_csym) X(cap_crkc) X(qsym)                                              \
  X(vform) X(vch) X(ncol) X(ncol_pad) X(prepared) X(d_V) X(d_epsinv) X(d_W) X(d_loss) X(d_pi11) X(d_sig)          \
  X(d_qeff) X(d_status) X(cap_V) X(cap_epsinv) X(cap_W) X(cap_loss) X(cap_pi11) X(cap_sig) X(cap_qeff)           \
  X(cap_status)

struct rpa_qslot {
#define RPA_QSLOT_FIELD(f) decltype(rpa_ctx::f) f{};
  RPA_Q_FIELDS(RPA_QSLOT_FIELD)
#undef RPA_QSLOT_FIELD
  double q[3] = {0.0, 0.0, 0.0};
};

// Exchanges the q-dependent fields of the context with those of a slot.
inline void swap_qslot(rpa_ctx *c, rpa_qslot &s) {
#define RPA_QSLOT_SWAP(f) std::swap(c->f, s.f);
  RPA_Q_FIELDS(RPA_QSLOT_SWAP)
#undef RPA_QSLOT_SWAP
  for (int a = 0; a < 3; ++a) std::swap(c->q[a], s.q[a]);
}

namespace rpa {

int fail(rpa_ctx *c, int code, const std::string &msg);
int prepare(rpa_ctx *c);
int chain(rpa_ctx *c);
void free_q_buffers(rpa_ctx *c);
int autotune_vertex(rpa_ctx *c);
int hip_fail(rpa_ctx *c, hipError_t e, const char *what);

#define RPA_HIP(ctx, call)                                 \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return rpa::hip_fail(ctx, e_, #call); \
  } while (0)

template <class T>
int ensure(rpa_ctx *c, T *&ptr, size_t &cap, size_t n) {
  if (n <= cap && ptr) return 0;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
  hipError_t e = hipMalloc((void **)&ptr, sizeof(T) * (n ? n : 1));
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc");
  cap = n;
  return 0;
}

// One host-to-device upload of a staged batch: each part is copied into the
// context's pinned staging buffer and sent with an asynchronous copy on the
// context's stream, ahead of the kernels that read it.  The buffer is reused
// only after the stream has drained.
struct Upload {
  void *dst;
  const void *src;
  size_t bytes;
};
int upload(rpa_ctx *c, const Upload *parts, int n);

// kernel launchers (defined in the .hip files)
int launch_build_kbox(rpa_ctx *c);
int launch_transitions(rpa_ctx *c);
int launch_vertices(rpa_ctx *c, void *rho1, void *rho2, const int *perm);
int launch_direct_map(rpa_ctx *c);
int build_col_plan(rpa_ctx *c);
int launch_vertices_col(rpa_ctx *c, const VBlock *vb, int nvb, void *rho, void *rho2, const int *perm);
void free_col_plan(rpa_ctx *c);
int launch_w_kk(rpa_ctx *c);
int launch_accumulate(rpa_ctx *c);
int launch_kstencil(rpa_ctx *c);
int launch_kk(rpa_ctx *c);
int launch_coulomb(rpa_ctx *c);
int launch_dyson(rpa_ctx *c);
int configure_dyson(rpa_ctx *c);
int configure_accumulate(rpa_ctx *c);
int configure_vertex(rpa_ctx *c);
int launch_epilogue(rpa_ctx *c);
int noalloc_init_device(int device);
int launch_kk_noalloc(rpa_ctx *c, double *chi0_dev);
// The Optabs Kramers-Kronig transform (stencil, product, epilogue) of ncol
// = 2 nlf^2 interleaved complex S0 columns s0[2no+1][ncol] on a stream, with
// no context: kmat holds (no-1)(2no+1) and x (no-1) ncol doubles, chi0 gets
// [no-1][nlf*nlf] complex.
hipError_t kk_optabs_columns(hipStream_t s, int no, int nlf, int ncol, double dw, const double *s0, double *kmat,
                             double *x, double *chi0);
// GW propagators of the resident W (all pairs): ws holds gamma_ws_doubles,
// S no * nlf^2 doubles, gp / gm [no-1][nlf^2] complex.
size_t gamma_ws_doubles(int no, int nser);
int launch_gamma_w(rpa_ctx *c, double *ws, double *S, double *gp, double *gm);
// The message of the last failed call that takes no context (this thread).
void set_free_error(const std::string &msg);

}  // namespace rpa

// ---- device helpers -------------------------------------------------------
#if defined(__HIPCC__)
typedef double rpa_d4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16) in FP64 on the matrix cores.
// Operand layout (wave64, measured on gfx950): lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; element e of lane l's result is D[4*e + l/16][l%16].
__device__ inline rpa_d4 mfma_f64_16x16x4(double a, double b, rpa_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

typedef float rpa_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 rpa_bf8 __attribute__((ext_vector_type(8)));

// D(16x16) += A(16x4) B(4x16) in FP32.  Lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; element e of lane l's result is D[4*(l/16) + e][l%16]
// (measured on gfx950; note the different row order from the FP64 form).
__device__ inline rpa_f4 mfma_f32_16x16x4(float a, float b, rpa_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// D(16x16) += A(16x32) B(32x16), BF16 inputs, FP32 accumulation.  Lane l
// supplies row l%16 of A and column l%16 of B, 8 values each for the K
// slice 8*(l/16) .. 8*(l/16)+7 (the same slice in A and B); element e of
// lane l's result is D[4*(l/16) + e][l%16] (measured on gfx950).
__device__ inline rpa_f4 mfma_bf16_16x16x32(rpa_bf8 a, rpa_bf8 b, rpa_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
#endif
