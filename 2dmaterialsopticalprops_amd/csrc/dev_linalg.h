// Complex FP64 helpers, a workgroup Gauss-Jordan inversion and a Gaussian
// solve shared by the per-point solvers of the downstream programs
// (k_photon.hip, k_ladder.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace rpa {

__device__ inline double2 c_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 c_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 c_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 c_scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline double2 c_conj(double2 a) { return make_double2(a.x, -a.y); }
// a / b by Smith's method
__device__ inline double2 c_div(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.x * r + b.y;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
// principal square root (branch cut on the negative real axis)
__device__ inline double2 c_sqrt(double2 z) {
  if (z.x == 0.0 && z.y == 0.0) return make_double2(0.0, z.y);
  const double t = sqrt(0.5 * (fabs(z.x) + hypot(z.x, z.y)));
  if (z.x >= 0.0) return make_double2(t, z.y / (2.0 * t));
  return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}
__device__ inline double2 c_expi(double x) {  // exp(i x)
  double s, c;
  sincos(x, &s, &c);
  return make_double2(c, s);
}
__device__ inline double2 c_exp(double2 z) {
  const double m = exp(z.x);
  double s, c;
  sincos(z.y, &s, &c);
  return make_double2(m * c, m * s);
}
__device__ inline double2 c_sin(double2 z) {  // sin(x + iy) = sin x cosh y + i cos x sinh y
  double s, c;
  sincos(z.x, &s, &c);
  return make_double2(s * cosh(z.y), c * sinh(z.y));
}


// (a, b) = (x / d, x % d) for x = x0, x0 + 256, ...: one division at the
// start, then a step of (256 / d, 256 % d) with a carry.  An integer division
// by a run-time value costs tens of instructions, several times the complex
// multiply-add of the element it indexes.
struct Walk256 {
  int a, b, qa, qb, d;
  __device__ Walk256(int x0, int d_) : d(d_) {
    a = x0 / d;
    b = x0 - a * d;
    qa = 256 / d;
    qb = 256 - qa * d;
  }
  __device__ void next() {
    a += qa;
    b += qb;
    if (b >= d) {
      b -= d;
      ++a;
    }
  }
};

// In-place Gauss-Jordan inversion of the n x n matrix sa (row-major) with
// row pivoting on |Re| + |Im| (lowest row on ties); a zero pivot sets the
// status flag and goes on without a swap.
template <bool GLOBAL>
__device__ void gj_invert(double2 *sa, int n, double2 *fcol, int *piv, double *red_v, int *red_i, int *status) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = n;
    for (int i = k + tid; i < n; i += 256) {
      const double2 z = sa[i * n + k];
      const double m = fabs(z.x) + fabs(z.y);
      if (m > best) {
        best = m;
        bi = i;
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
      red_v[wave] = best;
      red_i[wave] = bi;
    }
    __syncthreads();
    double pv = red_v[0];
    int p = red_i[0];
    for (int w = 1; w < 4; ++w)
      if (red_v[w] > pv || (red_v[w] == pv && red_i[w] < p)) {
        pv = red_v[w];
        p = red_i[w];
      }
    if (!(pv > 0.0)) {
      if (tid == 0) atomicOr(status, 1);
      p = k;
    }
    if (tid == 0) piv[k] = p;
    if (p != k)
      for (int j = tid; j < n; j += 256) {
        const double2 t = sa[k * n + j];
        sa[k * n + j] = sa[p * n + j];
        sa[p * n + j] = t;
      }
    __syncthreads();
    const double2 inv = c_div(make_double2(1.0, 0.0), sa[k * n + k]);
    __syncthreads();
    for (int j = tid; j < n; j += 256) sa[k * n + j] = (j == k) ? inv : c_mul(sa[k * n + j], inv);
    for (int i = tid; i < n; i += 256) fcol[i] = sa[i * n + k];
    __syncthreads();
    for (int i = tid; i < n; i += 256)
      if (i != k) sa[i * n + k] = make_double2(0.0, 0.0);
    __syncthreads();
    Walk256 wk(tid, n);  // (row, column) of x
    for (int x = tid; x < n * n; x += 256, wk.next())
      if (wk.a != k) sa[x] = c_sub(sa[x], c_mul(fcol[wk.a], sa[k * n + wk.b]));
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {  // row interchanges undone as column interchanges
    const int p = piv[k];
    if (p != k)
      for (int i = tid; i < n; i += 256) {
        const double2 t = sa[i * n + k];
        sa[i * n + k] = sa[i * n + p];
        sa[i * n + p] = t;
      }
    __syncthreads();
  }
}

// Solves A z = y in place for nr right-hand sides y[nr][n] (row r at y + r n)
// by Gaussian elimination with row pivoting on |Re| + |Im| (lowest row on
// ties), A n x n row-major in sa: per step the rows below the pivot lose
// their column-k entries (the multipliers are not kept) and the right-hand
// sides follow, then back substitution column by column; z overwrites y.
// A zero pivot sets the status flag and goes on without a swap.  n^3/3
// multiply-adds and traffic for A where its inverse takes n^3.  Indices are
// 32-bit (n <= 128 in every caller; a slice's offset is in sa already), so
// the LDS addresses need no 64-bit arithmetic.
template <bool GLOBAL>
__device__ void gauss_solve(double2 *sa, int n, double2 *y, int nr, double2 *fcol, double *red_v, int *red_i,
                            int *status) {
  __shared__ double2 s_inv;  // 1 / pivot, formed once per step
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = n;
    for (int i = k + tid; i < n; i += 256) {
      const double2 z = sa[i * n + k];
      const double m = fabs(z.x) + fabs(z.y);
      if (m > best) {
        best = m;
        bi = i;
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
      red_v[wave] = best;
      red_i[wave] = bi;
    }
    __syncthreads();
    double pv = red_v[0];
    int p = red_i[0];
    for (int w = 1; w < 4; ++w)
      if (red_v[w] > pv || (red_v[w] == pv && red_i[w] < p)) {
        pv = red_v[w];
        p = red_i[w];
      }
    if (!(pv > 0.0)) {
      if (tid == 0) atomicOr(status, 1);
      p = k;
    }
    // thread 0 swaps column k itself, so it reads the pivot before any write
    if (tid == 0) s_inv = c_div(make_double2(1.0, 0.0), sa[p * n + k]);
    if (p != k) {
      for (int j = k + tid; j < n; j += 256) {
        const double2 t = sa[k * n + j];
        sa[k * n + j] = sa[p * n + j];
        sa[p * n + j] = t;
      }
      for (int r = tid; r < nr; r += 256) {
        const double2 t = y[r * n + k];
        y[r * n + k] = y[r * n + p];
        y[r * n + p] = t;
      }
    }
    __syncthreads();
    const double2 inv = s_inv;
    for (int i = k + 1 + tid; i < n; i += 256) fcol[i] = c_mul(sa[i * n + k], inv);
    __syncthreads();
    const int m = n - 1 - k;  // the trailing block is m x m
    if (m > 0) {
      Walk256 wk(tid, m);  // (row, column) of the block
      for (int x = tid; x < m * m; x += 256, wk.next()) {
        const int i = k + 1 + wk.a, j = k + 1 + wk.b;
        sa[i * n + j] = c_sub(sa[i * n + j], c_mul(fcol[i], sa[k * n + j]));
      }
      Walk256 wr(tid, m);  // (right-hand side, row)
      for (int x = tid; x < nr * m; x += 256, wr.next()) {
        const int r = wr.a, i = k + 1 + wr.b;
        y[r * n + i] = c_sub(y[r * n + i], c_mul(fcol[i], y[r * n + k]));
      }
    }
    __syncthreads();
  }
  for (int i = n - 1; i >= 0; --i) {
    for (int r = tid; r < nr; r += 256) y[r * n + i] = c_div(y[r * n + i], sa[i * n + i]);
    __syncthreads();
    if (i > 0) {
      Walk256 wb(tid, i);  // (right-hand side, row)
      for (int x = tid; x < nr * i; x += 256, wb.next()) {
        const int r = wb.a, j = wb.b;
        y[r * n + j] = c_sub(y[r * n + j], c_mul(sa[j * n + i], y[r * n + i]));
      }
    }
    __syncthreads();
  }
}

// NV doubles summed over the workgroup into red[4 NV ..] (thread 0 reads them)
template <int NV>
__device__ inline void block_sum(const double (&v)[NV], double *red) {
  constexpr int nv = NV;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int e = 0; e < nv; ++e) {
    double s = v[e];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) red[wave * nv + e] = s;
  }
  __syncthreads();
  if (tid < nv) red[4 * nv + tid] = red[tid] + red[nv + tid] + red[2 * nv + tid] + red[3 * nv + tid];
  __syncthreads();
}

}  // namespace rpa
