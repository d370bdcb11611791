// Batched explicit inverse of complex128 blocks by blocked Gauss-Jordan
// elimination with partial pivoting (the diagonal-block inversion of RGF).
//
// Algorithm (per instance, n x n work matrix X, column panels of NB):
//   panel kernel (one workgroup per instance, the n x NB panel in registers,
//   RPT rows per thread):  NB exchange steps; step j picks the pivot row p_j
//   among rows not yet pivoted by max |re| + |im| (LAPACK's izamax rule, so
//   the pivot sequence is getrf's), scales it, eliminates column j from every
//   other row (Gauss-Jordan: rows above too).  The composite row operation T
//   of the NB steps acts on every other column c as
//       X[:,c] <- X[:,c] + (P - E) X[p_J, c],
//   P = the finished panel, E = unit columns at the pivot rows.  The kernel
//   writes W = P - E, gathers R = X[p_J, :] (identity in the panel columns)
//   and sets X[:,J] = E, so one MFMA GEMM  X <- X + W R  finishes the step.
//   No physical row swaps: after all panels X[p_j, k] = inv(A)[j, p_k], and
//   the final unscramble kernel writes inv(A)[r, c] = X[p_r, q_c] with
//   q = p^{-1}.
// Flops: 8 n^3 (the nominal getrf + getri count).
#include "negf_internal.hpp"
#include <climits>

namespace negf {

namespace {

// PT threads per panel workgroup, RPT rows per thread (n <= PT * RPT).
// One barrier per column: every wave reduces its own pivot candidate with
// shuffles and the candidate's owner publishes the whole unscaled row in a
// per-wave LDS slot; after the barrier each thread picks the winning slot and
// scales on the fly.  Slots are double-buffered by column parity, so a fast
// wave writing column j+1 never overwrites what a slow one still reads for j.
template <int PT, int RPT, int NB>
__global__ __launch_bounds__(PT) void gj_panel_kernel(int n, int k0, zc* __restrict__ X,
                                                      zc* __restrict__ W, zc* __restrict__ R,
                                                      int* __restrict__ piv,
                                                      int* __restrict__ used,
                                                      int* __restrict__ info) {
  constexpr int NW = PT / 64;
  const int inst = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  zc* Xi = X + (long long)inst * n * n;
  zc* Wi = W + (long long)inst * n * NB;
  zc* Ri = R + (long long)inst * NB * n;
  int* usedi = used + (long long)inst * n;
  const int nbe = (n - k0) < NB ? (n - k0) : NB;

  __shared__ double red_v[2][NW];
  __shared__ int red_i[2][NW];
  __shared__ zc cand[2][NW][NB];

  zc P[RPT][NB];
  bool usd[RPT];
  int pc[NB];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + PT * q;
    usd[q] = true;
#pragma unroll
    for (int c = 0; c < NB; ++c) P[q][c] = zmake(0.0, 0.0);
    if (r < n) {
      usd[q] = usedi[r] != 0;
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < nbe) P[q][c] = Xi[r + (long long)(k0 + c) * n];
    }
  }

#pragma unroll
  for (int jj = 0; jj < NB; ++jj) {
    pc[jj] = 0;
    if (jj < nbe) {
      const int buf = jj & 1;
      // pivot search: max |re|+|im| over rows not yet pivoted, lowest index on ties
      double bv = -1.0;
      int bi = INT_MAX;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int r = tid + PT * q;
        if (r < n && !usd[q]) {
          const double v = zabs1(P[q][jj]);
          if (v > bv) { bv = v; bi = r; }
        }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(bv, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
      }
      // the wave's candidate row, published by its owner
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        if (tid + PT * q == bi) {
#pragma unroll
          for (int c = 0; c < NB; ++c) cand[buf][wave][c] = P[q][c];
        }
      }
      if (lane == 0) { red_v[buf][wave] = bv; red_i[buf][wave] = bi; }
      __syncthreads();
      int ws = 0;
      bv = red_v[buf][0]; bi = red_i[buf][0];
#pragma unroll
      for (int w = 1; w < NW; ++w) {
        const double ov = red_v[buf][w];
        const int oi = red_i[buf][w];
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; ws = w; }
      }
      const int p = bi;
      pc[jj] = p;
      if (tid == 0 && bv <= 0.0 && info[inst] == 0) info[inst] = k0 + jj + 1;
      const zc pv = cand[buf][ws][jj];
      const zc inv = (pv.x == 0.0 && pv.y == 0.0) ? zmake(0.0, 0.0) : zrcp(pv);
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int r = tid + PT * q;
        if (r < n) {
          if (r == p) {
#pragma unroll
            for (int c = 0; c < NB; ++c) P[q][c] = (c == jj) ? inv : zmul(P[q][c], inv);
            usd[q] = true;
          } else {
            const zc fi = zmul(P[q][jj], inv);
#pragma unroll
            for (int c = 0; c < NB; ++c)
              if (c != jj) P[q][c] = zsub(P[q][c], zmul(fi, cand[buf][ws][c]));
            P[q][jj] = zmake(-fi.x, -fi.y);
          }
        }
      }
    }
  }

  // W = P - E ; X[:, J] = E ; used flags back
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + PT * q;
    if (r < n) {
      usedi[r] = usd[q] ? 1 : 0;
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c < nbe) {
          const bool e = (r == pc[c]);
          zc w = P[q][c];
          if (e) w.x -= 1.0;
          Wi[r + (long long)c * n] = w;
          Xi[r + (long long)(k0 + c) * n] = zmake(e ? 1.0 : 0.0, 0.0);
        }
      }
    }
  }
  if (tid == 0) {
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c < nbe) piv[(long long)inst * n + k0 + c] = pc[c];
  }
  // R = X[p_J, :] outside the panel, identity inside (ld NB).  Fixed trip
  // count so the gathers are issued back to back (NB * n <= NB * PT * RPT).
  zc g[NB * RPT];
#pragma unroll
  for (int t = 0; t < NB * RPT; ++t) {
    const int idx = tid + PT * t;
    const int c = idx % NB, col = idx / NB;
    g[t] = zmake(0.0, 0.0);
    int prow = 0;
#pragma unroll
    for (int cc = 0; cc < NB; ++cc)
      if (cc == c) prow = pc[cc];
    if (col < n && c < nbe) {
      if (col >= k0 && col < k0 + nbe) g[t] = zmake(col - k0 == c ? 1.0 : 0.0, 0.0);
      else g[t] = Xi[prow + (long long)col * n];
    }
  }
#pragma unroll
  for (int t = 0; t < NB * RPT; ++t) {
    const int idx = tid + PT * t;
    const int c = idx % NB, col = idx / NB;
    if (col < n && c < nbe) Ri[idx] = g[t];
  }
}

__global__ void gj_init_kernel(int n, int* used, int* info, int batch) {
  const long long total = (long long)n * batch;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    used[i] = 0;
    if (i < batch) info[i] = 0;
  }
}

// out[r, c] = X[p_r, q_c], q = p^{-1}; grid.x = column chunks of 64, grid.y = instance
__global__ __launch_bounds__(256) void gj_unscramble_kernel(int n, const zc* __restrict__ X,
                                                            const int* __restrict__ piv,
                                                            zc* __restrict__ out, long long ldo,
                                                            long long strideO) {
  extern __shared__ int sh[];
  int* p = sh;
  int* qinv = sh + n;
  const int inst = blockIdx.y;
  const int* pv = piv + (long long)inst * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) p[j] = pv[j];
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) qinv[p[j]] = j;
  __syncthreads();
  const zc* Xi = X + (long long)inst * n * n;
  zc* Oi = out + inst * strideO;
  const int c0 = blockIdx.x * 64;
  const int c1 = (c0 + 64 < n) ? c0 + 64 : n;
  for (int c = c0; c < c1; ++c) {
    const zc* col = Xi + (long long)qinv[c] * n;
    for (int r = threadIdx.x; r < n; r += blockDim.x) Oi[r + c * ldo] = col[p[r]];
  }
}

template <int PT, int RPT, int NB>
int run(int n, const zc* A, long long lda, long long strideA, zc* Ainv, long long ldinv,
        long long strideInv, int batch, const InvWork& w, int* info, hipStream_t s) {
  if (A != w.X || lda != n || strideA != (long long)n * n) {
    int rc = zcopy2d_launch(n, n, A, lda, strideA, w.X, n, (long long)n * n, batch, s);
    if (rc) return rc;
  }
  {
    long long total = (long long)n * batch;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(gj_init_kernel, dim3(blocks), dim3(256), 0, s, n, w.used, info, batch);
  }
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nbe = (n - k0) < NB ? (n - k0) : NB;
    hipLaunchKernelGGL((gj_panel_kernel<PT, RPT, NB>), dim3(batch), dim3(PT), 0, s, n, k0, w.X, w.W,
                       w.R, w.piv, w.used, info);
    GemmArgs g{};
    g.opA = OP_N; g.opB = OP_N;
    g.m = n; g.n = n; g.k = nbe;
    g.alpha = zmake(1.0, 0.0); g.beta = zmake(1.0, 0.0);
    g.A = w.W; g.lda = n; g.strideA = (long long)n * NB;
    g.B = w.R; g.ldb = NB; g.strideB = (long long)NB * n;
    g.C = w.X; g.ldc = n; g.strideC = (long long)n * n;
    g.D = w.X; g.ldd = n; g.strideD = (long long)n * n;
    g.batch = batch;
    int rc = zgemm_launch(g, s);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(gj_unscramble_kernel, dim3((n + 63) / 64, batch), dim3(256),
                     2 * n * sizeof(int), s, n, w.X, w.piv, Ainv, ldinv, strideInv);
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

// Panel width: the n x NB panel lives in registers (RPT * NB complex per
// thread <= 32), so wider panels need more threads per row block.  The
// workspace is sized for the widest panel of any variant.
int nb_for(int n) { return n <= 1024 ? 32 : 8; }

// Panel variant (autotuner choice, negf_inv_variant): 0 = 16-column panels
// of 256 threads, 1 = 32-column panels of 512 threads (half the rank-NB
// update passes over the work matrix, a wider barrier per column).
int g_inv_variant = 0;

}  // namespace

int zinv_set_variant(int v) {
  if (v < 0 || v > 1) return Status::BAD_ARG;
  g_inv_variant = v;
  return Status::OK;
}

size_t zinv_work_bytes(int n, int batch) {
  const size_t nb = nb_for(n);
  size_t b = 0;
  b += sizeof(zc) * (size_t)n * n * batch;   // X
  b += sizeof(zc) * (size_t)n * nb * batch;  // W
  b += sizeof(zc) * nb * (size_t)n * batch;  // R
  b += sizeof(int) * (size_t)n * batch * 2;  // piv, used
  return b + 1024;
}

void zinv_work_carve(void* base, int n, int batch, InvWork* w) {
  const size_t nb = nb_for(n);
  char* p = (char*)base;
  auto take = [&](size_t bytes) {
    char* r = p;
    p += (bytes + 255) / 256 * 256;
    return r;
  };
  w->X = (zc*)take(sizeof(zc) * (size_t)n * n * batch);
  w->W = (zc*)take(sizeof(zc) * (size_t)n * nb * batch);
  w->R = (zc*)take(sizeof(zc) * nb * (size_t)n * batch);
  w->piv = (int*)take(sizeof(int) * (size_t)n * batch);
  w->used = (int*)take(sizeof(int) * (size_t)n * batch);
}

int zinv_launch(int n, const zc* A, long long lda, long long strideA, zc* Ainv, long long ldinv,
                long long strideInv, int batch, const InvWork& w, int* info, hipStream_t s) {
  if (n < 0 || batch < 0 || lda < (n > 1 ? n : 1) || ldinv < (n > 1 ? n : 1)) return Status::BAD_ARG;
  if (n == 0 || batch == 0) return Status::OK;
  if (n > 2048) return Status::TOO_LARGE;
  if (g_inv_variant == 1) {
    if (n <= 256) return run<256, 1, 32>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
    if (n <= 512) return run<512, 1, 32>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  }
  if (n <= 256) return run<256, 1, 16>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  if (n <= 512) return run<256, 2, 16>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  if (n <= 1024) return run<512, 2, 16>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  return run<512, 4, 8>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
}

}  // namespace negf
