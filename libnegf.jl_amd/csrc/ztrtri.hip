// Batched inverse of a triangular complex128 block (LAPACK ztrtri semantics,
// out of place): X = tri(A)^{-1}, uplo 'U' (non-unit or unit diagonal) or 'L',
// with X zero outside the triangle.  The selected inverse uses it to turn its
// triangular solves with U_ii and L_ii into single MFMA GEMMs.
//
// Recursive 2 x 2 blocking, bottom up:
//   leaves: the 32 x 32 diagonal blocks, one wave each: the block in LDS,
//           lane j computes column j of its inverse by substitution (its
//           serial chain grows with the square of the leaf: 32 beat 64 by
//           6 % on the selected inverse at b = 512, 16 was within noise of 32);
//   level s (s = 32, 64, 128, ...): every pair of neighbouring diagonal blocks
//           of size s (the second possibly ragged) is merged,
//             upper:  X12 = -X11 A12 X22     lower:  X21 = -X22 A21 X11,
//           as two GEMMs batched over (pair, instance) with the 2-D
//           instance grid of the GEMM kernel.
// Flops: n^3 / 3 complex multiply-adds (the ztrtri count), in log2(n / 32)
// GEMM levels plus the leaves.
#include "negf_internal.hpp"

namespace negf {

namespace {

constexpr int LEAF = 32;

// One wave per (leaf block, instance).  Zeroes the leaf's strictly
// off-triangle part of X and writes the leaf inverse.
__global__ __launch_bounds__(64) void trtri_leaf_kernel(int upper, int unit, int n,
                                                        const zc* __restrict__ A, long long lda,
                                                        long long strideA, zc* __restrict__ X,
                                                        long long ldx, long long strideX) {
  const int leaf = blockIdx.x, inst = blockIdx.y;
  const int o = leaf * LEAF;
  const int m = (n - o) < LEAF ? (n - o) : LEAF;
  const zc* Ai = A + inst * strideA + o + (long long)o * lda;
  zc* Xi = X + inst * strideX + o + (long long)o * ldx;
  // rows padded to 65 entries (column j of a row-major copy read across lanes)
  __shared__ zc As[LEAF][LEAF + 1];   // As[r][c] = A[o + r, o + c]
  __shared__ zc Xs[LEAF][LEAF + 1];   // Xs[r][j] = X[o + r, o + j]
  const int j = threadIdx.x;
  for (int c = 0; c < m; ++c)
    As[j][c] = j < m ? Ai[j + (long long)c * lda] : zmake(0.0, 0.0);
  __syncthreads();
  if (j < m) {
    if (upper) {
      // column j of U^{-1}: x_jj = 1 / u_jj, x_ij = -(sum_{k=i+1..j} u_ik x_kj) / u_ii
      const zc djj = unit ? zmake(1.0, 0.0) : zrcp(As[j][j]);
      Xs[j][j] = djj;
      for (int i = j - 1; i >= 0; --i) {
        zc acc = zmake(0.0, 0.0);
        for (int k = i + 1; k <= j; ++k) acc = zadd(acc, zmul(As[i][k], Xs[k][j]));
        const zc d = unit ? zmake(1.0, 0.0) : zrcp(As[i][i]);
        Xs[i][j] = zmake(-(acc.x * d.x - acc.y * d.y), -(acc.x * d.y + acc.y * d.x));
      }
      for (int i = j + 1; i < m; ++i) Xs[i][j] = zmake(0.0, 0.0);
    } else {
      // column j of L^{-1}: x_jj = 1 / l_jj, x_ij = -(sum_{k=j..i-1} l_ik x_kj) / l_ii
      Xs[j][j] = unit ? zmake(1.0, 0.0) : zrcp(As[j][j]);
      for (int i = j + 1; i < m; ++i) {
        zc acc = zmake(0.0, 0.0);
        for (int k = j; k < i; ++k) acc = zadd(acc, zmul(As[i][k], Xs[k][j]));
        const zc d = unit ? zmake(1.0, 0.0) : zrcp(As[i][i]);
        Xs[i][j] = zmake(-(acc.x * d.x - acc.y * d.y), -(acc.x * d.y + acc.y * d.x));
      }
      for (int i = 0; i < j; ++i) Xs[i][j] = zmake(0.0, 0.0);
    }
  }
  __syncthreads();
  for (int c = 0; c < m; ++c)
    if (j < m) Xi[j + (long long)c * ldx] = Xs[j][c];
}

GemmArgs pair_gemm(int opA, int m, int nn, int k, zc alpha, const zc* A, long long lda, const zc* B,
                   long long ldb, zc* D, long long ldd, long long pstrideA, long long pstrideB,
                   long long pstrideD, int pairs, long long sA, long long sB, long long sD,
                   int batch) {
  GemmArgs g{};
  g.opA = opA; g.opB = OP_N;
  g.m = m; g.n = nn; g.k = k;
  g.alpha = alpha; g.beta = zmake(0.0, 0.0);
  g.A = A; g.lda = lda; g.strideA = pstrideA; g.outerA = sA;
  g.B = B; g.ldb = ldb; g.strideB = pstrideB; g.outerB = sB;
  g.D = D; g.ldd = ldd; g.strideD = pstrideD; g.outerD = sD;
  g.inner = pairs;
  g.batch = pairs * batch;
  return g;
}

}  // namespace

size_t ztrtri_work_bytes(int n, int batch) {
  return sizeof(zc) * (size_t)n * n * batch / 2 + sizeof(zc) * (size_t)LEAF * n * batch + 256;
}

int ztrtri_launch(char uplo, char diag, int n, const zc* A, long long lda, long long strideA,
                  zc* X, long long ldx, long long strideX, int batch, zc* work, hipStream_t s) {
  if (n < 0 || batch < 0 || lda < (n > 1 ? n : 1) || ldx < (n > 1 ? n : 1)) return Status::BAD_ARG;
  if ((uplo != 'U' && uplo != 'L') || (diag != 'N' && diag != 'U')) return Status::BAD_ARG;
  if (n == 0 || batch == 0) return Status::OK;
  if (batch > 65535) return Status::TOO_LARGE;
  const int upper = uplo == 'U', unit = diag == 'U';
  // X = 0 everywhere, then the leaves, then the merges
  int rc = zfill_launch(n, n, zmake(0.0, 0.0), zmake(0.0, 0.0), X, ldx, strideX, batch, s);
  if (rc) return rc;
  const int nleaf = (n + LEAF - 1) / LEAF;
  hipLaunchKernelGGL(trtri_leaf_kernel, dim3(nleaf, batch), dim3(64), 0, s, upper, unit, n, A, lda,
                     strideA, X, ldx, strideX);
  const zc ONE = zmake(1.0, 0.0), MONE = zmake(-1.0, 0.0);
  for (int sz = LEAF; sz < n; sz *= 2) {
    // pairs (p): first block rows [2p sz, 2p sz + sz), second [2p sz + sz, 2p sz + sz + s2)
    const int npair_all = (n + 2 * sz - 1) / (2 * sz);
    const int last_s2 = n - (npair_all - 1) * 2 * sz - sz;   // second size of the last pair
    const long long dstep = 2LL * sz;                          // diagonal step between pairs
    for (int part = 0; part < 2; ++part) {
      // part 0: the full pairs (s2 = sz); part 1: the last pair if ragged
      int p0, np, s2;
      if (part == 0) {
        p0 = 0;
        np = (last_s2 == sz) ? npair_all : npair_all - 1;
        s2 = sz;
      } else {
        if (last_s2 == sz || last_s2 <= 0) continue;
        p0 = npair_all - 1;
        np = 1;
        s2 = last_s2;
      }
      if (np <= 0) continue;
      const long long o = (long long)p0 * dstep;
      // element (r, c) of block (R, C) of pair p at offset (o + p dstep + R sz + r) +
      // (o + p dstep + C sz + c) ld
      auto at = [&](const zc* base, long long ld, int R, int C) {
        return base + (o + R * sz) + (o + C * sz) * ld;
      };
      const long long psA = dstep * (lda + 1), psX = dstep * (ldx + 1);
      GemmArgs g1, g2;
      if (upper) {
        // T = A12 X22 (sz x s2), X12 = -X11 T (sz x s2)
        g1 = pair_gemm(OP_N, sz, s2, s2, ONE, at(A, lda, 0, 1), lda, at(X, ldx, 1, 1), ldx, work, sz,
                       psA, psX, (long long)sz * s2, np, strideA, strideX,
                       (long long)sz * s2 * np, batch);
        g2 = pair_gemm(OP_N, sz, s2, sz, MONE, at(X, ldx, 0, 0), ldx, work, sz,
                       const_cast<zc*>(at(X, ldx, 0, 1)),
                       ldx, psX, (long long)sz * s2, psX, np, strideX, (long long)sz * s2 * np,
                       strideX, batch);
      } else {
        // T = A21 X11 (s2 x sz), X21 = -X22 T (s2 x sz)
        g1 = pair_gemm(OP_N, s2, sz, sz, ONE, at(A, lda, 1, 0), lda, at(X, ldx, 0, 0), ldx, work, s2,
                       psA, psX, (long long)s2 * sz, np, strideA, strideX,
                       (long long)s2 * sz * np, batch);
        g2 = pair_gemm(OP_N, s2, sz, s2, MONE, at(X, ldx, 1, 1), ldx, work, s2,
                       const_cast<zc*>(at(X, ldx, 1, 0)),
                       ldx, psX, (long long)s2 * sz, psX, np, strideX, (long long)s2 * sz * np,
                       strideX, batch);
      }
      rc = zgemm_launch(g1, s);
      if (rc) return rc;
      rc = zgemm_launch(g2, s);
      if (rc) return rc;
    }
  }
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

}  // namespace negf
