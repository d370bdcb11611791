// Batched LU with partial pivoting, triangular solves, getrs and getri on
// complex128 blocks: the per-op kernels behind be_lu!/be_mldivide!/be_getrf!/
// be_trsm!/be_getri! (src/backend_cpu.jl:147-265).
//
// zgetrf: right-looking, NB-column panels.  The panel (rows k0..n-1 of NB
//   columns) lives in registers of one 256-thread workgroup per instance;
//   each column takes the izamax pivot (max |re|+|im|, first index on ties),
//   swaps through LDS, scales by the reciprocal and applies the rank-1
//   update inside the panel.  Then the row interchanges are applied to the
//   columns left and right of the panel, U12 = L11^{-1} A12 by the blocked
//   trsm, and A22 -= L21 U12 on the MFMA GEMM.  ipiv is LAPACK's (1-based).
// ztrsm: blocked, 16-wide diagonal blocks solved by one thread per right-hand
//   side (row or column) with the block in LDS, the rest as MFMA GEMMs.
#include "negf_internal.hpp"
#include <climits>

namespace negf {

namespace {

constexpr int PT = 256;
constexpr int TB = 16;   // trsm diagonal block

template <int RPT, int NB, int NT = PT>
__global__ __launch_bounds__(NT) void lu_panel_kernel(int n, int k0, zc* __restrict__ A,
                                                      long long lda, long long strideA,
                                                      int* __restrict__ ipiv, long long strideP,
                                                      int* __restrict__ info, bool pivot) {
  const int inst = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  zc* Ai = A + inst * strideA;
  int* ip = ipiv + inst * strideP;
  const int nbe = (n - k0) < NB ? (n - k0) : NB;

  __shared__ double red_v[NT / 64];
  __shared__ int red_i[NT / 64];
  __shared__ zc rowcur[NB];
  __shared__ zc rowpiv[NB];

  zc P[RPT][NB];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = k0 + tid + NT * q;
#pragma unroll
    for (int c = 0; c < NB; ++c)
      P[q][c] = (r < n && c < nbe) ? Ai[r + (long long)(k0 + c) * lda] : zmake(0.0, 0.0);
  }

#pragma unroll
  for (int jj = 0; jj < NB; ++jj) {
    if (jj < nbe) {
      const int cur = k0 + jj;
      // each lane's best row (the lowest on ties: rows grow with q), the
      // wave's maximum by DPP (no LDS round trips), the lowest row holding it
      // by ballot (smallest q, then lowest lane)
      double bv = -1.0;
      int bq = 0;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int r = k0 + tid + NT * q;
        // without pivoting the "search" sees only the diagonal row
        if (r < n && (pivot ? r >= cur : r == cur)) {
          const double v = zabs1(P[q][jj]);
          if (v > bv) { bv = v; bq = q; }
        }
      }
      const double wmax = wave_max(bv);
      int win = INT_MAX;
#pragma unroll
      for (int q = RPT - 1; q >= 0; --q) {
        const unsigned long long m = __ballot(bv >= 0.0 && bv == wmax && bq == q);
        if (m != 0ull) win = k0 + 64 * wave + (__ffsll(m) - 1) + NT * q;
      }
      if (lane == 0) { red_v[wave] = wmax; red_i[wave] = win; }
      __syncthreads();
      // all waves' pairs in one round of LDS reads, then a branch-free pick
      double rv[NT / 64];
      int rix[NT / 64];
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) {
        rv[w] = red_v[w];
        rix[w] = red_i[w];
      }
      // the largest value, then the lowest row holding it (LAPACK's first
      // maximum), as a max and a min: compares chained on the running pick
      // compile to scalar branches, the values being wave-uniform
      bv = rv[0];
#pragma unroll
      for (int w = 1; w < NT / 64; ++w) bv = fmax(bv, rv[w]);
      int bi = INT_MAX;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) bi = min(bi, rv[w] == bv ? rix[w] : INT_MAX);
      const int p = bi;
      if (tid == 0) {
        ip[cur] = p + 1;
        if (bv <= 0.0 && info[inst] == 0) info[inst] = cur + 1;
      }
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int r = k0 + tid + NT * q;
        if (r == cur) {
#pragma unroll
          for (int c = 0; c < NB; ++c) rowcur[c] = P[q][c];
        }
        if (r == p) {
#pragma unroll
          for (int c = 0; c < NB; ++c) rowpiv[c] = P[q][c];
        }
      }
      __syncthreads();
      // the pivot row in one round of reads, its pivot entry among them
      zc rp[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) rp[c] = rowpiv[c];
      const zc pv = rp[jj];
      const bool nz = (pv.x != 0.0 || pv.y != 0.0);
      const zc inv = nz ? zrcp(pv) : zmake(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int r = k0 + tid + NT * q;
        if (r < n && r >= cur) {
          if (r == cur) {
#pragma unroll
            for (int c = 0; c < NB; ++c) P[q][c] = rp[c];
          } else {
            if (r == p) {
#pragma unroll
              for (int c = 0; c < NB; ++c) P[q][c] = rowcur[c];
            }
            if (nz) {
              const zc l = zmul(P[q][jj], inv);
              P[q][jj] = l;
#pragma unroll
              for (int c = 0; c < NB; ++c)
                if (c > jj) P[q][c] = zsub(P[q][c], zmul(l, rp[c]));
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = k0 + tid + NT * q;
    if (r < n) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < nbe) Ai[r + (long long)(k0 + c) * lda] = P[q][c];
    }
  }
}

__global__ void laswp_kernel(int ncols, zc* __restrict__ B, long long ldb, long long strideB,
                             const int* __restrict__ ipiv, long long strideP, int k1, int k2,
                             int incx) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int inst = blockIdx.y;
  if (col >= ncols) return;
  zc* Bc = B + inst * strideB + (long long)col * ldb;
  const int* ip = ipiv + inst * strideP;
  if (incx > 0) {
    for (int i = k1; i < k2; ++i) {
      const int p = ip[i] - 1;
      if (p != i) { zc t = Bc[i]; Bc[i] = Bc[p]; Bc[p] = t; }
    }
  } else {
    for (int i = k2 - 1; i >= k1; --i) {
      const int p = ip[i] - 1;
      if (p != i) { zc t = Bc[i]; Bc[i] = Bc[p]; Bc[p] = t; }
    }
  }
}

// The interchanges of rows k1..k2-1 on every column outside [skip, skip +
// nskip): one launch for the columns left and right of a panel.
__global__ void laswp_outside_kernel(int ncols, int skip, int nskip, zc* __restrict__ A,
                                     long long lda, long long strideA,
                                     const int* __restrict__ ipiv, long long strideP, int k1,
                                     int k2) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int inst = blockIdx.y;
  if (col >= ncols - nskip) return;
  if (col >= skip) col += nskip;
  zc* Ac = A + inst * strideA + (long long)col * lda;
  const int* ip = ipiv + inst * strideP;
  for (int i = k1; i < k2; ++i) {
    const int p = ip[i] - 1;
    if (p != i) { zc t = Ac[i]; Ac[i] = Ac[p]; Ac[p] = t; }
  }
}

__global__ void zero_info_kernel(int* info, int batch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < batch) info[i] = 0;
}

// Effective triangle T = op(A)[o:o+nb, o:o+nb] staged into LDS as T[i][j].
__device__ inline zc op_elem(const zc* A, long long lda, int op, int i, int j) {
  if (op == OP_N) return A[i + (long long)j * lda];
  zc v = A[j + (long long)i * lda];
  return op == OP_C ? zconj(v) : v;
}

// Left side: X = T^{-1} B on rows [o, o+nb) of B, one thread per column.
// lower: forward substitution, else backward.
__global__ __launch_bounds__(256) void trsm_left_diag_kernel(int o, int nb, int ncols,
                                                             const zc* __restrict__ A,
                                                             long long lda, long long strideA,
                                                             int op, bool lower, bool unit,
                                                             zc* __restrict__ B, long long ldb,
                                                             long long strideB) {
  __shared__ zc T[TB][TB + 1];
  const int inst = blockIdx.y;
  const zc* Ai = A + inst * strideA;
  for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
    const int i = e % TB, j = e / TB;
    T[i][j] = (i < nb && j < nb) ? op_elem(Ai, lda, op, o + i, o + j) : zmake(0.0, 0.0);
  }
  __syncthreads();
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncols) return;
  zc* Bc = B + inst * strideB + (long long)col * ldb + o;
  zc x[TB];
#pragma unroll
  for (int i = 0; i < TB; ++i) x[i] = i < nb ? Bc[i] : zmake(0.0, 0.0);
  if (lower) {
#pragma unroll
    for (int i = 0; i < TB; ++i) {
      if (i < nb) {
        zc s = x[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s = zsub(s, zmul(T[i][k], x[k]));
        x[i] = unit ? s : zmul(s, zrcp(T[i][i]));
      }
    }
  } else {
#pragma unroll
    for (int i = TB - 1; i >= 0; --i) {
      if (i < nb) {
        zc s = x[i];
#pragma unroll
        for (int k = i + 1; k < TB; ++k)
          if (k < nb) s = zsub(s, zmul(T[i][k], x[k]));
        x[i] = unit ? s : zmul(s, zrcp(T[i][i]));
      }
    }
  }
#pragma unroll
  for (int i = 0; i < TB; ++i)
    if (i < nb) Bc[i] = x[i];
}

// Right side: X = B T^{-1} on columns [o, o+nb) of B, one thread per row.
// upper: forward over columns, else backward.
__global__ __launch_bounds__(256) void trsm_right_diag_kernel(int o, int nb, int nrows,
                                                              const zc* __restrict__ A,
                                                              long long lda, long long strideA,
                                                              int op, bool upper, bool unit,
                                                              zc* __restrict__ B, long long ldb,
                                                              long long strideB) {
  __shared__ zc T[TB][TB + 1];
  const int inst = blockIdx.y;
  const zc* Ai = A + inst * strideA;
  for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
    const int i = e % TB, j = e / TB;
    T[i][j] = (i < nb && j < nb) ? op_elem(Ai, lda, op, o + i, o + j) : zmake(0.0, 0.0);
  }
  __syncthreads();
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  zc* Br = B + inst * strideB + row + (long long)o * ldb;
  zc x[TB];
#pragma unroll
  for (int j = 0; j < TB; ++j) x[j] = j < nb ? Br[(long long)j * ldb] : zmake(0.0, 0.0);
  if (upper) {
#pragma unroll
    for (int j = 0; j < TB; ++j) {
      if (j < nb) {
        zc s = x[j];
#pragma unroll
        for (int k = 0; k < j; ++k) s = zsub(s, zmul(x[k], T[k][j]));
        x[j] = unit ? s : zmul(s, zrcp(T[j][j]));
      }
    }
  } else {
#pragma unroll
    for (int j = TB - 1; j >= 0; --j) {
      if (j < nb) {
        zc s = x[j];
#pragma unroll
        for (int k = j + 1; k < TB; ++k)
          if (k < nb) s = zsub(s, zmul(x[k], T[k][j]));
        x[j] = unit ? s : zmul(s, zrcp(T[j][j]));
      }
    }
  }
#pragma unroll
  for (int j = 0; j < TB; ++j)
    if (j < nb) Br[(long long)j * ldb] = x[j];
}

int op_code(char c) {
  switch (c) {
    case 'N': case 'n': return OP_N;
    case 'T': case 't': return OP_T;
    case 'C': case 'c': return OP_C;
  }
  return -1;
}

// GEMM helper: B_sub -= opA_sub * X_sub (all column-major pointers).
int gemm_sub(int opA, int m, int n, int k, const zc* A, long long lda, long long strideA,
             const zc* Bx, long long ldbx, long long strideBx, int opB, zc* C, long long ldc,
             long long strideC, int batch, hipStream_t s) {
  GemmArgs g{};
  g.opA = opA; g.opB = opB;
  g.m = m; g.n = n; g.k = k;
  g.alpha = zmake(-1.0, 0.0); g.beta = zmake(1.0, 0.0);
  g.A = A; g.lda = lda; g.strideA = strideA;
  g.B = Bx; g.ldb = ldbx; g.strideB = strideBx;
  g.C = C; g.ldc = ldc; g.strideC = strideC;
  g.D = C; g.ldd = ldc; g.strideD = strideC;
  g.batch = batch;
  return zgemm_launch(g, s);
}

template <int RPT, int NB>
int getrf_run(int n, zc* A, long long lda, long long strideA, int* ipiv, long long strideP,
              int* info, int batch, bool pivot, hipStream_t s) {
  hipLaunchKernelGGL(zero_info_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, info, batch);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nbe = (n - k0) < NB ? (n - k0) : NB;
    hipLaunchKernelGGL((lu_panel_kernel<RPT, NB>), dim3(batch), dim3(PT), 0, s, n, k0, A, lda,
                       strideA, ipiv, strideP, info, pivot);
    // interchanges on the columns left and right of the panel, one launch
    if (pivot && n > nbe)
      hipLaunchKernelGGL(laswp_outside_kernel, dim3((n - nbe + 255) / 256, batch), dim3(256), 0, s,
                         n, k0, nbe, A, lda, strideA, ipiv, strideP, k0, k0 + nbe);
    const int rest = n - k0 - nbe;
    if (rest > 0) {
      zc* A12 = A + k0 + (long long)(k0 + nbe) * lda;
      int rc = ztrsm_launch('L', 'L', 'N', 'U', nbe, rest, zmake(1.0, 0.0),
                            A + k0 + (long long)k0 * lda, lda, strideA, A12, lda, strideA, batch,
                            s);
      if (rc) return rc;
      rc = gemm_sub(OP_N, rest, rest, nbe, A + (k0 + nbe) + (long long)k0 * lda, lda, strideA,
                    A12, lda, strideA, OP_N, A + (k0 + nbe) + (long long)(k0 + nbe) * lda, lda,
                    strideA, batch, s);
      if (rc) return rc;
    }
  }
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

// Two-level blocking for n > 2048, where the register panel of one
// NT-thread workgroup holds RPT x NB elements per thread: outer panels of
// OB columns are factored NB columns at a time (the panel kernel on rows
// j0..n-1, the interchanges and the trsm / rank-NB update confined to the
// outer panel's columns), then the outer panel's interchanges go to the
// columns left and right of it, U12 = L11^{-1} A12 (OB rows) and the trailing
// update A22 -= L21 U12 runs as one GEMM with k = OB on the matrix cores.
// The rank-NB updates touch only n x OB elements, so the HBM traffic is that
// of the OB-wide blocking, not of NB.
template <int NT, int RPT, int NB, int OB>
int getrf_blocked(int n, zc* A, long long lda, long long strideA, int* ipiv, long long strideP,
                  int* info, int batch, bool pivot, hipStream_t s) {
  hipLaunchKernelGGL(zero_info_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, info, batch);
  for (int k0 = 0; k0 < n; k0 += OB) {
    const int ob = (n - k0) < OB ? (n - k0) : OB;
    for (int j0 = k0; j0 < k0 + ob; j0 += NB) {
      const int nbe = (k0 + ob - j0) < NB ? (k0 + ob - j0) : NB;
      hipLaunchKernelGGL((lu_panel_kernel<RPT, NB, NT>), dim3(batch), dim3(NT), 0, s, n, j0, A,
                         lda, strideA, ipiv, strideP, info, pivot);
      const int right = k0 + ob - (j0 + nbe);   // columns of the outer panel right of j0
      if (pivot && j0 > k0)
        hipLaunchKernelGGL(laswp_kernel, dim3((j0 - k0 + 255) / 256, batch), dim3(256), 0, s,
                           j0 - k0, A + (long long)k0 * lda, lda, strideA, ipiv, strideP, j0,
                           j0 + nbe, 1);
      if (right > 0) {
        zc* A12 = A + j0 + (long long)(j0 + nbe) * lda;
        if (pivot)
          hipLaunchKernelGGL(laswp_kernel, dim3((right + 255) / 256, batch), dim3(256), 0, s,
                             right, A + (long long)(j0 + nbe) * lda, lda, strideA, ipiv, strideP,
                             j0, j0 + nbe, 1);
        int rc = ztrsm_launch('L', 'L', 'N', 'U', nbe, right, zmake(1.0, 0.0),
                              A + j0 + (long long)j0 * lda, lda, strideA, A12, lda, strideA, batch,
                              s);
        if (rc) return rc;
        if (n - j0 - nbe > 0) {
          rc = gemm_sub(OP_N, n - j0 - nbe, right, nbe, A + (j0 + nbe) + (long long)j0 * lda, lda,
                        strideA, A12, lda, strideA, OP_N,
                        A + (j0 + nbe) + (long long)(j0 + nbe) * lda, lda, strideA, batch, s);
          if (rc) return rc;
        }
      }
    }
    if (pivot && k0 > 0)
      hipLaunchKernelGGL(laswp_kernel, dim3((k0 + 255) / 256, batch), dim3(256), 0, s, k0, A, lda,
                         strideA, ipiv, strideP, k0, k0 + ob, 1);
    const int rest = n - k0 - ob;
    if (rest > 0) {
      zc* A12 = A + k0 + (long long)(k0 + ob) * lda;
      if (pivot)
        hipLaunchKernelGGL(laswp_kernel, dim3((rest + 255) / 256, batch), dim3(256), 0, s, rest,
                           A + (long long)(k0 + ob) * lda, lda, strideA, ipiv, strideP, k0, k0 + ob,
                           1);
      int rc = ztrsm_launch('L', 'L', 'N', 'U', ob, rest, zmake(1.0, 0.0),
                            A + k0 + (long long)k0 * lda, lda, strideA, A12, lda, strideA, batch, s);
      if (rc) return rc;
      rc = gemm_sub(OP_N, rest, rest, ob, A + (k0 + ob) + (long long)k0 * lda, lda, strideA, A12,
                    lda, strideA, OP_N, A + (k0 + ob) + (long long)(k0 + ob) * lda, lda, strideA,
                    batch, s);
      if (rc) return rc;
    }
  }
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

}  // namespace

int zgetrf_launch(int n, zc* A, long long lda, long long strideA, int* ipiv, long long strideP,
                  int* info, int batch, hipStream_t s, bool pivot) {
  if (n < 0 || batch < 0 || lda < (n > 1 ? n : 1)) return Status::BAD_ARG;
  if (n == 0 || batch == 0) return Status::OK;
  if (n > kLuMaxN) return Status::TOO_LARGE;
  // 512-thread panels (two waves per SIMD, so the 32 complex values per
  // thread stay in registers)
  if (n > 8192) return getrf_blocked<512, 32, 1, 64>(n, A, lda, strideA, ipiv, strideP, info, batch, pivot, s);
  if (n > 4096) return getrf_blocked<512, 16, 2, 64>(n, A, lda, strideA, ipiv, strideP, info, batch, pivot, s);
  if (n > 2048) return getrf_blocked<512, 8, 4, 64>(n, A, lda, strideA, ipiv, strideP, info, batch, pivot, s);
  if (n <= 256) return getrf_run<1, 16>(n, A, lda, strideA, ipiv, strideP, info, batch, pivot, s);
  if (n <= 512) return getrf_run<2, 16>(n, A, lda, strideA, ipiv, strideP, info, batch, pivot, s);
  if (n <= 1024) return getrf_run<4, 8>(n, A, lda, strideA, ipiv, strideP, info, batch, pivot, s);
  return getrf_run<8, 4>(n, A, lda, strideA, ipiv, strideP, info, batch, pivot, s);
}

int zlaswp_launch(int ncols, zc* B, long long ldb, long long strideB, const int* ipiv,
                  long long strideP, int k1, int k2, int incx, int batch, hipStream_t s) {
  if (ncols <= 0 || batch <= 0 || k2 <= k1) return Status::OK;
  hipLaunchKernelGGL(laswp_kernel, dim3((ncols + 255) / 256, batch), dim3(256), 0, s, ncols, B, ldb,
                     strideB, ipiv, strideP, k1, k2, incx);
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

int ztrsm_launch(char side, char uplo, char trans, char diag, int m, int n, zc alpha,
                 const zc* A, long long lda, long long strideA, zc* B, long long ldb,
                 long long strideB, int batch, hipStream_t s) {
  const int op = op_code(trans);
  const bool left = (side == 'L' || side == 'l');
  const bool lo = (uplo == 'L' || uplo == 'l');
  const bool unit = (diag == 'U' || diag == 'u');
  if (op < 0 || (!left && side != 'R' && side != 'r') || (!lo && uplo != 'U' && uplo != 'u') ||
      (!unit && diag != 'N' && diag != 'n'))
    return Status::BAD_ARG;
  if (m < 0 || n < 0 || batch < 0) return Status::BAD_ARG;
  const int na = left ? m : n;
  if (lda < (na > 1 ? na : 1) || ldb < (m > 1 ? m : 1)) return Status::BAD_ARG;
  if (m == 0 || n == 0 || batch == 0) return Status::OK;
  if (alpha.x != 1.0 || alpha.y != 0.0) {
    int rc = zscale_launch(m, n, alpha, nullptr, 0, B, ldb, strideB, B, ldb, strideB, batch, s);
    if (rc) return rc;
  }
  // effective lower triangle of op(A)?
  const bool eff_lower = (lo != (op != OP_N));
  // op(A)[r0:r0+mr, c0:c0+mc] as a GEMM operand
  auto sub = [&](int r0, int c0) -> const zc* {
    return op == OP_N ? A + r0 + (long long)c0 * lda : A + c0 + (long long)r0 * lda;
  };
  int rc = 0;
  if (left) {
    const dim3 grid((n + 255) / 256, batch);
    if (eff_lower) {
      for (int r0 = 0; r0 < m; r0 += TB) {
        const int r1 = (r0 + TB < m) ? r0 + TB : m;
        hipLaunchKernelGGL(trsm_left_diag_kernel, grid, dim3(256), 0, s, r0, r1 - r0, n, A, lda,
                           strideA, op, true, unit, B, ldb, strideB);
        if (r1 < m)
          rc = gemm_sub(op, m - r1, n, r1 - r0, sub(r1, r0), lda, strideA, B + r0, ldb, strideB,
                        OP_N, B + r1, ldb, strideB, batch, s);
        if (rc) return rc;
      }
    } else {
      const int nblk = (m + TB - 1) / TB;
      for (int bI = nblk - 1; bI >= 0; --bI) {
        const int r0 = bI * TB;
        const int r1 = (r0 + TB < m) ? r0 + TB : m;
        hipLaunchKernelGGL(trsm_left_diag_kernel, grid, dim3(256), 0, s, r0, r1 - r0, n, A, lda,
                           strideA, op, false, unit, B, ldb, strideB);
        if (r0 > 0)
          rc = gemm_sub(op, r0, n, r1 - r0, sub(0, r0), lda, strideA, B + r0, ldb, strideB, OP_N,
                        B, ldb, strideB, batch, s);
        if (rc) return rc;
      }
    }
  } else {
    const dim3 grid((m + 255) / 256, batch);
    const bool eff_upper = !eff_lower;
    if (eff_upper) {
      for (int c0 = 0; c0 < n; c0 += TB) {
        const int c1 = (c0 + TB < n) ? c0 + TB : n;
        hipLaunchKernelGGL(trsm_right_diag_kernel, grid, dim3(256), 0, s, c0, c1 - c0, m, A, lda,
                           strideA, op, true, unit, B, ldb, strideB);
        if (c1 < n) {
          // B[:, c1:] -= X[:, c0:c1] op(A)[c0:c1, c1:]
          GemmArgs g{};
          g.opA = OP_N; g.opB = op;
          g.m = m; g.n = n - c1; g.k = c1 - c0;
          g.alpha = zmake(-1.0, 0.0); g.beta = zmake(1.0, 0.0);
          g.A = B + (long long)c0 * ldb; g.lda = ldb; g.strideA = strideB;
          g.B = op == OP_N ? A + c0 + (long long)c1 * lda : A + c1 + (long long)c0 * lda;
          g.ldb = lda; g.strideB = strideA;
          g.C = B + (long long)c1 * ldb; g.ldc = ldb; g.strideC = strideB;
          g.D = B + (long long)c1 * ldb; g.ldd = ldb; g.strideD = strideB;
          g.batch = batch;
          rc = zgemm_launch(g, s);
          if (rc) return rc;
        }
      }
    } else {
      const int nblk = (n + TB - 1) / TB;
      for (int bI = nblk - 1; bI >= 0; --bI) {
        const int c0 = bI * TB;
        const int c1 = (c0 + TB < n) ? c0 + TB : n;
        hipLaunchKernelGGL(trsm_right_diag_kernel, grid, dim3(256), 0, s, c0, c1 - c0, m, A, lda,
                           strideA, op, false, unit, B, ldb, strideB);
        if (c0 > 0) {
          // B[:, :c0] -= X[:, c0:c1] op(A)[c0:c1, :c0]
          GemmArgs g{};
          g.opA = OP_N; g.opB = op;
          g.m = m; g.n = c0; g.k = c1 - c0;
          g.alpha = zmake(-1.0, 0.0); g.beta = zmake(1.0, 0.0);
          g.A = B + (long long)c0 * ldb; g.lda = ldb; g.strideA = strideB;
          g.B = op == OP_N ? A + c0 : A + (long long)c0 * lda;
          g.ldb = lda; g.strideB = strideA;
          g.C = B; g.ldc = ldb; g.strideC = strideB;
          g.D = B; g.ldd = ldb; g.strideD = strideB;
          g.batch = batch;
          rc = zgemm_launch(g, s);
          if (rc) return rc;
        }
      }
    }
  }
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

int zgetrs_launch(char trans, int n, int nrhs, const zc* A, long long lda, long long strideA,
                  const int* ipiv, long long strideP, zc* B, long long ldb, long long strideB,
                  int batch, hipStream_t s) {
  const int op = op_code(trans);
  if (op < 0 || n < 0 || nrhs < 0 || batch < 0) return Status::BAD_ARG;
  if (lda < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1)) return Status::BAD_ARG;
  if (n == 0 || nrhs == 0 || batch == 0) return Status::OK;
  const zc one = zmake(1.0, 0.0);
  int rc;
  if (op == OP_N) {
    rc = zlaswp_launch(nrhs, B, ldb, strideB, ipiv, strideP, 0, n, 1, batch, s);
    if (rc) return rc;
    rc = ztrsm_launch('L', 'L', 'N', 'U', n, nrhs, one, A, lda, strideA, B, ldb, strideB, batch, s);
    if (rc) return rc;
    return ztrsm_launch('L', 'U', 'N', 'N', n, nrhs, one, A, lda, strideA, B, ldb, strideB, batch,
                        s);
  }
  const char t = op == OP_T ? 'T' : 'C';
  rc = ztrsm_launch('L', 'U', t, 'N', n, nrhs, one, A, lda, strideA, B, ldb, strideB, batch, s);
  if (rc) return rc;
  rc = ztrsm_launch('L', 'L', t, 'U', n, nrhs, one, A, lda, strideA, B, ldb, strideB, batch, s);
  if (rc) return rc;
  return zlaswp_launch(nrhs, B, ldb, strideB, ipiv, strideP, 0, n, -1, batch, s);
}

int zgetri_launch(int n, zc* A, long long lda, long long strideA, const int* ipiv,
                  long long strideP, int batch, zc* work, hipStream_t s) {
  if (n < 0 || batch < 0 || lda < (n > 1 ? n : 1)) return Status::BAD_ARG;
  if (n == 0 || batch == 0) return Status::OK;
  const long long ws = (long long)n * n;
  int rc = zcopy2d_launch(n, n, A, lda, strideA, work, n, ws, batch, s);
  if (rc) return rc;
  rc = zfill_launch(n, n, zmake(0.0, 0.0), zmake(1.0, 0.0), A, lda, strideA, batch, s);
  if (rc) return rc;
  return zgetrs_launch('N', n, n, work, n, ws, ipiv, strideP, A, lda, strideA, batch, s);
}

namespace {

// perm[j]: the column of inv(U) inv(L) that becomes column j of inv(A), from
// LAPACK's interchanges (getri applies them to the columns from the last to
// the first); one thread, n steps.
__global__ void perm_of_ipiv_kernel(int n, const int* __restrict__ ipiv, int* __restrict__ perm) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int j = 0; j < n; ++j) perm[j] = j;
  for (int j = n - 1; j >= 0; --j) {
    const int p = ipiv[j] - 1;
    if (p != j) {
      const int t = perm[j];
      perm[j] = perm[p];
      perm[p] = t;
    }
  }
}

// out[:, j] = X[:, perm[j]]; grid (row chunks of 256, columns)
__global__ __launch_bounds__(256) void gather_cols_kernel(int n, const zc* __restrict__ X,
                                                          const int* __restrict__ perm,
                                                          zc* __restrict__ out, long long ldo) {
  const int j = blockIdx.y;
  const zc* src = X + (long long)perm[j] * n;
  zc* dst = out + (long long)j * ldo;
  for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256) dst[r] = src[r];
}

}  // namespace

size_t zinv_large_work_bytes(int n) {
  const size_t nn = (size_t)n * n * sizeof(zc);
  const size_t ints = ((size_t)2 * n * sizeof(int) + 255) / 256 * 256;
  return 3 * ((nn + 255) / 256 * 256) + ints + ztrtri_work_bytes(n, 1) + 256;
}

int zinv_large_launch(int n, const zc* A, long long lda, zc* Ainv, long long ldinv, void* work,
                      int* info, hipStream_t s) {
  if (n <= 0 || n > kLuMaxN || lda < n || ldinv < n || !work) return Status::BAD_ARG;
  const size_t nn = ((size_t)n * n * sizeof(zc) + 255) / 256 * 256;
  char* p = (char*)work;
  zc* LU = (zc*)p;  p += nn;
  zc* Li = (zc*)p;  p += nn;
  zc* Ui = (zc*)p;  p += nn;
  int* ipiv = (int*)p;
  int* perm = ipiv + n;
  p += ((size_t)2 * n * sizeof(int) + 255) / 256 * 256;
  zc* tw = (zc*)p;
  int rc = zcopy2d_launch(n, n, A, lda, 0, LU, n, 0, 1, s);
  if (rc) return rc;
  rc = zgetrf_launch(n, LU, n, 0, ipiv, 0, info, 1, s, true);
  if (rc) return rc;
  rc = ztrtri_launch('L', 'U', n, LU, n, 0, Li, n, 0, 1, tw, s);
  if (rc) return rc;
  rc = ztrtri_launch('U', 'N', n, LU, n, 0, Ui, n, 0, 1, tw, s);
  if (rc) return rc;
  // inv(U) inv(L) into the LU buffer (no longer needed)
  GemmArgs g{};
  g.opA = OP_N; g.opB = OP_N;
  g.m = g.n = g.k = n;
  g.alpha = zmake(1.0, 0.0); g.beta = zmake(0.0, 0.0);
  g.A = Ui; g.lda = n; g.B = Li; g.ldb = n;
  g.C = nullptr; g.ldc = n; g.D = LU; g.ldd = n;
  g.batch = 1;
  rc = zgemm_launch(g, s);
  if (rc) return rc;
  hipLaunchKernelGGL(perm_of_ipiv_kernel, dim3(1), dim3(64), 0, s, n, ipiv, perm);
  hipLaunchKernelGGL(gather_cols_kernel, dim3((n + 255) / 256 < 16 ? (n + 255) / 256 : 16, n),
                     dim3(256), 0, s, n, LU, perm, Ainv, ldinv);
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}


namespace {

// pinv per instance (one workgroup each): LAPACK's swaps (1-based ipiv,
// applied in order) give the row order pi of P A (row r of P A is row pi[r]
// of A); one thread runs the swaps in LDS (position k is final after step k),
// then all threads write pinv = pi^{-1}.  n <= kPermMaxN.
__global__ __launch_bounds__(256) void pivot_pinv_kernel(int n, const int* __restrict__ ipiv,
                                                         long long sP, int* __restrict__ pinv,
                                                         long long sI) {
  __shared__ int pi[kPermMaxN], ip[kPermMaxN];
  const int e = blockIdx.x;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    pi[k] = k;
    ip[k] = ipiv[(long long)e * sP + k] - 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < n; ++k) {
      const int q = ip[k];
      const int t = pi[k];
      pi[k] = pi[q];
      pi[q] = t;
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < n; r += blockDim.x) pinv[(long long)e * sI + pi[r]] = r;
}

// XP[:, c] = X[:, pinv[c]] over (8-column chunk, instance)
__global__ __launch_bounds__(256) void gather_cols_batched_kernel(int n, const zc* __restrict__ X,
                                                                  long long sX, zc* __restrict__ XP,
                                                                  long long ldxp, long long sXP,
                                                                  const int* __restrict__ pinv,
                                                                  long long sI) {
  const int e = blockIdx.y;
  const zc* src = X + (long long)e * sX;
  zc* dst = XP + (long long)e * sXP;
  const int* pv = pinv + (long long)e * sI;
  for (int cc = 0; cc < 8; ++cc) {
    const int c = blockIdx.x * 8 + cc;
    if (c >= n) break;
    const zc* col = src + (long long)pv[c] * n;
    for (int r = threadIdx.x; r < n; r += blockDim.x) dst[r + (long long)c * ldxp] = col[r];
  }
}

size_t al256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

int ztimes_perm_launch(int n, const zc* X, long long sX, zc* XP, long long ldxp, long long sXP,
                       const int* ipiv, long long sP, int* pinv, int batch, hipStream_t s) {
  if (n < 0 || batch < 0 || ldxp < (n > 1 ? n : 1)) return Status::BAD_ARG;
  if (n == 0 || batch == 0) return Status::OK;
  if (n > kPermMaxN || batch > 65535) return Status::TOO_LARGE;
  hipLaunchKernelGGL(pivot_pinv_kernel, dim3(batch), dim3(256), 0, s, n, ipiv, sP, pinv,
                     (long long)n);
  hipLaunchKernelGGL(gather_cols_batched_kernel, dim3((n + 7) / 8, batch), dim3(256), 0, s, n, X,
                     sX, XP, ldxp, sXP, pinv, (long long)n);
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

size_t zgetrs_inv_work_bytes(int n, int nrhs, int batch) {
  const size_t nn = al256(sizeof(zc) * (size_t)n * n * batch);
  const size_t nr = al256(sizeof(zc) * (size_t)n * nrhs * batch);
  const size_t pv = al256(sizeof(int) * (size_t)n * batch);
  return 3 * nn + nr + pv + al256(ztrtri_work_bytes(n, batch)) + 256;
}

int zgetrs_inv_launch(char trans, int n, int nrhs, const zc* A, long long lda, long long strideA,
                      const int* ipiv, long long strideP, zc* B, long long ldb, long long strideB,
                      int batch, void* work, hipStream_t s) {
  const int op = op_code(trans);
  if (op < 0 || n < 0 || nrhs < 0 || batch < 0) return Status::BAD_ARG;
  if (lda < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1)) return Status::BAD_ARG;
  if (n == 0 || nrhs == 0 || batch == 0) return Status::OK;
  if (n > kPermMaxN || batch > 65535) return Status::TOO_LARGE;
  if (!work) return Status::BAD_ARG;
  const long long tn = (long long)n * n, tr = (long long)n * nrhs;
  char* p = (char*)work;
  zc* Li = (zc*)p;  p += al256(sizeof(zc) * (size_t)tn * batch);
  zc* Ui = (zc*)p;  p += al256(sizeof(zc) * (size_t)tn * batch);
  zc* M = (zc*)p;   p += al256(sizeof(zc) * (size_t)tn * batch);
  zc* T = (zc*)p;   p += al256(sizeof(zc) * (size_t)tr * batch);
  int* pinv = (int*)p;  p += al256(sizeof(int) * (size_t)n * batch);
  zc* tw = (zc*)p;
  int rc = ztrtri_launch('L', 'U', n, A, lda, strideA, Li, n, tn, batch, tw, s);
  if (rc) return rc;
  rc = ztrtri_launch('U', 'N', n, A, lda, strideA, Ui, n, tn, batch, tw, s);
  if (rc) return rc;
  // M = inv(L) P: inv(A) = inv(U) M
  rc = ztimes_perm_launch(n, Li, tn, M, n, tn, ipiv, strideP, pinv, batch, s);
  if (rc) return rc;
  // 'N': T = M B, B = inv(U) T.  'T' / 'C': op(inv(A)) = op(M) op(inv(U)), so
  // T = op(inv(U)) B, B = op(M) T.
  auto prod = [&](int opX, const zc* X, const zc* Y, long long ldy, long long sY, zc* Z,
                  long long ldz, long long sZ) {
    GemmArgs g{};
    g.opA = opX; g.opB = OP_N;
    g.m = n; g.n = nrhs; g.k = n;
    g.alpha = zmake(1.0, 0.0); g.beta = zmake(0.0, 0.0);
    g.A = X; g.lda = n; g.strideA = tn;
    g.B = Y; g.ldb = ldy; g.strideB = sY;
    g.C = nullptr; g.ldc = ldz; g.strideC = sZ;
    g.D = Z; g.ldd = ldz; g.strideD = sZ;
    g.batch = batch;
    return zgemm_launch(g, s);
  };
  if (op == OP_N) {
    rc = prod(OP_N, M, B, ldb, strideB, T, n, tr);
    if (rc) return rc;
    rc = prod(OP_N, Ui, T, n, tr, B, ldb, strideB);
  } else {
    rc = prod(op, Ui, B, ldb, strideB, T, n, tr);
    if (rc) return rc;
    rc = prod(op, M, T, n, tr, B, ldb, strideB);
  }
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

size_t ztrsm_inv_work_bytes(int na, int m, int n, int batch) {
  return al256(sizeof(zc) * (size_t)na * na * batch) + al256(sizeof(zc) * (size_t)m * n * batch) +
         al256(ztrtri_work_bytes(na, batch)) + 256;
}

int ztrsm_inv_launch(char side, char uplo, char trans, char diag, int m, int n, zc alpha,
                     const zc* A, long long lda, long long strideA, zc* B, long long ldb,
                     long long strideB, int batch, void* work, hipStream_t s) {
  const int op = op_code(trans);
  const bool left = (side == 'L' || side == 'l');
  const bool lo = (uplo == 'L' || uplo == 'l');
  const bool unit = (diag == 'U' || diag == 'u');
  if (op < 0 || (!left && side != 'R' && side != 'r') || (!lo && uplo != 'U' && uplo != 'u') ||
      (!unit && diag != 'N' && diag != 'n'))
    return Status::BAD_ARG;
  if (m < 0 || n < 0 || batch < 0) return Status::BAD_ARG;
  const int na = left ? m : n;
  if (lda < (na > 1 ? na : 1) || ldb < (m > 1 ? m : 1)) return Status::BAD_ARG;
  if (m == 0 || n == 0 || batch == 0) return Status::OK;
  if (batch > 65535) return Status::TOO_LARGE;
  if (!work) return Status::BAD_ARG;
  const long long ta = (long long)na * na, tb = (long long)m * n;
  char* p = (char*)work;
  zc* Ti = (zc*)p;  p += al256(sizeof(zc) * (size_t)ta * batch);
  zc* T = (zc*)p;   p += al256(sizeof(zc) * (size_t)tb * batch);
  zc* tw = (zc*)p;
  int rc = ztrtri_launch(lo ? 'L' : 'U', unit ? 'U' : 'N', na, A, lda, strideA, Ti, na, ta, batch,
                         tw, s);
  if (rc) return rc;
  rc = zcopy2d_launch(m, n, B, ldb, strideB, T, m, tb, batch, s);
  if (rc) return rc;
  // left: B = alpha op(inv(A)) B; right: B = alpha B op(inv(A))
  GemmArgs g{};
  g.m = m; g.n = n; g.k = na;
  g.alpha = alpha; g.beta = zmake(0.0, 0.0);
  if (left) {
    g.opA = op; g.opB = OP_N;
    g.A = Ti; g.lda = na; g.strideA = ta;
    g.B = T; g.ldb = m; g.strideB = tb;
  } else {
    g.opA = OP_N; g.opB = op;
    g.A = T; g.lda = m; g.strideA = tb;
    g.B = Ti; g.ldb = na; g.strideB = ta;
  }
  g.C = nullptr; g.ldc = ldb; g.strideC = strideB;
  g.D = B; g.ldd = ldb; g.strideD = strideB;
  g.batch = batch;
  rc = zgemm_launch(g, s);
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

size_t zgetri_inv_work_bytes(int n, int batch) {
  return 2 * al256(sizeof(zc) * (size_t)n * n * batch) + al256(sizeof(int) * (size_t)n * batch) +
         al256(ztrtri_work_bytes(n, batch)) + 256;
}

int zgetri_inv_launch(int n, zc* A, long long lda, long long strideA, const int* ipiv,
                      long long strideP, int batch, zc* W, void* work, hipStream_t s) {
  if (n < 0 || batch < 0 || lda < (n > 1 ? n : 1)) return Status::BAD_ARG;
  if (n == 0 || batch == 0) return Status::OK;
  if (n > kPermMaxN || batch > 65535) return Status::TOO_LARGE;
  if (!work || (ipiv && !W)) return Status::BAD_ARG;
  const long long tn = (long long)n * n;
  char* p = (char*)work;
  zc* Li = (zc*)p;  p += al256(sizeof(zc) * (size_t)tn * batch);
  zc* Ui = (zc*)p;  p += al256(sizeof(zc) * (size_t)tn * batch);
  int* pinv = (int*)p;  p += al256(sizeof(int) * (size_t)n * batch);
  zc* tw = (zc*)p;
  int rc = ztrtri_launch('L', 'U', n, A, lda, strideA, Li, n, tn, batch, tw, s);
  if (rc) return rc;
  rc = ztrtri_launch('U', 'N', n, A, lda, strideA, Ui, n, tn, batch, tw, s);
  if (rc) return rc;
  // inv(A) = inv(U) inv(L) P: the product into W (with interchanges) or A
  GemmArgs g{};
  g.opA = OP_N; g.opB = OP_N;
  g.m = g.n = g.k = n;
  g.alpha = zmake(1.0, 0.0); g.beta = zmake(0.0, 0.0);
  g.A = Ui; g.lda = n; g.strideA = tn;
  g.B = Li; g.ldb = n; g.strideB = tn;
  g.C = nullptr;
  if (ipiv) {
    g.ldc = n; g.strideC = tn;
    g.D = W; g.ldd = n; g.strideD = tn;
  } else {
    g.ldc = lda; g.strideC = strideA;
    g.D = A; g.ldd = lda; g.strideD = strideA;
  }
  g.batch = batch;
  rc = zgemm_launch(g, s);
  if (rc) return rc;
  if (ipiv) {
    rc = ztimes_perm_launch(n, W, tn, A, lda, strideA, ipiv, strideP, pinv, batch, s);
    if (rc) return rc;
  }
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

}  // namespace negf
