// Memory-bound helpers on complex128 blocks (batched, column-major):
// adjoint (be_ctranspose!, src/backend_cpu.jl:58-68), 2-D copy
// (be_copy_in_hw!), fill / identity (be_fill!, be_identity, be_zero_array),
// scaling, and the trace reduction of the transmission.
#include "negf_internal.hpp"
#include <climits>
#include "sweeps.hpp"

namespace negf {

namespace {

// B = A^H through a 32x32 LDS tile (coalesced reads and writes).
__global__ __launch_bounds__(256) void adjoint_kernel(int m, int n, const zc* __restrict__ A,
                                                      long long lda, long long strideA,
                                                      zc* __restrict__ B, long long ldb,
                                                      long long strideB) {
  __shared__ zc tile[32][33];
  const int inst = blockIdx.z;
  const zc* Ai = A + inst * strideA;
  zc* Bi = B + inst * strideB;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;  // tile of A: rows i0.., cols j0..
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int jj = ty; jj < 32; jj += 8) {
    const int i = i0 + tx, j = j0 + jj;
    if (i < m && j < n) tile[jj][tx] = Ai[i + (long long)j * lda];
  }
  __syncthreads();
  // B is n x m: B[j, i] = conj(A[i, j])
  for (int ii = ty; ii < 32; ii += 8) {
    const int j = j0 + tx, i = i0 + ii;
    if (i < m && j < n) Bi[j + (long long)i * ldb] = zconj(tile[tx][ii]);
  }
}

__global__ void copy2d_kernel(int m, int n, const zc* __restrict__ A, long long lda,
                              long long strideA, zc* __restrict__ B, long long ldb,
                              long long strideB) {
  const int inst = blockIdx.z;
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j0 = blockIdx.y * 16 + (threadIdx.x >> 6) * 4;
  if (i >= m) return;
  const zc* Ai = A + inst * strideA;
  zc* Bi = B + inst * strideB;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const int j = j0 + jj;
    if (j < n) Bi[i + (long long)j * ldb] = Ai[i + (long long)j * lda];
  }
}

__global__ void fill_kernel(int m, int n, zc value, zc dvalue, zc* __restrict__ A, long long lda,
                            long long strideA) {
  const int inst = blockIdx.z;
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j0 = blockIdx.y * 16 + (threadIdx.x >> 6) * 4;
  if (i >= m) return;
  zc* Ai = A + inst * strideA;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const int j = j0 + jj;
    if (j < n) Ai[i + (long long)j * lda] = (i == j) ? dvalue : value;
  }
}

// B = s A (accumulate: B += s A)
__global__ void scale_kernel(int m, int n, zc value, const zc* __restrict__ sdev, long long sstride,
                             const zc* A, long long lda, long long strideA, zc* B, long long ldb,
                             long long strideB, bool accumulate) {
  const int inst = blockIdx.z;
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j0 = blockIdx.y * 16 + (threadIdx.x >> 6) * 4;
  if (i >= m) return;
  const zc sc = sdev ? sdev[inst * sstride] : value;
  const zc* Ai = A + inst * strideA;
  zc* Bi = B + inst * strideB;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const int j = j0 + jj;
    if (j < n) {
      zc v = zmul(sc, Ai[i + (long long)j * lda]);
      if (accumulate) v = zadd(v, Bi[i + (long long)j * ldb]);
      Bi[i + (long long)j * ldb] = v;
    }
  }
}

// One workgroup per instance; fixed reduction order (deterministic).
__global__ __launch_bounds__(256) void dotc_re_kernel(int m, int n, const zc* __restrict__ P,
                                                      long long ldp, long long strideP,
                                                      const zc* __restrict__ Q, long long ldq,
                                                      long long strideQ, double* out) {
  __shared__ double red[256];
  const int inst = blockIdx.x;
  const zc* Pi = P + inst * strideP;
  const zc* Qi = Q + inst * strideQ;
  double acc = 0.0;
  const long long total = (long long)m * n;
  for (long long e = threadIdx.x; e < total; e += 256) {
    const int i = (int)(e % m), j = (int)(e / m);
    const zc p = Pi[i + j * ldp], q = Qi[i + j * ldq];
    acc += p.x * q.x + p.y * q.y;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[inst] = red[0];
}

inline int ok() { return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR; }

// grid: (chunks of the block, n1 * n2, n3)
__global__ void copy3d_kernel(long long count, int n1, zc alpha, const zc* __restrict__ src,
                              long long s1, long long s2, long long s3, zc* __restrict__ dst,
                              long long d1, long long d2, long long d3) {
  const int i12 = blockIdx.y;
  const long long i1 = i12 % n1, i2 = i12 / n1, i3 = blockIdx.z;
  const zc* sp = src + i1 * s1 + i2 * s2 + i3 * s3;
  zc* dp = dst + i1 * d1 + i2 * d2 + i3 * d3;
  const bool one = (alpha.x == 1.0 && alpha.y == 0.0);
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
       e += (long long)gridDim.x * blockDim.x)
    dp[e] = one ? sp[e] : zmul(alpha, sp[e]);
}

}  // namespace

int zcopy3d_launch(long long count, int n1, int n2, int n3, zc alpha, const zc* src, long long s1,
                   long long s2, long long s3, zc* dst, long long d1, long long d2, long long d3,
                   hipStream_t s) {
  if (count < 0 || n1 < 0 || n2 < 0 || n3 < 0) return Status::BAD_ARG;
  if (count == 0 || n1 == 0 || n2 == 0 || n3 == 0) return Status::OK;
  if ((long long)n1 * n2 > 65535 || n3 > 65535) return Status::TOO_LARGE;
  long long bx = (count + 255) / 256;
  if (bx > 64) bx = 64;
  hipLaunchKernelGGL(copy3d_kernel, dim3((unsigned)bx, n1 * n2, n3), dim3(256), 0, s, count, n1,
                     alpha, src, s1, s2, s3, dst, d1, d2, d3);
  return ok();
}

int zadjoint_launch(int m, int n, const zc* A, long long lda, long long strideA, zc* B,
                    long long ldb, long long strideB, int batch, hipStream_t s) {
  if (m < 0 || n < 0 || batch < 0) return Status::BAD_ARG;
  if (m == 0 || n == 0 || batch == 0) return Status::OK;
  if (batch > 65535) return Status::TOO_LARGE;
  hipLaunchKernelGGL(adjoint_kernel, dim3((m + 31) / 32, (n + 31) / 32, batch), dim3(256), 0, s, m,
                     n, A, lda, strideA, B, ldb, strideB);
  return ok();
}

int zcopy2d_launch(int m, int n, const zc* A, long long lda, long long strideA, zc* B,
                   long long ldb, long long strideB, int batch, hipStream_t s) {
  if (m < 0 || n < 0 || batch < 0) return Status::BAD_ARG;
  if (m == 0 || n == 0 || batch == 0) return Status::OK;
  if (batch > 65535) return Status::TOO_LARGE;
  hipLaunchKernelGGL(copy2d_kernel, dim3((m + 63) / 64, (n + 15) / 16, batch), dim3(256), 0, s, m,
                     n, A, lda, strideA, B, ldb, strideB);
  return ok();
}

int zfill_launch(int m, int n, zc value, zc diag_value, zc* A, long long lda, long long strideA,
                 int batch, hipStream_t s) {
  if (m < 0 || n < 0 || batch < 0) return Status::BAD_ARG;
  if (m == 0 || n == 0 || batch == 0) return Status::OK;
  if (batch > 65535) return Status::TOO_LARGE;
  hipLaunchKernelGGL(fill_kernel, dim3((m + 63) / 64, (n + 15) / 16, batch), dim3(256), 0, s, m, n,
                     value, diag_value, A, lda, strideA);
  return ok();
}

int zscale_launch(int m, int n, zc value, const zc* scale_dev, long long scale_stride,
                  const zc* A, long long lda, long long strideA, zc* B, long long ldb,
                  long long strideB, int batch, hipStream_t s) {
  if (m < 0 || n < 0 || batch < 0) return Status::BAD_ARG;
  if (m == 0 || n == 0 || batch == 0) return Status::OK;
  if (batch > 65535) return Status::TOO_LARGE;
  hipLaunchKernelGGL(scale_kernel, dim3((m + 63) / 64, (n + 15) / 16, batch), dim3(256), 0, s, m, n,
                     value, scale_dev, scale_stride, A, lda, strideA, B, ldb, strideB, false);
  return ok();
}

int zaxpy_launch(int m, int n, zc value, const zc* scale_dev, long long scale_stride,
                 const zc* A, long long lda, long long strideA, zc* B, long long ldb,
                 long long strideB, int batch, hipStream_t s) {
  if (m < 0 || n < 0 || batch < 0) return Status::BAD_ARG;
  if (m == 0 || n == 0 || batch == 0) return Status::OK;
  if (batch > 65535) return Status::TOO_LARGE;
  hipLaunchKernelGGL(scale_kernel, dim3((m + 63) / 64, (n + 15) / 16, batch), dim3(256), 0, s, m, n,
                     value, scale_dev, scale_stride, A, lda, strideA, B, ldb, strideB, true);
  return ok();
}

// M = z S - H - E over `count` contiguous elements per instance (the packed
// diag, upper or lower array of a block-tridiagonal batch), z = zv[instance];
// S, H, E may be shared (batch stride 0), E may be null.  Evaluated as
// ((z * S) - H) - E with the complex product rounded term by term (no FMA
// contraction), the order of `energVal * S - H - Se`.
__global__ __launch_bounds__(256) void assemble_kernel(long long count, const zc* __restrict__ zv,
                                                       const zc* __restrict__ S, long long sbs,
                                                       const zc* __restrict__ H, long long hbs,
                                                       const zc* __restrict__ E, long long ebs,
                                                       zc* __restrict__ M, long long mbs) {
#pragma clang fp contract(off)
  const long long inst = blockIdx.y;
  const zc z = zv[inst];
  const zc* Si = S + inst * sbs;
  const zc* Hi = H + inst * hbs;
  const zc* Ei = E ? E + inst * ebs : nullptr;
  zc* Mi = M + inst * mbs;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < count; i += (long long)gridDim.x * 256) {
    const zc sv = Si[i], hv = Hi[i];
    const double pr = z.x * sv.x, pi = z.y * sv.y, qr = z.x * sv.y, qi = z.y * sv.x;
    zc m = zmake((pr - pi) - hv.x, (qr + qi) - hv.y);
    if (Ei) {
      const zc ev = Ei[i];
      m.x = m.x - ev.x;
      m.y = m.y - ev.y;
    }
    Mi[i] = m;
  }
}

int zassemble_launch(long long count, const zc* z, const zc* S, long long sbs, const zc* H,
                     long long hbs, const zc* E, long long ebs, zc* M, long long mbs, int batch,
                     hipStream_t s) {
  if (count < 0 || batch < 0) return Status::BAD_ARG;
  if (count == 0 || batch == 0) return Status::OK;
  if (batch > 65535) return Status::TOO_LARGE;
  long long bx = (count + 255) / 256;
  if (bx > 4096) bx = 4096;
  hipLaunchKernelGGL(assemble_kernel, dim3((unsigned)bx, batch), dim3(256), 0, s, count, z, S, sbs, H,
                     hbs, E, ebs, M, mbs);
  return ok();
}

namespace {

// Dimensions and packed offset of block i of `kind` (see btd_pad_launch).
// i is uniform over the workgroup, so the prefix sum runs on scalar registers.
__device__ inline long long pad_block(const PadTable& t, int kind, int i, int* m, int* n) {
  long long o = 0;
  if (kind == 0) {
    for (int k = 0; k < i; ++k) o += (long long)t.b[k] * t.b[k];
    *m = *n = t.b[i];
  } else {
    for (int k = 0; k < i; ++k) o += (long long)t.b[k] * t.b[k + 1];
    *m = kind == 1 ? t.b[i] : t.b[i + 1];
    *n = kind == 1 ? t.b[i + 1] : t.b[i];
  }
  return o;
}

// grid: (chunks of a padded block, block i, instance)
__global__ __launch_bounds__(256) void btd_pad_kernel(PadTable t, int kind,
                                                      const zc* __restrict__ src, long long sbs,
                                                      zc* __restrict__ dst, long long dbs) {
  int m, n;
  const int i = blockIdx.y;
  const long long off = pad_block(t, kind, i, &m, &n);
  const int B = t.bmax, BB = t.bmax * t.bmax;
  const zc* sp = src + blockIdx.z * sbs + off;
  zc* dp = dst + blockIdx.z * dbs + (long long)i * BB;
  for (int p = blockIdx.x * 256 + threadIdx.x; p < BB; p += gridDim.x * 256) {
    const int r = p % B, c = p / B;
    zc v = zmake(0.0, 0.0);
    if (r < m && c < n)   // kind 3: the upper block i is n x m, read transposed
      v = kind == 3 ? zconj(sp[c + (long long)r * n]) : sp[r + (long long)c * m];
    else if (kind == 0 && r == c)
      v = zmake(1.0, 0.0);
    dp[p] = v;
  }
}

__global__ __launch_bounds__(256) void btd_unpad_kernel(PadTable t, int kind,
                                                        const zc* __restrict__ src, long long sbs,
                                                        zc* __restrict__ dst, long long dbs) {
  int m, n;
  const int i = blockIdx.y;
  const long long off = pad_block(t, kind, i, &m, &n);
  const int B = t.bmax;
  const zc* sp = src + blockIdx.z * sbs + (long long)i * B * B;
  zc* dp = dst + blockIdx.z * dbs + off;
  for (int p = blockIdx.x * 256 + threadIdx.x; p < m * n; p += gridDim.x * 256) {
    const int r = p % m, c = p / m;
    dp[p] = sp[r + (long long)c * B];
  }
}

int pad_grid(const PadTable& t, int kind, int batch, dim3* g) {
  if (t.N < 1 || t.N > kPadMaxBlocks || t.bmax < 1 || batch < 0 || kind < 0 || kind > 3)
    return Status::BAD_ARG;
  if (batch > 65535) return Status::TOO_LARGE;
  const int nblk = kind == 0 ? t.N : t.N - 1;
  long long bx = ((long long)t.bmax * t.bmax + 255) / 256;
  if (bx > 64) bx = 64;
  *g = dim3((unsigned)bx, (unsigned)nblk, (unsigned)batch);
  return Status::OK;
}

}  // namespace

int btd_pad_launch(const PadTable& t, int kind, const zc* packed, long long sbs, zc* padded,
                   long long dbs, int batch, hipStream_t s) {
  dim3 g;
  int rc = pad_grid(t, kind, batch, &g);
  if (rc || g.y == 0 || batch == 0) return rc;
  hipLaunchKernelGGL(btd_pad_kernel, g, dim3(256), 0, s, t, kind, packed, sbs, padded, dbs);
  return ok();
}

int btd_unpad_launch(const PadTable& t, int kind, const zc* padded, long long sbs, zc* packed,
                     long long dbs, int batch, hipStream_t s) {
  dim3 g;
  if (kind == 3) return Status::BAD_ARG;
  int rc = pad_grid(t, kind, batch, &g);
  if (rc || g.y == 0 || batch == 0) return rc;
  hipLaunchKernelGGL(btd_unpad_kernel, g, dim3(256), 0, s, t, kind, padded, sbs, packed, dbs);
  return ok();
}

int zdotc_re_launch(int m, int n, const zc* P, long long ldp, long long strideP, const zc* Q,
                    long long ldq, long long strideQ, double* out, int batch, hipStream_t s) {
  if (m < 0 || n < 0 || batch < 0) return Status::BAD_ARG;
  if (batch == 0) return Status::OK;
  hipLaunchKernelGGL(dotc_re_kernel, dim3(batch), dim3(256), 0, s, m, n, P, ldp, strideP, Q, ldq,
                     strideQ, out);
  return ok();
}

namespace {

// Pivoted Cholesky of a Hermitian positive semidefinite matrix (diagonal
// pivoting, as LAPACK's zpstrf): P^T A P = L L^H, stopped where the largest
// remaining diagonal of the Schur complement falls to n * 2^-52 times the
// largest diagonal of A (the rank r); F = P L (zero beyond column r), so
// F F^H = A + O(u ||A||).  One 1024-thread workgroup per instance, A
// factored in place in full storage (both triangles of the Schur complement
// kept, so a pivot swap is a plain row and column exchange); perm: n ints
// per instance.  bad[instance] = 1 when A is not Hermitian positive
// semidefinite to 2^-26 of its largest diagonal (|a_ij - conj(a_ji)|, or an
// entry of the Schur complement left at the rank, above it; NaN included),
// else 0: F F^H then does not stand for A.
constexpr int kCholThreads = 1024;
constexpr int kCholMaxN = 2048;

__global__ __launch_bounds__(kCholThreads) void zpchol_kernel(int n, zc* __restrict__ A,
                                                              zc* __restrict__ F,
                                                              int* __restrict__ perm_ws,
                                                              int* __restrict__ bad) {
  constexpr int NT = kCholThreads, NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  zc* a = A + (long long)blockIdx.x * n * n;
  zc* f = F + (long long)blockIdx.x * n * n;
  int* perm = perm_ws + (long long)blockIdx.x * n;
  __shared__ double rv[NW];
  __shared__ int ri[NW];
  __shared__ double dmax0;
  __shared__ int flag;
  __shared__ zc lk[kCholMaxN];   // the column of L of the current step
  for (int i = tid; i < n; i += NT) perm[i] = i;
  if (tid == 0) flag = 0;
  // the largest |a_ij| over the trailing block [k0, n)^2 (Hermitian part
  // when herm), a NaN anywhere giving NaN
  // (one column per wave, rows over the lanes)
  auto max_entry = [&](int k0, bool herm) {
    double v = 0.0;
    for (int j = k0 + wave; j < n; j += NW)
      for (int i = k0 + lane; i < n; i += 64) {
        zc x = a[i + (long long)j * n];
        if (herm) {
          const zc y = a[j + (long long)i * n];
          x.x -= y.x;
          x.y += y.y;
        }
        const double ax = fabs(x.x) > fabs(x.y) ? fabs(x.x) : fabs(x.y);
        v = (ax > v || ax != ax) ? ax : v;
      }
    return v;
  };
  // the largest diagonal from position k on, lowest index on ties
  auto argmax_diag = [&](int k, double* v_out, int* i_out) {
    double bv = -1.0;
    int bi = INT_MAX;
    for (int i = k + tid; i < n; i += NT) {
      const double d = a[i + (long long)i * n].x;
      if (d > bv) { bv = d; bi = i; }
    }
    const double wmax = wave_max(bv);
    const unsigned long long m = __ballot(bv == wmax && bi != INT_MAX);
    int wi = INT_MAX;
    if (m) wi = __shfl(bi, __ffsll(m) - 1);
    __syncthreads();   // the slots are free
    if (lane == 0) { rv[wave] = wmax; ri[wave] = wi; }
    __syncthreads();
    double gv = rv[0];
    for (int w = 1; w < NW; ++w) gv = fmax(gv, rv[w]);
    int gi = INT_MAX;
    for (int w = 0; w < NW; ++w) gi = min(gi, rv[w] == gv ? ri[w] : INT_MAX);
    *v_out = gv;
    *i_out = gi;
  };
  {
    double v;
    int i;
    argmax_diag(0, &v, &i);
    if (tid == 0) dmax0 = v;
    __syncthreads();
  }
  const double tol = (double)n * 0x1p-52 * dmax0;
  const double bound = 0x1p-26 * dmax0;
  if (!(max_entry(0, true) <= bound)) flag = 1;   // benign race: every writer stores 1
  int r = n;
  for (int k = 0; k < n; ++k) {
    double dv;
    int p;
    argmax_diag(k, &dv, &p);
    if (!(dv > tol)) {   // the rest of the Schur complement is numerically zero
      r = k;
      break;
    }
    if (p != k) {
      for (int c = tid; c < n; c += NT) {   // rows k and p
        const zc x = a[k + (long long)c * n];
        a[k + (long long)c * n] = a[p + (long long)c * n];
        a[p + (long long)c * n] = x;
      }
      __syncthreads();
      for (int rr = tid; rr < n; rr += NT) {   // columns k and p
        const zc x = a[rr + (long long)k * n];
        a[rr + (long long)k * n] = a[rr + (long long)p * n];
        a[rr + (long long)p * n] = x;
      }
      if (tid == 0) {
        const int t = perm[k];
        perm[k] = perm[p];
        perm[p] = t;
      }
      __syncthreads();
    }
    const double sq = sqrt(dv), rs = 1.0 / sq;
    for (int i = k + 1 + tid; i < n; i += NT) {
      const zc x = a[i + (long long)k * n];
      const zc y = zmake(x.x * rs, x.y * rs);
      a[i + (long long)k * n] = y;
      lk[i] = y;
    }
    if (tid == 0) a[k + (long long)k * n] = zmake(sq, 0.0);
    __syncthreads();
    // Schur complement of the trailing block (both triangles): one column
    // per wave, rows over the lanes, four rows per lane loaded before any
    // is stored (the column of L from LDS)
    for (int j = k + 1 + wave; j < n; j += NW) {
      const zc lj = lk[j];
      zc* aj = a + (long long)j * n;
      auto upd = [&](zc x, zc li) {
        x.x -= li.x * lj.x + li.y * lj.y;   // li conj(lj)
        x.y -= li.y * lj.x - li.x * lj.y;
        return x;
      };
      int i = k + 1 + lane;
      for (; i + 192 < n; i += 256) {
        const zc x0 = aj[i], x1 = aj[i + 64], x2 = aj[i + 128], x3 = aj[i + 192];
        aj[i] = upd(x0, lk[i]);
        aj[i + 64] = upd(x1, lk[i + 64]);
        aj[i + 128] = upd(x2, lk[i + 128]);
        aj[i + 192] = upd(x3, lk[i + 192]);
      }
      for (; i < n; i += 64) aj[i] = upd(aj[i], lk[i]);
    }
    __syncthreads();
  }
  __syncthreads();
  // what the rank leaves must be numerically zero (a PSD Schur complement
  // with every diagonal <= tol has every entry <= tol)
  if (r < n && !(max_entry(r, false) <= bound)) flag = 1;
  // F = P L: row perm[i] of F is row i of L
  for (int j = wave; j < n; j += NW)
    for (int i = lane; i < n; i += 64)
      f[perm[i] + (long long)j * n] =
          (j < r && i >= j) ? a[i + (long long)j * n] : zmake(0.0, 0.0);
  __syncthreads();
  if (tid == 0) bad[blockIdx.x] = flag;
}

}  // namespace

int zpchol_launch(int n, zc* A, zc* F, int* perm, int* bad, int batch, hipStream_t s) {
  if (n <= 0 || n > kCholMaxN || batch <= 0) return Status::BAD_ARG;
  hipLaunchKernelGGL(zpchol_kernel, dim3(batch), dim3(kCholThreads), 0, s, n, A, F, perm, bad);
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

}  // namespace negf
