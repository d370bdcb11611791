// Internal declarations shared by the gfx950 kernels and the C-ABI layer.
//
// Storage convention (same as the reference's Matrix{ComplexF64} blocks):
// column-major, interleaved complex128 (re, im) = HIP double2 (x = re, y = im).
// A "batch" of instances is addressed as base + idx * stride (stride 0
// broadcasts one block to every instance, used for energy-independent blocks).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

namespace negf {

typedef double2 zc;

// op codes for gemm/trsm operands, as the BLAS characters 'N', 'T', 'C'
enum Op : int { OP_N = 0, OP_T = 1, OP_C = 2 };

__host__ __device__ inline zc zmake(double r, double i) { zc z; z.x = r; z.y = i; return z; }
__host__ __device__ inline zc zadd(zc a, zc b) { return zmake(a.x + b.x, a.y + b.y); }
__host__ __device__ inline zc zsub(zc a, zc b) { return zmake(a.x - b.x, a.y - b.y); }
__host__ __device__ inline zc zmul(zc a, zc b) {
  return zmake(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline zc zconj(zc a) { return zmake(a.x, -a.y); }
__host__ __device__ inline double zabs1(zc a) { return fabs(a.x) + fabs(a.y); }

// Maximum of a double over the 64 lanes of a wave, returned to every lane:
// DPP row shifts by 1, 2, 4, 8 within rows of 16 lanes, then row broadcasts
// of lanes 15 and 31 (lanes without a source keep their own value), and the
// total read from lane 63.  No LDS traffic, a few cycles per step.
__device__ __forceinline__ double wave_max(double x) {
  union {
    double d;
    int i[2];
  } u, v;
#define NEGF_DPP_MAX(CTRL, RMASK)                                                  \
  u.d = x;                                                                         \
  v.i[0] = __builtin_amdgcn_update_dpp(u.i[0], u.i[0], CTRL, RMASK, 0xf, false);  \
  v.i[1] = __builtin_amdgcn_update_dpp(u.i[1], u.i[1], CTRL, RMASK, 0xf, false);  \
  x = fmax(x, v.d);
  NEGF_DPP_MAX(0x111, 0xf)   // row_shr:1
  NEGF_DPP_MAX(0x112, 0xf)   // row_shr:2
  NEGF_DPP_MAX(0x114, 0xf)   // row_shr:4
  NEGF_DPP_MAX(0x118, 0xf)   // row_shr:8
  NEGF_DPP_MAX(0x142, 0xa)   // row_bcast:15 into rows 1, 3
  NEGF_DPP_MAX(0x143, 0xc)   // row_bcast:31 into rows 2, 3
#undef NEGF_DPP_MAX
  u.d = x;
  u.i[0] = __builtin_amdgcn_readlane(u.i[0], 63);
  u.i[1] = __builtin_amdgcn_readlane(u.i[1], 63);
  return u.d;
}
// 1/z with Smith's scaling (R. L. Smith, "Algorithm 116: Complex division",
// CACM 5(8), 1962), which avoids overflow/underflow of |z|^2.
__host__ __device__ inline zc zrcp(zc z) {
  if (fabs(z.y) <= fabs(z.x)) {
    double r = z.y / z.x, d = z.x + z.y * r;
    return zmake(1.0 / d, -r / d);
  } else {
    double r = z.x / z.y, d = z.y + z.x * r;
    return zmake(r / d, -1.0 / d);
  }
}

struct Status {
  static constexpr int OK = 0;
  static constexpr int BAD_ARG = -1;
  static constexpr int TOO_LARGE = -2;
  static constexpr int HIP_ERROR = -3;
  static constexpr int NO_MEMORY = -4;
};

// ----------------------------------------------------------------------------
// Batched ZGEMM on FP64 MFMA:  D = alpha * op(A) op(B) + beta * C  (+ shift I)
// C may alias D (in-place update); C is not read when beta == 0.
// `shift` (optional, may be null) adds shift[idx * shift_stride] to every
// diagonal element of instance idx, which builds A(E) = zI - ... in the
// epilogue of the Schur-complement GEMM.
// Instances: idx in [0, batch).  With inner == 0 operand X of instance idx
// is X + idx * strideX.  With inner > 0 the batch is 2-D: X + (idx % inner) *
// strideX + (idx / inner) * outerX (e.g. layers within energies).
struct GemmArgs {
  int opA, opB;
  int m, n, k;
  zc alpha, beta;
  const zc* A; long long lda, strideA;
  const zc* B; long long ldb, strideB;
  const zc* C; long long ldc, strideC;
  zc* D; long long ldd, strideD;
  const zc* shift; long long shift_stride;
  // optional per-instance factor of the product: alpha * scale[idx * scale_stride]
  const zc* scale; long long scale_stride;
  // 1 / -1: the caller guarantees the result is Hermitian / anti-Hermitian
  // (m == n); kernels may then compute the tiles on and below the diagonal
  // only and write the rest as +-conj of its mirror.  0: every element.
  int herm;
  int batch;
  int inner;
  long long outerA, outerB, outerC, outerD;
  // arithmetic of the product (operands and results stay complex128 in
  // memory): 0 = FP64 MFMA; 1 = FP32 MFMA with FP32 accumulation (operands
  // rounded to FP32, the 4-real-product complex form) where the launch takes
  // the LDS-DMA kernel, FP64 otherwise; 2 = integer-emulated FP64 (Ozaki
  // scheme II on int8 MFMA, zgemm_emu.hip); 3 = FP64 with three real
  // products per complex product (Gauss) where the launch takes the LDS-DMA
  // kernel, the four-product FP64 form otherwise
  int prec;
  // prec = 2: device workspace for the residue planes (zgemm_emu_work_bytes)
  // and the number of moduli (0: 13)
  void* work;
  long long work_bytes;
  int nmod;
};
int zgemm_launch(const GemmArgs& g, hipStream_t s);
// The integer-emulated product (prec = 2, k > 0), one problem.
int zgemm_emu_launch(const GemmArgs& g, hipStream_t s);
// Workspace of zgemm_emu_launch for an m x n x k product whose A / B hold
// batchA / batchB distinct instances (1 for a broadcast operand).
size_t zgemm_emu_work_bytes(int m, int n, int k, int batchA, int batchB, int nmod);
int zgemm_emu_default_moduli();
// beta (bits kept per row / column scale) for nmod moduli at this k.
int zgemm_emu_params(int nmod, int k, int* beta);
// Tile variant of the k > 32 GEMMs (0..8, see zgemm.hip), process-wide.
int zgemm_set_variant(int v);

// Several independent GEMMs of one shape in one launch (fills the GPU when
// each problem alone is a handful of tiles).  Problems run concurrently; a
// problem whose D is another's A, B, C or D (same base pointer) is launched
// after it instead.
int zgemm_group_launch(const GemmArgs* g, int count, hipStream_t s);

// ----------------------------------------------------------------------------
// Batched explicit inverse by blocked Gauss-Jordan with partial pivoting.
// Ainv[idx] = inv(A[idx]).  `X` is the in-place work matrix (n x n, ld n,
// one per instance) and may be the same buffer as A when A may be destroyed.
// Second stream of the split-step inverse (the strip updates beside the
// factoring) with its fork / join events.  All null: the library's own slot
// for the calling stream (created on first use, kept until
// zinv_aux_release_all).  A caller that captures launches into a graph owns
// one for the graph's lifetime instead (zinv_aux_create / _destroy).
struct InvAux {
  hipStream_t aux;
  hipEvent_t fork, join;
};
int zinv_aux_create(InvAux* a);
void zinv_aux_destroy(InvAux* a);
// Destroys the library's per-stream slots (the device is synchronized first).
void zinv_aux_release_all();

struct InvWork {
  zc* X;        // n*n*batch
  zc* W;        // 3*zinv_w_elems(n)*batch (W of panels J, J+1, J+2)
  zc* R;        // zinv_w_elems(n)*batch: a panel's chained pivot rows (variant 6)
  int* piv;     // n*batch
  int* used;    // n*batch
  int nmax;     // the n and batch the buffers were carved for: zinv_launch
  int batchmax; // rejects larger problems
  InvAux aux;   // null members: the per-calling-stream slot
  int inplace_only;  // 1: never the split steps (no second stream is used)
  double* fro;       // null, or fro[inst * ceil(n / 64) + c] = sum of |x|^2 over
                     // columns 64 c .. 64 c + 63 of instance inst's inverse
};
size_t zinv_work_bytes(int n, int batch);
// Elements of one W buffer per instance that serve every block size <= n.
size_t zinv_w_elems(int n);
// Schedule of the inverse (0: automatic, 1 / 2: in place with panels of 1 /
// 2 strips, 3 / 4 / 5: split steps with panels of 1 / 2 / 4 strips, 6: split
// steps with panels of 4 strips whose bulk update is a batched GEMM),
// process-wide.
int zinv_set_variant(int v);
void zinv_work_carve(void* base, int n, int batch, InvWork* w);
// The schedule zinv_launch runs for (n, batch, variant): strip width, strips
// per panel and W buffers (each n * strip * panel per instance).  Returns 1
// for split steps, 0 in place, < 0 for invalid arguments.
int zinv_schedule(int n, int batch, int variant, int* strip, int* panel, int* nbuf);
// A: input (not modified unless A == X), Ainv: output; info[idx] = 0 or the
// 1-based column of the first exactly-zero pivot.
int zinv_launch(int n, const zc* A, long long lda, long long strideA,
                zc* Ainv, long long ldinv, long long strideInv, int batch,
                const InvWork& w, int* info, hipStream_t s);

// ----------------------------------------------------------------------------
// Batched LU with partial pivoting (LAPACK zgetrf layout, 1-based ipiv),
// triangular solves and inversion from the factors (per-op drop-in).
// pivot = false: no row interchanges (ipiv = identity), the LU the block
// elimination of the selected inverse needs.
int zgetrf_launch(int n, zc* A, long long lda, long long strideA, int* ipiv,
                  long long strideP, int* info, int batch, hipStream_t s, bool pivot = true);
int ztrsm_launch(char side, char uplo, char trans, char diag, int m, int n, zc alpha,
                 const zc* A, long long lda, long long strideA, zc* B, long long ldb,
                 long long strideB, int batch, hipStream_t s);
int zlaswp_launch(int ncols, zc* B, long long ldb, long long strideB, const int* ipiv,
                  long long strideP, int k1, int k2, int incx, int batch, hipStream_t s);
int zgetrs_launch(char trans, int n, int nrhs, const zc* A, long long lda, long long strideA,
                  const int* ipiv, long long strideP, zc* B, long long ldb, long long strideB,
                  int batch, hipStream_t s);
// X = tri(A)^{-1} (uplo 'U' / 'L', diag 'N' / 'U'), out of place, X zero
// outside the triangle; work: ztrtri_work_bytes.
size_t ztrtri_work_bytes(int n, int batch);
int ztrtri_launch(char uplo, char diag, int n, const zc* A, long long lda, long long strideA,
                  zc* X, long long ldx, long long strideX, int batch, zc* work, hipStream_t s);
int zgetri_launch(int n, zc* A, long long lda, long long strideA, const int* ipiv,
                  long long strideP, int batch, zc* work, hipStream_t s);
// Largest n of zgetrf_launch (above 2048: two-level blocking, 512-thread panels).
constexpr int kLuMaxN = 16384;
// Explicit inverse of one n x n block for n beyond the Gauss-Jordan kernels
// (n <= kLuMaxN): A = P L U by zgetrf, inv(U) and inv(L) by the recursive
// triangular inverses, their product on the matrix cores, then the column
// interchanges of P.  work: zinv_large_work_bytes(n).  info as zgetrf's.
size_t zinv_large_work_bytes(int n);
int zinv_large_launch(int n, const zc* A, long long lda, zc* Ainv, long long ldinv, void* work,
                      int* info, hipStream_t s);
// XP = X P per instance for LAPACK's interchanges ipiv (1-based, rows swapped
// first to last: P A = L U): XP[:, c] = X[:, pinv[c]], pinv the inverse of
// the row order they give.  X is n x n with ld n, XP has ld ldxp; pinv: n
// ints per instance of scratch.  n <= kPermMaxN.
constexpr int kPermMaxN = 2048;
int ztimes_perm_launch(int n, const zc* X, long long sX, zc* XP, long long ldxp, long long sXP,
                       const int* ipiv, long long sP, int* pinv, int batch, hipStream_t s);
// getrs by the factors' explicit inverses: inv(L) P and inv(U) (recursive
// triangular inverses, column gather), then B <- inv(U) (inv(L) P B) for 'N',
// op(inv(L) P) (op(inv(U)) B) for 'T' / 'C': two GEMMs on the matrix cores
// instead of 2 n / 16 dependent launches of the blocked substitution.
// n <= kPermMaxN; work: zgetrs_inv_work_bytes.
size_t zgetrs_inv_work_bytes(int n, int nrhs, int batch);
int zgetrs_inv_launch(char trans, int n, int nrhs, const zc* A, long long lda, long long strideA,
                      const int* ipiv, long long strideP, zc* B, long long ldb, long long strideB,
                      int batch, void* work, hipStream_t s);
// trsm by the triangle's explicit inverse: B = alpha op(inv(A)) B (left) or
// alpha B op(inv(A)) (right), one GEMM after the recursive triangular
// inverse; work: ztrsm_inv_work_bytes(na = m or n, m, n, batch).
size_t ztrsm_inv_work_bytes(int na, int m, int n, int batch);
int ztrsm_inv_launch(char side, char uplo, char trans, char diag, int m, int n, zc alpha,
                     const zc* A, long long lda, long long strideA, zc* B, long long ldb,
                     long long strideB, int batch, void* work, hipStream_t s);
// getri by the factors' explicit inverses: inv(A) = inv(U) inv(L) P (the
// product into W, n x n per instance, then the column interchanges into A;
// straight into A when ipiv is null: no interchanges).  n <= kPermMaxN;
// work: zgetri_inv_work_bytes.
size_t zgetri_inv_work_bytes(int n, int batch);
int zgetri_inv_launch(int n, zc* A, long long lda, long long strideA, const int* ipiv,
                      long long strideP, int batch, zc* W, void* work, hipStream_t s);

// ----------------------------------------------------------------------------
// Elementwise helpers (all batched, column-major).
int zadjoint_launch(int m, int n, const zc* A, long long lda, long long strideA, zc* B,
                    long long ldb, long long strideB, int batch, hipStream_t s);
int zcopy2d_launch(int m, int n, const zc* A, long long lda, long long strideA, zc* B,
                   long long ldb, long long strideB, int batch, hipStream_t s);
int zfill_launch(int m, int n, zc value, zc diag_value, zc* A, long long lda,
                 long long strideA, int batch, hipStream_t s);
// out[idx] = Re sum_ij P[i,j] conj(Q[i,j]) over an m x n block.
int zdotc_re_launch(int m, int n, const zc* P, long long ldp, long long strideP, const zc* Q,
                    long long ldq, long long strideQ, double* out, int batch, hipStream_t s);
// B = s * A with s = scale_dev[idx * scale_stride] when scale_dev is given
// (device array, stride 0 broadcasts), else the host value `value`.
// Contiguous blocks of `count` elements over a 3-D index (i1 < n1, i2 < n2,
// i3 < n3): dst[i1 d1 + i2 d2 + i3 d3] = alpha * src[i1 s1 + i2 s2 + i3 s3].
int zcopy3d_launch(long long count, int n1, int n2, int n3, zc alpha, const zc* src,
                   long long s1, long long s2, long long s3, zc* dst, long long d1, long long d2,
                   long long d3, hipStream_t s);
// M[idx][i] = z[idx] S[idx][i] - H[idx][i] - E[idx][i] over `count` elements
// per instance (batch strides may be 0; E may be null).
int zassemble_launch(long long count, const zc* z, const zc* S, long long sbs, const zc* H,
                     long long hbs, const zc* E, long long ebs, zc* M, long long mbs, int batch,
                     hipStream_t s);
// Pivoted Cholesky factor of Hermitian positive semidefinite matrices:
// F F^H = A (F = P L, zero beyond the numerical rank), n x n, ld n, `batch`
// instances back to back; A is overwritten; perm: n ints per instance;
// bad[instance] = 1 where A is not Hermitian positive semidefinite (to 2^-26
// of its largest diagonal), else 0.
int zpchol_launch(int n, zc* A, zc* F, int* perm, int* bad, int batch, hipStream_t s);
int zscale_launch(int m, int n, zc value, const zc* scale_dev, long long scale_stride,
                  const zc* A, long long lda, long long strideA, zc* B, long long ldb,
                  long long strideB, int batch, hipStream_t s);

// Tuned variants for a shape (0: none tuned; the launcher's thresholds apply).
int tune_gemm_variant(int m, int n, int k, long long problems);
int tune_inv_variant(int n, int batch);
int tune_getrs_route(int n, int nrhs);   // 0 none, 1 substitution, 2 inverses
int tune_trsm_route(int na, int nrhs);    // the same for trsm (na: the triangle's order)

// Uniform view of a packed block-tridiagonal matrix whose block sizes differ:
// block (i, j) of b_i x b_j becomes the leading part of a bmax x bmax block,
// the rest 0 except 1 on the diagonal of a diagonal block.  Up to a
// permutation the padded matrix is A (+) I, so the band of its inverse holds
// the band of A^{-1} in the same leading parts, and the padded rows never
// mix with the original ones.  The sizes travel as kernel arguments.
constexpr int kPadMaxBlocks = 1024;
struct PadTable {
  int N, bmax;
  unsigned short b[kPadMaxBlocks];
};
// kind: 0 diagonal blocks, 1 upper (i, i+1), 2 lower (i+1, i), 3 lower from
// the upper blocks of Hermitian storage (A_{i+1,i} = A_{i,i+1}^H).  Packed
// side: blocks back to back, tight ld, batch stride sbs (0 broadcasts);
// padded side: block i at i * bmax^2, batch stride dbs.
int btd_pad_launch(const PadTable& t, int kind, const zc* packed, long long sbs, zc* padded,
                   long long dbs, int batch, hipStream_t s);
// kind 0, 1, 2: the leading parts of the padded blocks back into packed form.
int btd_unpad_launch(const PadTable& t, int kind, const zc* padded, long long sbs, zc* packed,
                     long long dbs, int batch, hipStream_t s);

}  // namespace negf
