// Batched complex128 GEMM on gfx950 FP64 MFMA (v_mfma_f64_16x16x4_f64).
//
//   D[idx] = alpha * op(A[idx]) * op(B[idx]) + beta * C[idx]  (+ shift[idx] * I)
//
// Complex products as one real GEMM over a doubled K' = 2K ("K'-space"):
//   Re D = sum_k' A-[i,k'] B'[k',j],  A-[i,2k] = Re a, A-[i,2k+1] = -Im a,
//                                     B'[2k,j] = Re b, B'[2k+1,j] = Im b
//   Im D = sum_k' A'[i,k'] B''[k',j], A'[i,2k] = Re a, A'[i,2k+1] =  Im a,
//                                     B''[2k,j] = Im b, B''[2k+1,j] = Re b
// so the interleaved (re, im) storage maps straight onto the MFMA K lanes:
// one LDS row per K' index, the im-rows of B read with k' ^ 1, and the
// A- sign flip a per-lane constant.  Two MFMAs per 16x16 output block and two
// complex K steps.
//
// Tiling: a 256-thread workgroup = 2 x 2 waves; each wave owns FM x FN MFMA
// blocks of 16 x 16 (re and im accumulators), so the workgroup tile is
// (32 FM) x (32 FN) complex.  K is staged 16 complex at a time through LDS
// (K'-major rows), the next stage prefetched into registers while the MFMAs
// run; with DB the LDS is double-buffered (one barrier per stage).
// Staging is bank-conflict free: every wave-wide LDS store writes 64
// consecutive doubles of one row (operands contiguous along K are loaded 4
// complex per lane and scattered to 4 rows), and rows are padded to 16 mod 32
// doubles so the MFMA operand reads of two rows fall in disjoint banks.
// op(A) / op(B) transposition and conjugation happen in the staging, so the
// MFMA loop is identical for all nine N/T/C combinations.
// Launches whose tiles all lie inside the output (the sweeps' b x b blocks
// with b a multiple of 64) take zgemm_lds_kernel instead: the same tile and
// MFMA loop, the operands loaded straight into LDS by the buffer unit
// (no register staging, one barrier per 8-deep stage); launches too small to
// fill the chip with 64 x 64 tiles take its 32 x 32 form.
//
// Launch: one flat grid over (problem, instance, tile).  Workgroups are
// dispatched round-robin over the 8 XCDs; the index is remapped so that the
// tiles of one instance run on one XCD and share its L2.
#include "negf_internal.hpp"

#include <type_traits>

namespace negf {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int TK = 16;            // complex K per stage
constexpr int KP = 2 * TK;        // K'-space rows per stage
constexpr int PAD = 16;           // row padding (doubles): rows 16 mod 32 apart
constexpr int MAXG = 4;           // problems per grouped launch

struct GemmGroup {
  GemmArgs p[MAXG];
  int count;
  int tiles_m, tiles_n;
  long long tiles_per_problem;    // tiles_m * tiles_n * batch
  long long total;                // tiles_per_problem * count
  long long grid_padded;          // multiple of 8
  int tri;                        // tiles on and below the diagonal only (herm)
};

// Epilogue of a wave's FM x FN MFMA blocks whose top-left corner is (ib, jb):
// D = alpha (scale) acc + beta C (+ shift on the diagonal), rows < m and
// columns < n.
// Output layout of the 16x16 f64 MFMA on gfx950 (measured): lane l, element
// e holds row l/16 + 4*e, column l % 16.
// With mirror (a tile strictly below the diagonal of a herm product) each
// element (i, j) is also written at (j, i) as herm * conj.
// The 16x16 f32 MFMA (the FP32 form) holds row 4 * (l / 16) + e instead
// (F32ROWS).
template <int FM, int FN, typename V = dbl4, bool F32ROWS = false>
__device__ __forceinline__ void store_tile(const GemmArgs& g, int inst, long long i_in,
                                           long long i_out, int ib, int jb, int lane,
                                           const V (&acc_re)[FM][FN],
                                           const V (&acc_im)[FM][FN], bool mirror = false) {
  const int kq = lane >> 4, li = lane & 15;
  zc* __restrict__ D = g.D + i_in * g.strideD + i_out * g.outerD;
  const zc* __restrict__ C = g.C ? g.C + i_in * g.strideC + i_out * g.outerC : nullptr;
  const bool use_c = (g.beta.x != 0.0 || g.beta.y != 0.0) && C != nullptr;
  zc sh = zmake(0.0, 0.0);
  const bool use_shift = g.shift != nullptr;
  if (use_shift) sh = g.shift[inst * g.shift_stride];
  const zc alpha = g.scale ? zmul(g.alpha, g.scale[inst * g.scale_stride]) : g.alpha;
#pragma unroll
  for (int r = 0; r < FM; ++r)
#pragma unroll
    for (int c = 0; c < FN; ++c) {
      const int j = jb + c * 16 + li;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = ib + r * 16 + (F32ROWS ? 4 * kq + e : kq + 4 * e);
        if (i < g.m && j < g.n) {
          const double vr = (double)acc_re[r][c][e], vi = (double)acc_im[r][c][e];
          zc out = zmake(alpha.x * vr - alpha.y * vi, alpha.x * vi + alpha.y * vr);
          if (use_c) out = zadd(out, zmul(g.beta, C[i + j * g.ldc]));
          if (use_shift && i == j) out = zadd(out, sh);
          D[i + j * g.ldd] = out;
          if (mirror)
            D[j + i * g.ldd] = g.herm > 0 ? zmake(out.x, -out.y) : zmake(-out.x, out.y);
        }
      }
    }
}

// Tile configuration (FM, FN: MFMA blocks per wave; DB: double-buffered LDS).
// SMALLK = 1 tags the instantiation used for k <= 32 (the rank-NB updates of
// the Gauss-Jordan / LU panels and k = 0 shifts): the same code, a separate
// symbol, so profiles separate the memory-bound updates from the full GEMMs.
template <int OPA, int OPB, int FM, int FN, bool DB, int SMALLK,
          int OCC = (FM * FN <= 4 && !DB) ? 2 : 1>
__global__ __launch_bounds__(256, OCC) void zgemm_mfma_kernel(
    const GemmGroup G) {
  constexpr int TM = 32 * FM, TN = 32 * FN;
  constexpr int SA = TM + PAD, SB = TN + PAD;
  constexpr int RA = TM / 64, RB = TN / 64;     // 64-wide row groups per thread
  constexpr int NBUF = DB ? 2 : 1;

  // XCD-aware remap of the flat workgroup index.
  const long long bid = blockIdx.x;
  const long long per = G.grid_padded >> 3;
  const long long logical = (bid & 7) * per + (bid >> 3);
  if (logical >= G.total) return;
  const int pidx = (int)(logical / G.tiles_per_problem);
  const long long rem = logical - (long long)pidx * G.tiles_per_problem;
  const int tiles_inst = G.tiles_m * G.tiles_n;
  const int inst = (int)(rem / tiles_inst);
  const int tile = (int)(rem - (long long)inst * tiles_inst);
  const int tm = tile % G.tiles_m, tn = tile / G.tiles_m;
  const GemmArgs& g = G.p[pidx];
  const int m = g.m, n = g.n, k = g.k;
  const int i0 = tm * TM, j0 = tn * TN;

  const long long i_in = g.inner > 0 ? inst % g.inner : inst;
  const long long i_out = g.inner > 0 ? inst / g.inner : 0;
  const zc* __restrict__ A = g.A + i_in * g.strideA + i_out * g.outerA;
  const zc* __restrict__ B = g.B + i_in * g.strideB + i_out * g.outerB;
  const long long lda = g.lda, ldb = g.ldb;

  __shared__ double As[NBUF][KP][SA];
  __shared__ double Bs[NBUF][KP][SB];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int tl = tid & 63, tw = tid >> 6;   // loader: column/row in a 64-group, wave

  dbl4 acc_re[FM][FN], acc_im[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      acc_re[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
      acc_im[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
    }

  // prefetch registers: per 64-group r and slot u (4 complex)
  zc ra[RA][4], rb[RB][4];

  // A: OP_N -> element (i = tl + 64r, k = tw + 4u); T/C -> (i = tl + 64r, k = 4tw + u)
  auto load_stage = [&](int k0) {
#pragma unroll
    for (int r = 0; r < RA; ++r)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = tl + 64 * r;
        const int kk = (OPA == OP_N) ? (tw + 4 * u) : (4 * tw + u);
        const int gi = i0 + i, gk = k0 + kk;
        zc v = zmake(0.0, 0.0);
        if (gi < m && gk < k) v = (OPA == OP_N) ? A[gi + gk * lda] : A[gk + gi * lda];
        ra[r][u] = v;
      }
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = tl + 64 * r;
        const int kk = (OPB == OP_N) ? (4 * tw + u) : (tw + 4 * u);
        const int gj = j0 + j, gk = k0 + kk;
        zc v = zmake(0.0, 0.0);
        if (gj < n && gk < k) v = (OPB == OP_N) ? B[gk + gj * ldb] : B[gj + gk * ldb];
        rb[r][u] = v;
      }
  };
  auto store_stage = [&](int buf) {
#pragma unroll
    for (int r = 0; r < RA; ++r)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = tl + 64 * r;
        const int kk = (OPA == OP_N) ? (tw + 4 * u) : (4 * tw + u);
        As[buf][2 * kk][i] = ra[r][u].x;
        As[buf][2 * kk + 1][i] = (OPA == OP_C) ? -ra[r][u].y : ra[r][u].y;
      }
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = tl + 64 * r;
        const int kk = (OPB == OP_N) ? (4 * tw + u) : (tw + 4 * u);
        Bs[buf][2 * kk][j] = rb[r][u].x;
        Bs[buf][2 * kk + 1][j] = (OPB == OP_C) ? -rb[r][u].y : rb[r][u].y;
      }
  };

  const int nk = (k + TK - 1) / TK;
  const int kq = lane >> 4;           // K' row within a 4-row MFMA step
  const double asign = (kq & 1) ? -1.0 : 1.0;
  const int li = lane & 15;

  auto compute_stage = [&](int buf) {
#pragma unroll
    for (int ks = 0; ks < TK / 2; ++ks) {
      const int kr = 4 * ks + kq;
      double a[FM], am[FM], b1[FN], b2[FN];
#pragma unroll
      for (int r = 0; r < FM; ++r) {
        a[r] = As[buf][kr][wm * 16 * FM + r * 16 + li];
        am[r] = asign * a[r];
      }
#pragma unroll
      for (int c = 0; c < FN; ++c) {
        b1[c] = Bs[buf][kr][wn * 16 * FN + c * 16 + li];
        b2[c] = Bs[buf][kr ^ 1][wn * 16 * FN + c * 16 + li];
      }
#pragma unroll
      for (int r = 0; r < FM; ++r)
#pragma unroll
        for (int c = 0; c < FN; ++c) {
          acc_re[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[r], b1[c], acc_re[r][c], 0, 0, 0);
          acc_im[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b2[c], acc_im[r][c], 0, 0, 0);
        }
    }
  };

  load_stage(0);
  store_stage(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = DB ? (kt & 1) : 0;
    if (kt + 1 < nk) load_stage((kt + 1) * TK);
    compute_stage(cur);
    if (DB) {
      // the other buffer was last read in stage kt-1, before the barrier below it
      if (kt + 1 < nk) store_stage(cur ^ 1);
      __syncthreads();
    } else if (kt + 1 < nk) {
      __syncthreads();
      store_stage(0);
      __syncthreads();
    }
  }

  store_tile<FM, FN>(g, inst, i_in, i_out, i0 + wm * 16 * FM, j0 + wn * 16 * FN, lane, acc_re, acc_im);
}

// Launches whose tiles all lie inside m x n (m, n multiples of 64, k a
// multiple of LTK, each operand under 2 GiB): the same 64 x 64 tile and MFMA
// loop, with the operands going straight from memory into LDS
// (buffer_load_dwordx4 ... lds, 16 bytes per lane), LTK complex k per stage,
// NBUF stages in flight and one barrier per stage -- no register staging and
// no LDS stores.  LDS per stage and operand: [k][x] complex, x the tile's row
// (A) or column (B), one 1 KB k-row per wave-wide load (lane l writes x = l).
// Operands contiguous along x load one cache line per 8 lanes; those
// contiguous along k one 16-byte piece per lane.  In the MFMA loop lane
// (l % 16, kq = l / 16) reads component kq & 1 of complex k = 2 ks + kq / 2
// (B'' the other component); the two 32-lane halves of each read fall in 64
// distinct banks.  Conjugation is a per-lane sign.
constexpr int LTK = 8;

// FORM (the arithmetic, GemmArgs::prec numbering):
//   0: FP64, the K'-form above (two real MFMAs per two complex k);
//   1: the FP32 form of the same kernel -- the staged FP64 operands rounded
//      to FP32 in registers and multiplied on v_mfma_f32_16x16x4_f32 (twice
//      the FP64 MFMA rate) with FP32 accumulation; same LDS traffic;
//   3: FP64 with Gauss's three-multiplication complex product: per four
//      complex k each lane reads whole complex a = A[i, k], b = B[k, j]
//      (ds_read_b128) and feeds three MFMAs, P += ar br, Q += ai bi,
//      R += (ar + ai)(br + bi) (conjugation folded into the sums' signs),
//      then Re = P - Q, Im = R - P - Q: 3 real products per complex product
//      instead of 4.  Error: normwise like the 4-product form within a small
//      factor (Higham, "Stability of a method for multiplying complex
//      matrices with three real matrix multiplications", 1992); the imaginary
//      part is not componentwise accurate.
template <int OPA, int OPB, int NBUF, int T, int FORM = 0>
__global__ __launch_bounds__(256, 2) void zgemm_lds_kernel(const GemmGroup G) {
  static_assert(T == 64 || T == 32, "tile of 64 or 32");
  static_assert(FORM == 0 || FORM == 1 || FORM == 3, "arithmetic form");
  constexpr bool F32 = FORM == 1;
  constexpr int FM = T / 32, FN = T / 32;   // 16 x 16 MFMA blocks per wave and direction
  constexpr int QL = T / 32;                // DMA loads per wave, operand and stage
  const long long bid = blockIdx.x;
  const long long per = G.grid_padded >> 3;
  const long long logical = (bid & 7) * per + (bid >> 3);
  if (logical >= G.total) return;
  const int pidx = (int)(logical / G.tiles_per_problem);
  const long long rem = logical - (long long)pidx * G.tiles_per_problem;
  const int tiles_inst = G.tri ? G.tiles_m * (G.tiles_m + 1) / 2 : G.tiles_m * G.tiles_n;
  const int inst = (int)(rem / tiles_inst);
  const int tile = (int)(rem - (long long)inst * tiles_inst);
  int tm, tn;
  if (G.tri) {   // tile = tm (tm + 1) / 2 + tn, tn <= tm
    tm = (int)((sqrtf(8.0f * (float)tile + 1.0f) - 1.0f) * 0.5f);
    while ((tm + 1) * (tm + 2) / 2 <= tile) ++tm;
    while (tm * (tm + 1) / 2 > tile) --tm;
    tn = tile - tm * (tm + 1) / 2;
  } else {
    tm = tile % G.tiles_m;
    tn = tile / G.tiles_m;
  }
  const GemmArgs& g = G.p[pidx];
  const int m = g.m, n = g.n, k = g.k;
  const int i0 = tm * T, j0 = tn * T;
  const long long i_in = g.inner > 0 ? inst % g.inner : inst;
  const long long i_out = g.inner > 0 ? inst / g.inner : 0;
  const zc* A = g.A + i_in * g.strideA + i_out * g.outerA;
  const zc* B = g.B + i_in * g.strideB + i_out * g.outerB;
  const int lda = (int)g.lda, ldb = (int)g.ldb;

  __shared__ zc Ls[NBUF][2][LTK][T];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;

  // byte extents of op(A) (rowsA x colsA, ld lda) and op(B); the host
  // guarantees they fit in 31 bits, so every offset below does too
  const int rowsA = OPA == OP_N ? m : k, colsA = OPA == OP_N ? k : m;
  const int rowsB = OPB == OP_N ? k : n, colsB = OPB == OP_N ? n : k;
  const __amdgpu_buffer_rsrc_t rsA = __builtin_amdgcn_make_buffer_rsrc(
      (void*)A, (short)0, (rowsA - 1 + (colsA - 1) * lda + 1) * 16, 0x00020000);
  const __amdgpu_buffer_rsrc_t rsB = __builtin_amdgcn_make_buffer_rsrc(
      (void*)B, (short)0, (rowsB - 1 + (colsB - 1) * ldb + 1) * 16, 0x00020000);
  // wave w loads k-rows 2w and 2w + 1 of each operand: with T = 64 one row
  // per load, with T = 32 both rows in one (lanes 32..63 land in the second,
  // since a row is 32 complex); lane offsets in bytes
  const int xl = lane & (T - 1), kl = T == 32 ? (lane >> 5) : 0;
  int voA[2], voB[2];   // QL used (a size of QL here drops the host stubs)
#pragma unroll
  for (int q = 0; q < QL; ++q) {
    const int kk = 2 * wave + q + kl;
    voA[q] = (OPA == OP_N ? (i0 + xl + kk * lda) : (kk + (i0 + xl) * lda)) * 16;
    voB[q] = (OPB == OP_N ? (kk + (j0 + xl) * ldb) : (j0 + xl + kk * ldb)) * 16;
  }
  auto issue = [&](int kt) {
    const int buf = kt % NBUF, k0 = kt * LTK;
    const int sA = OPA == OP_N ? k0 * lda * 16 : k0 * 16;
    const int sB = OPB == OP_N ? k0 * 16 : k0 * ldb * 16;
#pragma unroll
    for (int q = 0; q < QL; ++q) {
      const int kk = 2 * wave + q;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rsA, (__attribute__((address_space(3))) void*)&Ls[buf][0][kk][0], 16, voA[q], sA, 0, 0);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rsB, (__attribute__((address_space(3))) void*)&Ls[buf][1][kk][0], 16, voB[q], sB, 0, 0);
    }
  };

  typedef typename std::conditional<F32, flt4, dbl4>::type AccV;
  AccV acc_re[FM][FN], acc_im[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) {
      acc_re[a][b] = AccV{0, 0, 0, 0};
      acc_im[a][b] = AccV{0, 0, 0, 0};
    }

  const int kq = lane >> 4, li = lane & 15, ri = kq & 1;
  const double sgn = ri ? -1.0 : 1.0;   // the A- sign of K' row kq
  // double offsets within a stage's operand: component ri (B'': 1 - ri)
  const int offa = (wm * 16 * FM + li) * 2 + ri;
  const int offb1 = (wn * 16 * FN + li) * 2 + ri, offb2 = (wn * 16 * FN + li) * 2 + (1 - ri);
  // Signs, all carried by the A operand so that B goes to the MFMA as loaded:
  // Re term  A-[i,k'] B'[k',j]  = sgn^(1 + [A conj] + [B conj]) a_ri b_ri,
  // Im term  A'[i,k'] B''[k',j] = (A conj ? sgn : 1)(B conj ? -sgn : 1) a_ri b_(1-ri)
  // (conjugation negates the imaginary component, ri = 1 for A and B', ri = 0
  // for B'').
  constexpr bool re_sgn = ((1 + (OPA == OP_C) + (OPB == OP_C)) & 1) != 0;
  constexpr bool im_sgn = (OPA == OP_C) != (OPB == OP_C);
  constexpr bool im_neg = OPB == OP_C;

  auto compute_stage = [&](int buf) {
    const double* __restrict__ La = reinterpret_cast<const double*>(&Ls[buf][0][0][0]);
    const double* __restrict__ Lb = reinterpret_cast<const double*>(&Ls[buf][1][0][0]);
#pragma unroll
    for (int ks = 0; ks < LTK / 2; ++ks) {
      const int kr = (2 * ks + (kq >> 1)) * 2 * T;   // doubles per k-row: T complex
      double a[FM], am[FM], b1[FN], b2[FN];
#pragma unroll
      for (int r = 0; r < FM; ++r) {
        const double v = La[kr + offa + 32 * r];
        am[r] = re_sgn ? sgn * v : v;
        const double vi = im_sgn ? sgn * v : v;
        a[r] = im_neg ? -vi : vi;
      }
#pragma unroll
      for (int c = 0; c < FN; ++c) {
        b1[c] = Lb[kr + offb1 + 32 * c];
        b2[c] = Lb[kr + offb2 + 32 * c];
      }
#pragma unroll
      for (int r = 0; r < FM; ++r)
#pragma unroll
        for (int c = 0; c < FN; ++c) {
          if constexpr (F32) {
            acc_re[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)am[r], (float)b1[c],
                                                                acc_re[r][c], 0, 0, 0);
            acc_im[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a[r], (float)b2[c],
                                                                acc_im[r][c], 0, 0, 0);
          } else {
            acc_re[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[r], b1[c], acc_re[r][c], 0, 0, 0);
            acc_im[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b2[c], acc_im[r][c], 0, 0, 0);
          }
        }
    }
  };

  // the three-multiplication form: accumulators P, Q, R (acc_re, acc_im and
  // acc_r3); the signs of the conjugated operands' imaginary parts
  AccV acc_r3[FORM == 3 ? FM : 1][FORM == 3 ? FN : 1];
  if constexpr (FORM == 3) {
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b) acc_r3[a][b] = AccV{0, 0, 0, 0};
  }
  constexpr double sgA = OPA == OP_C ? -1.0 : 1.0, sgB = OPB == OP_C ? -1.0 : 1.0;
  auto compute_stage3 = [&](int buf) {
    if constexpr (FORM == 3) {
    const zc* __restrict__ La = &Ls[buf][0][0][0];
    const zc* __restrict__ Lb = &Ls[buf][1][0][0];
#pragma unroll
    for (int ks = 0; ks < LTK / 4; ++ks) {
      const int kr = (4 * ks + kq) * T;   // complex k-row 4 ks + kq of the stage
      zc a[FM], b[FN];
      double as[FM], bs[FN];
#pragma unroll
      for (int r = 0; r < FM; ++r) {
        a[r] = La[kr + wm * 16 * FM + 16 * r + li];
        as[r] = fma(sgA, a[r].y, a[r].x);
      }
#pragma unroll
      for (int c = 0; c < FN; ++c) {
        b[c] = Lb[kr + wn * 16 * FN + 16 * c + li];
        bs[c] = fma(sgB, b[c].y, b[c].x);
      }
#pragma unroll
      for (int r = 0; r < FM; ++r)
#pragma unroll
        for (int c = 0; c < FN; ++c) {
          acc_re[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r].x, b[c].x, acc_re[r][c], 0, 0, 0);
          acc_im[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r].y, b[c].y, acc_im[r][c], 0, 0, 0);
          acc_r3[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[r], bs[c], acc_r3[r][c], 0, 0, 0);
        }
    }
    }
  };

  // stage kt is waited for (vmcnt counts this wave's loads, in order, 2 QL a
  // stage: those of the NBUF - 2 later stages may stay in flight), then one
  // barrier makes it visible and frees the buffer of stage kt - 1 for stage
  // kt + NBUF - 1
  const int nk = k / LTK;
#pragma unroll
  for (int p = 0; p < NBUF - 1; ++p)
    if (p < nk) issue(p);
  for (int kt = 0; kt < nk; ++kt) {
    const int ahead = (nk - 1 - kt) < (NBUF - 2) ? (nk - 1 - kt) : (NBUF - 2);
    if (ahead >= 2) __builtin_amdgcn_s_waitcnt(0xF70 + 4 * QL);        // vmcnt(4 QL)
    else if (ahead == 1) __builtin_amdgcn_s_waitcnt(0xF70 + 2 * QL);   // vmcnt(2 QL)
    else __builtin_amdgcn_s_waitcnt(0xF70);                            // vmcnt(0)
    // a bare s_barrier: __syncthreads()' fence would wait for every load in
    // flight (vmcnt(0)), the later stages included; the clobbers keep the
    // compiler from moving LDS accesses across it
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (kt + NBUF - 1 < nk) issue(kt + NBUF - 1);
    if constexpr (FORM == 3) compute_stage3(kt % NBUF);
    else compute_stage(kt % NBUF);
  }
  if constexpr (FORM == 3) {
    // Re = P - sQ, Im = R - P - sQ with s = sgA sgB (Q holds ai bi unsigned)
    constexpr double sq = sgA * sgB;
#pragma unroll
    for (int r = 0; r < FM; ++r)
#pragma unroll
      for (int c = 0; c < FN; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double p = acc_re[r][c][e], q = sq * acc_im[r][c][e];
          acc_re[r][c][e] = p - q;
          acc_im[r][c][e] = acc_r3[r][c][e] - p - q;
        }
  }
  store_tile<FM, FN, AccV, F32>(g, inst, i_in, i_out, i0 + wm * 16 * FM, j0 + wn * 16 * FN, lane,
                                acc_re, acc_im, G.tri && tm > tn);
}

// Tile variants (autotuner choice, negf_gemm_variant).  0: zgemm_lds_kernel
// with two stages (double-buffered LDS) where the whole launch qualifies,
// 64 x 64 tiles for launches of at least 512 of them, else 32 x 32 (m, n
// multiples of 32), else 64x64 register-staged; 1: 64x64 register-staged,
// double buffer, 2: 64x128 single, 3: 64x128 double, 4: 128x64 double, 5:
// 64x64 register-staged held to 128 VGPRs (4 workgroups per CU); 6, 7:
// zgemm_lds_kernel with three, four stages; 8: 64x64 register-staged, single
// buffer, everywhere; 9, 10: zgemm_lds_kernel, two stages, 64 x 64 and
// 32 x 32 tiles wherever they qualify.  Variants 6, 7, 9, 10 fall back to 8.
// Measured on MI355X at 512^3 x 64, per-instance operands, after warm-up
// (TFLOP/s, NN/NC/CN/CC): 0: 69.5/69.2/69.0/68.6, 6: 69.0/68.9/68.6/68.1,
// 8: 62.1/62.7/61.3/61.0; variants 1-5 were at or below 8 (occupancy
// matters more than tile reuse at this FP64 MFMA rate).  MFMA issue alone
// (register operands, no memory) reaches 70.5 here.
int g_variant = 0;
constexpr int NVAR = 11;

template <int OPA, int OPB, int FM, int FN, bool DB, int SK, int OCC = (FM * FN <= 4 && !DB) ? 2 : 1>
void launch_cfg(GemmGroup G, hipStream_t s) {
  G.tiles_m = (G.p[0].m + 32 * FM - 1) / (32 * FM);
  G.tiles_n = (G.p[0].n + 32 * FN - 1) / (32 * FN);
  G.tiles_per_problem = (long long)G.tiles_m * G.tiles_n * G.p[0].batch;
  G.total = G.tiles_per_problem * G.count;
  G.grid_padded = (G.total + 7) / 8 * 8;
  hipLaunchKernelGGL((zgemm_mfma_kernel<OPA, OPB, FM, FN, DB, SK, OCC>), dim3((unsigned)G.grid_padded),
                     dim3(256), 0, s, G);
}

// Every tile of the group inside m x n, whole LTK stages, and 31-bit byte
// offsets into each operand (the buffer loads of zgemm_lds_kernel).
bool lds_eligible(const GemmGroup& G, int T) {
  const GemmArgs& g0 = G.p[0];
  if (g0.m % T || g0.n % T || g0.k % LTK) return false;
  for (int i = 0; i < G.count; ++i) {
    const GemmArgs& g = G.p[i];
    const long long rowsA = g.opA == OP_N ? g.m : g.k, colsA = g.opA == OP_N ? g.k : g.m;
    const long long rowsB = g.opB == OP_N ? g.k : g.n, colsB = g.opB == OP_N ? g.n : g.k;
    if ((rowsA + (colsA - 1) * g.lda) * 16 >= (1LL << 31)) return false;
    if ((rowsB + (colsB - 1) * g.ldb) * 16 >= (1LL << 31)) return false;
  }
  return true;
}

template <int OPA, int OPB, int NBUF, int T, int FORM = 0>
void launch_lds(GemmGroup G, hipStream_t s) {
  G.tiles_m = G.p[0].m / T;
  G.tiles_n = G.p[0].n / T;
  G.tri = G.count == 1 && G.p[0].herm != 0 && G.p[0].m == G.p[0].n;
  G.tiles_per_problem = (G.tri ? (long long)G.tiles_m * (G.tiles_m + 1) / 2
                               : (long long)G.tiles_m * G.tiles_n) * G.p[0].batch;
  G.total = G.tiles_per_problem * G.count;
  G.grid_padded = (G.total + 7) / 8 * 8;
  hipLaunchKernelGGL((zgemm_lds_kernel<OPA, OPB, NBUF, T, FORM>), dim3((unsigned)G.grid_padded),
                     dim3(256), 0, s, G);
}

template <int OPA, int OPB>
void launch_t(const GemmGroup& G, hipStream_t s) {
  if (G.p[0].k <= 2 * TK) {
    launch_cfg<OPA, OPB, 2, 2, false, 1>(G, s);
    return;
  }
  const GemmArgs& g0 = G.p[0];
  const long long tiles64 = (long long)(g0.m / 64) * (g0.n / 64) * g0.batch * G.count;
  if (g0.prec == 1) {   // the FP32 form where the LDS-DMA kernel applies
    if (tiles64 >= 512 && lds_eligible(G, 64)) return launch_lds<OPA, OPB, 2, 64, 1>(G, s);
    if (lds_eligible(G, 32)) return launch_lds<OPA, OPB, 2, 32, 1>(G, s);
  }
  if (g0.prec == 3) {   // FP64, three real products, where the LDS-DMA kernel applies
    if (tiles64 >= 512 && lds_eligible(G, 64)) return launch_lds<OPA, OPB, 2, 64, 3>(G, s);
    if (lds_eligible(G, 32)) return launch_lds<OPA, OPB, 2, 32, 3>(G, s);
  }
  // a forced variant, else the tuned one for this shape, else the thresholds
  const int var = g_variant ? g_variant
                            : tune_gemm_variant(g0.m, g0.n, g0.k, (long long)g0.batch * G.count);
  switch (var) {
    case 0:
      // fewer than two 64 x 64 tiles per CU (one instance of a 512^3 product
      // has 64), or m, n multiples of 32 only: 32 x 32 tiles
      if (tiles64 >= 512 && lds_eligible(G, 64)) return launch_lds<OPA, OPB, 2, 64>(G, s);
      if (lds_eligible(G, 32)) return launch_lds<OPA, OPB, 2, 32>(G, s);
      break;
    case 6: if (lds_eligible(G, 64)) return launch_lds<OPA, OPB, 3, 64>(G, s); break;
    case 7: if (lds_eligible(G, 64)) return launch_lds<OPA, OPB, 4, 64>(G, s); break;
    case 9: if (lds_eligible(G, 64)) return launch_lds<OPA, OPB, 2, 64>(G, s); break;
    case 10: if (lds_eligible(G, 32)) return launch_lds<OPA, OPB, 2, 32>(G, s); break;
    default: break;
  }
  switch (var) {
    case 0: case 6: case 7: case 8: case 9: case 10: launch_cfg<OPA, OPB, 2, 2, false, 0>(G, s); break;
    case 1: launch_cfg<OPA, OPB, 2, 2, true, 0>(G, s); break;
    case 2: launch_cfg<OPA, OPB, 2, 4, false, 0>(G, s); break;
    case 3: launch_cfg<OPA, OPB, 2, 4, true, 0>(G, s); break;
    case 4: launch_cfg<OPA, OPB, 4, 2, true, 0>(G, s); break;
    default: launch_cfg<OPA, OPB, 2, 2, false, 0, 4>(G, s); break;
  }
}

int check_args(const GemmArgs& g) {
  if (g.m < 0 || g.n < 0 || g.k < 0 || g.batch < 0) return Status::BAD_ARG;
  if (g.opA < 0 || g.opA > 2 || g.opB < 0 || g.opB > 2) return Status::BAD_ARG;
  const int rowsA = g.opA == OP_N ? g.m : g.k;
  const int rowsB = g.opB == OP_N ? g.k : g.n;
  if (g.m > 0 && g.n > 0 && g.k > 0) {
    if (g.lda < (rowsA > 1 ? rowsA : 1) || g.ldb < (rowsB > 1 ? rowsB : 1)) return Status::BAD_ARG;
    if (!g.A || !g.B) return Status::BAD_ARG;
  }
  if (g.m > 0 && g.n > 0) {
    if (g.ldd < (g.m > 1 ? g.m : 1) || !g.D) return Status::BAD_ARG;
    if ((g.beta.x != 0.0 || g.beta.y != 0.0) && (!g.C || g.ldc < (g.m > 1 ? g.m : 1)))
      return Status::BAD_ARG;
  }
  if (g.inner < 0) return Status::BAD_ARG;
  if (g.herm != 0 && (g.m != g.n || (g.herm != 1 && g.herm != -1))) return Status::BAD_ARG;
  if (g.prec < 0 || g.prec > 3) return Status::BAD_ARG;
  return Status::OK;
}

}  // namespace

int zgemm_set_variant(int v) {
  if (v < 0 || v >= NVAR) return Status::BAD_ARG;
  g_variant = v;
  return Status::OK;
}

int zgemm_group_launch(const GemmArgs* g, int count, hipStream_t s) {
  if (count <= 0) return Status::OK;
  if (count > MAXG) {
    int rc = zgemm_group_launch(g, MAXG, s);
    if (rc) return rc;
    return zgemm_group_launch(g + MAXG, count - MAXG, s);
  }
  for (int i = 0; i < count; ++i) {
    int rc = check_args(g[i]);
    if (rc) return rc;
  }
  if (g[0].prec == 2) {
    // the emulated product, one problem after another (they share the
    // residue workspace); k == 0 has no product and runs as FP64
    for (int i = 0; i < count; ++i) {
      GemmArgs gi = g[i];
      int rc;
      if (gi.k == 0) {
        gi.prec = 0;
        rc = zgemm_group_launch(&gi, 1, s);
      } else {
        rc = zgemm_emu_launch(gi, s);
      }
      if (rc) return rc;
    }
    return Status::OK;
  }
  // one launch per run of problems with the same shape and op codes; a
  // problem whose output another problem of the run reads (or writes) starts
  // a new run, since the problems of one launch run concurrently
  auto conflicts = [&](int i) {
    for (int j = 0; j < i; ++j) {
      const void* Di = g[i].D;
      const void* Dj = g[j].D;
      if (Di == Dj || Di == g[j].A || Di == g[j].B || Di == g[j].C || Dj == g[i].A ||
          Dj == g[i].B || Dj == g[i].C)
        return true;
    }
    return false;
  };
  for (int i = 1; i < count; ++i) {
    if (g[i].m != g[0].m || g[i].n != g[0].n || g[i].k != g[0].k || g[i].opA != g[0].opA ||
        g[i].opB != g[0].opB || g[i].batch != g[0].batch || g[0].herm != 0 || g[i].herm != 0 ||
        g[i].prec != g[0].prec || conflicts(i)) {
      int rc = zgemm_group_launch(g, i, s);
      if (rc) return rc;
      return zgemm_group_launch(g + i, count - i, s);
    }
  }
  const GemmArgs& g0 = g[0];
  if (g0.m == 0 || g0.n == 0 || g0.batch == 0) return Status::OK;
  GemmGroup G{};
  for (int i = 0; i < count; ++i) G.p[i] = g[i];
  G.count = count;
  // grid size bound with the smallest tile (32 x 32)
  const long long tiles = (long long)((g0.m + 31) / 32) * ((g0.n + 31) / 32) * g0.batch * count;
  if (tiles + 8 > 0x7fffffffLL) return Status::TOO_LARGE;
  // k == 0: D = beta C (+shift); the kernel handles it (no K stages, zero acc).
  const int key = g0.opA * 3 + g0.opB;
  switch (key) {
    case 0: launch_t<OP_N, OP_N>(G, s); break;
    case 1: launch_t<OP_N, OP_T>(G, s); break;
    case 2: launch_t<OP_N, OP_C>(G, s); break;
    case 3: launch_t<OP_T, OP_N>(G, s); break;
    case 4: launch_t<OP_T, OP_T>(G, s); break;
    case 5: launch_t<OP_T, OP_C>(G, s); break;
    case 6: launch_t<OP_C, OP_N>(G, s); break;
    case 7: launch_t<OP_C, OP_T>(G, s); break;
    case 8: launch_t<OP_C, OP_C>(G, s); break;
  }
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

int zgemm_launch(const GemmArgs& g, hipStream_t s) { return zgemm_group_launch(&g, 1, s); }

}  // namespace negf
