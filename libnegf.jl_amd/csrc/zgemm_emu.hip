// Batched complex GEMM with FP64-class accuracy on the int8 matrix cores
// (v_mfma_i32_16x16x64_i8): Ozaki scheme II (H. Ozaki, Y. Uchino,
// T. Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of floating-point
// matrix multiplication using an integer modular technique", 2025).
//
//   D[idx] = alpha * op(A[idx]) op(B[idx]) + beta * C[idx] (+ shift I), all
//   operands complex128 in memory, as zgemm_launch with GemmArgs::prec = 2.
//
// 1. Scaling (emu_rowmax_kernel and emu_resid_kernel, once per operand).  Row i of op(A) is scaled
//    by 2^eA_i so that its largest |re| or |im| lies in [2^(beta-1), 2^beta),
//    column j of op(B) likewise by 2^eB_j, and rounded to integers a, b
//    (|a|, |b| <= 2^beta).  This rounding is the only error of the product:
//    at most 2^-beta of the row / column maximum per element.
// 2. Residues.  For pairwise coprime moduli m_1..m_n <= 256 with product M,
//    the integers are stored as int8 residues in [-m/2, m/2] per modulus.
//    The complex product is two real products over K' = 2K:
//      Re = sum_k (ar br - ai bi) = [ar, -ai] . [br, bi]
//      Im = sum_k (ai br + ar bi) = [ai,  ar] . [br, bi]
//    so op(A) is stored in the forms R = (ar, -ai) and I = (ai, ar) and op(B)
//    as (br, bi), each k' contiguous per row of op(A) / column of op(B).
// 3. Products (emu_gemm_kernel).  Per modulus, the int32 products of the
//    residues (exact: |sum| <= K' 2^14 < 2^31) reduced mod m_l give the
//    residue of the exact integer product; the Chinese remainder theorem
//    recovers it: x = sum_l r_l w_l mod M, w_l = (M/m_l) ((M/m_l)^-1 mod m_l).
//    With K' 2^(2 beta) <= M/4 the symmetric representative is the product.
//    The sum is accumulated exactly in FP64 as S_hi = sum r_l w_hi_l (w_hi_l:
//    w_l rounded to a multiple of 2^g, so every partial sum is an integer
//    multiple of 2^g below 2^(53+g)) plus S_lo = sum r_l w_lo_l (rounded,
//    |S_lo| < 2^(g+11)); then q = rint((S_hi + S_lo) / M) and x = (S_hi -
//    q M_hi) + (S_lo - q M_lo), exact up to rounding far below the scaled
//    product.  Finally D = alpha x 2^-(eA_i + eB_j) + beta C.
//
// Tile: 64 x 64 complex outputs per 256-thread workgroup (2 x 2 waves of
// 32 x 32), K' staged 64 bytes at a time for the three operand forms, loaded
// by the buffer unit straight into LDS (two stages); each wave reads its
// 16-byte MFMA fragments with ds_read_b128, the 16-byte chunks of every LDS
// row XOR-swizzled by bits 2-3 of the row so that the 16 lanes of a read
// phase hit 64 distinct banks.  Per stage and wave: 8 int8 MFMAs (2 x 2
// blocks, Re and Im).  The CRT state (S_hi, S_lo of 32 real outputs per
// lane, 128 VGPRs) stays in registers across the moduli.
#include "negf_internal.hpp"

#include <cmath>
#include <cstring>

namespace negf {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

// Pairwise coprime moduli <= 256, largest first (255 = 3*5*17, 253 = 11*23,
// 247 = 13*19, 217 = 7*31, the rest prime or 2^8).
constexpr int kModuli[16] = {256, 255, 253, 251, 247, 241, 239, 233,
                             229, 227, 223, 217, 211, 199, 197, 193};
constexpr int kMaxMod = 16;

struct EmuConsts {
  int nmod, beta;
  int fp32_mod;                 // int32 products below 2^24: reduce in FP32
  float mf[kMaxMod], invmf[kMaxMod];
  double md[kMaxMod], invmd[kMaxMod];
  double whi[kMaxMod], wlo[kMaxMod];
  double Mhi, Mlo, invM;
};

// Residue planes of one operand: int8 [mod][inst][form][stage < Kp / 128]
// [row < Rp][128 bytes of k'] (k'-tiled: one stage of the product kernel for
// 64 rows is one contiguous 8 KB), and the scale exponent of every row,
// int32 [inst][row < Rp].
struct EmuPlanes {
  signed char* res;
  int* expo;
  int batch;      // instances stored (1: one operand shared by every instance)
  int forms;      // 2 (op(A): R and I) or 1 (op(B))
  int Rp, Kp;
};

struct PrepArgs {
  const zc* X;
  long long ld, stride, outer;
  int inner;
  int r_major;    // element (r, k) at X[r + k ld] (1) or X[k + r ld] (0)
  int conj;       // op C: negate the imaginary parts
  int rows, K;
  EmuPlanes P;
};

__device__ __forceinline__ zc prep_load(const PrepArgs& a, const zc* X, int r, int k) {
  zc v = a.r_major ? X[r + (long long)k * a.ld] : X[k + (long long)r * a.ld];
  if (a.conj) v.y = -v.y;
  return v;
}

// Symmetric residue of the integer-valued double x (|x| <= 2^52) mod m:
// x - m rint(x / m), exact (|x / m| rounds within 2^-50 relative).
__device__ __forceinline__ int residue(double x, double m, double invm) {
  return (int)fma(-rint(x * invm), m, x);
}

__device__ __forceinline__ unsigned pack4(int a, int b, int c, int d) {
  return (unsigned)(a & 0xff) | ((unsigned)(b & 0xff) << 8) | ((unsigned)(c & 0xff) << 16) |
         ((unsigned)(d & 0xff) << 24);
}

// Scaling and residues in two kernels.  emu_rowmax_kernel: one workgroup
// per 64 rows of one instance finds each row's largest |re|, |im| (loads
// coalesced for the operand's layout) and stores its scale exponent (kDead
// for rows below 2^-900, whose scale would overflow: they count as zero).
// emu_resid_kernel: one workgroup per (64 rows, 64 complex k = one 128-byte
// stage of the product kernel, instance) loads its chunk coalesced into an
// LDS tile [row][k], then thread t works on row t / 4, complex k 16 (t % 4)
// .. + 15 and writes its residues in the k'-tiled layout [stage][row][128
// bytes]: 32 bytes per thread, form and modulus, a wave's 16 rows one
// contiguous 2 KB.  Rows >= rows and k >= K are written as zeros up to Rp,
// Kp.  Each element is read twice from memory (once per kernel); the
// residues are the work (per element and modulus: a multiply, a rounding,
// a fused multiply-add and a conversion).
constexpr int PKC = 64;
constexpr int kDead = -(1 << 20);

__global__ __launch_bounds__(256) void emu_rowmax_kernel(const PrepArgs a, const EmuConsts c) {
  __shared__ double red[4][64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int inst = blockIdx.y;
  const int r0 = blockIdx.x * 64;
  const long long i_in = a.inner > 0 ? inst % a.inner : inst;
  const long long i_out = a.inner > 0 ? inst / a.inner : 0;
  const zc* X = a.X + i_in * a.stride + i_out * a.outer;
  if (a.r_major) {
    // X[r + k ld]: lanes along the 64 rows, the waves split k
    const int r = r0 + lane;
    double mx = 0.0;
    if (r < a.rows)
      for (int k = wave; k < a.K; k += 4) {
        const zc v = X[r + (long long)k * a.ld];
        mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
      }
    red[wave][lane] = mx;
  } else {
    // X[k + r ld]: lanes along k, each wave 16 rows
    for (int q = 0; q < 16; ++q) {
      const int lr = 16 * wave + q, r = r0 + lr;
      double mx = 0.0;
      if (r < a.rows)
        for (int k = lane; k < a.K; k += 64) {
          const zc v = X[k + (long long)r * a.ld];
          mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
        }
      mx = wave_max(mx);
      if (lane == 0) red[0][lr] = mx;
    }
  }
  __syncthreads();
  if (t < 64) {
    double mx = red[0][t];
    if (a.r_major) mx = fmax(fmax(mx, red[1][t]), fmax(red[2][t], red[3][t]));
    // largest |re|, |im| of the row in [2^(beta-1), 2^beta) after scaling
    const bool live = mx >= 0x1p-900;
    a.P.expo[(long long)inst * a.P.Rp + r0 + t] = live ? c.beta - (ilogb(mx) + 1) : kDead;
  }
}

__global__ __launch_bounds__(256) void emu_resid_kernel(const PrepArgs a, const EmuConsts c) {
  __shared__ zc T[64][PKC + 1];
  const int t = threadIdx.x;
  const int inst = blockIdx.z;
  const int r0 = blockIdx.x * 64;
  const int kt = blockIdx.y, k0 = kt * PKC;
  const long long i_in = a.inner > 0 ? inst % a.inner : inst;
  const long long i_out = a.inner > 0 ? inst / a.inner : 0;
  const zc* X = a.X + i_in * a.stride + i_out * a.outer;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    int lr, lk;
    if (a.r_major) {      // X[r + k ld]: lanes along rows
      lr = t & 63;
      lk = (t >> 6) + 4 * j;
    } else {              // X[k + r ld]: lanes along k
      lk = t & 63;
      lr = (t >> 6) + 4 * j;
    }
    const int r = r0 + lr, k = k0 + lk;
    zc v = zmake(0.0, 0.0);
    if (r < a.rows && k < a.K) v = prep_load(a, X, r, k);
    T[lr][lk] = v;
  }
  const int my_row = t >> 2, my_part = t & 3;
  const int row = r0 + my_row;
  const int e = a.P.expo[(long long)inst * a.P.Rp + row];
  const double sc = e == kDead ? 0.0 : ldexp(1.0, e);
  __syncthreads();
  const long long plane = (long long)a.P.Rp * a.P.Kp;
  signed char* base0 = a.P.res + (long long)inst * a.P.forms * plane +
                       ((long long)kt * a.P.Rp + row) * 128 + 32 * my_part;
#pragma unroll
  for (int h = 0; h < 2; ++h) {   // two halves of 8 complex: 16 bytes each
    double xr[8], xi[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const zc v = T[my_row][16 * my_part + 8 * h + q];
      xr[q] = rint(v.x * sc);
      xi[q] = rint(v.y * sc);
    }
    for (int l = 0; l < c.nmod; ++l) {
      const double m = c.md[l], im = c.invmd[l];
      int rr[8], ri[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        rr[q] = residue(xr[q], m, im);
        ri[q] = residue(xi[q], m, im);
      }
      signed char* base = base0 + (long long)l * a.P.batch * a.P.forms * plane + 16 * h;
      if (a.P.forms == 1) {   // (br, bi)
        *(uint4*)base =
            make_uint4(pack4(rr[0], ri[0], rr[1], ri[1]), pack4(rr[2], ri[2], rr[3], ri[3]),
                       pack4(rr[4], ri[4], rr[5], ri[5]), pack4(rr[6], ri[6], rr[7], ri[7]));
      } else {                // R = (ar, -ai), I = (ai, ar)
        *(uint4*)base = make_uint4(
            pack4(rr[0], -ri[0], rr[1], -ri[1]), pack4(rr[2], -ri[2], rr[3], -ri[3]),
            pack4(rr[4], -ri[4], rr[5], -ri[5]), pack4(rr[6], -ri[6], rr[7], -ri[7]));
        *(uint4*)(base + plane) = make_uint4(
            pack4(ri[0], rr[0], ri[1], rr[1]), pack4(ri[2], rr[2], ri[3], rr[3]),
            pack4(ri[4], rr[4], ri[5], rr[5]), pack4(ri[6], rr[6], ri[7], rr[7]));
      }
    }
  }
}

constexpr int EKS = 128;  // K' bytes per stage (two MFMA k-steps)
constexpr int ENB = 3;    // stages in flight

struct EmuGemmArgs {
  GemmArgs g;
  EmuPlanes A, B;         // A: forms R, I over rows of op(A); B: over columns of op(B)
  int tiles_m, tiles_n;
  long long total, grid_padded;
};

__global__ __launch_bounds__(256, 2) void emu_gemm_kernel(const EmuGemmArgs E, const EmuConsts c) {
  const GemmArgs& g = E.g;
  const long long bid = blockIdx.x;
  const long long per = E.grid_padded >> 3;
  const long long logical = (bid & 7) * per + (bid >> 3);   // an instance's tiles on one XCD
  if (logical >= E.total) return;
  const int tiles_inst = E.tiles_m * E.tiles_n;
  const int inst = (int)(logical / tiles_inst);
  const int tile = (int)(logical - (long long)inst * tiles_inst);
  const int tm = tile % E.tiles_m, tn = tile / E.tiles_m;
  const int i0 = tm * 64, j0 = tn * 64;
  const int ia = E.A.batch == 1 ? 0 : inst, ib = E.B.batch == 1 ? 0 : inst;

  // [stage][operand: R, I, B][row][128 bytes = 8 chunks of 16, swizzled by
  // chunk ^ ((row >> 1) & 7)]
  __shared__ __attribute__((aligned(16))) signed char Ls[ENB][3][64][EKS];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, kq = lane >> 4;

  const int nk = E.A.Kp / EKS;
  const int T = c.nmod * nk;
  const long long planeA = (long long)E.A.Rp * E.A.Kp, planeB = (long long)E.B.Rp * E.B.Kp;

  // DMA: 24 wave-loads of 1 KB (8 rows) per stage; wave w issues 6w .. 6w + 5
  // (operand idx / 8, row block idx % 8); lane l fills row 8 rb + l / 8,
  // slot l % 8 with the source chunk (l % 8) ^ ((row >> 1) & 7)
  int vo[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int idx = 6 * wave + q;
    const int o = idx >> 3, rb = idx & 7;
    const int row = 8 * rb + (lane >> 3);
    const int chunk = (lane & 7) ^ ((row >> 1) & 7);
    vo[q] = ((o == 2 ? j0 : i0) + row) * EKS + 16 * chunk;
  }
  auto issue = [&](int l, int kt, int buf) {
    const signed char* pa = E.A.res + ((long long)l * E.A.batch + ia) * 2 * planeA;
    const signed char* pb = E.B.res + ((long long)l * E.B.batch + ib) * planeB;
    const __amdgpu_buffer_rsrc_t rR =
        __builtin_amdgcn_make_buffer_rsrc((void*)pa, (short)0, (int)planeA, 0x00020000);
    const __amdgpu_buffer_rsrc_t rI =
        __builtin_amdgcn_make_buffer_rsrc((void*)(pa + planeA), (short)0, (int)planeA, 0x00020000);
    const __amdgpu_buffer_rsrc_t rB =
        __builtin_amdgcn_make_buffer_rsrc((void*)pb, (short)0, (int)planeB, 0x00020000);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int idx = 6 * wave + q;
      const int o = idx >> 3, rb = idx & 7;
      const __amdgpu_buffer_rsrc_t rs = o == 0 ? rR : (o == 1 ? rI : rB);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rs, (__attribute__((address_space(3))) void*)&Ls[buf][o][8 * rb][0], 16, vo[q],
          kt * EKS * (o == 2 ? E.B.Rp : E.A.Rp), 0, 0);
    }
  };

  double shr[2][2][4], slr[2][2][4], shi[2][2][4], sli[2][2][4];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int cc = 0; cc < 2; ++cc)
#pragma unroll
      for (int e = 0; e < 4; ++e) shr[r][cc][e] = slr[r][cc][e] = shi[r][cc][e] = sli[r][cc][e] = 0.0;
  v4i accr[2][2], acci[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) accr[r][cc] = acci[r][cc] = v4i{0, 0, 0, 0};

  // fragment offsets: row li of the 16-row fragment, k' chunk 4 s + kq of the
  // stage (s = 0, 1: its two MFMA k-steps)
  int offa[2][2], offb[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      const int row = wm * 32 + r * 16 + li;
      offa[r][st] = row * EKS + 16 * ((4 * st + kq) ^ ((row >> 1) & 7));
      const int col = wn * 32 + r * 16 + li;
      offb[r][st] = col * EKS + 16 * ((4 * st + kq) ^ ((col >> 1) & 7));
    }

  // stage t = (modulus l, k-step kt) lands in buffer t % 3; two stages ahead
  int il = 0, ikt = 0, ibuf = 0;   // next stage to issue
  auto advance = [&]() {
    if (++ikt == nk) { ikt = 0; ++il; }
    ibuf = ibuf == ENB - 1 ? 0 : ibuf + 1;
  };
  issue(il, ikt, ibuf);
  advance();
  if (T > 1) {
    issue(il, ikt, ibuf);
    advance();
  }
  int l = 0, kt = 0, buf = 0;
  for (int t = 0; t < T; ++t) {
    // stage t has landed once at most the 6 loads of stage t + 1 remain
    if (t + 1 < T) __builtin_amdgcn_s_waitcnt(0xF70 + 6);   // vmcnt(6)
    else __builtin_amdgcn_s_waitcnt(0xF70);                 // vmcnt(0)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    // buffer (t + 2) % 3 was last read in stage t - 1, before the barrier
    if (t + 2 < T) {
      issue(il, ikt, ibuf);
      advance();
    }
    const signed char* LR = &Ls[buf][0][0][0];
    const signed char* LI = &Ls[buf][1][0][0];
    const signed char* LB = &Ls[buf][2][0][0];
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      v4i fr[2], fi[2], fb[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        fr[r] = *(const v4i*)(LR + offa[r][st]);
        fi[r] = *(const v4i*)(LI + offa[r][st]);
        fb[r] = *(const v4i*)(LB + offb[r][st]);
      }
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
          accr[r][cc] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fr[r], fb[cc], accr[r][cc], 0, 0, 0);
          acci[r][cc] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fi[r], fb[cc], acci[r][cc], 0, 0, 0);
        }
    }
    if (kt == nk - 1) {
      // modulus l done: residues of the products, accumulated by the CRT
      const double whi = c.whi[l], wlo = c.wlo[l];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int cc = 0; cc < 2; ++cc)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            double rr, ri;
            if (c.fp32_mod) {
              const float m = c.mf[l], im = c.invmf[l];
              const float xr = (float)accr[r][cc][e], xi = (float)acci[r][cc][e];
              rr = (double)fmaf(-rintf(xr * im), m, xr);
              ri = (double)fmaf(-rintf(xi * im), m, xi);
            } else {
              const double m = c.md[l], im = c.invmd[l];
              const double xr = (double)accr[r][cc][e], xi = (double)acci[r][cc][e];
              rr = fma(-rint(xr * im), m, xr);
              ri = fma(-rint(xi * im), m, xi);
            }
            shr[r][cc][e] = fma(rr, whi, shr[r][cc][e]);
            slr[r][cc][e] = fma(rr, wlo, slr[r][cc][e]);
            shi[r][cc][e] = fma(ri, whi, shi[r][cc][e]);
            sli[r][cc][e] = fma(ri, wlo, sli[r][cc][e]);
            accr[r][cc][e] = 0;
            acci[r][cc][e] = 0;
          }
    }
    if (++kt == nk) { kt = 0; ++l; }
    buf = buf == ENB - 1 ? 0 : buf + 1;
  }

  // reconstruction, unscaling and the epilogue (output layout of the 16x16
  // int8 MFMA: lane l, element e -> row 4 (l / 16) + e, column l % 16)
  const long long i_in = g.inner > 0 ? inst % g.inner : inst;
  const long long i_out = g.inner > 0 ? inst / g.inner : 0;
  zc* __restrict__ D = g.D + i_in * g.strideD + i_out * g.outerD;
  const zc* __restrict__ C = g.C ? g.C + i_in * g.strideC + i_out * g.outerC : nullptr;
  const bool use_c = (g.beta.x != 0.0 || g.beta.y != 0.0) && C != nullptr;
  const bool use_shift = g.shift != nullptr;
  const zc sh = use_shift ? g.shift[inst * g.shift_stride] : zmake(0.0, 0.0);
  const zc alpha = g.scale ? zmul(g.alpha, g.scale[inst * g.scale_stride]) : g.alpha;
  const int* ea = E.A.expo + (long long)ia * E.A.Rp;
  const int* eb = E.B.expo + (long long)ib * E.B.Rp;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const int j = j0 + wn * 32 + cc * 16 + li;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = i0 + wm * 32 + r * 16 + 4 * kq + e;
        if (i < g.m && j < g.n) {
          const double qr = rint((shr[r][cc][e] + slr[r][cc][e]) * c.invM);
          const double qi = rint((shi[r][cc][e] + sli[r][cc][e]) * c.invM);
          const double xr = fma(-qr, c.Mhi, shr[r][cc][e]) + fma(-qr, c.Mlo, slr[r][cc][e]);
          const double xi = fma(-qi, c.Mhi, shi[r][cc][e]) + fma(-qi, c.Mlo, sli[r][cc][e]);
          const int ex = -(ea[i] + eb[j]);
          const double vr = ldexp(xr, ex), vi = ldexp(xi, ex);
          zc out = zmake(alpha.x * vr - alpha.y * vi, alpha.x * vi + alpha.y * vr);
          if (use_c) out = zadd(out, zmul(g.beta, C[i + (long long)j * g.ldc]));
          if (use_shift && i == j) out = zadd(out, sh);
          D[i + (long long)j * g.ldd] = out;
        }
      }
    }
}

// ---------------------------------------------------------------- host side
typedef unsigned __int128 u128;
typedef __int128 i128;

int modinv(int a, int m) {   // a^-1 mod m, gcd(a, m) = 1
  int t = 0, nt = 1, r = m, nr = a % m;
  while (nr != 0) {
    const int q = r / nr;
    int tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  return t < 0 ? t + m : t;
}

double i128_to_double(i128 v) { return (double)v; }

// Constants for nmod moduli and products over K' = 2k terms.
bool emu_consts(int nmod, int k, EmuConsts* c) {
  if (nmod < 2 || nmod > kMaxMod || k <= 0) return false;
  std::memset(c, 0, sizeof(*c));
  c->nmod = nmod;
  u128 M = 1;
  for (int l = 0; l < nmod; ++l) M *= (u128)kModuli[l];
  const double log2M = std::log2((double)M);
  const double Kp = 2.0 * k;
  // K' 2^(2 beta) <= M / 4 (the symmetric product leaves a factor 2 of
  // margin for the rounding of q), beta <= 52 (exact doubles)
  int beta = (int)std::floor((log2M - 2.0 - std::log2(Kp)) * 0.5 - 1e-9);
  if (beta > 52) beta = 52;
  if (beta < 8) return false;
  c->beta = beta;
  c->fp32_mod = Kp * 128.0 * 128.0 <= 16777216.0 ? 1 : 0;
  // grid 2^g of the exact part: nmod * 131 * M / 2^g < 2^53
  const int g = (int)std::ceil(std::log2((double)nmod * 131.0) + log2M - 53.0 + 1e-9);
  const u128 half = (u128)1 << (g - 1);
  for (int l = 0; l < nmod; ++l) {
    const int m = kModuli[l];
    const u128 Ml = M / (u128)m;
    const int inv = modinv((int)(Ml % (u128)m), m);
    const u128 w = (Ml * (u128)inv) % M;
    const u128 hi = ((w + half) >> g) << g;
    c->mf[l] = (float)m;
    c->invmf[l] = 1.0f / (float)m;
    c->md[l] = (double)m;
    c->invmd[l] = 1.0 / (double)m;
    c->whi[l] = std::ldexp((double)(unsigned long long)(hi >> g), g);
    c->wlo[l] = i128_to_double((i128)w - (i128)hi);
  }
  const u128 Mhi = ((M + half) >> g) << g;
  c->Mhi = std::ldexp((double)(unsigned long long)(Mhi >> g), g);
  c->Mlo = i128_to_double((i128)M - (i128)Mhi);
  c->invM = 1.0 / (double)M;
  return true;
}

long long round_up(long long x, long long a) { return (x + a - 1) / a * a; }

bool broadcast(long long stride, long long outer, int inner) {
  return stride == 0 && (inner == 0 || outer == 0);
}

}  // namespace

int zgemm_emu_default_moduli() { return 13; }

size_t zgemm_emu_work_bytes(int m, int n, int k, int batchA, int batchB, int nmod) {
  if (nmod <= 0) nmod = zgemm_emu_default_moduli();
  const long long Mp = round_up(m, 64), Np = round_up(n, 64), Kp = round_up(2LL * k, EKS);
  size_t b = 0;
  b += (size_t)nmod * batchA * 2 * Mp * Kp + 256;
  b += (size_t)nmod * batchB * Np * Kp + 256;
  b += sizeof(int) * (size_t)(batchA * Mp + batchB * Np) + 512;
  return b + 1024;
}

int zgemm_emu_launch(const GemmArgs& g, hipStream_t s) {
  if (g.m == 0 || g.n == 0 || g.batch == 0) return Status::OK;
  if (!g.work) return Status::BAD_ARG;
  const int nmod = g.nmod > 0 ? g.nmod : zgemm_emu_default_moduli();
  if (g.k == 0) return Status::BAD_ARG;   // the caller runs k == 0 in FP64
  EmuConsts c;
  if (!emu_consts(nmod, g.k, &c)) return Status::BAD_ARG;
  const bool bA = broadcast(g.strideA, g.outerA, g.inner);
  const bool bB = broadcast(g.strideB, g.outerB, g.inner);
  const int batchA = bA ? 1 : g.batch, batchB = bB ? 1 : g.batch;
  const size_t need = zgemm_emu_work_bytes(g.m, g.n, g.k, batchA, batchB, nmod);
  if ((size_t)g.work_bytes < need) return Status::BAD_ARG;
  const long long Mp = round_up(g.m, 64), Np = round_up(g.n, 64), Kp = round_up(2LL * g.k, EKS);
  if (Mp * Kp * 2 >= (1LL << 31) || Np * Kp >= (1LL << 31)) return Status::TOO_LARGE;
  char* p = (char*)g.work;
  auto take = [&](size_t bytes) {
    char* r = p;
    p += (bytes + 255) / 256 * 256;
    return r;
  };
  EmuPlanes PA{(signed char*)take((size_t)nmod * batchA * 2 * Mp * Kp), nullptr, batchA, 2,
               (int)Mp, (int)Kp};
  EmuPlanes PB{(signed char*)take((size_t)nmod * batchB * Np * Kp), nullptr, batchB, 1, (int)Np,
               (int)Kp};
  PA.expo = (int*)take(sizeof(int) * (size_t)batchA * Mp);
  PB.expo = (int*)take(sizeof(int) * (size_t)batchB * Np);
  // op(A) rows: N -> A[i + k lda] (rows contiguous), T / C -> A[k + i lda]
  PrepArgs pa{g.A, g.lda, g.strideA, g.outerA, g.inner, g.opA == OP_N ? 1 : 0,
              g.opA == OP_C ? 1 : 0, g.m, g.k, PA};
  // op(B) columns j as rows: N -> B[k + j ldb], T / C -> B[j + k ldb]
  PrepArgs pb{g.B, g.ldb, g.strideB, g.outerB, g.inner, g.opB == OP_N ? 0 : 1,
              g.opB == OP_C ? 1 : 0, g.n, g.k, PB};
  if (batchA > 65535 || batchB > 65535) return Status::TOO_LARGE;
  hipLaunchKernelGGL(emu_rowmax_kernel, dim3((unsigned)(Mp / 64), (unsigned)batchA), dim3(256), 0,
                     s, pa, c);
  hipLaunchKernelGGL(emu_rowmax_kernel, dim3((unsigned)(Np / 64), (unsigned)batchB), dim3(256), 0,
                     s, pb, c);
  hipLaunchKernelGGL(emu_resid_kernel, dim3((unsigned)(Mp / 64), (unsigned)(Kp / EKS),
                                            (unsigned)batchA), dim3(256), 0, s, pa, c);
  hipLaunchKernelGGL(emu_resid_kernel, dim3((unsigned)(Np / 64), (unsigned)(Kp / EKS),
                                            (unsigned)batchB), dim3(256), 0, s, pb, c);
  EmuGemmArgs E;
  E.g = g;
  E.A = PA;
  E.B = PB;
  E.tiles_m = (int)(Mp / 64);
  E.tiles_n = (int)(Np / 64);
  E.total = (long long)E.tiles_m * E.tiles_n * g.batch;
  E.grid_padded = (E.total + 7) / 8 * 8;
  if (E.grid_padded > 0x7fffffffLL) return Status::TOO_LARGE;
  hipLaunchKernelGGL(emu_gemm_kernel, dim3((unsigned)E.grid_padded), dim3(256), 0, s, E, c);
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

int zgemm_emu_params(int nmod, int k, int* beta) {
  EmuConsts c;
  if (!emu_consts(nmod > 0 ? nmod : zgemm_emu_default_moduli(), k, &c)) return Status::BAD_ARG;
  *beta = c.beta;
  return Status::OK;
}

}  // namespace negf
