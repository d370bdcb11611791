// Batched explicit inverse of complex128 blocks by blocked Gauss-Jordan
// elimination with partial pivoting (the diagonal-block inversion of RGF).
//
// Algorithm (per instance, n x n work matrix X, column strips of NB):
//   factoring a panel (the n x NB strip in registers of one workgroup, RPT
//   rows per thread): NB exchange steps; step j picks the pivot row p_j among
//   rows not yet pivoted by max |re| + |im| (LAPACK's izamax rule, so the
//   pivot sequence is getrf's), scales it, eliminates column j from every
//   other row (Gauss-Jordan: rows above too).  The composite row operation
//   of the NB steps acts on every other column c as
//       X[:,c] <- X[:,c] + (P - E) X[p_J, c],
//   P = the finished panel, E = unit columns at the pivot rows.
//   One launch per panel J applies that update to all other strips (each
//   strip in one workgroup: gather its NB x NB block of X[p_J, :], rank-NB
//   update in registers, store) and lets the strip of panel J+1 factor itself
//   right after its update, so the latency-bound elimination runs under the
//   HBM-bound updates of the same launch.  One pass over X per panel.
//   No physical row swaps: after all panels X[p_j, k] = inv(A)[j, p_k], and
//   the final unscramble kernel writes inv(A)[r, c] = X[p_r, q_c] with
//   q = p^{-1}.
// Flops: 8 n^3 (the nominal getrf + getri count).
#include "negf_internal.hpp"
#include <climits>

namespace negf {

namespace {

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

// Gauss-Jordan elimination of the NB columns held in registers: P[q][c] is
// row r = tid + PT * q of the strip's column c (k0 + c of the matrix).
// Pivot rule: max |re| + |im| over rows not yet pivoted, lowest row on ties
// (LAPACK's izamax, so the pivot sequence is getrf's).  Per column: each wave
// takes its maximum with DPP and finds the lowest row holding it by ballot
// (a lane's rows increase with q, so the smallest q wins, then the lowest
// lane); the winning lane publishes its whole unscaled row in the wave's LDS
// slot; one barrier; every thread picks the winning slot and eliminates.
// Slots are double-buffered by column parity, so a fast wave writing column
// j+1 never overwrites what a slow one still reads for j.
// The pivot row is scaled lazily: it keeps its unscaled values with a 1 in
// the pivot column and remembers s = 1 / pivot (its true values are s times
// the stored ones, and every later elimination step is linear in the row);
// so every row runs the same branch-free update (the pivot row with a zero
// multiplier), and the at most NB rows pivoted here are scaled once at the end.
template <int PT, int RPT, int NB>
__device__ __forceinline__ void gj_eliminate(int n, int nbe, int k0, zc (&P)[RPT][NB],
                                             bool (&usd)[RPT], int (&pc)[NB],
                                             double (&red_v)[2][PT / 64], int (&red_i)[2][PT / 64],
                                             zc (&cand)[2][PT / 64][NB], int* info_inst) {
  constexpr int NW = PT / 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  zc sc[RPT];
  bool piv_here[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    sc[q] = zmake(1.0, 0.0);
    piv_here[q] = false;
  }
#pragma unroll
  for (int jj = 0; jj < NB; ++jj) {
    pc[jj] = 0;
    if (jj < nbe) {
      const int buf = jj & 1;
      double bv = -1.0;
      int bq = 0;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int r = tid + PT * q;
        if (r < n && !usd[q]) {
          const double v = zabs1(P[q][jj]);
          if (v > bv) { bv = v; bq = q; }
        }
      }
      const double wmax = wave_max(bv);
      int win_lane = -1, win_q = 0;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const unsigned long long m = __ballot(bv >= 0.0 && bv == wmax && bq == q);
        if (win_lane < 0 && m != 0ull) {
          win_lane = __ffsll(m) - 1;
          win_q = q;
        }
      }
      if (lane == win_lane) {
#pragma unroll
        for (int q = 0; q < RPT; ++q)
          if (q == win_q) {
#pragma unroll
            for (int c = 0; c < NB; ++c) cand[buf][wave][c] = P[q][c];
          }
      }
      if (lane == 0) {
        red_v[buf][wave] = wmax;
        red_i[buf][wave] = win_lane >= 0 ? win_lane + 64 * wave + PT * win_q : INT_MAX;
      }
      __syncthreads();
      int ws = 0;
      double gv = red_v[buf][0];
      int bi = red_i[buf][0];
#pragma unroll
      for (int w = 1; w < NW; ++w) {
        const double ov = red_v[buf][w];
        const int oi = red_i[buf][w];
        if (ov > gv || (ov == gv && oi < bi)) { gv = ov; bi = oi; ws = w; }
      }
      const int p = bi;
      pc[jj] = p;
      if (tid == 0 && gv <= 0.0 && *info_inst == 0) *info_inst = k0 + jj + 1;
      const zc pv = cand[buf][ws][jj];
      const zc inv = (pv.x == 0.0 && pv.y == 0.0) ? zmake(0.0, 0.0) : zrcp(pv);
      zc cr[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) cr[c] = cand[buf][ws][c];
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int r = tid + PT * q;
        const bool isp = (r == p);
        const zc f = zmul(P[q][jj], inv);
        const zc fi = isp ? zmake(0.0, 0.0) : f;
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c != jj) {
            P[q][c].x = fma(-fi.x, cr[c].x, fma(fi.y, cr[c].y, P[q][c].x));
            P[q][c].y = fma(-fi.x, cr[c].y, fma(-fi.y, cr[c].x, P[q][c].y));
          }
        P[q][jj] = isp ? zmake(1.0, 0.0) : zmake(-fi.x, -fi.y);
        if (isp) {
          sc[q] = inv;
          usd[q] = true;
          piv_here[q] = true;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < RPT; ++q)
    if (piv_here[q]) {
#pragma unroll
      for (int c = 0; c < NB; ++c) P[q][c] = zmul(P[q][c], sc[q]);
    }
}

// One step J of the blocked Gauss-Jordan inverse, over all instances.  The
// matrix is cut into column strips of NB (strip s = columns s*NB ..).  Panel J
// (strip J) was factored by the previous launch, which left W_J = P_J - E_J
// (P_J the eliminated panel, E_J the unit columns at its pivot rows p_J) and
// the pivots.  Every other strip s gets the composite row operation of panel
// J:  X[:, s] <- X[:, s] + W_J X[p_J, s]  (its NB x NB block of pivot rows is
// gathered first, all inside this workgroup, so the in-place update has no
// race), and the strip of panel J+1, right after its own update, factors
// itself (look-ahead): its elimination overlaps the other strips' updates in
// the same launch.  The factoring workgroups come first in the grid (blocks
// 0 .. batch-1), the updates after them.  J = -1 only factors strip 0.
template <int PT, int RPT, int NB>
__global__ __launch_bounds__(PT) void gj_step_kernel(int n, int J, int nstrips, int batch,
                                                     zc* __restrict__ X,
                                                     const zc* __restrict__ Wcur,
                                                     zc* __restrict__ Wnext,
                                                     int* __restrict__ piv,
                                                     int* __restrict__ used,
                                                     int* __restrict__ info) {
  constexpr int NW = PT / 64;
  const int tid = threadIdx.x;
  const bool has_next = J + 1 < nstrips;
  // Work item k of the launch: k < batch factors instance k's next panel,
  // the rest update (instance, strip) pairs.  Workgroups are dealt to the 8
  // XCDs round-robin by block index; when batch is a multiple of 8 the items
  // are permuted so that all of one instance's strips land on one XCD (its
  // W_J is read into one L2) and every XCD starts with its factoring items.
  int k = blockIdx.x;
  if (J >= 0 && (batch & 7) == 0) {
    const int x = k & 7, kk = k >> 3, per = batch >> 3;
    if (kk < per) {
      k = x + 8 * kk;                                   // factoring item of instance x + 8 kk
    } else if (has_next) {
      const int u = kk - per;                           // update items of this XCD's instances
      k = batch + (u / per) * batch + x + 8 * (u % per);
    } else {
      const int u = kk;                                 // no factoring items in the last step
      k = (u / per) * batch + x + 8 * (u % per);
    }
  }
  int inst, s;
  bool factor;
  if (has_next && k < batch) {
    inst = k;
    s = J + 1;
    factor = true;
  } else {
    const int r = has_next ? k - batch : k;
    inst = r % batch;
    const int slot = r / batch;  // strips other than J (and J+1 when it exists)
    s = slot < J ? slot : slot + (has_next ? 2 : 1);
    factor = false;
  }
  const int k0 = s * NB;
  const int ncols = (n - k0) < NB ? (n - k0) : NB;
  zc* Xi = X + (long long)inst * n * n;

  __shared__ zc Rs[NB][NB];
  __shared__ double red_v[2][NW];
  __shared__ int red_i[2][NW];
  __shared__ zc cand[2][NW][NB];

  zc P[RPT][NB];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + PT * q;
#pragma unroll
    for (int c = 0; c < NB; ++c)
      P[q][c] = (r < n && c < ncols) ? Xi[r + (long long)(k0 + c) * n] : zmake(0.0, 0.0);
  }

  if (J >= 0) {
    const int kJ = J * NB;
    const int nbJ = (n - kJ) < NB ? (n - kJ) : NB;
    // the rows' W_J first (its latency overlaps the gather), then the NB x NB
    // block of pivot rows of this strip, then P += W_J R with R one column at
    // a time from LDS
    const zc* Wi = Wcur + (long long)inst * n * NB;
    zc w[RPT][NB];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + PT * q;
#pragma unroll
      for (int j = 0; j < NB; ++j)
        w[q][j] = (r < n && j < nbJ) ? Wi[r + (long long)j * n] : zmake(0.0, 0.0);
    }
    const int* pJ = piv + (long long)inst * n + kJ;
    for (int t = tid; t < NB * NB; t += PT) {
      const int j = t / NB, c = t % NB;
      Rs[j][c] = (j < nbJ && c < ncols) ? Xi[pJ[j] + (long long)(k0 + c) * n] : zmake(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      zc rc[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) rc[j] = Rs[j][c];
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        double re = P[q][c].x, im = P[q][c].y;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          re = fma(w[q][j].x, rc[j].x, re);
          re = fma(-w[q][j].y, rc[j].y, re);
          im = fma(w[q][j].x, rc[j].y, im);
          im = fma(w[q][j].y, rc[j].x, im);
        }
        P[q][c] = zmake(re, im);
      }
    }
  }

  if (!factor) {
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + PT * q;
      if (r < n) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c < ncols) Xi[r + (long long)(k0 + c) * n] = P[q][c];
      }
    }
    return;
  }

  int* usedi = used + (long long)inst * n;
  bool usd[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + PT * q;
    usd[q] = r < n ? usedi[r] != 0 : true;
  }
  int pc[NB];
  gj_eliminate<PT, RPT, NB>(n, ncols, k0, P, usd, pc, red_v, red_i, cand, info + inst);

  // X[:, strip] = P (its final value for this step); W_next = P - E; pivots
  zc* Wn = Wnext + (long long)inst * n * NB;
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + PT * q;
    if (r < n) {
      usedi[r] = usd[q] ? 1 : 0;
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c < ncols) {
          zc wv = P[q][c];
          if (r == pc[c]) wv.x -= 1.0;
          Wn[r + (long long)c * n] = wv;
          Xi[r + (long long)(k0 + c) * n] = P[q][c];
        }
      }
    }
  }
  if (tid == 0) {
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c < ncols) piv[(long long)inst * n + k0 + c] = pc[c];
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
  const int nstrips = (n + NB - 1) / NB;
  const long long wst = (long long)n * NB * batch;   // one W buffer
  hipLaunchKernelGGL((gj_step_kernel<PT, RPT, NB>), dim3(batch), dim3(PT), 0, s, n, -1, nstrips,
                     batch, w.X, w.W, w.W, w.piv, w.used, info);
  for (int J = 0; J < nstrips; ++J) {
    const long long blocks = (long long)batch * (nstrips - 1);
    if (blocks == 0) break;
    const zc* Wc = w.W + (J & 1) * wst;
    zc* Wn = w.W + ((J + 1) & 1) * wst;
    hipLaunchKernelGGL((gj_step_kernel<PT, RPT, NB>), dim3((unsigned)blocks), dim3(PT), 0, s, n, J,
                       nstrips, batch, w.X, Wc, Wn, w.piv, w.used, info);
  }
  hipLaunchKernelGGL(gj_unscramble_kernel, dim3((n + 63) / 64, batch), dim3(256),
                     2 * n * sizeof(int), s, n, w.X, w.piv, Ainv, ldinv, strideInv);
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

// Widest strip of any variant (the double-buffered W is sized for it).
int nb_for(int n) { return n <= 512 ? 16 : (n <= 1024 ? 8 : 4); }

// Strip variant (autotuner choice, negf_inv_variant): 0 = 16-column strips
// up to n = 512, 1 = 8-column strips (twice the passes over X, half the
// elimination work per column).  One matrix row per thread, so the strip,
// its W rows and one column of R stay in registers.
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
  b += 2 * sizeof(zc) * (size_t)n * nb * batch;  // W (two buffers: panel J and J+1)
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
  w->W = (zc*)take(2 * sizeof(zc) * (size_t)n * nb * batch);
  w->piv = (int*)take(sizeof(int) * (size_t)n * batch);
  w->used = (int*)take(sizeof(int) * (size_t)n * batch);
}

int zinv_launch(int n, const zc* A, long long lda, long long strideA, zc* Ainv, long long ldinv,
                long long strideInv, int batch, const InvWork& w, int* info, hipStream_t s) {
  if (n < 0 || batch < 0 || lda < (n > 1 ? n : 1) || ldinv < (n > 1 ? n : 1)) return Status::BAD_ARG;
  if (n == 0 || batch == 0) return Status::OK;
  if (n > 2048) return Status::TOO_LARGE;
  if (g_inv_variant == 1 && n <= 512)
    return run<512, 1, 8>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  if (n <= 256) return run<256, 1, 16>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  if (n <= 512) return run<512, 1, 16>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  if (n <= 1024) return run<1024, 1, 8>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  return run<1024, 2, 4>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
}

}  // namespace negf
