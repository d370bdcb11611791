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
#include <mutex>

namespace negf {

namespace {

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
// 1 / z = conj(z) / |z|^2 with z scaled by a power of two (exact) so that
// |z|^2 lies in [1, 8): one division on the column's critical path (Smith's
// form needs a branch and two dependent divisions); within 2 ulp per
// component.  z != 0.
__device__ __forceinline__ zc zrcp_scaled(zc z) {
  const int e = ilogb(fmax(fabs(z.x), fabs(z.y)));
  const double xs = ldexp(z.x, -e), ys = ldexp(z.y, -e);
  const double r = 1.0 / fma(xs, xs, ys * ys);
  return zmake(ldexp(xs * r, -e), ldexp(-ys * r, -e));
}

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
      // all waves' (maximum, row) pairs in one round of LDS reads, then a
      // branch-free pick (so the reads are not serialised behind compares)
      double rv[NW];
      int rix[NW];
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        rv[w] = red_v[buf][w];
        rix[w] = red_i[buf][w];
      }
      // the largest value, then the lowest row holding it (a max and a min
      // over the slots: compares chained on the running pick compile to
      // scalar branches, the values being uniform); the winner's slot is its
      // row's wave
      double gv = rv[0];
#pragma unroll
      for (int w = 1; w < NW; ++w) gv = fmax(gv, rv[w]);
      int bi = INT_MAX;
#pragma unroll
      for (int w = 0; w < NW; ++w) bi = min(bi, rv[w] == gv ? rix[w] : INT_MAX);
      const int ws = bi == INT_MAX ? 0 : (bi % PT) >> 6;
      const int p = bi;
      pc[jj] = p;
      if (tid == 0 && gv <= 0.0 && *info_inst == 0) *info_inst = k0 + jj + 1;
      // the winning row in one round of reads, its pivot entry among them
      zc cr[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) cr[c] = cand[buf][ws][c];
      const zc pv = cr[jj];
      const zc inv = (pv.x == 0.0 && pv.y == 0.0) ? zmake(0.0, 0.0) : zrcp_scaled(pv);
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

template <int PT, int RPT, int S>
__device__ __forceinline__ void load_strip(const zc* __restrict__ Xi, int n, int k0, int ncols,
                                           zc (&P)[RPT][S]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + PT * q;
#pragma unroll
    for (int c = 0; c < S; ++c)
      P[q][c] = (r < n && c < ncols) ? Xi[r + (long long)(k0 + c) * n] : zmake(0.0, 0.0);
  }
}

template <int PT, int RPT, int S>
__device__ __forceinline__ void store_strip(zc* __restrict__ Xi, int n, int k0, int ncols,
                                            const zc (&P)[RPT][S]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + PT * q;
    if (r < n) {
#pragma unroll
      for (int c = 0; c < S; ++c)
        if (c < ncols) Xi[r + (long long)(k0 + c) * n] = P[q][c];
    }
  }
}

// Row operation of one factored strip t applied to the strip held in P:
//   P <- P + W_t[rows, :] P[p_t, :]
// (W_t = P_t - E_t of that strip's factoring, n x S at Wt; p_t its wt pivot
// rows).  The pivot rows are taken from the current P: their owners copy them
// to LDS, then every row adds its W_t row times that S x S block, R one
// column at a time from LDS.  All threads of the workgroup must call it.
template <int PT, int RPT, int S>
__device__ __forceinline__ void apply_transform(int n, zc (&P)[RPT][S], const zc* __restrict__ Wt,
                                                const int* __restrict__ pt, int wt,
                                                zc (&Rs)[S][S]) {
  const int tid = threadIdx.x;
  // all wt pivots loaded at once (unconditional loads of in-range indices, so
  // the compiler issues them together instead of one wait per pivot)
  int pj[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int v = pt[j < wt ? j : 0];
    pj[j] = j < wt ? v : -1;
  }
  zc w[RPT][S];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + PT * q;
#pragma unroll
    for (int j = 0; j < S; ++j)
      w[q][j] = (r < n && j < wt) ? Wt[r + (long long)j * n] : zmake(0.0, 0.0);
  }
  __syncthreads();  // Rs is free
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = tid + PT * q;
#pragma unroll
    for (int j = 0; j < S; ++j)
      if (r == pj[j]) {
#pragma unroll
        for (int c = 0; c < S; ++c) Rs[j][c] = P[q][c];
      }
  }
  for (int t = wt * S + tid; t < S * S; t += PT) Rs[t / S][t % S] = zmake(0.0, 0.0);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < S; ++c) {
    zc rc[S];
#pragma unroll
    for (int j = 0; j < S; ++j) rc[j] = Rs[j][c];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      double re = P[q][c].x, im = P[q][c].y;
#pragma unroll
      for (int j = 0; j < S; ++j) {
        re = fma(w[q][j].x, rc[j].x, re);
        re = fma(-w[q][j].y, rc[j].y, re);
        im = fma(w[q][j].x, rc[j].y, im);
        im = fma(w[q][j].y, rc[j].x, im);
      }
      P[q][c] = zmake(re, im);
    }
  }
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// One launch of the blocked Gauss-Jordan inverse, over all instances.  The
// matrix is cut into column strips of S (strip s = columns s*S ..) and the
// strips into panels of PS consecutive strips.  Panel K was factored by the
// previous launch: for each of its strips t, W_t = P_t - E_t and the pivots
// p_t, where strip t's factoring already included the row operations of the
// panel's earlier strips.  This launch:
//  * update items: every strip outside panels K and K+1 gets the row
//    operations of panel K's strips in order, X[:, s] <- X[:, s] + W_t X[p_t, s]
//    (the pivot rows of each step taken from the strip as updated so far, all
//    in this workgroup, so the in-place update has no race);
//  * correction item (PS = 2): strip A of panel K gets strip B's operation,
//    the one it was factored without;
//  * factoring item (look-ahead, when panel K+1 exists): each strip of panel
//    K+1 gets panel K's operations and those of panel K+1's earlier strips,
//    then is eliminated; its elimination runs under the other items' HBM-bound
//    updates.
// One pass over X per panel, so PS = 2 halves the passes at the same
// registers (a strip is never wider than S).  K = -1 only factors panel 0.
template <int PT, int RPT, int S, int PS>
__global__ __launch_bounds__(PT) void gj_step_kernel(int n, int K, int nstrips, int batch,
                                                     zc* __restrict__ X,
                                                     const zc* __restrict__ Wcur,
                                                     zc* __restrict__ Wnext,
                                                     int* __restrict__ piv,
                                                     int* __restrict__ used,
                                                     int* __restrict__ info, int prev_applied) {
  constexpr int NW = PT / 64;
  const int tid = threadIdx.x;
  const int pk0 = PS * K, pk1 = (pk0 + PS) < nstrips ? pk0 + PS : nstrips;   // panel K
  const int nk0 = PS * (K + 1);
  const int nk1 = (nk0 + PS) < nstrips ? nk0 + PS : nstrips;                // panel K+1
  const bool has_next = nk0 < nstrips;
  const int nF = has_next ? 1 : 0;
  const int nC = (K >= 0 && PS == 2 && pk1 - pk0 == 2) ? 1 : 0;
  const int lo = pk0, hi = has_next ? nk1 : pk1;   // strips not updated as plain strips
  // Work item k of the launch: [factoring: batch][correction: batch][updates:
  // (strip slot, instance)].  Workgroups are dealt to the 8 XCDs round-robin
  // by block index; when batch is a multiple of 8 the items are permuted so
  // that all of one instance's items land on one XCD (its W stays in one L2)
  // and every XCD starts with its factoring items.
  int k = blockIdx.x;
  if (K >= 0 && (batch & 7) == 0) {
    const int x = k & 7, kk = k >> 3, per = batch >> 3;
    k = (kk / per) * batch + x + 8 * (kk % per);
  }
  const int cat = k / batch, inst = k - cat * batch;
  zc* Xi = X + (long long)inst * n * n;

  __shared__ zc Rs[S][S];
  __shared__ double red_v[2][NW];
  __shared__ int red_i[2][NW];
  __shared__ zc cand[2][NW][S];

  auto W_of = [&](const zc* base, int t_in_panel) {
    return base + ((long long)inst * PS + t_in_panel) * n * S;
  };
  auto width = [&](int t) { return (n - t * S) < S ? (n - t * S) : S; };

  zc P[RPT][S];
  if (cat >= nF) {
    int s, t0;
    if (cat < nF + nC) {   // correction of panel K's strip A: strip B's operation only
      s = pk0;
      t0 = pk0 + 1;
    } else {
      const int slot = cat - nF - nC;
      s = slot < lo ? slot : slot + (hi - lo);
      t0 = pk0;
    }
    const int k0 = s * S, ncols = width(s);
    load_strip<PT, RPT, S>(Xi, n, k0, ncols, P);
    for (int t = t0; t < pk1; ++t)
      apply_transform<PT, RPT, S>(n, P, W_of(Wcur, t - pk0), piv + (long long)inst * n + t * S,
                                  width(t), Rs);
    store_strip<PT, RPT, S>(Xi, n, k0, ncols, P);
    return;
  }

  // factoring item: the strips of panel K+1 in order
  int* usedi = used + (long long)inst * n;
  for (int s = nk0; s < nk1; ++s) {
    const int k0 = s * S, ncols = width(s);
    load_strip<PT, RPT, S>(Xi, n, k0, ncols, P);
    for (int t = pk0; K >= 0 && !prev_applied && t < pk1; ++t)   // panel K (none for K = -1)
      apply_transform<PT, RPT, S>(n, P, W_of(Wcur, t - pk0), piv + (long long)inst * n + t * S,
                                  width(t), Rs);
    for (int t = nk0; t < s; ++t)     // panel K+1's earlier strips
      apply_transform<PT, RPT, S>(n, P, W_of(Wnext, t - nk0), piv + (long long)inst * n + t * S,
                                  width(t), Rs);
    bool usd[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + PT * q;
      usd[q] = r < n ? usedi[r] != 0 : true;
    }
    int pc[S];
    gj_eliminate<PT, RPT, S>(n, ncols, k0, P, usd, pc, red_v, red_i, cand, info + inst);
    // X[:, strip] = P; W = P - E; pivots and used flags
    zc* Wn = Wnext + ((long long)inst * PS + (s - nk0)) * n * S;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + PT * q;
      if (r < n) {
        usedi[r] = usd[q] ? 1 : 0;
#pragma unroll
        for (int c = 0; c < S; ++c) {
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
      for (int c = 0; c < S; ++c)
        if (c < ncols) piv[(long long)inst * n + k0 + c] = pc[c];
    }
    // the next strip's transforms read these pivots (written by thread 0)
    __syncthreads();
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

// out[r, c] = X[p_r, q_c], q = p^{-1}; grid.x = column chunks of 64, grid.y =
// instance.  With fro: fro[inst * gridDim.x + chunk] = sum of |out|^2 over the
// chunk (fixed order).
__global__ __launch_bounds__(256) void gj_unscramble_kernel(int n, const zc* __restrict__ X,
                                                            const int* __restrict__ piv,
                                                            zc* __restrict__ out, long long ldo,
                                                            long long strideO,
                                                            double* __restrict__ fro) {
  extern __shared__ int sh[];
  __shared__ double red[256];
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
  double acc = 0.0;
  for (int c = c0; c < c1; ++c) {
    const zc* col = Xi + (long long)qinv[c] * n;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const zc v = col[p[r]];
      Oi[r + c * ldo] = v;
      acc = fma(v.x, v.x, fma(v.y, v.y, acc));
    }
  }
  if (fro) {   // uniform over the workgroup
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) fro[(long long)inst * gridDim.x + blockIdx.x] = red[0];
  }
}

// Item decoding shared by the prep and update kernels of one split step:
// per instance the strips outside panels K and K+1 (all of panel K's
// operations, t from pk0), then panel K's own strips (strip pk0 + u gets the
// operations of the later strips of its panel; none = a copy).
struct SplitStep {
  int pk0, pk1, hi, nupd, items;
  __host__ __device__ SplitStep(int K, int PS, int ns) {
    pk0 = PS * K;
    pk1 = (pk0 + PS) < ns ? pk0 + PS : ns;
    const int nk0 = PS * (K + 1), nk1 = (nk0 + PS) < ns ? nk0 + PS : ns;
    hi = nk0 < ns ? nk1 : pk1;     // strips [pk0, hi) are not plain update strips
    nupd = ns - (hi - pk0);
    items = nupd + (pk1 - pk0);
  }
  __device__ void item(int it, int* s, int* t_lo) const {
    if (it < nupd) {
      *s = it < pk0 ? it : it + (hi - pk0);
      *t_lo = pk0;
    } else {
      *s = pk0 + (it - nupd);
      *t_lo = *s + 1;
    }
  }
};

constexpr int MAXPS = 4;

// Update items of a split step, in place, one workgroup per (instance,
// strip) owning the whole strip (wave w: rows 64 w .. 64 w + 63, n <= PT).
// Strip s gets the row operations t = t_lo .. pk1 - 1 of panel K in order,
//   X[:, s] += sum_t W_t R_t,  R_t = X'[p_t, s] (X' = X after the operations
//   before t) = X[p_t, s] + sum_{t' < t} W_t'[p_t, :] R_t'.
// The pivot rows come from the registers of the waves that own them (through
// LDS), the chain is formed in LDS, and each wave applies the operations to
// its rows on the matrix cores: rows held transposed in the accumulator
// layout of v_mfma_f64_16x16x4_f64 (lane l, element e of block b: row
// 64 w + 16 b + l % 16, column l / 16 + 4 e), four real products per four
// complex k, step m covering k = 4 m + l / 16 so that each lane loads whole
// complex numbers of W and R and no two lanes the same one; W and the strip
// go through buffer operations (one VGPR of offset per lane) so that 4 waves
// per SIMD fit in 128 VGPRs.  Items per
// instance (next_panel = 0): the strips outside panels K and K+1 (all of
// panel K's operations), then panel K's strips but its last (the operations
// of the later strips of the panel); this runs beside the factoring of panel
// K+1.  With next_panel = 1: panel K+1's strips (all of panel K's
// operations), launched just before that factoring.  Four waves per SIMD
// (128 VGPRs) is the target at PT = 512, the sweep's n = 512; the 256-thread
// form (n <= 256, selected schedules only) is left to the compiler.
template <int PT>
__global__ __launch_bounds__(PT) __attribute__((amdgpu_waves_per_eu(PT == 512 ? 4 : 1))) void
gj_strip_update_kernel(int n, int K, int PS, int ns, int batch, int next_panel, int it_base,
                       int skip_at, int skip_len, zc* __restrict__ X,
                       const zc* __restrict__ Wcur, const int* __restrict__ piv) {
  constexpr int S = 16;
  const SplitStep st(K, PS, ns);
  // Workgroups are dealt to the 8 XCDs round-robin by block index.  With
  // batch a multiple of 8, XCD x takes instances x, x + 8, ... and runs
  // through all items of one instance before the next, so the workgroups in
  // flight on an XCD share the W of two or three instances in its L2 (with
  // instances fastest, all of its instances' W, 6 MB at 96 x 512, would
  // cycle through the 4 MB L2).
  const long long k = blockIdx.x;
  int itl, inst;
  if ((batch & 7) == 0) {
    const long long nit = (long long)gridDim.x / batch;   // items per instance
    const long long x = k & 7, kk = k >> 3;
    itl = (int)(kk % nit);
    inst = (int)(x + 8 * (kk / nit));
  } else {
    itl = (int)(k / batch);
    inst = (int)(k - (long long)itl * batch);
  }
  int s, t_lo;
  if (next_panel) {   // panel K+1's strips, ahead of their factoring
    s = st.pk1 + itl;
    t_lo = st.pk0;
  } else {            // items it_base + [0, skip_at) and it_base + skip_at + skip_len + ...
    st.item(it_base + (itl < skip_at ? itl : itl + skip_len), &s, &t_lo);
  }
  const int nt = st.pk1 - t_lo;   // >= 1 for every launched item
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int k0 = s * S, ncols = (n - k0) < S ? (n - k0) : S;
  zc* __restrict__ Xi = X + (long long)inst * n * n;
  const int* __restrict__ pv = piv + (long long)inst * n;
  const zc* __restrict__ Wp = Wcur + (long long)inst * PS * n * S;
  // rows padded to 17 entries: 16 threads touch 16 rows of one column.
  // Wg[u (u - 1) / 2 + v][j][q] = W_v[p_u[j], q] for the chain pairs v < u.
  constexpr int NPAIR = MAXPS * (MAXPS - 1) / 2;
  __shared__ zc Rs[MAXPS][S][S + 1];
  __shared__ zc Wg[NPAIR][S][S + 1];
  __shared__ int pjs[MAXPS][S];
  __shared__ int rowslot[PT];   // rowslot[r] = 16 u + j if row r is pivot j of operation u
  for (int x = tid; x < MAXPS * S; x += PT) {
    const int u = x / S, j = x - u * S, t = t_lo + u;
    const int wt = (n - t * S) < S ? (n - t * S) : S;
    pjs[u][j] = (u < nt && j < wt) ? pv[t * S + j] : -1;
  }
  rowslot[tid] = -1;
  for (int x = tid; x < MAXPS * S * S; x += PT) {
    const int u = x / (S * S), r = x - u * S * S;
    Rs[u][r / S][r % S] = zmake(0.0, 0.0);
  }
  __syncthreads();
  // every W block of the chain, issued together with the strip loads below
  const int npair = nt * (nt - 1) / 2;
  for (int x = tid; x < npair * S * S; x += PT) {
    const int pr = x / (S * S), r = x - pr * S * S, j = r & 15, q = r >> 4;
    int u = 1;
    while ((u + 1) * u / 2 <= pr) ++u;
    const int v = pr - u * (u - 1) / 2;
    const int tv = t_lo + v, wtv = (n - tv * S) < S ? (n - tv * S) : S;
    const int row = pjs[u][j];
    Wg[pr][j][q] = (row >= 0 && q < wtv) ? Wp[(tv - st.pk0) * n * S + row + q * n] : zmake(0.0, 0.0);
  }
  // the strip through buffer operations (per-lane offset of row 64 w + l % 16,
  // column k0 + l / 16; block b in the immediate, column group 4 e scalar)
  const __amdgpu_buffer_rsrc_t xs =
      __builtin_amdgcn_make_buffer_rsrc((void*)Xi, (short)0, n * n * (int)sizeof(zc), 0x00020000);
  const int xoff = (wave * 64 + li + (k0 + lg) * n) * (int)sizeof(zc);
  if (tid < MAXPS * S) {
    const int pr = pjs[tid / S][tid % S];
    if (pr >= 0) rowslot[pr] = tid;
  }
  dbl4 are[4], aim[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const bool rok = wave * 64 + b * 16 + li < n;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const bool ok = rok && lg + 4 * e < ncols;
      const auto raw = __builtin_amdgcn_raw_buffer_load_b128(
          xs, xoff + b * 16 * (int)sizeof(zc), ok ? 4 * e * n * (int)sizeof(zc) : 0, 0);
      zc v;
      __builtin_memcpy(&v, &raw, sizeof(zc));
      are[b][e] = ok ? v.x : 0.0;
      aim[b][e] = ok ? v.y : 0.0;
    }
  }
  __syncthreads();
  // the owners of the pivot rows publish them (values before this step)
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int r = wave * 64 + b * 16 + li;
    const int sl = r < n ? rowslot[r] : -1;
    if (sl >= 0) {
#pragma unroll
      for (int e = 0; e < 4; ++e) Rs[sl / S][sl % S][lg + 4 * e] = zmake(are[b][e], aim[b][e]);
    }
  }
  __syncthreads();
  // chain: block u gets the operations of blocks 0 .. u-1 (final by then)
  for (int u = 1; u < nt; ++u) {
    if (tid < S * S) {
      const int j = tid & 15, c = tid >> 4;
      double re = Rs[u][j][c].x, im = Rs[u][j][c].y;
      for (int v = 0; v < u; ++v) {
        const int pr = u * (u - 1) / 2 + v;
#pragma unroll
        for (int q = 0; q < S; ++q) {
          const zc wq = Wg[pr][j][q];
          const zc rq = Rs[v][q][c];
          re = fma(wq.x, rq.x, fma(-wq.y, rq.y, re));
          im = fma(wq.x, rq.y, fma(wq.y, rq.x, im));
        }
      }
      Rs[u][j][c] = zmake(re, im);
    }
    __syncthreads();
  }
  // W_t through buffer loads: one VGPR of per-lane byte offset (row
  // 64 w + l % 16, column l / 16), the 16-row block in the immediate offset
  // and the column group 4 m in the scalar offset, so the 16 loads of a
  // transform need no 64-bit address each (those spilled at 128 VGPRs).
  const int voff = (wave * 64 + li + lg * n) * (int)sizeof(zc);
  // W of (operation u, block b) loaded one block ahead of its products
  auto load_w = [&](int u, int b, zc (&w)[4]) {
    const int t = t_lo + u;
    const int wt = (n - t * S) < S ? (n - t * S) : S;
    const zc* __restrict__ Wt = Wp + (t - st.pk0) * n * S;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void*)Wt, (short)0, n * S * (int)sizeof(zc), 0x00020000);
    const bool rok = wave * 64 + b * 16 + li < n;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const auto raw = __builtin_amdgcn_raw_buffer_load_b128(rs, voff + b * 16 * (int)sizeof(zc),
                                                             4 * m * n * (int)sizeof(zc), 0);
      zc v;
      __builtin_memcpy(&v, &raw, sizeof(zc));
      w[m] = (rok && 4 * m + lg < wt) ? v : zmake(0.0, 0.0);
    }
  };
  zc wcur[4], wnxt[4];
  load_w(0, 0, wcur);
  for (int u = 0; u < nt; ++u) {
    zc rr[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) rr[m] = Rs[u][4 * m + lg][li];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b < 3)
        load_w(u, b + 1, wnxt);
      else if (u + 1 < nt)
        load_w(u + 1, 0, wnxt);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        are[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(rr[m].x, wcur[m].x, are[b], 0, 0, 0);
        aim[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(rr[m].x, wcur[m].y, aim[b], 0, 0, 0);
        are[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-rr[m].y, wcur[m].y, are[b], 0, 0, 0);
        aim[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(rr[m].y, wcur[m].x, aim[b], 0, 0, 0);
      }
#pragma unroll
      for (int m = 0; m < 4; ++m) wcur[m] = wnxt[m];
    }
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const bool rok = wave * 64 + b * 16 + li < n;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (rok && lg + 4 * e < ncols) {
        const zc v = zmake(are[b][e], aim[b][e]);
        __attribute__((ext_vector_type(4))) unsigned int raw;
        __builtin_memcpy(&raw, &v, sizeof(zc));
        __builtin_amdgcn_raw_buffer_store_b128(raw, xs, xoff + b * 16 * (int)sizeof(zc),
                                               4 * e * n * (int)sizeof(zc), 0);
      }
    }
  }
}

// Pivot rows of panel K in the bulk strips, chained (variant 6): for every
// strip s outside panels K, K+1 and K+2 (left of panel K, and right of panel
// K+2), R[u S + j, s S + c] = R_u[j, c] with
//   R_u = X[p_u, s] + sum_{v < u} W_v[p_u, :] R_v
// (p_u the pivots of panel K's strip u, W_v its operation), so that the
// whole panel's operation on those columns is one product,
//   X[:, bulk] += [W_0 .. W_{PS-1}] R[:, bulk]
// (n x PS S times PS S x columns), done by the batched GEMM next.  R is a
// separate buffer, so the in-place product never reads rows it has written.
// One 256-thread workgroup per (instance, strip); full strips only.
__global__ __launch_bounds__(256) void gj_bulk_chain_kernel(int n, int K, int PS, int batch,
                                                             int nleft, int rstart,
                                                             const zc* __restrict__ X,
                                                             const zc* __restrict__ Wcur,
                                                             const int* __restrict__ piv,
                                                             zc* __restrict__ R) {
  constexpr int S = 16;
  constexpr int NPAIR = MAXPS * (MAXPS - 1) / 2;
  // items instance-fastest within each XCD's share, as the update kernel
  const long long k = blockIdx.x;
  int itl, inst;
  if ((batch & 7) == 0) {
    const long long nit = (long long)gridDim.x / batch;
    const long long x = k & 7, kk = k >> 3;
    itl = (int)(kk % nit);
    inst = (int)(x + 8 * (kk / nit));
  } else {
    itl = (int)(k / batch);
    inst = (int)(k - (long long)itl * batch);
  }
  const int s = itl < nleft ? itl : rstart + (itl - nleft);
  const int pk0 = PS * K, k0 = s * S;
  const int tid = threadIdx.x;
  const zc* __restrict__ Xi = X + (long long)inst * n * n;
  const zc* __restrict__ Wp = Wcur + (long long)inst * PS * n * S;
  __shared__ int pjs[MAXPS][S];
  __shared__ zc Rs[MAXPS][S][S + 1];
  __shared__ zc Wg[NPAIR][S][S + 1];
  if (tid < PS * S) pjs[tid / S][tid % S] = piv[(long long)inst * n + pk0 * S + tid];
  __syncthreads();
  // the pivot rows of the strip, and the W rows at the pivots of later strips
  for (int x = tid; x < PS * S * S; x += 256) {
    const int u = x / (S * S), r = x - u * S * S, j = r & 15, c = r >> 4;
    Rs[u][j][c] = Xi[pjs[u][j] + (long long)(k0 + c) * n];
  }
  const int npair = PS * (PS - 1) / 2;
  for (int x = tid; x < npair * S * S; x += 256) {
    const int pr = x / (S * S), r = x - pr * S * S, j = r & 15, q = r >> 4;
    int u = 1;
    while ((u + 1) * u / 2 <= pr) ++u;
    const int v = pr - u * (u - 1) / 2;
    Wg[pr][j][q] = Wp[(long long)v * n * S + pjs[u][j] + (long long)q * n];
  }
  __syncthreads();
  // the chain, block u after blocks 0 .. u-1
  for (int u = 1; u < PS; ++u) {
    const int j = tid & 15, c = tid >> 4;
    double re = Rs[u][j][c].x, im = Rs[u][j][c].y;
    for (int v = 0; v < u; ++v) {
      const int pr = u * (u - 1) / 2 + v;
#pragma unroll
      for (int q = 0; q < S; ++q) {
        const zc wq = Wg[pr][j][q];
        const zc rq = Rs[v][q][c];
        re = fma(wq.x, rq.x, fma(-wq.y, rq.y, re));
        im = fma(wq.x, rq.y, fma(wq.y, rq.x, im));
      }
    }
    Rs[u][j][c] = zmake(re, im);
    __syncthreads();
  }
  // R is (PS S) x n per instance, column-major
  zc* __restrict__ Ri = R + (long long)inst * PS * S * n;
  for (int x = tid; x < PS * S * S; x += 256) {
    const int c = x / (PS * S), kk = x - c * PS * S;
    Ri[kk + (long long)(k0 + c) * PS * S] = Rs[kk / S][kk % S][c];
  }
}

// Second stream per calling stream for the split steps (update kernel beside
// the factoring kernel), with its fork / join events; created on first use,
// default priority, non-blocking.  A slot lives until zinv_aux_release_all:
// keyed by the stream's handle, it is looked up again on every call from that
// stream.  When the table is full a new calling stream gets an error, never
// another stream's slot (two callers recording and waiting on one fork event
// from different host threads could pair the wrong fork with a launch).
struct AuxSlot {
  int device;
  hipStream_t key;
  InvAux a;
};
constexpr int MAX_AUX = 64;
AuxSlot g_aux[MAX_AUX];
int g_naux = 0;
std::mutex g_aux_mu;

int aux_for(hipStream_t s, InvAux* out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return Status::HIP_ERROR;
  std::lock_guard<std::mutex> lk(g_aux_mu);
  for (int i = 0; i < g_naux; ++i)
    if (g_aux[i].device == dev && g_aux[i].key == s) {
      *out = g_aux[i].a;
      return Status::OK;
    }
  if (g_naux == MAX_AUX) return Status::HIP_ERROR;
  AuxSlot a{dev, s, {nullptr, nullptr, nullptr}};
  const int rc = zinv_aux_create(&a.a);
  if (rc) return rc;
  g_aux[g_naux++] = a;
  *out = a.a;
  return Status::OK;
}

// Split steps in place.  Per panel K, on the caller's stream: panel K's
// operations applied to panel K+1's strips on the matrix cores
// (gj_strip_update_kernel, next_panel = 1), then the factoring of panel K+1
// (gj_step_kernel, one workgroup per instance, with panel K's operations
// already applied).  On a second stream, forked after that first launch so
// it does not queue ahead of it: panel K's operations on panel K+2's strips
// (part A), an event, then on every other strip (part B).  The next step's
// first launch waits for part A only, so part B overlaps it and the next
// factoring.  The two streams touch disjoint strips at any time; the
// factoring of panel K+2 writes its W while part B of step K still reads
// panel K's, hence three W buffers.  The second stream is joined at the end.
template <int PT, int PS>
int run_split(int n, const zc* A, long long lda, long long strideA, zc* Ainv, long long ldinv,
              long long strideInv, int batch, const InvWork& w, int* info, hipStream_t s,
              bool bulk_gemm = false) {
  constexpr int S = 16;
  InvAux axv = w.aux;
  int rc = Status::OK;
  if (!axv.aux || !axv.fork || !axv.join) rc = aux_for(s, &axv);
  if (rc) return rc;
  const InvAux* ax = &axv;
  if (A != w.X || lda != n || strideA != (long long)n * n) {
    rc = zcopy2d_launch(n, n, A, lda, strideA, w.X, n, (long long)n * n, batch, s);
    if (rc) return rc;
  }
  {
    long long total = (long long)n * batch;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(gj_init_kernel, dim3(blocks), dim3(256), 0, s, n, w.used, info, batch);
  }
  const int ns = (n + S - 1) / S;
  const int np = (ns + PS - 1) / PS;
  const long long wst = (long long)n * S * PS * batch;
  hipLaunchKernelGGL((gj_step_kernel<PT, 1, S, PS>), dim3(batch), dim3(PT), 0, s, n, -1, ns,
                     batch, w.X, w.W + wst, w.W, w.piv, w.used, info, 0);
  bool partA_pending = false;   // the join event carries the last part A
  for (int K = 0; K < np; ++K) {
    const bool has_next = PS * (K + 1) < ns;
    const SplitStep st(K, PS, ns);
    const int items = st.items - 1;   // panel K's last strip has no later operation
    const zc* Wc = w.W + (K % 3) * wst;
    zc* Wn = w.W + ((K + 1) % 3) * wst;
    // panel K+1's strips got panel K-1's operations in the last part A
    if (partA_pending && hipStreamWaitEvent(s, ax->join, 0) != hipSuccess) return Status::HIP_ERROR;
    partA_pending = false;
    const int nnext = has_next ? (st.pk1 + PS < ns ? st.pk1 + PS : ns) - st.pk1 : 0;
    if (has_next)
      hipLaunchKernelGGL((gj_strip_update_kernel<PT>), dim3((unsigned)((long long)batch * nnext)),
                         dim3(PT), 0, s, n, K, PS, ns, batch, 1, 0, 0, 0, w.X, Wc, w.piv);
    if (hipEventRecord(ax->fork, s) != hipSuccess || hipStreamWaitEvent(ax->aux, ax->fork, 0) != hipSuccess)
      return Status::HIP_ERROR;
    if (has_next)
      hipLaunchKernelGGL((gj_step_kernel<PT, 1, S, PS>), dim3(batch), dim3(PT), 0, s, n, K, ns,
                         batch, w.X, Wc, Wn, w.piv, w.used, info, 1);
    // part A: panel K+2's strips, items pk0 .. pk0 + nA - 1 (the first plain
    // strips after panel K+1)
    const int nA = has_next ? (ns - st.hi < PS ? ns - st.hi : PS) : 0;
    if (nA > 0) {
      hipLaunchKernelGGL((gj_strip_update_kernel<PT>), dim3((unsigned)((long long)batch * nA)),
                         dim3(PT), 0, ax->aux, n, K, PS, ns, batch, 0, st.pk0, INT_MAX, 0, w.X,
                         Wc, w.piv);
      if (hipEventRecord(ax->join, ax->aux) != hipSuccess) return Status::HIP_ERROR;
      partA_pending = true;
    }
    if (!bulk_gemm) {
      // part B: every other item
      const int nB = items - nA;
      if (nB > 0)
        hipLaunchKernelGGL((gj_strip_update_kernel<PT>), dim3((unsigned)((long long)batch * nB)),
                           dim3(PT), 0, ax->aux, n, K, PS, ns, batch, 0, 0, st.pk0, nA, w.X, Wc,
                           w.piv);
      continue;
    }
    // part B as one product: the strips left of panel K and right of panel
    // K+2 get panel K's whole operation, X[:, bulk] += W R[:, bulk], with R
    // the chained pivot rows (gj_bulk_chain_kernel); then panel K's own
    // strips get the operations of their panel's later strips
    const int nleft = st.pk0, rstart = st.hi + nA, nright = ns - rstart;
    if (nleft + nright > 0) {
      hipLaunchKernelGGL(gj_bulk_chain_kernel, dim3((unsigned)((long long)batch * (nleft + nright))),
                         dim3(256), 0, ax->aux, n, K, PS, batch, nleft, rstart, w.X, Wc, w.piv,
                         w.R);
      for (int side = 0; side < 2; ++side) {
        const int c0 = side == 0 ? 0 : rstart * S, nc = (side == 0 ? nleft : nright) * S;
        if (nc == 0) continue;
        GemmArgs g{};
        g.opA = OP_N;
        g.opB = OP_N;
        g.m = n;
        g.n = nc;
        g.k = PS * S;
        g.alpha = zmake(1.0, 0.0);
        g.beta = zmake(1.0, 0.0);
        g.A = Wc;
        g.lda = n;
        g.strideA = (long long)PS * n * S;
        g.B = w.R + (long long)c0 * PS * S;
        g.ldb = PS * S;
        g.strideB = (long long)PS * S * n;
        g.C = w.X + (long long)c0 * n;
        g.ldc = n;
        g.strideC = (long long)n * n;
        g.D = w.X + (long long)c0 * n;
        g.ldd = n;
        g.strideD = (long long)n * n;
        g.batch = batch;
        rc = zgemm_launch(g, ax->aux);
        if (rc) return rc;
      }
    }
    const int ncorr = items - st.nupd;
    if (ncorr > 0)
      hipLaunchKernelGGL((gj_strip_update_kernel<PT>), dim3((unsigned)((long long)batch * ncorr)),
                         dim3(PT), 0, ax->aux, n, K, PS, ns, batch, 0, st.nupd, INT_MAX, 0, w.X,
                         Wc, w.piv);
  }
  if (hipEventRecord(ax->join, ax->aux) != hipSuccess || hipStreamWaitEvent(s, ax->join, 0) != hipSuccess)
    return Status::HIP_ERROR;
  hipLaunchKernelGGL(gj_unscramble_kernel, dim3((n + 63) / 64, batch), dim3(256),
                     2 * n * sizeof(int), s, n, w.X, w.piv, Ainv, ldinv, strideInv, w.fro);
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

template <int PT, int RPT, int S, int PS>
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
  const int ns = (n + S - 1) / S;
  const int np = (ns + PS - 1) / PS;
  const long long wst = (long long)n * S * PS * batch;   // one W buffer (a panel)
  hipLaunchKernelGGL((gj_step_kernel<PT, RPT, S, PS>), dim3(batch), dim3(PT), 0, s, n, -1, ns,
                     batch, w.X, w.W + wst, w.W, w.piv, w.used, info, 0);
  for (int K = 0; K < np; ++K) {
    const int pk0 = PS * K, pk1 = (pk0 + PS) < ns ? pk0 + PS : ns;
    const int nk0 = PS * (K + 1), nk1 = (nk0 + PS) < ns ? nk0 + PS : ns;
    const bool has_next = nk0 < ns;
    const int nF = has_next ? 1 : 0;
    const int nC = (PS == 2 && pk1 - pk0 == 2) ? 1 : 0;
    const int hi = has_next ? nk1 : pk1;
    const int m = nF + nC + (ns - (hi - pk0));
    if (m == 0) continue;
    const zc* Wc = w.W + (K & 1) * wst;
    zc* Wn = w.W + ((K + 1) & 1) * wst;
    hipLaunchKernelGGL((gj_step_kernel<PT, RPT, S, PS>), dim3((unsigned)((long long)batch * m)),
                       dim3(PT), 0, s, n, K, ns, batch, w.X, Wc, Wn, w.piv, w.used, info, 0);
  }
  hipLaunchKernelGGL(gj_unscramble_kernel, dim3((n + 63) / 64, batch), dim3(256),
                     2 * n * sizeof(int), s, n, w.X, w.piv, Ainv, ldinv, strideInv, w.fro);
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

// Widest panel (strips x strip width) of any variant: the two W buffers hold
// one panel each.
int nb_for(int n) { return n <= 512 ? 64 : (n <= 1024 ? 16 : 8); }

// Variant (autotuner choice, negf_inv_variant).  In place, one launch per
// panel with the update items in gj_step_kernel: 1 = one-strip panels, 2 =
// two-strip panels.  Split steps (run_split, strips of 16, n <= 512): 3, 4,
// 5 = panels of 1, 2, 4 strips; 6 = panels of 4 strips with the strips
// outside panels K .. K+2 updated by gj_bulk_chain_kernel and a batched GEMM
// (n a multiple of 64, else as 5).  0 = automatic: split with four-strip panels
// for 256 < n <= 512 and at least 32 instances, else in place with two-strip
// panels from 32 instances up, one-strip below.  Measured on MI355X (ms per
// batched inverse): 64 x 512: split/4 2.98, split/2 3.80, in place/2 3.90,
// in place/1 5.39; 1 x 512: in place/1 1.48, in place/2 1.75, split/4 1.96;
// 64 x 256: in place/2 0.79, split/2 0.84, split/4 0.85.  Strips are 16
// columns up to n = 512, 8 up to 1024, 4 beyond; one matrix row per thread up
// to n = 1024 (the factoring keeps a strip, its W rows and one column of R in
// registers).
int g_inv_variant = 0;

}  // namespace

int zinv_aux_create(InvAux* a) {
  a->aux = nullptr;
  a->fork = a->join = nullptr;
  if (hipStreamCreateWithFlags(&a->aux, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&a->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&a->join, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    zinv_aux_destroy(a);
    return Status::HIP_ERROR;
  }
  return Status::OK;
}

void zinv_aux_destroy(InvAux* a) {
  if (a->aux) (void)hipStreamSynchronize(a->aux);
  if (a->fork) (void)hipEventDestroy(a->fork);
  if (a->join) (void)hipEventDestroy(a->join);
  if (a->aux) (void)hipStreamDestroy(a->aux);
  a->aux = nullptr;
  a->fork = a->join = nullptr;
}

void zinv_aux_release_all() {
  std::lock_guard<std::mutex> lk(g_aux_mu);
  int cur = 0;
  const bool have_dev = hipGetDevice(&cur) == hipSuccess;
  for (int i = 0; i < g_naux; ++i) {
    if (have_dev) (void)hipSetDevice(g_aux[i].device);
    zinv_aux_destroy(&g_aux[i].a);
  }
  if (have_dev) (void)hipSetDevice(cur);
  g_naux = 0;
}

int zinv_set_variant(int v) {
  if (v < 0 || v > 6) return Status::BAD_ARG;
  g_inv_variant = v;
  return Status::OK;
}

size_t zinv_w_elems(int n) {
  // one W buffer per instance: the widest panel any schedule uses at this n
  // is nb_for(n) columns.  nb_for falls as n grows (64 -> 16 -> 8), so a
  // smaller block can need more than a larger one: a workspace carved for n
  // holds the largest need over every n' <= n.
  size_t c = (size_t)n * nb_for(n);
  if (n > 512) c = c > (size_t)512 * nb_for(512) ? c : (size_t)512 * nb_for(512);
  if (n > 1024) c = c > (size_t)1024 * nb_for(1024) ? c : (size_t)1024 * nb_for(1024);
  return c;
}

size_t zinv_work_bytes(int n, int batch) {
  size_t b = 0;
  b += sizeof(zc) * (size_t)n * n * batch;   // X
  b += 3 * sizeof(zc) * zinv_w_elems(n) * batch;  // W (three buffers: panels J .. J+2)
  b += sizeof(zc) * zinv_w_elems(n) * batch + 256;  // R (variant 6)
  b += sizeof(int) * (size_t)n * batch * 2;  // piv, used
  return b + 1024;
}

void zinv_work_carve(void* base, int n, int batch, InvWork* w) {
  char* p = (char*)base;
  auto take = [&](size_t bytes) {
    char* r = p;
    p += (bytes + 255) / 256 * 256;
    return r;
  };
  w->X = (zc*)take(sizeof(zc) * (size_t)n * n * batch);
  w->W = (zc*)take(3 * sizeof(zc) * zinv_w_elems(n) * batch);
  w->R = (zc*)take(sizeof(zc) * zinv_w_elems(n) * batch);
  w->piv = (int*)take(sizeof(int) * (size_t)n * batch);
  w->used = (int*)take(sizeof(int) * (size_t)n * batch);
  w->nmax = n;
  w->batchmax = batch;
  w->aux = InvAux{nullptr, nullptr, nullptr};
  w->inplace_only = 0;
  w->fro = nullptr;
}

int zinv_schedule(int n, int batch, int variant, int* strip, int* panel, int* nbuf) {
  if (n <= 0 || n > 2048 || batch <= 0 || variant < 0 || variant > 6) return Status::BAD_ARG;
  const int v = variant;
  const bool split = n <= 512 && (v >= 3 || (v == 0 && n > 256 && batch >= 32));
  *strip = n <= 512 ? 16 : (n <= 1024 ? 8 : 4);
  if (split) {
    *panel = v == 3 ? 1 : (v == 4 ? 2 : 4);
    // variant 6 (full panels only, else variant 5): the R buffer beside the W's
    *nbuf = (v == 6 && n % 64 == 0) ? 4 : 3;
  } else {
    // in place (v >= 3 with n > 512 falls back to the automatic panel choice)
    const bool pairs = v == 2 || ((v == 0 || v >= 3) && batch >= 32);
    *panel = pairs ? 2 : 1;
    *nbuf = 2;
  }
  return split ? 1 : 0;
}

int zinv_launch(int n, const zc* A, long long lda, long long strideA, zc* Ainv, long long ldinv,
                long long strideInv, int batch, const InvWork& w, int* info, hipStream_t s) {
  if (n < 0 || batch < 0 || lda < (n > 1 ? n : 1) || ldinv < (n > 1 ? n : 1)) return Status::BAD_ARG;
  if (n == 0 || batch == 0) return Status::OK;
  if (n > 2048) return Status::TOO_LARGE;
  // the workspace was carved for at most nmax x nmax blocks and batchmax
  // instances (zinv_work_carve); the schedule's W buffers must fit in it
  if (n > w.nmax || batch > w.batchmax) return Status::BAD_ARG;
  int S = 0, PS = 0, nbuf = 0;
  const int var = g_inv_variant ? g_inv_variant : tune_inv_variant(n, batch);
  int split = zinv_schedule(n, batch, var, &S, &PS, &nbuf);
  if (split < 0) return split;
  if (split && w.inplace_only)   // the automatic in-place panel choice instead
    split = zinv_schedule(n, batch, batch >= 32 ? 2 : 1, &S, &PS, &nbuf);
  if ((size_t)nbuf * n * S * PS > 4 * zinv_w_elems(w.nmax)) return Status::BAD_ARG;
  if (split && nbuf == 4 && n > 256)
    return run_split<512, 4>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s, true);
  if (split && nbuf == 4)
    return run_split<256, 4>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s, true);
  if (split) {
    if (n <= 256) {
      if (PS == 1) return run_split<256, 1>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
      if (PS == 2) return run_split<256, 2>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
      return run_split<256, 4>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
    }
    if (PS == 1) return run_split<512, 1>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
    if (PS == 2) return run_split<512, 2>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
    return run_split<512, 4>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  }
  if (PS == 1) {
    if (n <= 256) return run<256, 1, 16, 1>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
    if (n <= 512) return run<512, 1, 16, 1>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
    if (n <= 1024) return run<1024, 1, 8, 1>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
    return run<1024, 2, 4, 1>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  }
  if (n <= 256) return run<256, 1, 16, 2>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  if (n <= 512) return run<512, 1, 16, 2>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  if (n <= 1024) return run<1024, 1, 8, 2>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
  return run<1024, 2, 4, 2>(n, A, lda, strideA, Ainv, ldinv, strideInv, batch, w, info, s);
}

}  // namespace negf
