// Block-tridiagonal NEGF sweeps, batched over instances (energies), built from
// the MFMA GEMM and the Gauss-Jordan diagonal-block inverse.
//
// Every step of a sweep is one launch over all instances of the batch, so a
// wave of energies turns each b x b block operation into a batched GEMM with
// (wave * tiles) workgroups.  The sequence per layer follows the reference's
// algorithms with the LU solves replaced by products with the explicit
// inverse g_i = S_i^{-1} (computed once per layer and reused by every later
// pass):
//
//   RGF upward  (src/matinvertndiag_rgf.jl:143-162)
//     g_{N-1} = A_{N-1,N-1}^{-1};  for i = N-2..0:
//       t_i = A_{i,i+1} g_{i+1},  l_{i+1} = g_{i+1} A_{i+1,i},
//       g_i = (A_ii - A_{i,i+1} l_{i+1})^{-1}
//   RGF downward (:169-182)
//     G_00 = g_0;  G_{i-1,i} = -G_{i-1,i-1} t_{i-1},  G_{i,i-1} = -l_i G_{i-1,i-1},
//     G_ii = g_i - l_i G_{i-1,i}
//   RKD (src/keldyshndiag.jl:221-318): upward, central (X recursion), the
//     diagonal re-scaling g_i S_ii g_i^H, downward, as in keldyshndiag_v3!.
//   Transmission: G_{0,j} = -G_{0,j-1} t_{j-1} (:185-199),
//     T = Re Tr(Gamma_L G_{0,N-1} Gamma_R G_{0,N-1}^H).
#include "negf_internal.hpp"
#include "sweeps.hpp"

#include <vector>

namespace negf {

namespace {

const zc ONE = {1.0, 0.0}, ZERO = {0.0, 0.0}, MONE = {-1.0, 0.0};

struct M {           // one block position in a batch: base pointer, ld, batch stride
  const zc* p;
  long long ld, bs;
};
struct MW {
  zc* p;
  long long ld, bs;
};
inline M ro(const MW& w) { return M{w.p, w.ld, w.bs}; }

// Arithmetic of the products issued by this host thread (GemmArgs::prec and
// the emulation's workspace), set for the duration of an energy job by
// PrecScope; FP64 otherwise.
struct PrecCtx {
  int prec, nmod;
  void* work;
  long long bytes;
};
thread_local PrecCtx t_prec{0, 0, nullptr, 0};
struct PrecScope {
  PrecCtx saved;
  explicit PrecScope(const PrecCtx& c) : saved(t_prec) { t_prec = c; }
  ~PrecScope() { t_prec = saved; }
};
inline void apply_prec(GemmArgs& g) {
  g.prec = t_prec.prec;
  g.nmod = t_prec.nmod;
  g.work = t_prec.work;
  g.work_bytes = t_prec.bytes;
}

int gemm(int opA, int opB, int m, int n, int k, zc alpha, M A, M B, zc beta, M C, MW D, int batch,
         hipStream_t s, const zc* shift = nullptr, long long shift_stride = 0,
         const zc* scale = nullptr, long long scale_stride = 0, int herm = 0) {
  GemmArgs g{};
  apply_prec(g);
  g.opA = opA; g.opB = opB; g.m = m; g.n = n; g.k = k;
  g.alpha = alpha; g.beta = beta;
  g.A = A.p; g.lda = A.ld; g.strideA = A.bs;
  g.B = B.p; g.ldb = B.ld; g.strideB = B.bs;
  g.C = C.p; g.ldc = C.ld; g.strideC = C.bs;
  g.D = D.p; g.ldd = D.ld; g.strideD = D.bs;
  g.shift = shift; g.shift_stride = shift_stride;
  g.scale = scale; g.scale_stride = scale_stride;
  g.herm = herm;
  g.batch = batch;
  return zgemm_launch(g, s);
}

GemmArgs ga(int opA, int opB, int m, int n, int k, zc alpha, M A, M B, zc beta, M C, MW D,
            int batch) {
  GemmArgs g{};
  apply_prec(g);
  g.opA = opA; g.opB = opB; g.m = m; g.n = n; g.k = k;
  g.alpha = alpha; g.beta = beta;
  g.A = A.p; g.lda = A.ld; g.strideA = A.bs;
  g.B = B.p; g.ldb = B.ld; g.strideB = B.bs;
  g.C = C.p; g.ldc = C.ld; g.strideC = C.bs;
  g.D = D.p; g.ldd = D.ld; g.strideD = D.bs;
  g.batch = batch;
  return g;
}

#define RC(x)                 \
  do {                        \
    int _rc = (x);            \
    if (_rc) return _rc;      \
  } while (0)

// The device constant 1 the shift GEMMs read (a kernel store, not a host copy,
// so the sweeps contain only kernels and can be captured into a graph).
static int set_one(zc* p, hipStream_t s) { return zfill_launch(1, 1, ONE, ONE, p, 1, 1, 1, s); }

__global__ void fold_info_kernel(const int* layer_info, int* sweep_info, int layer, int batch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < batch && layer_info[i] != 0 && sweep_info[i] == 0) sweep_info[i] = layer + 1;
}

__global__ void zero_int_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// z_e = E_e + i eta, fL_e = f(E_e; mu_L), fR_e = f(E_e; mu_R) as complex scales,
// fermi = 1 / (1 + exp((E - mu) / kT)) in the overflow-safe tanh form.
__global__ void energy_scalars_kernel(int n, const double* E, double eta, double muL, double muR,
                                      double kT, zc* z, zc* fL, zc* fR, zc* sfL, zc* sfR) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double e = E[i];
  z[i] = zmake(e, eta);
  const double l = 0.5 * (1.0 - tanh(0.5 * (e - muL) / kT));
  const double r = 0.5 * (1.0 - tanh(0.5 * (e - muR) / kT));
  fL[i] = zmake(l, 0.0);
  fR[i] = zmake(r, 0.0);
  sfL[i] = zmake(sqrt(l), 0.0);
  sfR[i] = zmake(sqrt(r), 0.0);
}

// nrm[e * N + i] = ||g_i||_F for one layer i of any size (one workgroup per instance).
__global__ __launch_bounds__(256) void layer_fro_kernel(const zc* g, long long gbs, int bi, int i,
                                                         int N, double* nrm) {
  const int e = blockIdx.x;
  const zc* x = g + (long long)e * gbs;
  const long long n = (long long)bi * bi;
  double acc = 0.0;
  for (long long k = threadIdx.x; k < n; k += 256) acc = fma(x[k].x, x[k].x, fma(x[k].y, x[k].y, acc));
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) nrm[(long long)e * N + i] = sqrt(red[0]);
}

// est[e] = c sum_i sqrt(sum_c part[(i * nb + e) * nch + c]): the Frobenius
// norms of the inverses g_i from the partial sums their final pass wrote
// (InvWork::fro), layer by layer in a fixed order.
__global__ void fro_estimate_kernel(const double* part, int N, int nb, int nch, double c,
                                    double* est) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nb) return;
  double s = 0.0;
  for (int i = 0; i < N; ++i) {
    double q = 0.0;
    for (int k = 0; k < nch; ++k) q += part[((long long)i * nb + e) * nch + k];
    s += sqrt(q);
  }
  est[e] = c * s;
}

// est[e] = c sum_i nrm[e * N + i]
__global__ void energy_estimate_kernel(const double* nrm, int N, int nb, double c, double* est) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nb) return;
  double s = 0.0;
  for (int i = 0; i < N; ++i) s += nrm[(long long)e * N + i];
  est[e] = c * s;
}

}  // namespace

ProductScope::ProductScope(int prec)
    : prec_(t_prec.prec), nmod_(t_prec.nmod), work_(t_prec.work), bytes_(t_prec.bytes) {
  t_prec = PrecCtx{prec, 0, nullptr, 0};
}
ProductScope::~ProductScope() { t_prec = PrecCtx{prec_, nmod_, work_, bytes_}; }

double energy_prec_unit(int prec, int nmod, int k) {
  if (prec == 1) return 0x1p-24;
  // FP64 with three real products: the imaginary part's bound carries
  // (|ar| + |ai|)(|br| + |bi|) instead of |ar||bi| + |ai||br|, at most 2x;
  // two more bits of margin for the extra additions
  if (prec == 3) return 0x1p-51;
  if (prec == 2) {
    int beta = 0;
    if (zgemm_emu_params(nmod, k, &beta)) return 1.0;
    return ldexp(1.0, -beta);
  }
  return 0.0;
}

// ---------------------------------------------------------------------------
Layout::Layout(int N_, const int* sizes) : N(N_), b(sizes, sizes + N_) {
  od.resize(N + 1);
  oo.resize(N > 0 ? N : 1);
  od[0] = 0;
  for (int i = 0; i < N; ++i) od[i + 1] = od[i] + (long long)b[i] * b[i];
  oo[0] = 0;
  for (int i = 0; i + 1 < N; ++i) oo[i + 1] = oo[i] + (long long)b[i] * b[i + 1];
  nd = od[N];
  no = N > 1 ? oo[N - 1] : 0;
  bmax = 0;
  for (int x : b) bmax = x > bmax ? x : bmax;
}

// ---------------------------------------------------------------------------
// Workspace layout of one sweep wave (all in units of zc unless stated).
struct SweepWork {
  zc* g;      // N diag blocks (packed diag layout) per instance
  zc* t;      // N-1 offdiag (b_i x b_{i+1}) per instance
  zc* l;      // N-1 offdiag (b_{i+1} x b_i) per instance
  zc* T1;     // bmax^2 per instance
  zc* T2;     // bmax^2 per instance
  zc* X;      // bmax^2 per instance (RKD X recursion)
  zc* ones;   // 1 element (the shift that adds I)
  int* linfo; // batch
  InvWork inv;
  double* fro;  // null, or per layer i the inverses' partial sums of |g_i|^2
                // (InvWork::fro) at fro + i * batch * ceil(bmax / 64)
};

size_t sweep_work_bytes(const Layout& L, int batch, bool keldysh) {
  const size_t zb = sizeof(zc);
  size_t bytes = 0;
  auto add = [&](size_t b) { bytes += (b + 255) / 256 * 256; };
  add(zb * L.nd * batch);
  add(zb * L.no * batch);
  add(zb * L.no * batch);
  const size_t bb = (size_t)L.bmax * L.bmax;
  add(zb * bb * batch);
  add(zb * bb * batch);
  add(zb * bb * batch);
  add(zb);
  add(sizeof(int) * (size_t)batch);
  add(zinv_work_bytes(L.bmax, batch));
  (void)keldysh;
  return bytes + 4096;
}

static void carve(void* base, const Layout& L, int batch, SweepWork* w) {
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* r = p;
    p += (b + 255) / 256 * 256;
    return r;
  };
  const size_t zb = sizeof(zc);
  w->g = (zc*)take(zb * L.nd * batch);
  w->t = (zc*)take(zb * L.no * batch);
  w->l = (zc*)take(zb * L.no * batch);
  const size_t bb = (size_t)L.bmax * L.bmax;
  w->T1 = (zc*)take(zb * bb * batch);
  w->T2 = (zc*)take(zb * bb * batch);
  w->X = (zc*)take(zb * bb * batch);
  w->ones = (zc*)take(zb);
  w->linfo = (int*)take(sizeof(int) * (size_t)batch);
  zinv_work_carve(p, L.bmax, batch, &w->inv);
  w->fro = nullptr;
}

// ---------------------------------------------------------------------------
// Upward pass.  Source of the matrix blocks:
//   Adiag(i): A_ii   (or, with H-form, -H'_ii with the energy shift z added)
//   Aup(i):   A_{i,i+1};  Alow(i): A_{i+1,i} (opLow = OP_N), or the adjoint of
//   Aup(i) when the matrix is given in Hermitian-coupling form (opLow = OP_C,
//   sign sLow).
struct Source {
  // diag blocks
  const zc* diag; long long diag_bs;
  const zc* up; long long up_bs;
  const zc* low; long long low_bs;   // may be null: A_{i+1,i} = sLow * A_{i,i+1}^H
  zc sLow;
  // energy form: A_ii = shift - diag (diag holds H'_ii); shift per instance
  const zc* shift; long long shift_stride;
  zc sUp;      // A_{i,i+1} = sUp * up
};

// skip_top: stop after forming X = S_0 (in w.inv.X), before its inversion.
static int upward(const Layout& L, const Source& S, const SweepWork& w, int batch, int* info,
                  hipStream_t s, bool skip_top = false) {
  const int N = L.N;
  const long long gbs = L.nd, obs = L.no;
  hipLaunchKernelGGL(zero_int_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, info, batch);
  auto diagM = [&](int i) { return M{S.diag + L.od[i], L.b[i], S.diag_bs}; };
  auto invert = [&](int i, const zc* Xsrc, long long ldx, long long xs) -> int {
    InvWork iw = w.inv;
    if (w.fro) iw.fro = w.fro + (size_t)i * batch * ((L.bmax + 63) / 64);
    RC(zinv_launch(L.b[i], Xsrc, ldx, xs, w.g + L.od[i], L.b[i], gbs, batch, iw, w.linfo, s));
    hipLaunchKernelGGL(fold_info_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, w.linfo, info,
                       i, batch);
    return Status::OK;
  };
  // bottom layer
  {
    const int i = N - 1, bi = L.b[i];
    if (S.shift) {
      // X = shift * I - H'_{N-1}
      MW X{w.inv.X, bi, (long long)bi * bi};
      RC(gemm(OP_N, OP_N, bi, bi, 0, ONE, M{nullptr, bi, 0}, M{nullptr, 1, 0}, MONE, diagM(i), X,
              batch, s, S.shift, S.shift_stride));
      RC(invert(i, w.inv.X, bi, (long long)bi * bi));
    } else {
      RC(invert(i, S.diag + L.od[i], bi, S.diag_bs));
    }
  }
  for (int i = N - 2; i >= 0; --i) {
    const int bi = L.b[i], bj = L.b[i + 1];
    M up{S.up + L.oo[i], bi, S.up_bs};
    M gj{w.g + L.od[i + 1], bj, gbs};
    MW t{w.t + L.oo[i], bi, obs};
    MW l{w.l + L.oo[i], bj, obs};
    GemmArgs two[2];
    // t_i = A_{i,i+1} g_{i+1}
    two[0] = ga(OP_N, OP_N, bi, bj, bj, S.sUp, up, gj, ZERO, M{nullptr, 1, 0}, t, batch);
    // l_{i+1} = g_{i+1} A_{i+1,i}
    if (S.low) {
      M low{S.low + L.oo[i], bj, S.low_bs};
      two[1] = ga(OP_N, OP_N, bj, bi, bj, ONE, gj, low, ZERO, M{nullptr, 1, 0}, l, batch);
    } else {
      two[1] = ga(OP_N, OP_C, bj, bi, bj, S.sLow, gj, up, ZERO, M{nullptr, 1, 0}, l, batch);
    }
    if (bi == bj) {
      RC(zgemm_group_launch(two, 2, s));
    } else {
      RC(zgemm_launch(two[0], s));
      RC(zgemm_launch(two[1], s));
    }
    // X = A_ii - A_{i,i+1} l_{i+1}
    MW X{w.inv.X, bi, (long long)bi * bi};
    zc a = zmake(-S.sUp.x, -S.sUp.y);
    if (S.shift) {
      RC(gemm(OP_N, OP_N, bi, bi, bj, a, up, ro(l), MONE, diagM(i), X, batch, s, S.shift,
              S.shift_stride));
    } else {
      RC(gemm(OP_N, OP_N, bi, bi, bj, a, up, ro(l), ONE, diagM(i), X, batch, s));
    }
    if (i == 0 && skip_top) break;
    RC(invert(i, w.inv.X, bi, (long long)bi * bi));
  }
  return Status::OK;
}

// Downward RGF.  Writes G_ii into gd (packed diag layout, batch stride gd_bs)
// and G_{i,i+1} / G_{i+1,i} into gu / gl (packed offdiag, may be null: then
// two bmax^2 scratch blocks T1/T2 hold them one layer at a time).
static int downward(const Layout& L, const SweepWork& w, int batch, zc* gd, long long gd_bs,
                    zc* gu, zc* gl, long long go_bs, hipStream_t s) {
  const int N = L.N;
  const long long gbs = L.nd, obs = L.no;
  const long long xbs = (long long)L.bmax * L.bmax;
  RC(zcopy2d_launch(L.b[0], L.b[0], w.g, L.b[0], gbs, gd, L.b[0], gd_bs, batch, s));
  for (int i = 1; i < N; ++i) {
    const int bp = L.b[i - 1], bi = L.b[i];
    M Gpp{gd + L.od[i - 1], bp, gd_bs};
    MW Gup = gu ? MW{gu + L.oo[i - 1], bp, go_bs} : MW{w.T1, bp, xbs};
    MW Glo = gl ? MW{gl + L.oo[i - 1], bi, go_bs} : MW{w.T2, bi, xbs};
    M t{w.t + L.oo[i - 1], bp, obs};
    M l{w.l + L.oo[i - 1], bi, obs};
    GemmArgs two[2];
    two[0] = ga(OP_N, OP_N, bp, bi, bp, MONE, Gpp, t, ZERO, M{nullptr, 1, 0}, Gup, batch);
    two[1] = ga(OP_N, OP_N, bi, bp, bp, MONE, l, Gpp, ZERO, M{nullptr, 1, 0}, Glo, batch);
    if (!gl) {
      RC(zgemm_launch(two[0], s));       // G_{i,i-1} not wanted: only G_{i-1,i} feeds G_ii
    } else if (bp == bi) {
      RC(zgemm_group_launch(two, 2, s));
    } else {
      RC(zgemm_launch(two[0], s));
      RC(zgemm_launch(two[1], s));
    }
    // G_ii = g_i - l_i G_{i-1,i}
    MW Gii{gd + L.od[i], bi, gd_bs};
    RC(gemm(OP_N, OP_N, bi, bi, bp, MONE, l, ro(Gup), ONE, M{w.g + L.od[i], bi, gbs}, Gii, batch,
            s));
  }
  return Status::OK;
}

// Transmission T = Re Tr(Gamma_L G_{0,N-1} Gamma_R G_{0,N-1}^H), with the
// first block row from G_{0,j} = -G_{0,j-1} t_{j-1} starting at G_00 = g_0.
static int transmission(const Layout& L, const SweepWork& w, int batch, M gammaL, M gammaR,
                        double* T, hipStream_t s) {
  const int N = L.N;
  const long long xbs = (long long)L.bmax * L.bmax;
  const int b0 = L.b[0], bl = L.b[N - 1];
  // row chain ping-pongs between X and T1 (T2 is free here)
  MW cur{w.X, b0, xbs};
  RC(zcopy2d_launch(b0, b0, w.g, b0, L.nd, cur.p, b0, xbs, batch, s));
  for (int j = 1; j < N; ++j) {
    MW nxt{cur.p == w.X ? w.T1 : w.X, b0, xbs};
    RC(gemm(OP_N, OP_N, b0, L.b[j], L.b[j - 1], MONE, ro(cur), M{w.t + L.oo[j - 1], L.b[j - 1], L.no},
            ZERO, M{nullptr, 1, 0}, nxt, batch, s));
    cur = nxt;
  }
  // P = Gamma_L G (b0 x bl) into T2; Q = G Gamma_R (b0 x bl) into the other buffer
  MW P{w.T2, b0, xbs};
  MW Q{cur.p == w.X ? w.T1 : w.X, b0, xbs};
  GemmArgs two[2];
  two[0] = ga(OP_N, OP_N, b0, bl, b0, ONE, gammaL, ro(cur), ZERO, M{nullptr, 1, 0}, P, batch);
  two[1] = ga(OP_N, OP_N, b0, bl, bl, ONE, ro(cur), gammaR, ZERO, M{nullptr, 1, 0}, Q, batch);
  if (b0 == bl) {
    RC(zgemm_group_launch(two, 2, s));
  } else {
    RC(zgemm_launch(two[0], s));
    RC(zgemm_launch(two[1], s));
  }
  return zdotc_re_launch(b0, bl, P.p, b0, xbs, Q.p, b0, xbs, T, batch, s);
}

// RKD passes on Sd (packed diag, batch stride sd_bs) and Su (packed offdiag,
// stride su_bs), in place: input Sigma (Hermitian storage), output the band of
// A^{-1} Sigma A^{-H}.  lowSrc/sLow give A_{i+1,i} for the X recursion.
static int rkd(const Layout& L, const SweepWork& w, int batch, zc* Sd, long long sd_bs, zc* Su,
               long long su_bs, hipStream_t s) {
  const int N = L.N;
  const long long gbs = L.nd, obs = L.no;
  const long long xbs = (long long)L.bmax * L.bmax;
  auto sd = [&](int i) { return MW{Sd + L.od[i], L.b[i], sd_bs}; };
  auto su = [&](int i) { return MW{Su + L.oo[i], L.b[i], su_bs}; };
  auto g = [&](int i) { return M{w.g + L.od[i], L.b[i], gbs}; };
  auto t = [&](int i) { return M{w.t + L.oo[i], L.b[i], obs}; };
  auto l = [&](int i) { return M{w.l + L.oo[i - 1], L.b[i], obs}; };  // l_i = S_i^{-1} A_{i,i-1}
  // upward RKD (keldyshndiag.jl:230-234)
  for (int i = N - 2; i >= 0; --i) {
    const int bi = L.b[i], bj = L.b[i + 1];
    // Sd_i -= t_i Su_i^H ; Su_i -= t_i Sd_{i+1}  (two launches: the second
    // overwrites Su_i, which the first reads)
    RC(gemm(OP_N, OP_C, bi, bi, bj, MONE, t(i), ro(su(i)), ONE, ro(sd(i)), sd(i), batch, s));
    RC(gemm(OP_N, OP_N, bi, bj, bj, MONE, t(i), ro(sd(i + 1)), ONE, ro(su(i)), su(i), batch, s));
    // Sd_i -= Su_i t_i^H
    RC(gemm(OP_N, OP_C, bi, bi, bj, MONE, ro(su(i)), t(i), ONE, ro(sd(i)), sd(i), batch, s));
  }
  // central RKD (keldyshndiag.jl:256-301)
  if (N > 1) {
    const int b0 = L.b[0], b1 = L.b[1];
    MW T1{w.T1, b0, xbs};
    RC(gemm(OP_N, OP_N, b0, b1, b0, ONE, g(0), ro(su(0)), ZERO, M{nullptr, 1, 0}, T1, batch, s));
    RC(gemm(OP_N, OP_C, b0, b1, b1, ONE, ro(T1), g(1), ZERO, M{nullptr, 1, 0}, su(0), batch, s));
  }
  // X_i = I + A_{i,i-1} g_{i-1} X_{i-1} t_{i-1} (X_0 = I) enters only as
  // Y_i = g_i X_i = g_i + l_i (Y_{i-1} t_{i-1})  (l_i = g_i A_{i,i-1} from the
  // upward pass), which rescales Su_i and is kept in T1 for the next step.
  for (int i = 1; i + 1 < N; ++i) {
    const int bp = L.b[i - 1], bi = L.b[i], bn = L.b[i + 1];
    MW T2{w.T2, bp, xbs};
    // T2 = Y_{i-1} t_{i-1}     (bp x bi), Y_0 = g_0
    const M Y = (i == 1) ? g(0) : M{w.T1, bp, xbs};
    RC(gemm(OP_N, OP_N, bp, bi, bp, ONE, Y, t(i - 1), ZERO, M{nullptr, 1, 0}, T2, batch, s));
    // Y_i = T1 = g_i + l_i T2  (bi x bi)
    MW T1b{w.T1, bi, xbs}, T2b{w.T2, bi, xbs};
    RC(gemm(OP_N, OP_N, bi, bi, bp, ONE, l(i), ro(T2), ONE, g(i), T1b, batch, s));
    // T2 = T1 Su_i ; Su_i = T2 g_{i+1}^H
    RC(gemm(OP_N, OP_N, bi, bn, bi, ONE, ro(T1b), ro(su(i)), ZERO, M{nullptr, 1, 0}, T2b, batch,
            s));
    RC(gemm(OP_N, OP_C, bi, bn, bn, ONE, ro(T2b), g(i + 1), ZERO, M{nullptr, 1, 0}, su(i), batch,
            s));
  }
  // diagonal re-scaling: Sd_i = g_i Sd_i g_i^H
  for (int i = 0; i < N; ++i) {
    const int bi = L.b[i];
    MW T1{w.T1, bi, xbs};
    RC(gemm(OP_N, OP_N, bi, bi, bi, ONE, g(i), ro(sd(i)), ZERO, M{nullptr, 1, 0}, T1, batch, s));
    RC(gemm(OP_N, OP_C, bi, bi, bi, ONE, ro(T1), g(i), ZERO, M{nullptr, 1, 0}, sd(i), batch, s));
  }
  // downward RKD (keldyshndiag.jl:313-317), l = l_{i+1}
  for (int i = 0; i + 1 < N; ++i) {
    const int bi = L.b[i], bj = L.b[i + 1];
    GemmArgs two[2];
    // Sd_{i+1} -= Su_i^H l^H ; Su_i -= Sd_i l^H
    two[0] = ga(OP_C, OP_C, bj, bj, bi, MONE, ro(su(i)), l(i + 1), ONE, ro(sd(i + 1)), sd(i + 1),
                batch);
    two[1] = ga(OP_N, OP_C, bi, bj, bi, MONE, ro(sd(i)), l(i + 1), ONE, ro(su(i)), su(i), batch);
    if (bi == bj) {
      RC(zgemm_group_launch(two, 2, s));
    } else {
      RC(zgemm_launch(two[0], s));
      RC(zgemm_launch(two[1], s));
    }
    // Sd_{i+1} -= l Su_i
    RC(gemm(OP_N, OP_N, bj, bj, bi, MONE, l(i + 1), ro(su(i)), ONE, ro(sd(i + 1)), sd(i + 1), batch,
            s));
  }
  return Status::OK;
}

// Keldysh diagonal of the energy driver in injection form, factored.  With
// Sigma^n = f_L Gamma_L on layer 0 and f_R Gamma_R on layer N-1 only (its
// off-diagonal blocks zero), the reference's passes (src/keldyshndiag.jl:
// 221-318) give
//   G^n_jj = f_L G_j0 Gamma_L G_j0^H + f_R G_j,N-1 Gamma_R G_j,N-1^H,
// and RGF's own recursions give both columns: G_j0 = -l_j G_j-1,0 and
// G_j,N-1 = G_jj prod_k>=j (-t_k).  So the left term is C_j = l_j C_j-1 l_j^H
// (C_0 = f_L G_00 Gamma_L G_00^H) and the right one f_R G_jj B_j G_jj^H with
// B_j = t_j B_j+1 t_j^H, B_N-1 = Gamma_R -- the reference's upward Keldysh
// step (keldyshndiag.jl:230-234) run on Gamma_R alone -- whose B_0 gives
// T = Re Tr(Gamma_L G_00 B_0 G_00^H) (= Tr(Gamma_L G_0,N-1 Gamma_R
// G_0,N-1^H)) without the first-row chain.  Carried as factors: with the
// contacts' Gamma_R = F_R F_R^H and Gamma_L = F_L F_L^H (pivoted Cholesky,
// once per job), B_j = E_j E_j^H with E_j = t_j E_j+1, E_N-1 = F_R, and
// C_j = f_L F_j F_j^H with F_j = l_j F_j-1, F_0 = G_00 F_L, so
//   G^n_jj = Z_j Z_j^H,  Z_j = [sqrt(f_R) G_jj E_j | sqrt(f_L) F_j]  (b x 2b),
// one Hermitian product with k = 2b (its upper tiles mirrored: the output is
// read by nothing later).  Per layer E_j, F_j, W_j = sqrt(f_R) G_jj E_j and
// G^< = i Z Z^H: 4 GEMMs, the last at twice the depth and half the tiles,
// against 6 for B_j (2), C_j (2), V and G^< in the unfactored form; the same
// recursions, so the same products of t's and l's (a factor carried instead
// of its Gram matrix), and the Gram matrices come out Hermitian positive
// semidefinite by construction.  T = Re Tr(Gamma_L V V^H), V = G_00 E_0.
// Z_j alternates between two buffers (F_j is formed from F_j-1).
static int rkd_factored(const Layout& L, const SweepWork& w, int batch, zc* Ed, long long ed_bs,
                        const zc* FL, const zc* FR, const zc* Gd, long long gd_bs,
                        const zc* gammaL, const zc* sfL, const zc* sfR, zc* Z0, zc* Z1,
                        double* T, zc* Gless, long long gl_bs, hipStream_t s) {
  const int N = L.N, b = L.b[0];
  const long long xbs = (long long)L.bmax * L.bmax;
  const long long zbs = 2LL * b * b;
  const zc I1 = zmake(0.0, 1.0);
  const M none{nullptr, 1, 0};
  auto G = [&](int i) { return M{Gd + L.od[i], b, gd_bs}; };
  auto t = [&](int i) { return M{w.t + L.oo[i], b, L.no}; };
  auto l = [&](int i) { return M{w.l + L.oo[i - 1], b, L.no}; };   // l_i = S_i^{-1} A_{i,i-1}
  auto E = [&](int i) { return i == N - 1 ? M{FR, b, 0} : M{Ed + L.od[i], b, ed_bs}; };
  for (int i = N - 2; i >= 0; --i)
    RC(gemm(OP_N, OP_N, b, b, b, ONE, t(i), E(i + 1), ZERO, none, MW{Ed + L.od[i], b, ed_bs},
            batch, s));
  if (T) {
    MW V{w.T1, b, xbs}, P{w.T2, b, xbs};
    RC(gemm(OP_N, OP_N, b, b, b, ONE, G(0), E(0), ZERO, none, V, batch, s));
    RC(gemm(OP_N, OP_N, b, b, b, ONE, M{gammaL, b, 0}, ro(V), ZERO, none, P, batch, s));
    RC(zdotc_re_launch(b, b, P.p, b, xbs, V.p, b, xbs, T, batch, s));
  }
  for (int j = 0; j < N; ++j) {
    zc* Zc = (j & 1) ? Z1 : Z0;
    const zc* Zp = (j & 1) ? Z0 : Z1;
    const long long fo = (long long)b * b;   // F_j: the second b columns of Z_j
    if (j == 0)
      RC(gemm(OP_N, OP_N, b, b, b, ONE, G(0), M{FL, b, 0}, ZERO, none, MW{Zc + fo, b, zbs}, batch,
              s, nullptr, 0, sfL, 1));
    else
      RC(gemm(OP_N, OP_N, b, b, b, ONE, l(j), M{Zp + fo, b, zbs}, ZERO, none, MW{Zc + fo, b, zbs},
              batch, s));
    RC(gemm(OP_N, OP_N, b, b, b, ONE, G(j), E(j), ZERO, none, MW{Zc, b, zbs}, batch, s, nullptr,
            0, sfR, 1));
    RC(gemm(OP_N, OP_C, b, b, 2 * b, I1, M{Zc, b, zbs}, M{Zc, b, zbs}, ZERO, none,
            MW{Gless + L.od[j], b, gl_bs}, batch, s, nullptr, 0, nullptr, 0, -1));
  }
  return Status::OK;
}

// ---------------------------------------------------------------------------
// Two-sided RGF: band(A^{-1}) of a chain of uniform blocks with the upward
// recursion run from both ends at once.  With m = N / 2, the right part
// (layers m .. N-1) is the usual recursion (g_i connected to the layers
// above i) and the left part, layers m, m-1, .. 0 read in reverse with its
// couplings swapped, the same recursion on the mirrored chain (g_i connected
// to the layers below i).  Both parts go into one batch of 2B instances of
// n2 = N / 2 + 1 layers (the shorter padded at its far end by an identity
// layer with zero coupling, which changes nothing), so the dependent
// inversions number n2 instead of N.  At layer m both Schur terms meet:
// S_m = X_right + X_left - A_mm (each X = A_mm minus its side's term), whose
// inverse is G_mm, seeds both halves, and the usual downward pass runs
// outward in both at once.  The same RGF identities on each side (no
// pivoting across blocks is introduced), rounding in a different order.

int g_rgf_schedule = 0;   // 0 automatic, 1 one-sided always, 2 two-sided wherever it applies
constexpr int kTwoSidedMaxBatch = 16;

// Automatic while the inversions at 2B instances are still latency-bound:
// up to 16 instances of blocks up to 512, 4 of blocks up to 1024, 1 above.
// Measured (one-sided / two-sided, ms): 64 x 512 at 1, 4, 8 instances
// 96.6 / 54.2, 118.3 / 76.4, 143.9 / 109.6; 128 x 256 at 8 87.1 / 54.2;
// 32 x 1024 at 1, 4, 8 156.1 / 99.2, 230.4 / 186.0, 343.2 / 355.9.
bool two_sided_applies(const Layout& L, int batch) {
  if (g_rgf_schedule == 1 || L.N < 4) return false;
  for (int i = 1; i < L.N; ++i)
    if (L.b[i] != L.b[0]) return false;
  const int bmax_batch = L.b[0] <= 512 ? kTwoSidedMaxBatch : (L.b[0] <= 1024 ? 4 : 1);
  return g_rgf_schedule == 2 || batch <= bmax_batch;
}

namespace {

struct TwoSidedShape {
  int N, m, n2, b, B;
  long long bb, pbs;   // block elements; elements of one packed instance (3 n2 - 2 blocks)
  TwoSidedShape(const Layout& L, int batch)
      : N(L.N), m(L.N / 2), n2(L.N / 2 + 1), b(L.b[0]), B(batch),
        bb((long long)L.b[0] * L.b[0]), pbs((3LL * (L.N / 2 + 1) - 2) * bb) {}
};

// The 2B half-chains, per instance [diag n2][up n2-1][low n2-1] blocks.
// grid (3 n2 - 2 slots, 2B instances, column chunks of 16).
__global__ __launch_bounds__(256) void two_sided_pack_kernel(int N, int m, int n2, int b, int B,
                                                             const zc* __restrict__ Ad,
                                                             long long dbs,
                                                             const zc* __restrict__ Au,
                                                             long long ubs,
                                                             const zc* __restrict__ Al,
                                                             long long lbs, zc* __restrict__ P) {
  const int slot = blockIdx.x, k2 = blockIdx.y;
  const bool left = k2 >= B;
  const int k = left ? k2 - B : k2;
  const long long bb = (long long)b * b;
  zc* dst = P + (long long)k2 * (3LL * n2 - 2) * bb + (long long)slot * bb;
  int mode = 0;   // 0 zero, 1 identity, 2 copy, 3 adjoint of src
  const zc* src = nullptr;
  if (slot < n2) {
    const int layer = left ? m - slot : m + slot;
    if (layer >= 0 && layer < N) {
      mode = 2;
      src = Ad + (long long)k * dbs + (long long)layer * bb;
    } else {
      mode = 1;
    }
  } else {
    const bool up = slot < 2 * n2 - 1;
    const int p = up ? slot - n2 : slot - (2 * n2 - 1);
    // original coupling pair (i, i+1): right part positions p, p+1 are
    // layers m+p, m+p+1 (i = m+p); left part they are m-p, m-p-1 (i = m-p-1),
    // so its up block is A_{i+1,i} and its low block A_{i,i+1}
    const int i = left ? m - p - 1 : m + p;
    if (i >= 0 && i + 1 < N) {
      const bool want_upper = up != left;   // A_{i,i+1} wanted
      if (want_upper) {
        mode = 2;
        src = Au + (long long)k * ubs + (long long)i * bb;
      } else if (Al) {
        mode = 2;
        src = Al + (long long)k * lbs + (long long)i * bb;
      } else {   // Hermitian storage: A_{i+1,i} = A_{i,i+1}^H
        mode = 3;
        src = Au + (long long)k * ubs + (long long)i * bb;
      }
    }
  }
  const int c0 = blockIdx.z * 16, c1 = (c0 + 16) < b ? c0 + 16 : b;
  for (int c = c0; c < c1; ++c)
    for (int r = threadIdx.x; r < b; r += 256) {
      zc v;
      if (mode == 2) {
        v = src[r + (long long)c * b];
      } else if (mode == 3) {
        const zc x = src[c + (long long)r * b];
        v = zmake(x.x, -x.y);
      } else {
        v = zmake(mode == 1 && r == c ? 1.0 : 0.0, 0.0);
      }
      dst[r + (long long)c * b] = v;
    }
}

// X[k] += X[B + k] - A_mm[k] (A_mm: position 0 of the packed right part).
__global__ __launch_bounds__(256) void two_sided_combine_kernel(long long bb, int B, long long pbs,
                                                                const zc* __restrict__ P,
                                                                zc* __restrict__ X) {
  const int k = blockIdx.y;
  zc* x = X + (long long)k * bb;
  const zc* xl = X + (long long)(B + k) * bb;
  const zc* a = P + (long long)k * pbs;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < bb; e += 256LL * gridDim.x) {
    const zc u = x[e], v = xl[e], w = a[e];
    x[e] = zmake(u.x + v.x - w.x, u.y + v.y - w.y);
  }
}

// The half-chains' band back into the caller's layout.  grid (3N - 2 slots:
// diag N, up N-1, low N-1; B instances; column chunks of 16).
__global__ __launch_bounds__(256) void two_sided_unpack_kernel(
    int N, int m, int n2, int b, int B, const zc* __restrict__ Gd2, const zc* __restrict__ Gu2,
    const zc* __restrict__ Gl2, zc* __restrict__ Gd, long long dbs, zc* __restrict__ Gu,
    zc* __restrict__ Gl, long long obs) {
  const int slot = blockIdx.x, k = blockIdx.y;
  const long long bb = (long long)b * b;
  const long long d2 = (long long)n2 * bb, o2 = (long long)(n2 - 1) * bb;
  const zc* src;
  zc* dst;
  if (slot < N) {
    const int i = slot;
    src = i >= m ? Gd2 + (long long)k * d2 + (long long)(i - m) * bb
                 : Gd2 + (long long)(B + k) * d2 + (long long)(m - i) * bb;
    dst = Gd + (long long)k * dbs + (long long)i * bb;
  } else if (slot < 2 * N - 1) {   // G_{i,i+1}
    const int i = slot - N;
    if (!Gu) return;
    src = i >= m ? Gu2 + (long long)k * o2 + (long long)(i - m) * bb
                 : Gl2 + (long long)(B + k) * o2 + (long long)(m - i - 1) * bb;
    dst = Gu + (long long)k * obs + (long long)i * bb;
  } else {                         // G_{i+1,i}
    const int i = slot - (2 * N - 1);
    if (!Gl) return;
    src = i >= m ? Gl2 + (long long)k * o2 + (long long)(i - m) * bb
                 : Gu2 + (long long)(B + k) * o2 + (long long)(m - i - 1) * bb;
    dst = Gl + (long long)k * obs + (long long)i * bb;
  }
  const int c0 = blockIdx.z * 16, c1 = (c0 + 16) < b ? c0 + 16 : b;
  for (int c = c0; c < c1; ++c)
    for (int r = threadIdx.x; r < b; r += 256) dst[r + (long long)c * b] = src[r + (long long)c * b];
}

// info[k] = 1 + the original layer of instance k's first singular block
// found (right part first), or 0.
__global__ void two_sided_info_kernel(const int* __restrict__ info2, int m, int B,
                                      int* __restrict__ info) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= B) return;
  const int r = info2[k], l = info2[B + k];
  info[k] = r ? m + r : (l ? m - (l - 1) + 1 : 0);
}

size_t up256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace

size_t sweep_rgf_two_sided_bytes(const Layout& L, int batch) {
  if (!two_sided_applies(L, batch)) return 0;
  const TwoSidedShape t(L, batch);
  std::vector<int> sz(t.n2, t.b);
  Layout Lh(t.n2, sz.data());
  const size_t zb = sizeof(zc);
  return up256(sweep_work_bytes(Lh, 2 * batch, false)) +
         up256(zb * 2 * batch * (size_t)t.pbs) +               // packed halves
         up256(zb * 2 * batch * (size_t)t.n2 * t.bb) +          // G_ii of both halves
         2 * up256(zb * 2 * batch * (size_t)(t.n2 - 1) * t.bb) + // G_{i,i+1}, G_{i+1,i}
         up256(sizeof(int) * 2 * (size_t)batch);
}

static int sweep_rgf_two_sided(const Layout& L, const BtdIn& A, const BtdOut& G, int batch,
                               int* info, void* work, hipStream_t s, bool inplace_inverse) {
  const TwoSidedShape t(L, batch);
  const int B = batch, B2 = 2 * batch;
  std::vector<int> sz(t.n2, t.b);
  Layout Lh(t.n2, sz.data());
  const size_t zb = sizeof(zc);
  char* p = (char*)work;
  SweepWork w;
  carve(p, Lh, B2, &w);
  w.inv.inplace_only = inplace_inverse ? 1 : 0;
  p += up256(sweep_work_bytes(Lh, B2, false));
  zc* P = (zc*)p;
  p += up256(zb * B2 * (size_t)t.pbs);
  zc* Gd2 = (zc*)p;
  p += up256(zb * B2 * (size_t)t.n2 * t.bb);
  zc* Gu2 = (zc*)p;
  p += up256(zb * B2 * (size_t)(t.n2 - 1) * t.bb);
  zc* Gl2 = (zc*)p;
  p += up256(zb * B2 * (size_t)(t.n2 - 1) * t.bb);
  int* info2 = (int*)p;
  RC(set_one(w.ones, s));
  const unsigned chunks = (unsigned)((t.b + 15) / 16);
  hipLaunchKernelGGL(two_sided_pack_kernel, dim3(3 * t.n2 - 2, B2, chunks), dim3(256), 0, s, t.N,
                     t.m, t.n2, t.b, B, A.diag, A.diag_bs, A.up, A.up_bs, A.low, A.low_bs, P);
  Source S{};
  S.diag = P; S.diag_bs = t.pbs;
  S.up = P + (long long)t.n2 * t.bb; S.up_bs = t.pbs;
  S.low = P + (2LL * t.n2 - 1) * t.bb; S.low_bs = t.pbs;
  S.sLow = ONE; S.sUp = ONE;
  RC(upward(Lh, S, w, B2, info2, s, /*skip_top=*/true));
  // layer m: both sides' Schur terms, one inverse, G_mm into both halves
  const long long blocks = (t.bb + 255) / 256;
  hipLaunchKernelGGL(two_sided_combine_kernel, dim3((unsigned)(blocks < 1024 ? blocks : 1024), B),
                     dim3(256), 0, s, t.bb, B, t.pbs, P, w.inv.X);
  RC(zinv_launch(t.b, w.inv.X, t.b, t.bb, w.g, t.b, Lh.nd, B, w.inv, w.linfo, s));
  hipLaunchKernelGGL(fold_info_kernel, dim3((B + 255) / 256), dim3(256), 0, s, w.linfo, info2, 0,
                     B);
  RC(zcopy2d_launch(t.b, t.b, w.g, t.b, Lh.nd, w.g + (long long)B * Lh.nd, t.b, Lh.nd, B, s));
  RC(downward(Lh, w, B2, Gd2, Lh.nd, Gu2, G.low ? Gl2 : nullptr, Lh.no, s));
  hipLaunchKernelGGL(two_sided_unpack_kernel, dim3(3 * t.N - 2, B, chunks), dim3(256), 0, s, t.N,
                     t.m, t.n2, t.b, B, Gd2, Gu2, Gl2, G.diag, L.nd, G.up, G.low, L.no);
  hipLaunchKernelGGL(two_sided_info_kernel, dim3((B + 255) / 256), dim3(256), 0, s, info2, t.m, B,
                     info);
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

// ---------------------------------------------------------------------------
// Public sweep drivers (called by the C-ABI layer with validated arguments).

int sweep_rgf(const Layout& L, const BtdIn& A, const BtdOut& G, const BtdHerm* sigma,
              const BtdOut* gn, const GammaIn* gam, double* T, int batch, int* info, void* work,
              hipStream_t s, bool inplace_inverse, bool two_sided) {
  if (two_sided && G.diag && !(sigma && gn) && !(gam && T) && two_sided_applies(L, batch))
    return sweep_rgf_two_sided(L, A, G, batch, info, work, s, inplace_inverse);
  SweepWork w;
  carve(work, L, batch, &w);
  w.inv.inplace_only = inplace_inverse ? 1 : 0;
  RC(set_one(w.ones, s));
  Source S{};
  S.diag = A.diag; S.diag_bs = A.diag_bs;
  S.up = A.up; S.up_bs = A.up_bs;
  S.low = A.low; S.low_bs = A.low_bs;
  S.sLow = ONE; S.sUp = ONE;
  RC(upward(L, S, w, batch, info, s));
  if (G.diag) RC(downward(L, w, batch, G.diag, L.nd, G.up, G.low, L.no, s));
  if (gam && T)
    RC(transmission(L, w, batch, M{gam->gammaL, L.b[0], gam->gammaL_bs},
                    M{gam->gammaR, L.b[L.N - 1], gam->gammaR_bs}, T, s));
  if (sigma && gn) {
    // copy Sigma into the outputs and run RKD in place there
    RC(zcopy2d_launch((int)L.nd, 1, sigma->diag, L.nd, sigma->diag_bs, gn->diag, L.nd, L.nd, batch,
                      s));
    if (L.N > 1)
      RC(zcopy2d_launch((int)L.no, 1, sigma->up, L.no, sigma->up_bs, gn->up, L.no, L.no, batch, s));
    RC(rkd(L, w, batch, gn->diag, L.nd, gn->up, L.no, s));
  }
  return Status::OK;
}

// sweep_rgf with its products in a reduced arithmetic (prec 1: FP32 MFMA,
// 2: FP64 emulated on the int8 matrix cores with nmod moduli), the diagonal
// inverses in FP64, and per instance the estimate est[e] = 8 u sum_i
// ||g_i||_F of the energy job (u the products' unit roundoff) computed right
// after the upward pass.  work: sweep_work_bytes + sweep_mixed_extra_bytes.
size_t sweep_mixed_extra_bytes(const Layout& L, int batch, int prec, int nmod) {
  size_t b = sizeof(double) * (size_t)batch * L.N + 256;
  if (prec == 2) b += (zgemm_emu_work_bytes(L.bmax, L.bmax, L.bmax, batch, batch, nmod) + 255) / 256 * 256;
  return b;
}

int sweep_rgf_mixed(const Layout& L, const BtdIn& A, const BtdOut& G, const BtdHerm* sigma,
                    const BtdOut* gn, const GammaIn* gam, double* T, int batch, int* info,
                    void* work, int prec, int nmod, double* est, hipStream_t s) {
  const size_t main = (sweep_work_bytes(L, batch, sigma && gn) + 255) / 256 * 256;
  char* extra = (char*)work + main;
  double* nrm = (double*)extra;
  char* emu = extra + (sizeof(double) * (size_t)batch * L.N + 255) / 256 * 256;
  const size_t emu_bytes =
      prec == 2 ? zgemm_emu_work_bytes(L.bmax, L.bmax, L.bmax, batch, batch, nmod) : 0;
  PrecScope scope(PrecCtx{prec, nmod, prec == 2 ? (void*)emu : nullptr, (long long)emu_bytes});
  SweepWork w;
  carve(work, L, batch, &w);
  RC(set_one(w.ones, s));
  Source S{};
  S.diag = A.diag; S.diag_bs = A.diag_bs;
  S.up = A.up; S.up_bs = A.up_bs;
  S.low = A.low; S.low_bs = A.low_bs;
  S.sLow = ONE; S.sUp = ONE;
  RC(upward(L, S, w, batch, info, s));
  if (est) {
    for (int i = 0; i < L.N; ++i)
      hipLaunchKernelGGL(layer_fro_kernel, dim3(batch), dim3(256), 0, s, w.g + L.od[i], L.nd,
                         L.b[i], i, L.N, nrm);
    hipLaunchKernelGGL(energy_estimate_kernel, dim3((batch + 63) / 64), dim3(64), 0, s, nrm, L.N,
                       batch, 8.0 * energy_prec_unit(prec, nmod, L.bmax), est);
  }
  if (G.diag) RC(downward(L, w, batch, G.diag, L.nd, G.up, G.low, L.no, s));
  if (gam && T)
    RC(transmission(L, w, batch, M{gam->gammaL, L.b[0], gam->gammaL_bs},
                    M{gam->gammaR, L.b[L.N - 1], gam->gammaR_bs}, T, s));
  if (sigma && gn) {
    RC(zcopy2d_launch((int)L.nd, 1, sigma->diag, L.nd, sigma->diag_bs, gn->diag, L.nd, L.nd, batch,
                      s));
    if (L.N > 1)
      RC(zcopy2d_launch((int)L.no, 1, sigma->up, L.no, sigma->up_bs, gn->up, L.no, L.no, batch, s));
    RC(rkd(L, w, batch, gn->diag, L.nd, gn->up, L.no, s));
  }
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

// Every block of A^{-1} (the reference's buildFullInv, src/matinvertndiag_rgf.jl:
// 185-199) into a dense column-major n x n matrix per instance (batch stride
// n^2): the band from the downward pass written in place, then block
// diagonal d = 2 .. N-1 from diagonal d - 1:
//   G_{i,i+d} = -G_{i,i+d-1} t_{i+d-1},   G_{i+d,i} = -l_{i+d} G_{i+d-1,i},
// one launch per diagonal and direction over (i, instance) when the blocks
// are uniform, one per block otherwise.
int sweep_rgf_full(const Layout& L, const BtdIn& A, zc* Gf, int batch, int* info, void* work,
                   hipStream_t s) {
  SweepWork w;
  carve(work, L, batch, &w);
  RC(set_one(w.ones, s));
  Source S{};
  S.diag = A.diag; S.diag_bs = A.diag_bs;
  S.up = A.up; S.up_bs = A.up_bs;
  S.low = A.low; S.low_bs = A.low_bs;
  S.sLow = ONE; S.sUp = ONE;
  RC(upward(L, S, w, batch, info, s));
  const int N = L.N;
  std::vector<long long> off(N + 1, 0);
  for (int i = 0; i < N; ++i) off[i + 1] = off[i] + L.b[i];
  const long long n = off[N], nn = n * n;
  auto Gb = [&](int i, int j) { return MW{Gf + off[i] + off[j] * n, n, nn}; };
  auto t = [&](int i) { return M{w.t + L.oo[i], L.b[i], L.no}; };        // t_i: b_i x b_{i+1}
  auto l = [&](int i) { return M{w.l + L.oo[i - 1], L.b[i], L.no}; };    // l_i: b_i x b_{i-1}
  const M none{nullptr, 1, 0};
  // band: G_00 = g_0, then G_{i-1,i}, G_{i,i-1}, G_ii = g_i - l_i G_{i-1,i}
  RC(zcopy2d_launch(L.b[0], L.b[0], w.g, L.b[0], L.nd, Gf, n, nn, batch, s));
  for (int i = 1; i < N; ++i) {
    const int bp = L.b[i - 1], bi = L.b[i];
    RC(gemm(OP_N, OP_N, bp, bi, bp, MONE, ro(Gb(i - 1, i - 1)), t(i - 1), ZERO, none, Gb(i - 1, i),
            batch, s));
    RC(gemm(OP_N, OP_N, bi, bp, bp, MONE, l(i), ro(Gb(i - 1, i - 1)), ZERO, none, Gb(i, i - 1),
            batch, s));
    RC(gemm(OP_N, OP_N, bi, bi, bp, MONE, l(i), ro(Gb(i - 1, i)), ONE,
            M{w.g + L.od[i], bi, L.nd}, Gb(i, i), batch, s));
  }
  bool uniform = true;
  for (int i = 1; i < N; ++i) uniform = uniform && L.b[i] == L.b[0];
  for (int d = 2; d < N; ++d) {
    if (uniform) {
      const int b = L.b[0], cnt = N - d;
      const long long diag_step = (long long)b * (n + 1), bb = (long long)b * b;
      GemmArgs g{};
      apply_prec(g);
      g.opA = OP_N; g.opB = OP_N; g.m = g.n = g.k = b;
      g.alpha = MONE; g.beta = ZERO;
      g.batch = cnt * batch;
      g.inner = cnt;
      // upper: A = G_{i,i+d-1}, B = t_{i+d-1}, D = G_{i,i+d}
      g.A = Gf + (d - 1) * b * n; g.lda = n; g.strideA = diag_step; g.outerA = nn;
      g.B = w.t + (d - 1) * bb; g.ldb = b; g.strideB = bb; g.outerB = L.no;
      g.D = Gf + (long long)d * b * n; g.ldd = n; g.strideD = diag_step; g.outerD = nn;
      RC(zgemm_launch(g, s));
      // lower: A = l_{i+d}, B = G_{i+d-1,i}, D = G_{i+d,i}
      g.A = w.l + (d - 1) * bb; g.lda = b; g.strideA = bb; g.outerA = L.no;
      g.B = Gf + (long long)(d - 1) * b; g.ldb = n; g.strideB = diag_step; g.outerB = nn;
      g.D = Gf + (long long)d * b; g.ldd = n; g.strideD = diag_step; g.outerD = nn;
      RC(zgemm_launch(g, s));
    } else {
      for (int i = 0; i + d < N; ++i) {
        const int j = i + d;
        RC(gemm(OP_N, OP_N, L.b[i], L.b[j], L.b[j - 1], MONE, ro(Gb(i, j - 1)), t(j - 1), ZERO, none,
                Gb(i, j), batch, s));
        RC(gemm(OP_N, OP_N, L.b[j], L.b[i], L.b[j - 1], MONE, l(j), ro(Gb(j - 1, i)), ZERO, none,
                Gb(j, i), batch, s));
      }
    }
  }
  return Status::OK;
}

// ---------------------------------------------------------------------------
// Block (2w+1)-diagonal RGF, natively: bndiag_of_inv_general_rgf!
// (src/matinvertndiag_general_rgf.jl:31-146) with the LU solves as products
// with explicit inverses g_i.  Band layout: blocks (i, j), |i - j| <= w, in
// row-major order of (i, j), each b_i x b_j column-major (band_offsets).
//   upward, ix = N-2 .. 0, kmax = min(w, ix + 1), layers r, c in
//   ix - kmax + 1 .. ix:
//     T_{r,ix+1} = M_{r,ix+1} g_{ix+1},  L_{ix+1,c} = g_{ix+1} M_{ix+1,c}   (:58-69)
//     M_{r,c} -= M_{r,ix+1} L_{ix+1,c};  g_ix = M_{ix,ix}^{-1}             (:72-89)
//   downward, ix = 1 .. N-1, m, p over the kmax = min(w, ix) layers above:
//     G_{p,ix} = -sum_m G_{p,m} T_{m,ix},  G_{ix,p} = -sum_m L_{ix,m} G_{m,p}
//     G_{ix,ix} = g_ix - sum_m L_{ix,m} G_{m,ix}                           (:100-128)
// M is a working copy of the input band (the reference updates its input in
// place, :86); T and L share one band buffer (T above the diagonal, L below).
std::vector<long long> band_offsets(int N, const int* b, int w, long long* total) {
  std::vector<long long> off((size_t)N * (2 * w + 1), -1);
  long long o = 0;
  for (int i = 0; i < N; ++i)
    for (int j = (i - w > 0 ? i - w : 0); j <= (i + w < N - 1 ? i + w : N - 1); ++j) {
      off[(size_t)i * (2 * w + 1) + (j - i + w)] = o;
      o += (long long)b[i] * b[j];
    }
  *total = o;
  return off;
}

size_t general_rgf_work_bytes(int N, const int* b, int w, int batch) {
  long long band = 0;
  band_offsets(N, b, w, &band);
  int bmax = 0;
  long long nd = 0;
  for (int i = 0; i < N; ++i) {
    bmax = b[i] > bmax ? b[i] : bmax;
    nd += (long long)b[i] * b[i];
  }
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  return up(sizeof(zc) * 2 * band * batch) + up(sizeof(zc) * nd * batch) +
         up(sizeof(int) * batch) + up(zinv_work_bytes(bmax, batch)) + 4096;
}

int sweep_general_rgf(int N, const int* b, int w, const zc* Ain, long long a_bs, zc* Gout,
                      int batch, int* info, void* work, hipStream_t s) {
  long long band = 0;
  const std::vector<long long> off = band_offsets(N, b, w, &band);
  auto at = [&](int i, int j) { return off[(size_t)i * (2 * w + 1) + (j - i + w)]; };
  std::vector<long long> od(N + 1, 0);
  int bmax = 0;
  for (int i = 0; i < N; ++i) {
    od[i + 1] = od[i] + (long long)b[i] * b[i];
    bmax = b[i] > bmax ? b[i] : bmax;
  }
  const long long nd = od[N];
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  char* p = (char*)work;
  zc* Mw = (zc*)p;                  // working band copy
  zc* TL = Mw + band * batch;       // multipliers
  p += up(sizeof(zc) * 2 * band * batch);
  zc* g = (zc*)p;                   // g_i, packed diagonal layout
  p += up(sizeof(zc) * nd * batch);
  int* linfo = (int*)p;
  p += up(sizeof(int) * batch);
  InvWork iw;
  zinv_work_carve(p, bmax, batch, &iw);
  auto MB = [&](zc* base, int i, int j) { return MW{base + at(i, j), b[i], band}; };
  auto Mr = [&](const zc* base, int i, int j) { return M{base + at(i, j), b[i], band}; };
  auto G = [&](int i, int j) { return MW{Gout + at(i, j), b[i], band}; };
  auto gi = [&](int i) { return M{g + od[i], b[i], nd}; };
  const M none{nullptr, 1, 0};
  hipLaunchKernelGGL(zero_int_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, info, batch);
  RC(zcopy2d_launch((int)band, 1, Ain, band, a_bs, Mw, band, band, batch, s));
  auto invert = [&](int i, const zc* src) -> int {
    RC(zinv_launch(b[i], src, b[i], band, g + od[i], b[i], nd, batch, iw, linfo, s));
    hipLaunchKernelGGL(fold_info_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, linfo, info,
                       i, batch);
    return Status::OK;
  };
  RC(invert(N - 1, Mw + at(N - 1, N - 1)));
  for (int ix = N - 2; ix >= 0; --ix) {
    const int nx = ix + 1, kmax = (w < ix + 1 ? w : ix + 1);
    for (int k = 1; k <= kmax; ++k) {
      const int t = ix - k + 1;
      RC(gemm(OP_N, OP_N, b[t], b[nx], b[nx], ONE, Mr(Mw, t, nx), gi(nx), ZERO, none, MB(TL, t, nx),
              batch, s));
      RC(gemm(OP_N, OP_N, b[nx], b[t], b[nx], ONE, gi(nx), Mr(Mw, nx, t), ZERO, none, MB(TL, nx, t),
              batch, s));
    }
    for (int i = 1; i <= kmax; ++i)
      for (int j = 1; j <= kmax; ++j) {
        const int r = ix - i + 1, c = ix - j + 1;
        RC(gemm(OP_N, OP_N, b[r], b[c], b[nx], MONE, Mr(Mw, r, nx), Mr(TL, nx, c), ONE, Mr(Mw, r, c),
                MB(Mw, r, c), batch, s));
      }
    RC(invert(ix, Mw + at(ix, ix)));
  }
  RC(zcopy2d_launch(b[0], b[0], g, b[0], nd, Gout + at(0, 0), b[0], band, batch, s));
  for (int ix = 1; ix < N; ++ix) {
    const int kmax = (w < ix ? w : ix);
    for (int k = 1; k <= kmax; ++k) {
      const int tp = ix - k;
      for (int k2 = 1; k2 <= kmax; ++k2) {
        const int m = ix - k2;
        const zc beta = k2 == 1 ? ZERO : ONE;
        RC(gemm(OP_N, OP_N, b[tp], b[ix], b[m], MONE, ro(G(tp, m)), Mr(TL, m, ix), beta,
                ro(G(tp, ix)), G(tp, ix), batch, s));
        RC(gemm(OP_N, OP_N, b[ix], b[tp], b[m], MONE, Mr(TL, ix, m), ro(G(m, tp)), beta,
                ro(G(ix, tp)), G(ix, tp), batch, s));
      }
    }
    for (int k2 = 1; k2 <= kmax; ++k2) {
      const int m = ix - k2;
      const M cin = k2 == 1 ? gi(ix) : ro(G(ix, ix));
      RC(gemm(OP_N, OP_N, b[ix], b[ix], b[m], MONE, Mr(TL, ix, m), ro(G(m, ix)), ONE, cin, G(ix, ix),
              batch, s));
    }
  }
  return Status::OK;
}

// One stream's share of the energy grid: energies [e_begin, e_end) in waves.
// Per wave: the sweep's workspace, then Sd (the Keldysh pass's E_j) and G^R
// when the caller does not keep it (nd each per energy), then the factored
// pass's two Z buffers (2 b^2 each per energy).
static size_t energies_group_fp64_bytes(const Layout& L, int wave) {
  return (sweep_work_bytes(L, wave, true) + 255) / 256 * 256 +
         sizeof(zc) * (size_t)wave * (2 * L.nd) + sizeof(zc) * (size_t)wave * 4 * L.bmax * L.bmax +
         4096;
}

// The job's scalars and contact factors (sweep_energies' header).
struct EnergyJob {
  const zc *z, *fL, *fR, *sfL, *sfR;
  const zc *FL, *FR;   // F F^H = Gamma_L / Gamma_R (null: no G^< asked for)
  const int* gbad;     // 2 flags: Gamma_L / Gamma_R not Hermitian PSD
  int* einfo;
};

// info[e] = -1 for every energy of the wave when a contact's Gamma is not
// Hermitian positive semidefinite (the factored Keldysh pass needs Gamma =
// F F^H); it overrides a singular layer's code.
__global__ void gamma_info_kernel(const int* __restrict__ gbad, int* __restrict__ info, int nb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nb && (gbad[0] | gbad[1])) info[e] = -1;
}

static int energies_group(const Layout& L, const EnergyIn& in, const EnergyOut& out, int wave,
                          int e_begin, int e_end, const EnergyJob& job, const InvAux* aux,
                          void* work, hipStream_t s, const EnergyPrec* pr) {
  const int N = L.N, b = L.b[0];
  const zc* z = job.z;
  int* einfo = job.einfo;
  SweepWork w;
  carve(work, L, wave, &w);
  if (aux) w.inv.aux = *aux;
  zc* Sdw = (zc*)((char*)work + (sweep_work_bytes(L, wave, true) + 255) / 256 * 256);
  zc* Gdw = Sdw + (size_t)wave * L.nd;   // G^R diagonal when the caller does not keep it
  zc* Z0 = Gdw + (size_t)wave * L.nd;    // the factored Keldysh pass's Z_j, two buffers
  zc* Z1 = Z0 + (size_t)wave * 2 * L.bmax * L.bmax;
  // reduced-precision products: the emulation's residue workspace and the
  // per-layer norms of the estimate behind the FP64 group workspace
  const int prec = pr ? pr->prec : 0;
  char* extra = (char*)work + (energies_group_fp64_bytes(L, wave) + 255) / 256 * 256;
  size_t emu_bytes = 0;
  // (the factored Keldysh pass's G^< product has k = 2 bmax)
  if (prec == 2)
    emu_bytes = zgemm_emu_work_bytes(L.bmax, L.bmax, 2 * L.bmax, wave, wave, pr->nmod);
  // the inverses' partial sums of |g_i|^2, written by their final pass
  double* nrm = (double*)(extra + (emu_bytes + 255) / 256 * 256);
  const int nch = (L.bmax + 63) / 64;
  if (pr && pr->est) w.fro = nrm;
  PrecScope scope(PrecCtx{prec, pr ? pr->nmod : 0, prec == 2 ? (void*)extra : nullptr,
                          (long long)emu_bytes});
  // est = 8 u sum_i ||g_i||_F: measured on the 512-energy grid at N = 64,
  // b = 512, eta = 1e-4 (emulated products), error / (u sum ||g_i||_F) lies
  // in [0.10, 1.6] (median 0.13); 8 keeps a factor 5 over the largest
  const double est_c = pr && pr->est ? 8.0 * energy_prec_unit(prec, pr->nmod, L.bmax) : 0.0;
  RC(set_one(w.ones, s));
  for (int e0 = e_begin; e0 < e_end; e0 += wave) {
    const int nb = (e_end - e0) < wave ? (e_end - e0) : wave;
    Source S{};
    S.diag = in.Hp_diag; S.diag_bs = 0;        // H' (contacts folded in), shared
    S.up = in.H_up; S.up_bs = 0;
    S.low = nullptr; S.low_bs = 0;
    S.sUp = MONE;                              // A_{i,i+1} = -H_{i,i+1}
    S.sLow = MONE;                             // A_{i+1,i} = -H_{i,i+1}^H
    S.shift = z + e0; S.shift_stride = 1;
    int* info = out.info ? out.info + e0 : einfo + e0;
    RC(upward(L, S, w, nb, info, s));
    if (job.gbad)
      hipLaunchKernelGGL(gamma_info_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, job.gbad,
                         info, nb);
    if (pr && pr->est)
      hipLaunchKernelGGL(fro_estimate_kernel, dim3((nb + 63) / 64), dim3(64), 0, s, nrm, N, nb, nch,
                         est_c, pr->est + e0);
    // G^R diagonal blocks (the Keldysh pass reads them)
    zc* grd = out.GR_diag ? out.GR_diag + (size_t)e0 * L.nd : Gdw;
    RC(downward(L, w, nb, grd, L.nd, nullptr, nullptr, 0, s));
    if (out.T && !out.Gless_diag)
      RC(transmission(L, w, nb, M{in.gammaL, b, 0}, M{in.gammaR, b, 0}, out.T + e0, s));
    if (out.Gless_diag) {
      // Sigma^n = fL Gamma_L (+) fR Gamma_R on the diagonal, zero elsewhere:
      // G^< = i G^n layer by layer in the factored injection form (E_j in Sd),
      // T from E_0
      RC(rkd_factored(L, w, nb, Sdw, L.nd, job.FL, job.FR, grd, L.nd, in.gammaL, job.sfL + e0,
                      job.sfR + e0, Z0, Z1, out.T ? out.T + e0 : nullptr,
                      out.Gless_diag + (size_t)e0 * L.nd, L.nd, s));
    }
  }
  return Status::OK;
}

size_t energies_group_bytes(const Layout& L, int wave, const EnergyPrec* pr) {
  size_t b = (energies_group_fp64_bytes(L, wave) + 255) / 256 * 256;
  if (pr && pr->prec == 2)
    b += (zgemm_emu_work_bytes(L.bmax, L.bmax, 2 * L.bmax, wave, wave, pr->nmod) + 255) / 256 *
         256;
  if (pr && pr->est) b += sizeof(double) * (size_t)wave * L.N * ((L.bmax + 63) / 64) + 256;
  return b;
}

// The job's header: z, f_L, f_R, sqrt(f_L), sqrt(f_R) and info per energy,
// then the contact factors F_L, F_R, their Cholesky scratch, pivots and the
// two PSD flags.
static size_t energies_header_bytes(const Layout& L, int nE) {
  const size_t bb = (size_t)L.bmax * L.bmax;
  return (sizeof(zc) * 5 * (size_t)nE + sizeof(int) * (size_t)nE + 255) / 256 * 256 +
         (sizeof(zc) * 4 * bb + 255) / 256 * 256 + sizeof(int) * (2 * (size_t)L.bmax + 2) + 4096;
}

size_t energies_work_bytes(const Layout& L, int nE, int wave, int streams, const EnergyPrec* pr) {
  return (energies_header_bytes(L, nE) + 255) / 256 * 256 +
         (size_t)streams * ((energies_group_bytes(L, wave, pr) + 255) / 256 * 256);
}

// Energies are split into `streams` contiguous groups, each processed on its
// own stream in waves of `wave`, so one group's latency-bound steps (the
// Gauss-Jordan panels, one workgroup per energy) overlap another group's
// GEMMs.  The caller's stream waits for all groups.
int sweep_energies(const Layout& L, const EnergyIn& in, const EnergyOut& out, int wave,
                   int streams, const EnergyStreams& es, void* work, hipStream_t s,
                   const EnergyPrec* pr) {
  const int nE = in.nE, b = L.bmax;
  char* p = (char*)work;
  zc* z = (zc*)p;
  zc* fL = z + nE;
  zc* fR = fL + nE;
  zc* sfL = fR + nE;
  zc* sfR = sfL + nE;
  int* einfo = (int*)(sfR + nE);
  zc* F = (zc*)(p + (sizeof(zc) * 5 * (size_t)nE + sizeof(int) * (size_t)nE + 255) / 256 * 256);
  zc* Fw = F + 2 * (size_t)b * b;
  int* perm = (int*)(Fw + 2 * (size_t)b * b);
  int* gbad = perm + 2 * (size_t)b;
  char* groups = (char*)work + (energies_header_bytes(L, nE) + 255) / 256 * 256;
  const size_t gbytes = (energies_group_bytes(L, wave, pr) + 255) / 256 * 256;
  hipLaunchKernelGGL(energy_scalars_kernel, dim3((nE + 255) / 256), dim3(256), 0, s, nE,
                     in.energies, in.eta, in.muL, in.muR, in.kT, z, fL, fR, sfL, sfR);
  EnergyJob job{z, fL, fR, sfL, sfR, nullptr, nullptr, nullptr, einfo};
  if (out.Gless_diag) {
    // the contacts' factors, once per job: F F^H = Gamma (pivoted Cholesky)
    RC(zcopy2d_launch(b, b, in.gammaL, b, 0, Fw, b, 0, 1, s));
    RC(zcopy2d_launch(b, b, in.gammaR, b, 0, Fw + (size_t)b * b, b, 0, 1, s));
    RC(zpchol_launch(b, Fw, F, perm, gbad, 2, s));
    job.FL = F;
    job.FR = F + (size_t)b * b;
    job.gbad = gbad;
  }
  if (streams <= 1)
    return energies_group(L, in, out, wave, 0, nE, job, es.aux, groups, s, pr);
  hipStream_t* pool = es.pool;
  // fork / join events: the caller's (kept for a captured graph's lifetime),
  // else created here and destroyed on return
  std::vector<hipEvent_t> own;
  hipEvent_t* ev = es.events;
  if (!ev) {
    own.assign(1 + streams, nullptr);
    for (auto& e : own)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        for (auto& x : own)
          if (x) (void)hipEventDestroy(x);
        return Status::HIP_ERROR;
      }
    ev = own.data();
  }
  int rc = hipEventRecord(ev[0], s) == hipSuccess ? Status::OK : Status::HIP_ERROR;
  // contiguous chunks, each a whole number of waves where possible
  const int per = ((nE + streams - 1) / streams + wave - 1) / wave * wave;
  for (int g = 0; g < streams && rc == Status::OK; ++g) {
    const int e0 = g * per, e1 = (g + 1) * per < nE ? (g + 1) * per : nE;
    if (e0 >= e1) break;
    if (hipStreamWaitEvent(pool[g], ev[0], 0) != hipSuccess) {
      rc = Status::HIP_ERROR;
      break;
    }
    rc = energies_group(L, in, out, wave, e0, e1, job, es.aux ? es.aux + g : nullptr,
                        groups + g * gbytes, pool[g], pr);
    // join group g back into the caller's stream; where the event cannot
    // carry the edge, wait for the group on the host so later work on `s`
    // still follows it
    const bool joined = hipEventRecord(ev[1 + g], pool[g]) == hipSuccess &&
                        hipStreamWaitEvent(s, ev[1 + g], 0) == hipSuccess;
    if (!joined) {
      (void)hipStreamSynchronize(pool[g]);
      if (rc == Status::OK) rc = Status::HIP_ERROR;
    }
  }
  for (auto& x : own) (void)hipEventDestroy(x);
  return rc;
}

// ---------------------------------------------------------------------------
// DD-RGF (src/matinvertndiag_ddrgf_core.jl:719-758), uniform block size b.
//
// Chunk c = [separator s_c, L_c interior layers] (utils.jl:278-312, 478-575).
// The interiors of all chunks of all instances are solved as one batch:
// chunks 0..T-2 (length bs, a separator on each side) form the "inner"
// group, chunk T-1 (the leftover length, no right separator) the "last"
// group.  Per chunk instance, from one RGF pass over its interior:
//   band of g = T11^{-1}, first / last block column zF / zL and first / last
//   block row rF / rL of g (column j of g = L^{-1} D^{-1} U^{-1} e_j and row j =
//   e_j^T L^{-1} D^{-1} U^{-1} in the UDL factors of the upward pass), then
//   YF = zF A[f,s_c], YL = zL A[e,s_{c+1}] (T11^{-1} T12, core:48-114),
//   ZF = A[s_c,f] rF,  ZL = A[s_{c+1},e] rL (T21 T11^{-1}, core:203-266).
// Schur complement over separators (core:500-654), band of its inverse by
// RGF or by the next DD-RGF level (core:656-680), then
//   G21 = -S^{-1} [ZF; ZL] (core:268-344), G12 band blocks, and the band of
//   G11 = g - [YF YL] G21 (core:346-435), scattered to the original order.
namespace {

struct Bump {
  char* base;
  size_t off = 0;
  zc* take(size_t elems) {
    char* r = base ? base + off : nullptr;
    off += (elems * sizeof(zc) + 255) / 256 * 256;
    return (zc*)r;
  }
  int* take_int(size_t n) {
    char* r = base ? base + off : nullptr;
    off += (n * sizeof(int) + 255) / 256 * 256;
    return (int*)r;
  }
  void* take_bytes(size_t n) {
    char* r = base ? base + off : nullptr;
    off += (n + 255) / 256 * 256;
    return r;
  }
};

bool dd_plan(int N, int bs, int* T, int* lastL) {
  if (bs < 1) return false;
  const int per = bs + 1;
  const int nr = N / per + 1;
  const int left = N - (nr - 1) * per;
  if (left == 0) { *T = nr - 1; *lastL = bs; return true; }
  if (left - 1 == 0) return false;
  *T = nr; *lastL = left - 1;
  return true;
}

// GEMM over a 2-D instance grid (inner index i < inner, outer o): operand X
// at X + i * sX + o * oX.
int gemm2(int opA, int opB, int b, zc alpha, const zc* A, long long sA, long long oA, const zc* B,
          long long sB, long long oB, zc beta, const zc* C, long long sC, long long oC, zc* D,
          long long sD, long long oD, int inner, int outer, hipStream_t s) {
  GemmArgs g{};
  apply_prec(g);
  g.opA = opA; g.opB = opB; g.m = g.n = g.k = b;
  g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = b; g.strideA = sA; g.outerA = oA;
  g.B = B; g.ldb = b; g.strideB = sB; g.outerB = oB;
  g.C = C; g.ldc = b; g.strideC = sC; g.outerC = oC;
  g.D = D; g.ldd = b; g.strideD = sD; g.outerD = oD;
  g.batch = inner * outer;
  g.inner = inner;
  return zgemm_launch(g, s);
}

__global__ void dd_fold_info_kernel(const int* inst_info, int nC, int c0, int bs, int batch,
                                    int* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= batch) return;
  for (int ci = 0; ci < nC; ++ci) {
    const int v = inst_info[e * nC + ci];
    if (v != 0 && out[e] == 0) out[e] = (c0 + ci) * (bs + 1) + 1 + v;
  }
}

__global__ void dd_fold_sep_info_kernel(const int* sep_info, int bs, int batch, int* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < batch && sep_info[e] != 0 && out[e] == 0) out[e] = (sep_info[e] - 1) * (bs + 1) + 1;
}

struct Group {
  int c0, nC, L;
  bool right;
};

}  // namespace

// A: packed BTD (diag / up / low, all given, tight per-block ld = b, batch
// strides as in BtdIn); G: outputs, tight strides.  `bw` null measures only.
int ddrgf_run(int N, int b, const BtdIn& A, const BtdOut& G, int batch, const int* levels,
              int nlev, int* info, Bump& bw, hipStream_t s) {
  const bool run = bw.base != nullptr;
  const long long bb = (long long)b * b;
  const long long nd = (long long)N * bb, no = (long long)(N - 1) * bb;
  int T = 0, lastL = 0;
  if (nlev <= 0 || !dd_plan(N, levels[0], &T, &lastL)) return Status::BAD_ARG;
  const int bs = levels[0];
  std::vector<int> sizes(N, b);
  if (T <= 1) {
    // one task: sequential RGF (core:724-729)
    Layout L(N, sizes.data());
    SweepWork w;
    void* wp = bw.take_bytes(sweep_work_bytes(L, batch, false));
    if (!run) return Status::OK;
    carve(wp, L, batch, &w);
    RC(set_one(w.ones, s));
    Source S{};
    S.diag = A.diag; S.diag_bs = A.diag_bs; S.up = A.up; S.up_bs = A.up_bs;
    S.low = A.low; S.low_bs = A.low_bs; S.sLow = ONE; S.sUp = ONE;
    RC(upward(L, S, w, batch, info, s));
    return downward(L, w, batch, G.diag, nd, G.up, G.low, no, s);
  }
  std::vector<Group> groups;
  if (T >= 2) groups.push_back(Group{0, T - 1, bs, true});
  groups.push_back(Group{T - 1, 1, lastL, false});
  if (run)
    hipLaunchKernelGGL(zero_int_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, info, batch);

  // reduced system buffers (T layers per instance)
  zc* Sd = bw.take((size_t)T * bb * batch);
  zc* Su = bw.take((size_t)(T - 1) * bb * batch);
  zc* Sl = bw.take((size_t)(T - 1) * bb * batch);
  zc* SId = bw.take((size_t)T * bb * batch);
  zc* SIu = bw.take((size_t)(T - 1) * bb * batch);
  zc* SIl = bw.take((size_t)(T - 1) * bb * batch);
  int* sep_info = bw.take_int(batch);

  struct GBuf {
    zc *Cd, *Cu, *Cl, *Gd, *Gu, *Gl, *zF, *rF, *zL, *rL, *Xs, *YF, *ZF, *YL, *ZL, *G21F, *G21L;
    zc *Kls, *Kus, *Kur, *Klr, *tmp;
    int* ginfo;
    SweepWork w;
  };
  std::vector<GBuf> gb(groups.size());
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const Group& g = groups[gi];
    const size_t NI = (size_t)batch * g.nC, L = g.L;
    GBuf& q = gb[gi];
    q.Cd = bw.take(NI * L * bb);
    q.Cu = bw.take(NI * (L > 1 ? L - 1 : 1) * bb);
    q.Cl = bw.take(NI * (L > 1 ? L - 1 : 1) * bb);
    q.Gd = bw.take(NI * L * bb);
    q.Gu = bw.take(NI * (L > 1 ? L - 1 : 1) * bb);
    q.Gl = bw.take(NI * (L > 1 ? L - 1 : 1) * bb);
    q.zF = bw.take(NI * L * bb);
    q.rF = bw.take(NI * L * bb);
    q.YF = bw.take(NI * L * bb);
    q.ZF = bw.take(NI * L * bb);
    q.G21F = bw.take(NI * L * bb);
    q.Kls = bw.take(NI * bb);
    q.Kus = bw.take(NI * bb);
    q.tmp = bw.take(NI * bb);
    if (g.right) {
      q.zL = bw.take(NI * L * bb);
      q.rL = bw.take(NI * L * bb);
      q.Xs = bw.take(NI * L * bb);
      q.YL = bw.take(NI * L * bb);
      q.ZL = bw.take(NI * L * bb);
      q.G21L = bw.take(NI * L * bb);
      q.Kur = bw.take(NI * bb);
      q.Klr = bw.take(NI * bb);
    } else {
      q.zL = q.rL = q.Xs = q.YL = q.ZL = q.G21L = q.Kur = q.Klr = nullptr;
    }
    q.ginfo = bw.take_int(NI);
    std::vector<int> lsz(g.L, b);
    Layout Lg(g.L, lsz.data());
    void* wp = bw.take_bytes(sweep_work_bytes(Lg, (int)NI, false));
    if (run) carve(wp, Lg, (int)NI, &q.w);
  }
  // reduced-system solver workspace (RGF at the last level, next level otherwise)
  std::vector<int> tsz(T, b);
  Layout LT(T, tsz.data());
  SweepWork wT;
  Bump sub_bw{nullptr};
  void* wTp = nullptr;
  if (nlev > 1) {
    // measure the next level, then take that much
    Bump meas{nullptr};
    BtdIn SA{nullptr, 0, nullptr, 0, nullptr, 0};
    BtdOut SG{nullptr, nullptr, nullptr};
    int rc = ddrgf_run(T, b, SA, SG, batch, levels + 1, nlev - 1, nullptr, meas, s);
    if (rc) return rc;
    sub_bw.base = (char*)bw.take_bytes(meas.off);
  } else {
    wTp = bw.take_bytes(sweep_work_bytes(LT, batch, false));
  }
  if (!run) return Status::OK;

  for (auto& q : gb) RC(set_one(q.w.ones, s));

  // ---- per group: chunk RGF, columns / rows of T11^{-1}, couplings
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const Group& g = groups[gi];
    GBuf& q = gb[gi];
    const int NI = batch * g.nC, L = g.L;
    const long long Lb = (long long)L * bb, L1b = (long long)(L > 1 ? L - 1 : 1) * bb;
    const long long cstride = (long long)(bs + 1) * bb;   // chunk to chunk in the packed arrays
    const int f0 = g.c0 * (bs + 1) + 1, s0 = g.c0 * (bs + 1);
    // gather the interiors and the coupling blocks (instance = e * nC + ci)
    RC(zcopy3d_launch(bb, L, g.nC, batch, ONE, A.diag + f0 * bb, bb, cstride, A.diag_bs, q.Cd, bb,
                      Lb, (long long)g.nC * Lb, s));
    if (L > 1) {
      RC(zcopy3d_launch(bb, L - 1, g.nC, batch, ONE, A.up + f0 * bb, bb, cstride, A.up_bs, q.Cu, bb,
                        L1b, (long long)g.nC * L1b, s));
      RC(zcopy3d_launch(bb, L - 1, g.nC, batch, ONE, A.low + f0 * bb, bb, cstride, A.low_bs, q.Cl,
                        bb, L1b, (long long)g.nC * L1b, s));
    }
    RC(zcopy3d_launch(bb, 1, g.nC, batch, ONE, A.low + s0 * bb, 0, cstride, A.low_bs, q.Kls, 0, bb,
                      (long long)g.nC * bb, s));
    RC(zcopy3d_launch(bb, 1, g.nC, batch, ONE, A.up + s0 * bb, 0, cstride, A.up_bs, q.Kus, 0, bb,
                      (long long)g.nC * bb, s));
    if (g.right) {
      const int e0 = f0 + L - 1;
      RC(zcopy3d_launch(bb, 1, g.nC, batch, ONE, A.up + e0 * bb, 0, cstride, A.up_bs, q.Kur, 0, bb,
                        (long long)g.nC * bb, s));
      RC(zcopy3d_launch(bb, 1, g.nC, batch, ONE, A.low + e0 * bb, 0, cstride, A.low_bs, q.Klr, 0,
                        bb, (long long)g.nC * bb, s));
    }
    // chunk RGF: UDL factors and the band of T11^{-1}
    std::vector<int> lsz(L, b);
    Layout Lg(L, lsz.data());
    Source S{};
    S.diag = q.Cd; S.diag_bs = Lb; S.up = q.Cu; S.up_bs = L1b; S.low = q.Cl; S.low_bs = L1b;
    S.sLow = ONE; S.sUp = ONE;
    RC(upward(Lg, S, q.w, NI, q.ginfo, s));
    hipLaunchKernelGGL(dd_fold_info_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, q.ginfo,
                       g.nC, g.c0, bs, batch, info);
    RC(downward(Lg, q.w, NI, q.Gd, Lb, q.Gu, q.Gl, L1b, s));
    const zc* gg = q.w.g;   // g_j at j*bb + inst*Lb
    const zc* tt = q.w.t;   // t_j at j*bb + inst*L1b
    const zc* ll = q.w.l;   // l_{j+1} at j*bb + inst*L1b
    // first column zF_0 = g_0, zF_i = -l_i zF_{i-1};  first row rF_0 = g_0, rF_k = -rF_{k-1} t_{k-1}
    RC(zcopy3d_launch(bb, 1, 1, NI, ONE, gg, 0, 0, Lb, q.zF, 0, 0, Lb, s));
    RC(zcopy3d_launch(bb, 1, 1, NI, ONE, gg, 0, 0, Lb, q.rF, 0, 0, Lb, s));
    for (int i = 1; i < L; ++i) {
      GemmArgs two[2];
      two[0] = ga(OP_N, OP_N, b, b, b, MONE, M{ll + (i - 1) * bb, b, L1b}, M{q.zF + (i - 1) * bb, b, Lb},
                  ZERO, M{nullptr, 1, 0}, MW{q.zF + i * bb, b, Lb}, NI);
      two[1] = ga(OP_N, OP_N, b, b, b, MONE, M{q.rF + (i - 1) * bb, b, Lb}, M{tt + (i - 1) * bb, b, L1b},
                  ZERO, M{nullptr, 1, 0}, MW{q.rF + i * bb, b, Lb}, NI);
      RC(zgemm_group_launch(two, 2, s));
    }
    if (g.right) {
      // last column: x_{L-1} = I, x_i = -t_i x_{i+1}; y_i = g_i x_i; z_i = y_i - l_i z_{i-1}
      RC(zfill_launch(b, b, ZERO, ONE, q.Xs + (L - 1) * bb, b, Lb, NI, s));
      for (int i = L - 2; i >= 0; --i)
        RC(gemm(OP_N, OP_N, b, b, b, MONE, M{tt + i * bb, b, L1b}, M{q.Xs + (i + 1) * bb, b, Lb},
                ZERO, M{nullptr, 1, 0}, MW{q.Xs + i * bb, b, Lb}, NI, s));
      RC(gemm2(OP_N, OP_N, b, ONE, gg, bb, Lb, q.Xs, bb, Lb, ZERO, nullptr, 0, 0, q.zL, bb, Lb, L,
               NI, s));
      for (int i = 1; i < L; ++i)
        RC(gemm(OP_N, OP_N, b, b, b, MONE, M{ll + (i - 1) * bb, b, L1b}, M{q.zL + (i - 1) * bb, b, Lb},
                ONE, M{q.zL + i * bb, b, Lb}, MW{q.zL + i * bb, b, Lb}, NI, s));
      // last row: w_{L-1} = I, w_k = -w_{k+1} l_{k+1}; v_k = w_k g_k; r_k = v_k - r_{k-1} t_{k-1}
      RC(zfill_launch(b, b, ZERO, ONE, q.Xs + (L - 1) * bb, b, Lb, NI, s));
      for (int k = L - 2; k >= 0; --k)
        RC(gemm(OP_N, OP_N, b, b, b, MONE, M{q.Xs + (k + 1) * bb, b, Lb}, M{ll + k * bb, b, L1b},
                ZERO, M{nullptr, 1, 0}, MW{q.Xs + k * bb, b, Lb}, NI, s));
      RC(gemm2(OP_N, OP_N, b, ONE, q.Xs, bb, Lb, gg, bb, Lb, ZERO, nullptr, 0, 0, q.rL, bb, Lb, L,
               NI, s));
      for (int k = 1; k < L; ++k)
        RC(gemm(OP_N, OP_N, b, b, b, MONE, M{q.rL + (k - 1) * bb, b, Lb}, M{tt + (k - 1) * bb, b, L1b},
                ONE, M{q.rL + k * bb, b, Lb}, MW{q.rL + k * bb, b, Lb}, NI, s));
    }
    // couplings: YF = zF A[f,s_c], ZF = A[s_c,f] rF (and YL, ZL)
    RC(gemm2(OP_N, OP_N, b, ONE, q.zF, bb, Lb, q.Kls, 0, bb, ZERO, nullptr, 0, 0, q.YF, bb, Lb, L, NI,
             s));
    RC(gemm2(OP_N, OP_N, b, ONE, q.Kus, 0, bb, q.rF, bb, Lb, ZERO, nullptr, 0, 0, q.ZF, bb, Lb, L, NI,
             s));
    if (g.right) {
      RC(gemm2(OP_N, OP_N, b, ONE, q.zL, bb, Lb, q.Kur, 0, bb, ZERO, nullptr, 0, 0, q.YL, bb, Lb, L,
               NI, s));
      RC(gemm2(OP_N, OP_N, b, ONE, q.Klr, 0, bb, q.rL, bb, Lb, ZERO, nullptr, 0, 0, q.ZL, bb, Lb, L,
               NI, s));
    }
  }

  // ---- Schur complement over the separators (core:500-654)
  const long long Tb = (long long)T * bb, T1b = (long long)(T - 1) * bb;
  const long long cstride = (long long)(bs + 1) * bb;
  RC(zcopy3d_launch(bb, T, 1, batch, ONE, A.diag, cstride, 0, A.diag_bs, Sd, bb, 0, Tb, s));
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const Group& g = groups[gi];
    GBuf& q = gb[gi];
    const int L = g.L;
    const long long Lb = (long long)L * bb;
    const long long nCb = (long long)g.nC * bb, nCLb = (long long)g.nC * Lb;
    // Sd_c -= A[s_c,f] YF_0
    RC(gemm2(OP_N, OP_N, b, MONE, q.Kus, bb, nCb, q.YF, Lb, nCLb, ONE, Sd + g.c0 * bb, bb, Tb,
             Sd + g.c0 * bb, bb, Tb, g.nC, batch, s));
    if (g.right) {
      // Sd_{c+1} -= A[s_{c+1},e] YL_{L-1};  Su_c = -A[s_c,f] YL_0;  Sl_c = -A[s_{c+1},e] YF_{L-1}
      RC(gemm2(OP_N, OP_N, b, MONE, q.Klr, bb, nCb, q.YL + (L - 1) * bb, Lb, nCLb, ONE,
               Sd + (g.c0 + 1) * bb, bb, Tb, Sd + (g.c0 + 1) * bb, bb, Tb, g.nC, batch, s));
      RC(gemm2(OP_N, OP_N, b, MONE, q.Kus, bb, nCb, q.YL, Lb, nCLb, ZERO, nullptr, 0, 0,
               Su + g.c0 * bb, bb, T1b, g.nC, batch, s));
      RC(gemm2(OP_N, OP_N, b, MONE, q.Klr, bb, nCb, q.YF + (L - 1) * bb, Lb, nCLb, ZERO, nullptr,
               0, 0, Sl + g.c0 * bb, bb, T1b, g.nC, batch, s));
    }
  }

  // ---- band of S^{-1} (core:656-680): RGF at the last level, else the next level
  {
    BtdIn SA{Sd, Tb, Su, T1b, Sl, T1b};
    BtdOut SG{SId, SIu, SIl};
    if (nlev > 1) {
      RC(ddrgf_run(T, b, SA, SG, batch, levels + 1, nlev - 1, sep_info, sub_bw, s));
    } else {
      carve(wTp, LT, batch, &wT);
      RC(set_one(wT.ones, s));
      Source S{};
      S.diag = Sd; S.diag_bs = Tb; S.up = Su; S.up_bs = T1b; S.low = Sl; S.low_bs = T1b;
      S.sLow = ONE; S.sUp = ONE;
      RC(upward(LT, S, wT, batch, sep_info, s));
      RC(downward(LT, wT, batch, SId, Tb, SIu, SIl, T1b, s));
    }
    hipLaunchKernelGGL(dd_fold_sep_info_kernel, dim3((batch + 255) / 256), dim3(256), 0, s,
                       sep_info, bs, batch, info);
  }
  // separator diagonal blocks of G
  RC(zcopy3d_launch(bb, T, 1, batch, ONE, SId, bb, 0, Tb, G.diag, cstride, 0, nd, s));

  // ---- back-substitution per group
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const Group& g = groups[gi];
    GBuf& q = gb[gi];
    const int L = g.L;
    const int Mi = batch * g.nC;
    const long long Lb = (long long)L * bb, L1b = (long long)(L > 1 ? L - 1 : 1) * bb;
    const long long nCb = (long long)g.nC * bb, nCLb = (long long)g.nC * Lb;
    const zc* SIcc = SId + g.c0 * bb;            // S^{-1}_{c,c}: (ci, e) -> ci*bb + e*Tb
    // G21F_k = -(SI_cc ZF_k + SI_{c,c+1} ZL_k), G21L_k = -(SI_{c+1,c} ZF_k + SI_{c+1,c+1} ZL_k)
    for (int k = 0; k < L; ++k) {
      RC(gemm2(OP_N, OP_N, b, MONE, SIcc, bb, Tb, q.ZF + k * bb, Lb, nCLb, ZERO, nullptr, 0, 0,
               q.G21F + k * bb, Lb, nCLb, g.nC, batch, s));
      if (g.right) {
        RC(gemm2(OP_N, OP_N, b, MONE, SIu + g.c0 * bb, bb, T1b, q.ZL + k * bb, Lb, nCLb, ONE,
                 q.G21F + k * bb, Lb, nCLb, q.G21F + k * bb, Lb, nCLb, g.nC, batch, s));
        RC(gemm2(OP_N, OP_N, b, MONE, SIl + g.c0 * bb, bb, T1b, q.ZF + k * bb, Lb, nCLb, ZERO,
                 nullptr, 0, 0, q.G21L + k * bb, Lb, nCLb, g.nC, batch, s));
        RC(gemm2(OP_N, OP_N, b, MONE, SId + (g.c0 + 1) * bb, bb, Tb, q.ZL + k * bb, Lb, nCLb, ONE,
                 q.G21L + k * bb, Lb, nCLb, q.G21L + k * bb, Lb, nCLb, g.nC, batch, s));
      }
    }
    // boundary blocks: G[f_c, s_c] = -(YF_0 SI_cc + YL_0 SI_{c+1,c}) -> low[s_c]
    RC(gemm2(OP_N, OP_N, b, MONE, q.YF, Lb, nCLb, SIcc, bb, Tb, ZERO, nullptr, 0, 0, q.tmp, bb, nCb,
             g.nC, batch, s));
    if (g.right)
      RC(gemm2(OP_N, OP_N, b, MONE, q.YL, Lb, nCLb, SIl + g.c0 * bb, bb, T1b, ONE, q.tmp, bb, nCb,
               q.tmp, bb, nCb, g.nC, batch, s));
    RC(zcopy3d_launch(bb, 1, g.nC, batch, ONE, q.tmp, 0, bb, nCb, G.low + g.c0 * cstride, 0,
                      cstride, no, s));
    // G[s_c, f_c] = G21F_0 -> up[s_c]
    RC(zcopy3d_launch(bb, 1, g.nC, batch, ONE, q.G21F, 0, Lb, nCLb, G.up + g.c0 * cstride, 0,
                      cstride, no, s));
    if (g.right) {
      const long long eoff = (long long)(g.c0 * (bs + 1) + L) * bb;   // layer e_c
      // G[e_c, s_{c+1}] = -(YF_{L-1} SI_{c,c+1} + YL_{L-1} SI_{c+1,c+1}) -> up[e_c]
      RC(gemm2(OP_N, OP_N, b, MONE, q.YF + (L - 1) * bb, Lb, nCLb, SIu + g.c0 * bb, bb, T1b, ZERO,
               nullptr, 0, 0, q.tmp, bb, nCb, g.nC, batch, s));
      RC(gemm2(OP_N, OP_N, b, MONE, q.YL + (L - 1) * bb, Lb, nCLb, SId + (g.c0 + 1) * bb, bb, Tb,
               ONE, q.tmp, bb, nCb, q.tmp, bb, nCb, g.nC, batch, s));
      RC(zcopy3d_launch(bb, 1, g.nC, batch, ONE, q.tmp, 0, bb, nCb, G.up + eoff, 0, cstride, no,
                        s));
      // G[s_{c+1}, e_c] = G21L_{L-1} -> low[e_c]
      RC(zcopy3d_launch(bb, 1, g.nC, batch, ONE, q.G21L + (L - 1) * bb, 0, Lb, nCLb, G.low + eoff,
                        0, cstride, no, s));
    }
    // 11 part on the band: G_jk = g_jk - YF_j G21F_k - YL_j G21L_k
    RC(gemm2(OP_N, OP_N, b, MONE, q.YF, bb, Lb, q.G21F, bb, Lb, ONE, q.Gd, bb, Lb, q.Gd, bb, Lb, L,
             Mi, s));
    if (g.right)
      RC(gemm2(OP_N, OP_N, b, MONE, q.YL, bb, Lb, q.G21L, bb, Lb, ONE, q.Gd, bb, Lb, q.Gd, bb, Lb,
               L, Mi, s));
    if (L > 1) {
      // upper (j, j+1) and lower (j+1, j), j < L-1
      RC(gemm2(OP_N, OP_N, b, MONE, q.YF, bb, Lb, q.G21F + bb, bb, Lb, ONE, q.Gu, bb, L1b, q.Gu, bb,
               L1b, L - 1, Mi, s));
      RC(gemm2(OP_N, OP_N, b, MONE, q.YF + bb, bb, Lb, q.G21F, bb, Lb, ONE, q.Gl, bb, L1b, q.Gl, bb,
               L1b, L - 1, Mi, s));
      if (g.right) {
        RC(gemm2(OP_N, OP_N, b, MONE, q.YL, bb, Lb, q.G21L + bb, bb, Lb, ONE, q.Gu, bb, L1b, q.Gu,
                 bb, L1b, L - 1, Mi, s));
        RC(gemm2(OP_N, OP_N, b, MONE, q.YL + bb, bb, Lb, q.G21L, bb, Lb, ONE, q.Gl, bb, L1b, q.Gl,
                 bb, L1b, L - 1, Mi, s));
      }
    }
    // scatter the interior band to the original order
    const long long foff = (long long)(g.c0 * (bs + 1) + 1) * bb;
    RC(zcopy3d_launch(bb, L, g.nC, batch, ONE, q.Gd, bb, Lb, nCLb, G.diag + foff, bb, cstride, nd,
                      s));
    if (L > 1) {
      RC(zcopy3d_launch(bb, L - 1, g.nC, batch, ONE, q.Gu, bb, L1b, (long long)g.nC * L1b,
                        G.up + foff, bb, cstride, no, s));
      RC(zcopy3d_launch(bb, L - 1, g.nC, batch, ONE, q.Gl, bb, L1b, (long long)g.nC * L1b,
                        G.low + foff, bb, cstride, no, s));
    }
  }
  return hipGetLastError() == hipSuccess ? Status::OK : Status::HIP_ERROR;
}

size_t ddrgf_work_bytes(int N, int b, int batch, const int* levels, int nlev) {
  Bump meas{nullptr};
  BtdIn A{nullptr, 0, nullptr, 0, nullptr, 0};
  BtdOut G{nullptr, nullptr, nullptr};
  if (ddrgf_run(N, b, A, G, batch, levels, nlev, nullptr, meas, nullptr) != Status::OK) return 0;
  return meas.off + 4096;
}

int ddrgf(int N, int b, const BtdIn& A, const BtdOut& G, int batch, const int* levels, int nlev,
          int* info, void* work, hipStream_t s) {
  Bump bw{(char*)work};
  return ddrgf_run(N, b, A, G, batch, levels, nlev, info, bw, s);
}

}  // namespace negf
