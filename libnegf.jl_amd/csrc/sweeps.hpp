// Sweep-level drivers: packed block-tridiagonal layout and entry points.
#pragma once

#include "negf_internal.hpp"
#include <utility>
#include <vector>

namespace negf {

// Packed block-tridiagonal layout: diag blocks D_i (b_i x b_i) at od[i];
// off-diagonal blocks (A_{i,i+1}: b_i x b_{i+1}; A_{i+1,i}: b_{i+1} x b_i) at
// oo[i]; every block column-major with ld = its row count.
struct Layout {
  int N;
  std::vector<int> b;
  std::vector<long long> od, oo;
  long long nd, no;
  int bmax;
  Layout(int N, const int* sizes);
};

struct BtdIn {                 // A: diag / upper / lower (lower may be null = Hermitian)
  const zc* diag; long long diag_bs;
  const zc* up; long long up_bs;
  const zc* low; long long low_bs;
};
struct BtdOut {                // outputs with tight batch strides (nd / no)
  zc* diag; zc* up; zc* low;
};
struct BtdHerm {               // Hermitian block-tridiagonal Sigma (diag + upper)
  const zc* diag; long long diag_bs;
  const zc* up; long long up_bs;
};
struct GammaIn {
  const zc* gammaL; long long gammaL_bs;
  const zc* gammaR; long long gammaR_bs;
};
struct EnergyIn {
  const zc* Hp_diag;           // H'_ii = H_ii - (i/2) Gamma on the contact layers
  const zc* H_up;              // H_{i,i+1}
  const zc* gammaL; const zc* gammaR;
  const double* energies;      // device, nE
  int nE;
  double eta, muL, muR, kT;
};
struct EnergyOut {
  zc* GR_diag;                 // nE * nd or null
  zc* Gless_diag;              // nE * nd or null
  double* T;                   // nE or null
  int* info;                   // nE or null
};

// Arithmetic of the energy job's products (GemmArgs::prec; the diagonal-block
// inverses stay FP64) and the per-energy error estimate of a reduced-precision
// run: est[e] = 8 u sum_i ||g_i||_F, u the unit of the products (2^-24 for
// FP32, 2^-beta for the emulation), g_i = S_i^{-1} the upward pass's inverses.
struct EnergyPrec {
  int prec;        // 0 = FP64, 1 = FP32 MFMA, 2 = int8-emulated FP64, 3 = FP64 3M
  int nmod;        // moduli of the emulation (0: 13)
  double* est;     // device [nE] or null
};
// u of a precision form at inner dimension k (0 for FP64).
double energy_prec_unit(int prec, int nmod, int k);

size_t sweep_work_bytes(const Layout& L, int batch, bool keldysh);
size_t sweep_mixed_extra_bytes(const Layout& L, int batch, int prec, int nmod);
int sweep_rgf_mixed(const Layout& L, const BtdIn& A, const BtdOut& G, const BtdHerm* sigma,
                    const BtdOut* gn, const GammaIn* gam, double* T, int batch, int* info,
                    void* work, int prec, int nmod, double* est, hipStream_t s);
size_t energies_work_bytes(const Layout& L, int nE, int wave, int streams,
                           const EnergyPrec* pr = nullptr);

// Block (2w+1)-diagonal RGF (band layout: blocks (i, j), |i - j| <= w, in
// row-major order, each b_i x b_j column-major; band_offsets gives offsets).
std::vector<long long> band_offsets(int N, const int* b, int w, long long* total);
size_t general_rgf_work_bytes(int N, const int* b, int w, int batch);
int sweep_general_rgf(int N, const int* b, int w, const zc* A, long long a_bstride, zc* G,
                      int batch, int* info, void* work, hipStream_t s);
// Every block of A^{-1} into a dense column-major n x n matrix per instance
// (n = sum of the block sizes, batch stride n^2).
int sweep_rgf_full(const Layout& L, const BtdIn& A, zc* Gfull, int batch, int* info, void* work,
                   hipStream_t s);
// inplace_inverse: the diagonal-block inverses never take the split steps
// (which run on a second stream the library creates on first use).
// two_sided: band(A^{-1}) alone (no Keldysh pass, no transmission) of a chain
// of uniform blocks at small batch runs the two-sided recursion; the
// workspace must then hold max(sweep_work_bytes, sweep_rgf_two_sided_bytes).
int sweep_rgf(const Layout& L, const BtdIn& A, const BtdOut& G, const BtdHerm* sigma,
              const BtdOut* gn, const GammaIn* gam, double* T, int batch, int* info, void* work,
              hipStream_t s, bool inplace_inverse = false, bool two_sided = false);
// Whether the two-sided recursion applies to (L, batch) under the current
// schedule (uniform blocks, N >= 4, at most 16 instances of blocks <= 512,
// 4 of blocks <= 1024, 1 above, unless forced), and its
// workspace (0 when it does not apply).
extern int g_rgf_schedule;   // 0 automatic, 1 one-sided always, 2 two-sided wherever it applies
bool two_sided_applies(const Layout& L, int batch);
size_t sweep_rgf_two_sided_bytes(const Layout& L, int batch);
// Streams and events of the energy job.  `pool`: the group streams (streams
// > 1).  `aux`: one inverse aux set per group, or null for the library's
// per-stream slots.  `events`: 1 + streams events (fork, then one join per
// group), or null to create and destroy them within the call.  A caller that
// captures the job into a graph passes its own aux sets and events and keeps
// them until the graph is destroyed.
struct EnergyStreams {
  hipStream_t* pool;
  const InvAux* aux;
  hipEvent_t* events;
};
int sweep_energies(const Layout& L, const EnergyIn& in, const EnergyOut& out, int wave,
                   int streams, const EnergyStreams& es, void* work, hipStream_t s,
                   const EnergyPrec* pr = nullptr);

// DD-RGF (uniform block size b; A.low must be given)
size_t ddrgf_work_bytes(int N, int b, int batch, const int* levels, int nlev);
int ddrgf(int N, int b, const BtdIn& A, const BtdOut& G, int batch, const int* levels, int nlev,
          int* info, void* work, hipStream_t s);

// selected inverse on a block-sparse pattern (selinv.hip)
std::vector<std::pair<int, int>> block_fill_pattern(int npl,
                                                    const std::vector<std::pair<int, int>>& in);
size_t selinv_work_bytes(int b, int batch, int max_row_nnz);
// mode: 0 no pivoting, 1 the reference's (pivots discarded), 2 partial
// pivoting within each diagonal block (pivots into ipiv [batch][npl][b],
// which block_inverse then takes; null there = no pivots)
int block_factorization(int npl, int b, int nnzb, const int* brow, const int* bcol,
                        void* const* blocks, long long ld, long long stride, int batch, int mode,
                        int* ipiv, int* info, void* work, hipStream_t s);
int block_inverse(int npl, int b, int nnzb, const int* brow, const int* bcol, void* const* blocks,
                  long long ld, long long stride, int batch, const int* ipiv, void* work,
                  hipStream_t s);

// The arithmetic of every product the sweeps issue from this host thread
// while the scope lives: 0 = FP64 (four real products per complex product),
// 3 = FP64 with three real products (Gauss); the previous setting comes back
// on exit.  (The reduced forms 1 and 2 go through the energy job's own
// scope, which also carries their workspace.)
class ProductScope {
 public:
  explicit ProductScope(int prec);
  ~ProductScope();
  ProductScope(const ProductScope&) = delete;
  ProductScope& operator=(const ProductScope&) = delete;

 private:
  int prec_, nmod_;
  void* work_;
  long long bytes_;
};

int zaxpy_launch(int m, int n, zc value, const zc* scale_dev, long long scale_stride,
                 const zc* A, long long lda, long long strideA, zc* B, long long ldb,
                 long long strideB, int batch, hipStream_t s);

}  // namespace negf
