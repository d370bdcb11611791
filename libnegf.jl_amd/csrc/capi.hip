// extern "C" entry points declared in include/negf_hip.h: argument checks,
// workspace management, and dispatch to the gfx950 kernels.
#include "../../include/negf_hip.h"
#include "negf_internal.hpp"
#include "sweeps.hpp"

#include <map>
#include <mutex>
#include <utility>
#include <algorithm>
#include <vector>

using namespace negf;

namespace {

inline zc zz(const double* v) { return v ? zmake(v[0], v[1]) : zmake(0.0, 0.0); }
inline hipStream_t st(void* s) { return (hipStream_t)s; }

// Cached device workspace per (device, stream); grows on demand.
struct WsKey {
  int dev;
  void* stream;
  int tag;   // 0: sweep / inverse workspace, 1: identity pivots
  bool operator<(const WsKey& o) const {
    if (dev != o.dev) return dev < o.dev;
    if (stream != o.stream) return stream < o.stream;
    return tag < o.tag;
  }
};
struct WsBuf {
  void* p = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mu;
std::map<WsKey, WsBuf> g_ws;

int workspace(void* stream, size_t bytes, void** out, int tag = 0) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return NEGF_HIP_ERROR;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  WsBuf& b = g_ws[WsKey{dev, stream, tag}];
  if (b.bytes < bytes) {
    if (b.p) {
      if (hipStreamSynchronize(st(stream)) != hipSuccess) return NEGF_HIP_ERROR;
      (void)hipFree(b.p);
      b.p = nullptr;
      b.bytes = 0;
    }
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      b.p = nullptr;
      return NEGF_NO_MEMORY;
    }
    b.bytes = bytes;
  }
  *out = b.p;
  return NEGF_OK;
}

bool valid_op(char c) { return c == 'N' || c == 'T' || c == 'C' || c == 'n' || c == 't' || c == 'c'; }
int opc(char c) { return (c == 'N' || c == 'n') ? OP_N : ((c == 'T' || c == 't') ? OP_T : OP_C); }

// Route of the per-op solves (getrs, trsm, getri): the tune table's choice
// (1 substitution, 2 explicit inverses), else the explicit inverses of the
// triangular factors and GEMMs (the blocked substitution is 2 n / 16
// dependent launches) for n >= 128, or n >= 64 with 64 or more right-hand
// sides, when their workspace stays under 1 GiB.  Measured on MI355X (getrs,
// one instance): 2.3x (n = 512, 1 column) to 5x (512 columns) faster from
// n = 128 on, 7 % slower at n = 64 with <= 16 columns.
bool inverse_route(int tuned, int n, int nrhs, size_t bytes) {
  const int route = tuned ? tuned : ((n >= 128 || (n >= 64 && nrhs >= 64)) ? 2 : 1);
  return route == 2 && bytes <= ((size_t)1 << 30);
}

bool valid_sizes(int32_t nblocks, const int32_t* sizes) {
  if (nblocks <= 0 || !sizes) return false;
  for (int i = 0; i < nblocks; ++i)
    if (sizes[i] <= 0 || sizes[i] > 2048) return false;
  return true;
}

}  // namespace

extern "C" {

int negf_version(void) { return 1 * 10000 + 0 * 100 + 0; }

const char* negf_status_string(int status) {
  switch (status) {
    case NEGF_OK: return "ok";
    case NEGF_BAD_ARG: return "invalid argument";
    case NEGF_TOO_LARGE: return "problem size beyond the supported limits";
    case NEGF_HIP_ERROR: return "HIP runtime error";
    case NEGF_NO_MEMORY: return "out of device memory";
  }
  return status > 0 ? "singular pivot" : "unknown status";
}

int negf_gemm_variant(int32_t variant) { return zgemm_set_variant(variant); }
int negf_inv_variant(int32_t variant) { return zinv_set_variant(variant); }

int negf_rgf_schedule(int32_t schedule) {
  if (schedule < 0 || schedule > 2) return NEGF_BAD_ARG;
  g_rgf_schedule = schedule;
  return NEGF_OK;
}

int negf_zinv_schedule(int32_t n, int32_t batch, int32_t variant, int32_t out[3]) {
  if (!out) return NEGF_BAD_ARG;
  int S = 0, PS = 0, nbuf = 0;
  const int rc = zinv_schedule(n, batch, variant, &S, &PS, &nbuf);
  if (rc < 0) return rc;
  out[0] = S;
  out[1] = PS;
  out[2] = nbuf;
  return rc;
}

int64_t negf_zinv_workspace(int32_t n, int32_t batch) {
  if (n <= 0 || n > 2048 || batch <= 0) return 0;
  return (int64_t)zinv_work_bytes(n, batch);
}

}  // extern "C"

namespace {
// Streams of negf_bndiag_devices, per (device, calling stream): one per job
// slot, forked from the caller's stream by an event and joined back into it.
constexpr int kMaxJobs = 8;
struct JobPool {
  hipStream_t s[kMaxJobs] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kMaxJobs] = {};
};
std::mutex g_jobpool_mu;
std::map<std::pair<int, void*>, JobPool> g_jobpools;

void pool_destroy(JobPool& p) {
  for (int i = 0; i < kMaxJobs; ++i) {
    if (p.s[i]) (void)hipStreamDestroy(p.s[i]);
    if (p.join[i]) (void)hipEventDestroy(p.join[i]);
  }
  if (p.fork) (void)hipEventDestroy(p.fork);
  p = JobPool{};
}

int pool_for(void* stream, JobPool** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return NEGF_HIP_ERROR;
  std::lock_guard<std::mutex> lk(g_jobpool_mu);
  auto it = g_jobpools.find({dev, stream});
  if (it != g_jobpools.end()) {
    *out = &it->second;
    return NEGF_OK;
  }
  JobPool p;
  bool ok = hipEventCreateWithFlags(&p.fork, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; i < kMaxJobs && ok; ++i)
    ok = hipStreamCreateWithFlags(&p.s[i], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&p.join[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    pool_destroy(p);
    return NEGF_HIP_ERROR;
  }
  *out = &(g_jobpools[{dev, stream}] = p);
  return NEGF_OK;
}
}  // namespace

extern "C" {

int negf_workspace_release(void) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  (void)hipDeviceSynchronize();
  for (auto& kv : g_ws)
    if (kv.second.p) (void)hipFree(kv.second.p);
  g_ws.clear();
  zinv_aux_release_all();
  std::lock_guard<std::mutex> lk2(g_jobpool_mu);
  for (auto& kv : g_jobpools) pool_destroy(kv.second);
  g_jobpools.clear();
  return NEGF_OK;
}

// ------------------------------------------------------------------- per-op
int negf_zgemm_strided_batched(char transA, char transB, int32_t m, int32_t n, int32_t k,
                               const double alpha[2], const void* A, int64_t lda, int64_t strideA,
                               const void* B, int64_t ldb, int64_t strideB, const double beta[2],
                               void* C, int64_t ldc, int64_t strideC, int32_t batch,
                               void* stream) {
  if (!valid_op(transA) || !valid_op(transB) || !alpha || !beta) return NEGF_BAD_ARG;
  GemmArgs g{};
  g.opA = opc(transA); g.opB = opc(transB);
  g.m = m; g.n = n; g.k = k;
  g.alpha = zz(alpha); g.beta = zz(beta);
  g.A = (const zc*)A; g.lda = lda; g.strideA = strideA;
  g.B = (const zc*)B; g.ldb = ldb; g.strideB = strideB;
  g.C = (const zc*)C; g.ldc = ldc; g.strideC = strideC;
  g.D = (zc*)C; g.ldd = ldc; g.strideD = strideC;
  g.batch = batch;
  return zgemm_launch(g, st(stream));
}

int negf_zgemm_prec_strided_batched(char transA, char transB, int32_t m, int32_t n, int32_t k,
                                    const double alpha[2], const void* A, int64_t lda,
                                    int64_t strideA, const void* B, int64_t ldb, int64_t strideB,
                                    const double beta[2], void* C, int64_t ldc, int64_t strideC,
                                    int32_t batch, int32_t precision, int32_t n_moduli,
                                    void* stream) {
  if (!valid_op(transA) || !valid_op(transB) || !alpha || !beta) return NEGF_BAD_ARG;
  if (precision < 0 || precision > 3 || n_moduli < 0 || n_moduli == 1 || n_moduli > 16)
    return NEGF_BAD_ARG;
  GemmArgs g{};
  g.opA = opc(transA); g.opB = opc(transB);
  g.m = m; g.n = n; g.k = k;
  g.alpha = zz(alpha); g.beta = zz(beta);
  g.A = (const zc*)A; g.lda = lda; g.strideA = strideA;
  g.B = (const zc*)B; g.ldb = ldb; g.strideB = strideB;
  g.C = (const zc*)C; g.ldc = ldc; g.strideC = strideC;
  g.D = (zc*)C; g.ldd = ldc; g.strideD = strideC;
  g.batch = batch;
  g.prec = precision;
  g.nmod = n_moduli;
  if (precision == 2 && m > 0 && n > 0 && k > 0 && batch > 0) {
    // the residue planes live in the per-stream cached workspace; C is read
    // in the epilogue, after the operands' conversion, so C may not alias
    // the workspace (it cannot: the workspace is the library's own)
    const size_t bytes = zgemm_emu_work_bytes(m, n, k, strideA == 0 ? 1 : batch,
                                              strideB == 0 ? 1 : batch, n_moduli);
    void* w = nullptr;
    int rc = workspace(stream, bytes, &w, 2);
    if (rc) return rc;
    g.work = w;
    g.work_bytes = (long long)bytes;
  }
  return zgemm_launch(g, st(stream));
}

int negf_zgemm(char transA, char transB, int32_t m, int32_t n, int32_t k, const double alpha[2],
               const void* A, int64_t lda, const void* B, int64_t ldb, const double beta[2],
               void* C, int64_t ldc, void* stream) {
  return negf_zgemm_strided_batched(transA, transB, m, n, k, alpha, A, lda, 0, B, ldb, 0, beta, C,
                                    ldc, 0, 1, stream);
}

int negf_zgetrf_strided_batched(int32_t n, void* A, int64_t lda, int64_t strideA, int32_t* ipiv,
                                int64_t strideP, int32_t* info, int32_t batch, void* stream) {
  if (!A || !ipiv || !info) return n == 0 || batch == 0 ? NEGF_OK : NEGF_BAD_ARG;
  return zgetrf_launch(n, (zc*)A, lda, strideA, ipiv, strideP, info, batch, st(stream));
}

int negf_zgetrf(int32_t n, void* A, int64_t lda, int32_t* ipiv, int32_t* info, void* stream) {
  return negf_zgetrf_strided_batched(n, A, lda, 0, ipiv, 0, info, 1, stream);
}

int negf_zgetrs_strided_batched(char trans, int32_t n, int32_t nrhs, const void* A, int64_t lda,
                                int64_t strideA, const int32_t* ipiv, int64_t strideP, void* B,
                                int64_t ldb, int64_t strideB, int32_t batch, void* stream) {
  if (!valid_op(trans)) return NEGF_BAD_ARG;
  if ((!A || !ipiv || !B) && n > 0 && nrhs > 0 && batch > 0) return NEGF_BAD_ARG;
  if (n > 0 && nrhs > 0 && batch > 0 && batch <= 65535 && lda >= n && ldb >= n) {
    const size_t bytes = zgetrs_inv_work_bytes(n, nrhs, batch);
    if (n <= kPermMaxN && inverse_route(tune_getrs_route(n, nrhs), n, nrhs, bytes)) {
      void* w = nullptr;
      int rc = workspace(stream, bytes, &w, 3);
      if (rc) return rc;
      return zgetrs_inv_launch(trans, n, nrhs, (const zc*)A, lda, strideA, ipiv, strideP, (zc*)B,
                               ldb, strideB, batch, w, st(stream));
    }
  }
  return zgetrs_launch(trans, n, nrhs, (const zc*)A, lda, strideA, ipiv, strideP, (zc*)B, ldb,
                       strideB, batch, st(stream));
}

int negf_zgetrs(char trans, int32_t n, int32_t nrhs, const void* A, int64_t lda,
                const int32_t* ipiv, void* B, int64_t ldb, void* stream) {
  return negf_zgetrs_strided_batched(trans, n, nrhs, A, lda, 0, ipiv, 0, B, ldb, 0, 1, stream);
}

int negf_ztrsm_strided_batched(char side, char uplo, char transA, char diag, int32_t m, int32_t n,
                               const double alpha[2], const void* A, int64_t lda, int64_t strideA,
                               void* B, int64_t ldb, int64_t strideB, int32_t batch,
                               void* stream) {
  if (!alpha) return NEGF_BAD_ARG;
  if ((!A || !B) && m > 0 && n > 0 && batch > 0) return NEGF_BAD_ARG;
  const bool left = side == 'L' || side == 'l';
  const int na = left ? m : n, nrhs = left ? n : m;
  if (m > 0 && n > 0 && batch > 0 && batch <= 65535 && na <= kLuMaxN && lda >= na && ldb >= m &&
      valid_op(transA) && (left || side == 'R' || side == 'r')) {
    const size_t bytes = ztrsm_inv_work_bytes(na, m, n, batch);
    if (inverse_route(tune_trsm_route(na, nrhs), na, nrhs, bytes)) {
      void* w = nullptr;
      int rc = workspace(stream, bytes, &w, 3);
      if (rc) return rc;
      return ztrsm_inv_launch(side, uplo, transA, diag, m, n, zz(alpha), (const zc*)A, lda,
                              strideA, (zc*)B, ldb, strideB, batch, w, st(stream));
    }
  }
  return ztrsm_launch(side, uplo, transA, diag, m, n, zz(alpha), (const zc*)A, lda, strideA,
                      (zc*)B, ldb, strideB, batch, st(stream));
}

int negf_ztrsm(char side, char uplo, char transA, char diag, int32_t m, int32_t n,
               const double alpha[2], const void* A, int64_t lda, void* B, int64_t ldb,
               void* stream) {
  return negf_ztrsm_strided_batched(side, uplo, transA, diag, m, n, alpha, A, lda, 0, B, ldb, 0, 1,
                                    stream);
}

int negf_zgetri_strided_batched(int32_t n, void* A, int64_t lda, int64_t strideA,
                                const int32_t* ipiv, int64_t strideP, void* work, int32_t batch,
                                void* stream) {
  if (n < 0 || batch < 0) return NEGF_BAD_ARG;
  if (n == 0 || batch == 0) return NEGF_OK;
  if (!A || !work) return NEGF_BAD_ARG;
  if (batch <= 65535 && n <= kPermMaxN && lda >= n) {
    // getri as getrs against the identity: the same route
    const size_t bytes = zgetri_inv_work_bytes(n, batch);
    if (inverse_route(tune_getrs_route(n, n), n, n, bytes)) {
      void* w = nullptr;
      int rc = workspace(stream, bytes, &w, 3);
      if (rc) return rc;
      return zgetri_inv_launch(n, (zc*)A, lda, strideA, ipiv, strideP, batch, (zc*)work, w,
                               st(stream));
    }
  }
  const int32_t* ip = ipiv;
  void* ident = nullptr;
  if (!ip) {
    // identity pivots (be_getri!): 1..n, shared by all instances
    std::vector<int32_t> h(n);
    for (int i = 0; i < n; ++i) h[i] = i + 1;
    int rc = workspace(stream, sizeof(int32_t) * (size_t)n, &ident, 1);
    if (rc) return rc;
    if (hipMemcpyAsync(ident, h.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st(stream)) !=
        hipSuccess)
      return NEGF_HIP_ERROR;
    if (hipStreamSynchronize(st(stream)) != hipSuccess) return NEGF_HIP_ERROR;
    ip = (const int32_t*)ident;
    strideP = 0;
  }
  return zgetri_launch(n, (zc*)A, lda, strideA, ip, strideP, batch, (zc*)work, st(stream));
}

int negf_zgetri(int32_t n, void* A, int64_t lda, const int32_t* ipiv, void* work, void* stream) {
  return negf_zgetri_strided_batched(n, A, lda, 0, ipiv, 0, work, 1, stream);
}

int negf_zinv_strided_batched(int32_t n, const void* A, int64_t lda, int64_t strideA, void* Ainv,
                              int64_t ldinv, int64_t strideInv, int32_t batch, int32_t* info,
                              void* stream) {
  if (n < 0 || batch < 0) return NEGF_BAD_ARG;
  if (n == 0 || batch == 0) return NEGF_OK;
  if (!A || !Ainv || !info) return NEGF_BAD_ARG;
  if (n > kLuMaxN) return NEGF_TOO_LARGE;
  if (n > 2048) {
    // beyond the Gauss-Jordan kernels: LU, triangular inverses, product
    if (batch > 1 && (strideA < (int64_t)lda * n || strideInv < (int64_t)ldinv * n))
      return NEGF_BAD_ARG;
    void* w = nullptr;
    int rc = workspace(stream, zinv_large_work_bytes(n), &w);
    if (rc) return rc;
    for (int e = 0; e < batch && rc == NEGF_OK; ++e)
      rc = zinv_large_launch(n, (const zc*)A + e * strideA, lda, (zc*)Ainv + e * strideInv, ldinv,
                             w, info + e, st(stream));
    return rc;
  }
  void* w = nullptr;
  int rc = workspace(stream, zinv_work_bytes(n, batch), &w);
  if (rc) return rc;
  InvWork iw;
  zinv_work_carve(w, n, batch, &iw);
  return zinv_launch(n, (const zc*)A, lda, strideA, (zc*)Ainv, ldinv, strideInv, batch, iw, info,
                     st(stream));
}

int negf_zadjoint_strided_batched(int32_t m, int32_t n, const void* A, int64_t lda,
                                  int64_t strideA, void* B, int64_t ldb, int64_t strideB,
                                  int32_t batch, void* stream) {
  if (lda < (m > 1 ? m : 1) || ldb < (n > 1 ? n : 1)) return NEGF_BAD_ARG;
  return zadjoint_launch(m, n, (const zc*)A, lda, strideA, (zc*)B, ldb, strideB, batch,
                         st(stream));
}

int negf_zcopy_strided_batched(int32_t m, int32_t n, const void* A, int64_t lda, int64_t strideA,
                               void* B, int64_t ldb, int64_t strideB, int32_t batch,
                               void* stream) {
  if (lda < (m > 1 ? m : 1) || ldb < (m > 1 ? m : 1)) return NEGF_BAD_ARG;
  return zcopy2d_launch(m, n, (const zc*)A, lda, strideA, (zc*)B, ldb, strideB, batch,
                        st(stream));
}

int negf_zfill_strided_batched(int32_t m, int32_t n, const double value[2],
                               const double diag_value[2], void* A, int64_t lda, int64_t strideA,
                               int32_t batch, void* stream) {
  if (!value || !diag_value || lda < (m > 1 ? m : 1)) return NEGF_BAD_ARG;
  return zfill_launch(m, n, zz(value), zz(diag_value), (zc*)A, lda, strideA, batch, st(stream));
}

// ------------------------------------------------------------------- sweeps
int negf_rgf_rkd(int32_t nblocks, const int32_t* block_sizes, const void* diag,
                 int64_t diag_bstride, const void* upper, int64_t upper_bstride, const void* lower,
                 int64_t lower_bstride, const void* sigma_diag, int64_t sigma_diag_bstride,
                 const void* sigma_upper, int64_t sigma_upper_bstride, void* g_diag,
                 void* g_upper, void* g_lower, void* gn_diag, void* gn_upper,
                 const void* gamma_L, int64_t gamma_L_bstride, const void* gamma_R,
                 int64_t gamma_R_bstride, double* transmission, int32_t batch, int32_t* info,
                 void* stream) {
  if (!valid_sizes(nblocks, block_sizes) || batch < 0) return NEGF_BAD_ARG;
  if (batch == 0) return NEGF_OK;
  if (!diag || (nblocks > 1 && !upper)) return NEGF_BAD_ARG;
  if ((g_upper || g_lower) && !g_diag) return NEGF_BAD_ARG;
  const bool want_rkd = gn_diag != nullptr;
  if (want_rkd && (!sigma_diag || (nblocks > 1 && (!sigma_upper || !gn_upper))))
    return NEGF_BAD_ARG;
  const bool want_T = transmission != nullptr;
  if (want_T && (!gamma_L || !gamma_R)) return NEGF_BAD_ARG;
  Layout L(nblocks, block_sizes);
  void* w = nullptr;
  // band(A^{-1}) alone may take the two-sided recursion (its own workspace)
  const bool band_only = !want_rkd && !want_T;
  // the per-sweep info accumulator lives behind the sweep workspace
  size_t sw = (sweep_work_bytes(L, batch, want_rkd) + 255) / 256 * 256;
  if (band_only) sw = std::max(sw, (sweep_rgf_two_sided_bytes(L, batch) + 255) / 256 * 256);
  int rc = workspace(stream, sw + sizeof(int32_t) * (size_t)batch + 256, &w);
  if (rc) return rc;
  int32_t* inf = info ? info : (int32_t*)((char*)w + sw);
  BtdIn A{(const zc*)diag, diag_bstride, (const zc*)upper, upper_bstride, (const zc*)lower,
          lower_bstride};
  BtdOut G{(zc*)g_diag, (zc*)g_upper, (zc*)g_lower};
  BtdHerm S{(const zc*)sigma_diag, sigma_diag_bstride, (const zc*)sigma_upper,
            sigma_upper_bstride};
  BtdOut Gn{(zc*)gn_diag, (zc*)gn_upper, nullptr};
  GammaIn gam{(const zc*)gamma_L, gamma_L_bstride, (const zc*)gamma_R, gamma_R_bstride};
  return sweep_rgf(L, A, G, want_rkd ? &S : nullptr, want_rkd ? &Gn : nullptr,
                   want_T ? &gam : nullptr, transmission, batch, inf, w, st(stream), false,
                   band_only);
}

int negf_rgf_rkd_mixed(int32_t nblocks, const int32_t* block_sizes, const void* diag,
                       int64_t diag_bstride, const void* upper, int64_t upper_bstride,
                       const void* lower, int64_t lower_bstride, const void* sigma_diag,
                       int64_t sigma_diag_bstride, const void* sigma_upper,
                       int64_t sigma_upper_bstride, void* g_diag, void* g_upper, void* g_lower,
                       void* gn_diag, void* gn_upper, const void* gamma_L,
                       int64_t gamma_L_bstride, const void* gamma_R, int64_t gamma_R_bstride,
                       double* transmission, int32_t batch, int32_t* info, int32_t precision,
                       int32_t n_moduli, double tol, double* est_host, int32_t* fallback_host,
                       int32_t* n_fallback, void* stream) {
  if (precision < 0 || precision > 3 || n_moduli < 0 || n_moduli == 1 || n_moduli > 16 ||
      !(tol >= 0.0))
    return NEGF_BAD_ARG;
  if (n_fallback) *n_fallback = 0;
  if (precision == 0 || batch <= 0) {
    for (int e = 0; e < batch; ++e) {
      if (est_host) est_host[e] = 0.0;
      if (fallback_host) fallback_host[e] = 0;
    }
    return negf_rgf_rkd(nblocks, block_sizes, diag, diag_bstride, upper, upper_bstride, lower,
                        lower_bstride, sigma_diag, sigma_diag_bstride, sigma_upper,
                        sigma_upper_bstride, g_diag, g_upper, g_lower, gn_diag, gn_upper, gamma_L,
                        gamma_L_bstride, gamma_R, gamma_R_bstride, transmission, batch, info,
                        stream);
  }
  if (!valid_sizes(nblocks, block_sizes)) return NEGF_BAD_ARG;
  if (!diag || (nblocks > 1 && !upper)) return NEGF_BAD_ARG;
  if ((g_upper || g_lower) && !g_diag) return NEGF_BAD_ARG;
  const bool want_rkd = gn_diag != nullptr;
  if (want_rkd && (!sigma_diag || (nblocks > 1 && (!sigma_upper || !gn_upper))))
    return NEGF_BAD_ARG;
  const bool want_T = transmission != nullptr;
  if (want_T && (!gamma_L || !gamma_R)) return NEGF_BAD_ARG;
  Layout L(nblocks, block_sizes);
  const size_t sw = (sweep_work_bytes(L, batch, want_rkd) + 255) / 256 * 256;
  const size_t sx = (sweep_mixed_extra_bytes(L, batch, precision, n_moduli) + 255) / 256 * 256;
  void* w = nullptr;
  int rc = workspace(stream, sw + sx + sizeof(double) * (size_t)batch +
                                 sizeof(int32_t) * (size_t)batch + 512, &w);
  if (rc) return rc;
  double* est_dev = (double*)((char*)w + sw + sx);
  int32_t* inf = info ? info : (int32_t*)(est_dev + batch);
  hipStream_t s = st(stream);
  BtdIn A{(const zc*)diag, diag_bstride, (const zc*)upper, upper_bstride, (const zc*)lower,
          lower_bstride};
  BtdOut G{(zc*)g_diag, (zc*)g_upper, (zc*)g_lower};
  BtdHerm S{(const zc*)sigma_diag, sigma_diag_bstride, (const zc*)sigma_upper,
            sigma_upper_bstride};
  BtdOut Gn{(zc*)gn_diag, (zc*)gn_upper, nullptr};
  GammaIn gam{(const zc*)gamma_L, gamma_L_bstride, (const zc*)gamma_R, gamma_R_bstride};
  // 1. every instance with reduced-precision products
  rc = sweep_rgf_mixed(L, A, G, want_rkd ? &S : nullptr, want_rkd ? &Gn : nullptr,
                       want_T ? &gam : nullptr, transmission, batch, inf, w, precision, n_moduli,
                       est_dev, s);
  if (rc) return rc;
  std::vector<double> est(batch, 0.0);
  if (hipMemcpyAsync(est.data(), est_dev, sizeof(double) * batch, hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return NEGF_HIP_ERROR;
  // 2. instances whose estimate exceeds tol (or is not finite) again in FP64, one at a time
  int nfb = 0;
  for (int e = 0; e < batch; ++e) {
    const bool fb = !(est[e] <= tol);
    if (est_host) est_host[e] = est[e];
    if (fallback_host) fallback_host[e] = fb ? 1 : 0;
    if (!fb) continue;
    ++nfb;
    auto at = [&](const void* p, int64_t bs) -> const zc* {
      return p ? (const zc*)p + (long long)e * bs : nullptr;
    };
    auto out = [&](void* p, long long bs) -> zc* { return p ? (zc*)p + (long long)e * bs : nullptr; };
    BtdIn Ae{at(diag, diag_bstride), diag_bstride, at(upper, upper_bstride), upper_bstride,
             at(lower, lower_bstride), lower_bstride};
    BtdOut Ge{out(g_diag, L.nd), out(g_upper, L.no), out(g_lower, L.no)};
    BtdHerm Se{at(sigma_diag, sigma_diag_bstride), sigma_diag_bstride,
               at(sigma_upper, sigma_upper_bstride), sigma_upper_bstride};
    BtdOut Gne{out(gn_diag, L.nd), out(gn_upper, L.no), nullptr};
    GammaIn game{at(gamma_L, gamma_L_bstride), gamma_L_bstride, at(gamma_R, gamma_R_bstride),
                 gamma_R_bstride};
    rc = sweep_rgf(L, Ae, Ge, want_rkd ? &Se : nullptr, want_rkd ? &Gne : nullptr,
                   want_T ? &game : nullptr, want_T ? transmission + e : nullptr, 1, inf + e, w, s,
                   false);
    if (rc) return rc;
  }
  if (n_fallback) *n_fallback = nfb;
  return NEGF_OK;
}

int negf_rgf(int32_t nblocks, const int32_t* block_sizes, const void* diag, int64_t diag_bstride,
             const void* upper, int64_t upper_bstride, const void* lower, int64_t lower_bstride,
             void* g_diag, void* g_upper, void* g_lower, int32_t batch, int32_t* info,
             void* stream) {
  if (!g_diag) return NEGF_BAD_ARG;
  return negf_rgf_rkd(nblocks, block_sizes, diag, diag_bstride, upper, upper_bstride, lower,
                      lower_bstride, nullptr, 0, nullptr, 0, g_diag, g_upper, g_lower, nullptr,
                      nullptr, nullptr, 0, nullptr, 0, nullptr, batch, info, stream);
}

int negf_rgf_full(int32_t nblocks, const int32_t* block_sizes, const void* diag,
                  int64_t diag_bstride, const void* upper, int64_t upper_bstride,
                  const void* lower, int64_t lower_bstride, void* g_full, int32_t batch,
                  int32_t* info, void* stream) {
  if (!valid_sizes(nblocks, block_sizes) || batch < 0) return NEGF_BAD_ARG;
  if (batch == 0) return NEGF_OK;
  if (!diag || !g_full || (nblocks > 1 && !upper)) return NEGF_BAD_ARG;
  Layout L(nblocks, block_sizes);
  void* w = nullptr;
  const size_t sw = sweep_work_bytes(L, batch, false);
  int rc = workspace(stream, sw + sizeof(int32_t) * (size_t)batch + 256, &w);
  if (rc) return rc;
  int32_t* inf = info ? info : (int32_t*)((char*)w + sw);
  BtdIn A{(const zc*)diag, diag_bstride, (const zc*)upper, upper_bstride, (const zc*)lower,
          lower_bstride};
  return sweep_rgf_full(L, A, (zc*)g_full, batch, inf, w, st(stream));
}

int negf_general_rgf(int32_t nblocks, const int32_t* block_sizes, int32_t w, const void* A_band,
                     int64_t A_bstride, void* G_band, int32_t batch, int32_t* info,
                     void* stream) {
  if (!valid_sizes(nblocks, block_sizes) || w < 1 || batch < 0) return NEGF_BAD_ARG;
  if (batch == 0) return NEGF_OK;
  if (!A_band || !G_band || A_bstride < 0) return NEGF_BAD_ARG;
  void* wk = nullptr;
  const size_t bytes = general_rgf_work_bytes(nblocks, block_sizes, w, batch);
  int rc = workspace(stream, bytes + sizeof(int32_t) * (size_t)batch + 256, &wk);
  if (rc) return rc;
  int32_t* inf = info ? info : (int32_t*)((char*)wk + bytes);
  return sweep_general_rgf(nblocks, block_sizes, w, (const zc*)A_band, A_bstride, (zc*)G_band,
                           batch, inf, wk, st(stream));
}

int64_t negf_general_rgf_band_elements(int32_t nblocks, const int32_t* block_sizes, int32_t w) {
  if (!valid_sizes(nblocks, block_sizes) || w < 1) return 0;
  long long total = 0;
  band_offsets(nblocks, block_sizes, w, &total);
  return (int64_t)total;
}

int64_t negf_rgf_keldysh_energies_workspace(int32_t nblocks, int32_t b, int32_t n_energies,
                                            int32_t wave, int32_t streams) {
  if (nblocks <= 0 || b <= 0 || n_energies <= 0 || wave <= 0 || streams <= 0) return 0;
  std::vector<int32_t> sz(nblocks, b);
  Layout L(nblocks, sz.data());
  return (int64_t)(energies_work_bytes(L, n_energies, wave, streams) + sizeof(zc) * (size_t)L.nd +
                   4096);
}

namespace {
// Worker streams of the energy pipeline, created once per device (default
// priority, non-blocking with respect to the null stream).
constexpr int MAX_STREAMS = 4;
std::mutex g_pool_mu;
std::map<int, std::vector<hipStream_t>> g_pool;

int stream_pool(int n, hipStream_t** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return NEGF_HIP_ERROR;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  auto& v = g_pool[dev];
  while ((int)v.size() < n) {
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return NEGF_HIP_ERROR;
    v.push_back(s);
  }
  *out = v.data();
  return NEGF_OK;
}
}  // namespace

}  // extern "C"

namespace {
// Argument checks, wave size and stream count of the energy pipeline.  The
// wave is the largest of 96, 64, 48, 32, ... whose workspace fits 90 % of
// free memory.  Measured on MI355X (N = 64, b = 512, 256 energies,
// energies/s): one stream with waves of 96 73.5, of 64 72.7; two streams of
// 48 69.6, of 32 60.4 (each batched GEMM already fills the GPU, and the
// Gauss-Jordan factoring, one workgroup per energy and latency-bound, gains
// from more energies per step).
int energies_plan(int32_t nblocks, int32_t b, const void* H_diag, const void* H_upper,
                  const void* gamma_L, const void* gamma_R, const double* energies,
                  int32_t n_energies, double kT, int32_t* wave, int32_t* streams) {
  if (nblocks <= 0 || b <= 0 || b > 2048 || n_energies < 0 || *wave < 0 || *streams < 0)
    return NEGF_BAD_ARG;
  if (n_energies == 0) return NEGF_OK;
  if (!H_diag || (nblocks > 1 && !H_upper) || !gamma_L || !gamma_R || !energies)
    return NEGF_BAD_ARG;
  if (!(kT > 0.0)) return NEGF_BAD_ARG;
  if (*streams == 0) *streams = 1;
  if (*streams > MAX_STREAMS) *streams = MAX_STREAMS;
  if (*wave == 0) {
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) return NEGF_HIP_ERROR;
    const int per = (n_energies + *streams - 1) / *streams;
    static const int sizes[] = {96, 64, 48, 32, 16, 8, 4, 2, 1};
    *wave = 1;
    for (int w : sizes) {
      const int cand = w < per ? w : per;
      if ((size_t)negf_rgf_keldysh_energies_workspace(nblocks, b, n_energies, cand, *streams) <=
          freeb / 10 * 9) {
        *wave = cand;
        break;
      }
    }
  }
  if (*wave > n_energies) *wave = n_energies;
  while (*streams > 1 && (long long)(*streams - 1) * *wave >= n_energies) --*streams;
  return NEGF_OK;
}

// Everything the energy job puts on the stream: H' = H - (i/2) Gamma on the
// contact layers (behind the sweep workspace), then the waves.  Kernels and
// event fork/join only, so it can be captured into a graph.
int energies_enqueue(int32_t nblocks, int32_t b, const void* H_diag, const void* H_upper,
                     const void* gamma_L, const void* gamma_R, const double* energies,
                     int32_t n_energies, double eta, double mu_L, double mu_R, double kT,
                     void* GR_diag, void* Gless_diag, double* transmission, int32_t* info,
                     int32_t wave, int32_t streams, const InvAux* aux, hipEvent_t* events,
                     void* w, hipStream_t s, const EnergyPrec* pr = nullptr) {
  std::vector<int32_t> sz(nblocks, b);
  Layout L(nblocks, sz.data());
  const size_t ew = energies_work_bytes(L, n_energies, wave, streams, pr);
  hipStream_t* pool = nullptr;
  if (streams > 1) {
    int rc = stream_pool(streams, &pool);
    if (rc) return rc;
  }
  zc* Hp = (zc*)((char*)w + (ew + 255) / 256 * 256);
  int rc = zcopy2d_launch((int)L.nd, 1, (const zc*)H_diag, L.nd, 0, Hp, L.nd, 0, 1, s);
  if (rc) return rc;
  rc = zaxpy_launch(b, b, zmake(0.0, -0.5), nullptr, 0, (const zc*)gamma_L, b, 0, Hp, b, 0, 1, s);
  if (rc) return rc;
  rc = zaxpy_launch(b, b, zmake(0.0, -0.5), nullptr, 0, (const zc*)gamma_R, b, 0,
                    Hp + L.od[nblocks - 1], b, 0, 1, s);
  if (rc) return rc;
  EnergyIn in{Hp, (const zc*)H_upper, (const zc*)gamma_L, (const zc*)gamma_R, energies,
              n_energies, eta, mu_L, mu_R, kT};
  EnergyOut out{(zc*)GR_diag, (zc*)Gless_diag, transmission, info};
  EnergyStreams es{pool, aux, events};
  return sweep_energies(L, in, out, wave, streams, es, w, s, pr);
}
}  // namespace

// A captured energy job: its own workspace (held until destroy) and the
// instantiated graph.
// The streams and events the capture records (the inverse's second stream
// and its fork / join events per group, the groups' fork / join events) are
// the graph's own, created before the capture and destroyed with the graph,
// so nothing recorded into it is destroyed or shared while it exists.
struct negf_energies_graph {
  void* work = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<InvAux> aux;
  std::vector<hipEvent_t> events;
  double* est = nullptr;   // per-energy estimates (three-product form), in work
  int32_t n_est = 0;
};

namespace {
// Capture of the energy job into a graph (precision 0: FP64; 3: three real
// products per complex product, with the per-energy estimates written by
// the graph itself); defined after the mixed-precision layout.
int graph_create_impl(int32_t nblocks, int32_t b, const void* H_diag, const void* H_upper,
                      const void* gamma_L, const void* gamma_R, const double* energies,
                      int32_t n_energies, double eta, double mu_L, double mu_R, double kT,
                      void* GR_diag, void* Gless_diag, double* transmission, int32_t* info,
                      int32_t wave, int32_t streams, int32_t precision,
                      negf_energies_graph** out);
}  // namespace

extern "C" {

int negf_rgf_keldysh_energies(int32_t nblocks, int32_t b, const void* H_diag,
                              const void* H_upper, const void* gamma_L, const void* gamma_R,
                              const double* energies, int32_t n_energies, double eta, double mu_L,
                              double mu_R, double kT, void* GR_diag, void* Gless_diag,
                              double* transmission, int32_t* info, int32_t wave, int32_t streams,
                              void* stream) {
  int rc = energies_plan(nblocks, b, H_diag, H_upper, gamma_L, gamma_R, energies, n_energies, kT,
                         &wave, &streams);
  if (rc || n_energies == 0) return rc;
  void* w = nullptr;
  rc = workspace(stream, (size_t)negf_rgf_keldysh_energies_workspace(nblocks, b, n_energies, wave,
                                                                      streams), &w);
  if (rc) return rc;
  return energies_enqueue(nblocks, b, H_diag, H_upper, gamma_L, gamma_R, energies, n_energies,
                          eta, mu_L, mu_R, kT, GR_diag, Gless_diag, transmission, info, wave,
                          streams, nullptr, nullptr, w, st(stream));
}

int negf_rgf_keldysh_energies_graph_create(
    int32_t nblocks, int32_t b, const void* H_diag, const void* H_upper, const void* gamma_L,
    const void* gamma_R, const double* energies, int32_t n_energies, double eta, double mu_L,
    double mu_R, double kT, void* GR_diag, void* Gless_diag, double* transmission, int32_t* info,
    int32_t wave, int32_t streams, negf_energies_graph** out) {
  return graph_create_impl(nblocks, b, H_diag, H_upper, gamma_L, gamma_R, energies, n_energies,
                           eta, mu_L, mu_R, kT, GR_diag, Gless_diag, transmission, info, wave,
                           streams, 0, out);
}

int negf_rgf_keldysh_energies_graph_create_prec(
    int32_t nblocks, int32_t b, const void* H_diag, const void* H_upper, const void* gamma_L,
    const void* gamma_R, const double* energies, int32_t n_energies, double eta, double mu_L,
    double mu_R, double kT, void* GR_diag, void* Gless_diag, double* transmission, int32_t* info,
    int32_t wave, int32_t streams, int32_t precision, negf_energies_graph** out) {
  if (precision != 0 && precision != 3) return NEGF_BAD_ARG;
  return graph_create_impl(nblocks, b, H_diag, H_upper, gamma_L, gamma_R, energies, n_energies,
                           eta, mu_L, mu_R, kT, GR_diag, Gless_diag, transmission, info, wave,
                           streams, precision, out);
}

int negf_energies_graph_estimates(negf_energies_graph* g, double* est_host, void* stream) {
  if (!g || !est_host) return NEGF_BAD_ARG;
  if (!g->est) return NEGF_BAD_ARG;
  hipStream_t s = st(stream);
  if (hipMemcpyAsync(est_host, g->est, sizeof(double) * g->n_est, hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return NEGF_HIP_ERROR;
  return NEGF_OK;
}

int negf_energies_graph_launch(negf_energies_graph* g, void* stream) {
  if (!g || !g->exec) return NEGF_BAD_ARG;
  return hipGraphLaunch(g->exec, st(stream)) == hipSuccess ? NEGF_OK : NEGF_HIP_ERROR;
}

int negf_energies_graph_destroy(negf_energies_graph* g) {
  if (!g) return NEGF_OK;
  (void)hipDeviceSynchronize();
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (auto& a : g->aux) zinv_aux_destroy(&a);
  for (auto& e : g->events)
    if (e) (void)hipEventDestroy(e);
  if (g->work) (void)hipFree(g->work);
  delete g;
  return NEGF_OK;
}

// ------------------------------------------------------------- mixed precision
namespace {

// The fallback energies' outputs, computed in FP64 into temporaries, copied
// to their places in the caller's outputs (idx: their positions).
__global__ void scatter_energy_outputs_kernel(long long nd, const int* idx, const zc* tGR,
                                              const zc* tGL, const double* tT, const int* tI,
                                              zc* GR, zc* GL, double* T, int* I) {
  const int k = blockIdx.y;
  const long long dst = (long long)idx[k];
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < nd; i += 256LL * gridDim.x) {
    if (GR) GR[dst * nd + i] = tGR[k * nd + i];
    if (GL) GL[dst * nd + i] = tGL[k * nd + i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (T) T[dst] = tT[k];
    if (I) I[dst] = tI[k];
  }
}

constexpr int kFallbackWave = 32;

// Workspace: [main | est | fallback energies and indices | (temporaries)].
// The FP64 re-runs start after the whole grid's pass has finished (the host
// reads the estimates first), so they reuse the main region: their own
// energy job from its start, their temporary outputs right behind it when
// both fit there (waves of 32 or more), else in a region of their own at
// the end.
struct MixedLayout {
  size_t main;       // the reduced-precision job, later the FP64 re-runs
  size_t est;        // per-energy estimates
  size_t fb;         // fallback: energies and indices
  size_t tmp_off;    // offset of the re-runs' temporary outputs from the base
  size_t total;
  int fw;            // energies per FP64 re-run
};

MixedLayout mixed_layout(int32_t nblocks, int32_t b, int32_t nE, int32_t wave, int32_t streams,
                         const EnergyPrec& pr) {
  std::vector<int32_t> sz(nblocks, b);
  Layout L(nblocks, sz.data());
  MixedLayout m;
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  m.main = up(energies_work_bytes(L, nE, wave, streams, &pr) + sizeof(zc) * (size_t)L.nd + 4096);
  m.est = up(sizeof(double) * (size_t)nE);
  m.fb = up(sizeof(double) * nE) + up(sizeof(int) * nE);
  m.fw = wave < kFallbackWave ? wave : kFallbackWave;
  // the re-run's energy job: its workspace, then its copy of H' (as in
  // energies_enqueue)
  const size_t rerun =
      up(up(energies_work_bytes(L, m.fw, m.fw, 1, nullptr)) + sizeof(zc) * (size_t)L.nd + 4096);
  const size_t tmp = up(sizeof(zc) * (size_t)m.fw * 2 * L.nd) + up(sizeof(double) * m.fw) +
                     up(sizeof(int) * m.fw);
  if (rerun + tmp <= m.main) {
    m.tmp_off = rerun;
    m.total = m.main + m.est + m.fb;
  } else {
    m.tmp_off = m.main + m.est + m.fb;
    m.total = m.tmp_off + tmp;
  }
  return m;
}

int graph_create_impl(int32_t nblocks, int32_t b, const void* H_diag, const void* H_upper,
                      const void* gamma_L, const void* gamma_R, const double* energies,
                      int32_t n_energies, double eta, double mu_L, double mu_R, double kT,
                      void* GR_diag, void* Gless_diag, double* transmission, int32_t* info,
                      int32_t wave, int32_t streams, int32_t precision,
                      negf_energies_graph** out) {
  if (!out) return NEGF_BAD_ARG;
  *out = nullptr;
  if (n_energies <= 0) return NEGF_BAD_ARG;
  int rc = energies_plan(nblocks, b, H_diag, H_upper, gamma_L, gamma_R, energies, n_energies, kT,
                         &wave, &streams);
  if (rc) return rc;
  // the three-product form: the mixed layout (its estimates after the main
  // region; the FP64 re-run temporaries are not used by a graph)
  double dummy = 0.0;
  EnergyPrec pr{precision, 0, &dummy};
  const MixedLayout ml = mixed_layout(nblocks, b, n_energies, wave, streams, pr);
  const size_t bytes =
      precision == 0
          ? (size_t)negf_rgf_keldysh_energies_workspace(nblocks, b, n_energies, wave, streams)
          : ml.total;
  negf_energies_graph* g = new negf_energies_graph();
  if (hipMalloc(&g->work, bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete g;
    return NEGF_NO_MEMORY;
  }
  if (precision != 0) {
    g->est = (double*)((char*)g->work + ml.main);
    g->n_est = n_energies;
    pr.est = g->est;
  }
  g->aux.assign(streams, InvAux{nullptr, nullptr, nullptr});
  for (auto& a : g->aux)
    if (zinv_aux_create(&a) != NEGF_OK) {
      negf_energies_graph_destroy(g);
      return NEGF_HIP_ERROR;
    }
  g->events.assign(1 + streams, nullptr);
  for (auto& e : g->events)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      negf_energies_graph_destroy(g);
      return NEGF_HIP_ERROR;
    }
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
    negf_energies_graph_destroy(g);
    return NEGF_HIP_ERROR;
  }
  rc = NEGF_OK;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = NEGF_HIP_ERROR;
  if (rc == NEGF_OK) {
    rc = energies_enqueue(nblocks, b, H_diag, H_upper, gamma_L, gamma_R, energies, n_energies, eta,
                          mu_L, mu_R, kT, GR_diag, Gless_diag, transmission, info, wave, streams,
                          g->aux.data(), g->events.data(), g->work, cs,
                          precision == 0 ? nullptr : &pr);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &graph);
    if (rc == NEGF_OK && (e != hipSuccess || !graph)) rc = NEGF_HIP_ERROR;
    g->graph = graph;
    if (rc == NEGF_OK && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
      rc = NEGF_HIP_ERROR;
  }
  (void)hipStreamDestroy(cs);
  if (rc != NEGF_OK) {
    (void)hipGetLastError();
    negf_energies_graph_destroy(g);
    return rc;
  }
  *out = g;
  return NEGF_OK;
}

}  // namespace

extern "C" {

int64_t negf_rgf_keldysh_energies_mixed_workspace(int32_t nblocks, int32_t b, int32_t n_energies,
                                                  int32_t wave, int32_t streams, int32_t precision,
                                                  int32_t n_moduli) {
  if (nblocks <= 0 || b <= 0 || n_energies <= 0 || wave <= 0 || streams <= 0) return 0;
  if (precision < 0 || precision > 3) return 0;
  double dummy = 0.0;
  EnergyPrec pr{precision, n_moduli, &dummy};
  return (int64_t)mixed_layout(nblocks, b, n_energies, wave, streams, pr).total;
}

int negf_rgf_keldysh_energies_mixed(int32_t nblocks, int32_t b, const void* H_diag,
                                    const void* H_upper, const void* gamma_L,
                                    const void* gamma_R, const double* energies,
                                    int32_t n_energies, double eta, double mu_L, double mu_R,
                                    double kT, void* GR_diag, void* Gless_diag,
                                    double* transmission, int32_t* info, int32_t wave,
                                    int32_t streams, int32_t precision, int32_t n_moduli,
                                    double tol, double* est_host, int32_t* fallback_host,
                                    int32_t* n_fallback, void* stream) {
  if (precision < 0 || precision > 3 || n_moduli < 0 || n_moduli == 1 || n_moduli > 16 ||
      !(tol >= 0.0))
    return NEGF_BAD_ARG;
  int32_t w0 = wave, s0 = streams;
  int rc = energies_plan(nblocks, b, H_diag, H_upper, gamma_L, gamma_R, energies, n_energies, kT,
                         &w0, &s0);
  if (rc || n_energies == 0) return rc;
  double dummy = 0.0;
  EnergyPrec pr{precision, n_moduli, &dummy};
  if (wave == 0) {
    // the largest wave whose workspace (with the emulation's planes and the
    // fallback's temporaries) fits 90 % of free memory
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) return NEGF_HIP_ERROR;
    static const int sizes[] = {96, 64, 48, 32, 16, 8, 4, 2, 1};
    w0 = 1;
    for (int c : sizes) {
      const int cand = c < n_energies ? c : n_energies;
      if (mixed_layout(nblocks, b, n_energies, cand, s0, pr).total <= freeb / 10 * 9) {
        w0 = cand;
        break;
      }
    }
  }
  const MixedLayout ml = mixed_layout(nblocks, b, n_energies, w0, s0, pr);
  void* w = nullptr;
  rc = workspace(stream, ml.total, &w);
  if (rc) return rc;
  hipStream_t s = st(stream);
  char* base = (char*)w;
  double* est_dev = (double*)(base + ml.main);
  pr.est = est_dev;
  // 1. the whole grid in the reduced-precision form
  rc = energies_enqueue(nblocks, b, H_diag, H_upper, gamma_L, gamma_R, energies, n_energies, eta,
                        mu_L, mu_R, kT, GR_diag, Gless_diag, transmission, info, w0, s0, nullptr,
                        nullptr, w, s, precision == 0 ? nullptr : &pr);
  if (rc) return rc;
  std::vector<double> est(n_energies, 0.0);
  if (precision != 0) {
    if (hipMemcpyAsync(est.data(), est_dev, sizeof(double) * n_energies, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return NEGF_HIP_ERROR;
  }
  // 2. energies whose estimate exceeds tol (or is not finite) again in FP64
  std::vector<double> hE(n_energies);
  if (hipMemcpy(hE.data(), energies, sizeof(double) * n_energies, hipMemcpyDeviceToHost) !=
      hipSuccess)
    return NEGF_HIP_ERROR;
  std::vector<double> fbE;
  std::vector<int32_t> fbI;
  for (int e = 0; e < n_energies; ++e) {
    const bool fb = precision != 0 && !(est[e] <= tol);
    if (fallback_host) fallback_host[e] = fb ? 1 : 0;
    if (est_host) est_host[e] = est[e];
    if (fb) {
      fbE.push_back(hE[e]);
      fbI.push_back(e);
    }
  }
  if (n_fallback) *n_fallback = (int32_t)fbE.size();
  if (fbE.empty()) return NEGF_OK;
  std::vector<int32_t> sz(nblocks, b);
  Layout L(nblocks, sz.data());
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  char* fb = base + ml.main + ml.est;
  double* dE = (double*)fb;
  int32_t* dI = (int32_t*)(fb + up(sizeof(double) * n_energies));
  zc* tG = (zc*)(base + ml.tmp_off);
  const int fw = ml.fw;
  zc* tGL = tG + (size_t)fw * L.nd;
  double* tT = (double*)((char*)tG + up(sizeof(zc) * (size_t)fw * 2 * L.nd));
  int32_t* tI = (int32_t*)((char*)tT + up(sizeof(double) * fw));
  const int nfb = (int)fbE.size();
  if (hipMemcpyAsync(dE, fbE.data(), sizeof(double) * nfb, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(dI, fbI.data(), sizeof(int32_t) * nfb, hipMemcpyHostToDevice, s) != hipSuccess)
    return NEGF_HIP_ERROR;
  for (int c0 = 0; c0 < nfb; c0 += fw) {
    const int cnt = nfb - c0 < fw ? nfb - c0 : fw;
    rc = energies_enqueue(nblocks, b, H_diag, H_upper, gamma_L, gamma_R, dE + c0, cnt, eta, mu_L,
                          mu_R, kT, GR_diag ? (void*)tG : nullptr, Gless_diag ? (void*)tGL : nullptr,
                          transmission ? tT : nullptr, info ? tI : nullptr, cnt, 1, nullptr,
                          nullptr, w, s, nullptr);
    if (rc) return rc;
    const int gx = (int)((L.nd + 255) / 256 < 64 ? (L.nd + 255) / 256 : 64);
    hipLaunchKernelGGL(scatter_energy_outputs_kernel, dim3(gx, cnt), dim3(256), 0, s, L.nd,
                       dI + c0, tG, tGL, tT, tI, (zc*)GR_diag, (zc*)Gless_diag, transmission, info);
    if (hipGetLastError() != hipSuccess) return NEGF_HIP_ERROR;
  }
  return hipStreamSynchronize(s) == hipSuccess ? NEGF_OK : NEGF_HIP_ERROR;
}

}  // extern "C"

// ---------------------------------------------------------------------- DD-RGF
int64_t negf_ddrgf_workspace(int32_t nblocks, int32_t b, int32_t batch,
                             const int32_t* block_sizes_d1, int32_t nlevels) {
  if (nblocks <= 0 || b <= 0 || batch <= 0 || !block_sizes_d1 || nlevels <= 0) return 0;
  const size_t w = ddrgf_work_bytes(nblocks, b, batch, block_sizes_d1, nlevels);
  if (w == 0) return 0;
  return (int64_t)(w + sizeof(zc) * (size_t)(nblocks - 1) * b * b * batch + 4096);
}

int negf_ddrgf(int32_t nblocks, int32_t b, const void* diag, int64_t diag_bstride,
               const void* upper, int64_t upper_bstride, const void* lower, int64_t lower_bstride,
               void* g_diag, void* g_upper, void* g_lower, int32_t batch,
               const int32_t* block_sizes_d1, int32_t nlevels, int32_t* info, void* stream) {
  if (nblocks <= 0 || b <= 0 || b > 2048 || batch <= 0 || !block_sizes_d1 || nlevels <= 0 ||
      !info)
    return NEGF_BAD_ARG;
  if (!diag || !g_diag || (nblocks > 1 && (!upper || !g_upper || !g_lower))) return NEGF_BAD_ARG;
  const int64_t need = negf_ddrgf_workspace(nblocks, b, batch, block_sizes_d1, nlevels);
  if (need == 0) return NEGF_BAD_ARG;   // a level's block size does not split its chain
  void* w = nullptr;
  int rc = workspace(stream, (size_t)need, &w);
  if (rc) return rc;
  hipStream_t s = st(stream);
  const long long bb = (long long)b * b;
  const void* low = lower;
  long long low_bs = lower_bstride;
  if (!lower && nblocks > 1) {
    // Hermitian storage: A_{i+1,i} = A_{i,i+1}^H, materialized behind the workspace
    zc* lw = (zc*)((char*)w + ((size_t)ddrgf_work_bytes(nblocks, b, batch, block_sizes_d1,
                                                        nlevels) + 255) / 256 * 256);
    for (int e = 0; e < batch; ++e) {
      rc = zadjoint_launch(b, b, (const zc*)upper + e * upper_bstride, b, bb,
                           lw + (long long)e * (nblocks - 1) * bb, b, bb, nblocks - 1, s);
      if (rc) return rc;
    }
    low = lw;
    low_bs = (long long)(nblocks - 1) * bb;
  }
  BtdIn A{(const zc*)diag, diag_bstride, (const zc*)upper, upper_bstride, (const zc*)low, low_bs};
  BtdOut G{(zc*)g_diag, (zc*)g_upper, (zc*)g_lower};
  return ddrgf(nblocks, b, A, G, batch, block_sizes_d1, nlevels, info, w, s);
}

namespace {
#define CAPI_RC(x)         \
  do {                     \
    const int rc_ = (x);   \
    if (rc_) return rc_;   \
  } while (0)

bool pad_table(int32_t nblocks, const int32_t* sizes, PadTable* t, bool* uniform) {
  if (!valid_sizes(nblocks, sizes)) return false;
  int bmax = 0;
  *uniform = true;
  for (int i = 0; i < nblocks; ++i) {
    bmax = sizes[i] > bmax ? sizes[i] : bmax;
    *uniform = *uniform && sizes[i] == sizes[0];
  }
  if (*uniform) {
    t->N = nblocks;
    t->bmax = bmax;
    return true;
  }
  if (nblocks > kPadMaxBlocks) return false;
  t->N = nblocks;
  t->bmax = bmax;
  for (int i = 0; i < nblocks; ++i) t->b[i] = (unsigned short)sizes[i];
  return true;
}
}  // namespace

int64_t negf_ddrgf_blocksizes_workspace(int32_t nblocks, const int32_t* block_sizes, int32_t batch,
                                        const int32_t* block_sizes_d1, int32_t nlevels) {
  PadTable t;
  bool uniform = true;
  if (batch <= 0 || !pad_table(nblocks, block_sizes, &t, &uniform)) return 0;
  if (uniform) return negf_ddrgf_workspace(nblocks, t.bmax, batch, block_sizes_d1, nlevels);
  if (!block_sizes_d1 || nlevels <= 0) return 0;
  const size_t w = ddrgf_work_bytes(nblocks, t.bmax, batch, block_sizes_d1, nlevels);
  if (w == 0) return 0;
  const size_t BB = (size_t)t.bmax * t.bmax;
  const size_t padded = sizeof(zc) * (size_t)(3 * nblocks - 2) * BB * batch;   // A or G
  return (int64_t)(w + 2 * padded + 3 * 256 + 4096);
}

int negf_ddrgf_blocksizes(int32_t nblocks, const int32_t* block_sizes, const void* diag,
                          int64_t diag_bstride, const void* upper, int64_t upper_bstride,
                          const void* lower, int64_t lower_bstride, void* g_diag, void* g_upper,
                          void* g_lower, int32_t batch, const int32_t* block_sizes_d1,
                          int32_t nlevels, int32_t* info, void* stream) {
  PadTable t;
  bool uniform = true;
  if (batch <= 0 || !block_sizes_d1 || nlevels <= 0 || !info ||
      !pad_table(nblocks, block_sizes, &t, &uniform))
    return NEGF_BAD_ARG;
  if (uniform)
    return negf_ddrgf(nblocks, t.bmax, diag, diag_bstride, upper, upper_bstride, lower,
                      lower_bstride, g_diag, g_upper, g_lower, batch, block_sizes_d1, nlevels,
                      info, stream);
  if (!diag || !g_diag || (nblocks > 1 && (!upper || !g_upper || !g_lower)) || batch > 65535)
    return NEGF_BAD_ARG;
  const int64_t need = negf_ddrgf_blocksizes_workspace(nblocks, block_sizes, batch,
                                                       block_sizes_d1, nlevels);
  if (need == 0) return NEGF_BAD_ARG;   // a level's block size does not split its chain
  void* w = nullptr;
  int rc = workspace(stream, (size_t)need, &w);
  if (rc) return rc;
  hipStream_t s = st(stream);
  const int N = nblocks;
  const long long BB = (long long)t.bmax * t.bmax;
  const long long pd = (long long)N * BB, po = (long long)(N - 1) * BB;   // padded batch strides
  long long nd = 0, no = 0;                                               // packed output strides
  for (int i = 0; i < N; ++i) nd += (long long)block_sizes[i] * block_sizes[i];
  for (int i = 0; i + 1 < N; ++i) no += (long long)block_sizes[i] * block_sizes[i + 1];
  auto align = [](size_t x) { return (x + 255) / 256 * 256; };
  char* p = (char*)w;
  zc* Ad = (zc*)p;  p += align(sizeof(zc) * (size_t)pd * batch);
  zc* Au = (zc*)p;  p += align(sizeof(zc) * (size_t)po * batch);
  zc* Al = (zc*)p;  p += align(sizeof(zc) * (size_t)po * batch);
  zc* Gd = (zc*)p;  p += align(sizeof(zc) * (size_t)pd * batch);
  zc* Gu = (zc*)p;  p += align(sizeof(zc) * (size_t)po * batch);
  zc* Gl = (zc*)p;  p += align(sizeof(zc) * (size_t)po * batch);
  CAPI_RC(btd_pad_launch(t, 0, (const zc*)diag, diag_bstride, Ad, pd, batch, s));
  if (N > 1) {
    CAPI_RC(btd_pad_launch(t, 1, (const zc*)upper, upper_bstride, Au, po, batch, s));
    if (lower)
      CAPI_RC(btd_pad_launch(t, 2, (const zc*)lower, lower_bstride, Al, po, batch, s));
    else
      CAPI_RC(btd_pad_launch(t, 3, (const zc*)upper, upper_bstride, Al, po, batch, s));
  }
  BtdIn A{Ad, pd, Au, po, Al, po};
  BtdOut G{Gd, Gu, Gl};
  CAPI_RC(ddrgf(N, t.bmax, A, G, batch, block_sizes_d1, nlevels, info, p, s));
  CAPI_RC(btd_unpad_launch(t, 0, Gd, pd, (zc*)g_diag, nd, batch, s));
  if (N > 1) {
    CAPI_RC(btd_unpad_launch(t, 1, Gu, po, (zc*)g_upper, no, batch, s));
    CAPI_RC(btd_unpad_launch(t, 2, Gl, po, (zc*)g_lower, no, batch, s));
  }
  return NEGF_OK;
}

int negf_bndiag_devices(int32_t njobs, const NegfBandJob* jobs, void* stream) {
  if (njobs < 0 || njobs > kMaxJobs || (njobs > 0 && !jobs)) return NEGF_BAD_ARG;
  if (njobs == 0) return NEGF_OK;
  for (int j = 0; j < njobs; ++j)
    if (!jobs[j].info || jobs[j].batch <= 0 || !valid_sizes(jobs[j].nblocks, jobs[j].block_sizes) ||
        (jobs[j].block_sizes_d1 && jobs[j].nlevels <= 0))
      return NEGF_BAD_ARG;
  JobPool* pool = nullptr;
  int rc = pool_for(stream, &pool);
  if (rc) return rc;
  hipStream_t s = st(stream);
  if (hipEventRecord(pool->fork, s) != hipSuccess) return NEGF_HIP_ERROR;
  int first_err = NEGF_OK;
  for (int j = 0; j < njobs; ++j) {
    const NegfBandJob& J = jobs[j];
    hipStream_t js = pool->s[j];
    if (hipStreamWaitEvent(js, pool->fork, 0) != hipSuccess) return NEGF_HIP_ERROR;
    // the method: the given levels ({0}: RGF), else the tune table's choice
    int32_t lv[4] = {0, 0, 0, 0};
    int nlv = 0;
    if (J.block_sizes_d1) {
      nlv = J.nlevels < 4 ? J.nlevels : 4;
      for (int i = 0; i < nlv; ++i) lv[i] = J.block_sizes_d1[i];
    } else {
      int bmax = 0;
      for (int i = 0; i < J.nblocks; ++i) bmax = J.block_sizes[i] > bmax ? J.block_sizes[i] : bmax;
      const int32_t key[3] = {J.nblocks, bmax, J.batch};
      nlv = negf_tune_lookup(2, key, 3, lv, 4);
      if (nlv < 0) nlv = 0;
    }
    if (nlv > 0 && lv[0] == 0) nlv = 0;
    const int r = nlv > 0
        ? negf_ddrgf_blocksizes(J.nblocks, J.block_sizes, J.diag, J.diag_bstride, J.upper,
                                J.upper_bstride, J.lower, J.lower_bstride, J.g_diag, J.g_upper,
                                J.g_lower, J.batch, lv, nlv, J.info, (void*)js)
        : negf_rgf(J.nblocks, J.block_sizes, J.diag, J.diag_bstride, J.upper, J.upper_bstride,
                   J.lower, J.lower_bstride, J.g_diag, J.g_upper, J.g_lower, J.batch, J.info,
                   (void*)js);
    if (r != NEGF_OK && first_err == NEGF_OK) first_err = r;
    // joined whatever happened, so that the caller's stream never runs ahead of queued work
    if (hipEventRecord(pool->join[j], js) != hipSuccess ||
        hipStreamWaitEvent(s, pool->join[j], 0) != hipSuccess)
      return NEGF_HIP_ERROR;
  }
  return first_err;
}

int negf_bndiag_devices_prec(int32_t njobs, const NegfBandJob* jobs, int32_t precision,
                             void* stream) {
  if (precision != 0 && precision != 3) return NEGF_BAD_ARG;
  ProductScope scope(precision);
  return negf_bndiag_devices(njobs, jobs, stream);
}

// ------------------------------------------------------------ selected inverse
int negf_block_fill_pattern(int32_t npl, int32_t nnzb, const int32_t* brow, const int32_t* bcol,
                            int32_t capacity, int32_t* out_brow, int32_t* out_bcol) {
  if (npl <= 0 || nnzb < 0 || (nnzb > 0 && (!brow || !bcol))) return NEGF_BAD_ARG;
  std::vector<std::pair<int, int>> in;
  for (int k = 0; k < nnzb; ++k) {
    if (brow[k] < 0 || brow[k] >= npl || bcol[k] < 0 || bcol[k] >= npl) return NEGF_BAD_ARG;
    in.push_back({brow[k], bcol[k]});
  }
  auto out = block_fill_pattern(npl, in);
  if (out_brow && out_bcol) {
    if ((int)out.size() > capacity) return NEGF_TOO_LARGE;
    for (size_t k = 0; k < out.size(); ++k) {
      out_brow[k] = out[k].first;
      out_bcol[k] = out[k].second;
    }
  }
  return (int)out.size();
}

int negf_block_factorization(int32_t npl, int32_t b, int32_t nnzb, const int32_t* brow,
                             const int32_t* bcol, void* const* blocks, int64_t ld, int64_t stride,
                             int32_t batch, int32_t ref_pivoting, int32_t* info, void* stream) {
  if (!info || b <= 0 || b > 2048 || batch <= 0) return NEGF_BAD_ARG;
  void* w = nullptr;
  int rc = workspace(stream, selinv_work_bytes(b, batch, 1), &w);
  if (rc) return rc;
  return block_factorization(npl, b, nnzb, brow, bcol, blocks, ld, stride, batch,
                             ref_pivoting != 0 ? 1 : 0, nullptr, info, w, st(stream));
}

int negf_block_factorization_pivoted(int32_t npl, int32_t b, int32_t nnzb, const int32_t* brow,
                                     const int32_t* bcol, void* const* blocks, int64_t ld,
                                     int64_t stride, int32_t batch, int32_t mode, int32_t* ipiv,
                                     int32_t* info, void* stream) {
  if (!info || b <= 0 || b > 2048 || batch <= 0 || mode < 0 || mode > 2 || (mode == 2 && !ipiv))
    return NEGF_BAD_ARG;
  void* w = nullptr;
  int rc = workspace(stream, selinv_work_bytes(b, batch, 1), &w);
  if (rc) return rc;
  return block_factorization(npl, b, nnzb, brow, bcol, blocks, ld, stride, batch, mode, ipiv, info,
                             w, st(stream));
}

int negf_block_inverse(int32_t npl, int32_t b, int32_t nnzb, const int32_t* brow,
                       const int32_t* bcol, void* const* blocks, int64_t ld, int64_t stride,
                       int32_t batch, void* stream) {
  if (b <= 0 || b > 2048 || batch <= 0 || npl <= 0 || nnzb <= 0 || !brow || !bcol) return NEGF_BAD_ARG;
  int maxr = 0;
  std::vector<int> cnt(npl, 0);
  for (int k = 0; k < nnzb; ++k)
    if (bcol[k] > brow[k] && brow[k] >= 0 && brow[k] < npl) maxr = ++cnt[brow[k]] > maxr ? cnt[brow[k]] : maxr;
  void* w = nullptr;
  int rc = workspace(stream, selinv_work_bytes(b, batch, maxr), &w);
  if (rc) return rc;
  return block_inverse(npl, b, nnzb, brow, bcol, blocks, ld, stride, batch, nullptr, w,
                       st(stream));
}

int negf_block_inverse_pivoted(int32_t npl, int32_t b, int32_t nnzb, const int32_t* brow,
                               const int32_t* bcol, void* const* blocks, int64_t ld,
                               int64_t stride, int32_t batch, const int32_t* ipiv, void* stream) {
  if (b <= 0 || b > 2048 || batch <= 0 || npl <= 0 || nnzb <= 0 || !brow || !bcol)
    return NEGF_BAD_ARG;
  int maxr = 0;
  std::vector<int> cnt(npl, 0);
  for (int k = 0; k < nnzb; ++k)
    if (bcol[k] > brow[k] && brow[k] >= 0 && brow[k] < npl) maxr = ++cnt[brow[k]] > maxr ? cnt[brow[k]] : maxr;
  void* w = nullptr;
  int rc = workspace(stream, selinv_work_bytes(b, batch, maxr), &w);
  if (rc) return rc;
  return block_inverse(npl, b, nnzb, brow, bcol, blocks, ld, stride, batch, ipiv, w, st(stream));
}

// ------------------------------------------------------------ input assembly
int negf_build_M_from_HS(int32_t nblocks, const int32_t* block_sizes, const double* energies,
                         int32_t batch, const void* H_diag, int64_t H_diag_bstride,
                         const void* H_upper, int64_t H_upper_bstride, const void* H_lower,
                         int64_t H_lower_bstride, const void* S_diag, int64_t S_diag_bstride,
                         const void* S_upper, int64_t S_upper_bstride, const void* S_lower,
                         int64_t S_lower_bstride, const void* Se_diag, int64_t Se_diag_bstride,
                         const void* Se_upper, int64_t Se_upper_bstride, const void* Se_lower,
                         int64_t Se_lower_bstride, void* M_diag, void* M_upper, void* M_lower,
                         void* stream) {
  if (!valid_sizes(nblocks, block_sizes) || batch <= 0 || !energies) return NEGF_BAD_ARG;
  if (!H_diag || !S_diag || !M_diag) return NEGF_BAD_ARG;
  if (nblocks > 1 && (!H_upper || !H_lower || !S_upper || !S_lower || !M_upper || !M_lower))
    return NEGF_BAD_ARG;
  if (H_diag_bstride < 0 || H_upper_bstride < 0 || H_lower_bstride < 0 || S_diag_bstride < 0 ||
      S_upper_bstride < 0 || S_lower_bstride < 0 || Se_diag_bstride < 0 || Se_upper_bstride < 0 ||
      Se_lower_bstride < 0)
    return NEGF_BAD_ARG;
  Layout L(nblocks, block_sizes);
  hipStream_t s = st(stream);
  const zc* z = (const zc*)energies;
  int rc = zassemble_launch(L.nd, z, (const zc*)S_diag, S_diag_bstride, (const zc*)H_diag,
                            H_diag_bstride, (const zc*)Se_diag, Se_diag_bstride, (zc*)M_diag, L.nd,
                            batch, s);
  if (rc || nblocks == 1) return rc;
  rc = zassemble_launch(L.no, z, (const zc*)S_upper, S_upper_bstride, (const zc*)H_upper,
                        H_upper_bstride, (const zc*)Se_upper, Se_upper_bstride, (zc*)M_upper, L.no,
                        batch, s);
  if (rc) return rc;
  return zassemble_launch(L.no, z, (const zc*)S_lower, S_lower_bstride, (const zc*)H_lower,
                          H_lower_bstride, (const zc*)Se_lower, Se_lower_bstride, (zc*)M_lower,
                          L.no, batch, s);
}

}  // extern "C"
