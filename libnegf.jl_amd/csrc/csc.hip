// The reference's C ABI (src/c_interface/libnegf_c_interface.h:12-38,
// libnegf_c_interface.jl:85-142): a 0-based CSC matrix in, the CSC matrix of
// the block band of its inverse out (malloc'd arrays the caller frees).
// Host-side conversion CSC -> packed blocks (bm_convert, src/blockmatrix.jl:
// 34-72) and back (:82-127), the inverse on the GPU by DD-RGF (uniform
// blocks; the reference's choice, libnegf_c_interface.jl:119) or RGF.
#include "../../include/negf_hip.h"
#include "negf_internal.hpp"
#include "sweeps.hpp"

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace negf;

namespace {

int pick_ddrgf_bs(int N) {
  // aim at ~8 tasks (enough parallel partitions); 0 = plain RGF
  if (N < 8) return 0;
  for (int bs = 1; bs <= N; ++bs) {
    const int per = bs + 1, nr = N / per + 1, left = N - (nr - 1) * per;
    int T = left == 0 ? nr - 1 : (left - 1 == 0 ? -1 : nr);
    if (T > 0 && T <= 8) return T >= 2 ? bs : 0;
  }
  return 0;
}

}  // namespace

extern "C" {

int negf_run_bndiag_csc(int32_t m, int32_t n, int32_t nnz_in, const int32_t* colptr_in,
                        const int32_t* rowval_in, const double* nzval_in, int32_t n_blocks,
                        const int32_t* block_sizes_in, int32_t n_dict, const char** dict_keys_in,
                        const int32_t* dict_vals_in, int32_t is_hermitian, int32_t method,
                        CCscMatrix* out, void* stream) {
  if (!out) return NEGF_BAD_ARG;
  std::memset(out, 0, sizeof(*out));
  if (m <= 0 || m != n || nnz_in < 0 || !colptr_in || (nnz_in > 0 && (!rowval_in || !nzval_in)) ||
      n_blocks <= 0 || !block_sizes_in || method < 0 || method > 2)
    return NEGF_BAD_ARG;
  for (int k = 0; k < n_dict; ++k) {
    if (!dict_keys_in || !dict_vals_in || !dict_keys_in[k]) return NEGF_BAD_ARG;
    const std::string key(dict_keys_in[k]);
    if ((key == "in" || key == "out") && dict_vals_in[k] != 3) return NEGF_BAD_ARG;  // tridiagonal
  }
  const int N = n_blocks;
  std::vector<int> sizes(block_sizes_in, block_sizes_in + N);
  std::vector<long long> off(N + 1, 0);
  for (int i = 0; i < N; ++i) {
    if (sizes[i] <= 0 || sizes[i] > 2048) return NEGF_BAD_ARG;
    off[i + 1] = off[i] + sizes[i];
  }
  if (off[N] != n) return NEGF_BAD_ARG;
  Layout L(N, sizes.data());
  std::vector<int> blk(n);
  for (int i = 0; i < N; ++i)
    for (long long r = off[i]; r < off[i + 1]; ++r) blk[r] = i;
  const size_t nd = (size_t)L.nd, no = (size_t)(L.no > 0 ? L.no : 0);
  std::vector<zc> hd(nd, zmake(0, 0)), hu(no > 0 ? no : 1, zmake(0, 0)),
      hl(no > 0 ? no : 1, zmake(0, 0));
  for (int c = 0; c < n; ++c) {
    if (colptr_in[c] < 0 || colptr_in[c + 1] < colptr_in[c] || colptr_in[c + 1] > nnz_in)
      return NEGF_BAD_ARG;
    const int bj = blk[c];
    for (int k = colptr_in[c]; k < colptr_in[c + 1]; ++k) {
      const int r = rowval_in[k];
      if (r < 0 || r >= n) return NEGF_BAD_ARG;
      const int bi = blk[r];
      const zc v = zmake(nzval_in[2 * k], nzval_in[2 * k + 1]);
      const long long lr = r - off[bi], lc = c - off[bj];
      if (bi == bj) hd[L.od[bi] + lr + lc * sizes[bi]] = v;
      else if (bj == bi + 1) hu[L.oo[bi] + lr + lc * sizes[bi]] = v;
      else if (bi == bj + 1 && !is_hermitian) hl[L.oo[bj] + lr + lc * sizes[bi]] = v;
      // entries outside the band are not read (bm_convert copies band blocks only)
    }
  }
  if (is_hermitian) {
    // Hermitian storage: A_{i+1,i} = A_{i,i+1}^H (blockmatrix.jl:105-109)
    for (int i = 0; i + 1 < N; ++i) {
      const int bi = sizes[i], bj = sizes[i + 1];
      for (int c = 0; c < bi; ++c)
        for (int r = 0; r < bj; ++r) hl[L.oo[i] + r + (long long)c * bj] = zconj(hu[L.oo[i] + c + (long long)r * bi]);
    }
  }
  // device buffers: inputs then outputs
  const size_t tot = nd + 2 * (no > 0 ? no : 1);
  zc* dev = nullptr;
  if (hipMalloc(&dev, 2 * tot * sizeof(zc)) != hipSuccess) {
    (void)hipGetLastError();
    return NEGF_NO_MEMORY;
  }
  zc* din = dev;
  zc* dout = dev + tot;
  const size_t nos = (no > 0 ? no : 1);
  hipStream_t s = (hipStream_t)stream;
  int rc = NEGF_OK;
  int32_t* dinfo = nullptr;
  if (hipMemcpy(din, hd.data(), nd * sizeof(zc), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(din + nd, hu.data(), nos * sizeof(zc), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(din + nd + nos, hl.data(), nos * sizeof(zc), hipMemcpyHostToDevice) != hipSuccess ||
      hipMalloc(&dinfo, sizeof(int32_t)) != hipSuccess)
    rc = NEGF_HIP_ERROR;
  bool uniform = true;
  for (int i = 1; i < N; ++i) uniform = uniform && sizes[i] == sizes[0];
  const int bs = (method == 1 || !uniform) ? 0 : pick_ddrgf_bs(N);
  if (rc == NEGF_OK) {
    if (bs > 0)
      rc = negf_ddrgf(N, sizes[0], din, (int64_t)nd, din + nd, (int64_t)nos, din + nd + nos,
                      (int64_t)nos, dout, dout + nd, dout + nd + nos, 1, &bs, 1, dinfo, stream);
    else
      rc = negf_rgf(N, sizes.data(), din, (int64_t)nd, din + nd, (int64_t)nos, din + nd + nos,
                    (int64_t)nos, dout, dout + nd, dout + nd + nos, 1, dinfo, stream);
  }
  int32_t hinfo = 0;
  if (rc == NEGF_OK) {
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpy(&hinfo, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hd.data(), dout, nd * sizeof(zc), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hu.data(), dout + nd, nos * sizeof(zc), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hl.data(), dout + nd + nos, nos * sizeof(zc), hipMemcpyDeviceToHost) != hipSuccess)
      rc = NEGF_HIP_ERROR;
  }
  (void)hipFree(dev);
  if (dinfo) (void)hipFree(dinfo);
  if (rc != NEGF_OK) return rc;
  if (hinfo != 0) return hinfo;   // > 0: singular layer
  // band -> CSC, column by column in increasing row order, exact zeros dropped
  std::vector<int32_t> colptr(n + 1, 0), rowval;
  std::vector<zc> vals;
  for (int c = 0; c < n; ++c) {
    const int bj = blk[c];
    const long long lc = c - off[bj];
    for (int bi = bj - 1; bi <= bj + 1; ++bi) {
      if (bi < 0 || bi >= N) continue;
      for (long long lr = 0; lr < sizes[bi]; ++lr) {
        zc v;
        if (bi == bj) v = hd[L.od[bi] + lr + lc * sizes[bi]];
        else if (bj == bi + 1) v = hu[L.oo[bi] + lr + lc * sizes[bi]];
        else v = hl[L.oo[bj] + lr + lc * sizes[bi]];
        if (v.x != 0.0 || v.y != 0.0) {
          rowval.push_back((int32_t)(off[bi] + lr));
          vals.push_back(v);
        }
      }
    }
    colptr[c + 1] = (int32_t)rowval.size();
  }
  const size_t nnz = rowval.size();
  out->m = m;
  out->n = n;
  out->nnz = (int32_t)nnz;
  out->colptr = (int32_t*)std::malloc(sizeof(int32_t) * (n + 1));
  out->rowval = (int32_t*)std::malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
  out->nzval = (double*)std::malloc(sizeof(double) * 2 * (nnz > 0 ? nnz : 1));
  if (!out->colptr || !out->rowval || !out->nzval) {
    std::free(out->colptr); std::free(out->rowval); std::free(out->nzval);
    std::memset(out, 0, sizeof(*out));
    return NEGF_NO_MEMORY;
  }
  std::memcpy(out->colptr, colptr.data(), sizeof(int32_t) * (n + 1));
  std::memcpy(out->rowval, rowval.data(), sizeof(int32_t) * nnz);
  std::memcpy(out->nzval, vals.data(), sizeof(double) * 2 * nnz);
  return NEGF_OK;
}

CCscMatrix run_bndiag_wrapper(int32_t m, int32_t n, int32_t nnz_in, int32_t* colptr_in,
                              int32_t* rowval_in, double* nzval_in, int32_t n_blocks,
                              int32_t* block_sizes_in, int32_t n_dict, const char** dict_keys_in,
                              int32_t* dict_vals_in, int32_t is_hermitian) {
  CCscMatrix out;
  const int rc = negf_run_bndiag_csc(m, n, nnz_in, colptr_in, rowval_in, nzval_in, n_blocks,
                                     block_sizes_in, n_dict, dict_keys_in, dict_vals_in,
                                     is_hermitian, 0, &out, nullptr);
  if (rc != NEGF_OK) {
    std::fprintf(stderr, "run_bndiag_wrapper: %s (%d)\n", negf_status_string(rc), rc);
    std::memset(&out, 0, sizeof(out));
    out.m = m;
    out.n = n;
  }
  return out;
}

}  // extern "C"
