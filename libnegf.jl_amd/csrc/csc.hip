// The reference's C ABI (src/c_interface/libnegf_c_interface.h:12-38,
// libnegf_c_interface.jl:85-142): a 0-based CSC matrix in, the CSC matrix of
// the block band of its inverse out (malloc'd arrays the caller frees).
// The CSC arrays go to the device as they are; the conversion to packed
// blocks (bm_convert, src/blockmatrix.jl:34-72) and back (:82-127) runs there
// (one wave per column), the inverse by DD-RGF (uniform blocks; the
// reference's choice, libnegf_c_interface.jl:119) or RGF.  The host only
// checks the arguments, sums the output's per-column counts into colptr and
// copies the result into malloc'd arrays.
#include "../../include/negf_hip.h"
#include "negf_internal.hpp"
#include "sweeps.hpp"

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace negf;

namespace {

// Block geometry on the device: blk[c] = block of row / column c, off[i] =
// first row of block i, sizes[i], od / oo = packed offsets of diagonal /
// off-diagonal block i.
struct Geo {
  const int* blk;
  const long long* off;
  const int* sizes;
  const long long* od;
  const long long* oo;
  int N;
};

// CSC -> packed band: wave w of the grid takes column c = w, its lanes the
// column's entries; entries outside the band are not read (bm_convert copies
// band blocks only); with Hermitian storage the lower blocks are skipped here
// and filled as adjoints afterwards.
__global__ __launch_bounds__(256) void csc_to_band_kernel(int n, const int32_t* colptr,
                                                          const int32_t* rowval,
                                                          const double* nzval, Geo g, int herm,
                                                          zc* d, zc* u, zc* l) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= n) return;
  const int bj = g.blk[c];
  const long long lc = c - g.off[bj];
  for (int k = colptr[c] + lane; k < colptr[c + 1]; k += 64) {
    const int r = rowval[k];
    const int bi = g.blk[r];
    const long long lr = r - g.off[bi];
    const zc v = zmake(nzval[2LL * k], nzval[2LL * k + 1]);
    if (bi == bj) d[g.od[bi] + lr + lc * g.sizes[bi]] = v;
    else if (bj == bi + 1) u[g.oo[bi] + lr + lc * g.sizes[bi]] = v;
    else if (bi == bj + 1 && !herm) l[g.oo[bj] + lr + lc * g.sizes[bi]] = v;
  }
}

// Value of band entry (row r, column c) for rows of the blocks bj - 1 .. bj + 1.
__device__ __forceinline__ zc band_at(const Geo& g, const zc* d, const zc* u, const zc* l, int bi,
                                      int bj, long long lr, long long lc) {
  if (bi == bj) return d[g.od[bi] + lr + lc * g.sizes[bi]];
  if (bj == bi + 1) return u[g.oo[bi] + lr + lc * g.sizes[bi]];
  return l[g.oo[bj] + lr + lc * g.sizes[bi]];
}

// Per column: the band's nonzero count (cnt) or, with colptr given, its
// nonzeros written in increasing row order at colptr[c] (exact zeros
// dropped, as sparse() of a dense block does).  Wave per column; the lanes
// walk the column's band rows in chunks of 64 and compact with a ballot.
__global__ __launch_bounds__(256) void band_to_csc_kernel(int n, Geo g, const zc* d, const zc* u,
                                                          const zc* l, int32_t* cnt,
                                                          const int32_t* colptr, int32_t* rowval,
                                                          double* nzval) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= n) return;
  const int bj = g.blk[c];
  const long long lc = c - g.off[bj];
  const int b0 = bj > 0 ? bj - 1 : bj, b1 = bj + 1 < g.N ? bj + 1 : bj;
  const long long r0 = g.off[b0], r1 = g.off[b1] + g.sizes[b1];
  long long pos = colptr ? colptr[c] : 0;
  for (long long rb = r0; rb < r1; rb += 64) {
    const long long r = rb + lane;
    zc v = zmake(0.0, 0.0);
    if (r < r1) {
      const int bi = g.blk[r];
      v = band_at(g, d, u, l, bi, bj, r - g.off[bi], lc);
    }
    const bool nz = v.x != 0.0 || v.y != 0.0;
    const unsigned long long m = __ballot(nz);
    if (colptr && nz) {
      const long long p = pos + __popcll(m & ((1ull << lane) - 1ull));
      rowval[p] = (int32_t)r;
      nzval[2 * p] = v.x;
      nzval[2 * p + 1] = v.y;
    }
    pos += __popcll(m);
  }
  if (!colptr && lane == 0) cnt[c] = (int32_t)pos;
}

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
  // the CSC structure is checked here (O(n + nnz) integer reads, no values)
  if (colptr_in[0] != 0 || colptr_in[n] != nnz_in) return NEGF_BAD_ARG;
  for (int c = 0; c < n; ++c)
    if (colptr_in[c + 1] < colptr_in[c]) return NEGF_BAD_ARG;
  for (int k = 0; k < nnz_in; ++k)
    if (rowval_in[k] < 0 || rowval_in[k] >= n) return NEGF_BAD_ARG;
  Layout L(N, sizes.data());
  std::vector<int> blk(n);
  for (int i = 0; i < N; ++i)
    for (long long r = off[i]; r < off[i + 1]; ++r) blk[r] = i;
  const size_t nd = (size_t)L.nd, nos = (size_t)(L.no > 0 ? L.no : 1);
  hipStream_t s = (hipStream_t)stream;
  // one device allocation: band in / out, CSC in, geometry, counts, CSC out
  // (the output's nnz is at most the band's element count)
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t band = nd + 2 * nos;
  size_t max_out = 0;
  for (int j = 0; j < N; ++j) {
    const int b0 = j > 0 ? j - 1 : j, b1 = j + 1 < N ? j + 1 : j;
    max_out += (size_t)(off[b1 + 1] - off[b0]) * sizes[j];
  }
  const size_t nnz_alloc = nnz_in > 0 ? (size_t)nnz_in : 1;
  const size_t bytes = up(2 * band * sizeof(zc)) + up(sizeof(int32_t) * (n + 1)) +
                       up(sizeof(int32_t) * nnz_alloc) + up(sizeof(double) * 2 * nnz_alloc) +
                       up(sizeof(int) * n) + 3 * up(sizeof(long long) * (N + 1)) +
                       up(sizeof(int) * N) + up(sizeof(int32_t)) + up(sizeof(int32_t) * (n + 1)) +
                       up(sizeof(int32_t) * max_out) + up(sizeof(double) * 2 * max_out);
  char* dev = nullptr;
  if (hipMalloc((void**)&dev, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return NEGF_NO_MEMORY;
  }
  char* p = dev;
  auto take = [&](size_t b) {
    char* r = p;
    p += up(b);
    return r;
  };
  zc* din = (zc*)take(2 * band * sizeof(zc));
  zc* dout = din + band;
  int32_t* dcp = (int32_t*)take(sizeof(int32_t) * (n + 1));
  int32_t* drv = (int32_t*)take(sizeof(int32_t) * nnz_alloc);
  double* dnz = (double*)take(sizeof(double) * 2 * nnz_alloc);
  int* dblk = (int*)take(sizeof(int) * n);
  long long* doff = (long long*)take(sizeof(long long) * (N + 1));
  long long* dod = (long long*)take(sizeof(long long) * (N + 1));
  long long* doo = (long long*)take(sizeof(long long) * (N + 1));
  int* dsz = (int*)take(sizeof(int) * N);
  int32_t* dinfo = (int32_t*)take(sizeof(int32_t));
  int32_t* dcnt = (int32_t*)take(sizeof(int32_t) * (n + 1));
  int32_t* orv = (int32_t*)take(sizeof(int32_t) * max_out);
  double* onz = (double*)take(sizeof(double) * 2 * max_out);
  std::vector<long long> oo(N + 1, 0), od(L.od.begin(), L.od.end());
  for (int i = 0; i + 1 < N; ++i) oo[i] = L.oo[i];
  int rc = NEGF_OK;
  auto H2D = [&](void* dst, const void* src, size_t b) {
    if (rc == NEGF_OK && b > 0 && hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, s) != hipSuccess)
      rc = NEGF_HIP_ERROR;
  };
  H2D(dcp, colptr_in, sizeof(int32_t) * (n + 1));
  H2D(drv, rowval_in, sizeof(int32_t) * nnz_in);
  H2D(dnz, nzval_in, sizeof(double) * 2 * nnz_in);
  H2D(dblk, blk.data(), sizeof(int) * n);
  H2D(doff, off.data(), sizeof(long long) * (N + 1));
  H2D(dod, od.data(), sizeof(long long) * (N + 1));
  H2D(doo, oo.data(), sizeof(long long) * (N + 1));
  H2D(dsz, sizes.data(), sizeof(int) * N);
  const Geo geo{dblk, doff, dsz, dod, doo, N};
  const unsigned gcol = (unsigned)((n + 3) / 4);
  if (rc == NEGF_OK)
    rc = zfill_launch((int)band, 1, zmake(0, 0), zmake(0, 0), din, (long long)band, 0, 1, s);
  if (rc == NEGF_OK) {
    hipLaunchKernelGGL(csc_to_band_kernel, dim3(gcol), dim3(256), 0, s, n, dcp, drv, dnz, geo,
                       is_hermitian ? 1 : 0, din, din + nd, din + nd + nos);
    if (hipGetLastError() != hipSuccess) rc = NEGF_HIP_ERROR;
  }
  if (rc == NEGF_OK && is_hermitian) {
    // Hermitian storage: A_{i+1,i} = A_{i,i+1}^H (blockmatrix.jl:105-109)
    for (int i = 0; i + 1 < N && rc == NEGF_OK; ++i)
      rc = zadjoint_launch(sizes[i], sizes[i + 1], din + nd + L.oo[i], sizes[i], 0,
                           din + nd + nos + L.oo[i], sizes[i + 1], 0, 1, s);
  }
  // method 0: the tuned choice for (N, bmax, one instance) when the tune
  // table holds one, else ~8 partitions; 1: RGF; 2: DD-RGF (tuned levels,
  // else ~8 partitions).  Uneven sizes run DD-RGF on the padded chain.
  int bmax = 0;
  for (int i = 0; i < N; ++i) bmax = sizes[i] > bmax ? sizes[i] : bmax;
  int32_t lv[4] = {0, 0, 0, 0};
  int nlv = 0;
  if (method != 1) {
    const int32_t key[3] = {N, bmax, 1};
    nlv = negf_tune_lookup(2, key, 3, lv, 4);
    if (nlv <= 0 || (lv[0] == 0 && method == 2)) {
      lv[0] = pick_ddrgf_bs(N);
      nlv = 1;
    }
    if (lv[0] == 0) nlv = 0;
  }
  if (rc == NEGF_OK) {
    if (nlv > 0)
      rc = negf_ddrgf_blocksizes(N, sizes.data(), din, (int64_t)nd, din + nd, (int64_t)nos,
                                 din + nd + nos, (int64_t)nos, dout, dout + nd, dout + nd + nos, 1,
                                 lv, nlv, dinfo, stream);
    else
      rc = negf_rgf(N, sizes.data(), din, (int64_t)nd, din + nd, (int64_t)nos, din + nd + nos,
                    (int64_t)nos, dout, dout + nd, dout + nd + nos, 1, dinfo, stream);
  }
  int32_t hinfo = 0;
  std::vector<int32_t> colptr(n + 1, 0);
  if (rc == NEGF_OK) {
    hipLaunchKernelGGL(band_to_csc_kernel, dim3(gcol), dim3(256), 0, s, n, geo, dout, dout + nd,
                       dout + nd + nos, dcnt, nullptr, nullptr, nullptr);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&hinfo, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(colptr.data() + 1, dcnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = NEGF_HIP_ERROR;
  }
  if (rc == NEGF_OK && hinfo != 0) rc = hinfo;   // > 0: singular layer
  size_t nnz = 0;
  if (rc == NEGF_OK) {
    for (int c = 0; c < n; ++c) colptr[c + 1] += colptr[c];
    nnz = (size_t)colptr[n];
    if (hipMemcpyAsync(dcp, colptr.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s) !=
        hipSuccess)
      rc = NEGF_HIP_ERROR;
  }
  if (rc == NEGF_OK) {
    hipLaunchKernelGGL(band_to_csc_kernel, dim3(gcol), dim3(256), 0, s, n, geo, dout, dout + nd,
                       dout + nd + nos, nullptr, dcp, orv, onz);
    if (hipGetLastError() != hipSuccess) rc = NEGF_HIP_ERROR;
  }
  if (rc == NEGF_OK) {
    out->colptr = (int32_t*)std::malloc(sizeof(int32_t) * (n + 1));
    out->rowval = (int32_t*)std::malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
    out->nzval = (double*)std::malloc(sizeof(double) * 2 * (nnz > 0 ? nnz : 1));
    if (!out->colptr || !out->rowval || !out->nzval) {
      rc = NEGF_NO_MEMORY;
    } else {
      std::memcpy(out->colptr, colptr.data(), sizeof(int32_t) * (n + 1));
      if (nnz > 0 &&
          (hipMemcpyAsync(out->rowval, orv, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, s) !=
               hipSuccess ||
           hipMemcpyAsync(out->nzval, onz, sizeof(double) * 2 * nnz, hipMemcpyDeviceToHost, s) !=
               hipSuccess))
        rc = NEGF_HIP_ERROR;
      if (hipStreamSynchronize(s) != hipSuccess) rc = NEGF_HIP_ERROR;
    }
    if (rc != NEGF_OK) {
      std::free(out->colptr); std::free(out->rowval); std::free(out->nzval);
      std::memset(out, 0, sizeof(*out));
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(dev);
  if (rc != NEGF_OK) return rc;
  out->m = m;
  out->n = n;
  out->nnz = (int32_t)nnz;
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
