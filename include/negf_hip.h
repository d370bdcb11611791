/*
 * negf_hip.h -- C ABI of the MI355X (gfx950) block-tridiagonal NEGF library.
 *
 * This is the drop-in boundary for LibNEGF.jl's compute backend: a new
 * src/backend_amd.jl ccalls these entry points where backend_cpu.jl calls
 * LAPACK/BLAS, and the sweep-level entry points replace whole
 * matinvertndiag_* / keldyshndiag calls.  Every citation below is to the
 * reference file and line whose behaviour the entry point reproduces.
 *
 * Conventions
 *   - complex128 = two doubles (re, im), i.e. C99 `double complex` / Julia
 *     ComplexF64, column-major, exactly the reference's Matrix{ComplexF64}.
 *   - All matrix pointers are DEVICE pointers; `stream` is a hipStream_t
 *     (NULL = default stream).  Calls are asynchronous on that stream.
 *   - `*_strided_batched`: instance k of an operand lives at ptr + k*stride
 *     (elements); stride 0 broadcasts one operand to every instance.
 *   - Return value: NEGF_OK (0) or a negative NEGF_* error.  Numerical
 *     singularity is reported per instance through a device `info` array,
 *     LAPACK style: 0 = ok, k > 0 = the 1-based column (per-op) or layer
 *     (sweeps) of the first exactly-zero pivot.  The reference prints and
 *     calls exit() there (src/backend_cpu.jl:151-155); the Julia shim maps
 *     info > 0 to that behaviour.
 *
 * Packed block-tridiagonal layout (the sweeps): N blocks of sizes b_0..b_{N-1}
 *   diag  : A_ii        (b_i x b_i)      at element offset sum_{j<i} b_j^2
 *   upper : A_{i,i+1}   (b_i x b_{i+1})  at element offset sum_{j<i} b_j b_{j+1}
 *   lower : A_{i+1,i}   (b_{i+1} x b_i)  same offsets as upper
 * every block column-major with leading dimension = its row count.  This is
 * BlockMatrix with ndiag = 3 (src/blockmatrix.jl:10-17) flattened into three
 * device arrays; `lower` = NULL means Hermitian storage (isHermitian = true,
 * A_{i+1,i} = A_{i,i+1}^H, src/blockmatrix.jl:105-109).  Outputs use the
 * same layout with tight batch strides (nd = sum b_i^2, no = sum b_i b_{i+1}).
 */
#ifndef NEGF_HIP_H
#define NEGF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NEGF_OK 0
#define NEGF_BAD_ARG (-1)
#define NEGF_TOO_LARGE (-2)
#define NEGF_HIP_ERROR (-3)
#define NEGF_NO_MEMORY (-4)

/* Library version (major*10000 + minor*100 + patch). */
int negf_version(void);
/* Static description of a status code. */
const char* negf_status_string(int status);
/* Frees the cached sweep workspaces and the per-stream second streams of the
 * split-step inverse (the next call recreates them).  At most 64 calling
 * streams per process hold such a second stream at a time; beyond that the
 * split-step inverse returns NEGF_HIP_ERROR until this is called. */
int negf_workspace_release(void);
/* Autotuner knob: tile variant of the MFMA ZGEMM for k > 32, process-wide
 * (0 (default): operands loaded straight into two LDS stages when every tile
 * lies inside the output (k a multiple of 8), with 64x64 tiles (m, n multiples
 * of 64) for launches of at least 512 such tiles, else 32x32 tiles (m, n
 * multiples of 32), else variant 8; 1: 64x64 register-staged,
 * double-buffered; 2: 64x128 one buffer; 3: 64x128 double-buffered; 4: 128x64
 * double-buffered; 5: 64x64 at four workgroups per CU; 6, 7: 64x64 with three,
 * four LDS stages; 8: 64x64 register-staged, one LDS buffer, for every shape;
 * 9, 10: two LDS stages with 64x64, 32x32 tiles wherever they fit).  Results
 * of all variants agree to rounding. */
int negf_gemm_variant(int32_t variant);
/* Autotuner knob: schedule of the batched Gauss-Jordan inverse, process-wide
 * (0: automatic (default); in place, one launch per panel: 1: one column strip
 * per panel, 2: two strips per panel; split steps for n <= 512 (the factoring
 * of the next panel on the caller's stream beside matrix-core updates of the
 * other strips on a second stream): 3, 4, 5: panels of 1, 2, 4 strips).
 * Wider panels: fewer passes over the matrix, more latency-bound factoring
 * per pass. */
int negf_inv_variant(int32_t variant);
/* Autotuner knob: schedule of the band sweep negf_rgf (and negf_rgf_noalloc,
 * negf_bndiag_devices' RGF devices), process-wide.  0: automatic (default):
 * the two-sided recursion for band(A^{-1}) alone on a chain of uniform
 * blocks, N >= 4, at most 16 instances of blocks up to 512 (4 up to 1024, 1
 * above) -- the upward recursion from both
 * ends at once as 2 x batch instances of N/2 + 1 layers, meeting at layer
 * N/2 (half as many dependent inversions, the same RGF identities on each
 * side); 1: the one-sided recursion of bndiag_of_inv_rgf_local!
 * (src/matinvertndiag_rgf.jl:114-200) always; 2: two-sided wherever the
 * chain allows, any batch.  Other values: NEGF_BAD_ARG. */
int negf_rgf_schedule(int32_t schedule);
/* The schedule the inverse runs for (n, batch, variant): out[0] = column strip
 * width, out[1] = strips per panel, out[2] = W buffers of n * out[0] * out[1]
 * elements per instance.  Returns 1 for split steps, 0 in place, < 0 for bad
 * arguments.  No GPU work. */
int negf_zinv_schedule(int32_t n, int32_t batch, int32_t variant, int32_t out[3]);
/* Device workspace bytes of a batched n x n inverse; a workspace for n serves
 * every smaller block size at the same batch.  No GPU work. */
int64_t negf_zinv_workspace(int32_t n, int32_t batch);

/* ------------------------------------------------------------------ per-op */

/* C = alpha op(A) op(B) + beta C, op in {'N','T','C'} (zgemm).
 * Replaces be_gemm! (src/backend_cpu.jl:208-219). */
int negf_zgemm(char transA, char transB, int32_t m, int32_t n, int32_t k,
               const double alpha[2], const void* A, int64_t lda,
               const void* B, int64_t ldb, const double beta[2],
               void* C, int64_t ldc, void* stream);
int negf_zgemm_strided_batched(char transA, char transB, int32_t m, int32_t n, int32_t k,
                               const double alpha[2], const void* A, int64_t lda, int64_t strideA,
                               const void* B, int64_t ldb, int64_t strideB,
                               const double beta[2], void* C, int64_t ldc, int64_t strideC,
                               int32_t batch, void* stream);

/* negf_zgemm_strided_batched in a chosen arithmetic (operands and results
 * stay complex128 in memory): precision 0 = FP64 MFMA (the same call);
 * 1 = FP32 MFMA with FP32 accumulation (operands rounded to FP32; the
 * 4-real-product complex form) where the LDS-DMA kernel applies, FP64
 * otherwise; 2 = FP64 emulated on the int8 matrix cores (Ozaki scheme II:
 * per row of op(A) / column of op(B) a power-of-two scale keeping beta bits,
 * residues modulo n_moduli coprime moduli <= 256, 0 = 13, beta = 45 at
 * k = 512; error at most 2^-beta of the row / column maxima per term);
 * 3 = FP64 with Gauss's three real products per complex product (Re = P - Q,
 * Im = R - P - Q; 25 % fewer MFMAs, normwise error within a small factor of
 * the four-product form) where the LDS-DMA kernel applies, FP64 otherwise.
 * No counterpart in the reference: the reduced-precision products of the
 * mixed-precision energy path. */
int negf_zgemm_prec_strided_batched(char transA, char transB, int32_t m, int32_t n, int32_t k,
                                    const double alpha[2], const void* A, int64_t lda,
                                    int64_t strideA, const void* B, int64_t ldb, int64_t strideB,
                                    const double beta[2], void* C, int64_t ldc, int64_t strideC,
                                    int32_t batch, int32_t precision, int32_t n_moduli,
                                    void* stream);

/* LU with partial pivoting in place, LAPACK layout; ipiv (device, n int32,
 * 1-based) and info (device, 1 int32).  Replaces the getrf inside be_lu!
 * (src/backend_cpu.jl:147-169) and be_getrf! (:255-265, which discards ipiv). */
int negf_zgetrf(int32_t n, void* A, int64_t lda, int32_t* ipiv, int32_t* info, void* stream);
int negf_zgetrf_strided_batched(int32_t n, void* A, int64_t lda, int64_t strideA,
                                int32_t* ipiv, int64_t strideP, int32_t* info,
                                int32_t batch, void* stream);

/* B <- op(LU)^{-1} B, trans in {'N','T','C'} (zgetrs).
 * Replaces be_mldivide! (src/backend_cpu.jl:193-206; the copy Min->Mout is
 * the caller's negf_zcopy).  For 64 <= n <= 2048 (or as the tune table's
 * kind 3 says) B is formed as inv(U) (inv(L) P B), by the recursive
 * triangular inverses and two GEMMs, in a cached per-stream workspace;
 * otherwise by row interchanges and two blocked substitutions. */
int negf_zgetrs(char trans, int32_t n, int32_t nrhs, const void* A, int64_t lda,
                const int32_t* ipiv, void* B, int64_t ldb, void* stream);
int negf_zgetrs_strided_batched(char trans, int32_t n, int32_t nrhs, const void* A, int64_t lda,
                                int64_t strideA, const int32_t* ipiv, int64_t strideP,
                                void* B, int64_t ldb, int64_t strideB, int32_t batch,
                                void* stream);

/* Triangular solve with n or m right-hand sides (ztrsm).
 * Replaces be_trsm! (src/backend_cpu.jl:241-253).  Routed like negf_zgetrs
 * (tune kind 4, key {order of the triangle, right-hand sides}): one GEMM with
 * the triangle's recursive inverse, or the blocked substitution. */
int negf_ztrsm(char side, char uplo, char transA, char diag, int32_t m, int32_t n,
               const double alpha[2], const void* A, int64_t lda, void* B, int64_t ldb,
               void* stream);
int negf_ztrsm_strided_batched(char side, char uplo, char transA, char diag, int32_t m,
                               int32_t n, const double alpha[2], const void* A, int64_t lda,
                               int64_t strideA, void* B, int64_t ldb, int64_t strideB,
                               int32_t batch, void* stream);

/* Inverse from the LU factors in place (zgetri); `work` is a device buffer of
 * n*n complex128 per instance.  ipiv = NULL means identity pivots, which is
 * be_getri! (src/backend_cpu.jl:229-239).  Routed as negf_zgetrs with n
 * right-hand sides: inv(U) inv(L) by the recursive triangular inverses and
 * one GEMM, then the column interchanges, or getrs against the identity. */
int negf_zgetri(int32_t n, void* A, int64_t lda, const int32_t* ipiv, void* work, void* stream);
int negf_zgetri_strided_batched(int32_t n, void* A, int64_t lda, int64_t strideA,
                                const int32_t* ipiv, int64_t strideP, void* work,
                                int32_t batch, void* stream);

/* Explicit inverse by Gauss-Jordan with partial pivoting: Ainv = A^{-1}
 * (A untouched).  The AMD backend's be_lu! stores this inverse in its LU
 * holder, so be_mldivide! becomes one negf_zgemm (src/backend_cpu.jl:147-206);
 * also be_inv (:143-145). */
int negf_zinv_strided_batched(int32_t n, const void* A, int64_t lda, int64_t strideA,
                              void* Ainv, int64_t ldinv, int64_t strideInv, int32_t batch,
                              int32_t* info, void* stream);

/* B = A^H (m x n -> n x m).  Replaces be_ctranspose! (src/backend_cpu.jl:58-68). */
int negf_zadjoint_strided_batched(int32_t m, int32_t n, const void* A, int64_t lda,
                                  int64_t strideA, void* B, int64_t ldb, int64_t strideB,
                                  int32_t batch, void* stream);
/* B = A.  Replaces be_copy_in_hw! (src/backend_cpu.jl:42-44). */
int negf_zcopy_strided_batched(int32_t m, int32_t n, const void* A, int64_t lda,
                               int64_t strideA, void* B, int64_t ldb, int64_t strideB,
                               int32_t batch, void* stream);
/* A[i,j] = (i == j ? diag_value : value).  be_fill! (value = diag_value),
 * be_zero_array, be_identity (src/backend_cpu.jl:50-56, 86-88). */
int negf_zfill_strided_batched(int32_t m, int32_t n, const double value[2],
                               const double diag_value[2], void* A, int64_t lda,
                               int64_t strideA, int32_t batch, void* stream);

/* ------------------------------------------------------------------ sweeps */

/* band(A^{-1}) of a block-tridiagonal A by RGF, batched over `batch`
 * instances of identical block sizes (host array block_sizes[nblocks]).
 * Outputs (may be NULL to skip): g_diag (nd per instance), g_upper / g_lower
 * (no per instance).  info: device int32[batch] or NULL.
 * Replaces bndiag_of_inv_rgf_local! (src/matinvertndiag_rgf.jl:114-200). */
int negf_rgf(int32_t nblocks, const int32_t* block_sizes,
             const void* diag, int64_t diag_bstride,
             const void* upper, int64_t upper_bstride,
             const void* lower, int64_t lower_bstride,
             void* g_diag, void* g_upper, void* g_lower,
             int32_t batch, int32_t* info, void* stream);

/* The w-band of A^{-1} for a block (2w+1)-diagonal A by RGF, natively
 * (bndiag_of_inv_general_rgf!, src/matinvertndiag_general_rgf.jl:31-146:
 * kmax x kmax Schur updates of the band above each eliminated layer, then
 * the downward w-band fill).  Band layout of A_band and G_band: the blocks
 * (i, j) with |i - j| <= w in row-major order of (i, j), each b_i x b_j
 * column-major, packed; G_band's batch stride is the band's element count
 * (negf_general_rgf_band_elements).  The input is not modified (the
 * reference writes its Schur updates into Min, :86). */
int negf_general_rgf(int32_t nblocks, const int32_t* block_sizes, int32_t w, const void* A_band,
                     int64_t A_bstride, void* G_band, int32_t batch, int32_t* info,
                     void* stream);
/* Complex elements of one instance's band in that layout (0 on bad input). */
int64_t negf_general_rgf_band_elements(int32_t nblocks, const int32_t* block_sizes, int32_t w);

/* Every block of A^{-1} by RGF, the reference's buildFullInv option
 * (src/matinvertndiag_rgf.jl:185-199): the band, then G_{i,j} =
 * -G_{i,j-1} t_{j-1} and G_{j,i} = -l_j G_{j-1,i} for j >= i + 2.  g_full:
 * per instance a dense column-major n x n matrix (n = sum of block_sizes,
 * ld n, batch stride n^2).  Other arguments as negf_rgf. */
int negf_rgf_full(int32_t nblocks, const int32_t* block_sizes, const void* diag,
                  int64_t diag_bstride, const void* upper, int64_t upper_bstride,
                  const void* lower, int64_t lower_bstride, void* g_full, int32_t batch,
                  int32_t* info, void* stream);

/* negf_rgf without any device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels and results as negf_rgf; the workspace is a static device array of
 * NEGF_RGF_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.
 * Returns NEGF_TOO_LARGE when the sweep's workspace (about
 * 16 * batch * (sum b_i^2 + 2 sum b_i b_{i+1} + 4 bmax^2 + 3 * 64 * bmax)
 * bytes) does not fit.  Calls that overlap in time on one GPU share the
 * array and race: the caller orders them.  The diagonal-block inverses
 * always take the in-place schedule (no second stream). */
#define NEGF_RGF_NOALLOC_SCRATCH_BYTES (64ll << 20)
int negf_rgf_noalloc(int32_t nblocks, const int32_t* block_sizes,
                     const void* diag, int64_t diag_bstride,
                     const void* upper, int64_t upper_bstride,
                     const void* lower, int64_t lower_bstride,
                     void* g_diag, void* g_upper, void* g_lower,
                     int32_t batch, int32_t* info, void* stream);

/* Recursive Keldysh: band(A^{-1} Sigma A^{-H}) for a Hermitian
 * block-tridiagonal Sigma given in Hermitian storage (diag + upper); output
 * gn_diag / gn_upper in the same storage.  The optional RGF outputs reuse
 * the shared upward pass; gamma_L / gamma_R (b_0 x b_0, b_{N-1} x b_{N-1},
 * may be NULL) add T = Re Tr(Gamma_L G_{1N} Gamma_R G_{1N}^H) into
 * transmission[batch] (device).
 * Replaces keldyshndiag! (src/keldyshndiag.jl:120-131 -> keldyshndiag_v3!
 * :166-318) and, with g_* given, the preceding bndiag_of_inv_rgf_local!. */
int negf_rgf_rkd(int32_t nblocks, const int32_t* block_sizes,
                 const void* diag, int64_t diag_bstride,
                 const void* upper, int64_t upper_bstride,
                 const void* lower, int64_t lower_bstride,
                 const void* sigma_diag, int64_t sigma_diag_bstride,
                 const void* sigma_upper, int64_t sigma_upper_bstride,
                 void* g_diag, void* g_upper, void* g_lower,
                 void* gn_diag, void* gn_upper,
                 const void* gamma_L, int64_t gamma_L_bstride,
                 const void* gamma_R, int64_t gamma_R_bstride,
                 double* transmission,
                 int32_t batch, int32_t* info, void* stream);

/* negf_rgf_rkd with the block products in a reduced arithmetic: precision
 * 1 = FP32 MFMA (FP32 accumulation), 2 = FP64 emulated on the int8 matrix
 * cores with n_moduli moduli (0: 13), 3 = FP64 with three real products per
 * complex product, 0 = negf_rgf_rkd itself; the diagonal
 * inverses stay FP64.  Per instance the estimate est = 8 u sum_i ||g_i||_F
 * (u the products' unit roundoff, g_i the upward pass's inverses) of
 * negf_rgf_keldysh_energies_mixed; instances with est > tol (or not finite)
 * are computed again in FP64 and flagged in fallback.  est_host /
 * fallback_host: host arrays of batch entries, or NULL; n_fallback: host.
 * The mixed-precision variant of the bndiag_of_inv_rgf_local! /
 * keldyshndiag! replacement (src/matinvertndiag_rgf.jl:114-200,
 * src/keldyshndiag.jl:120-131), with the reduced precision the reference's
 * tests name (test/common_to_test.jl:20-26). */
int negf_rgf_rkd_mixed(int32_t nblocks, const int32_t* block_sizes, const void* diag,
                       int64_t diag_bstride, const void* upper, int64_t upper_bstride,
                       const void* lower, int64_t lower_bstride, const void* sigma_diag,
                       int64_t sigma_diag_bstride, const void* sigma_upper,
                       int64_t sigma_upper_bstride, void* g_diag, void* g_upper, void* g_lower,
                       void* gn_diag, void* gn_upper, const void* gamma_L,
                       int64_t gamma_L_bstride, const void* gamma_R, int64_t gamma_R_bstride,
                       double* transmission, int32_t batch, int32_t* info, int32_t precision,
                       int32_t n_moduli, double tol, double* est_host, int32_t* fallback_host,
                       int32_t* n_fallback, void* stream);

/* Energy-batched NEGF workload on one device Hamiltonian (no reference
 * counterpart; the reference runs one energy per call): uniform blocks of
 * size b, H_diag (N blocks, Hermitian), H_upper (N-1 blocks H_{i,i+1}),
 * contact broadenings gamma_L, gamma_R (b x b, wide band).  For every energy
 * E_e (device array) with A(E) = (E + i eta) I - H + (i/2)(Gamma_L (+) Gamma_R):
 *   GR_diag   : diagonal blocks of G^R = A^{-1}        (nE * N*b*b, or NULL)
 *   Gless_diag: diagonal blocks of G^< = G^R Sigma^< G^A with
 *               Sigma^< = i f_L Gamma_L (+) i f_R Gamma_R   (or NULL)
 *   transmission: T(E) = Tr(Gamma_L G^R_{1N} Gamma_R G^A_{N1})  (or NULL)
 *   f(E) = 1 / (1 + exp((E - mu) / kT)).
 * The grid is split into `streams` contiguous groups (0 = auto, at most 4),
 * each run on its own internal stream in waves of `wave` energies (0 = the
 * largest of 96, 64, 48, 32, ... whose workspace fits 90 % of free memory),
 * so one group's latency-bound steps overlap another's
 * GEMMs; `stream` waits for all of them.  info: device int32[nE] or NULL
 * (layer of the first singular Schur complement).  gamma_L and gamma_R must
 * be Hermitian positive semidefinite, as the broadening i(Sigma^R - Sigma^A)
 * of a contact is: G^< is computed from their factors Gamma = F F^H, and
 * when G^< is asked for and either is not (an entry of Gamma - Gamma^H, or of
 * what its pivoted Cholesky factorization leaves at the numerical rank,
 * above 2^-26 of its largest diagonal; NaN included) info[e] = -1 for every
 * energy and G^< and T are not valid. */
int negf_rgf_keldysh_energies(int32_t nblocks, int32_t b,
                              const void* H_diag, const void* H_upper,
                              const void* gamma_L, const void* gamma_R,
                              const double* energies, int32_t n_energies,
                              double eta, double mu_L, double mu_R, double kT,
                              void* GR_diag, void* Gless_diag, double* transmission,
                              int32_t* info, int32_t wave, int32_t streams, void* stream);

/* Bytes of device workspace the energy-batched call uses for a given wave
 * and stream count (for sizing `wave` against free memory). */
int64_t negf_rgf_keldysh_energies_workspace(int32_t nblocks, int32_t b, int32_t n_energies,
                                            int32_t wave, int32_t streams);

/* The energy job of negf_rgf_keldysh_energies with its complex products in
 * a reduced or emulated precision (the diagonal-block inverses stay FP64):
 * precision 1 = FP32 MFMA (4 real products, FP32 accumulation), 2 = FP64
 * emulated on the int8 matrix cores (Ozaki scheme II, n_moduli moduli, 0 =
 * 13), 3 = FP64 with three real products per complex product (Gauss), 0 =
 * FP64.  After the pass, each energy's error estimate
 * est = 8 u sum_i ||g_i||_F (u = 2^-24 for FP32, 2^-beta for the emulation,
 * 2^-51 for the three-product FP64 form,
 * g_i the upward pass's diagonal-block inverses; calibrated against the
 * oracle, see the tests) decides: energies with est > tol (or a non-finite
 * est) are computed again in FP64 and their outputs replace the first ones.
 * est_host[n_energies], fallback_host[n_energies] (1 = re-run) and
 * *n_fallback are host outputs (each may be NULL); the call synchronizes the
 * stream.  Other arguments as negf_rgf_keldysh_energies.  No counterpart in
 * the reference (its reduced-precision intent: FieldType, test/common_to_test.jl:20-26). */
int negf_rgf_keldysh_energies_mixed(int32_t nblocks, int32_t b, const void* H_diag,
                                    const void* H_upper, const void* gamma_L,
                                    const void* gamma_R, const double* energies,
                                    int32_t n_energies, double eta, double mu_L, double mu_R,
                                    double kT, void* GR_diag, void* Gless_diag,
                                    double* transmission, int32_t* info, int32_t wave,
                                    int32_t streams, int32_t precision, int32_t n_moduli,
                                    double tol, double* est_host, int32_t* fallback_host,
                                    int32_t* n_fallback, void* stream);
/* Device workspace bytes negf_rgf_keldysh_energies_mixed uses for a wave. */
int64_t negf_rgf_keldysh_energies_mixed_workspace(int32_t nblocks, int32_t b, int32_t n_energies,
                                                  int32_t wave, int32_t streams, int32_t precision,
                                                  int32_t n_moduli);

/* The same energy job captured once into a HIP graph (no reference
 * counterpart).  _create records every launch of negf_rgf_keldysh_energies
 * with these exact pointers and scalars into a graph with its own workspace
 * (held until _destroy) and instantiates it; nothing runs until _launch.
 * Each _launch replays the whole job on `stream` as one submission: new
 * energies, H or Gamma are taken by writing new values into the same device
 * arrays between launches (eta, mu_L, mu_R, kT are fixed at capture).
 * Launches of one graph must not overlap in time (they share the workspace). */
typedef struct negf_energies_graph negf_energies_graph;
int negf_rgf_keldysh_energies_graph_create(int32_t nblocks, int32_t b,
                                           const void* H_diag, const void* H_upper,
                                           const void* gamma_L, const void* gamma_R,
                                           const double* energies, int32_t n_energies,
                                           double eta, double mu_L, double mu_R, double kT,
                                           void* GR_diag, void* Gless_diag, double* transmission,
                                           int32_t* info, int32_t wave, int32_t streams,
                                           negf_energies_graph** graph);
/* The same capture with the complex products in `precision` (0: FP64, the
 * call above; 3: three real products per complex product, Gauss, in FP64 --
 * the form of negf_rgf_keldysh_energies_mixed with precision 3).  With
 * precision 3 each launch also writes the per-energy estimates
 * est = 8 u sum_i ||g_i||_F (u = 2^-51) into the graph's workspace;
 * negf_energies_graph_estimates copies them to the host.  The graph does not
 * re-run energies itself: a caller whose tolerance an estimate exceeds
 * computes those energies again with negf_rgf_keldysh_energies.  No
 * counterpart in the reference. */
int negf_rgf_keldysh_energies_graph_create_prec(int32_t nblocks, int32_t b,
                                                const void* H_diag, const void* H_upper,
                                                const void* gamma_L, const void* gamma_R,
                                                const double* energies, int32_t n_energies,
                                                double eta, double mu_L, double mu_R, double kT,
                                                void* GR_diag, void* Gless_diag,
                                                double* transmission, int32_t* info, int32_t wave,
                                                int32_t streams, int32_t precision,
                                                negf_energies_graph** graph);
int negf_energies_graph_launch(negf_energies_graph* graph, void* stream);
/* est_host[n_energies] <- the last launch's estimates (synchronizes `stream`);
 * NEGF_BAD_ARG for a graph captured with precision 0. */
int negf_energies_graph_estimates(negf_energies_graph* graph, double* est_host, void* stream);
int negf_energies_graph_destroy(negf_energies_graph* graph);

/* ------------------------------------------------- CSC (the reference ABI) */

/* Memory mirror of the reference's CCscMatrix
 * (src/c_interface/libnegf_c_interface.h:12-19): 0-based CSC; nzval holds
 * 2 doubles (re, im) per entry, i.e. C99 `double complex`.  Arrays returned
 * by the calls below are malloc'd; the caller free()s them. */
typedef struct {
  int32_t m;
  int32_t n;
  int32_t nnz;
  int32_t* colptr;
  int32_t* rowval;
  double* nzval;
} CCscMatrix;

/* Drop-in for the reference's exported function
 * (src/c_interface/libnegf_c_interface.h:22-38, libnegf_c_interface.jl:
 * 85-142): the CSC matrix of the block band of A^{-1} for a CSC block-
 * tridiagonal A (ndiag keys "in"/"out" must be 3), computed on the GPU by
 * DD-RGF or RGF, whichever the tune table (negf_tune_*) holds for the shape,
 * else DD-RGF with ~8 partitions.  On failure it returns
 * nnz = 0 and null arrays (the reference has no error channel). */
CCscMatrix run_bndiag_wrapper(int32_t m, int32_t n, int32_t nnz_in, int32_t* colptr_in,
                              int32_t* rowval_in, double* nzval_in, int32_t n_blocks,
                              int32_t* block_sizes_in, int32_t n_dict, const char** dict_keys_in,
                              int32_t* dict_vals_in, int32_t is_hermitian);

/* The same with a status (NEGF_OK, negative error, > 0 singular layer), a
 * method (0 auto, 1 RGF, 2 DD-RGF) and a stream. */
int negf_run_bndiag_csc(int32_t m, int32_t n, int32_t nnz_in, const int32_t* colptr_in,
                        const int32_t* rowval_in, const double* nzval_in, int32_t n_blocks,
                        const int32_t* block_sizes_in, int32_t n_dict, const char** dict_keys_in,
                        const int32_t* dict_vals_in, int32_t is_hermitian, int32_t method,
                        CCscMatrix* out, void* stream);

/* ----------------------------------------------------------------- DD-RGF */

/* band(A^{-1}) by domain-decomposed RGF, uniform block size b, batched over
 * instances.  Level k splits its chain into tasks [1 separator +
 * block_sizes_d1[k] interior layers] (the last task may be shorter); the
 * interiors of every task of every instance are solved as one batch, the
 * separators' Schur complement by the next level or, after the last level,
 * by RGF.  A given size that leaves an empty last task is NEGF_BAD_ARG.
 * lower = NULL reads A_{i+1,i} as A_{i,i+1}^H.  Outputs and info as negf_rgf
 * (info: the layer of the first singular block).
 * Replaces bndiag_of_inv_ddrgf! (src/matinvertndiag_ddrgf_core.jl:719-758)
 * with the levels that allocate_aux_data_DDRGF's tuner would choose
 * (src/matinvertndiag_ddrgf_setup.jl:4-51). */
int negf_ddrgf(int32_t nblocks, int32_t b, const void* diag, int64_t diag_bstride,
               const void* upper, int64_t upper_bstride, const void* lower, int64_t lower_bstride,
               void* g_diag, void* g_upper, void* g_lower, int32_t batch,
               const int32_t* block_sizes_d1, int32_t nlevels, int32_t* info, void* stream);

/* Device workspace bytes negf_ddrgf needs (0 when the sizes do not split). */
int64_t negf_ddrgf_workspace(int32_t nblocks, int32_t b, int32_t batch,
                             const int32_t* block_sizes_d1, int32_t nlevels);

/* negf_ddrgf for a chain whose block sizes differ (the reference's own DD-RGF
 * test workload, bm_create_synthetic_random: src/blockmatrix.jl:878-919,
 * test/test_matinvertndiag_ddrgf.jl:13-14).  Inputs and outputs in the packed
 * layout of negf_rgf (block_sizes[i] per layer).  Uniform sizes run
 * negf_ddrgf as is; otherwise (nblocks <= 1024) each block is embedded in a
 * bmax x bmax block (identity on the diagonal blocks' padding), which leaves
 * the band of A^{-1} unchanged in the leading parts, and the padded chain is
 * solved by the same batched partitions.  Replaces bndiag_of_inv_ddrgf!
 * (src/matinvertndiag_ddrgf_core.jl:719-758) on uneven BlockMatrix inputs. */
int negf_ddrgf_blocksizes(int32_t nblocks, const int32_t* block_sizes, const void* diag,
                          int64_t diag_bstride, const void* upper, int64_t upper_bstride,
                          const void* lower, int64_t lower_bstride, void* g_diag, void* g_upper,
                          void* g_lower, int32_t batch, const int32_t* block_sizes_d1,
                          int32_t nlevels, int32_t* info, void* stream);

/* Device workspace bytes negf_ddrgf_blocksizes needs (0: invalid sizes or levels). */
int64_t negf_ddrgf_blocksizes_workspace(int32_t nblocks, const int32_t* block_sizes, int32_t batch,
                                        const int32_t* block_sizes_d1, int32_t nlevels);

/* Several devices (chains of different block sizes and counts) in one call:
 * the uneven batched workload of one GPU.  Each job is band(A^{-1}) over its
 * own batch, by RGF or by DD-RGF with the given levels (block_sizes_d1 =
 * NULL: the tune table's choice for (nblocks, max block size, batch), RGF
 * when it holds none; {0}: RGF).  The jobs run concurrently on streams the
 * library keeps per calling stream, forked from `stream` and joined back
 * into it, so work queued on `stream` afterwards sees every result.  At
 * most 8 jobs.  Arguments of a job as negf_ddrgf_blocksizes; returns the
 * first job's error (every job is still joined).  No counterpart in the
 * reference, which runs one BlockMatrix per call (bndiag_of_inv_ddrgf!,
 * src/matinvertndiag_ddrgf_core.jl:719-758): several devices of uneven
 * block size and count in one call. */
typedef struct {
  int32_t nblocks;
  const int32_t* block_sizes;
  const void* diag;
  int64_t diag_bstride;
  const void* upper;
  int64_t upper_bstride;
  const void* lower;
  int64_t lower_bstride;
  void* g_diag;
  void* g_upper;
  void* g_lower;
  int32_t batch;
  const int32_t* block_sizes_d1;
  int32_t nlevels;
  int32_t* info;
} NegfBandJob;

int negf_bndiag_devices(int32_t njobs, const NegfBandJob* jobs, void* stream);
/* negf_bndiag_devices with every block product of the jobs (RGF and DD-RGF
 * alike) in `precision`: 0 = FP64 (the call above), 3 = FP64 with three real
 * products per complex product (Gauss; normwise FP64-accurate, 3/4 of the
 * matrix-core work).  The diagonal-block inverses stay as they are.  No
 * counterpart in the reference (its reduced-precision intent: FieldType,
 * test/common_to_test.jl:20-26). */
int negf_bndiag_devices_prec(int32_t njobs, const NegfBandJob* jobs, int32_t precision,
                             void* stream);

/* -------------------------------------------------------------- tuning */

/* The tune table: measured choices per problem shape, consulted by the
 * launchers whenever no variant is forced (negf_gemm_variant /
 * negf_inv_variant = 0); shapes it does not hold keep the built-in
 * thresholds.  The role of the reference's timing-based tuner opt_params
 * (src/matinvertndiag_ddrgf_utils.jl:357-453), which picks DD-RGF levels by
 * timing them, extended to the GEMM tile and the inverse schedule.
 *   kind 0 (gemm):   key {m, n, k, problems per launch} -> {tile variant}
 *   kind 1 (inv):    key {n, batch}                     -> {inverse variant}
 *   kind 2 (bndiag): key {N, bmax, batch}               -> {blockSizeD1 per level}, {0} = RGF
 *   kind 3 (getrs):  key {n, nrhs}                      -> {1 substitution, 2 explicit inverses}
 *                    (getri: key {n, n})
 *   kind 4 (trsm):   key {order of the triangle, nrhs}  -> {1 substitution, 2 explicit inverse}
 * A lookup matches every key but the last exactly and takes the entry with
 * the largest last key <= the request, else the smallest above it.
 * negf_tune_set replaces an entry with the same key; negf_tune_lookup
 * returns the number of values written (0: nothing for the key); the file
 * of negf_tune_load holds one entry per line, "gemm|inv|bndiag|getrs|trsm keys...
 * values...", '#' starting a comment, and the call returns the entries read. */
int negf_tune_set(int32_t kind, const int32_t* key, int32_t nkey, const int32_t* choice,
                  int32_t nchoice);
int negf_tune_lookup(int32_t kind, const int32_t* key, int32_t nkey, int32_t* choice,
                     int32_t maxchoice);
int negf_tune_clear(void);
int negf_tune_count(void);
int negf_tune_load(const char* path);

/* ------------------------------------------------------- selected inverse */

/* Fill-in pattern of the block factorization (host arrays, 0-based block
 * indices): the input pattern plus every block Block_factorization! creates
 * (src/selected_inverse.jl:41-46).  Writes at most `capacity` entries when
 * out_brow/out_bcol are given; returns the count (or a negative status). */
int negf_block_fill_pattern(int32_t npl, int32_t nnzb, const int32_t* brow, const int32_t* bcol,
                            int32_t capacity, int32_t* out_brow, int32_t* out_bcol);

/* Block LU without inter-block pivoting, in place on a symmetric block
 * pattern that already holds the fill-in: the nnzb assigned b x b blocks are
 * given by host arrays brow/bcol (0-based) and `blocks` (host array of
 * device pointers, leading dimension ld; instance k of a batch at
 * blocks[j] + k*stride).  Diagonal blocks are factored without pivoting
 * (ref_pivoting = 0) or, as the reference does, with partial pivoting whose
 * ipiv is then discarded (ref_pivoting = 1).  info: device int32[batch],
 * 1-based block row of the first zero pivot.
 * Replaces Block_factorization! (src/selected_inverse.jl:9-55). */
int negf_block_factorization(int32_t npl, int32_t b, int32_t nnzb, const int32_t* brow,
                             const int32_t* bcol, void* const* blocks, int64_t ld, int64_t stride,
                             int32_t batch, int32_t ref_pivoting, int32_t* info, void* stream);

/* Selected inverse of the factored pattern, in place: every assigned block
 * becomes the matching block of A^{-1}.  Any number of off-diagonal blocks
 * per row.  Replaces Block_inverse! (src/selected_inverse.jl:139-206). */
int negf_block_inverse(int32_t npl, int32_t b, int32_t nnzb, const int32_t* brow,
                       const int32_t* bcol, void* const* blocks, int64_t ld, int64_t stride,
                       int32_t batch, void* stream);

/* Block_factorization! with a choice of pivoting inside the diagonal blocks:
 * mode 0 = none, 1 = the reference's (partial pivoting, ipiv discarded:
 * right only when no row swap happens), 2 = partial pivoting with the row
 * permutation applied to the block row: S_ii = P_i L_ii U_ii, the U blocks
 * become L_ii^{-1} P A_ij, and the pivots are kept in ipiv (device int32
 * [batch][npl][b], LAPACK's 1-based rows) for negf_block_inverse_pivoted.
 * Mode 2 is correct for any nonsingular diagonal Schur complements; the
 * reference's own routine (src/selected_inverse.jl:9-55 with
 * src/backend_cpu.jl:255-265) is mode 1. */
int negf_block_factorization_pivoted(int32_t npl, int32_t b, int32_t nnzb, const int32_t* brow,
                                     const int32_t* bcol, void* const* blocks, int64_t ld,
                                     int64_t stride, int32_t batch, int32_t mode, int32_t* ipiv,
                                     int32_t* info, void* stream);
/* Block_inverse! after negf_block_factorization_pivoted: ipiv from mode 2
 * (every L_ii^{-1} of the inversion becomes L_ii^{-1} P), or NULL after
 * modes 0 and 1 (then identical to negf_block_inverse). */
int negf_block_inverse_pivoted(int32_t npl, int32_t b, int32_t nnzb, const int32_t* brow,
                               const int32_t* bcol, void* const* blocks, int64_t ld,
                               int64_t stride, int32_t batch, const int32_t* ipiv, void* stream);

/* --------------------------------------------------------- input assembly */

/* M = E S - H - Se for a batch of energies, on packed block-tridiagonal
 * arrays (replaces build_M_from_HS, src/matloader.jl:118-122, which builds
 * it on the sparse host matrices before bm_convert).  energies: device
 * complex128 array [batch] (E, or E + i eta).  H, S: block tridiagonal with
 * both off-diagonal bands (non-Hermitian storage); Se: the contact
 * self-energy in the same layout, or NULL arrays.  Each input array has its
 * own batch stride (0 = shared by all energies).  Outputs M_* are packed per
 * energy (strides nd, no).  Evaluated element by element as
 * ((E * S) - H) - Se with a rounded complex product, as Julia evaluates it. */
int negf_build_M_from_HS(int32_t nblocks, const int32_t* block_sizes, const double* energies,
                         int32_t batch,
                         const void* H_diag, int64_t H_diag_bstride,
                         const void* H_upper, int64_t H_upper_bstride,
                         const void* H_lower, int64_t H_lower_bstride,
                         const void* S_diag, int64_t S_diag_bstride,
                         const void* S_upper, int64_t S_upper_bstride,
                         const void* S_lower, int64_t S_lower_bstride,
                         const void* Se_diag, int64_t Se_diag_bstride,
                         const void* Se_upper, int64_t Se_upper_bstride,
                         const void* Se_lower, int64_t Se_lower_bstride,
                         void* M_diag, void* M_upper, void* M_lower, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NEGF_HIP_H */
