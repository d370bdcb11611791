"""Per-op kernels of libnegf_hip.so vs numpy/scipy (FP64 references).

Covers every op the sweeps and the selected inverse call (src/backend_cpu.jl):
zgemm (all nine N/T/C combinations, ragged sizes, broadcast strides),
Gauss-Jordan inverse, getrf (pivots equal LAPACK's), getrs, trsm (all
side/uplo/trans/diag variants), getri, adjoint, fill.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import relfro

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from libnegf_amd import api  # noqa: E402
from libnegf_amd._lib import call, cplx  # noqa: E402
import ctypes as C  # noqa: E402


def _rand(rng, m, n):
    return np.asfortranarray(rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))


def _dev(x):
    return api.col_major(torch.from_numpy(np.ascontiguousarray(x)).to("cuda"))


def _host(t):
    return t.cpu().numpy()


def _op(x, o):
    return x if o == "N" else (x.T if o == "T" else x.conj().T)


@pytest.mark.parametrize("tA", ["N", "T", "C"])
@pytest.mark.parametrize("tB", ["N", "T", "C"])
@pytest.mark.parametrize("mnk", [(64, 64, 64), (37, 70, 19), (130, 5, 67), (1, 1, 1)])
def test_zgemm(tA, tB, mnk):
    m, n, k = mnk
    print(f"zgemm {tA}{tB} {mnk}", flush=True)
    rng = np.random.default_rng(m * 7 + n)
    A = _rand(rng, *((m, k) if tA == "N" else (k, m)))
    B = _rand(rng, *((k, n) if tB == "N" else (n, k)))
    Cm = _rand(rng, m, n)
    alpha, beta = 0.7 - 0.2j, -1.3 + 0.5j
    dC = _dev(Cm)
    api.be_gemm(tA, tB, alpha, _dev(A), _dev(B), beta, dC)
    ref = alpha * _op(A, tA) @ _op(B, tB) + beta * Cm
    assert relfro(_host(dC), ref) < 1e-14


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
@pytest.mark.parametrize("tA", ["N", "T", "C"])
@pytest.mark.parametrize("tB", ["N", "T", "C"])
def test_zgemm_padded_ld_interior_and_edge_tiles(variant, tA, tB):
    """Leading dimensions above the row count and per-instance strides, with
    tiles wholly inside m x n (unchecked staging) next to ragged edge tiles
    and a k that is, then is not, a multiple of the 16-deep K stage."""
    rng = np.random.default_rng(17 + variant)
    batch, pad = 2, 3
    call("negf_gemm_variant", variant)
    try:
        for m, n, k in ((160, 200, 48), (256, 128, 80), (96, 224, 40), (100, 70, 37)):
            print(f"zgemm padded variant={variant} {tA}{tB} {(m, n, k)}", flush=True)
            ra, ca = (m, k) if tA == "N" else (k, m)
            rb, cb = (k, n) if tB == "N" else (n, k)
            lda, ldb, ldc = ra + pad, rb + pad, m + pad
            As = [_rand(rng, lda, ca) for _ in range(batch)]
            Bs = [_rand(rng, ldb, cb) for _ in range(batch)]
            Cs = [_rand(rng, ldc, n) for _ in range(batch)]
            flat = lambda xs: torch.from_numpy(np.stack([x.ravel(order="F") for x in xs])).to("cuda")
            dA, dB, dC = flat(As), flat(Bs), flat(Cs)
            alpha, beta = 0.4 + 1.1j, 0.3 - 0.6j
            call("negf_zgemm_strided_batched", tA.encode(), tB.encode(), m, n, k, cplx(alpha),
                 C.c_void_p(dA.data_ptr()), lda, lda * ca, C.c_void_p(dB.data_ptr()), ldb, ldb * cb,
                 cplx(beta), C.c_void_p(dC.data_ptr()), ldc, ldc * n, batch, None)
            torch.cuda.synchronize()
            out = dC.cpu().numpy()
            for i in range(batch):
                got = out[i].reshape(n, ldc).T
                ref = alpha * _op(As[i][:ra, :], tA) @ _op(Bs[i][:rb, :], tB) + beta * Cs[i][:m, :]
                assert relfro(got[:m, :], ref) < 1e-14
                # rows past m inside the leading dimension stay untouched
                assert np.array_equal(got[m:, :], Cs[i][m:, :])
    finally:
        call("negf_gemm_variant", 0)


def test_zgemm_beta_zero_ignores_nan_and_batched_broadcast():
    print("zgemm batched broadcast", flush=True)
    rng = np.random.default_rng(3)
    b, batch = 96, 5
    A = _rand(rng, b, b)
    Bs = [_rand(rng, b, b) for _ in range(batch)]
    dA = _dev(A)
    dB = torch.from_numpy(np.stack([x.ravel(order="F") for x in Bs])).to("cuda")
    dC = torch.full((batch, b * b), float("nan"), dtype=torch.complex128, device="cuda")
    call("negf_zgemm_strided_batched", b"N", b"C", b, b, b, cplx(1), C.c_void_p(dA.data_ptr()), b, 0,
         C.c_void_p(dB.data_ptr()), b, b * b, cplx(0), C.c_void_p(dC.data_ptr()), b, b * b, batch,
         None)
    torch.cuda.synchronize()
    out = dC.cpu().numpy()
    for i in range(batch):
        assert relfro(out[i].reshape(b, b).T, A @ Bs[i].conj().T) < 1e-14


@pytest.mark.parametrize("n", [1, 7, 64, 100, 256, 300, 512, 600])
def test_zinv(n):
    print(f"zinv n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = _rand(rng, n, n) + 2 * np.sqrt(n) * np.eye(n) * (rng.random() > 0.5)
    inv = _host(api.be_inv(_dev(A)))
    ref = np.linalg.inv(A)
    cond = np.linalg.cond(A)
    assert relfro(inv, ref) < 1e-15 * cond * max(n, 10)


@pytest.mark.parametrize("n", [40, 512, 1000, 1500])
def test_zinv_needs_pivoting(n):
    """A row-permuted near-identity: every column has one entry ~1 and the
    rest ~1e-8, so any pivot other than the column maximum divides by 1e-8
    and the inverse loses ~8 digits; with the maximum chosen the result is
    exact to rounding (cond(A) ~ 1)."""
    print(f"zinv pivoting n={n}", flush=True)
    rng = np.random.default_rng(7 + n)
    perm = rng.permutation(n)
    A = np.asfortranarray((np.eye(n) + 1e-8 * _rand(rng, n, n))[perm])
    inv = _host(api.be_inv(_dev(A)))
    assert relfro(inv, np.linalg.inv(A)) < 1e-13


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("n,batch", [(300, 6), (512, 3), (64, 2), (300, 16), (512, 32),
                                     (256, 8), (448, 16)])
def test_zinv_every_schedule(variant, n, batch):
    """Every schedule of the batched inverse (in place with panels of 1 / 2
    strips, split steps with panels of 1 / 2 / 4 strips, the same with the
    bulk update as a GEMM, automatic) on ragged (n = 300: a 12-column last
    strip, a partial last 64-row chunk; variant 6 falls back to 5 there) and
    full shapes (256 and 448: the GEMM form's left-only and right-only steps
    at other panel counts), against numpy.  Batches that are multiples of 8
    take the split update's per-XCD item order; 512 x 32 is the automatic
    schedule's split regime (the sweep's)."""
    print(f"zinv variant={variant} n={n} batch={batch}", flush=True)
    rng = np.random.default_rng(n + batch)
    As = [_rand(rng, n, n) + (2 * np.sqrt(n) * np.eye(n) if k % 2 else 0) for k in range(batch)]
    dA = torch.from_numpy(np.stack([a.ravel(order="F") for a in As])).to("cuda")
    dI = torch.empty_like(dA)
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    call("negf_inv_variant", variant)
    try:
        call("negf_zinv_strided_batched", n, C.c_void_p(dA.data_ptr()), n, n * n,
             C.c_void_p(dI.data_ptr()), n, n * n, batch, C.c_void_p(info.data_ptr()), None)
        torch.cuda.synchronize()
    finally:
        call("negf_inv_variant", 0)
    out = dI.cpu().numpy()
    assert not info.cpu().numpy().any()
    for k, a in enumerate(As):
        inv = out[k].reshape(n, n).T
        assert relfro(inv, np.linalg.inv(a)) < 1e-15 * np.linalg.cond(a) * max(n, 10)


def test_zinv_singular_reports_info():
    print("zinv singular", flush=True)
    A = np.zeros((8, 8), complex, order="F")
    A[:4, :4] = np.eye(4)
    with pytest.raises(api.SingularError):
        api.be_inv(_dev(A))


@pytest.mark.parametrize("n", [5, 64, 200, 512])
def test_zgetrf_matches_lapack(n):
    print(f"zgetrf n={n}", flush=True)
    rng = np.random.default_rng(100 + n)
    A = _rand(rng, n, n)
    lu = api.be_lu(None, _dev(A))
    lu_ref, piv_ref, info = sla.lapack.zgetrf(A)
    assert np.array_equal(lu.piv.cpu().numpy(), piv_ref + 1)
    assert relfro(_host(lu.A), lu_ref) < 1e-12


def _hadamard_phased(n, perm_seed=None):
    """Sylvester-Hadamard matrix with the columns multiplied by 1, i, -1, -i
    in turn (rows optionally permuted): every entry has |re| + |im| = 1, and
    LU with partial pivoting keeps every intermediate a small dyadic rational,
    so it is exact in any operation order and its ties stay exact ties."""
    H = np.array([[1.0]])
    while H.shape[0] < n:
        H = np.block([[H, H], [H, -H]])
    A = H * np.array([1, 1j, -1, -1j])[np.arange(n) % 4][None, :]
    if perm_seed is not None:
        A = A[np.random.default_rng(perm_seed).permutation(n)]
    return np.asfortranarray(A)


@pytest.mark.parametrize("n", [64, 256, 512, 1024])
@pytest.mark.parametrize("perm_seed", [None, 5])
def test_zgetrf_exact_ties_match_lapack(n, perm_seed):
    """Pivot ties broken as LAPACK's izamax does (the first row among equal
    |re| + |im|), for each rows-per-thread layout of the panel kernel: on a
    matrix whose factorization is exact, ipiv and the factors equal
    LAPACK's bit for bit."""
    print(f"zgetrf ties n={n} perm={perm_seed}", flush=True)
    A = _hadamard_phased(n, perm_seed)
    lu = api.be_lu(None, _dev(A))
    lu_ref, piv_ref, info = sla.lapack.zgetrf(A)
    assert info == 0
    assert np.array_equal(lu.piv.cpu().numpy(), piv_ref + 1)
    assert np.array_equal(_host(lu.A), lu_ref)


@pytest.mark.slow
@pytest.mark.parametrize("n,perm_seed", [(2560, 3), (4096, None), (4100, 8)])
def test_zgetrf_beyond_2048_exact_ties_match_lapack(n, perm_seed):
    """The two-level blocked LU (n > 2048: 64-column outer panels factored 8 /
    4 / 2 columns at a time by one 512-thread workgroup, the trailing update
    on the matrix cores) on the exactly-factorable Hadamard matrix, padded
    to a non-power of two by an identity block where n is one: ipiv and
    the factors equal LAPACK's bit for bit."""
    print(f"zgetrf large ties n={n} perm={perm_seed}", flush=True)
    m = 1 << (n - 1).bit_length() if n & (n - 1) else n
    m = m if m == n else m // 2
    A = np.eye(n, dtype=complex)
    A[:m, :m] = _hadamard_phased(m, perm_seed)
    A = np.asfortranarray(A)
    lu = api.be_lu(None, _dev(A))
    lu_ref, piv_ref, info = sla.lapack.zgetrf(A)
    assert info == 0
    assert np.array_equal(lu.piv.cpu().numpy(), piv_ref + 1)
    assert np.array_equal(_host(lu.A), lu_ref)


@pytest.mark.slow
@pytest.mark.parametrize("n", [2100, 4500])
def test_zinv_beyond_gauss_jordan(n):
    """be_inv above n = 2048 (LU, triangular inverses, their product, the
    column interchanges), on a random matrix and on a row-permuted
    near-identity that needs every pivot right."""
    print(f"zinv large n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = _rand(rng, n, n) + 2 * np.sqrt(n) * np.eye(n)
    inv = _host(api.be_inv(_dev(A)))
    assert relfro(inv, np.linalg.inv(A)) < 1e-15 * np.linalg.cond(A) * n
    perm = rng.permutation(n)
    B = np.asfortranarray((np.eye(n) + 1e-8 * _rand(rng, n, n))[perm])
    inv = _host(api.be_inv(_dev(B)))
    assert relfro(inv, np.linalg.inv(B)) < 1e-13


@pytest.mark.slow
def test_bndiag_of_inv_direct_beyond_2048():
    """bndiag_of_inv_direct (src/matinvertndiag_direct.jl:17-86) on the
    reference's uneven generator at N = 40 blocks of 64 +- 4 (n ~ 2560, the
    large-n inverse): its band equals the oracle's RGF band."""
    import scipy.sparse as sp
    from libnegf_amd.synthetic import bm_create_synthetic_random
    from oracle.rgf import bndiag_of_inv_rgf_local
    A = bm_create_synthetic_random(40, 64, seed=40)
    n = sum(A.block_sizes)
    print(f"bndiag_of_inv_direct n={n}", flush=True)
    out = api.bndiag_of_inv_direct(sp.csc_matrix(A.to_dense()), A.block_sizes).toarray()
    ref = bndiag_of_inv_rgf_local(A)
    off = np.concatenate([[0], np.cumsum(A.block_sizes)])
    got = {k: out[off[k[0]]:off[k[0] + 1], off[k[1]]:off[k[1] + 1]] for k in ref}
    assert max(relfro(got[k], ref[k]) for k in ref) < 1e-11
    mask = np.ones((n, n), dtype=bool)
    for (i, j) in ref:
        mask[off[i]:off[i + 1], off[j]:off[j + 1]] = False
    assert not out[mask].any()


@pytest.mark.parametrize("n", [64, 512])
def test_zinv_exact_ties(n):
    """Gauss-Jordan on the same tie-laden matrix (rows permuted): A A^H = n I,
    so inv(A) = A^H / n."""
    print(f"zinv ties n={n}", flush=True)
    A = _hadamard_phased(n, 5)
    inv = _host(api.be_inv(_dev(A)))
    assert relfro(inv, A.conj().T / n) < 1e-15


@pytest.fixture
def getrs_route():
    """Force negf_zgetrs's route through the tune table (kind 3); the loaded
    table is restored afterwards."""
    from libnegf_amd import tune

    def force(n, nrhs, route):
        tune.record("getrs", (n, nrhs), [route], persist=False)
    yield force
    call("negf_tune_clear")
    tune.load_table()


@pytest.mark.parametrize("route", [1, 2])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("n,nrhs", [(150, 33), (512, 512), (64, 1), (200, 700)])
def test_zgetrs(trans, n, nrhs, route, getrs_route):
    """be_lu + be_mldivide as the unchanged reference sweeps call them, by the
    blocked substitution (route 1) and by the factors' explicit inverses
    (route 2, the default for 64 <= n <= 2048); 512 x 512 with 512
    right-hand sides is the production block's solve."""
    print(f"zgetrs {trans} n={n} nrhs={nrhs} route={route}", flush=True)
    getrs_route(n, nrhs, route)
    rng = np.random.default_rng(7)
    A = _rand(rng, n, n)
    B = _rand(rng, n, nrhs)
    lu = api.be_lu(None, _dev(A))
    X = api.be_mldivide(trans, api.be_zero_array((n, nrhs)), _dev(B), lu)
    ref = np.linalg.solve(_op(A, trans), B)
    assert relfro(_host(X), ref) < 1e-11


@pytest.mark.parametrize("route", [1, 2])
@pytest.mark.parametrize("trans", ["N", "C"])
def test_zgetrs_strided_batched_padded(trans, route, getrs_route):
    """Three instances in one call, leading dimensions above n and instance
    strides above the blocks, on matrices whose pivoting swaps rows (LAPACK's
    ipiv from getrf): each instance against numpy's solve."""
    n, nrhs, batch, lda, ldb = 96, 40, 3, 101, 99
    print(f"zgetrs batched {trans} route={route}", flush=True)
    getrs_route(n, nrhs, route)
    rng = np.random.default_rng(31)
    As = [_rand(rng, n, n) for _ in range(batch)]
    Bs = [_rand(rng, n, nrhs) for _ in range(batch)]
    sA, sB = lda * n + 7, ldb * nrhs + 5
    dA = torch.zeros(sA * batch, dtype=torch.complex128, device="cuda")
    dB = torch.zeros(sB * batch, dtype=torch.complex128, device="cuda")
    for e in range(batch):
        Ap = np.zeros((lda, n), complex, order="F"); Ap[:n] = As[e]
        Bp = np.zeros((ldb, nrhs), complex, order="F"); Bp[:n] = Bs[e]
        dA[e * sA:e * sA + lda * n] = torch.from_numpy(Ap.ravel(order="F"))
        dB[e * sB:e * sB + ldb * nrhs] = torch.from_numpy(Bp.ravel(order="F"))
    piv = torch.zeros(batch * n, dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    vp = lambda t: C.c_void_p(t.data_ptr())
    call("negf_zgetrf_strided_batched", n, vp(dA), lda, sA, vp(piv), n, vp(info), batch, None)
    call("negf_zgetrs_strided_batched", trans.encode(), n, nrhs, vp(dA), lda, sA, vp(piv), n,
         vp(dB), ldb, sB, batch, None)
    torch.cuda.synchronize()
    assert not info.any()
    hB = dB.cpu().numpy()
    for e in range(batch):
        X = hB[e * sB:e * sB + ldb * nrhs].reshape((ldb, nrhs), order="F")
        assert relfro(X[:n], np.linalg.solve(_op(As[e], trans), Bs[e])) < 1e-12
        # the padding rows of B are not written
        assert not X[n:].any()


@pytest.mark.parametrize("route", [1, 2])
@pytest.mark.parametrize("mn", [(70, 45), (200, 130)])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("tA", ["N", "T", "C"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_ztrsm(side, uplo, tA, diag, mn, route):
    """Every side / uplo / op / diag by the blocked substitution (route 1)
    and by one GEMM with the triangle's inverse (route 2), against BLAS."""
    from libnegf_amd import tune
    print(f"ztrsm {side}{uplo}{tA}{diag} {mn} route={route}", flush=True)
    rng = np.random.default_rng(11)
    m, n = mn
    na = m if side == "L" else n
    tune.record("trsm", (na, n if side == "L" else m), [route], persist=False)
    A = _rand(rng, na, na) + 4 * np.eye(na)
    B = _rand(rng, m, n)
    alpha = 1.5 - 0.5j
    X = api.be_trsm(side, uplo, tA, diag, alpha, _dev(A), _dev(B))
    ref = sla.blas.ztrsm(alpha, A, B, side=0 if side == "L" else 1, lower=uplo == "L",
                         trans_a={"N": 0, "T": 1, "C": 2}[tA], diag=diag == "U")
    call("negf_tune_clear")
    tune.load_table()
    assert relfro(_host(X), ref) < 1e-12


@pytest.mark.parametrize("route", [1, 2])
@pytest.mark.parametrize("pivots", [True, False])
def test_zgetri_routes(route, pivots, getrs_route):
    """getri from getrf's factors (LAPACK pivots, rows swapped) and with
    identity pivots (be_getri!), by both routes, batched with padded leading
    dimensions, against numpy's inverse."""
    n, batch, lda = 150, 2, 153
    print(f"zgetri route={route} pivots={pivots}", flush=True)
    getrs_route(n, n, route)
    rng = np.random.default_rng(41)
    As = [_rand(rng, n, n) + (0 if pivots else 3 * n * np.eye(n)) for _ in range(batch)]
    sA = lda * n
    dA = torch.zeros(sA * batch, dtype=torch.complex128, device="cuda")
    for e in range(batch):
        Ap = np.zeros((lda, n), complex, order="F"); Ap[:n] = As[e]
        dA[e * sA:(e + 1) * sA] = torch.from_numpy(Ap.ravel(order="F"))
    piv = torch.zeros(batch * n, dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    work = torch.zeros(batch * n * n, dtype=torch.complex128, device="cuda")
    vp = lambda t: C.c_void_p(t.data_ptr())
    call("negf_zgetrf_strided_batched", n, vp(dA), lda, sA, vp(piv), n, vp(info), batch, None)
    if not pivots:
        assert np.array_equal(piv.cpu().numpy(), np.tile(np.arange(1, n + 1), batch))
    call("negf_zgetri_strided_batched", n, vp(dA), lda, sA, vp(piv) if pivots else None, n,
         vp(work), batch, None)
    torch.cuda.synchronize()
    h = dA.cpu().numpy()
    for e in range(batch):
        X = h[e * sA:(e + 1) * sA].reshape((lda, n), order="F")
        assert relfro(X[:n], np.linalg.inv(As[e])) < 1e-12
        assert not X[n:].any()


def test_zgetri_identity_pivots_and_adjoint_fill():
    print("zgetri / adjoint / fill", flush=True)
    rng = np.random.default_rng(5)
    n = 90
    A = _rand(rng, n, n) + 3 * n * np.eye(n)        # diagonally dominant: no row swaps
    dA = _dev(A)
    api.be_getrf(dA)
    api.be_getri(dA)
    assert relfro(_host(dA), np.linalg.inv(A)) < 1e-13
    B = _rand(rng, 33, 71)
    out = api.be_zero_array((71, 33))
    api.be_ctranspose(out, _dev(B))
    assert np.array_equal(_host(out), B.conj().T)
    assert np.array_equal(_host(api.be_ctranspose(_dev(B))), B.conj().T)   # non-mutating form
    I = api.be_identity(17)
    assert np.array_equal(_host(I), np.eye(17))
    api.be_fill(I, 2 - 1j)
    assert np.all(_host(I) == 2 - 1j)
