"""Reduced-precision and emulated complex products on the GPU
(negf_zgemm_prec_strided_batched) against numpy.

* precision 2 (Ozaki scheme II on the int8 matrix cores): on inputs whose
  rows / columns are small integers the scaled integers hold them exactly, so
  the CRT-recovered product is the exact integer product -- equal bit for bit
  to numpy's (also exact there).  On random data the error is bounded by
  2^-beta of the row / column maxima per term (beta = 45 at k = 512 with 13
  moduli): relative Frobenius error below 1e-12.  Fewer moduli keep fewer bits:
  the error grows by about 2^(2 * 8 / 2) per modulus dropped, which pins that
  the moduli count is what sets the accuracy.
* precision 1 (FP32 MFMA): error of FP32 class (1e-7 .. 1e-5 relative), and
  measurably above FP64 rounding, so the FP32 path really ran.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import relfro

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from libnegf_amd._lib import call, cplx  # noqa: E402


def _op(X, o):
    return X if o == "N" else (X.T if o == "T" else X.conj().T)


def _gemm(prec, tA, tB, A, B, Cin=None, alpha=1.0, beta=0.0, nmod=0, shareA=False):
    """A, B, C: lists of host matrices (one per instance; shareA: one A for
    all).  Returns the list of results."""
    batch = len(B)
    m = _op(A[0], tA).shape[0]
    k = _op(A[0], tA).shape[1]
    n = _op(B[0], tB).shape[1]
    to_dev = lambda mats: torch.from_numpy(
        np.stack([np.asfortranarray(x).ravel(order="F") for x in mats])).cuda()
    dA, dB = to_dev(A), to_dev(B)
    Cmats = Cin if Cin is not None else [np.zeros((m, n), complex) for _ in range(batch)]
    dC = to_dev(Cmats)
    s = torch.cuda.current_stream()
    call("negf_zgemm_prec_strided_batched", tA.encode(), tB.encode(), m, n, k, cplx(alpha),
         C.c_void_p(dA.data_ptr()), A[0].shape[0], 0 if shareA else dA.shape[1],
         C.c_void_p(dB.data_ptr()), B[0].shape[0], dB.shape[1], cplx(beta),
         C.c_void_p(dC.data_ptr()), m, dC.shape[1], batch, prec, nmod, C.c_void_p(s.cuda_stream))
    torch.cuda.synchronize()
    h = dC.cpu().numpy()
    return [h[i].reshape(n, m).T for i in range(batch)]


def _rand(rng, m, n, ints=False):
    if ints:
        return (rng.integers(-100, 101, (m, n)) + 1j * rng.integers(-100, 101, (m, n))).astype(complex)
    return rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))


def _shape_of(o, rows, cols):
    return (rows, cols) if o == "N" else (cols, rows)


@pytest.mark.parametrize("tA,tB", [("N", "N"), ("N", "C"), ("C", "N"), ("T", "C"), ("C", "T")])
@pytest.mark.parametrize("m,n,k,batch", [(64, 64, 64, 1), (100, 70, 33, 2), (512, 512, 512, 3)])
def test_emulated_exact_on_small_integers(tA, tB, m, n, k, batch):
    print(f"emu exact ints {tA}{tB} {m}x{n}x{k} x{batch}", flush=True)
    rng = np.random.default_rng(m + n + k)
    A = [_rand(rng, *_shape_of(tA, m, k), ints=True) for _ in range(batch)]
    B = [_rand(rng, *_shape_of(tB, k, n), ints=True) for _ in range(batch)]
    got = _gemm(2, tA, tB, A, B)
    for a, b, d in zip(A, B, got):
        assert np.array_equal(d, _op(a, tA) @ _op(b, tB))


@pytest.mark.parametrize("tA,tB", [("N", "N"), ("N", "C"), ("C", "C")])
@pytest.mark.parametrize("m,n,k,batch,shareA", [(96, 96, 96, 2, False), (512, 512, 512, 4, True),
                                                (130, 64, 200, 1, False)])
def test_emulated_random_within_bound(tA, tB, m, n, k, batch, shareA):
    print(f"emu random {tA}{tB} {m}x{n}x{k} x{batch} shareA={shareA}", flush=True)
    rng = np.random.default_rng(7)
    A = [_rand(rng, *_shape_of(tA, m, k))] * batch if shareA else \
        [_rand(rng, *_shape_of(tA, m, k)) for _ in range(batch)]
    B = [_rand(rng, *_shape_of(tB, k, n)) for _ in range(batch)]
    Cin = [_rand(rng, m, n) for _ in range(batch)]
    alpha, beta = -1.0 + 0.5j, 0.25 - 1.0j
    got = _gemm(2, tA, tB, A, B, Cin, alpha, beta, shareA=shareA)
    for a, b, c0, d in zip(A, B, Cin, got):
        ref = alpha * (_op(a, tA) @ _op(b, tB)) + beta * c0
        err = relfro(d, ref)
        print(f"  relerr {err:.2e}")
        assert err < 1e-12
        # the elementwise bound: 2^-beta of the row / column maxima per term
        opa, opb = _op(a, tA), _op(b, tB)
        ra = np.maximum(np.abs(opa.real), np.abs(opa.imag)).max(axis=1)
        cb = np.maximum(np.abs(opb.real), np.abs(opb.imag)).max(axis=0)
        bound = abs(alpha) * 4 * k * 2.0 ** -45 * np.outer(ra, cb) * 2
        assert np.all(np.abs(d - ref) <= bound + 1e-13 * np.abs(ref))


def test_emulated_accuracy_follows_moduli():
    """13 moduli keep beta = 45 bits, 9 keep 30 (at k = 256): the error grows by
    about 2^15 -- the moduli count sets the accuracy, as designed."""
    rng = np.random.default_rng(3)
    A = [_rand(rng, 128, 256)]
    B = [_rand(rng, 256, 128)]
    ref = A[0] @ B[0]
    e13 = relfro(_gemm(2, "N", "N", A, B, nmod=13)[0], ref)
    e9 = relfro(_gemm(2, "N", "N", A, B, nmod=9)[0], ref)
    print(f"13 moduli {e13:.2e}, 9 moduli {e9:.2e}")
    assert e13 < 1e-13
    assert 1e3 * e13 < e9 < 1e-7


def test_emulated_zero_rows_and_tiny_values():
    rng = np.random.default_rng(4)
    A = _rand(rng, 64, 64)
    A[5, :] = 0.0
    A[9, :] *= 1e-200
    B = _rand(rng, 64, 64)
    B[:, 3] = 0.0
    d = _gemm(2, "N", "N", [A], [B])[0]
    ref = A @ B
    assert np.all(d[5, :] == 0) and np.all(d[:, 3] == 0)
    assert relfro(d[9, :], ref[9, :]) < 1e-12
    assert relfro(d, ref) < 1e-12


@pytest.mark.parametrize("tA,tB", [("N", "N"), ("C", "N"), ("N", "C")])
def test_fp32_form_is_fp32_accurate(tA, tB):
    rng = np.random.default_rng(11)
    m = n = k = 512
    A = [_rand(rng, m, k) for _ in range(8)]
    B = [_rand(rng, k, n) for _ in range(8)]
    got = _gemm(1, tA, tB, A, B)
    for a, b, d in zip(A, B, got):
        err = relfro(d, _op(a, tA) @ _op(b, tB))
        print(f"fp32 {tA}{tB}: relerr {err:.2e}")
        assert 1e-9 < err < 1e-5


_OPS = [(a, b) for a in "NTC" for b in "NTC"]


@pytest.mark.parametrize("tA,tB", _OPS)
@pytest.mark.parametrize("m,n,k,batch", [(512, 512, 512, 8), (96, 64, 40, 3), (70, 33, 24, 2)])
def test_fp64_3m_form_is_fp64_accurate(tA, tB, m, n, k, batch):
    """precision 3 (three real products per complex product): FP64-class
    normwise error on random data for every op pair, at the 64 x 64 tiles
    (512^3 x 8 = 512 tiles), the 32 x 32 tiles (96 x 64) and a ragged shape
    that takes the four-product kernel; alpha, beta and C applied."""
    rng = np.random.default_rng(5 + m + 3 * k)
    A = [_rand(rng, *_shape_of(tA, m, k)) for _ in range(batch)]
    B = [_rand(rng, *_shape_of(tB, k, n)) for _ in range(batch)]
    Cm = [_rand(rng, m, n) for _ in range(batch)]
    alpha, beta = 0.75 - 0.5j, -1.25 + 0.25j
    got = _gemm(3, tA, tB, A, B, Cin=Cm, alpha=alpha, beta=beta)
    for a, b, c, d in zip(A, B, Cm, got):
        ref = alpha * (_op(a, tA) @ _op(b, tB)) + beta * c
        err = relfro(d, ref)
        assert err < 1e-14, (tA, tB, m, n, k, err)


def test_fp64_3m_form_runs_three_products():
    """The three-product form's imaginary part is R - P - Q with R = (ar +
    ai)(br + bi): where |Re| >> |Im| entries cancel, so Im loses digits that
    the four-product form keeps (Higham 1992) -- measurably, which shows the
    three products ran -- while the normwise error stays FP64-class."""
    rng = np.random.default_rng(17)
    m = n = k = 512
    A = [1e6 * rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k)) for _ in range(8)]
    B = [1e6 * rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n)) for _ in range(8)]
    got3 = _gemm(3, "N", "N", A, B)
    got4 = _gemm(0, "N", "N", A, B)
    for a, b, d3, d4 in zip(A, B, got3, got4):
        ref = a @ b
        im3, im4 = relfro(d3.imag, ref.imag), relfro(d4.imag, ref.imag)
        print(f"3M: Im relerr {im3:.2e} (4M {im4:.2e}), normwise {relfro(d3, ref):.2e}")
        assert relfro(d3, ref) < 1e-14
        assert im4 < 1e-14 and im3 > 1e3 * im4


# ---------------------------------------------------------------- energy job
from libnegf_amd import api  # noqa: E402
from libnegf_amd.synthetic import (device_hamiltonian, assemble_A, fermi,  # noqa: E402
                                   band_edge_energies)
from oracle.rgf import bndiag_of_inv_rgf_local as rgf_ref, rgf_upward, first_row_chain  # noqa
from oracle.keldysh import keldyshndiag as rkd_ref  # noqa: E402

_KT, _MUL, _MUR = 0.025, 0.1, -0.1


def _oracle_energy(dev, E, eta):
    """G^R diag, G^< diag and T of one energy from the oracle (the reference's
    call sequence in FP64)."""
    N, b = dev.nblocks, dev.b
    A = assemble_A(dev, E, eta)
    buf = rgf_upward(A)
    G = rgf_ref(A, upward=buf)
    fL, fR = fermi(E, _MUL, _KT), fermi(E, _MUR, _KT)
    Sd = [np.zeros((b, b), complex, order="F") for _ in range(N)]
    Sd[0] = np.asfortranarray(fL * dev.gamma_L)
    Sd[-1] = np.asfortranarray(Sd[-1] + fR * dev.gamma_R)
    Su = [np.zeros((b, b), complex, order="F") for _ in range(N - 1)]
    gd, _ = rkd_ref(A, Sd, Su, upward=buf)
    g1n = first_row_chain(A, G, buf)[-1]
    T = float(np.real(np.trace(dev.gamma_L @ g1n @ dev.gamma_R @ g1n.conj().T)))
    return [G[(i, i)] for i in range(N)], [1j * x for x in gd], T


def _energy_err(GR, GL, T, e, ref, b):
    """Largest relative Frobenius error over energy e's G^R and G^< blocks."""
    gr, gl, tr = ref
    err = 0.0
    for i in range(len(gr)):
        err = max(err, relfro(GR[e, i * b * b:(i + 1) * b * b].reshape(b, b).T, gr[i]))
        err = max(err, relfro(GL[e, i * b * b:(i + 1) * b * b].reshape(b, b).T, gl[i]))
    return err, abs(T[e] - tr) / max(abs(tr), 1e-300)


def _run_job(dev, Es, eta, precision, tol):
    N, b = dev.nblocks, dev.b
    H = api.DeviceHamiltonianGPU(dev)
    nE = len(Es)
    GR = torch.zeros((nE, N * b * b), dtype=torch.complex128, device="cuda")
    GL = torch.zeros_like(GR)
    T = torch.zeros(nE, dtype=torch.float64, device="cuda")
    E = torch.tensor(np.asarray(Es, dtype=np.float64), device="cuda")
    if precision == "fp64":
        api.rgf_keldysh_energies(H, E, eta, _MUL, _MUR, _KT, GR, GL, T)
        est = fb = None
    else:
        _, est, fb = api.rgf_keldysh_energies_mixed(H, E, eta, _MUL, _MUR, _KT, precision, tol,
                                                     GR_diag=GR, Gless_diag=GL, transmission=T)
    torch.cuda.synchronize()
    return GR.cpu().numpy(), GL.cpu().numpy(), T.cpu().numpy(), est, fb


@pytest.mark.slow
def test_mixed_energy_job_vs_oracle_production_block():
    """N = 16, b = 512, eta = 1e-4 at two band energies and two energies
    within 1e-3 of the lead band edges.  The emulated form (no fallback: tol
    = inf) meets 1e-8 against the oracle at every energy, and its estimate
    bounds the measured error; so does the three-product FP64 form (with tol
    = 1e-8: no energy re-run); the FP32 form misses 1e-8 by orders, its
    estimate says so, and with tol = 1e-8 every energy falls back to FP64 and
    then meets the FP64 bar."""
    N, b, eta = 16, 512, 1e-4
    dev = device_hamiltonian(N, b, seed=1234)
    lo, hi = band_edge_energies(dev)
    Es = [-0.61, 0.26984126984126977, lo + 3e-4, hi - 5e-4]
    print(f"mixed energy job N={N} b={b} energies={Es}", flush=True)
    refs = [_oracle_energy(dev, E, eta) for E in Es]
    GR, GL, T, est, fb = _run_job(dev, Es, eta, "emulated", np.inf)
    assert not fb.any()
    for e in range(len(Es)):
        err, terr = _energy_err(GR, GL, T, e, refs[e], b)
        print(f"  emulated E={Es[e]:+.5f}: err {err:.2e} T err {terr:.2e} est {est[e]:.2e}")
        assert err < 1e-8 and terr < 1e-8
        assert est[e] >= err          # the estimate bounds the error
        assert est[e] < 1e4 * err     # and is not vacuous
    GR, GL, T, est, fb = _run_job(dev, Es, eta, "fp64_3m", 1e-8)
    assert not fb.any()
    for e in range(len(Es)):
        err, terr = _energy_err(GR, GL, T, e, refs[e], b)
        print(f"  fp64_3m E={Es[e]:+.5f}: err {err:.2e} T err {terr:.2e} est {est[e]:.2e}")
        assert err < 1e-10 and terr < 1e-10
        assert est[e] >= err
    GR, GL, T, est, fb = _run_job(dev, Es, eta, "fp32", np.inf)
    for e in range(len(Es)):
        err, _ = _energy_err(GR, GL, T, e, refs[e], b)
        print(f"  fp32 E={Es[e]:+.5f}: err {err:.2e} est {est[e]:.2e}")
        assert err > 1e-6 and est[e] > 1e-8
    GR, GL, T, est, fb = _run_job(dev, Es, eta, "fp32", 1e-8)
    assert fb.all()
    for e in range(len(Es)):
        err, terr = _energy_err(GR, GL, T, e, refs[e], b)
        print(f"  fp32 + FP64 fallback E={Es[e]:+.5f}: err {err:.2e}")
        assert err < 2e-11 and terr < 1e-10


def test_mixed_energy_job_fallback_selects_and_scatters():
    """A tolerance between the energies' estimates re-runs exactly the
    energies above it, in FP64, and their outputs land at their own
    positions: equal to the plain FP64 job there; the rest stay emulated
    (close to FP64, not equal)."""
    N, b, eta = 6, 64, 1e-4
    dev = device_hamiltonian(N, b, seed=9)
    Es = list(np.linspace(-0.9, 0.9, 7))
    GRf, GLf, Tf, _, _ = _run_job(dev, Es, eta, "fp64", 0)
    _, _, _, est, _ = _run_job(dev, Es, eta, "emulated", np.inf)
    tol = float(np.median(est))
    GR, GL, T, est2, fb = _run_job(dev, Es, eta, "emulated", tol)
    assert np.array_equal(est, est2)
    assert np.array_equal(fb, (est > tol).astype(np.int32)) and 0 < fb.sum() < len(Es)
    for e in range(len(Es)):
        d = max(relfro(GR[e], GRf[e]), relfro(GL[e], GLf[e]))
        print(f"  E={Es[e]:+.3f} fallback={fb[e]} diff vs FP64 {d:.2e}")
        if fb[e]:
            assert d < 1e-13 and abs(T[e] - Tf[e]) <= 1e-13 * abs(Tf[e])
        else:
            assert 0 < d < 1e-9
