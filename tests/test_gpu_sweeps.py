"""RGF / Keldysh / transmission sweeps on the GPU vs the CPU oracle.

Tolerances are relative Frobenius errors per returned block.  1e-12 at the
small shapes (the problem's own FP64 noise floor there is ~1e-14..1e-13).  At
the production shape (N = 64, b = 512, eta = 1e-4) the floor itself is
~3e-12 (two FP64 formulations, tests/test_fp64_floors.py); the
production-shape checks with bars derived from it are in
test_gpu_production.py.
"""
import numpy as np
import pytest

from conftest import relfro, max_block_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from libnegf_amd import api  # noqa: E402
from libnegf_amd.btd import DeviceBlockTridiag  # noqa: E402
from libnegf_amd.synthetic import (bm_create_synthetic_random, device_hamiltonian,  # noqa: E402
                                   assemble_A, random_hermitian_btd, fermi)
from oracle.rgf import bndiag_of_inv_rgf_local as rgf_ref, rgf_upward, first_row_chain  # noqa
from oracle.keldysh import keldyshndiag as rkd_ref  # noqa: E402


def _gpu_rgf(mats):
    Min = DeviceBlockTridiag.from_host(mats)
    Mout = Min.similar()
    api.bndiag_of_inv_rgf_local(Mout, Min, api.allocate_aux_data_RGF(Min))
    return Mout


def _band(host):
    G = {}
    for i in range(host.nblocks):
        G[(i, i)] = host.diag[i]
    for i in range(host.nblocks - 1):
        G[(i, i + 1)] = host.upper[i]
        G[(i + 1, i)] = host.lower[i]
    return G


@pytest.mark.parametrize("N,b,seed", [(16, 64, 1234), (1, 40, 1), (2, 33, 2), (5, 130, 3)])
def test_rgf_negf_workload(N, b, seed):
    print(f"rgf workload N={N} b={b}", flush=True)
    dev = device_hamiltonian(N, b, seed=seed)
    A = assemble_A(dev, 0.3, 1e-3)
    G = _band(_gpu_rgf(A).to_host())
    err = max_block_err(G, rgf_ref(A))
    print(f"  max block relerr {err:.2e}")
    assert err < 1e-12


def test_rgf_reference_generator_uneven_blocks():
    print("rgf uneven (reference generator N=10 b~32)", flush=True)
    A = bm_create_synthetic_random(10, 32, seed=17)
    assert len(set(A.block_sizes)) > 1
    G = _band(_gpu_rgf(A).to_host())
    assert max_block_err(G, rgf_ref(A)) < 1e-12


def test_rgf_batched_energies_match_single():
    print("rgf batch of energies", flush=True)
    dev = device_hamiltonian(6, 48, seed=8)
    Es = [-0.5, 0.0, 0.25, 0.9]
    mats = [assemble_A(dev, E, 1e-3) for E in Es]
    out = _gpu_rgf(mats)
    for k, A in enumerate(mats):
        assert max_block_err(_band(out.to_host(k)), rgf_ref(A)) < 1e-12


def test_rgf_singular_block_raises():
    print("rgf singular", flush=True)
    A = bm_create_synthetic_random(3, 8, seed=2, uneven=False)
    A.diag[2][:] = 0.0
    A.upper[1][:] = 0.0
    A.lower[1][:] = 0.0
    Min = DeviceBlockTridiag.from_host(A)
    with pytest.raises(api.SingularError):
        api.bndiag_of_inv_rgf_local(Min.similar(), Min)


@pytest.mark.parametrize("N,b", [(10, 32), (3, 20)])
def test_keldysh_vs_oracle(N, b):
    print(f"keldysh N={N} b={b}", flush=True)
    A = bm_create_synthetic_random(N, b, seed=31)
    S = random_hermitian_btd(A.block_sizes, seed=32)
    M = DeviceBlockTridiag.from_host(A)
    Sd = DeviceBlockTridiag.from_host(S)
    aux = api.allocate_aux_data_Keldysh(M, Sd)
    Gr = M.similar()
    out = api.keldyshndiag(M, Sd, aux, Gr=Gr).to_host()
    rd, ru = rkd_ref(A, S.diag, S.upper)
    for i in range(N):
        assert relfro(out.diag[i], rd[i]) < 1e-12
    for i in range(N - 1):
        assert relfro(out.upper[i], ru[i]) < 1e-12
    assert max_block_err(_band(Gr.to_host()), rgf_ref(A)) < 1e-12


_ENERGIES = [-0.4, -0.05, 0.1, 0.33, 0.8]


@pytest.mark.parametrize("N,b,eta,Es,seed,tol,streams,gemm", [
    (1, 64, 1e-3, _ENERGIES, 4, 1e-12, 1, 0),
    (2, 64, 1e-3, _ENERGIES, 4, 1e-12, 1, 0),
    (6, 64, 1e-3, _ENERGIES, 4, 1e-12, 1, 0),
    (6, 64, 1e-3, _ENERGIES, 4, 1e-12, 2, 0),
    (6, 64, 1e-3, _ENERGIES, 4, 1e-12, 1, 9),
    (4, 160, 1e-3, _ENERGIES, 4, 1e-12, 1, 0),
    (3, 256, 1e-3, _ENERGIES, 4, 1e-12, 1, 0),
    (3, 256, 1e-3, _ENERGIES, 4, 1e-12, 1, 9),
    (3, 256, 1e-3, _ENERGIES, 4, 1e-12, 1, 8),
    (16, 512, 1e-4, [-0.61, 0.26984126984126977], 1234, 2e-11, 1, 0)])
def test_energies_workload_vs_oracle(N, b, eta, Es, seed, tol, streams, gemm):
    """N = 1: both contacts on the one layer, no coupling blocks; N = 2: no
    interior layer.  Two streams split the 5 energies into groups of 4 and 1 (waves of 2),
    each group forked from and joined back into the caller's stream.
    b = 160 (ragged last 64-tile) and 256 span several GEMM tiles.  Waves of
    2 energies at these sizes take the 32 x 32 GEMM tiles; gemm = 9 forces
    the 64 x 64 LDS tiles (the form of full-size waves), gemm = 8 the
    register-staged GEMM, each including the Hermitian-mirrored output.  The
    N = 16, b = 512, eta = 1e-4 case is the production block size and
    broadening at a depth where the Gamma_R propagation B_i = t_i B_{i+1}
    t_i^H is strongly contracting inside an expanding map (|t_i| ~ 40, |B_i|
    ~ 100): a forward perturbation of the computed B_i that is not of the
    backward-stable form (e.g. forcing exact Hermitian symmetry elementwise)
    grows ~|t|^2 per layer and fails this case by many orders."""
    kT, muL, muR = 0.025, 0.1, -0.1
    print(f"energies workload N={N} b={b} eta={eta} streams={streams} gemm={gemm}", flush=True)
    dev = device_hamiltonian(N, b, seed=seed)
    Es = np.array(Es)
    H = api.DeviceHamiltonianGPU(dev)
    nE = len(Es)
    GR = torch.zeros((nE, N * b * b), dtype=torch.complex128, device="cuda")
    GL = torch.zeros_like(GR)
    T = torch.zeros(nE, dtype=torch.float64, device="cuda")
    api.call("negf_gemm_variant", gemm)
    try:
        api.rgf_keldysh_energies(H, torch.tensor(Es, device="cuda"), eta, muL, muR, kT, GR, GL, T,
                                 wave=2, streams=streams)
        torch.cuda.synchronize()
    finally:
        api.call("negf_gemm_variant", 0)
    _check_energies_vs_oracle(dev, Es, eta, muL, muR, kT, GR, GL, T, tol)


def _check_energies_vs_oracle(dev, Es, eta, muL, muR, kT, GR, GL, T, tol):
    """Every energy's G^R and G^< diagonal blocks and T against the oracle's
    RGF, keldyshndiag (Sigma^n = f_L Gamma_L on layer 0, f_R Gamma_R on
    layer N-1) and Tr(Gamma_L G_1N Gamma_R G_1N^H)."""
    N, b = dev.nblocks, dev.b
    GR, GL, T = GR.cpu().numpy(), GL.cpu().numpy(), T.cpu().numpy()
    for e, E in enumerate(Es):
        A = assemble_A(dev, E, eta)
        buf = rgf_upward(A)
        G = rgf_ref(A, upward=buf)
        fL, fR = fermi(E, muL, kT), fermi(E, muR, kT)
        Sd = [np.zeros((b, b), complex, order="F") for _ in range(N)]
        Sd[0] = np.asfortranarray(fL * dev.gamma_L)
        Sd[-1] = np.asfortranarray(Sd[-1] + fR * dev.gamma_R)
        Su = [np.zeros((b, b), complex, order="F") for _ in range(N - 1)]
        gd, _ = rkd_ref(A, Sd, Su, upward=buf)
        g1n = first_row_chain(A, G, buf)[-1]
        Tref = np.real(np.trace(dev.gamma_L @ g1n @ dev.gamma_R @ g1n.conj().T))
        for i in range(N):
            blk = GR[e, i * b * b:(i + 1) * b * b].reshape(b, b).T
            assert relfro(blk, G[(i, i)]) < tol
            blk = GL[e, i * b * b:(i + 1) * b * b].reshape(b, b).T
            assert relfro(blk, 1j * gd[i]) < tol
        assert abs(T[e] - Tref) < 10 * tol * max(1.0, abs(Tref))


def _energies(dev, Es, eta, muL, muR, kT, check=True):
    N, b, nE = dev.nblocks, dev.b, len(Es)
    H = api.DeviceHamiltonianGPU(dev)
    GR = torch.zeros((nE, N * b * b), dtype=torch.complex128, device="cuda")
    GL = torch.zeros_like(GR)
    T = torch.zeros(nE, dtype=torch.float64, device="cuda")
    info = api.rgf_keldysh_energies(H, torch.tensor(Es, dtype=torch.float64, device="cuda"), eta,
                                    muL, muR, kT, GR, GL, T, wave=2, check=check)
    torch.cuda.synchronize()
    return GR, GL, T, info


def test_energies_rank_deficient_and_zero_broadening():
    """The contacts' Gamma enter G^< through their pivoted Cholesky factors
    Gamma = F F^H: a rank-3 Gamma_L (factor zero past column 3) and a zero
    Gamma_R (factor zero, no right injection, T = 0) still match the
    oracle, which uses Gamma itself.  eta = 1e-2: with no right broadening
    the last layer keeps only eta, and at eta = 1e-3 the oracle's RGF and a
    dense numpy inverse of A already differ by 1.3e-12 per block (2.2e-13 at
    1e-2), above the 1e-12 bar."""
    kT, muL, muR, eta = 0.025, 0.1, -0.1, 1e-2
    dev = device_hamiltonian(4, 64, seed=8)
    rng = np.random.default_rng(8)
    X = rng.standard_normal((64, 3)) + 1j * rng.standard_normal((64, 3))
    g = X @ X.conj().T / 3
    dev.gamma_L = np.asfortranarray((g + g.conj().T) / 2)
    dev.gamma_R = np.zeros((64, 64), complex, order="F")
    print("energies: rank-3 Gamma_L, zero Gamma_R, N=4 b=64", flush=True)
    GR, GL, T, info = _energies(dev, _ENERGIES, eta, muL, muR, kT)
    assert not bool((info != 0).any())
    _check_energies_vs_oracle(dev, _ENERGIES, eta, muL, muR, kT, GR, GL, T, 1e-12)
    assert float(T.abs().max()) == 0.0


@pytest.mark.parametrize("kind", ["negative", "indefinite", "non_hermitian"])
def test_energies_broadening_not_psd_is_reported(kind):
    """A Gamma that is not Hermitian positive semidefinite cannot be written
    F F^H: the job sets info = -1 for every energy and the Python host
    raises ValueError (the outputs are not valid)."""
    kT, muL, muR, eta = 0.025, 0.1, -0.1, 1e-3
    dev = device_hamiltonian(3, 64, seed=9)
    g = dev.gamma_R.copy()
    if kind == "negative":
        g = -g
    elif kind == "indefinite":
        g[:, :] = 0.0
        g[0, 1] = g[1, 0] = 1.0   # eigenvalues +1 and -1, zero diagonal
    else:
        g[0, 1] += 1e-3 * abs(g).max()
    dev.gamma_R = np.asfortranarray(g)
    print(f"energies: Gamma_R {kind}", flush=True)
    *_, info = _energies(dev, _ENERGIES, eta, muL, muR, kT, check=False)
    assert info.cpu().tolist() == [-1] * len(_ENERGIES)
    with pytest.raises(ValueError, match="positive semidefinite"):
        _energies(dev, _ENERGIES, eta, muL, muR, kT)


@pytest.mark.parametrize("streams", [1, 2])
def test_energies_graph_replays_match_eager(streams):
    """The graph-captured job replays bit for bit what the eager call
    computes, for new energies written in place between replays, with eager
    work and synchronisations between the replays."""
    N, b, eta, kT, muL, muR = 5, 48, 1e-3, 0.025, 0.1, -0.1
    print(f"energies graph N={N} b={b} streams={streams}", flush=True)
    dev = device_hamiltonian(N, b, seed=11)
    H = api.DeviceHamiltonianGPU(dev)
    nE = 6
    E_graph = torch.zeros(nE, dtype=torch.float64, device="cuda")
    outs = [torch.zeros((nE, N * b * b), dtype=torch.complex128, device="cuda") for _ in range(2)]
    T = torch.zeros(nE, dtype=torch.float64, device="cuda")
    graph = api.EnergiesGraph(H, E_graph, eta, muL, muR, kT, outs[0], outs[1], T, wave=2,
                              streams=streams)
    try:
        rng = np.random.default_rng(3)
        for rep in range(3):
            Es = torch.tensor(np.sort(rng.uniform(-1, 1, nE)), device="cuda")
            GR = torch.zeros_like(outs[0])
            GL = torch.zeros_like(outs[0])
            Te = torch.zeros_like(T)
            api.rgf_keldysh_energies(H, Es, eta, muL, muR, kT, GR, GL, Te, wave=2,
                                     streams=streams)
            torch.cuda.synchronize()
            E_graph.copy_(Es)
            for o in outs:
                o.zero_()
            T.zero_()
            graph.launch()
            torch.cuda.synchronize()
            assert torch.equal(outs[0], GR), f"replay {rep}: G^R differs"
            assert torch.equal(outs[1], GL), f"replay {rep}: G^< differs"
            assert torch.equal(T, Te), f"replay {rep}: T differs"
        # and the last replay against the oracle for one energy
        E = float(Es[2])
        A = assemble_A(dev, E, eta)
        G = rgf_ref(A)
        for i in range(N):
            blk = outs[0][2, i * b * b:(i + 1) * b * b].cpu().numpy().reshape(b, b).T
            assert relfro(blk, G[(i, i)]) < 1e-12
    finally:
        graph.close()


def test_energies_graph_three_product_form_replays_match_eager():
    """The energy job in the three-product FP64 form (the benchmark's),
    captured into a graph with its per-energy estimates: replays equal the
    eager call (negf_rgf_keldysh_energies_mixed, precision 3, no energy
    re-run) bit for bit, outputs and estimates, for new energies written in
    place between replays; b = 320 takes the LDS-DMA three-product kernel and
    the split inverse."""
    N, b, eta, kT, muL, muR = 3, 320, 1e-3, 0.025, 0.1, -0.1
    nE = 32
    print(f"energies graph three-product N={N} b={b} wave={nE}", flush=True)
    dev = device_hamiltonian(N, b, seed=23)
    H = api.DeviceHamiltonianGPU(dev)
    E_graph = torch.zeros(nE, dtype=torch.float64, device="cuda")
    outs = [torch.zeros((nE, N * b * b), dtype=torch.complex128, device="cuda") for _ in range(2)]
    T = torch.zeros(nE, dtype=torch.float64, device="cuda")
    graph = api.EnergiesGraph(H, E_graph, eta, muL, muR, kT, outs[0], outs[1], T, wave=nE,
                              precision="fp64_3m")
    try:
        rng = np.random.default_rng(7)
        for rep in range(2):
            Es = torch.tensor(np.sort(rng.uniform(-1, 1, nE)), device="cuda")
            GR = torch.zeros_like(outs[0])
            GL = torch.zeros_like(outs[0])
            Te = torch.zeros_like(T)
            _, est, fb = api.rgf_keldysh_energies_mixed(H, Es, eta, muL, muR, kT, "fp64_3m", 1e-8,
                                                        GR_diag=GR, Gless_diag=GL,
                                                        transmission=Te, wave=nE)
            torch.cuda.synchronize()
            assert not fb.any(), "no energy should need its FP64 re-run"
            E_graph.copy_(Es)
            for o in outs:
                o.zero_()
            T.zero_()
            graph.launch()
            torch.cuda.synchronize()
            assert torch.equal(outs[0], GR), f"replay {rep}: G^R differs"
            assert torch.equal(outs[1], GL), f"replay {rep}: G^< differs"
            assert torch.equal(T, Te), f"replay {rep}: T differs"
            assert np.array_equal(graph.estimates(), est), f"replay {rep}: estimates differ"
        k = 11
        A = assemble_A(dev, float(Es[k]), eta)
        G = rgf_ref(A)
        for i in range(N):
            blk = outs[0][k, i * b * b:(i + 1) * b * b].cpu().numpy().reshape(b, b).T
            assert relfro(blk, G[(i, i)]) < 1e-12
    finally:
        graph.close()


def test_energies_graph_split_inverse_replays_match_eager():
    """Waves of 32 energies at b = 320 take the batched inverse's split steps
    (the factoring on the job's stream, the strip updates on the library's
    second stream, joined by events every step); captured into a graph, the
    fork/join becomes graph edges and replays must equal the eager call bit for
    bit, and the oracle for a sampled energy."""
    N, b, eta, kT, muL, muR = 3, 320, 1e-3, 0.025, 0.1, -0.1
    nE = 32
    print(f"energies graph split inverse N={N} b={b} wave={nE}", flush=True)
    dev = device_hamiltonian(N, b, seed=19)
    H = api.DeviceHamiltonianGPU(dev)
    E_graph = torch.zeros(nE, dtype=torch.float64, device="cuda")
    outs = [torch.zeros((nE, N * b * b), dtype=torch.complex128, device="cuda") for _ in range(2)]
    T = torch.zeros(nE, dtype=torch.float64, device="cuda")
    graph = api.EnergiesGraph(H, E_graph, eta, muL, muR, kT, outs[0], outs[1], T, wave=nE)
    try:
        rng = np.random.default_rng(5)
        for rep in range(2):
            Es = torch.tensor(np.sort(rng.uniform(-1, 1, nE)), device="cuda")
            GR = torch.zeros_like(outs[0])
            GL = torch.zeros_like(outs[0])
            Te = torch.zeros_like(T)
            api.rgf_keldysh_energies(H, Es, eta, muL, muR, kT, GR, GL, Te, wave=nE)
            torch.cuda.synchronize()
            E_graph.copy_(Es)
            for o in outs:
                o.zero_()
            T.zero_()
            graph.launch()
            torch.cuda.synchronize()
            assert torch.equal(outs[0], GR), f"replay {rep}: G^R differs"
            assert torch.equal(outs[1], GL), f"replay {rep}: G^< differs"
            assert torch.equal(T, Te), f"replay {rep}: T differs"
        k = 7
        A = assemble_A(dev, float(Es[k]), eta)
        G = rgf_ref(A)
        for i in range(N):
            blk = outs[0][k, i * b * b:(i + 1) * b * b].cpu().numpy().reshape(b, b).T
            assert relfro(blk, G[(i, i)]) < 1e-12
    finally:
        graph.close()


def test_ward_identity_on_gpu_outputs():
    """i (G - G^H) = G (2 eta I + Gamma_L + Gamma_R) G^H on the band, both
    sides from the GPU (RKD fed with the broadening, RGF for G)."""
    N, b, eta = 12, 96, 1e-3
    print(f"ward identity N={N} b={b}", flush=True)
    dev = device_hamiltonian(N, b, seed=6)
    A = assemble_A(dev, 0.15, eta)
    Sd = [2 * eta * np.eye(b, dtype=complex) for _ in range(N)]
    Sd[0] = Sd[0] + dev.gamma_L
    Sd[-1] = Sd[-1] + dev.gamma_R
    from libnegf_amd.synthetic import BlockTridiag
    S = BlockTridiag([np.asfortranarray(x) for x in Sd],
                     [np.zeros((b, b), complex, order="F") for _ in range(N - 1)], None)
    M = DeviceBlockTridiag.from_host(A)
    Gr = M.similar()
    out = api.keldyshndiag(M, DeviceBlockTridiag.from_host(S),
                           api.allocate_aux_data_Keldysh(M, DeviceBlockTridiag.from_host(S)),
                           Gr=Gr).to_host()
    G = Gr.to_host()
    for i in range(N):
        assert relfro(out.diag[i], 1j * (G.diag[i] - G.diag[i].conj().T)) < 1e-11
    for i in range(N - 1):
        assert relfro(out.upper[i], 1j * (G.upper[i] - G.lower[i].conj().T)) < 1e-11


def _uneven_instances(sizes, shifts, seed):
    """One block-tridiagonal matrix with the given block sizes, shifted by
    z_k I per instance: A_k = z_k I - H, H a random Hermitian-coupled band."""
    rng = np.random.default_rng(seed)
    N = len(sizes)
    cn = lambda m, n: (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))) / np.sqrt(2 * n)
    H_d = []
    for s in sizes:
        X = cn(s, s)
        H_d.append((X + X.conj().T) / 2 + np.diag(rng.uniform(-2, 2, s)))
    H_u = [cn(sizes[i], sizes[i + 1]) for i in range(N - 1)]
    from libnegf_amd.synthetic import BlockTridiag
    out = []
    for z in shifts:
        out.append(BlockTridiag(
            [np.asfortranarray(z * np.eye(s) - h) for s, h in zip(sizes, H_d)],
            [np.asfortranarray(-u) for u in H_u],
            [np.asfortranarray(-u.conj().T) for u in H_u]))
    return out


@pytest.mark.slow
@pytest.mark.parametrize("sizes,batch,inv_variant", [
    ([600, 512, 600], 32, 0),    # the 512 layer takes split steps (64-column panels)
    ([600, 512, 600], 1, 4),     # forced split steps at one instance
    ([1100, 1024], 32, 0),       # the 1024 layer: 8-column strips, two per panel
])
def test_rgf_uneven_layers_smaller_block_wider_panel(sizes, batch, inv_variant):
    """Regression: the inverse workspace of a sweep is carved for the largest
    block, but a smaller layer can take a schedule with wider W panels.
    Every instance is checked against the oracle on a sample."""
    print(f"rgf uneven sizes={sizes} batch={batch} inv_variant={inv_variant}", flush=True)
    shifts = [complex(E, 1e-2) for E in np.linspace(-0.9, 0.9, batch)]
    mats = _uneven_instances(sizes, shifts, seed=5)
    api.call("negf_inv_variant", inv_variant)
    try:
        out = _gpu_rgf(mats)
        torch.cuda.synchronize()
    finally:
        api.call("negf_inv_variant", 0)
    for k in sorted({0, batch // 2, batch - 1}):
        err = max_block_err(_band(out.to_host(k)), rgf_ref(mats[k]))
        print(f"  instance {k}: max block relerr {err:.2e}")
        assert err < 1e-12


@pytest.mark.parametrize("sizes,batch", [([32] * 8, 2), ([20, 24, 17, 30, 22, 19], 1), ([40, 40], 1)])
def test_rgf_build_full_inverse(sizes, batch):
    """auxData.buildFullInv (src/matinvertndiag_rgf.jl:185-199): every block of
    the inverse, uniform blocks (one launch per block diagonal over all
    blocks and instances) and uneven blocks; against the oracle's
    full_inverse and the dense inverse; the band in Mout as without the flag."""
    from libnegf_amd.synthetic import BlockTridiag
    print(f"rgf full inverse sizes={sizes} batch={batch}", flush=True)
    mats = _uneven_instances(sizes, [complex(E, 5e-2) for E in np.linspace(-0.5, 0.5, batch)], 9)
    Min = DeviceBlockTridiag.from_host(mats)
    Mout = Min.similar()
    aux = api.allocate_aux_data_RGF(Min)
    aux.buildFullInv = True
    api.bndiag_of_inv_rgf_local(Mout, Min, aux)
    n = sum(sizes)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for e, A in enumerate(mats):
        F = aux.fullInv[e].cpu().numpy().T           # [column, row] -> dense
        assert relfro(F, np.linalg.inv(A.to_dense())) < 1e-12
        ref = rgf_ref(A, full_inverse=True)
        for (i, j), blk in ref.items():
            assert relfro(F[off[i]:off[i + 1], off[j]:off[j + 1]], blk) < 1e-12, (i, j)
        band = _band(Mout.to_host(e))
        for k, blk in band.items():
            assert np.array_equal(blk, F[off[k[0]]:off[k[0] + 1], off[k[1]]:off[k[1] + 1]])
