"""GEMM, AO->MO transform, get_eri and RMP2 on the GPU against the CPU oracle."""
import numpy as np
import pytest

import jues_oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 2e-5)])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (5, 7, 3), (127, 129, 33), (256, 128, 64), (300, 200, 257),
                                   (64, 1000, 17), (60, 60, 50000), (200, 150, 9000),
                                   # skinny tiles (64x256 / 256x64): full tiles, ragged edges, K not a slice multiple
                                   (40, 9000, 256), (64, 8192, 33), (9000, 40, 300), (8195, 17, 64),
                                   # short, wide FP64 products: the 64x256 kernel, a transposed A copied first
                                   (17, 9000, 64), (33, 10000, 96), (64, 16400, 512), (1, 8192, 16),
                                   (33, 9001, 272), (49, 8200, 160),
                                   # 3, 5 and 11 sixteen-deep k-slices: the short kernels' pipeline tails
                                   (20, 8300, 48), (45, 8500, 80), (48, 8200, 176)])
def test_gemm(ctx, prec, tol, ta, tb, M, N, K):
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.normal(size=(K, M) if ta == "T" else (M, K))
    B = rng.normal(size=(N, K) if tb == "T" else (K, N))
    C0 = rng.normal(size=(M, N))
    alpha, beta = 0.7, -1.3
    ref = alpha * ((A.T if ta == "T" else A) @ (B.T if tb == "T" else B)) + beta * C0
    out = ctx.gemm_test(prec, ta, tb, alpha, A, B, beta, C0)
    assert rel_err(out, ref) < tol * max(1.0, np.sqrt(K) / 4)


def test_gemm_beta_zero_ignores_nan(ctx):
    rng = np.random.default_rng(1)
    A, B = rng.normal(size=(70, 40)), rng.normal(size=(40, 90))
    C0 = np.full((70, 90), np.nan)
    out = ctx.gemm_test(0, "N", "N", 1.0, A, B, 0.0, C0)
    assert rel_err(out, A @ B) < 1e-13


@pytest.mark.parametrize("dense", [True, False])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 3e-6)])
def test_tei_transform(cfg1, dense, prec, tol):
    from jues import Transformation
    from jues.Wavefunction import FactoredERI
    s = cfg1
    g = s.ao_eri()
    eri = g if dense else FactoredERI(s.B)
    Co, Cv = s.Co, s.Cv
    for Cs in [(Co, Cv, Co, Cv), (Cv, Cv, Co, Co), (Co, Co, Co, Cv), (Cv, Co, Cv, Co), (s.C,) * 4]:
        ref = O.tei_transform(g, *Cs)
        out = Transformation.tei_transform(eri, *Cs, prec=prec)
        assert out.shape == ref.shape
        assert rel_err(out, ref) < tol, Cs


def test_tei_transform_general_tensor(ctx):
    """No symmetry assumed: random non-symmetric g and four different C's."""
    from jues import Transformation
    rng = np.random.default_rng(5)
    n = 9
    g = rng.normal(size=(n, n, n, n))
    Cs = [rng.normal(size=(n, d)) for d in (3, 4, 5, 2)]
    ref = O.tei_transform_bruteforce(g, *Cs) if n <= 4 else O.tei_transform(g, *Cs)
    out = Transformation.tei_transform(np.asfortranarray(g), *Cs)
    assert rel_err(out, ref) < 1e-12


def test_tei_transform_dense_n144(ctx):
    """Dense route at n = 144 (K = 144 > 128, a multiple of 16), where the quarter products run on
    the tall and the short batched kernels, with four coefficient blocks of no symmetry: both orders
    of the first two quarters (d1 <= d4 and d1 > d4) and of the last two (d3 <= d2 and d3 > d2), in
    FP64 and FP32-mixed.  The dense tensor is g = sum_P B_P (x) B_P with B_P not symmetric in
    (mu, nu), so the oracle's exact 3-index route gives the reference in seconds at this size."""
    from jues import Transformation
    n, o, v, naux = 144, 24, 120, 48
    rng = np.random.default_rng(144)
    B = rng.standard_normal((n, n, naux))
    B2 = B.reshape((n * n, naux), order="F")
    g = (B2 @ B2.T).T.reshape((n,) * 4, order="F")    # symmetric n^2 x n^2: its transpose is the same data
    Co, Cv = rng.normal(size=(n, o)), rng.normal(size=(n, v))
    Co2, Cv2 = rng.normal(size=(n, o)), rng.normal(size=(n, v))
    for Cs in [(Co, Cv, Co2, Cv2), (Co, Co2, Cv, Cv2), (Cv, Co, Cv2, Co2), (Cv, Cv2, Co, Co2)]:
        ref = O.mo_from_factors(B, *Cs)
        for prec, tol in [("fp64", 1e-12), ("fp32", 2e-5)]:
            out = Transformation.tei_transform(g, *Cs, prec=prec)
            print(f"dense n=144 dims={[c.shape[1] for c in Cs]} {prec}: rel err {rel_err(out, ref):.2e}", flush=True)
            assert rel_err(out, ref) < tol, ([c.shape[1] for c in Cs], prec)


@pytest.mark.parametrize("dims", [(3, 4, 5, 2), (4, 3, 2, 5), (6, 6, 6, 6)])
def test_tei_transform_nonsymmetric_factors(ctx, dims):
    """Factor route with B[mu,nu,P] != B[nu,mu,P]: both contraction orders of the
    half-transform (d1 <= d2 and d1 > d2) against the exact 3-index sum."""
    from jues import Transformation
    from jues.Wavefunction import FactoredERI
    rng = np.random.default_rng(6)
    n, naux = 9, 7
    B = np.asfortranarray(rng.normal(size=(n, n, naux)))
    Cs = [rng.normal(size=(n, d)) for d in dims]
    ref = O.mo_from_factors(B, *Cs)
    out = Transformation.tei_transform(FactoredERI(B), *Cs)
    assert rel_err(out, ref) < 1e-12


@pytest.mark.parametrize("dense", [True, False])
def test_get_eri(cfg1, dense):
    from jues import IntegralTransformation as IT
    from jues.Wavefunction import Wfn
    s = cfg1
    g = s.ao_eri()
    w = Wfn.from_synthetic(s, dense=dense)
    for blk in ["OOVV", "OVOV", "OVVO", "OOOO", "VVVV", "ooov", "ovvv"]:
        for note in ["phys", "chem"]:
            ref = O.get_eri(g, s.Co, s.Cv, blk, notation=note)
            out = IT.get_eri(w, blk, notation=note)
            assert rel_err(out, ref) < 1e-12, (blk, note)
    ref = O.get_eri(g, s.Co, s.Cv, "OOVV", fcn=1)
    assert rel_err(IT.get_eri(w, "OOVV", fcn=1), ref) < 1e-12
    from jues import JuesError
    with pytest.raises(JuesError):
        IT.get_eri(w, "OOV")
    with pytest.raises(JuesError):
        IT.get_eri(w, "OOXV")


@pytest.mark.parametrize("dense", [True, False])
def test_rmp2(cfg1, dense):
    from jues import MollerPlesset
    from jues.Wavefunction import Wfn
    s = cfg1
    g = s.ao_eri()
    ref = O.rmp2(O.get_eri(g, s.Co, s.Cv, "OOVV"), s.eps, s.o)
    w = Wfn.from_synthetic(s, dense=dense)
    E = MollerPlesset.do_rmp2(w)
    assert abs(E - ref) < 1e-10
    E32 = MollerPlesset.do_rmp2(w, prec="fp32_mixed")
    assert abs(E32 - ref) < 1e-5 * abs(ref)


@pytest.mark.parametrize("prec,tol", [(0, 1e-14), (1, 2e-6)])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("M,N,K", [(128, 128, 16), (129, 131, 33), (300, 257, 1000), (1000, 1000, 32),
                                   (256, 384, 48), (5, 7, 3), (1, 1, 16), (2048, 640, 1040), (40, 9000, 256),
                                   (9000, 40, 256), (257, 300, 17), (640, 520, 4111), (130, 129, 31),
                                   (2048, 2048, 256), (2100, 2300, 1041), (4099, 2049, 513),
                                   # the packed ladder's M = o(o+1)/2 at o = 60: a last tile row of 38 rows
                                   (1830, 8200, 4100),
                                   # tall, narrow FP64 NN products (256x48 and 256x64 tiles): ragged last tile
                                   # row, full-width tiles, one live column in the last 16-column block
                                   (8300, 60, 272), (8192, 48, 160), (8193, 49, 144), (10000, 64, 256)])
def test_gemm_nn_nt_occupancy_kernels(ctx, M, N, K, tb, prec, tol):
    """Products with A m-contiguous and B k- or n-contiguous on the high-occupancy kernels
    (FP32: four 256-thread workgroups per CU; FP64: two 512-thread ones): full 16-column
    slices, a zero-padded K-tail slice (1..15 columns), edge tiles with clamped rows, skinny shapes."""
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N)) if tb == "N" else rng.standard_normal((N, K))
    C0 = rng.standard_normal((M, N))
    ref = 0.7 * (A @ (B if tb == "N" else B.T)) - 0.3 * C0
    out = ctx.gemm_test(prec, "N", tb, 0.7, np.asfortranarray(A), np.asfortranarray(B), -0.3, np.asfortranarray(C0))
    assert rel_err(out, ref) < tol * max(1.0, np.sqrt(K) / 8)


@pytest.mark.parametrize("prec,tol", [(0, 1e-14), (1, 2e-6)])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("s0,s1", [((1830, 2000, 1000), (1770, 1990, 990)), ((130, 129, 31), (1000, 700, 1040)),
                                   ((1, 1, 16), (257, 300, 17)), ((300, 257, 100), (2, 3, 4)),
                                   ((2048, 640, 1040), (64, 9000, 33)),
                                   # few tiles, long K: both K ranges split into chunks (one launch) and reduced
                                   ((300, 257, 9000), (200, 150, 8000)), ((60, 60, 30000), (59, 58, 29001))])
def test_gemm_pair(ctx, s0, s1, tb, prec, tol):
    """Two products in one launch (the packed ladders' plus and minus parts): each against
    its own reference, with the first product's tile count padded to a multiple of 8, the
    split-K form for few tiles and long K, and the fallback to two launches (K < 16, skinny)."""
    rng = np.random.default_rng(sum(s0) + sum(s1))
    args, refs = [], []
    for (M, N, K) in (s0, s1):
        A = np.asfortranarray(rng.standard_normal((M, K)))
        B = np.asfortranarray(rng.standard_normal((K, N)) if tb == "N" else rng.standard_normal((N, K)))
        C0 = np.asfortranarray(rng.standard_normal((M, N)))
        args += [A, B, C0]
        refs.append(0.7 * (A @ (B if tb == "N" else B.T)) - 0.3 * C0)
    out0, out1 = ctx.gemm_pair_test(prec, tb, 0.7, *args, -0.3)
    for out, ref, (M, N, K) in zip((out0, out1), refs, (s0, s1)):
        assert rel_err(out, ref) < tol * max(1.0, np.sqrt(K) / 8)
