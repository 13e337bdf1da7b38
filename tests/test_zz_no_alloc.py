"""jues_rmp2_noalloc (the allocation-free FP64 RMP2 entry point) against jues_rmp2 on the
same inputs, through the C ABI: two calls in a row, the second on a smaller problem (fewer
occupied and virtual orbitals of the same factors), plus the size limit and the
requirement of factored integrals.  The device-memory counter of the library does not
move during the allocation-free calls."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _energies(ctx, s, o, v):
    from jues import _lib
    Co = _lib.f64(s.C[:, :o])
    Cv = _lib.f64(s.C[:, s.o:s.o + v])
    eps = _lib.f64(np.concatenate([s.eps[:o], s.eps[s.o:s.o + v]]))
    ref, got = np.zeros(1), np.zeros(1)
    ctx.check(ctx.lib.jues_rmp2(ctx.h, _lib.FP64, _lib.ptr(Co), o, _lib.ptr(Cv), v, s.n, _lib.ptr(eps),
                                _lib.ptr(ref)))
    before = ctx.lib.jues_device_bytes(ctx.h)
    ctx.check(ctx.lib.jues_rmp2_noalloc(ctx.h, _lib.FP64, _lib.ptr(Co), o, _lib.ptr(Cv), v, s.n, _lib.ptr(eps),
                                        _lib.ptr(got)))
    assert ctx.lib.jues_device_bytes(ctx.h) == before
    return float(ref[0]), float(got[0])


def test_noalloc_matches_rmp2_two_calls():
    from jues import _lib, synthetic
    from jues.Wavefunction import FactoredERI, ensure_ao
    s = synthetic.workload("cfg3_mp2grad_n256")
    ctx = _lib.context(0)
    ensure_ao(ctx, FactoredERI(s.B))
    for o, v in ((s.o, s.v), (11, 57)):          # the second call is smaller
        ref, got = _energies(ctx, s, o, v)
        print(f"o={o} v={v}: jues_rmp2 {ref:.15e} jues_rmp2_noalloc {got:.15e} diff {abs(got - ref):.1e}")
        assert abs(got - ref) < 1e-12 * abs(ref)


def test_noalloc_python_mirror_and_limits(cfg1):
    from jues import _lib, MollerPlesset
    from jues.Wavefunction import Wfn
    w = Wfn.from_synthetic(cfg1, dense=False)
    assert abs(MollerPlesset.do_rmp2_noalloc(w) - MollerPlesset.do_rmp2(w)) < 1e-12 * abs(MollerPlesset.do_rmp2(w))
    ctx = _lib.context(0)
    s = cfg1
    # beyond the scratch maximum: a clear error, nothing computed
    rng = np.random.default_rng(1)
    big = np.asfortranarray(rng.standard_normal((s.n, 6000)))
    eps = np.concatenate([-np.ones(6000), np.ones(6000)])
    E = np.zeros(1)
    st = ctx.lib.jues_rmp2_noalloc(ctx.h, _lib.FP64, _lib.ptr(big), 6000, _lib.ptr(big), 6000, s.n, _lib.ptr(eps),
                                   _lib.ptr(E))
    assert st == _lib.JUES_BAD_ARG
    msg = ctx.lib.jues_last_error(ctx.h).decode()
    assert "scratch" in msg and str(_lib.NOALLOC_SCRATCH_BYTES) in msg
    Co, Cv = _lib.f64(s.Co), _lib.f64(s.Cv)
    e = _lib.f64(s.eps)
    assert ctx.lib.jues_rmp2_noalloc(ctx.h, _lib.FP32_MIXED, _lib.ptr(Co), s.o, _lib.ptr(Cv), s.v, s.n,
                                     _lib.ptr(e), _lib.ptr(E)) == _lib.JUES_BAD_ARG
    # dense AO integrals are not accepted (their transform needs n^3-sized intermediates)
    wd = Wfn.from_synthetic(cfg1, dense=True)
    with pytest.raises(TypeError):
        MollerPlesset.do_rmp2_noalloc(wd)
    MollerPlesset.do_rmp2(wd)                       # makes the dense tensor current
    assert ctx.lib.jues_rmp2_noalloc(ctx.h, _lib.FP64, _lib.ptr(Co), s.o, _lib.ptr(Cv), s.v, s.n, _lib.ptr(e),
                                     _lib.ptr(E)) == _lib.JUES_BAD_ARG
    # the entry point still works afterwards
    assert abs(MollerPlesset.do_rmp2_noalloc(w) - MollerPlesset.do_rmp2(w)) < 1e-12 * abs(MollerPlesset.do_rmp2(w))
