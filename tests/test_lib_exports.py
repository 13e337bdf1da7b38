"""The C-ABI library loads and exports every function include/jues.h declares (no GPU needed)."""
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parents[1]


def header_functions():
    txt = (ROOT / "include" / "jues.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(jues_[A-Za-z0-9_]+)\s*\(", txt)
    return sorted(set(n for n in names if n not in ("jues_iter_cb",)))


def test_header_matches_binding_list():
    from jues import _lib
    assert set(header_functions()) == set(_lib.EXPORTS)


def test_library_exports_every_symbol():
    from jues import _lib
    lib = _lib.load()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_default_opts():
    """Every field of the ctypes mirror reads the documented default: a field added to or moved in
    jues_cc_opts without the mirror shifts the later ones and fails here (the last field included)."""
    from jues import _lib
    o = _lib.default_opts()
    assert o.maxit == 100 and o.diis == 1 and o.diis_max == 5 and o.e_conv == 1e-10
    assert o.r_conv == 1e-9 and o.guess == _lib.GUESS["mp2"] and o.ccsd_eqs == 0 and o.fcn == 0
    assert o.fp32_r_conv == 1e-5 and o.refine_maxit == 60 and o.verbose == 0
    assert o.coupling == 1.0 and o.continuation == 1 and o.ladder == _lib.LADDER["auto"]
    assert o.refine == _lib.REFINE["restart"]


def test_no_oracle_in_product_path():
    """The shipped package never imports the oracle."""
    pkg = ROOT / "jues.jl_amd"
    for f in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.jl")):
        assert "jues_oracle" not in f.read_text(), f
        assert "oracle/" not in f.read_text(), f
