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
    from jues import _lib
    o = _lib.default_opts()
    assert o.maxit == 100 and o.diis == 1 and o.diis_max == 5 and o.e_conv == 1e-10


def test_no_oracle_in_product_path():
    """The shipped package never imports the oracle."""
    pkg = ROOT / "jues.jl_amd"
    for f in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.jl")):
        assert "jues_oracle" not in f.read_text(), f
        assert "oracle/" not in f.read_text(), f
