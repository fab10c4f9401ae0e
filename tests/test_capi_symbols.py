"""CPU checks of the C-ABI boundary: the library loads and exports every
function that include/*.h declares (no compute call is made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
DECL = re.compile(r"^\s*(?:unsigned\s+int|int|void|double|const\s+char\s*\*)\s*\*?\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", re.M)


def declared():
    names = set()
    for f in os.listdir(INCLUDE):
        if f.endswith(".h"):
            text = open(os.path.join(INCLUDE, f)).read()
            names.update(DECL.findall(text))
    return sorted(names)


def test_headers_declare_the_boundary():
    names = declared()
    for must in ["cblas_dgemm", "cblas_dtrsm", "cblas_dsyrk", "cblas_idamax", "cblas_dscal",
                 "cblas_dswap", "LAPACKE_dgetrf", "LAPACKE_dgetrs", "LAPACKE_dgesv",
                 "LAPACKE_dpotrf", "LAPACKE_dpotrs", "LAPACKE_dposv", "LAPACKE_dlaswp",
                 "zmlhip_dgetrf", "zmlhip_dgemm"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from zml_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libzmlhip.so is not built (python -m zml_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes signature table covers the same set
    assert set(_lib.SIGNATURES) == set(declared())


def test_loader_fails_loudly_without_library(tmp_path):
    from zml_amd import _lib
    saved = _lib._lib
    _lib._lib = None
    try:
        with pytest.raises(RuntimeError, match="missing"):
            _lib.load(str(tmp_path / "nope.so"))
    finally:
        _lib._lib = saved


def test_product_never_touches_the_oracle():
    pkg = os.path.join(ROOT, "zml_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text, f
                assert "zml_oracle" not in text, f


def test_zig_declarations_match_the_c_abi():
    """The Zig front-end's extern declarations (zml_amd/zig/zmlhip.zig) name
    exported functions with the parameter counts include/zmlhip.h gives them
    (the Zig sources are not compiled here: no Zig toolchain)."""
    zig = open(os.path.join(ROOT, "zml_amd", "zig", "zmlhip.zig")).read()
    hdr = open(os.path.join(INCLUDE, "zmlhip.h")).read()
    decls = re.findall(r"pub extern fn (\w+)\(([^)]*)\)", zig, re.S)
    assert len(decls) >= 15
    for name, params in decls:
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", hdr, re.S)
        assert m, f"{name} is not declared in zmlhip.h"
        n_zig = len([p for p in params.split(",") if p.strip()])
        n_c = len([p for p in m.group(1).split(",") if p.strip() and p.strip() != "void"])
        assert n_zig == n_c, (name, n_zig, n_c)
    assert set(n for n, _ in decls) <= set(declared())
