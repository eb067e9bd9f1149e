"""The C-ABI: both HIP libraries load (no GPU needed) and export every
symbol include/*.h declares; the ctypes table covers the same drop-ins."""
import ctypes
import glob
import os
import re
import shutil
import subprocess

import pytest

from lsmlib_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBS = [os.path.join(ROOT, "lsmlib_amd", n) for n in ("liblsmhip.so", "liblsmhip_f32.so")]


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        macros = dict(re.findall(r"#define\s+(LSM[23]D_\w+)\s+(\w+_)\s*$", txt, flags=re.M))
        syms.update(macros.values())
        for name in re.findall(r"^(?:int|void|const char \*|void \*)\s*\*?\s*(lsm\w+)\(", txt, flags=re.M):
            syms.add(name)
    return syms


# the reference's C utilities, exported under their own names
REFERENCE_C = ["linearExtrapolationBC", "signedLinearExtrapolationBC", "copyExtrapolationBC", "homogeneousNeumannBC",
               "createGridSetDx", "createGridSetDxDyDz", "createGridSetGridDims", "destroyGrid",
               "writeGridToBinaryFile", "readGridFromBinaryFile", "writeDataArray", "writeDataArrayNoGrid",
               "readDataArray"]


def test_reference_named_utilities_declared_and_exported():
    txt = " ".join(open(h).read() for h in glob.glob(os.path.join(ROOT, "include", "*.h")))
    for name in REFERENCE_C:
        assert re.search(r"\b%s\(" % name, txt), name
    import torch  # noqa: F401
    for path in LIBS:
        if os.path.exists(path):
            lib = ctypes.CDLL(path)
            assert all(hasattr(lib, n) for n in REFERENCE_C), path


def test_headers_declare_the_spec_table():
    syms = declared_symbols()
    missing = [s for s in capi.SPECS if s not in syms]
    assert not missing, missing
    assert len(syms) > 90


@pytest.mark.parametrize("path", LIBS)
def test_library_exports_every_declared_symbol(path):
    if not os.path.exists(path):
        pytest.skip("not built")
    import torch  # noqa: F401  (the library shares torch's HIP runtime)
    lib = ctypes.CDLL(path)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, missing


def test_loader_reports_precision():
    import lsmlib_amd
    for prec, size in (("f64", 8), ("f32", 4)):
        lib = lsmlib_amd.load(prec)
        assert lib.cdll.lsmhip_real_size() == size
        assert lib.api.missing == []


def _expanded_prototypes():
    """Prototypes of include/lsmhip_all.h after the C preprocessor: name ->
    list of parameter types."""
    out = subprocess.run(["gcc", "-E", "-P", "-I", os.path.join(ROOT, "include"),
                          os.path.join(ROOT, "include", "lsmhip_all.h")], capture_output=True, text=True, check=True)
    txt = " ".join(out.stdout.split())
    protos = {}
    for name, params in re.findall(r"\b(lsm\w+)\s*\(([^;{}]*)\)\s*;", txt):
        types = []
        for prm in params.split(","):
            prm = prm.strip()
            if prm and prm != "void":
                types.append(re.sub(r"\s*\b\w+$", "", prm).replace(" ", ""))
        protos[name] = types
    return protos


# spec kind -> parameter type in the FP64 headers
_KIND_TYPE = {"R": ("double*", "constdouble*"), "S": ("constdouble*",), "SO": ("double*",),
              "I": ("constint*",), "IO": ("int*",), "IH": ("int*", "constint*"), "IA": ("int*", "constint*"),
              "M": ("unsignedchar*", "constunsignedchar*"), "U8": ("constunsignedchar*",)}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs the C preprocessor")
def test_prototypes_match_the_ctypes_table():
    """Every drop-in prototype has, position by position, the parameter kind
    the ctypes table passes (boxes are six const int pointers, 2D boxes four)."""
    protos = _expanded_prototypes()
    for sym, spec in capi.SPECS.items():
        assert sym in protos, sym
        want = []
        for kind, name in spec:
            if kind in ("B", "Q"):
                want.extend([("constint*",)] * (6 if kind == "B" else 4))
            else:
                want.append(_KIND_TYPE[kind])
        got = protos[sym]
        assert len(got) == len(want), (sym, len(got), len(want))
        for pos, (g, w) in enumerate(zip(got, want)):
            assert g in w, (sym, pos, g, w)
