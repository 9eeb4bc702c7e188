"""The C-ABI libraries load and export every function their headers declare.

No compute call is made here: this runs without a GPU.
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {
    "fkgpu.h": "fk_mc_amd/lib/libfkgpu.so",
    "fkmc_host.h": "fk_mc_amd/lib/libfkmc_host.so",
}
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(fk\w+)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = [n for n in DECL.findall(text) if not n.isupper()]
    return sorted(set(names))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_exports(header):
    path = os.path.join(ROOT, HEADERS[header])
    if not os.path.exists(path):
        pytest.fail(f"{path} not built")
    if header == "fkmc_host.h":  # the host layer resolves libfkgpu.so through its rpath
        ctypes.CDLL(os.path.join(ROOT, HEADERS["fkgpu.h"]))
    lib = ctypes.CDLL(path)
    names = declared(header)
    assert len(names) >= 3
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported {missing}"


def test_header_parse_sees_core_entry_points():
    names = declared("fkgpu.h")
    for n in ("fkgpu_create", "fkgpu_eig_logz", "fkgpu_sweep", "fkgpu_adjoint", "fkgpu_cheb_logz"):
        assert n in names
    assert "fkmc_run_chain" in declared("fkmc_host.h")


def test_product_has_no_cpu_fallback():
    """The product modules never import the oracle."""
    pkg = os.path.join(ROOT, "fk_mc_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", ".hpp")):
                src = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in src.replace("the oracle is", ""), os.path.join(dirpath, f)
