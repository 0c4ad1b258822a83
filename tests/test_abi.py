"""The C-ABI library loads and exports every entry point include/*.h declares (no GPU needed)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from sarbp import _lib


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(sarbp_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    # typedef'd callback type names are not functions
    return {n for n in names if not n.endswith("_fn")}


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert names, "no declarations found"
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declaration():
    assert declared_functions() == set(_lib.SIGNATURES)


def test_abi_version_and_strerror():
    L = _lib.load()
    assert L.sarbp_abi_version() == 1
    assert L.sarbp_strerror(0) == b"ok"
    assert L.sarbp_strerror(-1) == b"invalid argument"


def test_host_pp_count_matches_oracle():
    from oracle import oracle_lib as O
    rng = np.random.default_rng(4)
    n, naz = 2000, 500
    idx = rng.integers(-60, naz + 60, n).astype(np.int32)
    s = rng.integers(-8, 300, n).astype(np.int32)
    s[::13] = -9999
    m = (rng.random(n) < 0.9).astype(np.uint8)
    from sarbp.core import count_pp
    assert count_pp(m, idx, s, naz) == O.count_pp(m, idx, s, naz)


def test_null_arguments_rejected_without_device_work():
    L = _lib.load()
    h = ctypes.c_void_p()
    st = L.sarbp_dataset_create(None, 10, 10, None, None, None, 4000.0, 0.5, 0.031, ctypes.byref(h))
    assert st == _lib.EINVAL and not h.value
    assert b"null" in L.sarbp_last_error()
    assert L.sarbp_scene_create(None, None, None, None, None, None, None, 0, 0, ctypes.byref(h)) == _lib.EINVAL
    assert L.sarbp_dataset_destroy(None) == 0


STRUCTS = [("sarbp_request", "Request"), ("sarbp_server_config", "ServerConfig"),
           ("sarbp_request_info", "RequestInfo"), ("sarbp_server_stats", "ServerStats")]


def test_ctypes_struct_layouts_match_header(tmp_path):
    """Every structure the bindings pass across the ABI has the header's size and field offsets
    (a C program built against include/sarbp.h prints them)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "sarbp.h"', "int main(void) {"]
    for cname, pyname in STRUCTS:
        lines.append(f'  printf("{pyname} size %zu\\n", sizeof({cname}));')
        for f, _ in getattr(_lib, pyname)._fields_:
            lines.append(f'  printf("{pyname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = {}
    for ln in out:
        if ln:
            py, field, v = ln.split()
            want[(py, field)] = int(v)
    for _, pyname in STRUCTS:
        cls = getattr(_lib, pyname)
        assert ctypes.sizeof(cls) == want[(pyname, "size")], pyname
        for f, _ in cls._fields_:
            assert getattr(cls, f).offset == want[(pyname, f)], (pyname, f)


def test_product_path_fails_loudly_without_a_device():
    """No CPU fallback: with no HIP device visible every focusing entry point raises, and argument
    errors still raise ValueError first, as the reference does (sar_focuser_puntual.py:174-183)."""
    if _lib.device_count() > 0:
        pytest.skip("a HIP device is visible")
    from sarbp import synth
    from sarbp.core import Dataset
    from sarbp.sar_focuser_puntual import FocalizzatoreBpPx
    sc = synth.tiny_scene(seed=1)
    with pytest.raises(_lib.SarbpError) as e:
        Dataset(sc.raw, sc.tx, sc.ty, sc.tz, sc.near, sc.dra, sc.lam)
    assert e.value.status == _lib.ENODEV
    with pytest.raises(_lib.SarbpError):
        FocalizzatoreBpPx(*sc.args())
    args = list(sc.args())
    args[0] = np.asarray(args[0]).ravel()  # xxg1 not 2-D
    with pytest.raises(ValueError):
        FocalizzatoreBpPx(*args)


def test_integration_stub_matches_the_header():
    """The ctypes stub that the integration guide at the repository root gives a reference
    maintainer declares the entry points it binds with the header's parameter counts and loads
    against the built library (declarations only)."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = text.split("## 3.")[1].split("```python")[1].split("```")[0]
    decl = block.split("def focus_gpu")[0].replace('ctypes.CDLL("libsarbp.so")', "ctypes.CDLL(LIB)")
    ns = {"LIB": _lib.LIB_PATH}
    exec(decl, ns)  # noqa: S102 -- the repository's own documentation
    L = ns["_L"]
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "sarbp.h")).read(), flags=re.S)
    for name in ("sarbp_dataset_create", "sarbp_focus", "sarbp_dataset_destroy"):
        params = re.search(rf"\b{name}\s*\(([^)]*)\)", hdr).group(1)
        assert len(getattr(L, name).argtypes) == len([p for p in params.split(",") if p.strip()]), name
    assert "def focus_gpu" in block
