"""The C-ABI library loads and exports every function include/snapi_gpu.h
declares.  CPU only: no compute call is made."""
import ctypes as C
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(REPO, "include", "snapi_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(snapi_\w+)\s*\(", txt)))


@pytest.fixture(scope="module")
def libpath():
    from snapi_amd import build
    if not os.path.exists(build.LIB):
        build.build()
    return build.LIB


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("snapi_synth_stokes", "snapi_synth_stokes_dev", "snapi_stokes_responses", "snapi_lm_step",
              "snapi_formal_scalar"):
        assert n in names


def test_library_exports_every_declared_symbol(libpath):
    lib = C.CDLL(libpath)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_matches_header():
    """ctypes mirror sizes and field offsets agree with the C layout compiled
    by the host compiler."""
    import subprocess
    import tempfile
    from snapi_amd import abi
    pairs = [("snapi_line", abi.SnapiLine), ("snapi_component", abi.SnapiComponent),
             ("snapi_line_table", abi.SnapiLineTable), ("snapi_ray", abi.SnapiRay),
             ("snapi_ali_transition", abi.SnapiAliTransition), ("snapi_ali_atom", abi.SnapiAliAtom),
             ("snapi_ali_grid", abi.SnapiAliGrid), ("snapi_partition", abi.SnapiPartition),
             ("snapi_chem_element", abi.SnapiChemElement), ("snapi_chem_molecule", abi.SnapiChemMolecule),
             ("snapi_chem_level", abi.SnapiChemLevel), ("snapi_chem_bf", abi.SnapiChemBf),
             ("snapi_chem_table", abi.SnapiChemTable), ("snapi_state_map", abi.SnapiStateMap)]
    lines = []
    want = []
    for cname, cls in pairs:
        lines.append('printf("%%zu\\n", sizeof(%s));' % cname)
        want.append(C.sizeof(cls))
        for fname, _ in cls._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, fname))
            want.append(getattr(cls, fname).offset)
    src = ('#include <stdio.h>\n#include <stddef.h>\n#include "%s/include/snapi_gpu.h"\nint main(){%s return 0;}\n'
           % (REPO, "".join(lines)))
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "t.c"), "w").write(src)
        subprocess.run(["gcc", "-o", os.path.join(d, "t"), os.path.join(d, "t.c")], check=True)
        out = subprocess.run([os.path.join(d, "t")], capture_output=True, text=True, check=True).stdout.split()
    assert [int(x) for x in out] == want


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from snapi_amd import _native
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        _native.lib()
