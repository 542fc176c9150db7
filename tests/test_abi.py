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
    return sorted(set(re.findall(r"\b(?:int|size_t)\s+(snapi_\w+)\s*\(", txt)))


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
             ("snapi_chem_table", abi.SnapiChemTable), ("snapi_state_map", abi.SnapiStateMap),
             ("snapi_level", __import__("snapi_amd.lines", fromlist=["SnapiLevel"]).SnapiLevel),
             ("snapi_node_model", __import__("snapi_amd.invert", fromlist=["SnapiNodeModel"]).SnapiNodeModel),
             ("snapi_instrument", __import__("snapi_amd.invert", fromlist=["SnapiInstrument"]).SnapiInstrument),
             ("snapi_nlte_atom", abi.SnapiNlteAtom)]
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


def _cfg1_inputs(ncol=1):
    import numpy as np
    from snapi_amd import workloads
    t = workloads.template_cfg1()
    st = workloads.cube_state(t, 1, ncol, seed=3, dtype=np.float64)
    lam = np.ascontiguousarray(t["lam_vac"][::50])
    return t, st, lam


def test_responses_refuse_an_inclined_ray(libpath):
    """Per-depth responses and the adjoint exist only on the mu = 1 ray; any
    other ray is refused before a device is touched."""
    import math
    import numpy as np
    from snapi_amd import abi
    lib = C.CDLL(libpath)
    t, st, lam = _cfg1_inputs()
    ds = np.zeros((st.shape[0], 2) + st.shape[1:])
    ray = abi.SnapiRay(theta=math.pi - math.acos(0.6), lam_mid=t["ray"].lam_mid, tau_grid=1, hminus=1, hminus_mid=1)
    stokes = np.zeros((st.shape[0], lam.size, 4))
    resp = np.zeros((st.shape[0], 7, st.shape[2], lam.size, 4))
    vp = C.c_void_p
    rc = lib.snapi_stokes_responses(abi.SNAPI_FP64, C.byref(t["table"]), C.byref(ray), st.shape[0], st.shape[2],
                                    vp(st.ctypes.data), vp(ds.ctypes.data), lam.size, vp(lam.ctypes.data),
                                    vp(stokes.ctypes.data), vp(resp.ctypes.data))
    assert rc == abi.SNAPI_ERR_UNSUPPORTED
    ws = (C.c_double * 4)(1, 1, 1, 1)
    grad = np.zeros((st.shape[0], 7, st.shape[2]))
    rc = lib.snapi_stokes_adjoint(abi.SNAPI_FP64, C.byref(t["table"]), C.byref(ray), st.shape[0], st.shape[2],
                                  vp(st.ctypes.data), vp(ds.ctypes.data), lam.size, vp(lam.ctypes.data),
                                  vp(stokes.ctypes.data), ws, vp(grad.ctypes.data))
    assert rc == abi.SNAPI_ERR_UNSUPPORTED


def test_adjoint_rejects_negative_sizes_before_allocating(libpath):
    import math
    import numpy as np
    from snapi_amd import abi
    lib = C.CDLL(libpath)
    t, st, lam = _cfg1_inputs()
    ds = np.zeros((st.shape[0], 2) + st.shape[1:])
    ray = abi.SnapiRay(theta=math.pi, lam_mid=t["ray"].lam_mid, tau_grid=1, hminus=1, hminus_mid=1)
    vp = C.c_void_p
    ws = (C.c_double * 4)(1, 1, 1, 1)
    grad = np.zeros(16)
    for ncol, nl in ((-1, lam.size), (1, -5), (0, lam.size)):
        rc = lib.snapi_stokes_adjoint(abi.SNAPI_FP64, C.byref(t["table"]), C.byref(ray), ncol, st.shape[2],
                                      vp(st.ctypes.data), vp(ds.ctypes.data), nl, vp(lam.ctypes.data),
                                      vp(st.ctypes.data), ws, vp(grad.ctypes.data))
        assert rc == abi.SNAPI_ERR_ARG, (ncol, nl, rc)


def test_status_api_checks_its_arguments(libpath):
    from snapi_amd import abi
    lib = C.CDLL(libpath)
    assert lib.snapi_context_create(None) == abi.SNAPI_ERR_ARG
    assert lib.snapi_context_status(None, None, 0) == abi.SNAPI_ERR_ARG
    assert lib.snapi_context_destroy(None) == abi.SNAPI_ERR_ARG
    assert lib.snapi_context_bind(None) == abi.SNAPI_OK
