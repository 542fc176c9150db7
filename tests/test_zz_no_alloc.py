"""The allocation-free ALI iteration (snapi_ali_solve_dev_noalloc) against
snapi_ali_solve_dev on the same device inputs: two calls in a row, the
second on a smaller batch, each equal to the allocating entry point's
result; and a problem beyond the fixed area is refused."""
import ctypes as C

import numpy as np
import pytest

from snapi_amd import abi, ali, workloads
from snapi_amd._native import lib

pytestmark = pytest.mark.gpu


def _declare(L):
    f = L.snapi_ali_solve_dev_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.POINTER(abi.SnapiAliAtom), C.POINTER(abi.SnapiAliGrid), C.c_int, C.c_int] + \
        [C.c_void_p] * 6 + [C.c_double, C.c_int] + [C.c_void_p] * 4
    L.snapi_ali_noalloc_capacity.restype = C.c_size_t
    return f


def _device_problem(p, ncol, torch):
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(p[k][:ncol])).to(dev)
         for k in ("colstate", "vproj", "gamma", "bgop", "bgem", "coll")}
    t["pops"] = torch.from_numpy(np.ascontiguousarray(p["pops"][:ncol])).to(dev)
    return t


def _run_alloc(p, t, ncol, precision, torch):
    pops = t["pops"].clone()
    it = torch.zeros(ncol, dtype=torch.int32, device=pops.device)
    rc = torch.zeros(ncol, dtype=torch.float64, device=pops.device)
    ali.solve_dev(p["atom"], p["grid"], t["colstate"], t["vproj"], t["gamma"], t["bgop"], t["bgem"], t["coll"],
                  pops, it, rc, tol=1e-4, max_iter=60, precision=precision)
    torch.cuda.synchronize()
    return pops.cpu().numpy(), it.cpu().numpy(), rc.cpu().numpy()


def _run_noalloc(f, p, t, ncol, nd, precision, torch):
    pops = t["pops"].clone()
    it = torch.zeros(ncol, dtype=torch.int32, device=pops.device)
    rc = torch.zeros(ncol, dtype=torch.float64, device=pops.device)
    torch.cuda.synchronize()
    code = f(precision, C.byref(p["atom"]), C.byref(p["grid"]), ncol, nd, t["colstate"].data_ptr(),
             t["vproj"].data_ptr(), t["gamma"].data_ptr(), t["bgop"].data_ptr(), t["bgem"].data_ptr(),
             t["coll"].data_ptr(), 1e-4, 60, pops.data_ptr(), it.data_ptr(), rc.data_ptr(), None)
    torch.cuda.synchronize()
    assert code == abi.SNAPI_OK, code
    return pops.cpu().numpy(), it.cpu().numpy(), rc.cpu().numpy()


@pytest.mark.parametrize("precision", [abi.SNAPI_FP64, abi.SNAPI_MIXED])
def test_noalloc_matches_allocating_entry_point_twice(precision):
    import torch
    L = lib()
    f = _declare(L)
    nd = 61
    p = workloads.ali_problem(4, 4, nd=nd, nlam=120, nmu=2, seed=5)
    ncols = (16, 6)  # second call smaller: the area keeps the first call's contents
    for k, ncol in enumerate(ncols):
        t = _device_problem(p, ncol, torch)
        print("noalloc ALI call %d: precision %d, %d columns x %d depths x %d wavelengths x %d rays"
              % (k, precision, ncol, nd, p["nlam"], p["nray"]), flush=True)
        a_pops, a_it, a_rc = _run_alloc(p, t, ncol, precision, torch)
        n_pops, n_it, n_rc = _run_noalloc(f, p, t, ncol, nd, precision, torch)
        err = float(np.max(np.abs(n_pops / a_pops - 1.0)))
        print("  iterations %s / %s, max rel diff %.2e" % (n_it.tolist(), a_it.tolist(), err), flush=True)
        assert np.array_equal(n_it, a_it)
        # the same kernels in the same order on the same inputs
        assert err <= 1e-12
        assert np.allclose(n_rc, a_rc, rtol=1e-12, atol=0)


def test_noalloc_refuses_a_problem_beyond_its_area():
    import torch
    L = lib()
    f = _declare(L)
    cap = L.snapi_ali_noalloc_capacity()
    nd = 61
    p = workloads.ali_problem(1, 1, nd=nd, nlam=120, nmu=2, seed=5)
    t = _device_problem(p, 1, torch)
    need_per_col = nd * p["atom"].ntrans * 4 * 8
    ncol = int(cap // need_per_col) + 8  # more columns than the sums alone can hold
    dummy = t["pops"]
    print("noalloc ALI refusal: capacity %d bytes, %d columns" % (cap, ncol), flush=True)
    code = f(abi.SNAPI_FP64, C.byref(p["atom"]), C.byref(p["grid"]), ncol, nd, dummy.data_ptr(), dummy.data_ptr(),
             dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), 1e-4, 5, dummy.data_ptr(),
             dummy.data_ptr(), dummy.data_ptr(), None)
    assert code == abi.SNAPI_ERR_CAPACITY
