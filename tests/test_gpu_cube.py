"""The model cube to column states on the GPU (snapi_cube_to_state_dev) and
column spectra to the output cube (snapi_stokes_to_cube_dev): the cfg0
column written as a 1 x 1 cube and read back through the equation of state
reproduces the reference's own synthesis-path state (tests/golden/
cfg0_synth.npz); a seeded 3 x 2 cube against the oracle's restatement of
read_spinor3d's mapping, for the mu = 1 ray and an inclined one; the spectra
cube against a plain transpose, written and read back as an ANA file."""
import os

import numpy as np
import pytest

import _golden
import _oracle
from snapi_amd import abi, chemeq, cube, workloads

pytestmark = pytest.mark.gpu
K_B = 1.3806503E-16


def test_cfg0_cube_reproduces_reference_state(tmp_path):
    import torch
    g = _golden.load("cfg0_synth.npz")
    col = workloads.column_from_dump(os.path.join(_golden.HERE, "cfg0_synth.npz"))
    d = _golden.chemeq()
    dd = {k: v for k, v in d.items() if k != "table"}
    tb = abi.make_chem_table(dd, referent=0)
    smap = abi.make_state_map(tb, dd, g["lines"], g["bfsig"], col["ray"].lam_mid)
    T, Ne, Nt, Vt, Vr, Bm, Bi, Ba, opref, tauref, grid, nhi, hgt = g["depth"].T
    nz = len(T)
    c = np.zeros((12, 1, 1, nz))
    c[0, 0, 0], c[1, 0, 0], c[2, 0, 0] = np.log10(-tauref), hgt, T
    c[3, 0, 0], c[4, 0, 0], c[6, 0, 0] = Nt * K_B * T, Ne * K_B * T, opref
    c[7, 0, 0], c[8, 0, 0], c[9, 0, 0] = Bm, Vt, Vr
    c[10, 0, 0], c[11, 0, 0] = np.radians(Bi), np.radians(Ba)
    p = str(tmp_path / "cfg0_cube.f0")
    cube.ana_write(p, c)
    cc = cube.read_spinor3d(p)
    nf = abi.nfield(smap.nline)
    print("cfg0 as a 1 x 1 x %d cube -> state -> equation of state" % nz, flush=True)
    dev = torch.device("cuda")
    st = torch.zeros((1, nf, nz), dtype=torch.float64, device=dev)
    cube.cube_to_state_dev(torch.from_numpy(cc).to(dev), nf, st)
    chemeq.eos_state_dev(tb, smap, st)
    torch.cuda.synchronize()
    got, ref = st.cpu().numpy()[0], col["state"]
    fields = [abi.F_T, abi.F_VT, abi.F_VLOS, abi.F_BMAG, abi.F_BINC, abi.F_BAZI, abi.F_NE, abi.F_NHI, abi.F_OPREF,
              abi.F_GRID, abi.F_BGOP, abi.F_BGEM, abi.F_NT] + list(range(abi.NF_BASE, nf))
    for f in fields:
        e = np.max(np.abs(got[f] - ref[f]) / np.maximum(np.abs(ref[f]), 1e-30))
        assert e < 1e-12, (f, e)
    print("  all fields within 1e-12 of the reference's state", flush=True)


@pytest.mark.parametrize("theta,phi", [(np.pi, 0.0), (2.6, 0.7)])
def test_seeded_cube_matches_oracle(theta, phi):
    import torch
    rng = np.random.default_rng(8)
    nx, ny, nz, nf = 3, 2, 41, abi.nfield(2)
    c = np.zeros((12, nx, ny, nz))
    c[0] = np.linspace(-5, 1, nz)[None, None]
    c[2] = rng.uniform(3000, 9000, (nx, ny, nz))
    c[3] = 10 ** rng.uniform(2, 5, (nx, ny, nz))
    c[4] = 10 ** rng.uniform(-2, 2, (nx, ny, nz))
    c[6] = rng.uniform(1e-9, 1e-5, (nx, ny, nz))
    c[7] = rng.uniform(1, 3000, (nx, ny, nz))
    c[8] = rng.uniform(0, 2e5, (nx, ny, nz))
    c[9] = rng.normal(0, 2e5, (nx, ny, nz))
    c[10] = rng.uniform(0.05, 3.0, (nx, ny, nz))
    c[11] = rng.uniform(0.05, 3.0, (nx, ny, nz))
    print("seeded cube %d x %d x %d, theta %.2f phi %.2f" % (nx, ny, nz, theta, phi), flush=True)
    dev = torch.device("cuda")
    st = torch.zeros((nx * ny, nf, nz), dtype=torch.float64, device=dev)
    cube.cube_to_state_dev(torch.from_numpy(c).to(dev), nf, st, theta=theta, phi=phi)
    torch.cuda.synchronize()
    want = _oracle.cube_to_state(c, nf, theta, phi)
    e = np.max(np.abs(st.cpu().numpy() - want) / np.maximum(np.abs(want), 1e-300))
    print("  max rel diff %.2e" % e, flush=True)
    assert e < 1e-13


def test_spectra_cube(tmp_path):
    import torch
    rng = np.random.default_rng(4)
    nx, ny, nl = 4, 3, 17
    s = rng.standard_normal((nx * ny, nl, 4))
    dev = torch.device("cuda")
    out = torch.empty((nx, ny, 4, nl), dtype=torch.float64, device=dev)
    print("spectra cube %d x %d x 4 x %d" % (nx, ny, nl), flush=True)
    for dt in (torch.float64, torch.float32):
        st = torch.from_numpy(s).to(dev).to(dt)
        cube.stokes_to_cube_dev(st, nx, ny, out)
        torch.cuda.synchronize()
        want = st.cpu().double().numpy().reshape(nx, ny, nl, 4).transpose(0, 1, 3, 2)
        assert np.array_equal(out.cpu().numpy(), want)
    p = str(tmp_path / "spectra.f0")
    cube.ana_write(p, out.cpu().numpy())
    assert cube.ana_info(p)[:2] == (4, [nl, 4, ny, nx])
    assert np.array_equal(cube.ana_read(p), out.cpu().numpy())
