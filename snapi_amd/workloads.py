"""Column state from reference column dumps, and the synthetic cube workload.

`column_from_dump` turns a column recorded from the reference code (the
tests/golden/*.npz files: the state its atmosphere object holds after
popsetup/compute_op_referent/nltepops) into this library's inputs: the line
table, the column state [nfield][nd] and the wavelength grid.  This is the
conversion a reference-side caller performs before calling the C ABI
(atmos/atmos_obs.cc:23-58 computes the same quantities).

`cube_state` draws the 512 x 512 x 71 LTE cube workload around the 71-depth
template column: per-column temperature offset from a Gaussian random field
(sigma 150 K, correlation length 8 px), |B| log-uniform on 10-3000 G,
cos(inclination) uniform on [-1, 1], azimuth uniform on [0, 180) deg, vLOS
N(0, 2 km/s) plus a gradient N(0, 0.5 km/s per decade of tau), vmicro
uniform on [0, 2] km/s.  The thermodynamic fields of each column (electron
and neutral-hydrogen densities, level populations, referent opacity,
background) are the template's, rescaled with Boltzmann/Saha-type factors
for the column's temperature: synthetic inputs of the right shape and
magnitude, not a solution of the equation of state.  `cube_state_eos_dev`
draws the same columns and solves their thermodynamics on the GPU as the
reference builds a model atmosphere (hydrostatic equilibrium, the 12 x NX x
NY x NZ model cube, the equation of state); that is the synthesis workload
of bench.py.
"""
import os

import numpy as np

from . import abi
from .physics import THOMSON, planck, K_BOLTZ

EV = 1.602189e-12
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")


def column_from_dump(path):
    d = np.load(path)
    dep = d["depth"]
    nd = dep.shape[0]
    T, Ne, Nt, Vt, Vr, Bm, Bi, Ba, opref, tauref, grid, nhi, hgt = dep.T
    lam_vac = d["lam"][:, 1]
    lam_mid = 0.5 * (lam_vac[0] + lam_vac[-1])
    masses = dict((int(k), v) for k, v in d["masses"])
    levels = d["levels"]
    lpop = d["lpop"]
    zc = d["zc"]
    lines, pops = [], []
    for (s, z, lo, up, tr), ld, sp in zip(d["lines"], d["line_data"], d["line_split"]):
        lam0, A, Blu, Bul, grad, sigma, alpha = ld
        comps = [[], [], []]
        for row in zc:
            if int(row[0]) == s and int(row[1]) == tr:
                comps[int(row[2])].append((row[3], row[4]))
        elo = [r[3] for r in levels if int(r[0]) == s and int(r[1]) == z and int(r[2]) == lo][0]
        lines.append(dict(lam0=lam0, aul=A, blu=Blu, bul=Bul, grad=grad, sigma=sigma, alpha=alpha,
                          mass=masses[int(s)], split=int(sp), comps=comps, species=int(s), tr=int(tr),
                          e_low=elo))
        m = (lpop[:, 0] == s) & (lpop[:, 1] == tr)
        rows = lpop[m]
        rows = rows[np.argsort(rows[:, 2])]
        pops.append((rows[:, 3], rows[:, 4]))
    nf = abi.nfield(len(lines))
    st = np.zeros((nf, nd))
    st[abi.F_T] = T
    st[abi.F_VT] = Vt
    st[abi.F_VLOS] = Vr
    st[abi.F_BMAG] = Bm
    st[abi.F_BINC] = Bi
    st[abi.F_BAZI] = Ba
    st[abi.F_NE] = Ne
    st[abi.F_NHI] = nhi
    st[abi.F_OPREF] = opref
    st[abi.F_GRID] = grid
    st[abi.F_NT] = Nt
    bf = d["bfop"]
    bfop = np.zeros(nd)
    bfem = np.zeros(nd)
    l1 = bf[:, 1].min()
    for row in bf[bf[:, 1] == l1]:
        bfop[int(row[2]) - 1] += row[3]
        bfem[int(row[2]) - 1] += row[4]
    thop = Ne * THOMSON
    st[abi.F_BGOP] = thop + bfop
    st[abi.F_BGEM] = thop * planck(lam_mid, T) + bfem
    for i, (nl, nu) in enumerate(pops):
        st[abi.NF_BASE + 2 * i] = nl
        st[abi.NF_BASE + 2 * i + 1] = nu
    ray = abi.SnapiRay(theta=np.pi, lam_mid=lam_mid, tau_grid=1, hminus=1)
    return dict(lines=lines, table=abi.make_line_table(lines), ray=ray, state=st, lam_vac=lam_vac,
                lam_air=d["lam"][:, 0], stokes=d["stokes"], bfop=bfop, bfem=bfem, logtau=np.log10(-tauref))


def gaussian_random_field(nx, ny, sigma, corr, rng):
    """White noise filtered with a Gaussian kernel of the given correlation
    length (pixels), renormalised to standard deviation sigma."""
    # drawn on a periodic grid at least 8 correlation lengths wide, then cut
    mx, my = max(nx, int(8 * corr)), max(ny, int(8 * corr))
    w = rng.standard_normal((mx, my))
    kx = np.fft.fftfreq(mx)[:, None]
    ky = np.fft.fftfreq(my)[None, :]
    filt = np.exp(-2.0 * (np.pi * corr) ** 2 * (kx ** 2 + ky ** 2))
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    f -= f.mean()
    f *= sigma / f.std()
    return np.ascontiguousarray(f[:nx, :ny])


def cube_state(template, nx, ny, seed=1, dtype=np.float32):
    """[nx*ny][nfield][nd] column states of the cube workload (see module doc)."""
    rng = np.random.default_rng(seed)
    st0 = template["state"]
    nf, nd = st0.shape
    ncol = nx * ny
    dT = gaussian_random_field(nx, ny, 150.0, 8.0, rng).reshape(ncol)
    logB = rng.uniform(np.log10(10.0), np.log10(3000.0), ncol)
    cosi = rng.uniform(-1.0, 1.0, ncol)
    azi = rng.uniform(0.0, 180.0, ncol)
    v0 = rng.normal(0.0, 2e5, ncol)
    vg = rng.normal(0.0, 0.5e5, ncol)
    vmic = rng.uniform(0.0, 2e5, ncol)
    logtau = template["logtau"]
    out = np.empty((ncol, nf, nd), dtype=np.float64)
    T0 = st0[abi.F_T]
    T = T0[None, :] + dT[:, None]
    inv = 1.0 / (K_BOLTZ * T0[None, :]) - 1.0 / (K_BOLTZ * T)  # 1/kT0 - 1/kT
    out[:] = st0[None]
    out[:, abi.F_T] = T
    out[:, abi.F_VT] = vmic[:, None]
    out[:, abi.F_VLOS] = v0[:, None] + vg[:, None] * logtau[None, :]
    out[:, abi.F_BMAG] = (10.0 ** logB)[:, None]
    out[:, abi.F_BINC] = np.degrees(np.arccos(cosi))[:, None]
    # the reference's ray-frame azimuth is atan(By/Bx): folded into (-90, 90)
    out[:, abi.F_BAZI] = np.degrees(np.arctan(np.tan(np.radians(azi))))[:, None]
    # Saha-type rescaling of the electron density (metal ionisation, ~7.9 eV / 2)
    fne = np.exp(3.95 * EV * inv)
    out[:, abi.F_NE] = st0[abi.F_NE][None] * fne
    # referent (H-) opacity ~ Ne T^-1.5 10^(0.754 * 5040 / T)
    fop = fne * (T0[None] / T) ** 1.5 * 10.0 ** (0.754 * 5040.0 * (1.0 / T - 1.0 / T0[None]))
    out[:, abi.F_OPREF] = st0[abi.F_OPREF][None] * fop
    lam_mid = template["ray"].lam_mid
    thop = out[:, abi.F_NE] * THOMSON
    out[:, abi.F_BGOP] = thop + template["bfop"][None] * fne
    out[:, abi.F_BGEM] = thop * planck(lam_mid, T) + template["bfem"][None] * fne * planck(lam_mid, T) / planck(
        lam_mid, T0[None])
    for i, ln in enumerate(template["lines"]):
        # Boltzmann factor of the level energy, times the neutral fraction change
        fb = np.exp(ln["e_low"] * inv) / fne
        out[:, abi.NF_BASE + 2 * i] = st0[abi.NF_BASE + 2 * i][None] * fb
        out[:, abi.NF_BASE + 2 * i + 1] = st0[abi.NF_BASE + 2 * i + 1][None] * fb
    return np.ascontiguousarray(out.astype(dtype))


def cube_draws(template, nx, ny, seed=1):
    """The per-column draws of the cube workload (the same random sequence
    as cube_state): temperature T [ncol][nd] (template + Gaussian random
    field), |B| [G], inclination and azimuth [rad] in the atmosphere's frame,
    vertical velocity vz [ncol][nd] [cm/s] and vmicro [cm/s] per column."""
    rng = np.random.default_rng(seed)
    st0 = template["state"]
    ncol = nx * ny
    dT = gaussian_random_field(nx, ny, 150.0, 8.0, rng).reshape(ncol)
    logB = rng.uniform(np.log10(10.0), np.log10(3000.0), ncol)
    cosi = rng.uniform(-1.0, 1.0, ncol)
    azi = rng.uniform(0.0, 180.0, ncol)
    v0 = rng.normal(0.0, 2e5, ncol)
    vg = rng.normal(0.0, 0.5e5, ncol)
    vmic = rng.uniform(0.0, 2e5, ncol)
    logtau = template["logtau"]
    T = st0[abi.F_T][None, :] + dT[:, None]
    return dict(T=np.ascontiguousarray(T), B=10.0 ** logB, inc=np.arccos(cosi), azi=np.radians(azi),
                vz=np.ascontiguousarray(v0[:, None] + vg[:, None] * logtau[None, :]), vmic=vmic, logtau=logtau)


def model_cube(draws, nx, ny, Nt, ne, opref, rho, xp=np):
    """The 12 x nx x ny x nd model cube (log tau, height, T, Pg, Pe, rho,
    kappa_ref, |B|, vmicro, vz, inclination, azimuth: the parameters
    atmosphere::read_spinor3d reads, atmos/atmos_fio.cc:79-127) of the drawn
    columns with their hydrostatic densities Nt, ne, kappa_ref and rho
    [ncol][nd].  xp: numpy, or torch for device tensors (then every input
    array is a tensor on one device and the cube is float64 there)."""
    T = draws["T"]
    nd = T.shape[1]
    ncol = nx * ny
    if xp is np:
        c = np.zeros((12, ncol, nd))
        arr = lambda a: np.asarray(a, dtype=np.float64)  # noqa: E731
    else:
        import torch
        dev = T.device
        c = torch.zeros((12, ncol, nd), dtype=torch.float64, device=dev)
        arr = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev)  # noqa: E731
    c[0] = arr(draws["logtau"])[None, :]
    c[2] = T
    c[3] = Nt * K_BOLTZ * T
    c[4] = ne * K_BOLTZ * T
    c[5] = rho
    c[6] = opref
    c[7] = arr(draws["B"])[:, None]
    c[8] = arr(draws["vmic"])[:, None]
    c[9] = arr(draws["vz"])
    c[10] = arr(draws["inc"])[:, None]
    c[11] = arr(draws["azi"])[:, None]
    return c.reshape(12, nx, ny, nd)


def cube_state_eos_dev(template, nx, ny, seed=1, stream=None):
    """The cube workload with its thermodynamics solved on the GPU, as the
    reference builds a model atmosphere: the drawn temperatures on the
    template's log tau grid through hydrostatic equilibrium
    (snapi_hequilibrium_dev, atmos_pp::enforce_hequilibrium) for Nt, ne,
    kappa_ref and rho; the 12 x nx x ny x nd model cube in float32; the cube
    to column states for the template's ray (snapi_cube_to_state_dev,
    read_spinor3d with the ray projection); then the equation of state
    (snapi_eos_state_dev: Ne, neutral H, referent opacity, Thomson and
    bound-free background, LTE level populations of the lines, as popsetup
    and compute_op_referent).  Returns (state [ncol][nfield][nd] float64 and
    the float32 model cube [12][nx][ny][nd], both torch device tensors)."""
    import torch
    from . import chemeq, cube
    ray = template["ray"]
    dr = cube_draws(template, nx, ny, seed)
    chem, chem_he, smap = eos_tables("cfg1_template", ray.lam_mid)
    dev = torch.device("cuda")
    ncol, nd = dr["T"].shape
    T = torch.from_numpy(dr["T"]).to(dev)
    lt = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(dr["logtau"], (ncol, nd)))).to(dev)
    Nt, ne, kr, rho = (torch.empty((ncol, nd), dtype=torch.float64, device=dev) for _ in range(4))
    chemeq.hequilibrium_dev(chem_he, T, lt, Nt, ne, kr, rho, stream=stream)
    dd = dict(dr, T=T)
    c32 = model_cube(dd, nx, ny, Nt, ne, kr, rho, xp=torch).float()
    nf = abi.nfield(template["table"].nline)
    st = torch.zeros((ncol, nf, nd), dtype=torch.float64, device=dev)
    cube.cube_to_state_dev(c32.double(), nf, st, theta=ray.theta, phi=0.0, stream=stream)
    chemeq.eos_state_dev(chem, smap, st, stream=stream)
    return st, c32


# the non-thermodynamic node set of the 14-node inversion workload (7 T
# nodes, 2 vLOS, 1 vmicro, 2 B, 1 inclination, 1 azimuth), in log tau_500
INVERSION_NODES = [("vt", -2.0), ("vlos", -4.0), ("vlos", 0.0), ("B", -4.0), ("B", 0.0), ("theta", -2.0),
                   ("phi", -2.0)]


def inversion_problem(template, nx, ny, seed=3):
    """Pixels of the batched-inversion workload: cube columns (as cube_state),
    their node values (truth), starting values perturbed by up to +/-20 %,
    and per-pixel wavelength masks keeping a random 120..nlambda subset.
    Returns dict(state [npix][nfield][nd] float64, truth, start [npix][7],
    mask [npix][nlambda] float64)."""
    st = cube_state(template, nx, ny, seed=seed, dtype=np.float64)
    npix = st.shape[0]
    rng = np.random.default_rng(seed + 1000)
    logtau = template["logtau"]
    truth = np.empty((npix, len(INVERSION_NODES)))
    for k, (p, lt) in enumerate(INVERSION_NODES):
        z = int(np.argmin(np.abs(logtau - lt)))
        if p == "vt":
            truth[:, k] = st[:, abi.F_VT, z]
        elif p == "vlos":
            truth[:, k] = st[:, abi.F_VLOS, z]
        elif p == "B":
            truth[:, k] = st[:, abi.F_BMAG, z]
        elif p == "theta":
            truth[:, k] = np.radians(st[:, abi.F_BINC, z])
        else:
            truth[:, k] = np.radians(st[:, abi.F_BAZI, z])
    start = truth * rng.uniform(0.8, 1.2, truth.shape)
    nl = len(template["lam_vac"])
    mask = np.zeros((npix, nl))
    keep = rng.integers(120, nl + 1, npix)
    order = np.argsort(rng.random((npix, nl)), axis=1)
    for p in range(npix):
        mask[p, order[p, :keep[p]]] = 1.0
    return dict(state=st, truth=truth, start=start, mask=mask, nodes=INVERSION_NODES)


# the deepest temperature node at log tau 0.5, where the cube's columns stay
# below the reference's 12000 K bracket (atmos_model.cc:12-19)
TNODE_INVERSION_NODES = [("T", -5.0), ("T", -4.0), ("T", -3.0), ("T", -2.0), ("T", -1.0), ("T", 0.0), ("T", 0.5)] \
    + INVERSION_NODES


def inversion_problem_tnodes(template, nx, ny, seed=3):
    """The inversion workload with the full 14-node model (7 temperature
    nodes at log tau -5..0.5, then the 7 nodes of inversion_problem): truth
    temperatures are the cube columns' temperatures at the node positions,
    the other nodes as inversion_problem draws them; starting values
    perturbed by up to +/-20 % (temperatures included); the same wavelength
    masks.  The pixels' states are built from the truth nodes by the caller
    (snapi_build_from_nodes).  Returns dict(base [npix][nfield][nd], truth
    [npix][14], start, mask, nodes)."""
    p = inversion_problem(template, nx, ny, seed=seed)
    st = p["state"]
    npix = st.shape[0]
    logtau = template["logtau"]
    rng = np.random.default_rng(seed + 2000)
    tt = np.empty((npix, 7))
    for k, (_, lt) in enumerate(TNODE_INVERSION_NODES[:7]):
        z = int(np.argmin(np.abs(logtau - lt)))
        tt[:, k] = st[:, abi.F_T, z]
    truth = np.concatenate([tt, p["truth"]], axis=1)
    start = np.concatenate([tt * rng.uniform(0.8, 1.2, tt.shape), p["start"]], axis=1)
    start[:, :7] = np.clip(start[:, :7], 3400.0, 12000.0)
    return dict(base=st, truth=truth, start=start, mask=p["mask"], nodes=TNODE_INVERSION_NODES)


def ali_template():
    """The NLTE column recorded from the reference (tests/golden/nlte.npz):
    a Ca II 5-level + continuum atom (the reference's calcium_ir_rh.cfg) in
    the 61-depth photosphere, with every ALI input as the reference computes
    it (frequency grid, rays, background, damping, collision rates, LTE start
    populations)."""
    d = np.load(os.path.join(GOLDEN, "nlte.npz"))
    trans = []
    for row in d["trans"]:
        kind, lo, up = int(row[0]), int(row[1]), int(row[2])
        t = dict(kind=kind, lower=lo, upper=up)
        if kind == abi.TR_LINE:
            t.update(lam0=row[3], aul=row[4], blu=row[5], bul=row[6], energy=row[7])
        else:
            t.update(gratio=row[8], de=row[9])
        trans.append(t)
    dep = d["depth"]
    cs = np.zeros((abi.A_NF, int(d["nd"])))
    cs[abi.A_T], cs[abi.A_VT], cs[abi.A_NE] = dep[:, 0], dep[:, 1], dep[:, 2]
    cs[abi.A_OPREF], cs[abi.A_GRID], cs[abi.A_NA] = dep[:, 3], dep[:, 4], dep[:, 5]
    return dict(nlevel=int(d["nmap"]), mass=float(d["mass"]), trans=trans, lam=d["lam"], rays=d["rays"],
                sigma=d["sigma"], colstate=cs, vproj=d["vproj"], gamma=d["gamma"], bgop=d["bgop"],
                bgem=d["bgem"], coll=d["coll"], pop0=d["pop0"], boundary=float(d["boundary"]))


def gauss_radau_mu(n):
    """n-point Gauss-Radau nodes and weights on mu in (0, 1], mu = 1 included."""
    from numpy.polynomial import legendre as leg
    # interior nodes: roots of (P_{n-1}(x) + P_n(x)) / (1 + x) on [-1, 1], x = -1 fixed
    c = np.zeros(n + 1)
    c[n - 1] = 1.0
    c[n] = 1.0
    r = np.sort(np.real(leg.legroots(c)))
    x = np.concatenate([[-1.0], r[np.abs(r + 1.0) > 1e-9]])[:n]
    pn1 = leg.legval(x, np.eye(n)[n - 1])
    w = (1.0 - x) / (n * n * pn1 * pn1)
    w[0] = 2.0 / (n * n)
    return (1.0 - x) / 2.0, w / 2.0  # mapped to mu in (0, 1], weights summing to 1


def ali_problem(nx, ny, nd=81, nlam=220, nmu=5, seed=9):
    """The ALI workload: nx * ny plane-parallel columns of nd depths drawn
    around the reference's NLTE column, resampled in log tau, with a
    chromospheric temperature rise towards 8000 K above log tau = -4 and
    seeded +/-10 % perturbations of temperature, electron density and
    background opacity; nlam wavelengths taken from the reference's NLTE grid;
    nmu Gauss-Radau directions in each hemisphere (2 nmu rays).  The
    emissivity follows the temperature through the Planck function; collision
    rates and start populations are the template's at each depth.  Returns
    dict(atom, grid, colstate, vproj, gamma, bgop, bgem, coll, pops) with one
    background set for all rays."""
    t = ali_template()
    rng = np.random.default_rng(seed)
    nd0 = t["colstate"].shape[1]
    x0 = np.log10(-t["colstate"][abi.A_GRID])
    x = np.linspace(x0[0], x0[-1], nd)

    def rs(a, log=True):  # resample the last axis from the template depths
        a = np.asarray(a, dtype=np.float64)
        flat = a.reshape(-1, nd0)
        out = np.empty((flat.shape[0], nd))
        for i, row in enumerate(flat):
            if log and np.all(row > 0):
                out[i] = 10.0 ** np.interp(x, x0, np.log10(row))
            else:
                out[i] = np.interp(x, x0, row)
        return out.reshape(a.shape[:-1] + (nd,))
    li = np.unique(np.round(np.linspace(0, len(t["lam"]) - 1, nlam)).astype(int))
    lam = t["lam"][li, 0]
    lamw = np.empty_like(lam)
    lamw[0] = (lam[1] - lam[0]) * 0.5
    lamw[-1] = (lam[-1] - lam[-2]) * 0.5
    lamw[1:-1] = (lam[2:] - lam[:-2]) * 0.5
    mu, wmu = gauss_radau_mu(nmu)
    theta = np.concatenate([np.arccos(mu), np.pi - np.arccos(mu)])
    wray = np.concatenate([wmu, wmu]) * 2.0 * np.pi
    ncol = nx * ny
    cs0 = rs(t["colstate"])
    cs0[abi.A_GRID] = -10.0 ** x
    T0 = cs0[abi.A_T]
    rise = np.clip((-4.0 - x) / 2.0, 0.0, 1.0)  # 0 below log tau = -4, 1 above -6
    Tch = T0 + rise * np.maximum(8000.0 - T0, 0.0)
    cs = np.repeat(cs0[None], ncol, 0)
    fT = rng.uniform(0.9, 1.1, (ncol, 1))
    cs[:, abi.A_T] = Tch[None] * fT
    cs[:, abi.A_NE] *= rng.uniform(0.9, 1.1, (ncol, 1))
    bg0 = rs(t["bgop"][0].T).T[:, li]  # [nd][nlam] (the template's rays share the background at rest)
    be0 = rs(np.maximum(t["bgem"][0].T, 1e-300)).T[:, li]
    bgop = np.repeat(bg0[None, None], ncol, 0) * rng.uniform(0.9, 1.1, (ncol, 1, 1, 1))
    # emissivity = opacity x Planck at the column's temperature, scaled from the template's
    # B(lam, T) / B(lam, T0) = (e^a0 - 1) / (e^a - 1), a = h c / (lam k T), in a form
    # that neither overflows nor divides 0 by 0 in the far UV
    a = 1.4387769 / lam[None, None, :] / cs[:, abi.A_T][:, :, None]
    a0 = 1.4387769 / lam[None, None, :] / cs0[abi.A_T][None, :, None]
    ratio = np.exp(a0 - a) * (-np.expm1(-a0)) / (-np.expm1(-a))
    bgem = np.ascontiguousarray(np.repeat(be0[None, None], ncol, 0) * ratio[:, None] * (bgop / bg0[None, None]))
    trans = t["trans"]
    atom = abi.make_ali_atom(t["nlevel"], t["mass"], trans)
    sig = t["sigma"][:, li]
    grid = abi.make_ali_grid(lam, lamw, np.ones_like(lam), sig, theta, wray, 0, t["boundary"], 1)
    vproj = np.zeros((ncol, len(theta), nd))
    gamma = np.repeat(rs(t["gamma"])[None], ncol, 0)
    # collision rates at the template's temperature and electron density,
    # carried to each column's: excitation and ionisation i -> j scale as
    # Ne T^-1/2 exp(-dE/kT) (van Regemorter, Seaton), de-excitation as
    # Ne T^-1/2, three-body recombination as Ne^2 T^-2 (detailed balance with
    # the Saha-Boltzmann ratio), so every pair keeps detailed balance at the
    # column's temperature
    coll0 = np.moveaxis(rs(np.moveaxis(t["coll"], 0, -1), log=False), -1, 0)  # [nd][n][n]
    T0, N0 = cs0[abi.A_T], cs0[abi.A_NE]
    Tc, Nc = cs[:, abi.A_T], cs[:, abi.A_NE]  # [ncol][nd]
    n = t["nlevel"]
    de = np.zeros(n)  # binding energy of each level below the continuum [erg]
    cont = np.zeros(n, dtype=bool)
    for tr in trans:
        if tr["kind"] == abi.TR_BF:
            de[tr["lower"]] = tr["de"]
            cont[tr["upper"]] = True
    kb = 1.3806503e-16
    fT = np.sqrt(T0[None] / Tc)
    fN = Nc / N0[None]
    coll = np.empty((ncol, nd, n, n))
    for i in range(n):
        for j in range(n):
            if i == j:
                coll[:, :, i, j] = 0.0
                continue
            if cont[j] and not cont[i]:  # ionisation
                f = fN * fT * np.exp(-de[i] / kb * (1.0 / Tc - 1.0 / T0[None]))
            elif cont[i] and not cont[j]:  # three-body recombination
                f = fN * fN * fT ** 4
            elif j > i:  # excitation over de[i] - de[j]
                f = fN * fT * np.exp(-(de[i] - de[j]) / kb * (1.0 / Tc - 1.0 / T0[None]))
            else:  # de-excitation
                f = fN * fT
            coll[:, :, i, j] = coll0[None, :, i, j] * f
    coll = np.ascontiguousarray(coll)
    # start populations: LTE (Saha-Boltzmann) at each column's temperature and
    # electron density with the atom's total population, as nltepops starts
    # from ltepops
    ratio = np.zeros((ncol, nd, n))
    for tr in trans:
        if tr["kind"] == abi.TR_BF:
            ratio[:, :, tr["lower"]] = Nc * tr["gratio"] * 2.0706615382e-16 * Tc ** -1.5 * np.exp(tr["de"] / (kb * Tc))
    ratio[:, :, cont] = 1.0
    pops = np.ascontiguousarray(ratio * (cs[:, abi.A_NA] / ratio.sum(axis=2))[:, :, None])
    return dict(atom=atom, grid=grid, colstate=np.ascontiguousarray(cs), vproj=vproj,
                gamma=np.ascontiguousarray(gamma), bgop=np.ascontiguousarray(bgop), bgem=bgem, coll=coll,
                pops=pops, nlam=len(lam), nray=len(theta))


CA_ABUNDANCE = 2.19e-6  # Ca / H by number (log eps = 6.34)


def ca8542_synthesis(p, pops, seed=12, nlam=201, half_width=1.0e-8):
    """Polarized synthesis inputs of the Ca II 8542 A line (3d 2D5/2 - 4p
    2P3/2) in the ALI workload's columns with their converged populations
    pops [ncol][nd][nlevel]: the line built from the atom's levels
    (energies from the bound-free thresholds, g and 2J from the statistical
    weights, LS Lande factors 1.2 and 4/3, the atom's A values;
    snapi_line_add), the columns' temperature, microturbulence, electron
    density, referent opacity and grid, the population-independent
    background of the ALI grid point nearest the line centre (it holds H-,
    so the synthesis adds none), a seeded field with |B| uniform on
    0-500 G, cos(inclination) uniform on [-1, 1] and azimuth on [0, 180)
    deg, no velocity, neutral hydrogen taken as the total hydrogen density
    NA / A(Ca), and nlam vacuum wavelengths over +-half_width [cm].
    Returns (table, ray, state [ncol][nfield][nd], lam_vac)."""
    from . import lines as L
    atom_tr = []
    de = {}
    for k in range(p["atom"].ntrans):
        tr = p["atom"].tr[k]
        atom_tr.append(tr)
        if tr.kind == abi.TR_BF:
            de[tr.lower] = (tr.de, tr.gratio)
    n = p["atom"].nlevel
    e0 = de[0][0]
    energy = {i: e0 - de[i][0] for i in de}
    A = np.zeros((n, n))
    line = None
    for tr in atom_tr:
        if tr.kind == abi.TR_LINE:
            A[tr.upper, tr.lower] = tr.aul
            if abs(tr.lam0 - 8544.4e-8) < 2e-8:
                line = tr
    if line is None:
        raise ValueError("the atom has no line at 8542 A")
    lo, up = line.lower, line.upper
    glo, gup = de[lo][1], de[up][1]
    lower = L.level(energy[lo], glo, int(round(glo)) - 1, 2, 1.2)
    upper = L.level(energy[up], gup, int(round(gup)) - 1, 1, 4.0 / 3.0)
    t = L.empty_table()
    L.add_line(t, p["atom"].mass, 1, e0, lower, upper, line.aul, L.radiative_damping(A, lo, up))
    lam0 = t.line[0].lam0
    lam = lam0 + np.linspace(-half_width, half_width, nlam)
    cs = p["colstate"]
    ncol, _, nd = cs.shape
    rng = np.random.default_rng(seed)
    nf = abi.nfield(1)
    st = np.zeros((ncol, nf, nd))
    st[:, abi.F_T] = cs[:, abi.A_T]
    st[:, abi.F_VT] = cs[:, abi.A_VT]
    st[:, abi.F_NE] = cs[:, abi.A_NE]
    st[:, abi.F_OPREF] = cs[:, abi.A_OPREF]
    st[:, abi.F_GRID] = cs[:, abi.A_GRID]
    st[:, abi.F_NT] = cs[:, abi.A_NA] / CA_ABUNDANCE
    st[:, abi.F_NHI] = cs[:, abi.A_NA] / CA_ABUNDANCE
    st[:, abi.F_BMAG] = rng.uniform(0.0, 500.0, (ncol, 1))
    st[:, abi.F_BINC] = np.degrees(np.arccos(rng.uniform(-1.0, 1.0, (ncol, 1))))
    st[:, abi.F_BAZI] = rng.uniform(0.0, 180.0, (ncol, 1))
    glam = np.array([p["grid"].lam[i] for i in range(p["grid"].nlam)])
    il = int(np.argmin(np.abs(glam - lam0)))
    st[:, abi.F_BGOP] = p["bgop"][:, 0, :, il]
    st[:, abi.F_BGEM] = p["bgem"][:, 0, :, il]
    st[:, abi.NF_BASE] = pops[:, :, lo]
    st[:, abi.NF_BASE + 1] = pops[:, :, up]
    ray = abi.SnapiRay(theta=np.pi, lam_mid=0.5 * (lam[0] + lam[-1]), tau_grid=1, hminus=0, hminus_mid=0)
    return t, ray, np.ascontiguousarray(st), lam


def response_column_from_dump(path):
    """Column state as the reference's response path sees it, plus the
    thermodynamic derivatives dstate[2][nfield][nd] (d/dT, d/dNt of Ne,
    neutral H, referent opacity and line-level populations)."""
    g = column_from_dump(path)
    d = np.load(path)
    st = g["state"]
    nf, nd = st.shape
    ds = np.zeros((2, nf, nd))
    th = d["thermo"]  # z Nt [dNe dNH dkref] x 2
    for row in th:
        z = int(row[0]) - 1
        for p in range(2):
            ds[p, abi.F_NE, z] = row[2 + 3 * p]
            ds[p, abi.F_NHI, z] = row[3 + 3 * p]
            ds[p, abi.F_OPREF, z] = row[4 + 3 * p]
    dp = d["dpop"]  # s zi lo up z [dlo dup] x 2
    for i, ln in enumerate(g["lines"]):
        rows = dp[dp[:, 0] == ln["species"]]
        # match this line's lower/upper levels through the population dump order
        s_lines = [l for l in g["lines"] if l["species"] == ln["species"]]
        k = s_lines.index(ln)
        keys = sorted(set((int(r[1]), int(r[2]), int(r[3])) for r in rows), key=lambda t: (t[0], t[2], t[1]))
        zi, lo, up = keys[k]
        sel = rows[(rows[:, 1] == zi) & (rows[:, 2] == lo) & (rows[:, 3] == up)]
        sel = sel[np.argsort(sel[:, 4])]
        for p in range(2):
            ds[p, abi.NF_BASE + 2 * i] = sel[:, 5 + 2 * p]
            ds[p, abi.NF_BASE + 2 * i + 1] = sel[:, 6 + 2 * p]
    g["dstate"] = ds
    g["resp"] = d["resp"]
    return g


def template_cfg1():
    return column_from_dump(os.path.join(GOLDEN, "cfg1_template.npz"))


def cfg0_column():
    return column_from_dump(os.path.join(GOLDEN, "cfg0_synth.npz"))


def chemeq_problem(nx, ny, seed=5):
    """Depth points of the thermodynamics workload: the species table recorded
    from the reference (eight background atoms, the two-line Fe atom, 13
    molecules, tests/golden/chemeq.npz) and nx*ny columns of the cfg0 column's
    61 depths (T, Nt after the reference's hydrostatic integration), with a
    Gaussian random field of 150 K rms added to T and a per-column 0.1 dex rms
    scaling of Nt.  Returns dict(table, T [ncol][nd], Nt [ncol][nd])."""
    d = np.load(os.path.join(GOLDEN, "chemeq.npz"))
    table = abi.make_chem_table(d)
    n0, n1 = (int(x) for x in d["ncolumn_points"])
    T0, Nt0 = d["T"][n0:n0 + n1], d["Nt"][n0:n0 + n1]
    rng = np.random.default_rng(seed)
    ncol = nx * ny
    dT = gaussian_random_field(nx, ny, 150.0, 8.0, rng).reshape(ncol)
    dn = rng.normal(0.0, 0.1, ncol)
    T = np.ascontiguousarray(T0[None, :] + dT[:, None])
    Nt = np.ascontiguousarray(Nt0[None, :] * 10.0 ** dn[:, None])
    return dict(table=table, T=T, Nt=Nt)


def hequilibrium_problem(nx, ny, seed=6):
    """Columns for hydrostatic equilibrium: the species table recorded from the
    reference with the opacity taken at 5000 A as enforce_hequilibrium does,
    and nx*ny columns of the cfg0 column's temperatures on its 61-point
    log tau grid (-6 .. 1), with a Gaussian random field of 150 K rms per
    column plus 30 K rms independent noise per depth.  Returns dict(table,
    T [ncol][nd], logtau [ncol][nd])."""
    d = np.load(os.path.join(GOLDEN, "chemeq.npz"))
    table = abi.make_chem_table(d, referent=1)
    he = d["he"]
    rng = np.random.default_rng(seed)
    ncol, nd = nx * ny, he.shape[0]
    dT = gaussian_random_field(nx, ny, 150.0, 8.0, rng).reshape(ncol)
    T = np.ascontiguousarray(he[None, :, 0] + dT[:, None] + rng.normal(0.0, 30.0, (ncol, nd)))
    logtau = np.ascontiguousarray(np.broadcast_to(he[None, :, 1], (ncol, nd)))
    return dict(table=table, T=T, logtau=logtau)


def eos_tables(dump_name, lam_mid):
    """The species tables and state map that rebuild the thermodynamics of a
    column recorded in tests/golden/<dump_name>.npz: the table at
    airtovac(5000 A) (synthesis and response paths), the one at 5000 A
    (hydrostatic equilibrium), and the state map of the recorded lines and
    bound-free background.  Returns (chem, chem_he, smap)."""
    d = np.load(os.path.join(GOLDEN, "chemeq.npz"))
    g = np.load(os.path.join(GOLDEN, dump_name + ".npz"))
    chem = abi.make_chem_table(d, referent=0)
    chem_he = abi.make_chem_table(d, referent=1)
    smap = abi.make_state_map(chem, d, g["lines"], g["bfsig"], lam_mid)
    return chem, chem_he, smap


def nlte_system_case():
    """The NLTE response case recorded from the reference
    (tests/golden/nlterin.npz: the Ca II atom of the cfg0 column after
    nltepops, with everything compute_nlte_population_responses feeds its
    assembly) in the layouts of snapi_nlte_response_system: (atom, case) with
    case holding the quadrature, the wavelength grid, sigma and one column's
    depth, pops, damp, coll, vr and opem, plus the reference's own M, beta,
    profile norms and profile samples."""
    d = dict(np.load(os.path.join(GOLDEN, "nlterin.npz")))
    nd, nray, nlam, nmap, ntr, Z, tau_grid, damping = (int(x) for x in d["head"])
    mass, boundary = d["consts"][0], d["consts"][4]
    lev, tr = d["levels"], d["trans"]
    stage = lev[:, 1].astype(int)
    stage_ip = np.zeros(Z + 1)
    for i in range(nmap):
        stage_ip[stage[i]] = lev[i, 5]
    # transitions: index, kind (0 line, 1 bound-free), stage, lower and upper
    # within the stage, map index of the upper level, A, Blu, Bul, edge
    first = {z: int(np.argmax(stage == z)) for z in set(stage)}
    lower = np.array([first[int(t[2])] + int(t[3]) for t in tr])
    upper = np.array([int(t[5]) for t in tr])
    atom = abi.make_nlte_atom(stage, lev[:, 3], lev[:, 4], stage_ip, tr[:, 1].astype(int), lower, upper, tr[:, 6],
                              tr[:, 7], tr[:, 8], d["damp"][0, :, 0], tr[:, 9], mass, damping)
    case = dict(nd=nd, nray=nray, nlam=nlam, nmap=nmap, ntr=ntr, tau_grid=tau_grid, boundary=float(boundary),
                theta=d["rays"][:, 0].copy(), wray=d["rays"][:, 1].copy(), lam=d["lam"][:, 0].copy(),
                lamw=d["lam"][:, 1].copy(), sigma=d["sigma"].copy(),
                depth=np.ascontiguousarray(d["depth"][:, :abi.SNAPI_NLTE_NFIELD].T),
                pops=np.ascontiguousarray(d["pops"].T), damp=np.ascontiguousarray(d["damp"][1:]),
                coll=np.ascontiguousarray(d["coll"].transpose(3, 1, 2, 0)), vr=d["vr"].copy(),
                opem=np.ascontiguousarray(d["pp"].transpose(0, 1, 3, 2)),
                M=d["M"], beta=d["beta"], norm=d["norm"], prof=d["prof"])
    return atom, case


def nlte_system_problem(ncol, seed=14):
    """ncol columns around the recorded Ca II case: opacity and emissivity of
    every (ray, wavelength, depth) scaled by exp(N(0, 0.1)), the emissivity
    again by exp(N(0, 0.05)), level populations by exp(N(0, 0.05)); the
    per-depth state, damping and collision rates are the case's.  Returns
    (atom, case, inputs) with inputs the per-column arrays (leading column
    axis) and sigma."""
    atom, c = nlte_system_case()
    rng = np.random.default_rng(seed)
    rep = lambda a: np.repeat(a[None], ncol, axis=0)  # noqa: E731
    opem = c["opem"][None] * np.exp(rng.normal(0.0, 0.1, (ncol,) + c["opem"].shape[:2] + (1, c["nd"])))
    opem[:, :, :, 1] *= np.exp(rng.normal(0.0, 0.05, (ncol,) + c["opem"].shape[:2] + (c["nd"],)))
    pops = c["pops"][None] * np.exp(rng.normal(0.0, 0.05, (ncol,) + c["pops"].shape))
    inputs = dict(sigma=c["sigma"], depth=rep(c["depth"]), pops=pops, damp=rep(c["damp"]), coll=rep(c["coll"]),
                  vr=rep(c["vr"]), opem=np.ascontiguousarray(opem))
    return atom, c, inputs
