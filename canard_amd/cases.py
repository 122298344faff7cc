"""Grids, initial conditions and restart files for the solver's cases.

Host-side set-up only (numpy).  Conventions follow the reference: a block of
lxi x let x lze cells stores (lxi+1)(let+1)(lze+1) points, xi fastest;
fields are (ncomp, npoints) arrays.
"""
import struct

import numpy as np

GAM = 1.4

# Boundary codes, src/mo_parameters.f90:43-49
NONREF, WALL_INV, WALL_VIS, INTER_CURV, INTER_STRAIGHT, INTER_SUB, PERIODIC = 10, 20, 25, 30, 35, 40, 45


def periodic_block(n, inst=0, lxi=None, let=None, lze=None, me=0):
    """One triply periodic block of n cells per direction (its own partner)."""
    return dict(lxi=lxi or n, let=let or n, lze=lze or n, nbc=[[PERIODIC, PERIODIC]] * 3,
                mcd=[[me, me]] * 3, inst=inst)


def gridtoy(lxi, let, lze, doml0=0.0, doml1=2 * np.pi, domh=np.pi, span=2 * np.pi):
    """Uniform single-block grid of gridtoy (src/mo_gridgen.f90:510-567):
    x = -doml0 + i (doml1+doml0)/lxi, y = -domh + j 2 domh/let,
    z = span ((lze-k)/lze - 1/2) (z decreases with k)."""
    i = np.arange(lxi + 1, dtype=np.float64)
    j = np.arange(let + 1, dtype=np.float64)
    k = np.arange(lze + 1, dtype=np.float64)
    ra0 = (doml1 + doml0) / lxi
    ra1 = 2.0 * domh / let
    x = -doml0 + i * ra0
    y = -domh + j * ra1
    z = span * ((lze - k) / lze - 0.5)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()])


def gridtoy_multi(mb, lxi0, lxi1, lxi2, let0, lze0, doml0=0.0, doml1=2 * np.pi, domh=np.pi, span=2 * np.pi):
    """Block mb (0..5) of gridtoy_multi (src/mo_gridgen.f90:569-654): a 3 x 2
    arrangement of blocks with lxi0 | lxi1 | lxi2 cells along x and let0
    cells along y each, points spaced (doml1 + doml0)/(lxi0 + lxi1 + lxi2)
    in x and domh/let0 in y; z as gridtoy."""
    lxi = (lxi0, lxi1, lxi2)[mb % 3]
    col = mb % 3
    i0 = (0, lxi0 + 1, lxi0 + lxi1 + 2)[col]
    j0 = 0 if mb < 3 else let0 + 1
    i = np.arange(lxi + 1, dtype=np.float64) + i0
    j = np.arange(let0 + 1, dtype=np.float64) + j0
    k = np.arange(lze0 + 1, dtype=np.float64)
    ra0 = (doml1 + doml0) / (lxi0 + lxi1 + lxi2)
    ra1 = domh / (2.0 * let0) * 2.0
    x = -doml0 + (i - col) * ra0
    y = -domh + (j - (0 if mb < 3 else 1)) * ra1
    z = span * ((lze0 - k) / lze0 - 0.5)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()])


def splitmix64(x):
    """SplitMix64 finaliser on uint64 numpy arrays (counter-based RNG)."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return z ^ (z >> np.uint64(31))


def uniform01(seed, idx):
    """U[0,1) keyed by (seed, idx): idx any non-negative integer array."""
    with np.errstate(over="ignore"):
        key = np.uint64(seed) * np.uint64(0xD1B54A32D192ED03) + idx.astype(np.uint64)
        h = splitmix64(key.astype(np.uint64))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def tgv_state(xyz, shape, mach=0.1, pert=0.0, seed=20261202, modes=0.0, mode_seed=0):
    """Taylor-Green vortex: u = M sin x cos y cos z, v = -M cos x sin y cos z,
    w = 0, p = 1/gam + M^2/16 (cos 2x + cos 2y)(cos 2z + 2), rho = gam p
    (T = gam p / rho = 1).  pert > 0 adds a seeded uniform velocity
    perturbation of amplitude pert*M, keyed by the periodic point index
    (i mod n, j mod n, k mod n) so duplicated periodic points agree; modes >
    0 adds resolved-scale Fourier modes (resolved_modes) of relative
    amplitude modes drawn from mode_seed."""
    x, y, z = xyz
    u = mach * np.sin(x) * np.cos(y) * np.cos(z)
    v = -mach * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    if pert > 0.0:
        nx, ny, nz = shape   # points per direction
        kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        ii, jj, kk = ii.ravel() % (nx - 1), jj.ravel() % (ny - 1), kk.ravel() % (nz - 1)
        base = ((kk.astype(np.int64) * (ny - 1) + jj) * (nx - 1) + ii) * 3
        for c, arr in enumerate((u, v, w)):
            arr += pert * mach * (2.0 * uniform01(seed, base + c) - 1.0)
    if modes > 0.0:
        du = resolved_modes(xyz, mach, modes, mode_seed)
        u += du[0]
        v += du[1]
        w += du[2]
    p = 1.0 / GAM + (mach * mach / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    rho = GAM * p
    e = p / (GAM - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, e])


def resolved_modes(xyz, mach, amp, seed, nmodes=8, kmax=3):
    """Velocity perturbation at resolved scales on the [0, 2 pi]-periodic
    box: per component, nmodes Fourier modes sin(k . x + phi) with integer
    wavevectors |k_i| <= kmax and random phases, each of amplitude
    amp * mach / sqrt(nmodes), all drawn from seed.  Returns (3, points)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((3, xyz.shape[1]))
    for c in range(3):
        for _ in range(nmodes):
            k = rng.integers(-kmax, kmax + 1, size=3)
            if not k.any():
                k[c] = 1
            phi = rng.uniform(0.0, 2 * np.pi)
            out[c] += amp * mach / np.sqrt(nmodes) * np.sin(k[0] * xyz[0] + k[1] * xyz[1] + k[2] * xyz[2] + phi)
    return out


def uniform_state(npt, u=(0.0, 0.0, 0.0), rho=1.0, p=1.0 / GAM):
    """Free stream (mo_sponge.f90:83-85 non-dimensionalisation: rho=1, c=1)."""
    r = np.full(npt, rho)
    e = p / (GAM - 1.0) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    return np.stack([r, r * u[0], r * u[1], r * u[2], np.full(npt, e)])


# ---- restart files (src/mo_io.f90:184-271) -----------------------------------
# Direct-access unformatted FP64, record length 5 doubles: record 1 =
# (n, ndt, dt, dts, dte), record 2 = (timo, 0, 0, 0, 0), then one record
# per point of the block in i-fastest order.

def write_restart(path, q, clock):
    q = np.asarray(q, dtype=np.float64)
    with open(path, "wb") as f:
        f.write(struct.pack("<5d", float(clock["n"]), float(clock["ndt"]), clock["dt"], clock["dts"], clock["dte"]))
        f.write(struct.pack("<5d", clock["timo"], 0.0, 0.0, 0.0, 0.0))
        f.write(np.ascontiguousarray(q.T).astype("<f8").tobytes())


def read_restart(path, npt):
    with open(path, "rb") as f:
        h = struct.unpack("<5d", f.read(40))
        t = struct.unpack("<5d", f.read(40))
        data = np.frombuffer(f.read(40 * npt), dtype="<f8")
    if data.size != 5 * npt:
        raise ValueError(f"{path}: truncated restart file ({data.size // 5} of {npt} points)")
    clock = dict(n=int(h[0]), ndt=int(h[1]), dt=h[2], dts=h[3], dte=h[4], timo=t[0])
    return data.reshape(npt, 5).T.copy(), clock


# ---- multi-block and ensemble cases -------------------------------------------

def sheared_pair_blocks(sizes, amp=0.05, inst=0, first=0):
    """Two blocks joined on their xi faces (BC 35), non-reflective outer x/y
    faces (BC 10), periodic z (BC 45).  sizes = [(lxi, let, lze), (lxi, let, lze)]
    with equal let, lze.  Returns block dicts (partners offset by first)."""
    (a, b) = sizes
    assert a[1] == b[1] and a[2] == b[2]
    blk0 = dict(lxi=a[0], let=a[1], lze=a[2], nbc=[[NONREF, INTER_STRAIGHT], [NONREF, NONREF], [PERIODIC, PERIODIC]],
                mcd=[[-1, first + 1], [-1, -1], [first, first]], inst=inst)
    blk1 = dict(lxi=b[0], let=b[1], lze=b[2], nbc=[[INTER_STRAIGHT, NONREF], [NONREF, NONREF], [PERIODIC, PERIODIC]],
                mcd=[[first, -1], [-1, -1], [first + 1, first + 1]], inst=inst)
    return [blk0, blk1]


def sheared_pair_grids(sizes, h=1.0, amp=0.05):
    """Coordinates of a sheared pair: unit spacing h, x = X + amp*Lx*sin(2 pi Y/Ly)
    with Lx the pair's total length and Ly its height; block 1 starts where
    block 0 ends (shared interface points)."""
    (a, b) = sizes
    lx_tot = (a[0] + b[0]) * h
    ly = a[1] * h
    out = []
    x0 = 0.0
    for (lxi, let, lze) in sizes:
        i = np.arange(lxi + 1) * h + x0
        j = np.arange(let + 1) * h
        k = np.arange(lze + 1) * h
        Z, Y, X = np.meshgrid(k, j, i, indexing="ij")
        Xs = X + amp * lx_tot * np.sin(2 * np.pi * Y / ly)
        out.append(np.stack([Xs.ravel(), Y.ravel(), Z.ravel()]))
        x0 += lxi * h
    return out


def pulse_state(xyz, mach=0.5, centres=(), amp=1e-3, width=5.0):
    """Uniform flow at Mach `mach` along x with Gaussian pressure pulses
    p' = amp/gam exp(-r^2/(2 width^2)) and isentropic density."""
    x, y, z = xyz
    p0 = 1.0 / GAM
    dp = np.zeros_like(x)
    for (cx, cy, cz) in centres:
        r2 = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2
        dp += (amp / GAM) * np.exp(-r2 / (2.0 * width * width))
    p = p0 + dp
    rho = (p / p0) ** (1.0 / GAM)
    u = np.full_like(x, mach)
    e = p / (GAM - 1.0) + 0.5 * rho * u * u
    return np.stack([rho, rho * u, np.zeros_like(x), np.zeros_like(x), e])


def ensemble(seed=7, count=32, sizes=(32, 48, 64, 96, 128), mach=(0.05, 0.3), re=(400.0, 3200.0)):
    """Ensemble of independent periodic TGV instances: size n^3 drawn from
    sizes, Mach and Re uniform, all from one seed."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        m = float(rng.uniform(*mach))
        r = float(rng.uniform(*re))
        out.append(dict(n=n, mach=m, re=r))
    return out


# ---- the benchmark configurations -------------------------------------------

def config3(seed=20261202, sizes=((384, 256, 64), (192, 256, 64), (512, 128, 64), (128, 128, 64)), pulses=3):
    """Four uneven sheared blocks as two independent pairs joined on their xi
    faces (BC 35), non-reflective outer x/y faces, periodic z; uniform Mach 0.5
    flow along x with seeded Gaussian pressure pulses (amplitude 1e-3, width 5
    cells) and a sponge (szco 0.5).  Returns (blocks, grids, states, params)."""
    rng = np.random.default_rng(seed)
    blocks, grids, states = [], [], []
    for p, pair in enumerate((sizes[:2], sizes[2:])):
        blocks += sheared_pair_blocks(list(pair), first=2 * p)
        g = sheared_pair_grids(list(pair), h=1.0, amp=0.05)
        lx = pair[0][0] + pair[1][0]
        centres = [(rng.uniform(0.2, 0.8) * lx, rng.uniform(0.2, 0.8) * pair[0][1], rng.uniform(0, pair[0][2]))
                   for _ in range(pulses)]
        grids += g
        states += [pulse_state(x, mach=0.5, centres=centres, amp=1e-3, width=5.0) for x in g]
    lx_max = max(sizes[0][0] + sizes[1][0], sizes[2][0] + sizes[3][0])
    params = dict(reoo=1000.0, amach1=0.5, nsmf=1, cfl=0.8, viscous=1, sponge=1, szco=0.5,
                  szth0=0.1 * lx_max, szth1=0.1 * lx_max, doml0=0.0, doml1=float(lx_max))
    return blocks, grids, states, params


def config4(seed=20261202, count=32, sizes=(32, 48, 64, 96, 128)):
    """Ensemble of periodic Taylor-Green instances: n^3 cells with n drawn
    from sizes, Mach uniform in [0.05, 0.3], Re uniform in [400, 3200], a 1e-3
    velocity perturbation, all by seed.  Returns (blocks, grids, states,
    params list, draws)."""
    rng = np.random.default_rng(seed)
    blocks, grids, states, params, draws = [], [], [], [], []
    for i in range(count):
        n = int(rng.choice(sizes))
        mach = float(rng.uniform(0.05, 0.3))
        re = float(rng.uniform(400.0, 3200.0))
        xyz = gridtoy(n, n, n, doml0=0.0, doml1=2 * np.pi, domh=np.pi, span=2 * np.pi)
        blocks.append(periodic_block(n, inst=i, me=i))
        grids.append(xyz)
        states.append(tgv_state(xyz, (n + 1,) * 3, mach=mach, pert=1e-3, seed=seed + i))
        params.append(dict(fltk=0.49, fltrbc=0.9, reoo=re / mach, tempoo=300.0, cfl=0.8, viscous=1))
        draws.append(dict(n=n, mach=mach, re=re))
    return blocks, grids, states, params, draws
