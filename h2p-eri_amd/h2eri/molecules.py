"""Seeded synthetic molecules, basis sets and densities for the benchmark
workloads, written in the .mol format that CMS_read_mol_file reads
(format: src/CMS.c:14-123 of H2P-ERI; coordinates in bohr).

Workloads (all seeded with numpy.random.Generator(PCG64(20260000 + config))):
  * linear alkane C_nH_{2n+2}: zig-zag backbone, C-C 1.54 A, C-H 1.09 A,
    tetrahedral angles, uniform +-0.02 A jitter on every coordinate;
  * water cluster: O-H 0.96 A, H-O-H 104.5 deg, random orientations,
    centres on a jittered cubic lattice (3.1 A spacing, +-0.1 A jitter).
Basis: Dunning cc-pVDZ, Cartesian d (C 3s2p1d, H 2s1p, O 3s2p1d).  The C and
H blocks are those of H2P-ERI's tests/alk32_ccpvdz.mol; the O block is the
published cc-pVDZ set (T. H. Dunning, J. Chem. Phys. 90, 1007 (1989)).
Density: C ~ N(0,1) (nbf x nocc), Q = qr(C).Q, D0 = 2 Q Q^T,
D_ij = D0_ij exp(-0.5 |R_i - R_j|) with R the atom centre of each basis
function in bohr.
"""
import numpy as np

ANGSTROM = 1.0 / 0.529177210903   # bohr per angstrom

# (shell type, [(exponent, coefficient), ...]) per element
BASIS = {
    "C": [
        ("S", [(6665.0, 0.000692), (1000.0, 0.005329), (228.0, 0.027077), (64.71, 0.101718),
               (21.06, 0.27474), (7.495, 0.448564), (2.797, 0.285074), (0.5215, 0.015204)]),
        ("S", [(6665.0, -0.000146), (1000.0, -0.001154), (228.0, -0.005725), (64.71, -0.023312),
               (21.06, -0.063955), (7.495, -0.149981), (2.797, -0.127262), (0.5215, 0.544529)]),
        ("S", [(0.1596, 1.0)]),
        ("P", [(9.439, 0.038109), (2.002, 0.20948), (0.5456, 0.508557)]),
        ("P", [(0.1517, 1.0)]),
        ("D", [(0.55, 1.0)]),
    ],
    "H": [
        ("S", [(13.01, 0.019685), (1.962, 0.137977), (0.4446, 0.478148)]),
        ("S", [(0.122, 1.0)]),
        ("P", [(0.727, 1.0)]),
    ],
    "O": [
        ("S", [(11720.0, 0.00071), (1759.0, 0.00547), (400.8, 0.027837), (113.7, 0.1048),
               (37.03, 0.283062), (13.27, 0.448719), (5.025, 0.270952), (1.013, 0.015458)]),
        ("S", [(11720.0, -0.00016), (1759.0, -0.001263), (400.8, -0.006267), (113.7, -0.025716),
               (37.03, -0.070924), (13.27, -0.165411), (5.025, -0.116955), (1.013, 0.557368)]),
        ("S", [(0.3023, 1.0)]),
        ("P", [(17.7, 0.043018), (3.854, 0.228913), (1.046, 0.508728)]),
        ("P", [(0.2753, 1.0)]),
        ("D", [(1.185, 1.0)]),
    ],
}
NCART = {"S": 1, "P": 3, "D": 6}


def rng_for(config):
    return np.random.Generator(np.random.PCG64(20260000 + config))


def _unit(v):
    return v / np.linalg.norm(v)


def alkane(n_carbon, rng, jitter=0.02):
    """Atoms [(symbol, xyz in angstrom)] of a zig-zag C_nH_{2n+2}."""
    cc, ch = 1.54, 1.09
    theta = np.deg2rad(109.4712206)
    dx = cc * np.sin(theta / 2)
    dy = cc * np.cos(theta / 2)
    C = np.array([[i * dx, (i % 2) * dy, 0.0] for i in range(n_carbon)])
    atoms = [("C", c) for c in C]
    for i in range(n_carbon):
        nb = [C[j] for j in (i - 1, i + 1) if 0 <= j < n_carbon]
        if len(nb) == 2:
            u1, u2 = _unit(nb[0] - C[i]), _unit(nb[1] - C[i])
            b = _unit(-(u1 + u2))
            n = _unit(np.cross(u1, u2))
            a = theta / 2
            for s in (1, -1):
                atoms.append(("H", C[i] + ch * (np.cos(a) * b + s * np.sin(a) * n)))
        else:
            ax = _unit(nb[0] - C[i]) if nb else np.array([1.0, 0.0, 0.0])
            e1 = _unit(np.cross(ax, [0.0, 0.0, 1.0]))
            e2 = np.cross(ax, e1)
            nh = 3 if nb else 4
            if nh == 3:
                for k in range(3):
                    phi = 2 * np.pi * k / 3
                    d = np.cos(theta) * ax + np.sin(theta) * (np.cos(phi) * e1 + np.sin(phi) * e2)
                    atoms.append(("H", C[i] + ch * d))
            else:   # methane
                for d in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]):
                    atoms.append(("H", C[i] + ch * _unit(np.array(d, float))))
    xyz = np.array([a[1] for a in atoms])
    xyz = xyz + rng.uniform(-jitter, jitter, xyz.shape)
    return [(a[0], xyz[k]) for k, a in enumerate(atoms)]


def water_cluster(n_water, rng, spacing=3.1, jitter=0.1):
    """Atoms of n_water H2O on a jittered cubic lattice, random orientations."""
    side = int(np.ceil(n_water ** (1 / 3) - 1e-9))
    roh, ang = 0.96, np.deg2rad(104.5)
    local = np.array([[0.0, 0.0, 0.0],
                      [roh * np.sin(ang / 2), roh * np.cos(ang / 2), 0.0],
                      [-roh * np.sin(ang / 2), roh * np.cos(ang / 2), 0.0]])
    atoms = []
    cnt = 0
    for i in range(side):
        for j in range(side):
            for k in range(side):
                if cnt == n_water:
                    break
                q = rng.standard_normal(4)
                q /= np.linalg.norm(q)
                w, x, y, z = q
                R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                              [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                              [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
                c = np.array([i, j, k], float) * spacing + rng.uniform(-jitter, jitter, 3)
                for s, p in zip(("O", "H", "H"), local):
                    atoms.append((s, c + R @ p))
                cnt += 1
    return atoms


def mol_text(atoms):
    """.mol text (bohr) with the cc-pVDZ blocks above."""
    out = [str(len(atoms))]
    for sym, xyz in atoms:
        shells = BASIS[sym]
        nprim = sum(len(p) for _, p in shells)
        out.append(f"{sym} {len(shells)} {nprim} {nprim}")
        b = np.asarray(xyz) * ANGSTROM
        out.append(f"{float(b[0])!r} {float(b[1])!r} {float(b[2])!r}")
        for typ, prims in shells:
            out.append(f"{typ} {len(prims)} 1")
            for e, c in prims:
                out.append(f"{float(e)!r}    {float(c)!r}")
    return "\n".join(out) + "\n"


def bf_centers(atoms):
    """Centre (bohr) of every Cartesian basis function, in .mol order."""
    cs = []
    for sym, xyz in atoms:
        b = np.asarray(xyz) * ANGSTROM
        for typ, _ in BASIS[sym]:
            cs.extend([b] * NCART[typ])
    return np.array(cs)


def density(atoms, nocc, rng):
    """Symmetric damped density D = (2 Q Q^T) o exp(-0.5 |R_i - R_j|)."""
    R = bf_centers(atoms)
    nbf = len(R)
    C = rng.standard_normal((nbf, nocc))
    Q, _ = np.linalg.qr(C)
    D0 = 2.0 * Q @ Q.T
    dist = np.linalg.norm(R[:, None, :] - R[None, :, :], axis=-1)
    D = D0 * np.exp(-0.5 * dist)
    return 0.5 * (D + D.T)


# Workload definitions of the benchmark configurations
CONFIGS = {
    0: dict(kind="alkane", n=10, reltol=1e-6, nocc=41),
    1: dict(kind="alkane", n=30, reltol=1e-6, nocc=121),
    2: dict(kind="alkane", n=100, reltol=1e-8, nocc=401),
    4: dict(kind="water", n=64, reltol=1e-6, nocc=320),
}


def batch_densities(atoms, nocc, count=16):
    """The batching workload's densities on a molecule: `count` damped
    densities of nocc occupied columns, seeded with the batching
    configuration's generator (20260003)."""
    rng = rng_for(3)
    return np.stack([density(atoms, nocc, rng) for _ in range(count)])


def workload_name(config):
    """Short name of a configuration's molecule (C30H62, (H2O)64, ...)."""
    cfg = CONFIGS[config]
    if cfg["kind"] == "alkane":
        return f"C{cfg['n']}H{2 * cfg['n'] + 2}"
    return f"(H2O){cfg['n']}"


def workload(config):
    """(atoms, mol text, density, reltol) of a configuration, seeded."""
    cfg = CONFIGS[config]
    rng = rng_for(config)
    atoms = alkane(cfg["n"], rng) if cfg["kind"] == "alkane" else water_cluster(cfg["n"], rng)
    D = density(atoms, cfg["nocc"], rng)
    return atoms, mol_text(atoms), D, cfg["reltol"]
