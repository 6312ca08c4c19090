"""Synthetic, seeded systems for the nonbonded path.

United-atom chain systems (3,000 / 100,000 / 1,000,000 sites and the uneven
replicas of the replica-exchange workload) and C-alpha Go-model walkers.

Units follow the reference (defines.h:60-71): Angstrom, kJ/mol, amu, e, ps.
Parameter conventions follow the reference's STYPE tables:
  * ljset[ta][tb].eps = 4.184*sqrt(eps_a*eps_b)     (setlj.c:182, kcal -> kJ)
  * ljset[ta][tb].sig = sig_a/2 + sig_b/2           (setlj.c:183, Rmin form)
    used as E = eps*((sig/r)^12 - 2 (sig/r)^6)      (cnonbond_nblist.c:309-311)
  * ljset14 = 0.5 * ljset eps (1-4 LJ scaled by 0.5), E14FAC = 0.5 for charges
  * exclusions in xint order: bonds (xflag 1), bends (2), 1-4 (4)
    (read_topology.c:847-891)
  * bond energy k (r - r0)^2 (cbond.c:103-104), k = 450 kcal/mol/A^2

Chain geometry.  At 0.1007 sites/A^3 with Rmin 3-4 A the system is about
twice as dense as hard-sphere close packing, so a free random walk would put
sites within a few tenths of an Angstrom of each other, with r^-12 forces no
1 fs step can integrate.  Each chain is therefore an all-trans zigzag (bond
1.53 A, angle 112 deg) with a random direction, the zigzags lie in parallel
planes, and chains sit on a jittered lattice of rows along x; sites, types
and empty lattice slots are drawn from the seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

KCAL = 4.184
BOND_R0 = 1.53
BOND_K = 450.0 * KCAL
SITE_MASS = 14.027
BOND_ANGLE = math.radians(112.0)
DENSITY = 0.1007


@dataclass
class ChainSystem:
    """One periodic box of linear united-atom chains (struct-of-arrays, FP64)."""
    n: int
    box: np.ndarray            # (3,) box lengths
    pos: np.ndarray            # (n,3) wrapped into [-L/2, L/2]
    vel: np.ndarray            # (n,3) A/ps
    mass: np.ndarray           # (n,)
    q: np.ndarray              # (n,)
    type: np.ndarray           # (n,) int32 site type (atomid)
    ntypes: int
    lj_eps: np.ndarray         # (ntypes,ntypes) kJ/mol
    lj_sig: np.ndarray         # (ntypes,ntypes) A (Rmin)
    lj14_eps: np.ndarray
    lj14_sig: np.ndarray
    bonds: np.ndarray          # (nb,2) int32
    bond_r0: np.ndarray        # (nb,)
    bond_k: np.ndarray         # (nb,)
    xint: np.ndarray           # (nx,3) int32: a, b, xflag
    rc: float = 10.0
    rclist: float = 12.0
    rclist_ew: float = 12.0
    kappa: float = 0.32
    eps_rf: float = 1.0
    e14fac: float = 0.5
    order: int = 4
    grid: tuple = (32, 32, 32)
    dt: float = 0.001
    temperature: float = 300.0
    name: str = ""
    meta: dict = field(default_factory=dict)


def _type_tables(rng: np.random.Generator, ntypes: int):
    eps_t = rng.uniform(0.05, 0.20, ntypes)          # kcal/mol
    sig_t = rng.uniform(3.0, 4.0, ntypes)            # A
    eps = KCAL * np.sqrt(np.outer(eps_t, eps_t))
    sig = 0.5 * (sig_t[:, None] + sig_t[None, :])
    return eps, sig, 0.5 * eps, sig.copy()


def _zigzag(nsite: int) -> np.ndarray:
    """All-trans chain along +x centred on the origin."""
    half = 0.5 * BOND_ANGLE
    dx = BOND_R0 * math.sin(half)
    dy = BOND_R0 * math.cos(half)
    j = np.arange(nsite, dtype=np.float64)
    out = np.zeros((nsite, 3))
    out[:, 0] = dx * j - dx * (nsite - 1) / 2.0
    out[:, 1] = np.where(np.arange(nsite) % 2 == 0, -0.5 * dy, 0.5 * dy)
    return out


def wrap(pos: np.ndarray, box: np.ndarray) -> np.ndarray:
    """Wrap into [-L/2, L/2] as pbc_all.c:23-88 leaves positions."""
    return pos - box * np.round(pos / box)


def make_chain_system(nchains: int, box_len: float, grid: int, *, seed: int = 1234,
                      nsite: int = 10, ntypes: int = 32, rc: float = 10.0, skin: float = 2.0,
                      kappa: float = 0.32, order: int = 4, temperature: float = 300.0,
                      name: str = "") -> ChainSystem:
    rng = np.random.default_rng(seed)
    L = float(box_len)
    n = nchains * nsite
    eps, sig, eps14, sig14 = _type_tables(rng, ntypes)

    # Lattice of chain slots: rows of chains along x, a rectangular grid of
    # chain axes in (y,z).  The row count maximises the smallest of the axis
    # spacings and the end-to-end gap between chains of one row.
    span = (nsite - 1) * BOND_R0 * math.sin(0.5 * BOND_ANGLE)
    zx = BOND_R0 * math.sin(0.5 * BOND_ANGLE)          # x advance per bond
    zy = BOND_R0 * math.cos(0.5 * BOND_ANGLE)          # zigzag amplitude (peak to peak)
    best = None
    for rows_ in range(1, max(1, int(L // span)) + 1):
        per = -(-nchains // rows_)
        gap = L / rows_ - span
        for ny_ in range(1, per + 1):
            nz_ = -(-per // ny_)
            sy, sz = L / ny_, L / nz_
            # closest approach of in-phase zigzags one row apart in y
            dy = math.sqrt(max(sy - zy, 0.0) ** 2 + zx * zx)
            score = min(gap, sy, dy, sz)
            if best is None or score > best[0]:
                best = (score, rows_, ny_, nz_)
    _, rows, ny, nz = best
    slots = np.array([(r, iy, iz) for r in range(rows) for iy in range(ny) for iz in range(nz)])
    pick = rng.permutation(len(slots))[:nchains]
    slots = slots[np.sort(pick)]
    # All zigzags lie in planes parallel to xy and in phase.  A reversed chain
    # is the same point set rotated by 180 degrees about z (the zigzag of an
    # even number of sites is centrosymmetric), so reversal only reorders sites.
    base = _zigzag(nsite)
    flip = rng.integers(0, 2, nchains) * 2 - 1
    jit = rng.uniform(-0.5, 0.5, (nchains, 3)) * np.array([0.0, 0.1, 0.1])
    centre = np.column_stack([(slots[:, 0] + 0.5) * L / rows, (slots[:, 1] + 0.5) * L / ny,
                              (slots[:, 2] + 0.5) * L / nz]) - 0.5 * L
    b = np.broadcast_to(base, (nchains, nsite, 3)).copy()
    b[:, :, 0] *= flip[:, None]
    b[:, :, 1] *= flip[:, None]
    pos = b + (centre + jit)[:, None, :]
    pos = wrap(pos.reshape(n, 3), np.array([L, L, L]))
    self_meta = {"rows": rows, "ny": ny, "nz": nz, "min_spacing": best[0]}

    types = rng.integers(0, ntypes, n).astype(np.int32)
    q = np.tile(np.where(np.arange(nsite) % 2 == 0, 0.40, -0.40), nchains)
    mass = np.full(n, SITE_MASS)

    first = (np.arange(nchains) * nsite)[:, None]
    j = np.arange(nsite - 1)[None, :]
    bonds = np.stack([(first + j).ravel(), (first + j + 1).ravel()], axis=1).astype(np.int32)
    j2 = np.arange(nsite - 2)[None, :]
    bends = np.stack([(first + j2).ravel(), (first + j2 + 2).ravel()], axis=1)
    j3 = np.arange(nsite - 3)[None, :]
    p14 = np.stack([(first + j3).ravel(), (first + j3 + 3).ravel()], axis=1)
    xint = np.concatenate([
        np.column_stack([bonds, np.full(len(bonds), 1)]),
        np.column_stack([bends, np.full(len(bends), 2)]),
        np.column_stack([p14, np.full(len(p14), 4)]),
    ]).astype(np.int32)

    # Maxwell velocities at T (vinit.c:40-43), centre-of-mass momentum removed
    kT_over_m = 1.380662e-23 * temperature / (mass * 1.6605655e-27) * 1e-4
    vel = rng.standard_normal((n, 3)) * np.sqrt(kT_over_m)[:, None]
    vel -= (mass[:, None] * vel).sum(0) / mass.sum()

    return ChainSystem(
        n=n, box=np.array([L, L, L]), pos=np.ascontiguousarray(pos), vel=vel, mass=mass,
        q=q, type=types, ntypes=ntypes, lj_eps=eps, lj_sig=sig, lj14_eps=eps14, lj14_sig=sig14,
        bonds=bonds, bond_r0=np.full(len(bonds), BOND_R0), bond_k=np.full(len(bonds), BOND_K),
        xint=xint, rc=rc, rclist=rc + skin, rclist_ew=rc + skin, kappa=kappa, order=order,
        grid=(grid, grid, grid), temperature=temperature, name=name, meta=self_meta)


@dataclass
class AngleTerms:
    """Angle-type bonded terms of a chain system (read_topology.c's bend,
    urey, tors and impr arrays): bends (a,b,c) with angeq [rad] and krbend,
    Urey-Bradley 1-3 pairs (a,b) with Seq and k, torsions (a,b,c,d) with six
    Fourier entries each (kphi, nphi, delphi; a zero kphi ends the series),
    impropers (a,b,c,d) with kimpr and angeq."""
    bend: np.ndarray           # (nbend,3) int32
    bend_eq: np.ndarray
    bend_k: np.ndarray
    urey: np.ndarray           # (nurey,2) int32
    urey_s0: np.ndarray
    urey_k: np.ndarray
    tors: np.ndarray           # (ntors,4) int32
    kphi: np.ndarray           # (ntors,6)
    nphi: np.ndarray
    delphi: np.ndarray
    impr: np.ndarray           # (nimpr,4) int32
    impr_k: np.ndarray
    impr_eq: np.ndarray


def dihedral(pos: np.ndarray, box: np.ndarray, quads: np.ndarray) -> np.ndarray:
    """Signed dihedral angles of a-b-c-d quads (minimum image), radians, with
    the sign of (n1 x n2) . b2 as ctorsion.c:171-182 takes it."""
    def mi(d):
        return d - box * np.round(d / box)
    b1 = mi(pos[quads[:, 1]] - pos[quads[:, 0]])
    b2 = mi(pos[quads[:, 2]] - pos[quads[:, 1]])
    b3 = mi(pos[quads[:, 3]] - pos[quads[:, 2]])
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    y = (np.cross(n1, n2) * b2).sum(1) / np.linalg.norm(b2, axis=1)
    return np.arctan2(y, (n1 * n2).sum(1))


def angle_terms(s: ChainSystem, seed: int = 5, generic: bool = False) -> AngleTerms:
    """United-atom alkane-like angle terms for every chain of s: a bend and a
    Urey-Bradley pair on each consecutive triple (angle 112 deg, CHARMM-like
    constants), a three-term torsion on each consecutive quad, and a
    planarity improper (i, i+1, i+3, i+2), equilibrium 0, on every third
    quad.  generic: torsion phases and improper equilibria drawn at random
    instead (energies that are not even in the dihedral, for tests of its
    sign)."""
    rng = np.random.default_rng(seed)
    nb = len(s.bonds)
    # chains: runs of consecutive bonds i -> i+1
    first = [0]
    for k in range(1, nb):
        if s.bonds[k, 0] != s.bonds[k - 1, 1]:
            first.append(k)
    first.append(nb)
    tri, quad = [], []
    for c0, c1 in zip(first[:-1], first[1:]):
        sites = [int(s.bonds[c0, 0])] + [int(x) for x in s.bonds[c0:c1, 1]]
        tri += [sites[i:i + 3] for i in range(len(sites) - 2)]
        quad += [sites[i:i + 4] for i in range(len(sites) - 3)]
    tri = np.array(tri, np.int32).reshape(-1, 3)
    quad = np.array(quad, np.int32).reshape(-1, 4)
    nt = len(quad)
    kphi = np.zeros((nt, 6)); nphi = np.zeros((nt, 6)); delphi = np.zeros((nt, 6))
    kphi[:, :3] = KCAL * np.array([0.195, 0.10, 0.05]) * rng.uniform(0.8, 1.2, (nt, 3))
    nphi[:, :3] = [3.0, 2.0, 1.0]
    delphi[:, :3] = [0.0, math.pi, 0.0]
    imp = np.ascontiguousarray(quad[::3][:, [0, 1, 3, 2]])
    impr_eq = np.zeros(len(imp))
    if generic:
        delphi[:, :3] = rng.uniform(-math.pi, math.pi, (nt, 3))
        impr_eq = rng.uniform(-0.3, 0.3, len(imp))
    return AngleTerms(
        bend=tri, bend_eq=np.full(len(tri), BOND_ANGLE), bend_k=np.full(len(tri), 58.35 * KCAL),
        urey=np.ascontiguousarray(tri[:, [0, 2]]), urey_s0=np.full(len(tri), 2.561),
        urey_k=np.full(len(tri), 11.16 * KCAL),
        tors=quad, kphi=kphi, nphi=nphi, delphi=delphi,
        impr=imp, impr_k=np.full(len(imp), 2.0 * KCAL), impr_eq=impr_eq)


def jitter(s: ChainSystem, amp: float, seed: int = 9) -> np.ndarray:
    """Positions of s displaced by uniform noise of +-amp per component
    (wrapped): breaks the exact planarity of the all-trans start."""
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(wrap(s.pos + rng.uniform(-amp, amp, s.pos.shape), s.box))


def chains_3k(seed: int = 1234) -> ChainSystem:
    """300 chains x 10 sites, 31 A box, PME 32^3."""
    return make_chain_system(300, 31.0, 32, seed=seed, name="chains-3k")


def chains_100k(seed: int = 1234) -> ChainSystem:
    """10,000 chains x 10 sites, 100 A box, PME 96^3."""
    return make_chain_system(10000, 100.0, 96, seed=seed, name="chains-100k")


def chains_1m(seed: int = 1234) -> ChainSystem:
    """100,000 chains x 10 sites, 215 A box, PME 224^3, cutoff 12 A."""
    return make_chain_system(100000, 215.0, 224, seed=seed, rc=12.0, name="chains-1m")


def even_grid(box_len: float) -> int:
    """pme_setup.c:19-33: K = (int)L, made even."""
    k = int(box_len)
    return k + (k % 2)


def replica_sizes(nrep: int = 64, lo: int = 2000, hi: int = 20000, seed: int = 77) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(math.log(lo), math.log(hi), nrep))
    return (np.round(n / 10.0) * 10).astype(int)


def make_replicas(nrep: int = 64, seed: int = 77, t_lo: float = 300.0, t_hi: float = 600.0):
    """Uneven replicas at the chain density, each with a 1 A PME grid."""
    sizes = replica_sizes(nrep, seed=seed)
    temps = t_lo * (t_hi / t_lo) ** (np.arange(nrep) / max(1, nrep - 1))
    out = []
    for r, nr in enumerate(sizes):
        L = (nr / DENSITY) ** (1.0 / 3.0)
        out.append(make_chain_system(nr // 10, L, even_grid(L), seed=seed * 1000 + r,
                                     temperature=float(temps[r]), name=f"replica{r}"))
    return out


# --------------------------------------------------------------------------
# Go-model walkers
# --------------------------------------------------------------------------

GO_BOND = 3.8
GO_CONTACT = 6.5
GO_SIG_REP = 4.0
GO_EPS = 1.0 * KCAL
GO_KBOND = 100.0 * KCAL


@dataclass
class GoChain:
    n: int
    native: np.ndarray      # (n,3) native structure
    pos: np.ndarray         # (n,3) current configuration
    contacts: np.ndarray    # (nc,2) native contact pairs (a<b, b-a>=3)
    r0: np.ndarray          # (nc,) native distances


def _compact_saw(n: int, rng: np.random.Generator) -> np.ndarray:
    """Self-avoiding walk on a cubic lattice inside a cube sized for n sites,
    growing toward sites with many occupied neighbours (compact)."""
    side = int(math.ceil(n ** (1.0 / 3.0))) + 1
    moves = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    for _attempt in range(1000):
        occ = {(0, 0, 0)}
        path = [(0, 0, 0)]
        ok = True
        for _ in range(n - 1):
            cur = np.array(path[-1])
            cands = []
            for m in moves:
                nxt = tuple(cur + m)
                if nxt in occ or max(abs(v) for v in nxt) > side:
                    continue
                nb = sum((tuple(np.array(nxt) + mm) in occ) for mm in moves)
                free = sum((tuple(np.array(nxt) + mm) not in occ) for mm in moves)
                if free == 0 and len(path) < n - 1:
                    continue
                cands.append((nxt, nb))
            if not cands:
                ok = False
                break
            w = np.array([3.0 ** c[1] for c in cands])
            k = rng.choice(len(cands), p=w / w.sum())
            occ.add(cands[k][0])
            path.append(cands[k][0])
        if ok:
            p = np.array(path, dtype=np.float64) * GO_BOND
            return p - p.mean(0)
    raise RuntimeError("compact walk failed")


def make_go_chain(n: int, rng: np.random.Generator) -> GoChain:
    nat = _compact_saw(n, rng)
    d = np.linalg.norm(nat[:, None, :] - nat[None, :, :], axis=-1)
    ii, jj = np.triu_indices(n, 3)
    m = d[ii, jj] < GO_CONTACT
    contacts = np.stack([ii[m], jj[m]], axis=1).astype(np.int32)
    return GoChain(n=n, native=nat, pos=nat.copy(), contacts=contacts, r0=d[ii[m], jj[m]])


def make_go_walkers(nwalk: int = 256, lo: int = 60, hi: int = 300, seed: int = 4242):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, nwalk)
    return [make_go_chain(int(s), rng) for s in sizes]
