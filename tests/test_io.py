"""Reference-format inputs and .trr output (include/ttc_io.h).

The native readers (thecodecombo_7_amd/host/ttc_io.c) are compared array by
array, bit for bit, with what the reference's own read_sim.c / read_atom.c /
read_config.c / read_topology.c / setlj.c / pbc_init.c leave in its globals
for the same files (compiled by oracle/Makefile into oracle/_ref, driven by
tests/ref_io_dump.py in a child process), and .trr files byte for byte with
the reference's write_trr.c.  The cases cover both simul.input forms, two
molecule types, explicit exclusions overlapping the 1-4 pairs, bends,
torsions and impropers listed in either direction and out of order (every
orientation case of the 1-4 search), torsions with several exact Fourier
rows and wildcard rows shadowing later exact ones, NBFIX, and the
reference's exit codes.  Where oracle/_ref is absent (no reference tree) the
comparisons against it are skipped; the round-trip checks still run.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from thecodecombo_7_amd import refio, system  # noqa: E402

REF_IO = os.path.join(ROOT, "oracle", "_ref", "libttcref_io.so")
need_ref = pytest.mark.skipif(not os.path.exists(REF_IO), reason="oracle/_ref/libttcref_io.so not built")

ARRAYS = ["pos_nopbc", "pos", "q", "mass", "atomid", "resid", "type_atomid", "lj_eps", "lj_sig", "lj14_eps",
          "lj14_sig", "bonds", "bond_r0", "bond_k", "bends", "bend_eq", "bend_k", "ureys", "urey_s0", "urey_k",
          "tors", "kphi", "nphi", "delphi", "impr", "impr_k", "impr_eq", "exclusions", "in14", "xint"]


def scrambled_inputs(nchains=60, seed=3, ncomp=2, nbfix=True, exclusions=True):
    """A chain system whose psf lists bonded tuples reversed and out of
    order, with explicit exclusions (some of them 1-4 pairs) and NBFIX."""
    s = system.make_chain_system(nchains, 22.0, 24, seed=seed, ntypes=6)
    a = system.angle_terms(s)
    rng = np.random.default_rng(seed + 100)
    ex = None
    if exclusions:
        p14 = np.array([(q[0], q[3]) for q in a.tors[::5]])
        far = rng.integers(0, s.n, (25, 2))
        far = far[far[:, 0] != far[:, 1]]
        ex = np.concatenate([p14, far]).astype(np.int32)
    ri = refio.from_chain_system(s, a, ncomp=ncomp, exclusions=ex,
                                 nbfix=[(0, 1, -0.3, 3.7), (2, 2, -0.15, 4.1)] if nbfix else None)

    def scramble(arr):
        arr = np.array(arr)
        flip = rng.uniform(size=len(arr)) < 0.5
        arr[flip] = arr[flip][:, ::-1]
        return arr[rng.permutation(len(arr))]

    ri.bonds = scramble(ri.bonds)
    ri.bends = scramble(ri.bends)
    ri.tors = scramble(ri.tors)
    ri.impr = scramble(ri.impr)
    return s, a, ri


def ref_dump(case_dir, tmp_path, kw=False):
    out = str(tmp_path / ("ref_kw.npz" if kw else "ref.npz"))
    subprocess.run([sys.executable, os.path.join(ROOT, "tests", "ref_io_dump.py"), "read", case_dir, out]
                   + (["kw"] if kw else []), check=True, capture_output=True, timeout=300)
    return np.load(out)


def compare(r, z):
    for k in ARRAYS:
        a, b = np.asarray(r[k]), z[k]
        assert a.shape == b.shape, (k, a.shape, b.shape)
        assert np.array_equal(a, b), (k, np.argwhere(a != b)[:5])
    assert list(z["names"]) == r["names"]
    c = r["counts"]
    assert [c[k] for k in ("nsite", "ncomp", "nbond", "nbend", "nurey", "ntors", "nimpr", "nexcl", "n14",
                           "nxint")] == list(z["counts"])
    m = r["sim"]
    assert [m[k] for k in ("lx", "ly", "lz", "rc", "rclist", "rc_ew", "rclist_ew", "kappa", "eps_rf", "dt",
                           "temp", "nhc_q")] == list(z["sim_d"])
    assert [m["order"], *m["grid"], m["nhc_m"], m["respa"], m["nboxes"], m["mode"], m["intmode"]] == \
        list(z["sim_i"])


@need_ref
@pytest.mark.parametrize("kw", [False, True])
def test_readers_equal_the_reference_readers(tmp_path, kw):
    s, a, ri = scrambled_inputs()
    case = str(tmp_path / "case")
    ri.write(os.path.join(case, "INPUT"), keyword_sim=kw)
    r = refio.read_raw(os.path.join(case, "INPUT"))
    z = ref_dump(case, tmp_path, kw)
    compare(r, z)
    # the case exercises what it claims to
    assert r["counts"]["nexcl"] > 0 and r["counts"]["ncomp"] == 2
    assert (r["kphi"][:, 1] != 0).any(), "no torsion with several Fourier terms"
    nt = (r["kphi"] != 0).sum(1)
    assert (nt == 1).any() and (nt >= 2).any()
    b = r["bends"]
    assert (b[:, 0] > b[:, 2]).any() and (b[:, 0] < b[:, 2]).any()
    # 1-4 pairs excluded explicitly are not listed twice
    p14 = {tuple(sorted(p)) for p in r["in14"][:, [0, 3]]}
    exs = {tuple(sorted(p)) for p in r["exclusions"]}
    assert not (p14 & exs)
    assert len(p14) == r["counts"]["n14"]


def test_round_trip_of_a_chain_system(tmp_path):
    """Written then read back: positions, charges, types, LJ tables, bonds,
    the angle-type site tuples and the exclusion table (bonds, bend ends,
    1-4 pairs derived from the bends) equal the generating system's (the LJ
    tables within an ulp: eps passes through eps/4.184 and back)."""
    s = system.make_chain_system(40, 20.0, 20, seed=7, ntypes=5)
    a = system.angle_terms(s)
    ri = refio.from_chain_system(s, a)
    d = str(tmp_path / "INPUT")
    ri.write(d)
    s2, ang = refio.read_inputs(d, e14fac=0.5)
    for k in ("bend", "urey", "tors", "impr"):
        np.testing.assert_array_equal(getattr(ang, k), getattr(a, k))
    np.testing.assert_array_equal(s2.pos, s.pos)
    np.testing.assert_array_equal(s2.q, s.q)
    np.testing.assert_array_equal(s2.type, s.type)
    np.testing.assert_array_equal(s2.bonds, s.bonds)
    np.testing.assert_array_equal(s2.xint, s.xint)
    np.testing.assert_allclose(s2.lj_eps, s.lj_eps, rtol=4e-16)
    np.testing.assert_array_equal(s2.lj_sig, s.lj_sig)
    np.testing.assert_allclose(s2.bond_k, s.bond_k, rtol=4e-16)
    assert s2.grid == s.grid and s2.rc == s.rc and s2.kappa == s.kappa


def test_one_four_pairs_equal_a_brute_force_search(tmp_path):
    """The 1-4 list equals, as a set, the a-d pairs of bend pairs sharing an
    edge minus bonded, bent and explicitly excluded pairs (found here by a
    plain double loop over the bends)."""
    s, a, ri = scrambled_inputs(nchains=30, seed=5)
    d = str(tmp_path / "INPUT")
    ri.write(d)
    r = refio.read_raw(d)
    bends = r["bends"]
    taken = {tuple(sorted(p)) for p in r["bonds"]} | {tuple(sorted((b[0], b[2]))) for b in bends} | \
        {tuple(sorted(p)) for p in r["exclusions"]}
    want = set()
    for i in range(len(bends)):
        for j in range(i + 1, len(bends)):
            bi, bj = bends[i], bends[j]
            for x in (bi, bi[::-1]):
                for y in (bj, bj[::-1]):
                    if x[1] == y[0] and x[2] == y[1]:
                        p = tuple(sorted((x[0], y[2])))
                        if p not in taken:
                            want.add(p)
    got = {tuple(sorted((q[0], q[3]))) for q in r["in14"]}
    assert got == want


@need_ref
def test_trr_frames_equal_the_reference_writer(tmp_path):
    rng = np.random.default_rng(2)
    n, nfr = 37, 3
    x = rng.uniform(-15, 15, (nfr, n, 3))
    v = rng.normal(0, 3, (nfr, n, 3))
    f = rng.normal(0, 300, (nfr, n, 3))
    box = np.array([31.0, 29.5, 30.25])
    cycles = [0, 10, 20]
    for fx, fv, ff in ((1, 0, 0), (1, 1, 1)):
        tag = f"{fx}{fv}{ff}"
        npz = str(tmp_path / f"trr{tag}.npz")
        np.savez(npz, x=x, v=v, f=f, box=box, dt=0.001, cycle=np.array(cycles), flag=0, cyc_eq=0, fx=fx, fv=fv,
                 ff=ff, findex=7, nframes=nfr)
        out = tmp_path / f"ref{tag}"
        subprocess.run([sys.executable, os.path.join(ROOT, "tests", "ref_io_dump.py"), "trr", str(out), npz],
                       check=True, capture_output=True, timeout=120)
        ref = (out / "OUTPUT" / "GROMACS" / "gromacs0_7.trr").read_bytes()
        mine = str(tmp_path / f"mine{tag}.trr")
        for k in range(nfr):
            refio.write_trr(mine, x[k], box, vel=v[k] if fv else None, force=f[k] if ff else None,
                            step=cycles[k], time=np.float32(0.001 * cycles[k]))
        assert open(mine, "rb").read() == ref
        frames = refio.read_trr(mine)
        assert [fr["step"] for fr in frames] == cycles
        np.testing.assert_allclose(frames[1]["x"], x[1] / 10, rtol=1e-7)
        assert ("v" in frames[0]) == bool(fv)


def test_trr_layout_by_hand(tmp_path):
    """Header fields, big-endian reals rounded through float, nm units."""
    p = str(tmp_path / "a.trr")
    x = np.array([[1.0, 2.0, 3.0], [0.1, -0.2, 4.4]])
    refio.write_trr(p, x, (30.0, 31.0, 32.0), step=42, time=0.042)
    raw = open(p, "rb").read()
    hdr = np.frombuffer(raw, ">i4", 16)
    assert list(hdr) == [1993, 13, 0, 0, 0, 72, 0, 0, 0, 0, 48, 0, 0, 2, 42, 0]
    assert len(raw) == 64 + 16 + 72 + 48
    t = np.frombuffer(raw, ">f8", 1, 64)[0]
    assert t == float(np.float32(0.042))
    xs = np.frombuffer(raw, ">f8", 6, 64 + 16 + 72).reshape(2, 3)
    np.testing.assert_array_equal(xs, (x / 10).astype(np.float32).astype(np.float64))


def bad_mol(ri, d):
    """Rewrite mol.input with one molecule more than the psf holds."""
    ri.write(d)
    good = list(ri.comp_nmol)
    ri.comp_nmol = [good[0] + 1] + good[1:]
    with open(os.path.join(d, "mol.input"), "w") as fh:
        fh.write(ri._mol())
    ri.comp_nmol = good


def test_reference_exit_codes(tmp_path):
    s, a, ri = scrambled_inputs(nchains=20, seed=9, ncomp=1, nbfix=False, exclusions=False)
    d = str(tmp_path / "INPUT")
    ri.write(d)
    refio.read_raw(d)                                   # the intact case reads
    with pytest.raises(refio.IoError) as e:
        refio.read_raw(str(tmp_path / "missing"))
    assert e.value.code == 1
    # psf site count that disagrees with mol.input: 102
    bad_mol(ri, d)
    with pytest.raises(refio.IoError) as e:
        refio.read_raw(d)
    assert e.value.code == 102
    # a bond between types with no parameter row: 1
    rows = ri.bond_rows
    ri.bond_rows = rows[1:]
    ri.write(d)
    with pytest.raises(refio.IoError) as e:
        refio.read_raw(d)
    assert e.value.code == 1 and "bond parameters" in str(e.value)
    ri.bond_rows = rows
    # a crd site name that differs from the psf: 1
    names = list(ri.site_name)
    ri.site_name = names[:3] + ["ZZ"] + names[4:]
    ri.write(d)
    with open(os.path.join(d, "simul0.crd")) as fh:
        lines = fh.read().splitlines()
    ri.site_name = names
    ri.write(d)
    with open(os.path.join(d, "simul0.crd"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    with pytest.raises(refio.IoError) as e:
        refio.read_raw(d)
    assert e.value.code == 1 and "does not match" in str(e.value)
    # simul.input without a timestep (keyword form): 2
    ri.write(d, keyword_sim=True)
    with open(os.path.join(d, "simul.input")) as fh:
        txt = fh.read()
    with open(os.path.join(d, "simul.input"), "w") as fh:
        fh.write("\n".join(l for l in txt.splitlines() if not l.startswith("sim_timestep")) + "\n")
    with pytest.raises(refio.IoError) as e:
        refio.read_raw(d)
    assert e.value.code == 2


@need_ref
def test_reference_rejects_the_same_bad_site_count(tmp_path):
    """The psf/mol.input mismatch above is the reference's exit 102."""
    s, a, ri = scrambled_inputs(nchains=20, seed=9, ncomp=1, nbfix=False, exclusions=False)
    case = tmp_path / "case"
    bad_mol(ri, str(case / "INPUT"))
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "ref_io_dump.py"), "read", str(case),
                        str(tmp_path / "x.npz")], capture_output=True, timeout=120)
    assert r.returncode == 102
    with pytest.raises(refio.IoError) as e:
        refio.read_raw(str(case / "INPUT"))
    assert e.value.code == 102


def test_round_trip_of_bonds_only_chemistry(tmp_path):
    """A chain system without angle energies exports its bends with zero
    force constants (the reference derives the 1-3 exclusions and the 1-4
    pairs from them) and reads back with no angle terms and the same
    exclusion table."""
    s = system.make_chain_system(30, 18.0, 18, seed=4, ntypes=4)
    d = str(tmp_path / "INPUT")
    refio.from_chain_system(s).write(d)
    s2, ang = refio.read_inputs(d, e14fac=0.5)
    assert ang is None
    np.testing.assert_array_equal(s2.xint, s.xint)
    np.testing.assert_array_equal(s2.pos, s.pos)
    np.testing.assert_array_equal(s2.lj14_sig, s.lj14_sig)
    np.testing.assert_allclose(s2.lj14_eps, s.lj14_eps, rtol=4e-16)
