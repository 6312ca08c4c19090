"""The autotuners and their cache (ttc_tune, ttc_tune_spread,
ttc_tune_cached; no reference counterpart).

  * the spread tuner times 64/128/256 sites per pass and keeps one of them;
    the charge grid accumulates in int64 fixed point, so forces computed with
    any chunk are bit-identical;
  * the cache: the first call on a context shape tunes and appends one line,
    a second context of the same shape applies that line without timing,
    a context of another shape tunes again, and lines of another key or
    malformed lines are ignored.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_spread_chunk_does_not_change_forces():
    from thecodecombo_7_amd import system
    from thecodecombo_7_amd.nonbond import Box
    s = system.chains_3k()
    print("case: chains-3k, spread chunk tuning", flush=True)
    b = Box(s)
    try:
        b.nblist()
        f0, e0 = b.forces()
        ch, t = b.tune_spread()
        print("spread chunk", ch, "timings", t, flush=True)
        assert ch in (64, 128, 256) and t[[64, 128, 256].index(ch)] == t[t > 0].min()
        f1, e1 = b.forces()
        np.testing.assert_array_equal(f1, f0)
        assert e1 == e0
    finally:
        b.close()


def test_tuning_cache(tmp_path):
    from thecodecombo_7_amd import system
    from thecodecombo_7_amd.nonbond import Box
    path = tmp_path / "tune.txt"
    path.write_text("garbage line without a tab\nn=1 box=1,1,1\t8 64\n")
    s = system.chains_3k()
    print("case: chains-3k, tuning cache", flush=True)
    boxes = []
    try:
        b = Box(s); boxes.append(b)
        b.nblist()
        w, ch, hit = b.tune_cached(str(path))
        assert not hit and w in (4, 8, 16) and ch in (64, 128, 256)
        lines = path.read_text().splitlines()
        assert len(lines) == 3 and lines[-1].endswith(f"\t{w} {ch}")
        b2 = Box(s); boxes.append(b2)
        b2.nblist()
        w2, ch2, hit2 = b2.tune_cached(str(path))
        assert hit2 and (w2, ch2) == (w, ch)
        assert len(path.read_text().splitlines()) == 3
        f1, _ = b.forces()
        f2, _ = b2.forces()
        np.testing.assert_array_equal(f1, f2)
        # another shape: tuned and appended
        s3 = system.make_chain_system(120, 24.0, 24, seed=5)
        b3 = Box(s3); boxes.append(b3)
        b3.nblist()
        _, _, hit3 = b3.tune_cached(str(path))
        assert not hit3 and len(path.read_text().splitlines()) == 4
    finally:
        for b in boxes:
            b.close()
