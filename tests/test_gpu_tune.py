"""The tile-shape autotuner (mdft_tune_auto / save / load): it measures both
shapes of every kernel, the table round-trips through a file, and transforms
stay within the FP64-oracle bound and bitwise equal whatever the table holds
(the shape changes only how lines are grouped into workgroups)."""
import numpy as np
import pytest

from tests.helpers import check_against_fp64, rand_c64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mdft():
    from math_amd import mdft as M

    M.init(0)
    return M


# FFT shapes: 2^s points for s in 1..14, 3 * 2^(s - 32) points for s in 38..43
SHAPES = list(range(1, 15)) + list(range(38, 44))


def table(M):
    return {(l, b, c): M.tune_get(l, b, c) for l in SHAPES for b in (0, 1) for c in (0, 1)}


def test_autotune_save_load(mdft, tmp_path):
    before = table(mdft)
    rng = np.random.default_rng(12)
    x = rand_c64(rng, (700, 1500))
    ref = x.copy()
    mdft.Plan(1500, 700).exec(False, ref)
    sec = mdft.tune_auto()
    print(f"autotune {sec:.2f} s", flush=True)
    assert 0 < sec < 120
    tuned = table(mdft)
    assert set(tuned.values()) <= {0, 1}
    p = str(tmp_path / "tune.txt")
    mdft.tune_save(p)
    for k in tuned:
        mdft.tune_set(*k, 1 - tuned[k])
    mdft.tune_load(p)
    assert table(mdft) == tuned
    got = x.copy()
    mdft.Plan(1500, 700).exec(False, got)
    assert np.array_equal(got, ref), "tile shape changed the arithmetic"
    check_against_fp64(got, x, False, "tuned")
    for k, v in before.items():
        mdft.tune_set(*k, v)


def test_tune_errors(mdft, tmp_path):
    before = table(mdft)
    bad = tmp_path / "bad.txt"
    bad.write_text("# header\n10 0 1 1\n10 0 1 7\n")
    with pytest.raises(mdft.ErrorException, match="bad entry on line 3"):
        mdft.tune_load(str(bad))
    assert table(mdft) == before, "a rejected file must change nothing"
    with pytest.raises(mdft.ErrorException, match="cannot read"):
        mdft.tune_load(str(tmp_path / "missing.txt"))
    with pytest.raises(mdft.ErrorException, match="shape must be in"):
        mdft.tune_get(15, 0, 0)
    with pytest.raises(mdft.ErrorException, match="shape must be in"):
        mdft.tune_set(37, 1, 0, 0)
    with pytest.raises(mdft.ErrorException, match="variant must be 0 or 1"):
        mdft.tune_set(10, 0, 0, 2)
