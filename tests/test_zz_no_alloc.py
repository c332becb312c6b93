"""nsx_glob_cfl_noalloc, the entry point for callers that cannot allocate,
against nsx_glob_cfl on the same resident fields: two calls in a row, the
second on a smaller mesh, each bit for bit equal to the allocating entry
point's values (the same kernels, grid and order of summation, with the
partial sums in the library's static per-device scratch)."""
import pytest

from nsx.case import config_case, make_case

pytestmark = pytest.mark.gpu


def _stepped(case, steps):
    from nsx.device import Context
    G = Context(case)
    G.init_state()
    G.step(steps)
    return G


def test_glob_cfl_noalloc_matches_glob_cfl_twice_smaller_second():
    big = _stepped(config_case(0), 2)
    small = _stepped(make_case(8, 6, 6, seed=3, jitter=0.1, bed=True), 1)
    a = big.glob_cfl_noalloc()
    b = small.glob_cfl_noalloc()
    ref_a, ref_b = big.glob_cfl(), small.glob_cfl()
    for got, ref in ((a, ref_a), (b, ref_b)):
        assert got.keys() == ref.keys()
        for k in ref:
            assert got[k] == ref[k], (k, got[k], ref[k])
    # the scratch keeps the last call's contents; a third call on the big
    # mesh starts again from its own partial sums
    assert big.glob_cfl_noalloc() == ref_a
    big.close()
    small.close()
