"""Full-size parity: the benchmark workloads at their full size and step count on the HIP path,
against the oracle's digests in tests/golden/fullsize_<workload>.npz (written by
tests/golden/make_fullsize.py on the CPU).

* 2x4 blocks of 150x150, Roe/RK4 CFL 0.8, 200 steps (the reference's README.md:7-17 setup);
* 5x1 blocks of 500x500, HLLL/SSP-RK2 CFL 0.4, 500 steps (README.md:152-159's block layout and
  scheme, oblique-shock initial state);
* 64 seeded uneven explosion cases, Roe/RK4 CFL 0.8, 100 steps, per-case dt; cases the scheme
  cannot advance fail at the same step as in the oracle and keep the state that step produced.

The bar is bit-exact: every block's final state has the oracle's sha256, every step's dt equals the
oracle's, and the failing steps agree.  On a mismatch the test reports the first differing dt and
the relative differences of the per-block maxima, sums and sample cells that the digest keeps.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.golden import make_fullsize as MF

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _fixture(workload):
    path = MF.HERE / f"fullsize_{workload}.npz"
    if not path.exists():
        pytest.fail(f"missing fixture {path.name}: run tests/golden/make_fullsize.py {workload}")
    with np.load(path) as z:
        return {k: z[k] for k in z.files}


def _diagnose(states, f, nblocks):
    """Worst relative difference of the kept per-block figures (for the failure message)."""
    got = {"absmax": [], "sum": [], "samples": []}
    for case_states in states:
        d = MF.summarize(case_states)
        for k in got:
            got[k].extend(d[k])
    msg = []
    for k in got:
        g, r = np.array(got[k]), f[k]
        den = np.maximum(np.abs(r), 1e-300)
        nan_pos = bool((np.isnan(g) != np.isnan(r)).any())
        msg.append(f"{k}: max rel diff {np.nanmax(np.abs(g - r) / den):.3e}" + (", NaN positions differ" if nan_pos else ""))
    return "; ".join(msg)


@pytest.mark.parametrize("workload", MF.WORKLOADS)
def test_full_size_run_matches_the_oracle_bit_for_bit(workload):
    from pyhype_amd import Euler2D, Euler2DBatch

    f = _fixture(workload)
    cfgs, meshes, steps = MF.cases(workload)
    assert int(f["steps"]) == steps and int(f["ncases"]) == len(cfgs)
    print(f"full-size parity: {workload}, {len(cfgs)} case(s), {steps} steps", flush=True)
    solver = Euler2D(cfgs[0], meshes[0]) if len(cfgs) == 1 else Euler2DBatch(cfgs, meshes)
    solver.strict = False   # failures are compared with the oracle's, not raised
    solver.apply_initial_condition()
    st0 = solver.states()
    sha0 = [MF.sha(U) for case in st0 for U in case.values()]
    assert sha0 == list(f["sha0"]), "initial states differ from the oracle's"

    ncases = len(cfgs)
    dts = np.zeros((ncases, steps))
    failed = np.full(ncases, -1)
    for s in range(steps):
        solver.run_steps(1)
        bad, _, _, dt = solver.plan.status()
        for c in range(ncases):
            if failed[c] < 0:
                dts[c, s] = dt[c]
                if bad[c] >= 0:
                    failed[c] = bad[c]
    assert list(failed) == list(f["failed"]), f"failing steps {list(failed)} != oracle {list(f['failed'])}"
    for c in range(ncases):
        n = steps if failed[c] < 0 else failed[c]
        same = dts[c, :n] == f["dts"][c, :n]
        assert same.all(), f"case {c}: dt of step {int(np.argmin(same)) + 1} differs: " \
                           f"{dts[c, np.argmin(same)]!r} != {f['dts'][c, np.argmin(same)]!r}"
    states = solver.states()
    sha = [MF.sha(U) for case in states for U in case.values()]
    bad_blocks = [i for i, (a, b) in enumerate(zip(sha, f["sha"])) if a != b]
    assert not bad_blocks, f"final states of blocks {bad_blocks} differ: " + _diagnose(states, f, f["nblocks"])
    t = solver.t
    for c in range(ncases):
        if failed[c] < 0:
            assert t[c] == f["t"][c], f"case {c}: final time {t[c]!r} != {f['t'][c]!r}"
