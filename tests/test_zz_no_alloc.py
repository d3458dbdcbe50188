"""dgr_gauge_run_noalloc (the gauge run staged in the library's fixed
per-device workspace, no device allocation) against dgr_gauge_run: the same
depths bit for bit on an ensemble of two uneven coastal members, on two calls
in a row, the second with fewer gauges and steps (a smaller input, so stale
workspace contents from the first call would show), and a clear error when a
request does not fit the workspace."""
import numpy as np
import pytest

from dgrain import configs, native

pytestmark = pytest.mark.gpu


def _ensemble():
    wls = [configs.coastal(N=N, nsteps=0, seed=s) for N, s in ((12, 1), (7, 2))]
    e = native.Ensemble([w.mesh for w in wls], [w.solver_params() for w in wls])
    for i, w in enumerate(wls):
        e.set_state(i, w.zeta, w.qx, w.qy)
    return wls, e


def _gauges(wls, ng, seed):
    rng = np.random.default_rng(seed)
    return np.stack([np.sort(rng.choice(w.num_el, ng, replace=False)) for w in wls]).astype(np.int32)


def test_noalloc_gauge_run_matches(gpu):
    wls, ref = _ensemble()
    _, na = _ensemble()
    for nsteps, every, ng, seed in ((24, 3, 9, 0), (10, 5, 4, 1)):
        gel = _gauges(wls, ng, seed)
        a = ref.gauge_run(nsteps, every, gel)
        b = na.gauge_run_noalloc(nsteps, every, gel)
        assert a.shape == b.shape == (2, nsteps // every, ng)
        assert np.array_equal(a, b), np.abs(a - b).max()
        for i in range(2):   # both handles stepped the same way
            for x, y in zip(ref.get_state(i), na.get_state(i)):
                assert np.array_equal(x, y)
    ref.close()
    na.close()


def test_noalloc_gauge_run_rejects_oversize(gpu):
    wls, e = _ensemble()
    gel = _gauges(wls, 64, 3)
    nsamp = native.NOALLOC_SCRATCH_BYTES // (8 * 2 * 64) + 1   # one sample more than fits
    with pytest.raises(native.DgrError) as ex:
        e.gauge_run_noalloc(nsamp, 1, gel)
    assert ex.value.code == native.DGR_ERR_TOO_LARGE
    e.close()
