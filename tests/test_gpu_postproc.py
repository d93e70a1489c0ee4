"""GPU post-processing against the reference's numpy definitions
(python/boltzmann/utils/postproc.py:12-127), bit for bit on the float32 velocity."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ref_vmag(v):
    return np.sqrt(np.sum(v ** 2, axis=-1))


def ref_curl(v):
    g = np.gradient(v)
    if v.ndim == 3:
        return g[1][..., 0] - g[0][..., 1]
    return np.stack((g[2][..., 1] - g[1][..., 2], g[0][..., 2] - g[2][..., 0],
                     g[1][..., 0] - g[0][..., 1]), axis=-1)


def ref_stream(v, origin=(0, 0)):
    ivdx = np.cumsum(v[..., 1], axis=0)
    iudy = np.cumsum(v[..., 0], axis=1)
    phi = iudy[0, :][None, :] - ivdx
    phi -= phi[origin[0], origin[1]]
    return phi


def _sim(counts, storage="f32", seed=0):
    import boltzmann_amd
    from boltzmann_amd.core import CellFlags, bgk
    rng = np.random.default_rng(seed)
    sim = boltzmann_amd.Simulation("gpu", list(counts), bgk(0.7), storage=storage)
    d = len(counts)
    sim.fluid.vel[...] = (0.03 * rng.standard_normal(tuple(counts) + (d,))).astype(np.float32)
    sim.cells.flags[(rng.random(counts) < 0.05)] = CellFlags.WALL
    sim.iterate(4)
    return sim


@pytest.mark.parametrize("storage", ["f32", "f64", "f16s"])
def test_postproc_2d_bit_exact(storage):
    from boltzmann_amd import postproc
    sim = _sim((64, 48), storage)
    v = sim.fluid.vel.copy()
    np.testing.assert_array_equal(postproc.calc_vmag(sim), ref_vmag(v))
    np.testing.assert_array_equal(postproc.calc_curl(sim), ref_curl(v))
    np.testing.assert_array_equal(postproc.calc_stream_func(sim), ref_stream(v))
    np.testing.assert_array_equal(postproc.calc_stream_func(sim, origin=(17, 5)), ref_stream(v, (17, 5)))


def test_postproc_3d_bit_exact():
    from boltzmann_amd import postproc
    sim = _sim((12, 10, 9))
    v = sim.fluid.vel.copy()
    np.testing.assert_array_equal(postproc.calc_vmag(sim), ref_vmag(v))
    np.testing.assert_array_equal(postproc.calc_curl(sim), ref_curl(v))
    with pytest.raises(RuntimeError, match="2D"):
        postproc.calc_stream_func(sim)


def test_postproc_edges_and_errors():
    import boltzmann_amd
    from boltzmann_amd import postproc
    from boltzmann_amd.core import bgk
    sim = _sim((2, 5))                                   # two cells: one-sided differences only
    v = sim.fluid.vel.copy()
    np.testing.assert_array_equal(postproc.calc_curl(sim), ref_curl(v))
    with pytest.raises(RuntimeError, match="outside"):
        postproc.calc_stream_func(sim, origin=(2, 0))
    fresh = boltzmann_amd.Simulation("gpu", [8, 8], bgk(0.7))
    with pytest.raises(RuntimeError, match="after iterate"):
        postproc.calc_vmag(fresh)
    thin = _sim((1, 8))
    with pytest.raises(RuntimeError, match="at least 2"):
        postproc.calc_curl(thin)
    np.testing.assert_array_equal(postproc.calc_vmag(thin), ref_vmag(thin.fluid.vel))


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_count_nonfinite_matches_the_host_check(storage):
    """count_nonfinite counts on the device what np.isfinite finds in the read-back arrays
    (the check of the reference's run_sim, simulation.py:241-242)."""
    import boltzmann_amd
    from boltzmann_amd.core import bgk
    sim = boltzmann_amd.Simulation("gpu", [24, 16], bgk(0.6), storage=storage)
    sim.fluid.vel[..., 0] = 0.01
    sim.iterate(2)
    assert sim.count_nonfinite("vel") == 0 and sim.count_nonfinite("rho") == 0
    bad = boltzmann_amd.Simulation("gpu", [24, 16], bgk(0.6), storage=storage)
    bad.fluid.vel[5, 7, 0] = np.nan
    bad.fluid.rho[11, 3] = np.inf
    bad.iterate(4)                      # the non-finite values spread to the neighbours
    for field, host in (("vel", bad.fluid.vel), ("rho", bad.fluid.rho)):
        n = bad.count_nonfinite(field)
        assert n > 2, (field, n)
        assert n == np.count_nonzero(~np.isfinite(host)), field
    with pytest.raises(boltzmann_amd.BoltError):
        bad.count_nonfinite("C")
