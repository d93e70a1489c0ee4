"""The batch driver (boltzmann_amd.simulation): checkpoint gaters and IterInfo on the CPU
(python/tests/test_simulation.py:7-45 of the reference covers the same helpers), run_sim
on the GPU."""

import datetime
import os

import numpy as np
import pytest

from boltzmann_amd.simulation import (EveryN, EveryT, IterInfo, NoCheckpoints,
                                      parse_checkpoints, run_sim)


def test_every_n():
    g = EveryN(3)
    assert [g.allow() for _ in range(7)] == [False, True, False, False, True, False, False]
    with pytest.raises(ValueError):
        EveryN(0)


def test_every_t():
    g = EveryT(datetime.timedelta(seconds=3600))
    assert not g.allow()
    g._last -= datetime.timedelta(seconds=7200)
    assert g.allow()
    assert not g.allow()


def test_parse_checkpoints():
    assert isinstance(parse_checkpoints("off"), NoCheckpoints)
    assert parse_checkpoints("7").interval == 7
    assert parse_checkpoints("5m").interval == datetime.timedelta(minutes=5)
    for bad in ("", "5s", "m", "-1"):
        with pytest.raises(ValueError):
            parse_checkpoints(bad)


def test_iter_info():
    assert IterInfo.make(dt=0.1, dt_output=1.0, count=5) == IterInfo(10, 5)
    assert IterInfo.make(dt=0.1, t_max=10.0, count=5) == IterInfo(20, 5)
    assert IterInfo.make(dt=0.5, dt_output=2.0, t_max=10.0) == IterInfo(4, 5)
    with pytest.raises(ValueError):
        IterInfo.make(dt=0.1, dt_output=1.0)


@pytest.mark.gpu
def test_run_sim_batches_checkpoints_and_nan(tmp_path):
    import boltzmann_amd
    from boltzmann_amd.core import bgk
    sim = boltzmann_amd.Simulation("gpu", [64, 32], bgk(0.6))
    sim.fluid.vel[..., 0] = 0.02
    seen = []
    for i in run_sim(sim, IterInfo(interval=9, count=4), tmp_path, EveryN(1)):
        seen.append((i, sim.iteration))
    assert seen == [(0, 0), (1, 10), (2, 20), (3, 30)]   # 9 rounds up to an even 10
    assert os.path.exists(tmp_path / "checkpoint.mpk")
    assert not os.path.exists(tmp_path / "checkpoint.mpk.tmp")
    sim2 = boltzmann_amd.Simulation.load_checkpoint("gpu", str(tmp_path / "checkpoint.mpk"))
    assert sim2.iteration == 30
    # resume: starts at the stored batch index
    first = next(run_sim(sim2, IterInfo(interval=10, count=5), tmp_path, NoCheckpoints()))
    assert first == 3
    bad = boltzmann_amd.Simulation("gpu", [16, 16], bgk(0.6))
    bad.fluid.vel[3, 3, 0] = np.nan
    gen = run_sim(bad, IterInfo(interval=2, count=3), tmp_path, NoCheckpoints())
    next(gen)
    with pytest.raises(ValueError, match="Non-finite"):
        next(gen)
