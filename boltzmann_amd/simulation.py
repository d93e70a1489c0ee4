"""Batch driver around ``Simulation.iterate``: output cadence, MLUPS log, NaN check and
atomic checkpoints -- the caller side of the update path in the reference
(python/boltzmann/simulation.py:30-271).

``run_sim`` yields once before stepping (output 0, or the resumed index), then after each
batch of ``interval`` steps (rounded to an even count, simulation.py:211).  After every
batch it raises ``ValueError`` on a non-finite velocity (simulation.py:241-242), logs the
batch throughput in MLUP/s measured around ``iterate`` (simulation.py:244-246), and, when the
gater allows it, writes ``checkpoint.mpk`` through a temporary file and ``os.replace``
(simulation.py:253-258).
"""

from __future__ import annotations

import datetime
import logging
import os
import re
import time
from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from pathlib import Path
from typing import Generator


logger = logging.getLogger(__name__)


@dataclass
class IterInfo:
    """LBM steps per output (``interval``) and number of outputs (``count``)."""

    interval: int
    count: int

    @staticmethod
    def make(*, dt: float, dt_output: float | None = None, t_max: float | None = None,
             count: int | None = None) -> "IterInfo":
        """From ``dt`` and exactly two of ``dt_output``, ``t_max``, ``count``
        (simulation.py:44-80)."""
        given = (dt_output is not None, t_max is not None, count is not None)
        if given == (False, True, True):
            dt_output = t_max / count
        elif given == (True, False, True):
            pass
        elif given == (True, True, False):
            count = int(t_max / dt_output + 1e-8)
        else:
            raise ValueError("Must specify dt and exactly two other arguments! "
                             f"[{dt_output=}, {t_max=}, {count=}]")
        return IterInfo(int(dt_output / dt), count)


class CheckpointGater(ABC):
    """Decides, once per batch, whether ``run_sim`` writes a checkpoint."""

    @abstractmethod
    def allow(self) -> bool: ...


@dataclass
class NoCheckpoints(CheckpointGater):
    def allow(self) -> bool:
        return False


@dataclass
class EveryN(CheckpointGater):
    """Every ``interval``-th call, never the first (simulation.py:105-124)."""

    interval: int
    _curr: int = 1

    def __post_init__(self):
        if self.interval <= 0:
            raise ValueError("Checkpoint Interval must be positive!")

    def allow(self) -> bool:
        self._curr = (self._curr + 1) % self.interval
        return self._curr == 0


@dataclass
class EveryT(CheckpointGater):
    """When more than ``interval`` has passed since the last checkpoint (or creation)."""

    interval: datetime.timedelta
    _last: datetime.datetime = field(default_factory=datetime.datetime.now)

    def allow(self) -> bool:
        now = datetime.datetime.now()
        if now - self._last > self.interval:
            self._last = now
            return True
        return False


def parse_checkpoints(s: str) -> CheckpointGater:
    """``"off"``, ``"<N>"`` (every N batches) or ``"<N>m"`` (every N minutes)."""
    if s == "off":
        return NoCheckpoints()
    if (m := re.fullmatch(r"(\d+)", s)) is not None:
        return EveryN(int(m.group(1)))
    if (m := re.fullmatch(r"(\d+)m", s)) is not None:
        return EveryT(datetime.timedelta(minutes=int(m.group(1))))
    raise ValueError(f"Unable to parse checkpoint interval! [{s!r}]")


def run_sim(sim, meta: IterInfo, output_dir: Path | str,
            checkpoints: CheckpointGater | None = None) -> Generator[int, None, None]:
    """Steps ``sim`` in batches of ``meta.interval`` (made even) up to ``meta.count`` outputs,
    yielding the output index before the first batch and after each one.  After each batch
    a non-finite velocity raises ValueError, as in simulation.py:241-242; the values are
    counted on the device rather than by a pass over the host array."""
    output_dir = Path(output_dir)
    if output_dir.exists():
        if not output_dir.is_dir():
            raise ValueError(f"Base path must be a directory! [{output_dir=}]")
    else:
        output_dir.mkdir()
    checkpoints = checkpoints or EveryN(5)
    batch_iters = 2 * int(meta.interval / 2 + 0.5)
    logger.info("device %s, %d MB, %d iters / output, %.1fM cells (%s)",
                sim.device_info, sim.size_bytes // 1_000_000, batch_iters,
                sim.cells.count / 1e6, " x ".join(str(int(c)) for c in sim.cells.counts))
    start = sim.iteration // batch_iters if batch_iters else 0
    yield start
    t_all = time.perf_counter()
    for i in range(start + 1, meta.count):
        t0 = time.perf_counter()
        sim.iterate(batch_iters)
        if sim.count_nonfinite("vel"):
            raise ValueError("Non-finite value detected!")
        secs = time.perf_counter() - t0
        mlups = sim.cells.count * batch_iters / (1e6 * secs)
        t_out = time.perf_counter()
        yield i
        if checkpoints.allow():
            tmp = output_dir / "checkpoint.mpk.tmp"
            sim.write_checkpoint(str(tmp))
            os.replace(tmp, output_dir / "checkpoint.mpk")
            logger.info("Checkpoint written.")
        out_secs = time.perf_counter() - t_out
        per = (time.perf_counter() - t_all) / (i - start)
        eta = datetime.datetime.now() + datetime.timedelta(seconds=per * (meta.count - i))
        logger.info("Batch %d: %6.0f MLUP/s, sim %.1fs, out %.1fs, eta %s", i, mlups, secs,
                    out_secs, f"{eta:%H:%M:%S}")
