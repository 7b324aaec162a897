"""The CPU-baseline timing protocol of bench.py (host logic, on a simulated clock).

A synthetic workload whose time per iteration falls from 4 ms to 2 ms once the
simulated clock passes `drop_at` seconds stands for a host that is still settling
(clocks, page placement). The timed runs must measure the settled rate whenever the drop
happens within the warm-up or the settle phase."""
import pytest

import bench


class FakeHost:
  def __init__(self, drop_at, slow=4e-3, fast=2e-3):
    self.t = 0.0
    self.drop_at = drop_at
    self.slow, self.fast = slow, fast
    self.iterations = 0

  def clock(self):
    return self.t

  def step(self, k):
    for _ in range(k):
      self.t += self.slow if self.t < self.drop_at else self.fast
      self.iterations += 1


def run(monkeypatch, host, **kw):
  monkeypatch.setattr(bench.time, "perf_counter", host.clock)
  before = []
  orig = host.step

  def step(k):
    orig(k)
    before.append(k)

  prot = bench.cpu_protocol(step, **kw)
  return prot, before


@pytest.mark.parametrize("drop_at", [0.5, 5.0, 9.9])
def test_drop_inside_warmup_is_not_timed(monkeypatch, drop_at):
  host = FakeHost(drop_at)
  prot, _ = run(monkeypatch, host, warm_s=10.0, run_s=3.0, max_warm=1_000_000)
  per_it = prot["median_s"] / prot["iterations_per_run"]
  assert per_it == pytest.approx(2e-3, rel=1e-6)
  assert max(prot["run_s"]) == pytest.approx(min(prot["run_s"]), rel=1e-3)
  assert prot["warmup_s"] >= 10.0


def test_drop_after_warmup_is_caught_by_the_settle_phase(monkeypatch):
  host = FakeHost(drop_at=11.5)
  prot, _ = run(monkeypatch, host, warm_s=10.0, run_s=3.0, max_warm=1_000_000)
  assert prot["median_s"] / prot["iterations_per_run"] == pytest.approx(2e-3, rel=1e-6)
  assert max(prot["run_s"]) == pytest.approx(min(prot["run_s"]), rel=1e-3)


def test_without_settle_phase_the_ramp_is_timed(monkeypatch):
  # the same host with the settle phase off: the first timed run still sees the slow rate
  host = FakeHost(drop_at=11.5)
  prot, _ = run(monkeypatch, host, warm_s=10.0, run_s=3.0, max_warm=1_000_000, settle_s=0.0)
  assert max(prot["run_s"]) > 1.1 * min(prot["run_s"])


def test_warmup_count_and_run_lengths(monkeypatch):
  host = FakeHost(drop_at=0.0)
  prot, calls = run(monkeypatch, host, warm_s=2.0, run_s=1.0, max_warm=100)
  runs = prot["runs"]
  timed = calls[-runs:]
  assert all(k == prot["iterations_per_run"] for k in timed)
  assert sum(calls[:-runs]) == prot["warmup_iterations"]
  assert 499 <= prot["iterations_per_run"] <= 500   # about run_s at 2 ms per iteration
  assert 100 <= prot["warmup_iterations"] <= 100 + 1 + 500 + 1


def test_slow_iterations_keep_the_iteration_cap(monkeypatch):
  # one iteration takes longer than the warm-up budget: one calibration call, at most
  # max_warm warm-up iterations, runs of one iteration
  host = FakeHost(drop_at=1e9, slow=3.0)
  prot, calls = run(monkeypatch, host, warm_s=1.0, run_s=1.0, max_warm=2, settle_s=0.0)
  assert prot["iterations_per_run"] == 1
  assert prot["warmup_iterations"] <= 1 + 2 + 1
