"""Sample collection through a data collector (io.py:502-771), CPU only: the reference's
own doctest (io.py:650-672: every second of five iterations kept, three samples), the
npz output of MemoryCollector with its leaf names, and solver.mcmc handing its kept
samples to the collector."""
import numpy as np
import torch

import __graft_entry__

__graft_entry__.load()
from jax_sgmc_amd import io, scheduler, solver  # noqa: E402


def test_memory_collector_doctest():
  dc = io.MemoryCollector()
  init_save, save, postprocess_save = io.save(dc)
  static = scheduler.static_information(samples_collected=3)
  state = init_save({"iteration": torch.tensor(0, dtype=torch.int32)}, {}, static)
  for it in range(5):
    state = save(state, it % 2 == 0, {"iteration": torch.tensor(it, dtype=torch.int32)})
  out = postprocess_save(state, None)
  assert out["sample_count"] == 3
  assert out["samples"]["iteration"].dtype == np.int32
  np.testing.assert_array_equal(out["samples"]["iteration"], [0, 2, 4])


def test_memory_collector_writes_npz(tmp_path):
  dc = io.MemoryCollector(save_dir=str(tmp_path))
  init_save, save, post = io.save(dc)
  sample = {"w": torch.zeros(3), "nested": {"b": torch.zeros(2, 2)}}
  st0 = init_save(sample, {}, scheduler.static_information(samples_collected=2))
  st1 = init_save(sample, {}, scheduler.static_information(samples_collected=1))
  for k in range(2):
    st0 = save(st0, True, {"w": torch.full((3,), float(k)), "nested": {"b": torch.eye(2) * k}})
  st1 = save(st1, True, {"w": torch.ones(3), "nested": {"b": torch.ones(2, 2)}})
  out0, out1 = post(st0), post(st1)
  f = np.load(tmp_path / "chain_0.npz")
  assert sorted(f.files) == ["nested/b", "w"]
  np.testing.assert_array_equal(f["w"], [[0, 0, 0], [1, 1, 1]])
  np.testing.assert_array_equal(out0["samples"]["nested"]["b"][1], np.eye(2))
  np.testing.assert_array_equal(np.load(tmp_path / "chain_1.npz")["w"], [[1, 1, 1]])
  assert out1["sample_count"] == 1


def test_mcmc_hands_kept_samples_to_the_collector():
  # a solver whose state is the iteration number, run through solver.mcmc with a burn-in
  # of 3 of 8 iterations
  sv = (lambda s: s, lambda s, sched: (s + 1, None), lambda s: {"x": torch.tensor(float(s))})
  sched = scheduler.init_scheduler(burn_in=scheduler.initial_burn_in(3))
  dc = io.MemoryCollector()
  run = solver.mcmc(sv, sched, saving=io.save(dc))
  out = run(0, iterations=8)[0]
  assert out["sample_count"] == 5
  np.testing.assert_array_equal(out["samples"]["x"], [4.0, 5.0, 6.0, 7.0, 8.0])
