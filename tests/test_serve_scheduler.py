"""Host logic of the serving loop (no GPU): admission against the memory
budget, grouping of compatible operations across jobs, arrival order, groups
that advance independently and the latency statistics, with a CPU executor
standing in for the kernels."""
import threading
import time

import pytest

import numpy as np

from ecs_mhd_2d_amd import serve


def toy_job(n, steps, value, traj=64):
    """A job that asks for `steps` 'grad' ops (trajectories of `traj` steps)
    and sums what it receives."""
    total = 0.0
    for _ in range(steps):
        r = yield serve.Op("grad", n, steps=traj, x=value)
        total += r
    return total


class Exec:
    def __init__(self):
        self.calls = []
        self.lock = threading.Lock()

    def __call__(self, key, ops, lane):
        with self.lock:
            self.calls.append((key, len(ops), lane))
        return [2.0 * o.kw["x"] for o in ops]


def arrivals(k, mem=10.0, gap=0.0, steps=3):
    return [(gap * i, "adjoint", (lambda i=i: toy_job(32 if i % 2 else 64, steps, float(i))), mem, 32 if i % 2 else 64)
            for i in range(k)]


def test_results_and_batching():
    ex = Exec()
    done, stats = serve.Server(executor=ex).run(arrivals(6))
    assert stats["jobs"] == 6
    res = {j.id: j.result for j in done}
    for i in range(6):
        assert res[i] == 3 * 2.0 * i
    # all six jobs arrive at t=0: two keys (one per grid size), their batches
    # halved until the four lanes are busy
    assert {c[0] for c in ex.calls} == {("grad", 32), ("grad", 64)}
    assert sorted(c[1] for c in ex.calls[:4]) == [1, 1, 2, 2]
    assert sorted(c[2] for c in ex.calls[:4]) == [0, 1, 2, 3]
    assert stats["mean_batch"] > 1.0


def test_gradient_jobs_of_different_segment_counts_batch_together():
    """Trajectories of whole 64-step segments batch across step counts (each
    candidate runs its own count); other counts batch only with equal ones."""
    ex = Exec()
    arr = [(0.0, "adjoint", (lambda k=k: toy_job(32, 2, 1.0, traj=k)), 1.0, 32) for k in (64, 128, 192, 100, 100, 90)]
    serve.Server(executor=ex, max_groups=3).run(arr)   # three keys, three lanes: no batch is halved
    keys = sorted({(c[0], c[1]) for c in ex.calls}, key=str)
    assert keys == sorted({(("grad", 32), 3), (("grad", 32, 100), 2), (("grad", 32, 90), 1)}, key=str)


def test_admission_budget_limits_residency():
    ex = Exec()
    done, stats = serve.Server(hbm_budget=25.0, executor=ex).run(arrivals(6, mem=10.0))
    assert stats["jobs"] == 6
    assert stats["peak_resident_bytes"] <= 25.0
    assert max(c[1] for c in ex.calls) <= 1      # two resident jobs of different N: never batched together
    assert stats["ticks"] > 6


def test_latency_percentiles_follow_arrivals():
    ex = Exec()
    done, stats = serve.Server(executor=ex).run(arrivals(4, gap=0.05, steps=1))
    lat = sorted(j.finish - j.arrival for j in done)
    assert stats["p50_latency_s"] == np.percentile(lat, 50)
    assert stats["p99_latency_s"] <= max(lat) + 1e-12
    assert all(j.finish >= j.arrival for j in done)


def test_max_batch_splits_groups_into_lanes():
    ex = Exec()
    srv = serve.Server(executor=ex, max_batch=2, max_groups=3)
    done, stats = srv.run([(0.0, "adjoint", (lambda i=i: toy_job(64, 1, float(i))), 1.0, 64) for i in range(5)])
    assert stats["jobs"] == 5
    sizes = sorted(c[1] for c in ex.calls)
    assert sizes == [1, 2, 2]
    assert sorted(c[2] for c in ex.calls) == [0, 1, 2]   # concurrent tasks on one N get distinct lanes


def failing_job(n, fail_at):
    for k in range(3):
        r = yield serve.Op("grad", n, steps=64, x=1.0)
        if k == fail_at:
            raise FloatingPointError("non-finite matvec")
    return r


def test_a_failing_job_ends_alone():
    """A job whose own code raises is recorded as failed; the others finish
    and the latency statistics cover the completed jobs only."""
    ex = Exec()
    arr = arrivals(4) + [(0.0, "newton", (lambda: failing_job(32, 1)), 10.0, 32)]
    done, stats = serve.Server(executor=ex).run(arr)
    assert stats["jobs"] == 5 and stats["failed_jobs"] == 1
    bad = [j for j in done if j.error is not None]
    assert len(bad) == 1 and "non-finite" in bad[0].error and bad[0].result is None
    res = {j.id: j.result for j in done if j.error is None}
    assert sorted(res.values()) == [0.0, 6.0, 12.0, 18.0]
    assert np.isfinite(stats["p50_latency_s"])


@pytest.mark.timeout(60)
def test_a_job_larger_than_the_budget_is_rejected():
    """A job whose memory exceeds the whole budget ends at once with an error
    (it would never fit), and the jobs behind it are served."""
    ex = Exec()
    arr = arrivals(3, mem=10.0) + [(0.0, "newton", (lambda: toy_job(64, 2, 1.0)), 30.0, 64)]
    done, stats = serve.Server(hbm_budget=25.0, executor=ex).run(arr)
    assert stats["jobs"] == 4 and stats["failed_jobs"] == 1
    bad = [j for j in done if j.error is not None]
    assert len(bad) == 1 and "exceeds the HBM budget" in bad[0].error and bad[0].mem == 30.0
    assert sorted(j.result for j in done if j.error is None) == [0.0, 6.0, 12.0]


class SlowExec(Exec):
    def __init__(self, delay):
        super().__init__()
        self.delay = delay

    def __call__(self, key, ops, lane):
        time.sleep(self.delay.get(key[1], 0.0))
        return super().__call__(key, ops, lane)


@pytest.mark.timeout(60)
def test_smaller_jobs_pass_a_waiting_large_one():
    """Budget 25: a resident job of 20 keeps a second job of 20 waiting; a job
    of 5 that arrives later is admitted in the meantime."""
    ex = SlowExec({64: 0.05})
    arr = [(0.0, "adjoint", (lambda: toy_job(64, 4, 1.0)), 20.0, 64),
           (0.0, "adjoint", (lambda: toy_job(64, 1, 2.0)), 20.0, 64),
           (0.01, "adjoint", (lambda: toy_job(32, 1, 3.0)), 5.0, 32)]
    done, stats = serve.Server(hbm_budget=25.0, executor=ex).run(arr)
    by = {j.id: j for j in done}
    assert stats["failed_jobs"] == 0 and stats["peak_resident_bytes"] <= 25.0
    assert by[2].start < by[1].start and by[2].finish < by[1].start


@pytest.mark.timeout(60)
def test_groups_advance_independently():
    """A slow group (N=64, 0.5 s per operation) does not hold back a fast job
    on another grid size: the fast job's five operations all run while the
    slow one is in flight."""
    ex = SlowExec({64: 0.5})
    arr = [(0.0, "adjoint", (lambda: toy_job(64, 1, 1.0)), 1.0, 64),
           (0.0, "adjoint", (lambda: toy_job(32, 5, 2.0)), 1.0, 32)]
    done, stats = serve.Server(executor=ex).run(arr)
    by = {j.id: j for j in done}
    assert by[1].result == 20.0 and by[0].result == 2.0
    assert by[1].finish < by[0].finish and by[1].finish - by[1].arrival < 0.4


@pytest.mark.timeout(60)
def test_group_slots_bound_concurrency():
    """At most max_groups groups run at once."""
    running, peak = [0], [0]
    lock = threading.Lock()

    def ex(key, ops, lane):
        with lock:
            running[0] += 1
            peak[0] = max(peak[0], running[0])
        time.sleep(0.02)
        with lock:
            running[0] -= 1
        return [o.kw["x"] for o in ops]

    arr = [(0.0, "adjoint", (lambda i=i: toy_job(16 * (i + 1), 2, 1.0)), 1.0, 16 * (i + 1)) for i in range(6)]
    done, stats = serve.Server(executor=ex, max_groups=2).run(arr)
    assert stats["failed_jobs"] == 0 and peak[0] == 2


class StepsExec(Exec):
    def __call__(self, key, ops, lane):
        with self.lock:
            self.calls.append((key, sorted(o.kw["steps"] for o in ops), lane))
        return [2.0 * o.kw["x"] for o in ops]


def test_batches_keep_operations_of_similar_length_together():
    """Two lanes, four gradient jobs of one key with trajectories of 64..256
    steps: the batch is halved into its two shorter and its two longer
    members (a batch runs as long as its longest member)."""
    ex = StepsExec()
    arr = [(0.0, "adjoint", (lambda k=k: toy_job(32, 1, 1.0, traj=k)), 1.0, 32) for k in (256, 64, 192, 128)]
    done, stats = serve.Server(executor=ex, max_groups=2).run(arr)
    assert stats["failed_jobs"] == 0
    assert sorted(c[1] for c in ex.calls) == [[64, 128], [192, 256]]


@pytest.mark.timeout(60)
def test_free_lanes_take_single_jobs_of_one_kind():
    """Four jobs of one key and four lanes: each runs in a lane of its own and
    advances on its own (a job that finishes an operation starts the next at
    once instead of waiting for the others)."""
    ex = SlowExec({32: 0.02})
    arr = [(0.0, "adjoint", (lambda i=i: toy_job(32, 3, float(i))), 1.0, 32) for i in range(4)]
    done, stats = serve.Server(executor=ex).run(arr)
    assert stats["failed_jobs"] == 0 and stats["ticks"] == 12
    assert all(c[1] == 1 for c in ex.calls)
    assert sorted({c[2] for c in ex.calls}) == [0, 1, 2, 3]


def test_line_search_window_statistics_count_the_one_at_a_time_tries():
    """A Newton job's line-search window (one "objective" operation holding
    several tries) counts, in the served steps, only the tries the
    one-at-a-time search makes: up to and including the first completed try
    whose |f| is below the job's f_norm, all of them when none is."""
    import torch

    def out(scale, attempts, completed=True):
        f = {"fields": torch.full((2, 4, 4), scale, dtype=torch.float64), "T": torch.tensor(0.0),
             "sx": torch.tensor(0.0)}
        return f, {"attempts": attempts, "completed": completed}

    trials = [{"fields": None, "T": t, "sx": 0.0} for t in (1.0, 2.0, 3.0, 4.0)]
    op = serve.Op("objective", 4, trials=trials, f_norm=10.0, adaptive={"atol": 1e-8, "max_steps_per_checkpoint": 8,
                                                                       "checkpoints": 2})
    # |f| of a full((2,4,4), a) field is a * sqrt(32) = 5.66 a
    res = [out(3.0, 100), out(1.5, 110, completed=False), out(1.5, 120), out(0.1, 130)]
    assert serve.useful_trials(op, res) == 3
    assert serve.op_rk_steps(op, res) == 100 + 110 + 120
    none = [out(3.0, 100), out(2.0, 110), out(2.5, 120), out(4.0, 130)]
    assert serve.useful_trials(op, none) == 4 and serve.op_rk_steps(op, none) == 460
    assert op.length() == 4.0
