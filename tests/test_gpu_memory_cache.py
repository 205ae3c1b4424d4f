"""The engine's block cache (mm_memory_cached / mm_memory_trim,
include/mumax_engine.h): freeing a simulation or a batch reuses its memory
later and does not wait for work other batches have queued on their own
streams (hipFree would wait for the whole device)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def film(mm, n, seed):
    return mm.workloads.make_sim(dict(
        size=(1, n, n), cellsize=(3e-9, 3e-9, 3e-9), msat=8e5, aexch=1.3e-11, alpha=0.02, solver="rkdp",
        maxerror=1e-5, field=("static", (0.0, 0.01, 0.02)),
        m0=mm.workloads.normalize(np.random.default_rng(seed).standard_normal((3, 1, n, n))).astype(np.float32)))


def test_freed_memory_is_reused_and_trimmed(mm):
    lib = mm.load()
    lib.mm_memory_trim()
    a = film(mm, 64, 1)
    a.steps(2)
    a.free()
    held = lib.mm_memory_cached()
    assert held > 0
    b = film(mm, 64, 2)  # the same sizes: served from the cache
    b.steps(2)
    assert lib.mm_memory_cached() < held
    b.free()
    assert lib.mm_memory_trim() >= held
    assert lib.mm_memory_cached() == 0


def test_free_does_not_wait_for_other_batches(mm):
    """Batch A has a long stretch of trials queued on its stream; meanwhile a
    second batch and a lone simulation are built, run and freed.  A is still
    running afterwards, so none of those frees waited for it; A's members then
    finish with the same m as when run alone."""
    lib = mm.load()
    big = [film(mm, 512, 10 + j) for j in range(2)]
    ref = []
    for j in range(2):
        s = film(mm, 512, 10 + j)
        s.steps(1000)
        ref.append(s.m)
        s.free()
    A = mm.Batch(big)
    A.start(steps=[1000, 1000])
    A.trials(1)             # captures the batched trial
    A.enqueue(1000)         # about a second of work on A's stream
    small = [film(mm, 32, 20 + j) for j in range(3)]
    B = mm.Batch(small)
    B.start(steps=[5, 5, 5])
    B.trials(8)
    B.free()
    for s in small:
        s.free()
    c = film(mm, 48, 30)
    c.steps(3)
    c.free()
    busy = not A.ready()
    print("batch A still running after the frees: %s; cache holds %d bytes" % (busy, lib.mm_memory_cached()))
    assert busy
    while not A.flags().all():
        A.enqueue(100)      # trials of done members are skipped
    A.free()
    for s, m in zip(big, ref):
        assert np.array_equal(s.m, m)
        s.free()
