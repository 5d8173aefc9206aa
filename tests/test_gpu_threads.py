"""The host C ABI is re-entrant like the reference's (its interfaces are
declared pure and may be called from `do concurrent`): calls made at the same
time from several host threads give the same bits as the same calls made one
at a time. ctypes releases the GIL around each foreign call, so the calls
below really overlap (each thread has its own stream and staging buffers)."""
import threading

import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen

pytestmark = pytest.mark.gpu


def _run_concurrently(jobs):
    out = [None] * len(jobs)
    err = []
    start = threading.Barrier(len(jobs))

    def work(i):
        try:
            start.wait()
            out[i] = jobs[i]()
        except Exception as e:  # reported below, with the job index
            err.append((i, e))

    ts = [threading.Thread(target=work, args=(i,)) for i in range(len(jobs))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not err, err
    return out


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def test_concurrent_matmul_elementwise_and_sum(gpu):
    rng = np.random.default_rng(21)
    shapes = [(70, 33, 50), (128, 64, 96), (5, 300, 9), (200, 17, 130)]
    mats = [(datagen.matrix(rng, m, k, 10), datagen.matrix(rng, k, n, 10)) for m, k, n in shapes]
    vecs = [datagen.exp_inputs(rng, 50000 + 1000 * i) for i in range(4)]
    sums = [datagen.spread(rng, 70000 + 3 * i, 20) for i in range(4)]
    jobs = []
    for (A, B), x, s in zip(mats, vecs, sums):
        jobs.append(lambda A=A, B=B: mf.matmul(A, B, 8))
        jobs.append(lambda x=x: mf.exp(x))
        jobs.append(lambda s=s: mf.sum(s))
    alone = [j() for j in jobs]
    for rep in range(3):
        together = _run_concurrently(jobs)
        for i, (a, b) in enumerate(zip(alone, together)):
            assert np.array_equal(_bits(a), _bits(b)), (rep, i)


def test_concurrent_ozaki_keeps_per_thread_slice_counts(gpu):
    """Both products forms at once; each thread reads back its own call's
    slice count."""
    rng = np.random.default_rng(22)
    cases = [(96, 64, 80, 10, 12), (520, 70, 515, 10, 13), (64, 40, 70, 0, 0), (512, 48, 512, 20, 0)]
    data = [(datagen.matrix(rng, m, k, e or None), datagen.matrix(rng, k, n, e or None), S)
            for m, k, n, e, S in cases]

    def job(A, B, S):
        C = mf.matmul_ozaki(A, B, S)
        return C, mf.ozaki_slices_used()

    jobs = [lambda A=A, B=B, S=S: job(A, B, S) for A, B, S in data]
    alone = [j() for j in jobs]
    for rep in range(3):
        together = _run_concurrently(jobs)
        for i, ((ca, sa), (cb, sb)) in enumerate(zip(alone, together)):
            assert sa == sb, (rep, i, sa, sb)
            assert np.array_equal(_bits(ca), _bits(cb)), (rep, i)


def test_short_lived_threads_release_their_staging(gpu):
    """Every thread that calls a host-pointer entry point gets its own device
    buffers and stream; they are freed when the thread exits. 48 short-lived
    threads that each stage about 20 MB must leave the device's free memory
    where it was (a leak would hold about 1 GB)."""
    import torch
    torch.cuda.init()
    rng = np.random.default_rng(31)
    m = k = n = 512
    A = datagen.matrix(rng, m, k, 10)
    B = datagen.matrix(rng, k, n, 10)
    want = _bits(mf.matmul(A, B, 8))
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for wave in range(6):
        jobs = [lambda: mf.matmul(A, B, 8) for _ in range(8)]
        for got in _run_concurrently(jobs):
            assert np.array_equal(_bits(got), want)
    free1, _ = torch.cuda.mem_get_info()
    assert free0 - free1 < 256 << 20, (free0 - free1) / 2 ** 20
