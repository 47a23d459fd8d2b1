"""Concurrent reductions from two host threads on two streams.

The library's reduction scratch is owned per (device, stream), so calls that
run at the same time on different streams must not see each other's partial
sums.  Two host threads, each with its own HIP stream (lsm_amd_set_stream is
per host thread), repeat the CFL-dt and smoothed-volume reductions on
different fields; every result must equal the serial result bit for bit.
"""
import threading

import numpy as np
import pytest
import torch

from lsmlib_amd import toolbox

pytestmark = pytest.mark.gpu


def _fields(seed, n=96):
    rng = np.random.default_rng(seed)
    shape = (n + 6, n + 6, n + 6)
    vel = tuple(torch.from_numpy(rng.uniform(-1.0, 1.0, size=shape)).cuda().contiguous() for _ in range(3))
    phi = torch.from_numpy(rng.uniform(-0.2, 0.2, size=shape)).cuda().contiguous()
    gb = (-3, n + 2, -3, n + 2, -3, n + 2)
    ib = (0, n - 1, 0, n - 1, 0, n - 1)
    return vel, phi, gb, ib


def _reduce(lib, case):
    vel, phi, gb, ib = case
    h = 1.0 / 96
    dt = lib.lsm3dcomputestableadvectiondt_(*vel, gb, ib, h, h, h, 0.5)
    vol = lib.lsm3dvolumeregionphilessthanzero_(phi, gb, ib, h, h, h, 1.5 * h)
    return dt, vol


def test_two_threads_on_two_streams_match_serial_results():
    lib = toolbox("double")
    cases = [_fields(1), _fields(2)]
    serial = [_reduce(lib, c) for c in cases]
    assert serial[0] != serial[1]
    results = [[], []]
    errors = []

    def worker(q):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                lib.set_stream(s.cuda_stream)
                for _ in range(40):
                    results[q].append(_reduce(lib, cases[q]))
                lib.set_stream(torch.cuda.default_stream().cuda_stream)
        except Exception as e:  # reported in the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(q,)) for q in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for q in range(2):
        assert len(results[q]) == 40
        assert all(r == serial[q] for r in results[q]), (q, set(results[q]), serial[q])
