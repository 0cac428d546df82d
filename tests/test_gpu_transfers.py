"""Host <-> device copies of caller arrays.

The library moves a caller's (pageable) array through two pinned 4 MiB
buffers, host threads staging one chunk while the DMA engine moves the other,
with the renumbering to library order on the device.  These checks round-trip
states bitwise: a state of one chunk and one of six (the 1M-cell channel,
not a multiple of the chunk), and two contexts driven from two host threads
at once (the pinned buffers are shared by the process, one transfer at a
time).
"""
import sys
import threading

import numpy as np
import pytest

from hgswe import capi, cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("cells", [20_000, 1_000_000])
def test_state_round_trip_is_bitwise(cells):
    c = cases.channel(cells, seed=3, nsteps=10)
    assert c.N > 0
    ctx = capi.Context(c)
    rng = np.random.default_rng(cells)
    Q = rng.normal(size=3 * c.N)
    ctx.upload(Q)
    back = ctx.download()
    say(f"  {c.N} cells: {3 * c.N * 8 / 2**20:.1f} MiB each way, bitwise {bool(np.array_equal(back, Q))}")
    assert np.array_equal(back, Q)
    # the RHS of a host array equals the RHS of the same state stepped on the device
    r = ctx.rhs(c.Q0)
    ctx.upload(c.Q0)
    r2 = ctx.rhs(ctx.download())
    assert np.array_equal(r, r2)


def test_two_threads_two_contexts():
    cs = [cases.channel(30_000, seed=s, nsteps=10) for s in (4, 5)]
    ctxs = [capi.Context(c) for c in cs]
    want = [k.step(c.Q0, c.dt, 5) for k, c in zip(ctxs, cs)]
    errs = []

    def run(i):
        try:
            for _ in range(20):
                got = ctxs[i].step(cs[i].Q0, cs[i].dt, 5)
                if not np.array_equal(got, want[i]):
                    errs.append(f"context {i}: results differ")
                    return
        except Exception as e:  # reported below
            errs.append(f"context {i}: {e!r}")

    th = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th)
    assert not errs, errs
