"""The no-allocation CFL entry point (hg_cfl_dt_noalloc) against hg_cfl_dt
through the C ABI: bitwise the same dt on two calls in a row, the second on
a smaller mesh (the static scratch keeps the first call's partial minima
beyond the second's; none of them may leak into its result), then the first
mesh again.  States carry velocities and dry cells."""
import sys

import numpy as np
import pytest

from hgswe import capi, cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _state(c, seed):
    rng = np.random.default_rng(seed)
    N = c.N
    xi = c.Q0[:N] + rng.uniform(-0.2, 0.2, N)
    dry = rng.choice(N, max(1, N // 50), replace=False)
    xi[dry] = 0.5 * c.h_small - c.hstill[dry]
    return np.concatenate([xi, rng.normal(0.0, 1.5, N), rng.normal(0.0, 1.5, N)])


def test_noalloc_cfl_matches_cfl_dt_twice_second_smaller():
    big = cases.channel(400_000, seed=21, nsteps=1)
    small = cases.channel(6_000, seed=22, Lx=1_500.0, Ly=150.0, nsteps=1)
    cb, cs = capi.Context(big), capi.Context(small)
    Qb, Qs = _state(big, 1), _state(small, 2)
    say(f"  no-alloc CFL: big N={big.N}, small N={small.N}")
    for ctx, Q, name in ((cb, Qb, "big"), (cs, Qs, "small"), (cb, Qb, "big again")):
        for cfl in (0.5, 0.9):
            ref = ctx.cfl_dt(Q, cfl)
            got = ctx.cfl_dt_noalloc(Q, cfl)
            say(f"  {name} cfl={cfl}: hg_cfl_dt {ref!r} no-alloc {got!r}")
            assert got == ref and np.isfinite(got) and got > 0.0
    # the small mesh's dt is not the big one's (a stale partial would show)
    assert cs.cfl_dt_noalloc(Qs, 0.5) != cb.cfl_dt_noalloc(Qb, 0.5)
    # the state it leaves on the device is Q, as hg_cfl_dt leaves it
    assert np.array_equal(cs.download(), Qs)
