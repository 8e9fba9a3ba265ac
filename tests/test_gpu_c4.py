"""The uneven batch at its stated size: the 96 seeded C4 instances (meshes of
200x100 .. 1600x400 elements, 1-4 layers, each isotropic or tilted-axis
hexagonal up to 6%, their own ray parameter and polarisation) admitted into
one engine and stepped together, one persistent launch per step, for 25
steps.  The reference runs each of these as its own process
(applications/test/prem_epi83.pl:81-82); batching must not change any
result: the smallest, the largest and a tilted-axis member are bit-identical
(traces and final field) to the same instance run alone, and every trace of
every member is finite."""
import numpy as np
import pytest

import semsks
from semsks import configs

pytestmark = pytest.mark.gpu

NSTEP = 25


def admit(e, b, nt):
    prob = b.problem(nt)
    prob.nst = NSTEP + 1
    e.planewave(*b.planewave_args(nt), to_host=False)
    return e.admit(prob)


def same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_c4_uneven_batch_members_equal_solo_runs():
    print("case: C4, 96 instances 200x100..1600x400, %d steps" % NSTEP, flush=True)
    cases = configs.c4_cases(96)
    built = [configs.Built(c) for c in cases]
    npts = [b.nglob() for b in built]
    tilted = [k for k, c in enumerate(cases) if any(a is not None and a[1] is not None for a in c.aniso)]
    assert tilted, "the seeded C4 set holds tilted-axis layers"
    picks = {"smallest": int(np.argmin(npts)), "largest": int(np.argmax(npts)), "tilted": tilted[0]}
    assert min(npts) < 400_000 and max(npts) > 8_000_000
    nt = NSTEP + 8
    with semsks.Engine(max_slots=96) as e:
        slots = [admit(e, b, nt) for b in built]
        e.advance(NSTEP)
        print("batch stepped, tiles of %d rows" % e.tile_rows(), flush=True)
        for k, s in enumerate(slots):
            tr = e.traces(s)
            assert np.isfinite(tr).all(), k
        batch = {name: (e.traces(slots[k]), e.field(slots[k])) for name, k in picks.items()}
    for name, k in picks.items():
        with semsks.Engine() as e:
            s = admit(e, built[k], nt)
            e.advance(NSTEP)
            tr, f = e.traces(s), e.field(s)
        assert np.abs(f).max() > 0, name
        assert same(tr, batch[name][0]), name
        assert same(f, batch[name][1]), name
