"""hfx_eri_shells_noalloc (integral blocks staged in the library's fixed
device scratch array, no allocation) against hfx_eri_shells through the C ABI:
two calls in a row, the first close to the scratch size and the second smaller
(no stale contents may leak into it), and the size limit."""
import sys

import numpy as np
import pytest

from hfsolver_amd.basis import Basis
from hfsolver_amd.clusters import water_cluster
from hfsolver_amd.engine import Engine, HfxError, _ptr, lib

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 16 << 20


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def nc(L):
    return (L + 1) * (L + 2) // 2


def call(eng, name, quartets):
    q = np.ascontiguousarray(np.asarray(quartets, dtype=np.int32).reshape(-1, 4))
    Ls = eng.shells()[2]
    tot = int(sum(nc(Ls[a]) * nc(Ls[b]) * nc(Ls[c]) * nc(Ls[d]) for a, b, c, d in q))
    out = np.full(tot, np.nan)
    eng._check(getattr(lib(), name)(eng._h, len(q), _ptr(q), _ptr(out)))
    return out


def test_noalloc_matches_eri_shells_on_two_calls():
    atno, xyz = water_cluster(4, seed=4)
    eng = Engine(Basis(atno, xyz))
    ns, f0, Ls, _ = eng.shells()
    rng = np.random.default_rng(11)
    dsh = np.nonzero(Ls == 2)[0]
    big = dsh[rng.integers(0, len(dsh), size=(1500, 4))]          # (dd|dd) blocks: 15.6 MB of integrals
    small = rng.integers(0, ns, size=(97, 4))                      # every class combination, mixed orientations
    assert 1500 * (1296 * 8 + 20) < SCRATCH_BYTES
    for k, q in enumerate((big, small)):
        say("call %d: %d quartets" % (k, len(q)))
        ref = call(eng, "hfx_eri_shells", q)
        got = call(eng, "hfx_eri_shells_noalloc", q)
        assert not np.isnan(got).any()
        assert np.array_equal(got, ref)
    eng.close()


def test_noalloc_rejects_requests_beyond_the_scratch():
    atno, xyz = water_cluster(1, seed=0)
    eng = Engine(Basis(atno, xyz))
    ns, f0, Ls, _ = eng.shells()
    d = int(np.nonzero(Ls == 2)[0][0])
    too_big = np.full((1700, 4), d, dtype=np.int32)                # 17.6 MB of integrals
    with pytest.raises(HfxError, match="scratch"):
        call(eng, "hfx_eri_shells_noalloc", too_big)
    # the engine stays usable after the rejected call
    q = [(0, 1, d, d), (d, 0, 2, 3)]
    assert np.array_equal(call(eng, "hfx_eri_shells_noalloc", q), call(eng, "hfx_eri_shells", q))
    eng.close()
