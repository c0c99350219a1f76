"""The allocation-free lookup (mt_find_noalloc) against mt_find, through the C-ABI.

mt_find_noalloc stages keys and slots in a fixed per-GPU device array instead of
allocating; its results must be mt_find's, slot for slot, on consecutive calls that
reuse that array (the second with fewer keys, so stale entries would show), and it
must refuse more keys than the array holds before touching the device.
"""
import numpy as np
import pytest

from mortontrie_amr_amd import amr, cases as C

pytestmark = pytest.mark.gpu


def test_find_noalloc_matches_find():
    case = C.sphere_amr((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5))
    g = amr.AmrCase(case)
    rng = np.random.default_rng(3)
    # call 1: every key of the finest level, shuffled, plus keys that are absent
    k2 = case["levels"][2]["keys"]
    absent = C.encode_key(rng.integers(0, 4 * 32, 500), rng.integers(0, 4 * 24, 500),
                          rng.integers(0, 4 * 24, 500), 2, 2).astype(np.uint64)
    q1 = rng.permutation(np.concatenate([k2, absent]))
    want1 = g.find(2, q1)
    got1 = g.find_noalloc(2, q1)
    assert np.array_equal(got1, want1)
    assert (want1 >= 0).sum() >= k2.size and (want1 < 0).any()
    # call 2: fewer keys on another level; the scratch still holds call 1's tail
    k0 = case["levels"][0]["keys"]
    q2 = rng.permutation(k0)[: k0.size // 3]
    want2 = g.find(0, q2)
    got2 = g.find_noalloc(0, q2)
    assert np.array_equal(got2, want2)
    assert (want2 >= 0).all()
    # call 3: the first call again gives the same slots
    assert np.array_equal(g.find_noalloc(2, q1), want1)


def test_find_noalloc_rejects_more_keys_than_its_scratch():
    case = C.cavity(8)
    g = amr.AmrCase(case)
    n = amr.FIND_NOALLOC_MAX_KEYS + 1
    with pytest.raises(amr.MtError, match="scratch capacity"):
        g.find_noalloc(0, np.zeros(n, np.uint64))
    # an empty call is fine and the entry point still works afterwards
    assert g.find_noalloc(0, np.zeros(0, np.uint64)).size == 0
    k = case["levels"][0]["keys"]
    assert np.array_equal(g.find_noalloc(0, k), g.find(0, k))
