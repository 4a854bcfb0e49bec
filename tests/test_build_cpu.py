"""The GPU build's CPU half: inserting rows into the levels above 0 only gives the same upper
levels, entry point and element levels as the full insertion (addPoint, hnswalg.h:1152-1266),
because no step at a level above 0 reads level 0.  Checked on the saved bytes: everything but
the level-0 link lists is identical."""
import numpy as np

from hnsw_rabitq_amd import _lib, _ptr, datasets

HEADER = 96


def _blob(hg, upper_only, data):
    p = hg.Index("l2", data.shape[1])
    p.init_index(data.shape[0], M=16, ef_construction=200, random_seed=100)
    if upper_only:
        _lib.call("hg_hnsw_add_items_upper", p.cpu_handle(), _ptr(np.ascontiguousarray(data)), data.shape[0],
                  None, 1)
    else:
        p.add_items(data, num_threads=1)
    return np.frombuffer(p.serialize(), np.uint8)


def test_upper_levels_equal_full_build(hg):
    data = datasets.c1_uniform(n=4000)
    full = _blob(hg, False, data)
    upper = _blob(hg, True, data)
    assert full.shape == upper.shape
    hdr = full[:HEADER]
    np.testing.assert_array_equal(hdr, upper[:HEADER])
    n = int(hdr[8:16].view(np.uint64)[0])
    per = int(hdr[24:32].view(np.uint64)[0])
    links0 = int(hdr[40:48].view(np.uint64)[0])  # offsetData_: the level-0 link block ends there
    lv0_f = full[HEADER:HEADER + n * per].reshape(n, per)
    lv0_u = upper[HEADER:HEADER + n * per].reshape(n, per)
    np.testing.assert_array_equal(lv0_f[:, links0:], lv0_u[:, links0:])  # vectors and labels
    assert (lv0_u[:, :links0] == 0).all()  # level-0 lists left empty
    np.testing.assert_array_equal(full[HEADER + n * per:], upper[HEADER + n * per:])  # upper levels
