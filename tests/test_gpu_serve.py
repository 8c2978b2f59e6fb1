"""Serve mode (the reference's compactForServing, include/hnswlib/hrnn.h:299-364,
and hrnn_search --max_k_serve): the KNN lists are dropped for a dense
n x max_k_serve matrix of the first distances, FLT_MAX past a node's count.
Search after compaction reads r_k from that matrix and the reverse lists
from the CSR built before, so every answer for k <= max_k_serve is the
answer before compaction; k > max_k_serve is refused (the reference reads
past the row there), and so is saving the index."""
import numpy as np
import pytest

import hrnn_amd
from hrnn_amd.index import Index

pytestmark = pytest.mark.gpu


def test_search_after_compaction_equals_before(tmp_path):
    x = hrnn_amd.synth_mixture(41, 6000, 128, 40, 20.0, 0.0, 255.0, True, 0.0, 1041)
    q = hrnn_amd.synth_mixture(42, 200, 128, 40, 20.0, 0.0, 255.0, True, 0.0, 1041)
    ix = Index(6000, 128, 16, 100, 100, 48)
    ix.add_points_gpu(x)
    ids, d, cnt = ix.get_knng()
    cnt = cnt.copy()
    # short lists: r_k past the count is FLT_MAX in both modes.  Only lists
    # holding at least 5 entries are cut (the first points inserted have
    # fewer, and entries past a count were never written)
    cut = np.arange(0, len(cnt), 7)
    cut = cut[cnt[cut] >= 5]
    assert len(cut) > 500
    cnt[cut] = 5
    ix.set_knng(ids, d, cnt)
    ix.build_rknng()
    before = {}
    for k, m in ((1, 3), (10, 5), (16, 10)):
        before[k] = ix.search(q, m, 1.0, k, 48, 100)
    ix.compact_for_serving(16)
    for k, m in ((1, 3), (10, 5), (16, 10)):
        off, lab = ix.search(q, m, 1.0, k, 48, 100)
        assert np.array_equal(off, before[k][0]) and np.array_equal(lab, before[k][1]), k
        assert int(off[-1]) > 0
    with pytest.raises(hrnn_amd.HrnnError):
        ix.search(q, 5, 1.0, 17, 48, 100)
    with pytest.raises(hrnn_amd.HrnnError, match="serve mode"):
        ix.save(str(tmp_path / "s.idx"))
