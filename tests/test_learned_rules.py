"""Host-side rules of the learned early stop (no GPU): the class table's
ordering and the cumulative-probability class choice the GPU pick kernel
applies."""
import numpy as np

from hrnn_amd.learned import CLASSES, EF1, choose_class, label_classes, per_query_recall


def test_class_table_is_a_domination_chain():
    ef, m = CLASSES[:, 0].astype(int), CLASSES[:, 1].astype(int)
    assert CLASSES.shape == (16, 2)
    assert ef[0] == EF1 and np.all(np.diff(ef) >= 0) and np.all(np.diff(m) >= 0)
    assert np.all(np.diff(ef) + np.diff(m) > 0)          # no two classes alike
    assert ef.max() <= 512 and m.min() >= 1


def test_choose_class_cumulative_rule():
    p = np.full((3, 16), 1 / 16, np.float32)
    p[1] = 0.0
    p[1, 5] = 1.0                                       # all mass on class 5
    p[2] = 0.0
    p[2, 0], p[2, 15] = 0.4, 0.6
    assert choose_class(p, 0.5).tolist() == [7, 5, 15]  # 8/16 reaches 0.5 at index 7
    assert choose_class(p, 0.4).tolist() == [6, 5, 0]   # 7/16 = 0.4375 >= 0.4
    assert choose_class(p, 0.0).tolist()[1:] == [5, 15]  # argmax
    assert choose_class(p, 2.0).tolist() == [15, 15, 15]  # never reached: the last class


class _FakeDev:
    """stage1_features and rknn_search over a scripted answer table: with m
    proxies query i answers ids[i][:m] (as if each proxy added one id)."""

    def __init__(self, ids):
        self.ids = ids

    def stage1_features(self, q, ef, m, k):
        nq = q.shape[0]
        return np.zeros((nq, 32), np.float32), np.tile(np.arange(m, dtype=np.uint32), (nq, 1)), \
            np.full((nq, 2), ef, np.uint32)

    def rknn_search(self, q, prox, k, K_prime):
        m = prox.shape[1]
        lists = [np.asarray(r[:m], np.uint32) for r in self.ids]
        off = np.zeros(len(lists) + 1, np.uint64)
        off[1:] = np.cumsum([len(a) for a in lists])
        return off, np.concatenate(lists)


def test_per_query_recall_matches_sets():
    rng = np.random.default_rng(0)
    g = [rng.choice(50, rng.integers(0, 6), replace=False) for _ in range(40)]
    r = [rng.choice(50, rng.integers(0, 8), replace=False) for _ in range(40)]
    off = lambda ls: np.concatenate([[0], np.cumsum([len(a) for a in ls])])
    cat = lambda ls: np.concatenate(ls).astype(np.uint32)
    got = per_query_recall(off(g), cat(g), off(r), cat(r))
    want = [len(set(a) & set(b)) / len(a) if len(a) else np.nan for a, b in zip(g, r)]
    assert np.allclose(got, want, equal_nan=True)


def test_label_classes_best_and_reference_recall():
    """Three queries, groundtruth {0..3}; each extra proxy finds one more id.
    Classes m = 1, 2, 3, 4: the best-recall label is the first m reaching the
    query's best; with ref_recall (a fixed-m run's per-query recall) it is the
    first m reaching that instead, the best class where no class does."""
    gt = [np.arange(4)] * 3
    ids = [np.array([0, 1, 2, 3]), np.array([0, 9, 1, 9]), np.array([9, 9, 9, 9])]
    gt_off = np.array([0, 4, 8, 12])
    gt_ids = np.concatenate(gt).astype(np.uint32)
    cls = np.array([(16, 1), (16, 2), (16, 3), (16, 4)], np.uint32)
    q = np.zeros((3, 4), np.float32)
    lab, rec, _ = label_classes(_FakeDev(ids), q, gt_off, gt_ids, 10, 100, classes=cls)
    assert np.allclose(rec[:, 0], [0.25, 0.5, 0.75, 1.0]) and np.allclose(rec[:, 1], [0.25, 0.25, 0.5, 0.5])
    assert lab.tolist() == [3, 2, 0]
    lab2, _, _ = label_classes(_FakeDev(ids), q, gt_off, gt_ids, 10, 100, classes=cls,
                               ref_recall=np.array([0.5, 0.25, 1.0]))
    assert lab2.tolist() == [1, 0, 0]
