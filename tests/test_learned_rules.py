"""Host-side rules of the learned early stop (no GPU): the class table's
ordering and the cumulative-probability class choice the GPU pick kernel
applies."""
import numpy as np

from hrnn_amd.learned import CLASSES, EF1, choose_class


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
