"""Label rule of the chooser's training data (host logic, no GPU)."""
import numpy as np

from ducc_amd.chooser_train import candidates, kernel_db, pick_label


def _rows(db, nd, single, W):
    sel = (db["ndim"] == nd) & (db["single"] == single) & (db["W"] == W)
    r = np.nonzero(sel)[0]
    return r[np.argsort(db["ofactor"][r])]


def test_analytic_kept_unless_clearly_faster():
    db = kernel_db()
    a = int(_rows(db, 2, 1, 7)[3])
    b = int(_rows(db, 2, 1, 6)[3])
    # 20% faster than the analytic setting: not enough, analytic stays
    times = {(a, 5): 1.0, (b, 5): 0.80}
    assert pick_label(times, (1.0, a, 5), db, 2, False) == (a, 5)
    # 30% less time (1.43x): the candidate wins
    times = {(a, 5): 1.0, (b, 5): 0.70}
    assert pick_label(times, (1.0, a, 5), db, 2, False) == (b, 5)


def test_near_ties_prefer_small_support_then_small_sigma_then_default_tile():
    db = kernel_db()
    w6 = _rows(db, 2, 1, 6)
    w7 = _rows(db, 2, 1, 7)
    k6a, k6b, k7 = int(w6[2]), int(w6[5]), int(w7[2])
    times = {(k7, 5): 0.80, (k6b, 5): 0.81, (k6a, 6): 0.82, (k6a, 5): 0.823,
             (k6a, 4): 0.95}
    # all but the last are within 3% of the best; W=6 beats W=7, the smaller
    # oversampling factor beats the larger, the default tile (5) beats 6
    assert pick_label(times, (1.2, k7, 4), db, 2, False) == (k6a, 5)


def test_candidates_reach_eps_and_stay_in_sigma_range():
    db = kernel_db()
    req = dict(shape=(512, 512), M=10000, eps=1e-5, dist="uniform")
    c = candidates(req, False, db)
    assert c
    for k, t in c:
        assert db["ndim"][k] == 2 and db["single"][k] == 1
        assert db["epsilon"][k] <= 1e-5
        assert 1.2 <= db["ofactor"][k] <= 2.0
        assert t in (4, 5, 6)
