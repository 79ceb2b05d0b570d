"""The per-point class words (host code, no device): every decision the
serial model's stages take from the mask, restated here point by point from
its IF tests, must be the one the library's class word encodes.

  momentum_u  original/nemolite2d.f90:431-437 (skip tests), :458-470 (walls)
  momentum_v  :520-525 (skip tests, incl. the pt(ji+1,jj) test of :523),
              :546-558 (walls)
  bc_ssh      :631-643       bc_solid_u :648-652    bc_solid_v :656-660
  Flather u   :669-680       Flather v  :684-695
  next        :735-746 (sshn_u), :751-762 (sshn_v)
"""
import numpy as np
import pytest

import nemo2d
from oracle import Oracle
from test_gpu_parity import copy_grid_from_oracle, random_grid_oracle

# bit layout of a class word (the library's per-point classification)
U_WALL_S, U_WALL_N, U_FL_EAST = 1 << 2, 1 << 3, 1 << 4
V_WALL_W, V_WALL_E, V_FL_NORTH = 1 << 7, 1 << 8, 1 << 9
T_CLAMP = 1 << 10
NONE, COMP, SOLID, FLATHER = 0, 1, 2, 3
SKIP, AVG, HI, LO = 0, 1, 2, 3


def expected_mismatches(pt, words):
    """Compare words (mask layout) with the serial model's decisions; return
    a list of (what, ji, jj) that disagree."""
    jpi, jpj = pt.shape[1] - 2, pt.shape[0] - 2
    P = lambda i, j: int(pt[j, i])  # noqa: E731  the reference's pt(i, j)
    bad = []
    for jj in range(0, jpj + 2):
        for ji in range(0, jpi + 2):
            w = int(words[jj, ji])
            # ---- U point (ji, jj): ji = 0..jpi, jj = 1..jpj
            if 0 <= ji <= jpi and 1 <= jj <= jpj:
                a, b = P(ji, jj), P(ji + 1, jj)
                comp = 1 <= ji <= jpi - 1 and a + b > 0 and a > 0 and b > 0
                solid = a * b == 0
                flat = not (a + b <= -1) and (a < 0 or b < 0)
                cls = COMP if comp else SOLID if solid else FLATHER if flat else NONE
                if (w & 3) != cls:
                    bad.append(("u class", ji, jj))
                if comp:
                    ws = P(ji, jj - 1) <= 0 or P(ji + 1, jj - 1) <= 0
                    wn = P(ji, jj + 1) <= 0 or P(ji + 1, jj + 1) <= 0
                    if bool(w & U_WALL_S) != ws or bool(w & U_WALL_N) != wn:
                        bad.append(("u walls", ji, jj))
                if flat and bool(w & U_FL_EAST) != (a < 0):
                    bad.append(("u flather side", ji, jj))
                su = SKIP if a + b <= 0 else AVG if a * b > 0 else HI if a <= 0 else LO
                if (w >> 11) & 3 != su:
                    bad.append(("sshn_u rule", ji, jj))
            # ---- V point (ji, jj): ji = 1..jpi, jj = 0..jpj
            if 1 <= ji <= jpi and 0 <= jj <= jpj:
                a, b = P(ji, jj), P(ji, jj + 1)
                comp = (1 <= jj <= jpj - 1 and P(ji, jj) + P(ji + 1, jj) > 0 and a > 0 and b > 0)
                solid = a * b == 0
                flat = not (a + b <= -1) and (a < 0 or b < 0)
                cls = COMP if comp else SOLID if solid else FLATHER if flat else NONE
                if (w >> 5) & 3 != cls:
                    bad.append(("v class", ji, jj))
                if comp:
                    ww = P(ji - 1, jj) <= 0 or P(ji - 1, jj + 1) <= 0
                    we = P(ji + 1, jj) <= 0 or P(ji + 1, jj + 1) <= 0
                    if bool(w & V_WALL_W) != ww or bool(w & V_WALL_E) != we:
                        bad.append(("v walls", ji, jj))
                if flat and bool(w & V_FL_NORTH) != (a < 0):
                    bad.append(("v flather side", ji, jj))
                sv = SKIP if a + b <= 0 else AVG if a * b > 0 else HI if a <= 0 else LO
                if (w >> 13) & 3 != sv:
                    bad.append(("sshn_v rule", ji, jj))
            # ---- T point (ji, jj): bc_ssh's clamp
            if 1 <= ji <= jpi and 1 <= jj <= jpj:
                clamp = P(ji, jj) > 0 and (P(ji, jj - 1) < 0 or P(ji, jj + 1) < 0 or
                                           P(ji + 1, jj) < 0 or P(ji - 1, jj) < 0)
                if bool(w & T_CLAMP) != clamp:
                    bad.append(("ssh clamp", ji, jj))
    return bad


@pytest.mark.parametrize("jpi,jpj", [(1, 1), (2, 3), (7, 5), (16, 9)])
def test_reference_basin_classes(jpi, jpj):
    g = nemo2d.default_grid(jpi, jpj)
    w = nemo2d.classify(g)
    assert expected_mismatches(g["pt"], w) == []
    # the reference basin's closed form: the clamp row and the Flather row
    assert all(w[1, ji] & T_CLAMP for ji in range(1, jpi + 1))
    assert all((w[0, ji] >> 5) & 3 == FLATHER and w[0, ji] & V_FL_NORTH for ji in range(1, jpi + 1))


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_general_mask_classes(seed):
    o = random_grid_oracle(37, 29, seed)
    g = copy_grid_from_oracle(o)
    assert expected_mismatches(g["pt"], nemo2d.classify(g)) == []


def test_chained_flather_mask_classes():
    o = random_grid_oracle(24, 20, 7, islands=False, vary_metrics=False, chained=True)
    g = copy_grid_from_oracle(o)
    assert expected_mismatches(g["pt"], nemo2d.classify(g)) == []


def test_a_planted_mistake_is_seen():
    """The restatement notices a wrong word (one flipped decision)."""
    g = nemo2d.default_grid(7, 5)
    w = nemo2d.classify(g)
    w[1, 3] ^= T_CLAMP
    assert expected_mismatches(g["pt"], w) == [("ssh clamp", 3, 1)]


def test_classify_rejects_a_mask_the_reference_cannot_run():
    g = nemo2d.default_grid(8, 6)
    g["pt"][3, 0] = 1
    with pytest.raises(nemo2d.Nemo2dError):
        nemo2d.classify(g)


def test_oracle_grid_matches_default_grid():
    o = Oracle(9, 7)
    assert np.array_equal(np.array(o.pt), nemo2d.default_grid(9, 7)["pt"])
