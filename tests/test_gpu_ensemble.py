"""Batched ensembles: every FP64 member is bitwise the member run alone
through the single-model path (and the oracle); the graph-captured loop,
every tiling the autotuner may pick and the plain launch loop agree; the
FP32 ensemble stays close to the FP64 one."""
import numpy as np
import pytest

import nemo2d
from oracle import Oracle

pytestmark = pytest.mark.gpu

RDT = 12.0


def small_members():
    return {"jpi": np.array([1, 5, 17, 64, 80, 33], dtype=np.int32),
            "jpj": np.array([1, 9, 17, 40, 80, 70], dtype=np.int32),
            "dep": np.array([60.0, 150.0, 75.0, 199.0, 100.0, 55.0]),
            "amp": np.array([0.1, 0.25, 0.2, 0.3, 0.2, 0.15]),
            "period": np.array([11.0, 12.0, 12.42, 13.0, 12.42, 11.5])}


def single(k, m, nsteps):
    with nemo2d.Model(int(m["jpi"][k]), int(m["jpj"][k]), dep_const=float(m["dep"][k]), rdt=RDT,
                      cbfr=0.0015, visc=0.1) as s:
        s.set_tide(float(m["amp"][k]), float(m["period"][k]))
        s.record_checksums(nsteps)
        s.run(1, nsteps)
        return s.retrieve(), s.checksum_log(nsteps)


@pytest.mark.parametrize("tiling", [0, 8, 9, 10])
def test_fp64_members_bitwise_equal_single_model_and_oracle(tiling):
    m = small_members()
    nsteps = 80
    with nemo2d.Ensemble(m, rdt=RDT) as e:
        e.set_tiling(tiling)
        e.run(1, nsteps, graph=True)
        for k in range(len(m["jpi"])):
            r = e.retrieve(k)
            ref, log_ref = single(k, m, nsteps)
            for f in ("sshn", "un", "vn"):
                assert np.array_equal(r[f], ref[f]), (k, f)
            log = e.checksum_log(k, nsteps)
            assert np.allclose(log, log_ref, rtol=1e-13, atol=0), k
        # two members against the oracle directly
        for k in (1, 4):
            o = Oracle(int(m["jpi"][k]), int(m["jpj"][k]), dep_const=float(m["dep"][k]), rdt=RDT)
            o.set_tide(float(m["amp"][k]), float(m["period"][k]))
            o.run(1, nsteps)
            r = e.retrieve(k)
            for f in ("sshn", "un", "vn"):
                assert np.array_equal(r[f], o.arr(f)), (k, f)


def test_graph_plain_and_tilings_agree():
    m = small_members()
    nsteps = 40
    results = []
    with nemo2d.Ensemble(m, rdt=RDT) as e:
        for tiling, graph in ((0, True), (0, False), (3, True), (5, True), (7, False), (2, True),
                              (8, True), (8, False), (9, True), (9, False), (10, True), (10, False)):
            e.set_tiling(tiling)
            e.reset()
            e.run(1, nsteps, graph=graph)
            results.append([e.retrieve(k) for k in range(len(m["jpi"]))])
        choice, ms = e.tune(nsteps=5)
        assert 0 <= choice < 11 and ms > 0
        e.run(1, nsteps)
        results.append([e.retrieve(k) for k in range(len(m["jpi"]))])
    for other in results[1:]:
        for a, b in zip(results[0], other):
            for f in ("sshn", "un", "vn"):
                assert np.array_equal(a[f], b[f])


def test_fp32_ensemble_close_to_fp64():
    m = nemo2d.ensemble_members(12, seed=4, size=(32, 200))
    nsteps = 600
    with nemo2d.Ensemble(m, rdt=RDT, precision=64) as e64, \
            nemo2d.Ensemble(m, rdt=RDT, precision=32) as e32:
        e64.run(1, nsteps)
        e32.run(1, nsteps)
        for k in range(12):
            a, b = e32.retrieve(k), e64.retrieve(k)
            err = nemo2d.field_errors(a, b)
            assert err["ssh"]["ref_rms"] > 0
            for f in ("ssh", "u", "v"):
                assert np.isfinite(err[f]["rms"])
                assert err[f]["rms"] <= 1e-2 * err[f]["ref_rms"], (k, f, err[f])


def test_fp32_marching_tiling_equals_per_cell_tiling():
    m = nemo2d.ensemble_members(7, seed=9, size=(20, 150))
    out = []
    with nemo2d.Ensemble(m, rdt=RDT, precision=32) as e:
        for tiling in (0, 8, 9, 10):
            e.set_tiling(tiling)
            e.reset()
            e.run(1, 120)
            out.append([e.retrieve(k) for k in range(7)])
    for other in out[1:]:
        for a, b in zip(out[0], other):
            for f in ("sshn", "un", "vn"):
                assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("precision", [64, 32])
def test_switching_tilings_mid_run(precision):
    """Steps of the sshn-only marching tiling leave sshn_u/sshn_v unwritten;
    a following step of a tiling that reads them must see them filled in
    (graph or not), and the large members have plain work items."""
    m = nemo2d.ensemble_members(5, seed=11, size=(130, 260))
    with nemo2d.Ensemble(m, rdt=RDT, precision=precision) as e:
        e.set_tiling(0)
        e.run(1, 60)
        ref = [e.retrieve(k) for k in range(5)]
        e.reset()
        for tiling, (a, b), graph in ((9, (1, 20), True), (0, (21, 30), True), (10, (31, 45), False),
                                      (8, (46, 50), False), (9, (51, 55), True), (10, (56, 58), True),
                                      (3, (59, 60), True)):
            e.set_tiling(tiling)
            e.run(a, b, graph=graph)
        for k in range(5):
            r = e.retrieve(k)
            for f in ("sshn", "un", "vn"):
                assert np.array_equal(r[f], ref[k][f]), (k, f)


@pytest.mark.parametrize("tiling", [8, 9, 10])
def test_many_work_items_in_cost_order(tiling):
    """More marching work items than the kernel has waves (tilings 8 and 9
    deal the items to the waves in cost order, several per wave; tiling 10
    gives each its own wave), each item keeping its own checksum slot:
    members bitwise equal to their single runs, per-member checksum logs
    equal to theirs."""
    m = nemo2d.ensemble_members(40, seed=13, size=(180, 300))
    nsteps = 25
    with nemo2d.Ensemble(m, rdt=RDT) as e:
        e.set_tiling(tiling)
        e.run(1, nsteps, graph=True)
        for k in (0, 17, 39):
            r = e.retrieve(k)
            ref, log_ref = single(k, m, nsteps)
            for f in ("sshn", "un", "vn"):
                assert np.array_equal(r[f], ref[f]), (k, f)
            assert np.allclose(e.checksum_log(k, nsteps), log_ref, rtol=1e-13, atol=0), k


def test_ensemble_rejects_bad_members():
    m = small_members()
    m["jpi"] = m["jpi"].copy()
    m["jpi"][2] = 0
    with pytest.raises(nemo2d.Nemo2dError):
        nemo2d.Ensemble(m)
