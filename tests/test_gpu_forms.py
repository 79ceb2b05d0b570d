"""The single-precision form against the FP64 path, and the FP64
diagnostics against a numpy evaluation of the same formulas."""
import numpy as np
import pytest

import nemo2d

pytestmark = pytest.mark.gpu

GRAV = 9.80665


def run(prec, grid, nsteps, compensated=True, n=192):
    # cbfr = 0.0015 (the benchmark physics): with seamounts the weaker
    # friction of the shipped namelist (0.00015) lets the run blow up in
    # FP64 already
    with nemo2d.Model(n, n, grid=grid, precision=prec, compensated=compensated,
                      cbfr=0.0015) as m:
        m.run(1, nsteps)
        return m.retrieve(), m.diagnostics()


def test_fp32_error_against_fp64_and_compensation_ablation():
    n, nsteps = 192, 1500
    grid, _ = nemo2d.seamount_grid(n, n, seed=3, radius=(5e3, 40e3))
    ref, dref = run("fp64", grid, nsteps, n=n)
    f32, d32 = run("fp32", grid, nsteps, n=n)
    raw, draw = run("fp32", grid, nsteps, compensated=False, n=n)
    e = nemo2d.field_errors(f32, ref)
    e_raw = nemo2d.field_errors(raw, ref)
    print("fp32 vs fp64:", e)
    print("fp32 without FP64 ssh accumulation vs fp64:", e_raw)
    # the state is far from rest (tide, reflections); FP32 stays within a
    # small fraction of the signal
    for k in ("ssh", "u", "v"):
        assert np.isfinite(e[k]["rms"]) and e[k]["ref_rms"] > 0
        assert e[k]["rms"] < 1e-2 * e[k]["ref_rms"], (k, e[k])
    # keeping the ssh accumulation in FP64 matters
    assert e["ssh"]["rms"] < e_raw["ssh"]["rms"], (e["ssh"], e_raw["ssh"])
    # volume diagnostic: FP32 compensated tracks the FP64 volume closely
    assert abs(d32[0] - dref[0]) <= 1e-5 * abs(dref[0])


@pytest.mark.parametrize("compensated", [True, False])
def test_fp32_marching_kernel_equals_per_cell_kernel(compensated):
    """The single-precision form's marching kernels (shared face terms,
    per-point depths: the one reading sshn, un, vn and the depths, which
    forms hu, hv from ht on this basin, and the five-array one) give the
    per-cell kernel's fields bit for bit; 'auto' (per cell at this size)
    too."""
    ni, nj, nsteps = 211, 97, 60
    grid, _ = nemo2d.seamount_grid(ni, nj, seed=5, radius=(5e3, 30e3))
    out = {}
    for variant in ("march12_grd", "march8_grd", "march12_shd", "auto", "cell"):
        with nemo2d.Model(ni, nj, grid=grid, precision="fp32", compensated=compensated,
                          cbfr=0.0015) as m:
            m.set_variant(variant)
            m.run(1, nsteps)
            out[variant] = m.retrieve()
    for variant in ("march12_grd", "march8_grd", "march12_shd", "auto"):
        for k in ("sshn", "un", "vn", "sshn_u", "sshn_v"):
            a, b = out[variant][k], out["cell"][k]
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (variant, k,
                                                                          np.abs(a - b).max())


def test_fp32_rest_stays_at_rest_and_tide_row():
    grid = nemo2d.default_grid(40, 30)
    with nemo2d.Model(40, 30, grid=grid, precision="fp32") as m:
        m.step(1)
        r = m.retrieve()
    t = nemo2d.tide(20.0)
    assert np.all(r["sshn"][0, :] == t)            # FP64 ssh clamp (compensated form)
    assert np.all(r["sshn"][1:, :] == 0.0)
    assert np.all(r["un"] == 0.0) and np.all(r["vn"] == 0.0)


def test_fp32_stage_calls_refused():
    with nemo2d.Model(8, 8, precision="fp32") as m:
        with pytest.raises(nemo2d.Nemo2dError) as e:
            m.continuity()
        assert e.value.code == nemo2d.ERR_STATE


def test_fp32_needs_uniform_metrics():
    g = nemo2d.default_grid(8, 8)
    g["e1t"][2, 2] = 1100.0
    with pytest.raises(nemo2d.Nemo2dError) as e:
        nemo2d.Model(8, 8, grid=g, precision="fp32")
    assert e.value.code == nemo2d.ERR_ARG


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_diagnostics_match_numpy(prec):
    n = 64
    grid, _ = nemo2d.seamount_grid(n, n, seed=5, radius=(5e3, 20e3))
    with nemo2d.Model(n, n, grid=grid, precision=prec) as m:
        m.run(1, 300)
        r = m.retrieve()
        vol, en, area = m.diagnostics()
    e12t = grid["e12t"]
    z = r["sshn"]
    ub = 0.5 * (r["un"][:, :-1] + r["un"][:, 1:])
    vb = 0.5 * (r["vn"][:-1, :] + r["vn"][1:, :])
    ht = grid["ht"].astype(np.float32).astype(np.float64) if prec == "fp32" else grid["ht"]
    vol_ref = np.sum(e12t * z)
    en_ref = np.sum(e12t * (0.5 * (ht + z) * (ub * ub + vb * vb) + 0.5 * GRAV * z * z))
    assert area == pytest.approx(np.sum(e12t), rel=1e-14)
    assert vol == pytest.approx(vol_ref, rel=1e-10, abs=1e-6)
    assert en == pytest.approx(en_ref, rel=1e-10)
    assert vol != 0.0 and en > 0.0
