"""The closure's fine->coarse restriction on the C-grid (CPU)."""
import numpy as np

import nemo2d
from nemo2d import closure_pipeline as cp


def random_fine(nf, seed):
    rng = np.random.default_rng(seed)
    g, _ = nemo2d.seamount_grid(nf, nf, seed, radius=(3e3, 12e3))
    s = nemo2d.shapes(nf, nf)
    st = {"sshn": 0.1 * rng.standard_normal(s["T"]), "un": 0.05 * rng.standard_normal(s["U"]),
          "vn": 0.05 * rng.standard_normal(s["V"])}
    st["sshn_u"], st["sshn_v"] = cp.next_interp(st["sshn"])
    return g, st


def test_volume_is_conserved():
    g, st = random_fine(24, 1)
    cg = cp.coarse_grid(g, 2000.0, 2000.0)
    c = cp.restrict_state(st, g, cg)
    vf = np.sum(g["e12t"] * st["sshn"])
    vc = np.sum(cg["e12t"] * c["sshn"])
    assert abs(vf - vc) <= 1e-12 * np.sum(np.abs(g["e12t"] * st["sshn"]))


def test_face_transport_is_conserved():
    """The coarse face carries the two fine faces' volume transport, with the
    coarse face depth taken as the fine faces' mean."""
    g, st = random_fine(16, 2)
    cg = cp.coarse_grid(g, 2000.0, 2000.0)
    c = cp.restrict_state(st, g, cg)
    hu = g["hu"][:, 0::2]
    fine_q = hu[0::2] * st["un"][0::2, 0::2] * 1000.0 + hu[1::2] * st["un"][1::2, 0::2] * 1000.0
    coarse_q = 0.5 * (hu[0::2] + hu[1::2]) * c["un"] * 2000.0
    assert np.allclose(coarse_q, fine_q, rtol=1e-13, atol=1e-18)
    hv = g["hv"][0::2, :]
    fine_qv = hv[:, 0::2] * st["vn"][0::2, 0::2] * 1000.0 + hv[:, 1::2] * st["vn"][0::2, 1::2] * 1000.0
    coarse_qv = 0.5 * (hv[:, 0::2] + hv[:, 1::2]) * c["vn"] * 2000.0
    assert np.allclose(coarse_qv, fine_qv, rtol=1e-13, atol=1e-18)


def test_uniform_fields_restrict_to_themselves():
    g = nemo2d.default_grid(12, 12)
    st = nemo2d.initial_state(g, np.full((12, 12), 0.3))
    st["un"][:] = 0.2
    st["vn"][:] = -0.1
    cg = cp.coarse_grid(g, 2000.0, 2000.0)
    c = cp.restrict_state(st, g, cg)
    assert np.allclose(c["sshn"], 0.3) and np.allclose(c["un"], 0.2) and np.allclose(c["vn"], -0.1)
    assert np.allclose(cg["ht"], 100.0)


def test_increment_goes_to_interior_faces_only():
    g = nemo2d.default_grid(6, 5)
    st = nemo2d.initial_state(g)
    d = np.zeros((5, 6, 3))
    d[..., 1] = 1.0
    d[..., 2] = 2.0
    d[..., 0] = 0.5
    s = cp.apply_increment(st, d, g)
    assert np.all(s["un"][:, 1:-1] == 1.0) and np.all(s["un"][:, [0, -1]] == 0.0)
    assert np.all(s["vn"][1:-1, :] == 2.0) and np.all(s["vn"][[0, -1], :] == 0.0)
    assert np.all(s["sshn"] == 0.5)


def test_increment_of_field_difference_recovers_target():
    """apply_increment(a, t_fields(b) - t_fields(a)) puts b's values on every
    corrected point: the cell's ssh, its east face and its north face."""
    rng = np.random.default_rng(3)
    g = nemo2d.default_grid(7, 6)
    a = {k: rng.standard_normal(v.shape) for k, v in nemo2d.initial_state(g).items()}
    b = {k: rng.standard_normal(v.shape) for k, v in a.items()}
    s = cp.apply_increment(a, cp.t_fields(b) - cp.t_fields(a), g)
    assert np.allclose(s["sshn"], b["sshn"], rtol=0, atol=1e-14)
    assert np.allclose(s["un"][:, 1:-1], b["un"][:, 1:-1], rtol=0, atol=1e-14)
    assert np.allclose(s["vn"][1:-1, :], b["vn"][1:-1, :], rtol=0, atol=1e-14)
    assert np.array_equal(s["un"][:, [0, -1]], a["un"][:, [0, -1]])
    assert np.array_equal(s["vn"][[0, -1], :], a["vn"][[0, -1], :])
