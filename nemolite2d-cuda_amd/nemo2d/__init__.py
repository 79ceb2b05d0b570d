"""Python host binding of the nemo2d gfx950 library (include/nemo2d.h).

A thin ctypes layer over ``libnemo2d.so``, mirroring the reference's driver
interface (original/nemolite2d.f90: setup / grid / initialisation / step)
for scripts and tests.  Host arrays are numpy arrays in the reference's
Fortran layout: an array with Fortran bounds ``(lb1:ub1, lb2:ub2)`` is a C-
ordered numpy array of shape ``(ub2-lb2+1, ub1-lb1+1)``, so
``a[jj - lb2, ji - lb1]`` is the reference's ``a(ji, jj)``.

There is no CPU fallback: loading fails loudly if the library is missing.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnemo2d.so")

OK, ERR_ARG, ERR_HIP, ERR_MASK, ERR_STATE, ERR_NONFINITE = range(6)

# every entry point declared in include/nemo2d.h
SYMBOLS = (
    "nemo2d_last_error", "nemo2d_version", "nemo2d_create", "nemo2d_set_grid",
    "nemo2d_set_state", "nemo2d_continuity", "nemo2d_momentum_u", "nemo2d_momentum_v",
    "nemo2d_bc_ssh", "nemo2d_bc_solid_u", "nemo2d_bc_solid_v", "nemo2d_bc_flather_u",
    "nemo2d_bc_flather_v", "nemo2d_next", "nemo2d_step", "nemo2d_run",
    "nemo2d_record_checksums", "nemo2d_checksum_log", "nemo2d_checksums",
    "nemo2d_retrieve_results", "nemo2d_retrieve_aux", "nemo2d_timer_start",
    "nemo2d_timer_stop", "nemo2d_destroy", "nemo2d_create_mixed", "nemo2d_diagnostics",
    "nemo2d_set_variant", "nemo2d_set_tide", "nemo2d_ensemble_create", "nemo2d_ensemble_reset",
    "nemo2d_ensemble_tune", "nemo2d_ensemble_set_tiling", "nemo2d_ensemble_tiling",
    "nemo2d_ensemble_run", "nemo2d_ensemble_checksum_log", "nemo2d_ensemble_retrieve",
    "nemo2d_ensemble_destroy", "nemo2d_nn_conv_fwd", "nemo2d_nn_conv_bwd_data",
    "nemo2d_nn_wgrad_work_floats", "nemo2d_nn_conv_bwd_weight", "nemo2d_nn_gelu_bwd",
    "nemo2d_nn_mse", "nemo2d_nn_adam", "nemo2d_nn_pack_weights", "nemo2d_nn_to_bf16",
    "nemo2d_dev_alloc", "nemo2d_dev_free", "nemo2d_dev_upload", "nemo2d_dev_download",
    "nemo2d_dev_zero", "nemo2d_closure_features", "nemo2d_closure_apply",
    "nemo2d_stream_bandwidth", "nemo2d_format_go2d", "nemo2d_set_output_grid",
    "nemo2d_snapshot_begin", "nemo2d_snapshot_wait", "nemo2d_stage_timers", "nemo2d_stage_times",
    "nemo2d_set_stage_fusion", "nemo2d_retrieve_grid", "nemo2d_retrieve_classes",
    "nemo2d_classify", "nemo2d_nn_conv_bwd_data_gelu", "nemo2d_closure_restrict",
    "nemo2d_closure_target", "nemo2d_dev_synchronize", "nemo2d_nn_conv_bwd_weight_noalloc",
    "nemo2d_last_variant",
)

# nemo2d_set_variant choices
VARIANTS = {"auto": 0, "cell": 1, "cell_flin": 2, "march16": 3, "march12_shn": 4, "march12_shd": 5,
            "march12_dp": 6, "march12_dpf": 7, "march12_dpb": 8, "march12_dpfb": 9, "march12_rc": 10,
            "march12_grd": 11, "march8_grd": 12, "march8_rc": 13}


class Nemo2dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("nemo2d error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    """Load libnemo2d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("nemo2d: %s not built (run __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.nemo2d_last_error.restype = ctypes.c_char_p
        L.nemo2d_version.restype = ctypes.c_char_p
        L.nemo2d_create.argtypes = [ctypes.POINTER(vp), i, i, d, d, d, i, i, i, d, d, d]
        L.nemo2d_create_mixed.argtypes = [ctypes.POINTER(vp), i, i, d, d, d, i, i, i, d, d, d, i]
        L.nemo2d_diagnostics.argtypes = [vp, dp]
        L.nemo2d_set_grid.argtypes = [vp, ip] + [dp] * 14
        L.nemo2d_set_state.argtypes = [vp] + [dp] * 5
        for n in ("continuity", "momentum_u", "momentum_v", "bc_solid_u", "bc_solid_v",
                  "bc_flather_u", "bc_flather_v", "next", "timer_start", "destroy"):
            getattr(L, "nemo2d_" + n).argtypes = [vp]
        L.nemo2d_bc_ssh.argtypes = [vp, d]
        L.nemo2d_step.argtypes = [vp, i]
        L.nemo2d_set_variant.argtypes = [vp, i]
        L.nemo2d_last_variant.argtypes = [vp, ip]
        L.nemo2d_set_stage_fusion.argtypes = [vp, i]
        L.nemo2d_set_tide.argtypes = [vp, d, d]
        L.nemo2d_ensemble_create.argtypes = [ctypes.POINTER(vp), i, ip, ip, dp, dp, dp, d, d, d, d, d, i]
        L.nemo2d_ensemble_reset.argtypes = [vp]
        L.nemo2d_ensemble_tune.argtypes = [vp, i, ip, dp]
        L.nemo2d_ensemble_set_tiling.argtypes = [vp, i]
        L.nemo2d_ensemble_tiling.argtypes = [i, ip, ip]
        L.nemo2d_ensemble_run.argtypes = [vp, i, i, i, dp]
        L.nemo2d_ensemble_checksum_log.argtypes = [vp, i, dp, i, ip]
        L.nemo2d_ensemble_retrieve.argtypes = [vp, i, dp, dp, dp]
        L.nemo2d_ensemble_destroy.argtypes = [vp]
        L.nemo2d_stream_bandwidth.argtypes = [ctypes.c_long, i, dp]
        L.nemo2d_set_output_grid.argtypes = [vp, dp, dp, dp]
        L.nemo2d_stage_timers.argtypes = [vp, i]
        L.nemo2d_stage_times.argtypes = [vp, ip, dp, dp]
        L.nemo2d_snapshot_begin.argtypes = [vp, i, ctypes.c_char_p]
        L.nemo2d_snapshot_wait.argtypes = [vp]
        L.nemo2d_format_go2d.argtypes = [i, i, dp, dp, dp, dp, dp, dp, ctypes.c_char_p, ctypes.c_long,
                                         ctypes.POINTER(ctypes.c_long)]
        L.nemo2d_closure_features.argtypes = [vp, dp, dp, vp]
        L.nemo2d_closure_apply.argtypes = [vp, ctypes.POINTER(ctypes.c_float), i, dp]
        L.nemo2d_closure_restrict.argtypes = [vp, vp]
        L.nemo2d_closure_target.argtypes = [vp, vp, dp, vp]
        L.nemo2d_dev_alloc.argtypes = [ctypes.c_long, ctypes.POINTER(vp)]
        L.nemo2d_dev_free.argtypes = [vp]
        L.nemo2d_dev_upload.argtypes = [vp, vp, ctypes.c_long]
        L.nemo2d_dev_download.argtypes = [vp, vp, ctypes.c_long]
        L.nemo2d_dev_zero.argtypes = [vp, ctypes.c_long]
        L.nemo2d_run.argtypes = [vp, i, i]
        L.nemo2d_record_checksums.argtypes = [vp, i]
        L.nemo2d_checksum_log.argtypes = [vp, dp, i, ip]
        L.nemo2d_checksums.argtypes = [vp, dp]
        L.nemo2d_retrieve_results.argtypes = [vp] + [dp] * 5
        L.nemo2d_retrieve_aux.argtypes = [vp] + [dp] * 3
        L.nemo2d_timer_stop.argtypes = [vp, dp]
        u16p = ctypes.POINTER(ctypes.c_uint16)
        L.nemo2d_retrieve_grid.argtypes = [vp, i, dp]
        L.nemo2d_retrieve_classes.argtypes = [vp, u16p]
        L.nemo2d_classify.argtypes = [i, i, ip, dp, dp, u16p]
        for s in SYMBOLS:
            if s not in ("nemo2d_last_error", "nemo2d_version") and not s.startswith("nemo2d_nn_"):
                getattr(L, s).restype = i
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise Nemo2dError(rc, lib().nemo2d_last_error().decode())
    return rc


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def shapes(jpi, jpj):
    """numpy shapes of the T, U, V and pt arrays."""
    return {"T": (jpj, jpi), "U": (jpj, jpi + 1), "V": (jpj + 1, jpi), "pt": (jpj + 2, jpi + 2)}


def default_grid(jpi, jpj, dx=1000.0, dy=1000.0, dep_const=100.0):
    """The jphgr_msh=1 grid of original/nemolite2d.f90:180-279: land ring,
    open south edge, uniform spacing, f-plane at 50 N, constant depth."""
    s = shapes(jpi, jpj)
    pt = np.ones(s["pt"], dtype=np.int32)
    pt[:, 0] = 0
    pt[:, jpi + 1] = 0
    pt[jpj + 1, :] = 0
    pt[0, 1:jpi + 1] = -1
    g = {"pt": pt}
    for k, v in (("e1t", dx), ("e2t", dy), ("ht", dep_const)):
        g[k] = np.full(s["T"], v)
    for k, v in (("e1u", dx), ("e2u", dy), ("gphiu", 50.0), ("hu", dep_const)):
        g[k] = np.full(s["U"], v)
    for k, v in (("e1v", dx), ("e2v", dy), ("gphiv", 50.0), ("hv", dep_const)):
        g[k] = np.full(s["V"], v)
    g["e12t"] = g["e1t"] * g["e2t"]
    g["e12u"] = g["e1u"] * g["e2u"]
    g["e12v"] = g["e1v"] * g["e2v"]
    return g


GRID_FILE_ARRAYS = ("e1t", "e2t", "e1u", "e2u", "e1v", "e2v", "gphiu", "gphiv", "ht", "hu", "hv")


def write_grid_file(path, grid):
    """The grid as the Fortran driver reads it with jphgr_msh = 0: the
    interior sizes, then pt and the metrics, latitudes and depths, each in
    Fortran storage order (first index fastest), one value a line, 17
    significant digits (a double read back is the same double)."""
    pt = np.asarray(grid["pt"])
    jpj, jpi = pt.shape[0] - 2, pt.shape[1] - 2
    with open(path, "w") as f:
        f.write("%d %d\n" % (jpi, jpj))
        f.write("\n".join("%d" % v for v in pt.ravel()) + "\n")
        for k in GRID_FILE_ARRAYS:
            f.write("\n".join("%.17e" % v for v in np.asarray(grid[k], dtype=np.float64).ravel()) + "\n")


def initial_state(grid, sshn=None):
    """Initial now-state of original/nemolite2d.f90:324-371 (sshn given or 0;
    sshn_u, sshn_v interpolated with clamped indices; un = vn = 0)."""
    jpj, jpi = grid["e1t"].shape
    s = shapes(jpi, jpj)
    if sshn is None:  # at rest every interpolation is 0
        return {"sshn": np.zeros(s["T"]), "sshn_u": np.zeros(s["U"]), "sshn_v": np.zeros(s["V"]),
                "un": np.zeros(s["U"]), "vn": np.zeros(s["V"])}
    sshn = np.ascontiguousarray(sshn, dtype=np.float64)
    e12t = grid["e12t"]
    es = e12t * sshn
    # column ji takes T columns min(ji+1, jpi) and max(ji, 1) (1-based)
    i1 = np.minimum(np.arange(jpi + 1) + 1, jpi) - 1
    i2 = np.maximum(np.arange(jpi + 1), 1) - 1
    sshn_u = 0.5 * (es[:, i1] + es[:, i2]) / grid["e12u"]
    j1 = np.minimum(np.arange(jpj + 1) + 1, jpj) - 1
    j2 = np.maximum(np.arange(jpj + 1), 1) - 1
    sshn_v = 0.5 * (es[j1, :] + es[j2, :]) / grid["e12v"]
    return {"sshn": sshn, "sshn_u": np.ascontiguousarray(sshn_u), "sshn_v": np.ascontiguousarray(sshn_v),
            "un": np.zeros(s["U"]), "vn": np.zeros(s["V"])}


def classify(grid):
    """Class words the library derives from grid's mask and depths (host
    code, no device), in the mask's layout."""
    pt = np.ascontiguousarray(grid["pt"], dtype=np.int32)
    jpi, jpj = pt.shape[1] - 2, pt.shape[0] - 2
    hu = np.ascontiguousarray(grid["hu"], dtype=np.float64)
    hv = np.ascontiguousarray(grid["hv"], dtype=np.float64)
    out = np.zeros(pt.shape, dtype=np.uint16)
    _check(lib().nemo2d_classify(jpi, jpj, pt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(hu),
                                 _dp(hv), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))))
    return out


def tide(rtime, amp=0.2, period_hours=12.42):
    """Clamped-ssh tide of original/nemolite2d.f90:629-635 (host libm)."""
    omega_tide = 2.0 * 3.14159 / (period_hours * 3600.0)
    return amp * math.sin(omega_tide * rtime)


class Model:
    """One nemo2d model (device state + stream) on the current HIP device."""

    GRID_KEYS = ("e1t", "e2t", "e1u", "e2u", "e1v", "e2v", "e12t", "e12u", "e12v",
                 "gphiu", "gphiv", "ht", "hu", "hv")
    STATE_KEYS = ("sshn", "sshn_u", "sshn_v", "un", "vn")

    def __init__(self, jpi, jpj, dx=1000.0, dy=1000.0, dep_const=100.0, nit000=1, nitend=1,
                 irecord=1, rdt=20.0, cbfr=0.0015, visc=0.1, grid=None, state=None,
                 precision="fp64", compensated=True):
        """precision: "fp64" (the drop-in, bitwise the serial model) or "fp32"
        (single-precision state and stencils; compensated=True keeps the ssh
        accumulation in FP64)."""
        self.jpi, self.jpj = int(jpi), int(jpj)
        self.rdt = float(rdt)
        self.nit000, self.nitend = int(nit000), int(nitend)
        self.precision = precision
        h = ctypes.c_void_p()
        if precision == "fp64":
            _check(lib().nemo2d_create(ctypes.byref(h), self.jpi, self.jpj, dx, dy, dep_const,
                                       int(nit000), int(nitend), int(irecord), rdt, cbfr, visc))
        elif precision == "fp32":
            _check(lib().nemo2d_create_mixed(ctypes.byref(h), self.jpi, self.jpj, dx, dy, dep_const,
                                             int(nit000), int(nitend), int(irecord), rdt, cbfr,
                                             visc, 1 if compensated else 0))
        else:
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self._h = h
        g = default_grid(self.jpi, self.jpj, dx, dy, dep_const) if grid is None else grid
        self.set_grid(g)
        self.set_state(initial_state(g) if state is None else state)

    @classmethod
    def from_namelist(cls, nml, **kw):
        return cls(nml["jpiglo"] - 2, nml["jpjglo"] - 2, dx=nml["dx"], dy=nml["dy"],
                   dep_const=nml.get("dep", nml.get("dep_const")), nit000=nml["nit000"],
                   nitend=nml["nitend"], irecord=nml.get("irecord", 1), rdt=nml["rdt"],
                   cbfr=nml["cbfr"], visc=nml["visc"], **kw)

    def close(self):
        if getattr(self, "_h", None):
            lib().nemo2d_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _arr(self, g, k, shape):
        a = np.ascontiguousarray(g[k], dtype=np.float64)
        if a.shape != shape:
            raise ValueError("%s: shape %s, expected %s" % (k, a.shape, shape))
        return a

    def set_grid(self, g):
        s = shapes(self.jpi, self.jpj)
        kinds = {"e1t": "T", "e2t": "T", "e12t": "T", "ht": "T", "e1u": "U", "e2u": "U",
                 "e12u": "U", "gphiu": "U", "hu": "U", "e1v": "V", "e2v": "V", "e12v": "V",
                 "gphiv": "V", "hv": "V"}
        arrs = [self._arr(g, k, s[kinds[k]]) for k in self.GRID_KEYS]
        pt = np.ascontiguousarray(g["pt"], dtype=np.int32)
        if pt.shape != s["pt"]:
            raise ValueError("pt: shape %s, expected %s" % (pt.shape, s["pt"]))
        self._grid_keep = (pt, arrs)
        _check(lib().nemo2d_set_grid(self._h, pt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                     *[_dp(a) for a in arrs]))

    def set_state(self, st):
        s = shapes(self.jpi, self.jpj)
        kinds = {"sshn": "T", "sshn_u": "U", "sshn_v": "V", "un": "U", "vn": "V"}
        arrs = [self._arr(st, k, s[kinds[k]]) for k in self.STATE_KEYS]
        _check(lib().nemo2d_set_state(self._h, *[_dp(a) for a in arrs]))

    # the nine stages
    def continuity(self): _check(lib().nemo2d_continuity(self._h))
    def momentum_u(self): _check(lib().nemo2d_momentum_u(self._h))
    def momentum_v(self): _check(lib().nemo2d_momentum_v(self._h))
    def bc_ssh(self, rtime): _check(lib().nemo2d_bc_ssh(self._h, float(rtime)))
    def bc_solid_u(self): _check(lib().nemo2d_bc_solid_u(self._h))
    def bc_solid_v(self): _check(lib().nemo2d_bc_solid_v(self._h))
    def bc_flather_u(self): _check(lib().nemo2d_bc_flather_u(self._h))
    def bc_flather_v(self): _check(lib().nemo2d_bc_flather_v(self._h))
    def next(self): _check(lib().nemo2d_next(self._h))

    def step_stages(self, istp):
        """One step as the nine stage calls (original/nemolite2d.f90:375-383)."""
        rtime = float(istp) * self.rdt
        self.continuity(); self.momentum_u(); self.momentum_v()
        self.bc_ssh(rtime); self.bc_solid_u(); self.bc_solid_v()
        self.bc_flather_u(); self.bc_flather_v(); self.next()

    def step(self, istp):
        """One fused step."""
        _check(lib().nemo2d_step(self._h, int(istp)))

    def set_tide(self, amp, period_hours):
        """Tide amplitude (m) and period (hours) of the clamped-ssh boundary."""
        _check(lib().nemo2d_set_tide(self._h, float(amp), float(period_hours)))

    def set_stage_fusion(self, on):
        """Record the nine stage calls of a step and run them with next as one
        fused step (on, the default), or run each stage call at once (off)."""
        _check(lib().nemo2d_set_stage_fusion(self._h, 1 if on else 0))

    def set_variant(self, name):
        """Kernel of the fused step (see VARIANTS); results do not depend on it."""
        _check(lib().nemo2d_set_variant(self._h, VARIANTS[name] if isinstance(name, str) else int(name)))

    def last_variant(self):
        """Name of the kernel the last fused step ran (None before any)."""
        v = ctypes.c_int(0)
        _check(lib().nemo2d_last_variant(self._h, ctypes.byref(v)))
        return {n: k for k, n in VARIANTS.items()}.get(v.value) if v.value else None

    def run(self, nit000, nitend):
        _check(lib().nemo2d_run(self._h, int(nit000), int(nitend)))

    def record_checksums(self, capacity):
        _check(lib().nemo2d_record_checksums(self._h, int(capacity)))

    def checksum_log(self, max_steps):
        out = np.zeros((max(int(max_steps), 1), 3))
        n = ctypes.c_int(0)
        _check(lib().nemo2d_checksum_log(self._h, _dp(out), int(max_steps), ctypes.byref(n)))
        return out[:min(n.value, int(max_steps))]

    def checksums(self):
        out = np.zeros(3)
        rc = lib().nemo2d_checksums(self._h, _dp(out))
        if rc not in (OK, ERR_NONFINITE):
            _check(rc)
        return out

    def retrieve(self):
        s = shapes(self.jpi, self.jpj)
        r = {"sshn": np.zeros(s["T"]), "un": np.zeros(s["U"]), "vn": np.zeros(s["V"]),
             "ua": np.zeros(s["U"]), "va": np.zeros(s["V"]), "sshn_u": np.zeros(s["U"]),
             "sshn_v": np.zeros(s["V"]), "ssha": np.zeros(s["T"])}
        _check(lib().nemo2d_retrieve_results(self._h, _dp(r["sshn"]), _dp(r["un"]), _dp(r["vn"]),
                                             _dp(r["ua"]), _dp(r["va"])))
        _check(lib().nemo2d_retrieve_aux(self._h, _dp(r["sshn_u"]), _dp(r["sshn_v"]), _dp(r["ssha"])))
        return r

    # nemo2d_retrieve_grid's array order, then the Coriolis parameters and
    # the Flather coefficients
    GRID_READBACK = ("e1t", "e2t", "e1u", "e2u", "e1v", "e2v", "e12t", "e12u", "e12v",
                     "ht", "hu", "hv", "fcu", "fcv", "flu", "flv")

    def retrieve_grid(self):
        """{name: array} of every grid array as the device holds it."""
        s = shapes(self.jpi, self.jpj)
        out = {}
        for k, name in enumerate(self.GRID_READBACK):
            kind = "T" if name in ("e1t", "e2t", "e12t", "ht") else \
                "U" if name in ("e1u", "e2u", "e12u", "hu", "fcu", "flu") else "V"
            a = np.zeros(s[kind])
            _check(lib().nemo2d_retrieve_grid(self._h, k, _dp(a)))
            out[name] = a
        return out

    def retrieve_classes(self):
        """The per-point class words on the device, in the mask's layout."""
        out = np.zeros(shapes(self.jpi, self.jpj)["pt"], dtype=np.uint16)
        _check(lib().nemo2d_retrieve_classes(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))))
        return out

    def retrieve_ssh(self):
        """sshn alone (T layout)."""
        out = np.zeros(shapes(self.jpi, self.jpj)["T"])
        _check(lib().nemo2d_retrieve_results(self._h, _dp(out), None, None, None, None))
        return out

    def diagnostics(self):
        """(volume, energy, wet area) of the current state, FP64 reductions."""
        out = np.zeros(3)
        _check(lib().nemo2d_diagnostics(self._h, _dp(out)))
        return out

    REGIONS = ("Continuity", "Momentum", "BCs", "Next", "Fused step")

    def stage_timers(self, on=True):
        _check(lib().nemo2d_stage_timers(self._h, 1 if on else 0))

    def stage_times(self):
        """{region: (count, total seconds, standard error of the mean)}."""
        c = (ctypes.c_int * 5)()
        t, e = np.zeros(5), np.zeros(5)
        _check(lib().nemo2d_stage_times(self._h, c, _dp(t), _dp(e)))
        return {r: (int(c[k]), float(t[k]), float(e[k])) for k, r in enumerate(self.REGIONS)}

    def set_output_grid(self, xt, yt, ht):
        """T-point coordinates and depth for snapshot files."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (xt, yt, ht)]
        _check(lib().nemo2d_set_output_grid(self._h, *[_dp(a) for a in arrs]))

    def snapshot_begin(self, istp, directory="."):
        """Queue go2d_<istp>.dat of the current state (written behind later steps)."""
        _check(lib().nemo2d_snapshot_begin(self._h, int(istp), str(directory).encode()))

    def snapshot_wait(self):
        _check(lib().nemo2d_snapshot_wait(self._h))

    def closure_features(self, mean, scale, x, offset=0):
        """Closure input features of the now-state into the device array x
        (BF16 bits [jpj*jpi, 8], starting `offset` bytes in); mean, scale: 4
        values each."""
        m = np.ascontiguousarray(mean, dtype=np.float64)
        sc = np.ascontiguousarray(scale, dtype=np.float64)
        _check(lib().nemo2d_closure_features(self._h, _dp(m), _dp(sc),
                                             ctypes.c_void_p(x.ptr.value + int(offset))))

    def restrict_from(self, fine):
        """This (coarse) model's now-state becomes the restriction of the fine
        model's (on the device)."""
        _check(lib().nemo2d_closure_restrict(fine._h, self._h))

    def closure_target(self, pred, scale, y, offset=0):
        """FP32 (self - pred) / scale of the per-cell ssh, east-face u and
        north-face v into the device array y ([jpj*jpi, 8], `offset` bytes in)."""
        sc = np.ascontiguousarray(scale, dtype=np.float64)
        _check(lib().nemo2d_closure_target(self._h, pred._h, _dp(sc),
                                           ctypes.c_void_p(y.ptr.value + int(offset))))

    def closure_apply(self, z, scale):
        """Adds the device FP32 increments z [jpj*jpi, C] times scale (3
        values) to the now-state and re-derives sshn_u, sshn_v."""
        sc = np.ascontiguousarray(scale, dtype=np.float64)
        _check(lib().nemo2d_closure_apply(self._h, z.fptr(), int(z.shape[-1]), _dp(sc)))

    def timer_start(self):
        _check(lib().nemo2d_timer_start(self._h))

    def timer_stop(self):
        ms = ctypes.c_double(0.0)
        _check(lib().nemo2d_timer_stop(self._h, ctypes.byref(ms)))
        return ms.value


def format_go2d(xt, yt, ht, sshn, un, vn):
    """The text of a go2d snapshot file (original/nemolite2d.f90:773-799) for
    host arrays in the reference's layout: xt, yt, ht, sshn (jpj, jpi), un
    (jpj, jpi+1), vn (jpj+1, jpi).  Host code only."""
    jpj, jpi = np.shape(sshn)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (xt, yt, ht, sshn, un, vn)]
    n = ctypes.c_long(0)
    _check(lib().nemo2d_format_go2d(jpi, jpj, *[_dp(a) for a in arrs], None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    _check(lib().nemo2d_format_go2d(jpi, jpj, *[_dp(a) for a in arrs], buf, n.value, ctypes.byref(n)))
    return buf.raw[:n.value]


def stream_bandwidth(n, repeats=20):
    """GB/s of a 5-array-read, 5-array-write streaming kernel over n FP64
    elements (the fused step's traffic without the stencil)."""
    out = ctypes.c_double(0.0)
    _check(lib().nemo2d_stream_bandwidth(int(n), int(repeats), ctypes.byref(out)))
    return out.value


def seamount_grid(jpi, jpj, seed, n_mounts=32, dx=1000.0, dy=1000.0, dep_const=100.0,
                  height=(10.0, 60.0), radius=(20e3, 200e3), min_depth=10.0):
    """The reference grid (land ring, open south edge, f-plane) with seeded
    Gaussian seamounts: ht = max(dep_const - sum_k h_k exp(-r^2 / R_k^2),
    min_depth), h_k ~ U(height), R_k ~ U(radius), centres uniform over the
    basin; hu, hv are the means of the two adjacent ht (the outer faces take
    the adjacent cell's depth).  Returns (grid, parameters of the mounts)."""
    rng = np.random.default_rng(seed)
    g = default_grid(jpi, jpj, dx, dy, dep_const)
    hk = rng.uniform(*height, size=n_mounts)
    rk = rng.uniform(*radius, size=n_mounts)
    xk = rng.uniform(0.0, jpi * dx, size=n_mounts)
    yk = rng.uniform(0.0, jpj * dy, size=n_mounts)
    bump = np.zeros((jpj, jpi))
    for h, r, x0, y0 in zip(hk, rk, xk, yk):
        # only within 6 radii (exp(-36) ~ 2e-16 of the height beyond)
        i0 = max(int((x0 - 6 * r) / dx), 0)
        i1 = min(int((x0 + 6 * r) / dx) + 1, jpi)
        j0 = max(int((y0 - 6 * r) / dy), 0)
        j1 = min(int((y0 + 6 * r) / dy) + 1, jpj)
        if i0 >= i1 or j0 >= j1:
            continue
        x = (np.arange(i0, i1) + 0.5) * dx - x0
        y = (np.arange(j0, j1) + 0.5) * dy - y0
        # exp(-(x^2 + y^2)/r^2) as the product of its two factors (one
        # exponential per row and column instead of per cell)
        bump[j0:j1, i0:i1] += np.outer(h * np.exp(-y * y / (r * r)), np.exp(-x * x / (r * r)))
    ht = np.maximum(dep_const - bump, min_depth)
    hu = np.empty((jpj, jpi + 1))
    hu[:, 1:jpi] = 0.5 * (ht[:, :-1] + ht[:, 1:])
    hu[:, 0] = ht[:, 0]
    hu[:, jpi] = ht[:, -1]
    hv = np.empty((jpj + 1, jpi))
    hv[1:jpj, :] = 0.5 * (ht[:-1, :] + ht[1:, :])
    hv[0, :] = ht[0, :]
    hv[jpj, :] = ht[-1, :]
    g["ht"], g["hu"], g["hv"] = ht, hu, hv
    return g, {"height": hk, "radius": rk, "x": xk, "y": yk}


def field_errors(a, b):
    """RMS and max |a-b| of ssh (T), u (U interior faces), v (V interior faces)."""
    out = {}
    for k, sl in (("ssh", np.s_[:, :]), ("u", np.s_[:, 1:-1]), ("v", np.s_[1:-1, :])):
        key = {"ssh": "sshn", "u": "un", "v": "vn"}[k]
        d = (a[key] - b[key])[sl]
        out[k] = {"rms": float(np.sqrt(np.mean(d * d))), "max": float(np.max(np.abs(d))),
                  "ref_rms": float(np.sqrt(np.mean(b[key][sl] ** 2)))}
    return out


def ensemble_members(n, seed, size=(64, 1024), depth=(50.0, 200.0), amp=(0.1, 0.3),
                     period=(11.0, 13.0)):
    """Seeded ensemble: square basins of side round(exp(U(ln a, ln b))),
    constant depth U(depth) m, tide amplitude U(amp) m and period U(period) h."""
    rng = np.random.default_rng(seed)
    side = np.rint(np.exp(rng.uniform(np.log(size[0]), np.log(size[1]), n))).astype(np.int32)
    return {"jpi": side, "jpj": side.copy(), "dep": rng.uniform(*depth, n),
            "amp": rng.uniform(*amp, n), "period": rng.uniform(*period, n)}


class Ensemble:
    """Batched basins advanced by one persistent kernel per step."""

    def __init__(self, members, dx=1000.0, dy=1000.0, rdt=12.0, cbfr=0.0015, visc=0.1,
                 precision=64):
        self.members = {k: np.ascontiguousarray(v) for k, v in members.items()}
        self.n = len(self.members["jpi"])
        self.rdt = rdt
        m = self.members
        jpi = np.ascontiguousarray(m["jpi"], dtype=np.int32)
        jpj = np.ascontiguousarray(m["jpj"], dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        h = ctypes.c_void_p()
        _check(lib().nemo2d_ensemble_create(
            ctypes.byref(h), self.n, jpi.ctypes.data_as(ip), jpj.ctypes.data_as(ip),
            _dp(np.ascontiguousarray(m["dep"], dtype=np.float64)),
            _dp(np.ascontiguousarray(m["amp"], dtype=np.float64)),
            _dp(np.ascontiguousarray(m["period"], dtype=np.float64)),
            dx, dy, rdt, cbfr, visc, int(precision)))
        self._h = h
        self.cells = int(np.sum(jpi.astype(np.int64) * jpj))

    def close(self):
        if getattr(self, "_h", None):
            lib().nemo2d_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def reset(self):
        _check(lib().nemo2d_ensemble_reset(self._h))

    def tune(self, nsteps=20):
        c = ctypes.c_int(0)
        t = ctypes.c_double(0.0)
        _check(lib().nemo2d_ensemble_tune(self._h, int(nsteps), ctypes.byref(c), ctypes.byref(t)))
        return c.value, t.value

    def set_tiling(self, choice):
        _check(lib().nemo2d_ensemble_set_tiling(self._h, int(choice)))

    @staticmethod
    def tiling_shape(choice):
        tx, ty = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().nemo2d_ensemble_tiling(int(choice), ctypes.byref(tx), ctypes.byref(ty)))
        return tx.value, ty.value

    def run(self, nit000, nitend, graph=True):
        """Returns the device milliseconds of the steps."""
        ms = ctypes.c_double(0.0)
        _check(lib().nemo2d_ensemble_run(self._h, int(nit000), int(nitend), 1 if graph else 0,
                                         ctypes.byref(ms)))
        return ms.value

    def checksum_log(self, member, max_steps):
        out = np.zeros((max(int(max_steps), 1), 3))
        n = ctypes.c_int(0)
        _check(lib().nemo2d_ensemble_checksum_log(self._h, int(member), _dp(out), int(max_steps),
                                                  ctypes.byref(n)))
        return out[:min(n.value, int(max_steps))]

    def retrieve(self, member):
        ni, nj = int(self.members["jpi"][member]), int(self.members["jpj"][member])
        s = shapes(ni, nj)
        r = {"sshn": np.zeros(s["T"]), "un": np.zeros(s["U"]), "vn": np.zeros(s["V"])}
        _check(lib().nemo2d_ensemble_retrieve(self._h, int(member), _dp(r["sshn"]), _dp(r["un"]),
                                              _dp(r["vn"])))
        return r
