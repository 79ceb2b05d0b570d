"""ctypes view of the C oracle (oracle/nemo2d_oracle.c) -- test infrastructure.

Only the tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg
use this module, and only as the checker.  Arrays are exposed as numpy views
shaped ``(n2, n1)`` over the oracle's Fortran-ordered storage, so that
``view[jj - lb2, ji - lb1]`` is the reference's ``a(ji, jj)``.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnemo2d_oracle.so")


class _Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("dx", "dy", "dep_const", "rdt", "cbfr", "visc")]


_lib = None


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.o2d_create.restype = vp
        L.o2d_create.argtypes = [i, i]
        L.o2d_destroy.argtypes = [vp]
        L.o2d_set_params.argtypes = [vp, ctypes.POINTER(_Params)]
        L.o2d_array.restype = ctypes.POINTER(ctypes.c_double)
        L.o2d_array.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int * 4)]
        L.o2d_pt.restype = ctypes.POINTER(ctypes.c_int)
        L.o2d_pt.argtypes = [vp]
        for name in ("o2d_grid_default", "o2d_grid_areas", "o2d_initialise", "o2d_continuity",
                     "o2d_momentum_u", "o2d_momentum_v", "o2d_bc_solid_u", "o2d_bc_solid_v",
                     "o2d_bc_flather_u", "o2d_bc_flather_v", "o2d_next"):
            getattr(L, name).argtypes = [vp]
        L.o2d_bc_ssh.argtypes = [vp, d]
        L.o2d_bc.argtypes = [vp, d]
        L.o2d_tide.restype = d
        L.o2d_tide.argtypes = [d]
        L.o2d_set_tide.argtypes = [vp, d, d]
        L.o2d_step.argtypes = [vp, i]
        L.o2d_checksums.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.o2d_run.restype = d
        L.o2d_run.argtypes = [vp, i, i, ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


ARRAYS = ("e1t", "e2t", "e1u", "e2u", "e1f", "e2f", "e1v", "e2v", "e12t", "e12u", "e12v",
          "gphiu", "gphiv", "gphif", "xt", "yt", "ht", "hu", "hv", "hf",
          "sshb", "sshb_u", "sshb_v", "sshn", "sshn_u", "sshn_v", "ssha", "ssha_u", "ssha_v",
          "un", "vn", "ua", "va")


class Oracle:
    """One oracle model instance (the reference's PROGRAM state)."""

    def __init__(self, jpi, jpj, dx=1000.0, dy=1000.0, dep_const=100.0, rdt=20.0,
                 cbfr=0.0015, visc=0.1, default_grid=True):
        L = lib()
        self._m = L.o2d_create(int(jpi), int(jpj))
        if not self._m:
            raise ValueError("oracle: bad grid size %r x %r" % (jpi, jpj))
        self.jpi, self.jpj = int(jpi), int(jpj)
        self.params = dict(dx=dx, dy=dy, dep_const=dep_const, rdt=rdt, cbfr=cbfr, visc=visc)
        p = _Params(**self.params)
        L.o2d_set_params(self._m, ctypes.byref(p))
        self._views = {}
        if default_grid:
            L.o2d_grid_default(self._m)
            L.o2d_initialise(self._m)

    @classmethod
    def from_namelist(cls, nml, **kw):
        """nml: dict with the &namctl keys (jpiglo/jpjglo include the ring)."""
        return cls(nml["jpiglo"] - 2, nml["jpjglo"] - 2, dx=nml["dx"], dy=nml["dy"],
                   dep_const=nml.get("dep", nml.get("dep_const")), rdt=nml["rdt"],
                   cbfr=nml["cbfr"], visc=nml["visc"], **kw)

    def __del__(self):
        if getattr(self, "_m", None):
            lib().o2d_destroy(self._m)
            self._m = None

    def arr(self, name):
        """numpy view [jj - lb2, ji - lb1] of the named Fortran array, plus bounds."""
        if name not in self._views:
            b = (ctypes.c_int * 4)()
            p = lib().o2d_array(self._m, name.encode(), ctypes.byref(b))
            if not p:
                raise KeyError(name)
            lb1, ub1, lb2, ub2 = list(b)
            n1, n2 = ub1 - lb1 + 1, ub2 - lb2 + 1
            v = np.ctypeslib.as_array(p, shape=(n2 * n1,)).reshape(n2, n1)
            self._views[name] = (v, (lb1, ub1, lb2, ub2))
        return self._views[name][0]

    def bounds(self, name):
        self.arr(name)
        return self._views[name][1]

    @property
    def pt(self):
        p = lib().o2d_pt(self._m)
        return np.ctypeslib.as_array(p, shape=((self.jpj + 2) * (self.jpi + 2),)).reshape(
            self.jpj + 2, self.jpi + 2)

    # stages, in the reference's order
    def grid_areas(self): lib().o2d_grid_areas(self._m)
    def initialise(self): lib().o2d_initialise(self._m)
    def continuity(self): lib().o2d_continuity(self._m)
    def momentum_u(self): lib().o2d_momentum_u(self._m)
    def momentum_v(self): lib().o2d_momentum_v(self._m)
    def bc_ssh(self, rtime): lib().o2d_bc_ssh(self._m, float(rtime))
    def bc_solid_u(self): lib().o2d_bc_solid_u(self._m)
    def bc_solid_v(self): lib().o2d_bc_solid_v(self._m)
    def bc_flather_u(self): lib().o2d_bc_flather_u(self._m)
    def bc_flather_v(self): lib().o2d_bc_flather_v(self._m)
    def bc(self, rtime): lib().o2d_bc(self._m, float(rtime))
    def set_tide(self, amp, period_hours): lib().o2d_set_tide(self._m, float(amp), float(period_hours))
    def next(self): lib().o2d_next(self._m)
    def step(self, istp): lib().o2d_step(self._m, int(istp))

    def checksums(self):
        out = (ctypes.c_double * 3)()
        lib().o2d_checksums(self._m, out)
        return np.array(list(out))

    def run(self, nit000, nitend):
        """Step nit000..nitend; returns (per-step checksums (n,3), loop seconds)."""
        n = nitend - nit000 + 1
        sums = np.zeros((max(n, 0), 3))
        t = lib().o2d_run(self._m, int(nit000), int(nitend),
                          sums.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return sums, t

    def run_timed(self, nit000, nitend):
        """Step without checksums; returns loop seconds."""
        return lib().o2d_run(self._m, int(nit000), int(nitend), None)


def tide(rtime):
    return lib().o2d_tide(float(rtime))


def fortran_e16_8(x):
    """Format like Fortran's E16.8 edit descriptor (gocean_mod.F90:73-81)."""
    x = float(x)
    if x == 0.0:
        return "  0.00000000E+00"
    sign = "-" if x < 0 else ""
    s = "%.7E" % abs(x)           # d.dddddddE+xx, 8 significant digits
    mant, ex = s.split("E")
    e = int(ex) + 1
    digits = mant.replace(".", "")
    if abs(e) <= 99:
        body = "%s0.%sE%+03d" % (sign, digits, e)
    else:
        body = "%s0.%s%+04d" % (sign, digits, e)
    return body.rjust(16)
