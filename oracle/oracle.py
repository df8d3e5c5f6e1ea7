"""ctypes front end of the CPU FP64 oracle (oracle/fgpe_oracle.c).

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py import
it; the product path (fgpe_amd) never does.

Complex arrays are numpy complex128 of shape (nz, ny, nx) (x fastest, the
reference's ID(x,y,z) layout, hpp/header.hpp:20); two-component states are
shape (2, nz, ny, nx).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfgpe_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


class OrcGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("kx", _dp), ("ky", _dp), ("kz", _dp), ("dv", C.c_double)]


class OrcGeom(C.Structure):
    _fields_ = [("xmax", C.c_double * 3), ("dx", C.c_double * 3), ("n", C.c_int * 3)]


class OrcHam(C.Structure):
    _fields_ = [("ncomp", C.c_int), ("a", C.c_double * 3), ("quench_on", C.c_int),
                ("quench", C.c_double * 6), ("vtilde", _dp), ("vpx", C.c_double * 2),
                ("user_id", C.c_int), ("user", C.c_double * 8)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_kaxis.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, _dp]
        L.orc_twiddle.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_fft1d.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int]
        L.orc_laplace.argtypes = [C.POINTER(OrcGrid), _dp, _dp]
        L.orc_vdd.argtypes = [C.POINTER(OrcGrid), C.c_double, _dp, _dp]
        L.orc_vsoftcore.argtypes = [C.POINTER(OrcGrid), C.c_int, C.c_double, C.c_double, _dp]
        L.orc_convolve.argtypes = [C.POINTER(OrcGrid), _dp, _dp, _dp]
        L.orc_quench_a.argtypes = [C.c_double, _dp]
        L.orc_quench_a.restype = C.c_double
        L.orc_hpsi.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcHam), C.c_double,
                               _dp, _dp, _dp, _dp]
        L.orc_norm.argtypes = [C.POINTER(OrcGrid), _dp]
        L.orc_norm.restype = C.c_double
        L.orc_obs2.argtypes = [C.POINTER(OrcGrid), C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.orc_rk4.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcHam), _dp, _dp,
                              C.c_double, C.c_double, C.c_int, C.c_int, C.c_double]
        L.orc_imag_heavy_ball.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcHam), _dp, _dp,
                                          C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_int, _dp]
        L.orc_imag_heavy_ball.restype = C.c_int
        L.orc_num_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        _gp = C.POINTER(OrcGeom)
        _u32p = C.POINTER(C.c_uint32)
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_noise.argtypes = [C.c_longlong, C.c_uint32, C.c_uint64, _dp, _dp]
        L.orc_psi_init.argtypes = [_gp, C.c_int, _dp, C.c_double, C.c_double, C.c_uint64, _dp]
        L.orc_psi_add_noise.argtypes = [_gp, C.c_double, C.c_double, C.c_uint64, C.c_uint32, _dp]
        L.orc_psi_vortex.argtypes = [_gp, C.c_double, _dp]
        L.orc_force.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_force.restype = C.c_double
        L.orc_psi_force.argtypes = [_gp, C.c_int, C.c_double, C.c_double, _dp, C.c_double, _dp]
        L.orc_psi_normalize.argtypes = [_gp, C.c_double, _dp]
        L.orc_psi_normalize.restype = C.c_double
        L.orc_vext_init.argtypes = [_gp, C.c_int, _dp, _dp]
        L.orc_vext_box.argtypes = [_gp, C.c_double, C.c_double, _dp, _dp]
        L.orc_vext_add_cos.argtypes = [_gp, C.c_double, C.c_double, _dp]
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def num_threads() -> int:
    return lib().orc_num_threads()


def set_threads(n: int) -> None:
    lib().orc_set_threads(n)


class Grid:
    """Grid of the reference: k axes from InitializeKspace, DV = DX*DY*DZ."""

    def __init__(self, nx, ny, nz, kx, ky, kz, dv):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.kx = np.ascontiguousarray(kx, dtype=np.float64)
        self.ky = np.ascontiguousarray(ky, dtype=np.float64)
        self.kz = np.ascontiguousarray(kz, dtype=np.float64)
        self.dv = float(dv)
        self._s = OrcGrid(nx, ny, nz, _p(self.kx), _p(self.ky), _p(self.kz), self.dv)

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def N(self):
        return self.nx * self.ny * self.nz


class Ham:
    def __init__(self, ncomp=1, a=(0.0, 0.0, 0.0), quench=None, vtilde=None, vpx=(0.0, 0.0), user=None,
                 user_id=1):
        """user: parameters of the test user pointwise term user_id (None: no term)."""
        self.ncomp = ncomp
        self.vtilde = None if vtilde is None else np.ascontiguousarray(vtilde, dtype=np.float64)
        q = (0.0,) * 6 if quench is None else tuple(quench)
        aa = tuple(a) + (0.0,) * (3 - len(a))
        vv = tuple(vpx) + (0.0,) * (2 - len(vpx))
        self._s = OrcHam(ncomp, (C.c_double * 3)(*aa), 0 if quench is None else 1,
                         (C.c_double * 6)(*q),
                         _p(self.vtilde) if self.vtilde is not None else _dp(),
                         (C.c_double * 2)(*vv), 0 if user is None else int(user_id),
                         (C.c_double * 8)(*(tuple(user or ()) + (0.0,) * (8 - len(user or ())))))


def kaxis(n, dk, kmax, active=True):
    out = np.zeros(n)
    lib().orc_kaxis(n, dk, kmax, 1 if active else 0, _p(out))
    return out


def twiddle(n, m):
    c, s = C.c_double(), C.c_double()
    lib().orc_twiddle(n, m, C.byref(c), C.byref(s))
    return c.value, s.value


def fft1d(x, sign=-1):
    a = _c128(x).copy()
    lib().orc_fft1d(a.size, _p(a.view(np.float64)), sign)
    return a


def fft3d(x, sign=-1):
    a = _c128(x).copy()
    nz, ny, nx = a.shape
    lib().orc_fft3d(nx, ny, nz, _p(a.view(np.float64)), sign)
    return a


def laplace(g: Grid, psi):
    a = _c128(psi)
    out = np.empty_like(a)
    lib().orc_laplace(C.byref(g._s), _p(a.view(np.float64)), _p(out.view(np.float64)))
    return out


def vdd(g: Grid, a_dd, d=(0.0, 0.0, 1.0)):
    out = np.empty(g.shape)
    dd = np.ascontiguousarray(d, dtype=np.float64)
    lib().orc_vdd(C.byref(g._s), a_dd, _p(dd), _p(out))
    return out


def vsoftcore(g: Grid, dim, R, U):
    out = np.empty(g.shape)
    lib().orc_vsoftcore(C.byref(g._s), dim, R, U, _p(out))
    return out


def convolve(g: Grid, vt, dens):
    vt = np.ascontiguousarray(vt, dtype=np.float64)
    dens = np.ascontiguousarray(dens, dtype=np.float64)
    out = np.empty(g.shape)
    lib().orc_convolve(C.byref(g._s), _p(vt), _p(dens), _p(out))
    return out


def quench_a(t, prms):
    p = np.ascontiguousarray(prms, dtype=np.float64)
    return lib().orc_quench_a(t, _p(p))


def hpsi(g: Grid, h: Ham, t, psi, V, energy=True):
    psi = _c128(psi)
    V = _c128(V)
    hmu = np.empty_like(psi)
    hen = np.empty_like(psi) if energy else None
    lib().orc_hpsi(C.byref(g._s), C.byref(h._s), t, _p(psi.view(np.float64)),
                   _p(V.view(np.float64)), _p(hmu.view(np.float64)),
                   _p(hen.view(np.float64)) if energy else _dp())
    return (hmu, hen) if energy else hmu


def norm(g: Grid, psi):
    return lib().orc_norm(C.byref(g._s), _p(_c128(psi).view(np.float64)))


def obs2(g: Grid, nrm, psi, o1psi, o2psi):
    a, b = C.c_double(), C.c_double()
    lib().orc_obs2(C.byref(g._s), nrm, _p(_c128(psi).view(np.float64)),
                   _p(_c128(o1psi).view(np.float64)), _p(_c128(o2psi).view(np.float64)),
                   C.byref(a), C.byref(b))
    return a.value, b.value


def rk4(g: Grid, h: Ham, psi, V, dt, nsteps, t0=0.0, imaginary=False, npart=1.0):
    a = _c128(psi).copy()
    V = _c128(V)
    lib().orc_rk4(C.byref(g._s), C.byref(h._s), _p(a.view(np.float64)), _p(V.view(np.float64)),
                  dt, t0, nsteps, 1 if imaginary else 0, npart)
    return a


def imag_heavy_ball(g: Grid, h: Ham, psi, V, dt, beta, npart, eps, niter):
    a = _c128(psi).copy()
    V = _c128(V)
    hist = np.zeros((max(niter, 1), 3 * h.ncomp))
    n = lib().orc_imag_heavy_ball(C.byref(g._s), C.byref(h._s), _p(a.view(np.float64)),
                                  _p(V.view(np.float64)), dt, beta, npart, eps, niter,
                                  _p(hist))
    return a, n, hist[:n]


# ---- initial state and trap (fgpe_oracle_init.c) --------------------------
LATTICES = ("1d", "2d_triangular", "2d_square", "3d_hex", "3d_fcc", "3d_bcc", "3d_sc")


def geom(n, xmax, dx) -> OrcGeom:
    return OrcGeom((C.c_double * 3)(*map(float, xmax)), (C.c_double * 3)(*map(float, dx)),
                   (C.c_int * 3)(*map(int, n)))


def _d3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return tuple(o)


def noise(i, stream, seed):
    a, b = C.c_double(), C.c_double()
    lib().orc_noise(int(i), int(stream), int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(a), C.byref(b))
    return a.value, b.value


def _ntot(g: OrcGeom):
    return int(g.n[0]) * int(g.n[1]) * int(g.n[2])


def psi_init(g: OrcGeom, kind, ax=(1.0, 1.0, 1.0), rand=(0.0, 0.0), seed=0):
    """kind 0 harmonic, 1 tube, 2 void (src/psi.cu:80-131), Philox noise."""
    out = np.empty(_ntot(g), np.complex128)
    lib().orc_psi_init(C.byref(g), int(kind), _p(_d3(ax)), float(rand[0]), float(rand[1]),
                       int(seed) & 0xFFFFFFFFFFFFFFFF, _p(out.view(np.float64)))
    return out


def psi_add_noise(g: OrcGeom, psi, amp, seed, stream):
    out = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    lib().orc_psi_add_noise(C.byref(g), float(amp[0]), float(amp[1]), int(seed) & 0xFFFFFFFFFFFFFFFF,
                            int(stream), _p(out.view(np.float64)))
    return out


def psi_vortex(g: OrcGeom, psi, ncirc):
    out = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    lib().orc_psi_vortex(C.byref(g), float(ncirc), _p(out.view(np.float64)))
    return out


def force(lattice, x, y, z, amp, k):
    return lib().orc_force(LATTICES.index(lattice), float(x), float(y), float(z), float(amp), float(k))


def psi_force(g: OrcGeom, psi, lattice, amp, k, off, npart):
    out = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    lib().orc_psi_force(C.byref(g), LATTICES.index(lattice), float(amp), float(k), _p(_d3(off)),
                        float(npart), _p(out.view(np.float64)))
    return out


def psi_normalize(g: OrcGeom, psi, npart):
    out = np.ascontiguousarray(psi, dtype=np.complex128).copy()
    nrm = lib().orc_psi_normalize(C.byref(g), float(npart), _p(out.view(np.float64)))
    return out, nrm


def vext_init(g: OrcGeom, kind, ax=(1.0, 1.0, 1.0)):
    out = np.empty(_ntot(g), np.float64)
    lib().orc_vext_init(C.byref(g), int(kind), _p(_d3(ax)), _p(out))
    return out


def vext_box(g: OrcGeom, V, B, w):
    out = np.empty(_ntot(g), np.float64)
    lib().orc_vext_box(C.byref(g), float(V), float(B), _p(_d3(w)), _p(out))
    return out


def vext_add_cos(g: OrcGeom, v, A, n):
    out = np.ascontiguousarray(v, dtype=np.float64).copy()
    lib().orc_vext_add_cos(C.byref(g), float(A), float(n), _p(out))
    return out
