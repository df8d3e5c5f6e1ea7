"""ctypes binding of libfgpe.so (include/fgpe_c.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FGPE_LIBRARY selects another build of the library (e.g. one with a user
# pointwise term, libfgpe_user_example.so)
LIB_PATH = os.environ.get("FGPE_LIBRARY") or os.path.join(_HERE, "libfgpe.so")

FGPE_PREC_F64 = 0
FGPE_PREC_MIXED = 1
FGPE_IMAG_RK4_RENORM = 0
FGPE_IMAG_HEAVY_BALL = 1

# every symbol include/fgpe_c.h declares
EXPORTS = (
    "fgpe_ctx_create", "fgpe_ctx_destroy", "fgpe_last_error", "fgpe_ctx_set_stream",
    "fgpe_ctx_sync", "fgpe_ctx_device_bytes", "fgpe_set_kspace", "fgpe_set_hamiltonian",
    "fgpe_upload_psi", "fgpe_download_psi", "fgpe_upload_vext", "fgpe_psi_device_ptr",
    "fgpe_fft3d_c2c_f64", "fgpe_fft3d_c2c_f32", "fgpe_laplacian_f64", "fgpe_apply_h",
    "fgpe_observables", "fgpe_adjoint_forward", "fgpe_adjoint_backward", "fgpe_calc_px", "fgpe_calc_cos", "fgpe_rk4_steps", "fgpe_imag_steps", "fgpe_imag_converged",
    "fgpe_timing_enable",
    "fgpe_timing_read", "fgpe_autotune", "fgpe_set_variant", "fgpe_get_variant", "fgpe_set_slabs",
    "fgpe_get_slabs", "fgpe_set_order", "fgpe_get_order", "fgpe_set_kmul_axis", "fgpe_get_kmul_axis",
    "fgpe_group_create", "fgpe_group_destroy", "fgpe_group_autotune", "fgpe_group_imag_steps",
    "fgpe_group_imag_status", "fgpe_group_imag_finish", "fgpe_graph_capture_imag", "fgpe_graph_launch", "fgpe_graph_destroy",
    "fgpe_psi_init", "fgpe_psi_add_noise", "fgpe_psi_imprint_vortex_tube", "fgpe_psi_force_lattice",
    "fgpe_psi_normalize", "fgpe_vext_init", "fgpe_vext_init_box", "fgpe_vext_add_cos", "fgpe_vext_commit",
    "fgpe_user_term", "fgpe_user_term_ids", "fgpe_imag_steps_noalloc", "fgpe_upload_psi_dev",
    "fgpe_download_psi_dev", "fgpe_upload_vext_real_dev",
)
# initialiser kinds and lattices (fgpe_psi_init, fgpe_vext_init, fgpe_psi_force_lattice)
INIT_HARMONIC, INIT_TUBE, INIT_VOID = 0, 1, 2
LATTICES = ("1d", "2d_triangular", "2d_square", "3d_hex", "3d_fcc", "3d_bcc", "3d_sc")
KCLASS_NAMES = ("xstage", "strided", "last", "fft_x", "other")


class HParams(C.Structure):
    _fields_ = [("ncomp", C.c_int), ("a", C.c_double * 3), ("quench_on", C.c_int),
                ("quench", C.c_double * 6), ("a_dd", C.c_double), ("dip", C.c_double * 3),
                ("sc_U", C.c_double), ("sc_R", C.c_double), ("sc_dim", C.c_int), ("vpx", C.c_double * 2),
                ("user", C.c_double * 8), ("user_id", C.c_int)]


class Geom(C.Structure):
    """Cell centres x_i = -xmax + i dx + dx/2 per axis."""
    _fields_ = [("xmax", C.c_double * 3), ("dx", C.c_double * 3)]

    @classmethod
    def of(cls, xmax, dx):
        return cls((C.c_double * 3)(*map(float, xmax)), (C.c_double * 3)(*map(float, dx)))


class FgpeError(RuntimeError):
    pass


def user_term_enabled() -> bool:
    """True when the loaded library was built with a user pointwise term."""
    v = C.c_int()
    check(load().fgpe_user_term(C.byref(v)))
    return bool(v.value)


def user_term_ids() -> list:
    """The ids of the user pointwise terms the loaded library compiles."""
    ids = (C.c_int * 16)()
    n = C.c_int()
    check(load().fgpe_user_term_ids(ids, 16, C.byref(n)))
    return list(ids[:min(n.value, 16)])


_lib = None
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


def load() -> C.CDLL:
    """Load libfgpe.so and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FgpeError(f"{LIB_PATH} is missing: build it with `python -m fgpe_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.fgpe_last_error.restype = C.c_char_p
    L.fgpe_ctx_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]
    L.fgpe_ctx_destroy.argtypes = [_vp]
    L.fgpe_ctx_set_stream.argtypes = [_vp, _vp]
    L.fgpe_ctx_sync.argtypes = [_vp]
    L.fgpe_ctx_device_bytes.argtypes = [_vp, C.POINTER(C.c_size_t)]
    L.fgpe_set_kspace.argtypes = [_vp, _dp, _dp, _dp, C.c_double]
    L.fgpe_set_hamiltonian.argtypes = [_vp, C.POINTER(HParams)]
    L.fgpe_upload_psi.argtypes = [_vp, C.c_int, _dp]
    L.fgpe_download_psi.argtypes = [_vp, C.c_int, _dp]
    L.fgpe_upload_vext.argtypes = [_vp, C.c_int, _dp]
    L.fgpe_psi_device_ptr.argtypes = [_vp, C.c_int, C.POINTER(_vp)]
    L.fgpe_upload_psi_dev.argtypes = [_vp, C.c_int, _vp]
    L.fgpe_download_psi_dev.argtypes = [_vp, C.c_int, _vp]
    L.fgpe_upload_vext_real_dev.argtypes = [_vp, C.c_int, _vp]
    L.fgpe_fft3d_c2c_f64.argtypes = [_vp, _vp, _vp, C.c_int]
    L.fgpe_fft3d_c2c_f32.argtypes = [_vp, _vp, _vp, C.c_int]
    L.fgpe_laplacian_f64.argtypes = [_vp, _vp, _vp]
    L.fgpe_apply_h.argtypes = [_vp, C.c_double, _vp, _vp]
    L.fgpe_observables.argtypes = [_vp, C.c_double, _dp]
    L.fgpe_rk4_steps.argtypes = [_vp, C.c_double, C.c_double, C.c_int]
    L.fgpe_imag_steps.argtypes = [_vp, C.c_int, C.c_double, C.c_double, _dp, C.c_double,
                                  C.c_int, _dp, C.POINTER(C.c_int)]
    L.fgpe_imag_steps_noalloc.argtypes = [_vp, C.c_int, C.c_double, C.c_double, _dp, C.c_double,
                                          C.c_int, _dp, C.POINTER(C.c_int)]
    L.fgpe_adjoint_forward.argtypes = [_vp, C.c_double, C.c_double, C.c_int, C.c_int]
    L.fgpe_adjoint_backward.argtypes = [_vp, _vp, _vp, _vp, _dp]
    L.fgpe_calc_px.argtypes = [_vp, C.c_int, _dp]
    L.fgpe_calc_cos.argtypes = [_vp, C.c_int, _dp, C.c_double, _dp]
    L.fgpe_imag_converged.argtypes = [_vp, C.POINTER(C.c_int)]
    L.fgpe_autotune.argtypes = [_vp, C.c_int, _dp]
    L.fgpe_set_variant.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.fgpe_get_variant.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.fgpe_set_slabs.argtypes = [_vp, C.c_int]
    L.fgpe_get_slabs.argtypes = [_vp, C.POINTER(C.c_int)]
    L.fgpe_set_order.argtypes = [_vp, C.c_int]
    L.fgpe_get_order.argtypes = [_vp, C.POINTER(C.c_int)]
    L.fgpe_set_kmul_axis.argtypes = [_vp, C.c_int]
    L.fgpe_get_kmul_axis.argtypes = [_vp, C.POINTER(C.c_int)]
    L.fgpe_timing_enable.argtypes = [_vp, C.c_int]
    L.fgpe_timing_read.argtypes = [_vp, _dp]
    _ip = C.POINTER(C.c_int)
    L.fgpe_group_create.argtypes = [C.POINTER(_vp), C.c_int, C.POINTER(_vp)]
    L.fgpe_group_destroy.argtypes = [_vp]
    L.fgpe_group_autotune.argtypes = [_vp, C.c_int, _dp]
    L.fgpe_group_imag_steps.argtypes = [_vp, C.c_double, C.c_double, _dp, C.c_double, C.c_int]
    L.fgpe_group_imag_status.argtypes = [_vp, _ip, _ip]
    L.fgpe_group_imag_finish.argtypes = [_vp, _ip, _ip]
    L.fgpe_graph_capture_imag.argtypes = [C.POINTER(_vp), C.c_int, C.c_double, C.c_double, _dp, C.c_double,
                                          C.c_int, C.POINTER(_vp)]
    L.fgpe_graph_launch.argtypes = [_vp]
    L.fgpe_graph_destroy.argtypes = [_vp]
    _gp = C.POINTER(Geom)
    L.fgpe_psi_init.argtypes = [_vp, C.c_int, C.c_int, _gp, _dp, C.c_double, C.c_double, C.c_ulonglong]
    L.fgpe_psi_add_noise.argtypes = [_vp, C.c_int, C.c_double, C.c_double, C.c_ulonglong, C.c_uint]
    L.fgpe_psi_imprint_vortex_tube.argtypes = [_vp, C.c_int, _gp, C.c_double]
    L.fgpe_psi_force_lattice.argtypes = [_vp, C.c_int, C.c_int, _gp, C.c_double, C.c_double, _dp, C.c_double]
    L.fgpe_psi_normalize.argtypes = [_vp, C.c_int, C.c_double, _gp, _dp]
    L.fgpe_vext_init.argtypes = [_vp, C.c_int, C.c_int, _gp, _dp]
    L.fgpe_vext_init_box.argtypes = [_vp, C.c_int, _gp, C.c_double, C.c_double, _dp]
    L.fgpe_vext_add_cos.argtypes = [_vp, C.c_int, _gp, C.c_double, C.c_double]
    L.fgpe_vext_commit.argtypes = [_vp, C.c_int, _dp]
    L.fgpe_user_term.argtypes = [C.POINTER(C.c_int)]
    L.fgpe_user_term_ids.argtypes = [C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        raise FgpeError(f"fgpe error {rc}: {load().fgpe_last_error().decode()}")


def _dptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] and a.dtype in (np.float64, np.complex128)
    return a.ctypes.data_as(_dp)


class Context:
    """Owns one fgpe_ctx (device buffers + stream) for an nx*ny*nz grid."""

    def __init__(self, nx, ny, nz, precision=FGPE_PREC_F64):
        L = load()
        h = _vp()
        check(L.fgpe_ctx_create(nx, ny, nz, precision, C.byref(h)))
        self.h = h
        self.shape = (nz, ny, nx)
        self.N = nx * ny * nz
        self.precision = precision
        self.ncomp = 1

    def close(self):
        if self.h:
            load().fgpe_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_stream(self, stream_ptr: int | None):
        check(load().fgpe_ctx_set_stream(self.h, stream_ptr))

    def sync(self):
        check(load().fgpe_ctx_sync(self.h))

    def device_bytes(self) -> int:
        b = C.c_size_t()
        check(load().fgpe_ctx_device_bytes(self.h, C.byref(b)))
        return b.value

    def set_kspace(self, kx, ky, kz, dv):
        kx, ky, kz = (np.ascontiguousarray(k, dtype=np.float64) for k in (kx, ky, kz))
        check(load().fgpe_set_kspace(self.h, _dptr(kx), _dptr(ky), _dptr(kz), float(dv)))

    def set_hamiltonian(self, ncomp=1, a=(0.0,), quench=None, a_dd=0.0, dip=(0.0, 0.0, 1.0),
                        sc_U=0.0, sc_R=0.0, sc_dim=3, vpx=(0.0, 0.0), user=(), user_id=0):
        aa = tuple(a) + (0.0,) * (3 - len(a))
        q = (0.0,) * 6 if quench is None else tuple(quench)
        vv = tuple(vpx) + (0.0,) * (2 - len(vpx))
        uu = tuple(map(float, user)) + (0.0,) * (8 - len(user))
        hp = HParams(ncomp, (C.c_double * 3)(*aa), 0 if quench is None else 1,
                     (C.c_double * 6)(*q), a_dd, (C.c_double * 3)(*dip), sc_U, sc_R, sc_dim,
                     (C.c_double * 2)(*vv), (C.c_double * 8)(*uu), int(user_id))
        check(load().fgpe_set_hamiltonian(self.h, C.byref(hp)))
        self.ncomp = ncomp

    def upload_psi(self, psi, comp=0):
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        if a.size == 2 * self.N and comp == 0:   # both components at once
            for c in range(2):
                self.upload_psi(a.reshape(2, -1)[c], c)
            return
        a = a.reshape(-1)
        assert a.size == self.N
        check(load().fgpe_upload_psi(self.h, comp, _dptr(a)))

    def download_psi(self, comp=None):
        if comp is None:
            if self.ncomp == 2:
                return np.stack([self.download_psi(0), self.download_psi(1)])
            comp = 0
        out = np.empty(self.shape, dtype=np.complex128)
        check(load().fgpe_download_psi(self.h, comp, _dptr(out)))
        return out

    def upload_vext(self, V, comp=0):
        a = np.ascontiguousarray(V, dtype=np.complex128)
        if a.size == 2 * self.N and comp == 0:
            for c in range(2):
                self.upload_vext(a.reshape(2, -1)[c], c)
            return
        a = a.reshape(-1)
        assert a.size == self.N
        check(load().fgpe_upload_vext(self.h, comp, _dptr(a)))

    # ---- initial state and trap built on the device (fgpe_psi_*, fgpe_vext_*)
    def init_psi(self, geom: Geom, kind=INIT_HARMONIC, ax=(1.0, 1.0, 1.0), rand=(0.0, 0.0), seed=0, comp=0):
        axa = (C.c_double * 3)(*map(float, ax))
        check(load().fgpe_psi_init(self.h, comp, kind, C.byref(geom), axa, float(rand[0]), float(rand[1]),
                                   int(seed) & 0xFFFFFFFFFFFFFFFF))

    def psi_add_noise(self, amp=(0.0, 0.0), seed=0, stream=1, comp=0):
        check(load().fgpe_psi_add_noise(self.h, comp, float(amp[0]), float(amp[1]),
                                        int(seed) & 0xFFFFFFFFFFFFFFFF, int(stream)))

    def psi_imprint_vortex_tube(self, geom: Geom, ncirc: float, comp=0):
        check(load().fgpe_psi_imprint_vortex_tube(self.h, comp, C.byref(geom), float(ncirc)))

    def psi_force_lattice(self, geom: Geom, lattice: str, amp: float, k: float, off=(0.0, 0.0, 0.0),
                          npart=1.0, comp=0):
        o = (C.c_double * 3)(*map(float, off))
        check(load().fgpe_psi_force_lattice(self.h, comp, LATTICES.index(lattice), C.byref(geom), float(amp),
                                            float(k), o, float(npart)))

    def psi_normalize(self, geom: Geom, npart: float, comp=0) -> float:
        out = C.c_double()
        check(load().fgpe_psi_normalize(self.h, comp, float(npart), C.byref(geom), C.byref(out)))
        return out.value

    def vext_init(self, geom: Geom, kind=INIT_HARMONIC, ax=(1.0, 1.0, 1.0), comp=0):
        axa = (C.c_double * 3)(*map(float, ax))
        check(load().fgpe_vext_init(self.h, comp, kind, C.byref(geom), axa))

    def vext_init_box(self, geom: Geom, V: float, B: float, w, comp=0):
        wa = (C.c_double * 3)(*map(float, w))
        check(load().fgpe_vext_init_box(self.h, comp, C.byref(geom), float(V), float(B), wa))

    def vext_add_cos(self, geom: Geom, A: float, n: float, comp=0):
        check(load().fgpe_vext_add_cos(self.h, comp, C.byref(geom), float(A), float(n)))

    def vext_commit(self, comp=0, download=False):
        """Make the assembled trap component comp's; returns it (N complex) if asked."""
        out = np.empty(self.N, np.complex128) if download else None
        check(load().fgpe_vext_commit(self.h, comp, _dptr(out) if download else None))
        return out

    def upload_psi_dev(self, dev_psi: int, comp=0):
        """psi from device memory (N complex doubles), on the context's stream."""
        check(load().fgpe_upload_psi_dev(self.h, comp, dev_psi))

    def download_psi_dev(self, dev_out: int, comp=0):
        check(load().fgpe_download_psi_dev(self.h, comp, dev_out))

    def upload_vext_real_dev(self, dev_v: int, comp=0):
        """A real FP64 trap from device memory (N doubles)."""
        check(load().fgpe_upload_vext_real_dev(self.h, comp, dev_v))

    def psi_device_ptr(self, comp=0) -> int:
        p = _vp()
        check(load().fgpe_psi_device_ptr(self.h, comp, C.byref(p)))
        return p.value

    def fft3d_f64(self, dev_in: int, dev_out: int, direction: int):
        check(load().fgpe_fft3d_c2c_f64(self.h, dev_in, dev_out, direction))

    def fft3d_f32(self, dev_in: int, dev_out: int, direction: int):
        check(load().fgpe_fft3d_c2c_f32(self.h, dev_in, dev_out, direction))

    def laplacian_f64(self, dev_in: int, dev_out: int):
        check(load().fgpe_laplacian_f64(self.h, dev_in, dev_out))

    def apply_h(self, t: float, dev_hmu: int, dev_hen: int | None):
        check(load().fgpe_apply_h(self.h, t, dev_hmu, dev_hen))

    def observables(self, t: float = 0.0) -> np.ndarray:
        out = np.zeros(3 * self.ncomp)
        check(load().fgpe_observables(self.h, t, _dptr(out)))
        return out

    def rk4_steps(self, dt: float, nsteps: int, t0: float = 0.0):
        check(load().fgpe_rk4_steps(self.h, dt, t0, nsteps))

    def adjoint_forward(self, dt: float, nsteps: int, ckpt_every: int, t0: float = 0.0):
        check(load().fgpe_adjoint_forward(self.h, dt, t0, nsteps, ckpt_every))

    def adjoint_backward(self, dev_lamT: int, dev_lam0: int | None, dev_gradV: int):
        """Device pointers to ncomp x N arrays (component-major); returns
        (dJ/da or dJ/da_11, dJ/da_dd, dJ/dU, dJ/da_22, dJ/da_12)."""
        g = np.zeros(5)
        check(load().fgpe_adjoint_backward(self.h, dev_lamT, dev_lam0, dev_gradV, _dptr(g)))
        return g

    def calc_px(self, comp=0) -> float:
        v = C.c_double()
        check(load().fgpe_calc_px(self.h, comp, C.byref(v)))
        return v.value

    def calc_cos(self, x, L, comp=0) -> float:
        xx = np.ascontiguousarray(x, dtype=np.float64)
        v = C.c_double()
        check(load().fgpe_calc_cos(self.h, comp, _dptr(xx), L, C.byref(v)))
        return v.value

    def imag_converged(self) -> bool:
        f = C.c_int(0)
        check(load().fgpe_imag_converged(self.h, C.byref(f)))
        return bool(f.value)

    def autotune(self, reps=5):
        """Time the plan variants per kernel class, the block orders, the
        k-space multiply axis (3D) and 1/2/4 column slabs of the strided/last/
        strided sequence; returns {class: (ms v0, ms v1), "xstage": (ms v0,
        ms v1, ms v2 or 0), "slabs": (ms 1, 2, 4)}."""
        out = np.zeros(10)
        check(load().fgpe_autotune(self.h, reps, _dptr(out)))
        return {"strided": tuple(out[0:2]), "last": tuple(out[2:4]), "xstage": (out[4], out[5], out[9]),
                "slabs": tuple(out[6:9])}

    def set_slabs(self, nslab: int):
        check(load().fgpe_set_slabs(self.h, nslab))

    def slabs(self) -> int:
        v = C.c_int()
        check(load().fgpe_get_slabs(self.h, C.byref(v)))
        return v.value

    def set_order(self, mask: int):
        check(load().fgpe_set_order(self.h, mask))

    def order(self) -> int:
        v = C.c_int()
        check(load().fgpe_get_order(self.h, C.byref(v)))
        return v.value

    def set_kmul_axis(self, axis: int):
        check(load().fgpe_set_kmul_axis(self.h, axis))

    def kmul_axis(self) -> int:
        v = C.c_int()
        check(load().fgpe_get_kmul_axis(self.h, C.byref(v)))
        return v.value

    def set_variant(self, kclass: int, axis: int, variant: int):
        check(load().fgpe_set_variant(self.h, kclass, axis, variant))

    def variants(self) -> dict:
        out = {}
        for name, k in (("xstage", 0), ("strided", 1), ("last", 2)):
            for axis in range(3):
                v = C.c_int()
                check(load().fgpe_get_variant(self.h, k, axis, C.byref(v)))
                out[(name, axis)] = v.value
        return out

    def timing_enable(self, on=True):
        check(load().fgpe_timing_enable(self.h, 1 if on else 0))

    def timing_read(self) -> dict:
        out = np.zeros(2 * len(KCLASS_NAMES))
        check(load().fgpe_timing_read(self.h, _dptr(out)))
        return {n: (out[2 * i], int(out[2 * i + 1])) for i, n in enumerate(KCLASS_NAMES)}

    def imag_steps(self, dt, nsteps, scheme=FGPE_IMAG_RK4_RENORM, beta=0.0, npart=1.0, eps=0.0,
                   history=False):
        hist = np.zeros((max(nsteps, 1), 3 * self.ncomp)) if history else None
        done = C.c_int(0)
        np_ = np.ascontiguousarray(np.broadcast_to(np.asarray(npart, dtype=np.float64), (2,)))
        check(load().fgpe_imag_steps(self.h, scheme, dt, beta, _dptr(np_), eps, nsteps,
                                     _dptr(hist) if history else _dp(), C.byref(done)))
        return (done.value, hist[:done.value]) if history else done.value


class Group:
    """Several same-shape contexts run as batched launches (fgpe_group_*).

    Keeps references to its contexts; close it before them."""

    def __init__(self, ctxs):
        self.ctxs = list(ctxs)
        arr = (_vp * len(self.ctxs))(*[c.h for c in self.ctxs])
        h = _vp()
        check(load().fgpe_group_create(arr, len(self.ctxs), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            load().fgpe_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, reps=3):
        out = np.zeros(10)
        check(load().fgpe_group_autotune(self.h, reps, _dptr(out)))
        return {"strided": tuple(out[0:2]), "last": tuple(out[2:4]), "xstage": (out[4], out[5], out[9])}

    def imag_steps(self, dt, nsteps, beta=0.0, npart=1.0, eps=0.0):
        np_ = np.ascontiguousarray(np.broadcast_to(np.asarray(npart, dtype=np.float64), (2,)))
        check(load().fgpe_group_imag_steps(self.h, dt, beta, _dptr(np_), eps, nsteps))

    def _flags(self, fn):
        n = len(self.ctxs)
        it = (C.c_int * n)()
        cv = (C.c_int * n)()
        check(fn(self.h, it, cv))
        return np.array(it[:], dtype=np.int64), np.array(cv[:], dtype=bool)

    def imag_status(self):
        """(iterations run, converged) per instance; the run can go on."""
        return self._flags(load().fgpe_group_imag_status)

    def imag_finish(self):
        """(iterations run, converged) per instance, after applying the pending scales."""
        return self._flags(load().fgpe_group_imag_finish)


class Graph:
    """`iters` heavy-ball iterations of every group captured as one hipGraph."""

    def __init__(self, groups, dt, iters, beta=0.0, npart=1.0, eps=0.0):
        self.groups = list(groups)
        arr = (_vp * len(self.groups))(*[g.h for g in self.groups])
        np_ = np.ascontiguousarray(np.broadcast_to(np.asarray(npart, dtype=np.float64), (2,)))
        h = _vp()
        check(load().fgpe_graph_capture_imag(arr, len(self.groups), dt, beta, _dptr(np_), eps, iters,
                                             C.byref(h)))
        self.h = h
        self.iters = iters

    def launch(self):
        check(load().fgpe_graph_launch(self.h))

    def close(self):
        if self.h:
            load().fgpe_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
