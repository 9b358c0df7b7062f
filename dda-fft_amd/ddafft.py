"""ctypes binding of libddafft_amd.so (the C ABI declared in include/ddafft.h).

Arrays follow the reference's storage: a field is complex64 [ni, nj, nl, 3],
the occupancy map int8 [ni, nj, nl], Â complex64 [2ni, 2nj, 2nl, 6].

The library must have been built (``make -C dda-fft_amd`` or
``__graft_entry__.build()``); importing this module never falls back to any
other implementation: if the shared library is missing, it raises.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libddafft_amd.so")

SOLVERS = {"cg": 0, "aa": 1, "bicgstab": 2, "simple": 3}

EXPORTS = (
    "ddafft_last_error", "ddafft_version", "ddafft_create", "ddafft_destroy", "ddafft_get_ahat",
    "ddafft_set_target", "ddafft_set_directions", "ddafft_convolution", "ddafft_matvec", "ddafft_solve",
    "ddafft_get_dipoles", "ddafft_cabs_cext", "ddafft_far_field", "ddafft_profile", "ddafft_time_matvec",
)

_lib = None


class DdafftError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Loads the shared library (once) and declares the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise DdafftError("libddafft_amd.so not built: %s (run make -C dda-fft_amd)" % path)
    lib = C.CDLL(path)
    P, I, F = C.c_void_p, C.c_int, C.c_float
    lib.ddafft_last_error.restype = C.c_char_p
    lib.ddafft_version.restype = I
    lib.ddafft_create.argtypes = [I, I, I, F, I, C.POINTER(P)]
    lib.ddafft_destroy.argtypes = [P]
    lib.ddafft_get_ahat.argtypes = [P, P]
    lib.ddafft_set_target.argtypes = [P, P, I, P]
    lib.ddafft_set_directions.argtypes = [P, I, P]
    lib.ddafft_convolution.argtypes = [P, P, P, I]
    lib.ddafft_matvec.argtypes = [P, P, P, P, I]
    lib.ddafft_solve.argtypes = [P, I, I, F, I, F, P, P, P, I, P, P, P, P, P, P]
    lib.ddafft_get_dipoles.argtypes = [P, I, P]
    lib.ddafft_cabs_cext.argtypes = [P, P, P, P, P, P, P]
    lib.ddafft_far_field.argtypes = [P, P, I, P, P]
    lib.ddafft_profile.argtypes = [P, I, P, P]
    lib.ddafft_time_matvec.argtypes = [P, I, I, P, P]
    for name in EXPORTS:
        if name != "ddafft_last_error":
            getattr(lib, name).restype = I
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise DdafftError("ddafft error %d: %s" % (rc, _lib.ddafft_last_error().decode()))


class Context:
    """One lattice (ni, nj, nl) with spacing kd, Â resident on the GPU."""

    def __init__(self, dims, kd, nrhs_max=2):
        lib = load()
        self.dims = tuple(int(v) for v in dims)
        self.kd = float(kd)
        self.nrhs_max = int(nrhs_max)
        h = C.c_void_p()
        _check(lib.ddafft_create(*self.dims, C.c_float(kd), self.nrhs_max, C.byref(h)))
        self._h = h
        self.nmat = None

    def close(self):
        if getattr(self, "_h", None):
            _lib.ddafft_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------
    def ahat(self):
        ni, nj, nl = self.dims
        out = np.zeros((2 * ni, 2 * nj, 2 * nl, 6), np.complex64)
        _check(_lib.ddafft_get_ahat(self._h, _p(out)))
        return out

    def set_target(self, occ, mv):
        occ = np.ascontiguousarray(occ, np.int8)
        assert occ.shape == self.dims
        mv = np.ascontiguousarray(mv, np.int32)
        _check(_lib.ddafft_set_target(self._h, _p(occ), int(mv.size), _p(mv)))
        self.nmat = int(mv.size)

    def set_directions(self, dirs):
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        _check(_lib.ddafft_set_directions(self._h, dirs.shape[0], _p(dirs)))
        self.ndirs = dirs.shape[0]

    def convolution(self, x, conj=False):
        x = np.ascontiguousarray(x, np.complex64)
        y = np.zeros_like(x)
        _check(_lib.ddafft_convolution(self._h, _p(x), _p(y), int(conj)))
        return y

    def matvec(self, alpha, x, conj=False):
        x = np.ascontiguousarray(x, np.complex64)
        a = np.ascontiguousarray(alpha, np.complex64)
        y = np.zeros_like(x)
        _check(_lib.ddafft_matvec(self._h, _p(a), _p(x), _p(y), int(conj)))
        return y

    def solve(self, solver, alpha, kinc, einc, dirs=None, max_step=20, err_max=1e-3, tau=0.0, ndirs=None):
        """alpha: [nrhs, nmat] (or [nmat]); einc: [nrhs, 3] (or [3]) complex.
        Returns a dict of per-right-hand-side arrays."""
        solver = SOLVERS.get(solver, solver)
        alpha = np.ascontiguousarray(np.atleast_2d(alpha), np.complex64)
        einc = np.ascontiguousarray(np.atleast_2d(einc), np.complex64)
        nrhs = einc.shape[0]
        if alpha.shape[0] == 1 and nrhs > 1:
            alpha = np.ascontiguousarray(np.repeat(alpha, nrhs, axis=0))
        kinc = np.ascontiguousarray(kinc, np.float32)
        if dirs is not None:
            dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
            nd = dirs.shape[0]
        else:
            nd = int(ndirs if ndirs is not None else getattr(self, "ndirs", 0))
        esc = np.zeros((nrhs, max(nd, 1), 3), np.complex64)
        cabs = np.zeros(nrhs, np.float32)
        cext = np.zeros(nrhs, np.float32)
        hist = np.zeros((nrhs, max_step + 1), np.float64)
        info = np.zeros((nrhs, 4), np.int32)
        _check(_lib.ddafft_solve(self._h, nrhs, int(solver), C.c_float(tau), int(max_step), C.c_float(err_max),
                                 _p(alpha), _p(kinc), _p(einc), nd, _p(dirs), _p(esc), _p(cabs), _p(cext),
                                 _p(hist), _p(info)))
        out = []
        for r in range(nrhs):
            steps = int(info[r, 0])
            out.append(dict(esc=esc[r, :nd], cabs=cabs[r], cext=cext[r], err_hist=hist[r, : steps + 1],
                            steps=steps, step_min=int(info[r, 1]), restored=bool(info[r, 2]),
                            matvecs=int(info[r, 3])))
        return out

    def dipoles(self, rhs=0):
        p = np.zeros(self.dims + (3,), np.complex64)
        _check(_lib.ddafft_get_dipoles(self._h, int(rhs), _p(p)))
        return p

    def cabs_cext(self, alpha, kinc, einc, p):
        ca = np.zeros(1, np.float32)
        ce = np.zeros(1, np.float32)
        _check(_lib.ddafft_cabs_cext(self._h, _p(np.ascontiguousarray(alpha, np.complex64)),
                                     _p(np.ascontiguousarray(kinc, np.float32)),
                                     _p(np.ascontiguousarray(einc, np.complex64)),
                                     _p(np.ascontiguousarray(p, np.complex64)), _p(ca), _p(ce)))
        return ca[0], ce[0]

    def far_field(self, p, dirs):
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        esc = np.zeros((dirs.shape[0], 3), np.complex64)
        _check(_lib.ddafft_far_field(self._h, _p(np.ascontiguousarray(p, np.complex64)), dirs.shape[0], _p(dirs),
                                     _p(esc)))
        return esc

    def profile(self, enable):
        """Returns (summed pass durations [5] in ms, matvecs timed) since the
        last call, then enables or disables per-pass event timing."""
        ms = np.zeros(5, np.float64)
        n = C.c_longlong(0)
        _check(_lib.ddafft_profile(self._h, int(enable), _p(ms), C.byref(n)))
        return ms, int(n.value)

    def time_matvec(self, nrhs=1, reps=20):
        k = np.zeros(5, np.float32)
        t = np.zeros(1, np.float32)
        _check(_lib.ddafft_time_matvec(self._h, int(nrhs), int(reps), _p(k), _p(t)))
        return k, float(t[0])
