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
    "ddafft_gradient", "ddafft_polarizability", "ddafft_polarizability_deriv", "ddafft_target_ellipsoid",
    "ddafft_target_parallelepiped", "ddafft_target_bisphere", "ddafft_geometry_fixed", "ddafft_muller",
    "ddafft_muller_average", "ddafft_solve_ex", "ddafft_solve_batch", "ddafft_rng_worker_seed", "ddafft_rng_next",
    "ddafft_target_random_ellipsoid", "ddafft_target_random_parallelepiped", "ddafft_target_nibble_ellipsoid",
    "ddafft_target_hierarchical_ellipsoid", "ddafft_target_sphere_spec", "ddafft_worm_count", "ddafft_target_worms",
    "ddafft_target_file_read", "ddafft_target_file_write", "ddafft_orientation_mc", "ddafft_rotate",
    "ddafft_target_mixture", "ddafft_agglomerate_count", "ddafft_target_agglomerate",
    "ddafft_host_alloc", "ddafft_host_free", "ddafft_set_sum_mode",
    "ddafft_set_tuning", "ddafft_solve_noalloc", "ddafft_set_ahat_mode",
)


class Job(C.Structure):
    """ddafft_job of include/ddafft.h."""
    _fields_ = [("nrhs", C.c_int), ("alpha", C.c_void_p), ("kinc", C.c_void_p), ("einc", C.c_void_p),
                ("ndirs", C.c_int), ("dirs", C.c_void_p), ("dirs_per_rhs", C.c_int), ("esc", C.c_void_p),
                ("cabs", C.c_void_p), ("cext", C.c_void_p), ("info", C.c_void_p), ("exact_phases", C.c_int)]

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
    lib.ddafft_set_sum_mode.argtypes = [P, I]
    lib.ddafft_set_ahat_mode.argtypes = [P, I]
    lib.ddafft_set_tuning.argtypes = [P, I, I]
    lib.ddafft_convolution.argtypes = [P, P, P, I]
    lib.ddafft_matvec.argtypes = [P, P, P, P, I]
    lib.ddafft_solve.argtypes = [P, I, I, F, I, F, P, P, P, I, P, P, P, P, P, P]
    lib.ddafft_solve_noalloc.argtypes = [P, I, I, F, I, F, P, P, P, I, P, P, P, P, P, P]
    lib.ddafft_get_dipoles.argtypes = [P, I, P]
    lib.ddafft_cabs_cext.argtypes = [P, P, P, P, P, P, P]
    lib.ddafft_far_field.argtypes = [P, P, I, P, P]
    lib.ddafft_profile.argtypes = [P, I, P, P]
    lib.ddafft_time_matvec.argtypes = [P, I, I, P, P]
    lib.ddafft_gradient.argtypes = [P, I, I, I, F, I, F, P, P, P, P, P, I, P, P, P, P, P, P, P, P, P, P, P]
    lib.ddafft_solve_ex.argtypes = [P, C.POINTER(Job), I, F, I, F, P]
    lib.ddafft_solve_batch.argtypes = [I, P, P, I, F, I, F, I]
    lib.ddafft_orientation_mc.argtypes = [P, P]
    lib.ddafft_rotate.argtypes = [P, I, P, P]
    lib.ddafft_rng_worker_seed.argtypes = [C.c_ulonglong, I]
    lib.ddafft_rng_next.argtypes = [P]
    D = C.c_double
    lib.ddafft_target_mixture.argtypes = [I, P, I, D, P, P]
    lib.ddafft_agglomerate_count.argtypes = [I, I, I, D, D]
    lib.ddafft_target_agglomerate.argtypes = [I, I, I, D, D, C.c_longlong, P, P, P]
    lib.ddafft_host_alloc.argtypes = [C.c_size_t, C.POINTER(P)]
    lib.ddafft_host_free.argtypes = [P]
    lib.ddafft_polarizability.argtypes = [I, F, P, P, P, P]
    lib.ddafft_polarizability_deriv.argtypes = [I, F, P, P, P, P]
    D = C.c_double
    lib.ddafft_target_ellipsoid.argtypes = [I, I, I, D, D, P]
    lib.ddafft_target_parallelepiped.argtypes = [I, I, I, D, D, P]
    lib.ddafft_target_bisphere.argtypes = [I, I, I, D, P]
    lib.ddafft_geometry_fixed.argtypes = [I, I, D, D, D, D, D, P, P, P, P]
    lib.ddafft_muller.argtypes = [P] * 8
    lib.ddafft_muller_average.argtypes = [I, I, P, P, P, P, P, P, P]
    for name in EXPORTS:
        if name not in ("ddafft_last_error", "ddafft_rng_worker_seed"):
            getattr(lib, name).restype = I
    lib.ddafft_rng_worker_seed.restype = C.c_ulonglong
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

    def set_sum_mode(self, mode):
        """0: FP64 sums of the observables (default); 1: the reference's
        sequential float sums in its cell order (ddafft_set_sum_mode)."""
        _check(_lib.ddafft_set_sum_mode(self._h, int(mode)))

    def set_ahat_mode(self, mode):
        """0: the reference's Â over the whole padded grid (default,
        bit-identical convolutions); 1: one octant of the Â of A(r) with its
        Nyquist planes zeroed, mirrored in the X pass (ddafft_set_ahat_mode)."""
        _check(_lib.ddafft_set_ahat_mode(self._h, int(mode)))

    def set_tuning(self, x_blocks=None, x_rb=None, graph=None, fuse_zy=None, ff_mfma=None, ff_group=None):
        """Launch schedule (ddafft_set_tuning); None leaves a setting as it
        is.  x_blocks: X-pass launch size up to which right-hand-side groups
        are spread over blocks (0: the tuned table's); x_rb: right-hand sides
        per X tile (0: tuned, 1, 2 or 4); graph: iterations replayed as
        captured HIP graphs (True / False, or -1: tuned); fuse_zy: z- and
        y-forward transforms of an i-plane in one kernel where it fits (True /
        False, or -1: tuned); ff_mfma: the far field on the matrix cores
        (True / False, or -1: the library's choice); ff_group: the far field
        by groups of directions sharing their z component where the set
        splits so (True / False, or -1: the default, on)."""
        if x_blocks is not None:
            _check(_lib.ddafft_set_tuning(self._h, 0, int(x_blocks)))
        if x_rb is not None:
            _check(_lib.ddafft_set_tuning(self._h, 1, int(x_rb)))
        if graph is not None:
            _check(_lib.ddafft_set_tuning(self._h, 2, -1 if graph == -1 else int(bool(graph))))
        if fuse_zy is not None:
            _check(_lib.ddafft_set_tuning(self._h, 3, -1 if fuse_zy == -1 else int(bool(fuse_zy))))
        if ff_mfma is not None:
            _check(_lib.ddafft_set_tuning(self._h, 4, -1 if ff_mfma == -1 else int(bool(ff_mfma))))
        if ff_group is not None:
            _check(_lib.ddafft_set_tuning(self._h, 5, -1 if ff_group == -1 else int(bool(ff_group))))

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

    def solve(self, solver, alpha, kinc, einc, dirs=None, max_step=20, err_max=1e-3, tau=0.0, ndirs=None,
              noalloc=False):
        """alpha: [nrhs, nmat] (or [nmat]); einc: [nrhs, 3] (or [3]) complex.
        Returns a dict of per-right-hand-side arrays.  noalloc: through
        ddafft_solve_noalloc (no device or page-locked allocation)."""
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
        fn = _lib.ddafft_solve_noalloc if noalloc else _lib.ddafft_solve
        _check(fn(self._h, nrhs, int(solver), C.c_float(tau), int(max_step), C.c_float(err_max), _p(alpha), _p(kinc),
                  _p(einc), nd, _p(dirs), _p(esc), _p(cabs), _p(cext), _p(hist), _p(info)))
        out = []
        for r in range(nrhs):
            steps = int(info[r, 0])
            out.append(dict(esc=esc[r, :nd], cabs=cabs[r], cext=cext[r], err_hist=hist[r, : steps + 1],
                            steps=steps, step_min=int(info[r, 1]), restored=bool(info[r, 2]),
                            matvecs=int(info[r, 3])))
        return out

    def gradient(self, solver, alpha, dalpha_dm, m, kinc, einc, weights, dirs=None, normals=None, max_step=20,
                 err_max=1e-3, tau=0.0, ndirs=None, grad_chi_out=None):
        """Forward + adjoint gradients (ddafft_gradient).

        alpha, dalpha_dm: [npol, nmat] complex; m: [nmat] complex; einc: [npol, 3];
        weights: [nobj, 2 + ndirs] rows (w_ext, w_abs, w_s11...); normals: [ndirs, 3].
        Returns a dict: f [nobj], grad_m [nobj, nmat, 2], grad_chi [nobj, ni, nj, nl, 2],
        esc [npol, ndirs, 3], cabs/cext [npol], s11 [ndirs], info [npol*(1+nobj), 4].
        grad_chi_out: an optional float32 array of that shape to receive grad_chi
        (for example one from pinned_empty, reused across calls).
        """
        solver = SOLVERS.get(solver, solver)
        alpha = np.ascontiguousarray(np.atleast_2d(alpha), np.complex64)
        dadm = np.ascontiguousarray(np.atleast_2d(dalpha_dm), np.complex64)
        einc = np.ascontiguousarray(np.atleast_2d(einc), np.complex64)
        npol = einc.shape[0]
        mm = np.ascontiguousarray(m, np.complex64)
        kinc = np.ascontiguousarray(kinc, np.float32)
        if dirs is not None:
            dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
            nd = dirs.shape[0]
        else:
            nd = int(ndirs if ndirs is not None else getattr(self, "ndirs", 0))
        if normals is not None:
            normals = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(np.atleast_2d(weights), np.float32)
        nobj = w.shape[0]
        assert w.shape[1] == 2 + nd
        f = np.zeros(nobj, np.float32)
        esc = np.zeros((npol, max(nd, 1), 3), np.complex64)
        cabs = np.zeros(npol, np.float32)
        cext = np.zeros(npol, np.float32)
        s11 = np.zeros(max(nd, 1), np.float32)
        gm = np.zeros((nobj, self.nmat, 2), np.float32)
        shape = (nobj,) + self.dims + (2,)
        if grad_chi_out is not None:
            if grad_chi_out.shape != shape or grad_chi_out.dtype != np.float32 or not grad_chi_out.flags.c_contiguous:
                raise ValueError("grad_chi_out must be a C-contiguous float32 array of shape %s" % (shape,))
            gchi = grad_chi_out
        else:
            gchi = np.zeros(shape, np.float32)
        info = np.zeros((npol * (1 + nobj), 4), np.int32)
        want_s11 = normals is not None and npol == 2 and nd > 0
        _check(_lib.ddafft_gradient(self._h, npol, nobj, int(solver), C.c_float(tau), int(max_step),
                                    C.c_float(err_max), _p(alpha), _p(dadm), _p(mm), _p(kinc), _p(einc), nd,
                                    _p(dirs), _p(normals), _p(w), _p(f), _p(esc), _p(cabs), _p(cext),
                                    _p(s11) if want_s11 else None, _p(gm), _p(gchi), _p(info)))
        return dict(f=f, grad_m=gm, grad_chi=gchi, esc=esc[:, :nd], cabs=cabs, cext=cext, s11=s11[:nd],
                    info=info)

    def _job(self, alpha, kinc, einc, dirs, ndirs, exact_phases=False):
        """Builds a Job and its output arrays; keeps the inputs alive in the result."""
        alpha = np.ascontiguousarray(np.atleast_2d(alpha), np.complex64)
        einc = np.ascontiguousarray(np.atleast_2d(einc), np.complex64)
        nrhs = einc.shape[0]
        if alpha.shape[0] == 1 and nrhs > 1:
            alpha = np.ascontiguousarray(np.repeat(alpha, nrhs, axis=0))
        kinc = np.ascontiguousarray(np.broadcast_to(np.asarray(kinc, np.float32).reshape(-1, 3), (nrhs, 3)))
        per_rhs = 0
        if dirs is not None:
            dirs = np.ascontiguousarray(dirs, np.float32)
            if dirs.ndim == 3:
                assert dirs.shape[0] == nrhs
                per_rhs = 1
                nd = dirs.shape[1]
            else:
                dirs = dirs.reshape(-1, 3)
                nd = dirs.shape[0]
        else:
            nd = int(ndirs if ndirs is not None else getattr(self, "ndirs", 0))
        out = dict(esc=np.zeros((nrhs, max(nd, 1), 3), np.complex64), cabs=np.zeros(nrhs, np.float32),
                   cext=np.zeros(nrhs, np.float32), info=np.zeros((nrhs, 4), np.int32), nd=nd,
                   keep=(alpha, kinc, einc, dirs))
        j = Job(nrhs, _p(alpha), _p(kinc), _p(einc), nd, _p(dirs), per_rhs, _p(out["esc"]), _p(out["cabs"]),
                _p(out["cext"]), _p(out["info"]), int(bool(exact_phases)))
        return j, out

    @staticmethod
    def _results(out):
        res = []
        for r in range(out["info"].shape[0]):
            res.append(dict(esc=out["esc"][r, : out["nd"]], cabs=out["cabs"][r], cext=out["cext"][r],
                            steps=int(out["info"][r, 0]), step_min=int(out["info"][r, 1]),
                            restored=bool(out["info"][r, 2]), matvecs=int(out["info"][r, 3])))
        return res

    def solve_ex(self, solver, alpha, kinc, einc, dirs=None, max_step=20, err_max=1e-3, tau=0.0, ndirs=None,
                 exact_phases=False):
        """Like solve, with kinc [nrhs, 3] per right-hand side and dirs either
        [ndirs, 3] (shared) or [nrhs, ndirs, 3] (one set per right-hand side).
        exact_phases: incident phases of every incidence from the C library's
        cosf/sinf on the host, as the reference evaluates them."""
        solver = SOLVERS.get(solver, solver)
        j, out = self._job(alpha, kinc, einc, dirs, ndirs, exact_phases)
        hist = np.zeros((j.nrhs, max_step + 1), np.float64)
        _check(_lib.ddafft_solve_ex(self._h, C.byref(j), int(solver), C.c_float(tau), int(max_step),
                                    C.c_float(err_max), _p(hist)))
        res = self._results(out)
        for r, d in enumerate(res):
            d["err_hist"] = hist[r, : d["steps"] + 1]
        return res

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


# ----------------------------------------------------------------------
# Host-side functions of the library (no GPU needed)
# ----------------------------------------------------------------------

def _einc_arg(einc):
    if einc is None:
        return None
    e = np.asarray(einc, np.complex64).reshape(3)
    return np.ascontiguousarray(np.stack([e.real, e.imag], 1).reshape(-1), np.float32)


def polarizability(model, kd, m, kinc=None, einc=None):
    """alpha of model 1..5 (float, as the reference rounds it)."""
    load()
    mm = np.asarray([np.real(m), np.imag(m)], np.float32)
    k = None if kinc is None else np.ascontiguousarray(kinc, np.float32)
    e = _einc_arg(einc)
    out = np.zeros(2, np.float32)
    _check(_lib.ddafft_polarizability(int(model), C.c_float(kd), _p(mm), _p(k), _p(e), _p(out)))
    return np.complex64(out[0] + 1j * out[1])


def polarizability_deriv(model, kd, m, kinc=None, einc=None):
    """d alpha / d m of model 1..5."""
    load()
    mm = np.asarray([np.real(m), np.imag(m)], np.float32)
    k = None if kinc is None else np.ascontiguousarray(kinc, np.float32)
    e = _einc_arg(einc)
    out = np.zeros(2, np.float32)
    _check(_lib.ddafft_polarizability_deriv(int(model), C.c_float(kd), _p(mm), _p(k), _p(e), _p(out)))
    return np.complex64(out[0] + 1j * out[1])


def target_ellipsoid(dims, b_a=1.0, c_a=1.0):
    load()
    occ = np.zeros(tuple(dims), np.int8)
    cnt = _lib.ddafft_target_ellipsoid(*[int(v) for v in dims], float(b_a), float(c_a), _p(occ))
    if cnt < 0:
        raise DdafftError("ellipsoid does not fit the lattice")
    return occ, cnt


class _Pinned:
    """Owner of one page-locked host allocation (freed with the last array view)."""

    def __init__(self, nbytes):
        load()
        self.ptr = C.c_void_p()
        _check(_lib.ddafft_host_alloc(int(nbytes), C.byref(self.ptr)))
        self.nbytes = int(nbytes)

    def __del__(self):
        if _lib is not None and self.ptr:
            _lib.ddafft_host_free(self.ptr)
            self.ptr = C.c_void_p()


def pinned_empty(shape, dtype=np.float32):
    """A numpy array in page-locked host memory (ddafft_host_alloc): results
    copied into it arrive at DMA speed.  The memory lives as long as the array."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape)) * dtype.itemsize
    owner = _Pinned(max(n, 1))
    buf = (C.c_char * max(n, 1)).from_address(owner.ptr.value)
    buf._owner = owner
    arr = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)
    return arr


def target_mixture(base, material, fraction, state):
    """Each cell of material 1 in `base` becomes `material` with probability
    `fraction` (ddafft_target_mixture); `state` is a uint64[1] stream state,
    advanced.  Returns (occ, changed)."""
    load()
    base = np.ascontiguousarray(base, np.int8)
    occ = np.zeros_like(base)
    n = _lib.ddafft_target_mixture(int(base.size), _p(base), int(material), float(fraction), _p(state), _p(occ))
    if n < 0:
        raise DdafftError("bad mixture arguments")
    return occ, n


def target_agglomerate(dims, r_mono, packing, state, max_attempts=10 ** 7):
    """Random-packed agglomerate of sphere monomers (ddafft_target_agglomerate).
    Returns (occ, count, centres[K, 3])."""
    load()
    k = _lib.ddafft_agglomerate_count(*[int(v) for v in dims], float(r_mono), float(packing))
    if k < 0:
        raise DdafftError("bad agglomerate arguments")
    occ = np.zeros(tuple(dims), np.int8)
    cen = np.zeros((k, 3), np.float64)
    n = _lib.ddafft_target_agglomerate(*[int(v) for v in dims], float(r_mono), float(packing), int(max_attempts),
                                       _p(state), _p(occ), _p(cen))
    if n == -3:
        raise DdafftError("agglomerate placement gave up after %d draws" % max_attempts)
    if n < 0:
        raise DdafftError("bad agglomerate arguments")
    return occ, n, cen


def geometry_fixed(nplanes, nangles, theta_min, theta_step, psi=0.0, iota=0.0, kappa=0.0):
    """Directions [nplanes, nangles, 3], plane normals [nplanes, 3], kinc [3],
    einc [2, 3] complex of a fixed orientation."""
    load()
    dirs = np.zeros((nplanes, nangles, 3), np.float32)
    normals = np.zeros((nplanes, 3), np.float32)
    kinc = np.zeros(3, np.float32)
    einc = np.zeros((2, 3, 2), np.float32)
    _check(_lib.ddafft_geometry_fixed(int(nplanes), int(nangles), float(theta_min), float(theta_step), float(psi),
                                      float(iota), float(kappa), _p(dirs), _p(normals), _p(kinc), _p(einc)))
    return dirs, normals, kinc, (einc[..., 0] + 1j * einc[..., 1]).astype(np.complex64)


def solve_batch(ctxs, jobs, solver, max_step=20, err_max=1e-3, tau=0.0, width=4):
    """Solves several contexts together (ddafft_solve_batch).  jobs: one dict
    per context with keys alpha, kinc, einc and optionally dirs; returns one
    list of per-right-hand-side dicts per context."""
    load()
    solver = SOLVERS.get(solver, solver)
    built = [ctx._job(jb["alpha"], jb["kinc"], jb["einc"], jb.get("dirs"), jb.get("ndirs"),
                      jb.get("exact_phases", False)) for ctx, jb in zip(ctxs, jobs)]
    arr = (Job * len(built))(*[b[0] for b in built])
    hs = (C.c_void_p * len(ctxs))(*[c._h.value for c in ctxs])
    _check(_lib.ddafft_solve_batch(len(ctxs), hs, arr, int(solver), C.c_float(tau), int(max_step),
                                   C.c_float(err_max), int(width)))
    return [Context._results(b[1]) for b in built]



def orientation_mc(state):
    """Draws one Monte-Carlo orientation from the random stream `state` (a
    uint64 array of one element, advanced in place); returns the 3x3 matrix."""
    load()
    rot = np.zeros(9, np.float32)
    _check(_lib.ddafft_orientation_mc(_p(state), _p(rot)))
    return rot.reshape(3, 3)


def rotate(rot, v):
    """Applies an orientation matrix to vectors [..., 3] as main.cpp does (float)."""
    load()
    v = np.ascontiguousarray(v, np.float32)
    flat = v.reshape(-1, 3)
    out = np.zeros_like(flat)
    _check(_lib.ddafft_rotate(_p(np.ascontiguousarray(rot, np.float32)), flat.shape[0], _p(flat), _p(out)))
    return out.reshape(v.shape)
