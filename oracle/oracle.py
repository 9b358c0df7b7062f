"""ctypes front end for the oracle (test infrastructure only).

`Oracle("f32")` / `Oracle("f64")` load the C restatement built from
oracle/ddaoracle.c; `Reference()` loads oracle/_ref/libddaref.so, the
reference's own sources compiled in this container with oracle/ref_harness.cpp.
Both expose the same methods on numpy arrays so tests can compare them.

Array shapes follow the reference's storage: a field is complex64/complex128
[ni, nj, nl, 3]; Â is [2ni, 2nj, 2nl, 6] (e00 e01 e02 e11 e12 e22); the
occupancy map is int8 [ni, nj, nl].

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker.
"""

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_P = C.c_void_p
_I = C.c_int


def build(quiet=True):
    """Compile the restatement (and the reference when its tree is present)."""
    out = subprocess.run(["make", "-C", HERE, "-s"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)


def _ptr(a):
    return a.ctypes.data_as(_P)


class _Base:
    real = np.float32
    cplx = np.complex64
    prefix = "orc_"

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def _r(self, x):
        return self.real(x)

    # --- FFT pieces -----------------------------------------------------
    def fourier_phase(self, n):
        out = np.zeros(max(n, 2), self.cplx)
        self._f("fourier_phase")(_I(n), _ptr(out))
        return out[:n]

    def fft_line(self, x, inverse=False, legacy=False):
        x = np.ascontiguousarray(x, self.cplx)
        out = np.zeros_like(x)
        fn = self._f("fft_line_legacy" if legacy else "fft_line")
        rc = fn(_I(x.shape[0]), _ptr(x), _ptr(out), _I(int(inverse)))
        if rc != 0:
            raise ValueError("fft_line rejected n=%d" % x.shape[0])
        return out

    def amatrix(self, n, kd):
        ni, nj, nl = n
        out = np.zeros((2 * ni, 2 * nj, 2 * nl, 6), self.cplx)
        self._f("amatrix")(_I(ni), _I(nj), _I(nl), self._real_arg(kd), _ptr(out))
        return out

    def build_ahat(self, n, kd):
        ni, nj, nl = n
        out = np.zeros((2 * ni, 2 * nj, 2 * nl, 6), self.cplx)
        self._f("build_ahat")(_I(ni), _I(nj), _I(nl), self._real_arg(kd), _ptr(out))
        return out

    def build_ahat_compact(self, n, kd):
        """The Â grid of the compact mode: the transform of A(r) with its
        Nyquist planes zeroed, kept on the octant k <= n and mirrored to the
        whole grid, the components odd along a mirrored axis negated."""
        ni, nj, nl = n
        z = np.zeros((2 * ni, 2 * nj, 2 * nl, 6), self.cplx)
        self._f("build_ahat_ex")(_I(ni), _I(nj), _I(nl), self._real_arg(kd), _ptr(z), _I(1))
        out = np.empty_like(z)
        I = np.arange(2 * ni)
        J = np.arange(2 * nj)
        L = np.arange(2 * nl)
        fx, fy, fz = I > ni, J > nj, L > nl
        Ic, Jc, Lc = np.where(fx, 2 * ni - I, I), np.where(fy, 2 * nj - J, J), np.where(fz, 2 * nl - L, L)
        out[...] = z[Ic][:, Jc][:, :, Lc]
        FX, FY, FZ = np.meshgrid(fx, fy, fz, indexing="ij")
        for comp, (a, b) in ((1, (FX, FY)), (2, (FX, FZ)), (4, (FY, FZ))):
            flip = a != b
            out[..., comp][flip] = -out[..., comp][flip]
        return out

    def fft3d(self, x):
        x = np.ascontiguousarray(x, self.cplx)
        ni, nj, nl = x.shape[:3]
        spec = np.zeros((2 * ni, 2 * nj, 2 * nl, 3), self.cplx)
        self._f("fft3d")(_I(ni), _I(nj), _I(nl), _ptr(x), _ptr(spec))
        return spec

    def ifft3d(self, spec):
        spec = np.ascontiguousarray(spec, self.cplx)
        ni, nj, nl = (s // 2 for s in spec.shape[:3])
        y = np.zeros((ni, nj, nl, 3), self.cplx)
        self._f("ifft3d")(_I(ni), _I(nj), _I(nl), _ptr(spec), _ptr(y))
        return y

    def convolution(self, ahat, x, conj=False):
        x = np.ascontiguousarray(x, self.cplx)
        ahat = np.ascontiguousarray(ahat, self.cplx)
        ni, nj, nl = x.shape[:3]
        y = np.zeros_like(x)
        self._f("convolution")(_I(ni), _I(nj), _I(nl), _ptr(ahat), _ptr(x), _ptr(y), _I(int(conj)))
        return y

    # --- physics ---------------------------------------------------------
    def first_approx(self, occ, mv, alpha, kd, kinc, einc):
        occ = np.ascontiguousarray(occ, np.int8)
        ni, nj, nl = occ.shape
        mv = np.ascontiguousarray(mv, np.int32)
        alpha = np.ascontiguousarray(alpha, self.cplx)
        kinc = np.ascontiguousarray(kinc, self.real)
        einc = np.ascontiguousarray(einc, self.cplx)
        p0 = np.zeros((ni, nj, nl, 3), self.cplx)
        self._f("first_approx")(_I(ni), _I(nj), _I(nl), _ptr(mv), _ptr(alpha), _ptr(occ),
                                self._real_arg(kd), _ptr(kinc), _ptr(einc), _ptr(p0))
        return p0

    def cabs_cext(self, kd, kinc, einc, alpha, occ, p):
        occ = np.ascontiguousarray(occ, np.int8)
        ni, nj, nl = occ.shape
        kinc = np.ascontiguousarray(kinc, self.real)
        einc = np.ascontiguousarray(einc, self.cplx)
        alpha = np.ascontiguousarray(alpha, self.cplx)
        p = np.ascontiguousarray(p, self.cplx)
        ca = np.zeros(1, self.real)
        ce = np.zeros(1, self.real)
        self._f("cabs_cext")(self._real_arg(kd), _ptr(kinc), _ptr(einc), _I(ni), _I(nj), _I(nl),
                             _ptr(alpha), _ptr(occ), _ptr(p), _ptr(ca), _ptr(ce))
        return ca[0], ce[0]

    def far_field(self, occ, mv, kd, p, direction):
        occ = np.ascontiguousarray(occ, np.int8)
        ni, nj, nl = occ.shape
        mv = np.ascontiguousarray(mv, np.int32)
        p = np.ascontiguousarray(p, self.cplx)
        d = np.ascontiguousarray(direction, self.real)
        out = np.zeros(3, self.cplx)
        self._f("far_field")(_I(ni), _I(nj), _I(nl), _ptr(mv), _ptr(occ), self._real_arg(kd),
                             _ptr(p), _ptr(d), _ptr(out))
        return out

    def polarizability(self, model, kd, m, kinc=(0, 0, 1), einc=((1, 0), (0, 0), (0, 0))):
        mm = np.asarray([m.real, m.imag] if np.iscomplexobj(m) else m, self.real)
        k = np.asarray(kinc, self.real)
        e = np.asarray(einc, self.real).reshape(-1)
        if e.size == 3:
            e = np.stack([e, np.zeros(3, self.real)], axis=1).reshape(-1)
        out = np.zeros(2, self.real)
        self._f("polarizability")(_I(model), self._real_arg(kd), _ptr(mm), _ptr(k), _ptr(e), _ptr(out))
        return self.cplx(out[0] + 1j * out[1])

    def muller(self, kinc, ksc, normal, einc0, einc1, esc0, esc1):
        a = [np.ascontiguousarray(v, self.real) for v in (kinc, ksc, normal)]
        b = [np.ascontiguousarray(v, self.cplx) for v in (einc0, einc1, esc0, esc1)]
        out = np.zeros((4, 4), self.real)
        self._f("muller")(*[_ptr(v) for v in a], *[_ptr(v) for v in b], _ptr(out))
        return out


class Oracle(_Base):
    """The C restatement, float (`f32`) or double (`f64`) build."""

    def __init__(self, precision="f32"):
        path = os.path.join(HERE, "liboracle_%s.so" % precision)
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.precision = precision
        if precision == "f64":
            self.real, self.cplx = np.float64, np.complex128
        for name in ("orc_fourier_phase", "orc_amatrix", "orc_build_ahat", "orc_fft3d", "orc_ifft3d",
                     "orc_convolution", "orc_matvec", "orc_first_approx", "orc_cabs_cext",
                     "orc_far_field", "orc_polarizability", "orc_muller", "orc_dda_fft", "orc_solve_b"):
            getattr(self.lib, name).restype = C.c_int if name in ("orc_dda_fft", "orc_solve_b") else None

    def _real_arg(self, x):
        return C.c_float(x) if self.real is np.float32 else C.c_double(x)

    def set_phase_rounding(self, correctly_rounded):
        """Test switch: incident/scattered phases from correctly rounded
        double cos/sin (as the HIP library computes them) instead of the
        reference's float libm calls (the default)."""
        self.lib.orc_set_phase_rounding(_I(int(bool(correctly_rounded))))

    def matvec(self, ahat, occ, mv, alpha, x, conj=False):
        occ = np.ascontiguousarray(occ, np.int8)
        ni, nj, nl = occ.shape
        x = np.ascontiguousarray(x, self.cplx)
        y = np.zeros_like(x)
        self.lib.orc_matvec(_I(ni), _I(nj), _I(nl), _ptr(np.ascontiguousarray(ahat, self.cplx)),
                            _ptr(np.ascontiguousarray(mv, np.int32)),
                            _ptr(np.ascontiguousarray(alpha, self.cplx)), _ptr(occ), _ptr(x), _ptr(y),
                            _I(int(conj)))
        return y

    def solve_b(self, solver, occ, mv, alpha, ahat, b, max_step=20, err_max=1e-3, tau=0.0):
        """The solver iteration from the source b (P_0 = b); returns (P, info, err_hist)."""
        occ = np.ascontiguousarray(occ, np.int8)
        ni, nj, nl = occ.shape
        b = np.ascontiguousarray(b, self.cplx).reshape(ni, nj, nl, 3)
        p = np.zeros_like(b)
        hist = np.full(max_step + 1, np.nan)
        info = np.zeros(4, np.int32)
        rc = self.lib.orc_solve_b(
            _I(solver), self._real_arg(tau), _I(ni), _I(nj), _I(nl),
            _ptr(np.ascontiguousarray(alpha, self.cplx)), _ptr(np.ascontiguousarray(mv, np.int32)), _ptr(occ),
            _I(max_step), self._real_arg(err_max), _ptr(np.ascontiguousarray(ahat, self.cplx)), _ptr(b),
            _ptr(hist), _ptr(info), _ptr(p))
        if rc != 0:
            raise RuntimeError("orc_solve_b failed")
        return p, info, hist[: int(info[0]) + 1]

    def target_ellipsoid(self, n, b_a, c_a):
        occ = np.zeros(n, np.int8)
        cnt = self.lib.orc_target_ellipsoid(_I(n[0]), _I(n[1]), _I(n[2]), self._real_arg(b_a),
                                            self._real_arg(c_a), _ptr(occ))
        return occ, cnt

    def dda_fft(self, solver, occ, mv, alpha, kd, kinc, einc, ahat, dirs, max_step=20, err_max=1e-3,
                tau=0.0, return_p=False):
        """One DDA_FFT solve.  dirs: [ndirs, 3].  Returns a dict."""
        occ = np.ascontiguousarray(occ, np.int8)
        ni, nj, nl = occ.shape
        dirs = np.ascontiguousarray(dirs, self.real).reshape(-1, 3)
        nd = dirs.shape[0]
        esc = np.zeros((nd, 3), self.cplx)
        ca = np.zeros(1, self.real)
        ce = np.zeros(1, self.real)
        hist = np.full(max_step + 1, np.nan)
        info = np.zeros(4, np.int32)
        p = np.zeros((ni, nj, nl, 3), self.cplx) if return_p else None
        rc = self.lib.orc_dda_fft(
            _I(solver), self._real_arg(tau), _I(ni), _I(nj), _I(nl),
            _ptr(np.ascontiguousarray(alpha, self.cplx)), _ptr(np.ascontiguousarray(mv, np.int32)),
            _ptr(occ), _I(max_step), self._real_arg(err_max), self._real_arg(kd),
            _ptr(np.ascontiguousarray(kinc, self.real)), _ptr(np.ascontiguousarray(einc, self.cplx)),
            _ptr(np.ascontiguousarray(ahat, self.cplx)), _I(nd), _ptr(dirs), _ptr(esc), _ptr(ca),
            _ptr(ce), _ptr(hist), _ptr(info), _ptr(p) if p is not None else None)
        if rc != 0:
            raise RuntimeError("orc_dda_fft failed")
        steps = int(info[0])
        out = dict(esc=esc, cabs=ca[0], cext=ce[0], err_hist=hist[: steps + 1], steps=steps,
                   step_min=int(info[1]), restored=bool(info[2]), matvecs=int(info[3]))
        if return_p:
            out["p"] = p
        return out


class Reference(_Base):
    """The reference's own functions, compiled from its sources (oracle/_ref)."""

    prefix = "ref_"

    def __init__(self):
        path = os.path.join(REF_DIR, "libddaref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)

    def _real_arg(self, x):
        return C.c_float(x)

    def target(self, kind, n, p1=1.0, p2=1.0):
        occ = np.zeros(n, np.int8)
        cnt = self.lib.ref_target(_I(kind), _I(n[0]), _I(n[1]), _I(n[2]), C.c_float(p1), C.c_float(p2),
                                  _ptr(occ))
        return occ, cnt

    def dda_fft(self, solver, occ, mv, alpha, kd, kinc, einc, ahat, dirs, max_step=20, err_max=1e-3,
                tau=0.0, log_path=None):
        occ = np.ascontiguousarray(occ, np.int8)
        ni, nj, nl = occ.shape
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        nd = dirs.shape[0]
        esc = np.zeros((nd, 3), np.complex64)
        ca = np.zeros(1, np.float32)
        ce = np.zeros(1, np.float32)
        self.lib.ref_dda_fft(
            _I(solver), C.c_float(tau), _I(ni), _I(nj), _I(nl),
            _ptr(np.ascontiguousarray(alpha, np.complex64)), _ptr(np.ascontiguousarray(mv, np.int32)),
            _ptr(occ), _I(max_step), C.c_float(err_max), C.c_float(kd),
            _ptr(np.ascontiguousarray(kinc, np.float32)), _ptr(np.ascontiguousarray(einc, np.complex64)),
            _ptr(np.ascontiguousarray(ahat, np.complex64)), _I(1), _I(nd), _ptr(dirs), _ptr(esc),
            _ptr(ca), _ptr(ce), (log_path or "").encode())
        return dict(esc=esc, cabs=ca[0], cext=ce[0])


def parse_steps(log_text):
    """Discrepancies printed by DDA_FFT ('Step %4d discrepancy=%9.7Lf')."""
    out = []
    for line in log_text.splitlines():
        line = line.strip()
        if "replaced" in line:      # the line after it repeats the best step
            break
        if line.startswith("Step") and "discrepancy=" in line:
            head, rest = line.split("discrepancy=")
            out.append((int(head.split()[1]), float(rest.split()[0])))
    return out
