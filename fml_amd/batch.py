"""Prepared, HBM-resident calls for batching and HIP-graph capture (no
reference counterpart; the batching form of the hot path).

Each object owns its device inputs, output and workspace, so its ``run()``
can be enqueued any number of times, captured into a ``fml_amd.graphs.Graph``
together with other prepared calls, and replayed with one launch.  New input
values of the same shape are copied in with ``set_*`` between runs.

* ``CoulombCall``   sorted Coulomb matrices of a molecule batch
                    (frepresentations.f90:23-110), one workgroup per molecule.
* ``AradDescribeCall`` ARAD descriptors of a molecule batch
                    (arad.py:70-145), one wavefront per molecule; its
                    outputs are a ready AradKernelCall input, so descriptor
                    and kernel can share one graph.
* ``AradKernelCall`` a molecular ARAD kernel (fget_kernels_arad /
                    fget_symmetric_kernels_arad / fget_kernel_arad,
                    farad_kernels.f90:95-755) over size-sorted molecule-pair
                    tiles planned on the device; ``autotune()`` picks its
                    tile shape by timing candidates on the device.
* ``KRRCall``       plain FP64 KRR fit + predict of a molecule batch
                    (Coulomb matrices, kernels, cho_solve, K^T alpha).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native, tuning
from .arad import MAX_DEVICE_ATOMS, describe_into
from .device import check_offsets, empty, to_device, workspace

_MODES = {"global": _native.ARAD_GLOBAL, "symmetric": _native.ARAD_SYMMETRIC,
          "kernel": _native.ARAD_KERNEL}


class CoulombCall:
    """Sorted Coulomb matrices of a fixed batch of molecules.

    charges (total,), coords (total, 3), offsets (nmol+1,).  ``cm`` (nmol,
    nmax(nmax+1)/2) f64 (and ``cm32`` when float32) are resident outputs."""

    def __init__(self, charges, coords, offsets, nmax=23, float32=False):
        self.z = to_device(charges, np.float64)
        self.r = to_device(np.asarray(coords).reshape(-1, 3), np.float64)
        check_offsets(offsets, int(self.z.shape[0]))
        self.off = to_device(offsets, np.int32)
        self.nmol = int(self.off.shape[0]) - 1
        self.nmax = int(nmax)
        counts = np.diff(np.asarray(offsets, dtype=np.int64))
        if self.nmol and counts.max() > self.nmax:
            raise ValueError(f"molecule {int(np.argmax(counts > self.nmax))} has more than "
                             f"nmax={self.nmax} atoms")
        packed = self.nmax * (self.nmax + 1) // 2
        self.cm = empty((self.nmol, packed), np.float64)
        self.cm32 = empty((self.nmol, packed), np.float32) if float32 else None
        _native.call("fmlh_init")

    def set_coords(self, coords):
        """New coordinates for the same molecules (same atom order)."""
        self.r.copy_(to_device(np.asarray(coords).reshape(-1, 3), np.float64))

    def run(self):
        _native.call("fmlh_coulomb_sorted", _native.ptr(self.z), _native.ptr(self.r),
                     _native.ptr(self.off), self.nmol, self.nmax, _native.ptr(self.cm),
                     _native.ptr(self.cm32), None, None, _native.stream_ptr())


class AradDescribeCall:
    """ARAD descriptors of a fixed batch of molecules.

    charges (total,), coords (total, 3), offsets (nmol+1,).  Resident
    outputs ``q`` (nmol, amax, 5, amax), ``zt`` (nmol, amax, 2) and ``n``
    (nmol,) int32 in the layouts ``AradKernelCall`` takes (pass them
    straight in: no copy is made), and ``status`` (nmol,) int32, 1 for a
    molecule with a charge without a PTP entry."""

    def __init__(self, charges, coords, offsets, amax=23, cut=5.0):
        self.z = to_device(charges, np.float64)
        self.r = to_device(np.asarray(coords).reshape(-1, 3), np.float64)
        check_offsets(offsets, int(self.z.shape[0]))
        self.off = to_device(offsets, np.int32)
        self.nmol = int(self.off.shape[0]) - 1
        self.amax = int(amax)
        self.cut = float(cut)
        if not 1 <= self.amax <= MAX_DEVICE_ATOMS:
            raise ValueError(f"amax must be in 1..{MAX_DEVICE_ATOMS}, got {amax}")
        counts = np.diff(np.asarray(offsets, dtype=np.int64))
        if self.nmol and counts.max() > self.amax:
            raise ValueError(f"molecule {int(np.argmax(counts > self.amax))} has more than "
                             f"amax={self.amax} atoms")
        self.q = empty((self.nmol, self.amax, 5, self.amax), np.float64)
        self.zt = empty((self.nmol, self.amax, 2), np.float64)
        self.n = empty((self.nmol,), np.int32)
        self.status = empty((max(self.nmol, 1),), np.int32)
        _native.call("fmlh_init")

    def set_coords(self, coords):
        """New coordinates for the same molecules (same atom order)."""
        self.r.copy_(to_device(np.asarray(coords).reshape(-1, 3), np.float64))

    def run(self):
        if self.nmol:
            describe_into(self.z, self.r, self.off, self.nmol, self.amax, self.cut, self.q,
                          self.zt, self.n, self.status)


class AradKernelCall:
    """One molecular ARAD kernel call on resident descriptor sets.

    q (nm, amax, 5, amax) f64, z (nm, amax, 2) PTP types, n (nm,) atom
    counts, as ``fget_kernels_arad``.  mode "global" (two sets), "symmetric"
    or "kernel" (set 1 only).  ``out`` holds the result in C order
    (nm2, nm1, nsig) -- i.e. the reference's Fortran (nsig, nm1, nm2);
    ``result()`` returns that Fortran-order view."""

    def __init__(self, q1, z1, n1, q2=None, z2=None, n2=None, sigmas=(1.0,), mode="global",
                 width=0.2, cut_distance=6.0, r_width=1.0, c_width=0.5, precision=0, tiling=None):
        self.mode = _MODES[mode]
        self.q1, self.z1, self.n1 = to_device(q1), to_device(z1), to_device(n1, np.int32)
        self.nm1, self.amax = int(self.q1.shape[0]), int(self.q1.shape[1])
        if self.mode == _native.ARAD_GLOBAL:
            if q2 is None:
                raise ValueError("the global mode needs set 2")
            self.q2, self.z2, self.n2 = to_device(q2), to_device(z2), to_device(n2, np.int32)
            self.nm2 = int(self.q2.shape[0])
        else:
            self.q2 = self.z2 = self.n2 = None
            self.nm2 = self.nm1
        sig = np.asarray(sigmas, dtype=np.float64).reshape(-1)
        if self.mode == _native.ARAD_KERNEL:
            sig = sig[:1]
            width, cut_distance, r_width, c_width = 0.2, 6.0, 1.0, 0.5
        if not 1 <= sig.size <= 8:
            raise ValueError("one prepared call takes 1..8 sigmas")
        self.nsig = int(sig.size)
        self.sig = (ctypes.c_double * self.nsig)(*sig.tolist())
        self.width, self.cut, self.rw, self.cw = float(width), float(cut_distance), float(r_width), float(c_width)
        self.precision = int(precision)
        self.tiling = tiling
        lib = _native.load()
        self.wsb = lib.fmlh_arad_ws_bytes(self.nm1, self.nm2, self.amax, self.mode)
        self.ws = workspace(self.wsb)
        self.out = (empty((self.nm1, self.nm1), np.float64) if self.mode == _native.ARAD_KERNEL
                    else empty((self.nm2, self.nm1, self.nsig), np.float64))

    def _args(self, nm1=None):
        p = _native.ptr
        return (self.mode, p(self.q1), p(self.z1), p(self.n1), self.nm1 if nm1 is None else nm1,
                p(self.q2), p(self.z2), p(self.n2), self.nm2, self.amax, self.sig, self.nsig,
                self.width, self.cut, self.rw, self.cw, self.precision, p(self.out), p(self.ws),
                self.wsb)

    def set_descriptors(self, q1=None, q2=None):
        """New descriptor values with the same atom counts."""
        if q1 is not None:
            self.q1.copy_(to_device(q1))
        if q2 is not None:
            if self.q2 is None:
                raise ValueError("this call has no set 2")
            self.q2.copy_(to_device(q2))

    def run(self):
        tp = ctypes.byref(self.tiling) if self.tiling is not None else None
        _native.call("fmlh_arad_kernels", *self._args(), tp, _native.stream_ptr())

    def autotune(self, sample_rows=None, reps=3):
        """Time the candidate tilings on the first `sample_rows` set-1
        molecules (all of them by default) against the whole set 2; the
        fastest becomes this call's tiling and the library's choice for
        (mode, amax, precision, nsig).  Returns (tiling, ms per call on the
        sample)."""
        rows = self.nm1 if sample_rows is None else max(1, min(int(sample_rows), self.nm1))
        if self.mode != _native.ARAD_GLOBAL:
            rows = self.nm1  # the triangular forms need the whole set
        best = _native.AradTiling()
        ms = ctypes.c_double(0.0)
        _native.call("fmlh_arad_autotune", *self._args(rows), int(reps), ctypes.byref(best),
                     ctypes.byref(ms), _native.stream_ptr())
        self.tiling = best
        tuning.record(self.mode, self.amax, self.precision, self.nsig, best)
        if rows != self.nm1:
            self.run()  # out then holds the full result again
        return best, ms.value

    def result(self):
        """Fortran-order view: (nsig, nm1, nm2), or (nm1, nm1) for "kernel"."""
        if self.mode == _native.ARAD_KERNEL:
            return self.out.t()
        return self.out.permute(2, 1, 0)


class KRRCall:
    """Plain FP64 kernel-ridge regression of a fixed molecule batch, fit and
    predict, on resident buffers: the reference's chain (README.md:33-69)
    sorted Coulomb matrices (frepresentations.f90:23-110) -> Gaussian kernel
    of the training rows (fkernels.f90:317-349) -> K + lambda I ->
    cho_solve (fcho_solve.f90:23-47) -> Gaussian kernel of the test rows
    against the training rows -> K_test^T alpha.

    The first ``ntrain`` molecules train, the rest are predicted.  Every
    launch reads and writes buffers the object owns (``alpha``, ``pred``,
    ``info`` are the results), so ``run()`` can be captured into a
    ``fml_amd.graphs.Graph`` and replayed; ``set_coords`` / ``set_targets``
    change the inputs between runs.  ``info`` holds the factorisation's
    LAPACK info after a run (0: K + lambda I positive definite)."""

    def __init__(self, charges, coords, offsets, ntrain, y, sigma, lam, nmax=23):
        self.cm = CoulombCall(charges, coords, offsets, nmax=nmax)
        nmol = self.cm.nmol
        if not 0 < int(ntrain) <= nmol:
            raise ValueError(f"ntrain must be in 1..{nmol}")
        self.ntrain = n = int(ntrain)
        self.ntest = nmol - n
        self.sigma = float(sigma)
        self.lam = float(lam)
        t = _native.torch_mod()
        self.y = to_device(np.asarray(y, dtype=np.float64)[:n], np.float64)
        if int(self.y.shape[0]) != n:
            raise ValueError("y needs one target per training molecule")
        self.K = empty((n, n), np.float64)
        self.Kt = empty((max(self.ntest, 1), n), np.float64)
        self.alpha = empty((n,), np.float64)
        self.pred = empty((max(self.ntest, 1),), np.float64)
        self.info = t.zeros(1, dtype=t.int32, device="cuda")
        self.ws_bytes = int(_native.load().fmlh_potrs_ws_bytes(n, 1))
        self.ws = workspace(self.ws_bytes)

    def set_coords(self, coords):
        self.cm.set_coords(coords)

    def set_targets(self, y):
        self.y.copy_(to_device(np.asarray(y, dtype=np.float64)[:self.ntrain], np.float64))

    def run(self):
        from .krr import add_diagonal, predict_from_kernel
        n, st = self.ntrain, _native.stream_ptr()
        X = self.cm.cm
        d = int(X.shape[1])
        self.cm.run()
        _native.call("fmlh_gaussian_kernel", _native.ptr(X), X.stride(0), n, _native.ptr(X),
                     X.stride(0), n, d, self.sigma, _native.ptr(self.K), self.K.stride(0), st)
        add_diagonal(self.K, self.lam)
        _native.call("fmlh_potrf", _native.ptr(self.K), self.K.stride(0), n, _native.ptr(self.info),
                     None, 0, st)
        # alpha = y by an elementwise kernel (a graph then holds no memory-copy node)
        _native.torch_mod().mul(self.y, 1.0, out=self.alpha)
        _native.call("fmlh_potrs", _native.ptr(self.K), self.K.stride(0), n, _native.ptr(self.alpha),
                     n, 1, _native.ptr(self.ws), self.ws_bytes, st)
        if self.ntest:
            Xt = X[n:]
            _native.call("fmlh_gaussian_kernel", _native.ptr(X), X.stride(0), n, _native.ptr(Xt),
                         Xt.stride(0), self.ntest, d, self.sigma, _native.ptr(self.Kt),
                         self.Kt.stride(0), st)
            predict_from_kernel(self.Kt[:self.ntest], self.alpha, out=self.pred)
