"""ctypes bridge to the oracle's C restatement (liboracle.so, built from sht_ref.c).

TEST INFRASTRUCTURE ONLY: used by tests/ as a second restatement (FP32 loop order of
SpeedyTransforms/src/legendre.jl) and by bench.py's cpu_baseline leg. Tables come from
the FP64 numpy oracle (sht_oracle.OracleTransform).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .sht_oracle import OracleTransform

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Geom(C.Structure):
    _fields_ = [
        ("nlat_half", C.c_int32), ("lmax", C.c_int32), ("mmax", C.c_int32),
        ("nharm", C.c_int32), ("npoints", C.c_int64),
        ("nlons", C.c_void_p), ("ring_start", C.c_void_p), ("mtrunc", C.c_void_p),
        ("offsets", C.c_void_p), ("coslat_inv", C.c_void_p), ("solid_angles", C.c_void_p),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} missing: run __graft_entry__.build() (make -C oracle)")
        _LIB = C.CDLL(path)
        for fn in ("oref_synthesis_f32", "oref_synthesis_f64"):
            getattr(_LIB, fn).argtypes = [C.POINTER(_Geom), C.c_void_p, C.c_void_p, C.c_int,
                                          C.c_int, C.c_void_p, C.c_int]
        for fn in ("oref_analysis_f32", "oref_analysis_f64"):
            getattr(_LIB, fn).argtypes = [C.POINTER(_Geom), C.c_void_p, C.c_void_p, C.c_int,
                                          C.c_void_p, C.c_int]
        _LIB.oref_fft_f64.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    return _LIB


class CRefTransform:
    """The C restatement bound to the tables of an OracleTransform."""

    def __init__(self, oracle: OracleTransform):
        self.o = oracle
        g, sp = oracle.grid, oracle.spectrum
        n = g.nlat_half
        self._keep = dict(
            nlons=np.ascontiguousarray(g.nlons_half, np.int32),
            ring_start=np.ascontiguousarray(g.ring_start, np.int64),
            mtrunc=np.ascontiguousarray(oracle.mtrunc, np.int32),
            offsets=np.ascontiguousarray(sp.offsets, np.int64),
            coslat_inv=np.ascontiguousarray(oracle.coslat_inv[:n], np.float64),
            solid_angles=np.ascontiguousarray(oracle.solid_angles[:n], np.float64),
        )
        k = self._keep
        self.geom = _Geom(n, sp.lmax, sp.mmax, sp.nharm, g.npoints,
                          k["nlons"].ctypes.data, k["ring_start"].ctypes.data,
                          k["mtrunc"].ctypes.data, k["offsets"].ctypes.data,
                          k["coslat_inv"].ctypes.data, k["solid_angles"].ctypes.data)
        lam = np.ascontiguousarray(oracle.legendre.T)  # [j][lm]
        self.lam = {np.float64: lam, np.float32: lam.astype(np.float32)}

    def synthesis(self, coeffs: np.ndarray, dtype=np.float32, unscale_coslat=False, nthreads=0):
        coeffs = np.atleast_2d(coeffs)
        K = coeffs.shape[0]
        cdt = np.complex64 if dtype == np.float32 else np.complex128
        a = np.ascontiguousarray(coeffs, cdt)
        out = np.zeros((K, self.o.grid.npoints), dtype)
        fn = lib().oref_synthesis_f32 if dtype == np.float32 else lib().oref_synthesis_f64
        rc = fn(C.byref(self.geom), self.lam[dtype].ctypes.data, a.ctypes.data, K,
                int(unscale_coslat), out.ctypes.data, nthreads)
        if rc:
            raise MemoryError("oracle synthesis failed")
        return out

    def analysis(self, field: np.ndarray, dtype=np.float32, nthreads=0):
        field = np.atleast_2d(field)
        K = field.shape[0]
        cdt = np.complex64 if dtype == np.float32 else np.complex128
        f = np.ascontiguousarray(field, dtype)
        out = np.zeros((K, self.o.spectrum.nharm), cdt)
        fn = lib().oref_analysis_f32 if dtype == np.float32 else lib().oref_analysis_f64
        rc = fn(C.byref(self.geom), self.lam[dtype].ctypes.data, f.ctypes.data, K,
                out.ctypes.data, nthreads)
        if rc:
            raise MemoryError("oracle analysis failed")
        return out


def fft64(x: np.ndarray, sign: int) -> np.ndarray:
    """The C restatement's complex FFT (sum x_n exp(sign 2 pi i n k / N))."""
    x = np.ascontiguousarray(x, np.complex128)
    out = np.zeros_like(x)
    lib().oref_fft_f64(x.size, sign, x.ctypes.data, out.ctypes.data)
    return out
