"""ctypes wrapper of cpu_fp32.c, the FP32 multi-core CPU restatement of the
3D NCC GetValueAndDerivative (bench.py's cpu_baseline).

TEST INFRASTRUCTURE ONLY.  The library is compiled on the host that runs it
(-O3 -march=native -fopenmp), into a temporary directory, so that it uses
that host's vector units.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)


class Problem(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int * 3), ("fixed", c_float_p), ("coeffs", c_float_p),
                ("cp_size", ctypes.c_int * 3), ("cp_origin", ctypes.c_float * 3), ("cp_spacing", ctypes.c_float * 3),
                ("cp_aos", c_float_p), ("cpw", c_float_p), ("sigma", ctypes.c_float), ("tensor", ctypes.c_int),
                ("cov", c_float_p), ("max_eig_cov", ctypes.c_float * 9)]


_lib = None
COMPILER_FLAGS = ["-O3", "-march=native", "-fopenmp", "-fPIC", "-shared", "-std=c99"]


def load():
    global _lib
    if _lib is None:
        out = os.path.join(tempfile.mkdtemp(prefix="skm_cpu_fp32_"), "libskm_cpu_fp32.so")
        subprocess.run(["gcc", *COMPILER_FLAGS, "-o", out, os.path.join(HERE, "cpu_fp32.c"), "-lm"], check=True)
        lib = ctypes.CDLL(out)
        lib.skmf_set_threads.restype = ctypes.c_int
        lib.skmf_set_threads.argtypes = [ctypes.c_int]
        lib.skmf_value_and_derivative_ncc.restype = ctypes.c_int
        lib.skmf_value_and_derivative_ncc.argtypes = [ctypes.POINTER(Problem), ctypes.c_double, c_double_p,
                                                      c_double_p, c_float_p]
        lib.skmf_bilateral.restype = ctypes.c_int
        lib.skmf_bilateral.argtypes = [ctypes.POINTER(Problem), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       c_double_p, c_float_p]
        _lib = lib
    return _lib


def fp(a):
    return a.ctypes.data_as(c_float_p)


class CpuProblem:
    """fixed / moving (z, y, x) float arrays, spacing 1, origin 0; coeffs the
    padded (n+3)^3 B-spline coefficients of moving; cp grid (size, origin,
    spacing) in (x, y, z); cov per voxel (n, 9) or None."""

    def __init__(self, fixed, coeffs, cp_grid, cp_aos, cpw, sigma, cov=None, max_eig_cov=None):
        self.lib = load()
        self._keep = [np.ascontiguousarray(fixed, np.float32), np.ascontiguousarray(coeffs, np.float32),
                      np.ascontiguousarray(cp_aos, np.float32).reshape(-1), np.ascontiguousarray(cpw, np.float32)]
        p = Problem()
        for d, n in enumerate(reversed(fixed.shape)):
            p.size[d] = n
        p.fixed, p.coeffs, p.cp_aos, p.cpw = (fp(a) for a in self._keep)
        size, origin, spacing = cp_grid
        for d in range(3):
            p.cp_size[d] = int(size[d])
            p.cp_origin[d] = float(origin[d])
            p.cp_spacing[d] = float(spacing[d])
        p.sigma = float(sigma)
        if cov is not None:
            c = np.ascontiguousarray(cov, np.float32).reshape(-1)
            self._keep.append(c)
            p.tensor = 1
            p.cov = fp(c)
            m = np.asarray(max_eig_cov, np.float64).reshape(-1)
            for i in range(9):
                p.max_eig_cov[i] = float(m[i])
        self.p = p
        self.n_cp = int(np.prod(size))

    def evaluate(self, rd_weight=0.0, bilateral=None):
        """(NCC, RD value, derivative SoA) -- with bilateral = (weight, sigma_s,
        sigma_i) also the bilateral term: (NCC, RD value, bilateral value, derivative)."""
        der = np.zeros(self.n_cp * 3, np.float32)
        measure, rd = ctypes.c_double(), ctypes.c_double()
        if self.lib.skmf_value_and_derivative_ncc(ctypes.byref(self.p), rd_weight, ctypes.byref(measure),
                                                  ctypes.byref(rd), fp(der)) != 0:
            raise RuntimeError("cpu_fp32 evaluation failed")
        if bilateral is None:
            return measure.value, rd.value, der
        bv = ctypes.c_double()
        w, ss, si = bilateral
        if self.lib.skmf_bilateral(ctypes.byref(self.p), float(w), float(ss), float(si), ctypes.byref(bv), fp(der)) != 0:
            raise RuntimeError("cpu_fp32 bilateral failed")
        return measure.value, rd.value, bv.value, der
