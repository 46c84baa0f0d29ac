"""ctypes wrapper of oracle/liblbm_oracle.so -- TEST INFRASTRUCTURE ONLY.

Mirrors lbcuda_amd.lib.Box so tests drive the oracle and the HIP library with
the same calls.  Arrays come back Fortran-ordered as numpy (order='F') views:
pops (nx,ny,nz,19), rho (nx,ny,nz), pos (13,n), force (12,n).
"""
import ctypes
import os
import subprocess

import numpy as np

from lbcuda_amd.abi import LbmConfig, ERR_NAMES

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblbm_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.orc_create.argtypes = [P(LbmConfig), ctypes.c_int, P(vp)]
        L.orc_destroy.argtypes = [vp]
        L.orc_init_fields.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.orc_set_pops.argtypes = [vp, vp, vp]
        L.orc_set_particles.argtypes = [vp, ctypes.c_int, vp]
        L.orc_step.argtypes = [vp, ctypes.c_int]
        L.orc_get_pops.argtypes = [vp, ctypes.c_int, vp]
        L.orc_get_rho.argtypes = [vp, vp, vp]
        L.orc_get_flags.argtypes = [vp, vp]
        L.orc_get_particles.argtypes = [vp, vp, vp]
        L.orc_status.argtypes = [vp, P(ctypes.c_int), P(ctypes.c_long), ctypes.c_char_p, ctypes.c_int]
        L.orc_template.argtypes = [vp, P(ctypes.c_int), vp, ctypes.c_int]
        _LIB = L
    return _LIB


def _f32(a):
    return None if a is None else np.asfortranarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleBox:
    """One box of the CPU oracle; bits=64 (parity oracle) or 32 (CPU path)."""

    def __init__(self, cfg, bits=64):
        self.cfg = cfg
        self.n = (cfg.nx, cfg.ny, cfg.nz)
        self.h = ctypes.c_void_p()
        rc = lib().orc_create(ctypes.byref(cfg), bits, ctypes.byref(self.h))
        if rc:
            raise RuntimeError(f"orc_create: {ERR_NAMES.get(rc, rc)}")
        self.npart = 0

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what}: {ERR_NAMES.get(rc, rc)}")

    def init_fields(self, init_type, u0=(0.0, 0.0, 0.0), rho_r=None, rho_b=None, vel=None):
        u = np.asarray(u0, dtype=np.float32)
        rr, rb, v = _f32(rho_r), _f32(rho_b), _f32(vel)
        self._check(lib().orc_init_fields(self.h, init_type, _ptr(u), _ptr(rr), _ptr(rb), _ptr(v)),
                    "orc_init_fields")

    def set_pops(self, pops_r, pops_b):
        a, b = _f32(pops_r), _f32(pops_b)
        self._check(lib().orc_set_pops(self.h, _ptr(a), _ptr(b)), "orc_set_pops")

    def set_particles(self, pos13):
        p = np.asfortranarray(pos13, dtype=np.float64)
        self.npart = p.shape[1]
        self._check(lib().orc_set_particles(self.h, self.npart, _ptr(p)), "orc_set_particles")

    def step(self, n=1):
        """Returns the oracle's error code (0 if none) after n steps."""
        return lib().orc_step(self.h, n)

    def pops(self, color):
        out = np.zeros(self.n + (19,), dtype=np.float64, order="F")
        self._check(lib().orc_get_pops(self.h, color, _ptr(out)), "orc_get_pops")
        return out

    def rho(self):
        a = np.zeros(self.n, dtype=np.float64, order="F")
        b = np.zeros(self.n, dtype=np.float64, order="F")
        self._check(lib().orc_get_rho(self.h, _ptr(a), _ptr(b)), "orc_get_rho")
        return a, b

    def flags(self):
        out = np.zeros(self.n, dtype=np.int8, order="F")
        self._check(lib().orc_get_flags(self.h, _ptr(out)), "orc_get_flags")
        return out

    def particles(self):
        pos = np.zeros((13, self.npart), dtype=np.float64, order="F")
        frc = np.zeros((12, self.npart), dtype=np.float64, order="F")
        self._check(lib().orc_get_particles(self.h, _ptr(pos), _ptr(frc)), "orc_get_particles")
        return pos, frc

    def status(self):
        code = ctypes.c_int()
        step = ctypes.c_long()
        buf = ctypes.create_string_buffer(256)
        lib().orc_status(self.h, ctypes.byref(code), ctypes.byref(step), buf, 256)
        return code.value, step.value, buf.value.decode()

    def template(self):
        r = ctypes.c_int()
        lib().orc_template(self.h, ctypes.byref(r), None, 0)
        w = 2 * r.value + 1
        out = np.zeros((w, w, w), dtype=np.int8, order="F")
        lib().orc_template(self.h, ctypes.byref(r), _ptr(out), out.size)
        return r.value, out
