"""ctypes binding of liblbcuda_amd.so, the C ABI of include/lbcuda_amd.h.

Box mirrors the reference host program's view of one simulation box: create
(setupGPUgridAndAlloc + SetConstants, main.CUF:2133-2407, 1043-1162), initial
fields (setupRho_Pops, main.CUF:2531-2608), particles (read_input_atom +
SetupParts, main.CUF:1187-1300) and the time step (main.CUF:127-278).  Arrays
are numpy arrays in Fortran order: pops (nx,ny,nz,19), rho (nx,ny,nz),
vel (3,nx,ny,nz), pos (13,n), force (12,n).

There is no CPU fallback: if the shared library is missing or cannot be
loaded, every call raises.
"""
import ctypes
import os

import numpy as np

from .abi import LbmConfig, ERR_NAMES

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "liblbcuda_amd.so")
_LIB = None


class LbmError(RuntimeError):
    def __init__(self, code, what, msg=""):
        self.code = code
        super().__init__(f"{what}: {ERR_NAMES.get(code, code)} {msg}".strip())


def lib():
    """Load liblbcuda_amd.so (built by lbcuda_amd.build / __graft_entry__.build)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    i32 = ctypes.c_int32
    L.lbm_config_default.argtypes = [P(LbmConfig), i32]
    L.lbm_config_default.restype = None
    L.lbm_create.argtypes = [P(LbmConfig), P(vp)]
    L.lbm_destroy.argtypes = [vp]
    L.lbm_init_fields.argtypes = [vp, i32, vp, vp, vp, vp]
    L.lbm_set_pops.argtypes = [vp, vp, vp]
    L.lbm_set_particles.argtypes = [vp, i32, vp]
    L.lbm_step.argtypes = [vp, i32]
    L.lbm_get_pops.argtypes = [vp, i32, vp]
    L.lbm_get_rho.argtypes = [vp, vp, vp]
    L.lbm_get_moments.argtypes = [vp, vp, vp, vp, vp]
    L.lbm_get_flags.argtypes = [vp, vp]
    L.lbm_get_particles.argtypes = [vp, vp, vp]
    L.lbm_get_sums.argtypes = [vp, vp]
    L.lbm_checkpoint_size.argtypes = [vp, P(ctypes.c_int64)]
    L.lbm_checkpoint_save.argtypes = [vp, vp, ctypes.c_int64]
    L.lbm_checkpoint_load.argtypes = [vp, vp, ctypes.c_int64]
    L.lbm_status.argtypes = [vp, P(i32), P(i32), ctypes.c_char_p, i32]
    L.lbm_kernel_timing.argtypes = [vp, i32]
    L.lbm_get_kernel_timing.argtypes = [vp, P(ctypes.c_double), P(ctypes.c_int64)]
    L.lbm_kernel_info.argtypes = [vp, ctypes.c_char_p, i32, P(ctypes.c_double)]
    L.lbm_set_kernel.argtypes = [vp, i32, i32, i32]
    L.lbm_get_kernel.argtypes = [vp, P(i32), P(i32), P(i32)]
    L.lbm_num_tiles.argtypes = []
    L.lbm_num_tiles.restype = i32
    L.lbm_tile_shape.argtypes = [i32, P(i32), P(i32), P(i32)]
    L.lbm_autotune.argtypes = [vp, i32, P(i32), P(i32), P(ctypes.c_double)]
    L.lbm_copy_bandwidth.argtypes = [ctypes.c_int64, i32, P(ctypes.c_double)]
    L.lbm_stream_bandwidth.argtypes = [ctypes.c_int64, i32, P(ctypes.c_double)]
    L.lbm_mlp_forward.argtypes = [vp, vp, vp, i32]
    L.lbm_mlp_loss_grad.argtypes = [vp, vp, vp, i32, P(ctypes.c_double), vp]
    L.lbm_mlp_train.argtypes = [vp, vp, vp, i32, i32, i32, ctypes.c_float, ctypes.c_uint64, P(ctypes.c_double)]
    L.lbm_set_closure.argtypes = [vp, vp]
    L.lbm_ensemble_create.argtypes = [i32, P(LbmConfig), P(vp)]
    L.lbm_ensemble_destroy.argtypes = [vp]
    L.lbm_ensemble_box.argtypes = [vp, i32, P(vp)]
    L.lbm_ensemble_step.argtypes = [vp, i32]
    L.lbm_ensemble_set_tile.argtypes = [vp, i32, i32]
    L.lbm_ensemble_autotune.argtypes = [vp, i32, P(i32), P(i32), P(ctypes.c_double)]
    L.lbm_version.restype = ctypes.c_char_p
    L.lbm_fill_random_rho.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_double, i32, i32, i32, vp, vp]
    L.lbm_fill_random_rho.restype = ctypes.c_int
    for name in ("lbm_create", "lbm_destroy", "lbm_init_fields", "lbm_set_pops", "lbm_set_particles",
                 "lbm_step", "lbm_get_pops", "lbm_get_rho", "lbm_get_moments", "lbm_get_flags",
                 "lbm_get_particles", "lbm_get_sums", "lbm_status", "lbm_kernel_timing",
                 "lbm_checkpoint_size", "lbm_checkpoint_save", "lbm_checkpoint_load",
                 "lbm_get_kernel_timing", "lbm_kernel_info", "lbm_set_kernel", "lbm_get_kernel",
                 "lbm_tile_shape", "lbm_autotune", "lbm_copy_bandwidth", "lbm_stream_bandwidth",
                 "lbm_ensemble_create", "lbm_ensemble_destroy", "lbm_ensemble_box", "lbm_ensemble_step", "lbm_ensemble_set_tile",
                 "lbm_ensemble_autotune", "lbm_mlp_forward", "lbm_mlp_loss_grad", "lbm_mlp_train",
                 "lbm_set_closure"):
        getattr(L, name).restype = ctypes.c_int
    _LIB = L
    return L


EXPORTED = ["lbm_config_default", "lbm_create", "lbm_destroy", "lbm_init_fields", "lbm_set_pops",
            "lbm_set_particles", "lbm_step", "lbm_get_pops", "lbm_get_rho", "lbm_get_moments",
            "lbm_get_flags", "lbm_get_particles", "lbm_get_sums", "lbm_status", "lbm_kernel_timing",
            "lbm_get_kernel_timing", "lbm_kernel_info", "lbm_version", "lbm_fill_random_rho",
            "lbm_set_kernel", "lbm_get_kernel", "lbm_num_tiles", "lbm_tile_shape", "lbm_autotune",
            "lbm_copy_bandwidth", "lbm_stream_bandwidth", "lbm_ensemble_create", "lbm_ensemble_destroy",
            "lbm_ensemble_box",
            "lbm_ensemble_step", "lbm_ensemble_set_tile", "lbm_ensemble_autotune", "lbm_mlp_forward",
            "lbm_mlp_loss_grad", "lbm_mlp_train", "lbm_set_closure", "lbm_checkpoint_size",
            "lbm_checkpoint_save", "lbm_checkpoint_load"]


def _f32(a):
    return None if a is None else np.asfortranarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Box:
    """One periodic box on the GPU."""

    def __init__(self, cfg, handle=None):
        self.cfg = cfg
        self.n = (cfg.nx, cfg.ny, cfg.nz)
        self.h = ctypes.c_void_p()
        self.owned = handle is None
        if handle is None:
            self._check(lib().lbm_create(ctypes.byref(cfg), ctypes.byref(self.h)), "lbm_create")
        else:
            self.h = handle
        self.npart = 0

    def close(self):
        if self.h and self.owned:
            lib().lbm_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise LbmError(rc, what)

    def init_fields(self, init_type, u0=(0.0, 0.0, 0.0), rho_r=None, rho_b=None, vel=None):
        u = np.asarray(u0, dtype=np.float32)
        rr, rb, v = _f32(rho_r), _f32(rho_b), _f32(vel)
        self._check(lib().lbm_init_fields(self.h, init_type, _ptr(u), _ptr(rr), _ptr(rb), _ptr(v)),
                    "lbm_init_fields")

    def set_pops(self, pops_r, pops_b):
        a, b = _f32(pops_r), _f32(pops_b)
        self._check(lib().lbm_set_pops(self.h, _ptr(a), _ptr(b)), "lbm_set_pops")

    def set_particles(self, pos13):
        p = np.asfortranarray(pos13, dtype=np.float64)
        self.npart = p.shape[1]
        self._check(lib().lbm_set_particles(self.h, self.npart, _ptr(p)), "lbm_set_particles")

    def step(self, n=1):
        """Advance n steps; returns the device error code (0 if none)."""
        rc = lib().lbm_step(self.h, n)
        if rc in (1, 2, 3, 9):
            raise LbmError(rc, "lbm_step")
        return rc

    def pops(self, color):
        out = np.zeros(self.n + (19,), dtype=np.float32, order="F")
        self._check(lib().lbm_get_pops(self.h, color, _ptr(out)), "lbm_get_pops")
        return out

    def rho(self):
        a = np.zeros(self.n, dtype=np.float32, order="F")
        b = np.zeros(self.n, dtype=np.float32, order="F")
        self._check(lib().lbm_get_rho(self.h, _ptr(a), _ptr(b)), "lbm_get_rho")
        return a, b

    def moments(self):
        rr = np.zeros(self.n, dtype=np.float32, order="F")
        rb = np.zeros(self.n, dtype=np.float32, order="F")
        vel = np.zeros((3,) + self.n, dtype=np.float32, order="F")
        ph = np.zeros(self.n, dtype=np.float32, order="F")
        self._check(lib().lbm_get_moments(self.h, _ptr(rr), _ptr(rb), _ptr(vel), _ptr(ph)), "lbm_get_moments")
        return rr, rb, vel, ph

    def flags(self):
        out = np.zeros(self.n, dtype=np.int8, order="F")
        self._check(lib().lbm_get_flags(self.h, _ptr(out)), "lbm_get_flags")
        return out

    def particles(self):
        pos = np.zeros((13, self.npart), dtype=np.float64, order="F")
        frc = np.zeros((12, self.npart), dtype=np.float64, order="F")
        self._check(lib().lbm_get_particles(self.h, _ptr(pos), _ptr(frc)), "lbm_get_particles")
        return pos, frc

    def sums(self):
        out = np.zeros(5, dtype=np.float64)
        self._check(lib().lbm_get_sums(self.h, _ptr(out)), "lbm_get_sums")
        return out

    def checkpoint(self):
        """Whole box state as bytes (lbm_checkpoint_save)."""
        n = ctypes.c_int64()
        self._check(lib().lbm_checkpoint_size(self.h, ctypes.byref(n)), "lbm_checkpoint_size")
        buf = np.empty(n.value, dtype=np.uint8)
        self._check(lib().lbm_checkpoint_save(self.h, _ptr(buf), n.value), "lbm_checkpoint_save")
        return buf

    def restore(self, image):
        """Load an image of checkpoint() into this box (lbm_checkpoint_load)."""
        buf = np.ascontiguousarray(image, dtype=np.uint8)
        self._check(lib().lbm_checkpoint_load(self.h, _ptr(buf), buf.size), "lbm_checkpoint_load")

    def status(self):
        code = ctypes.c_int32()
        step = ctypes.c_int32()
        buf = ctypes.create_string_buffer(256)
        self._check(lib().lbm_status(self.h, ctypes.byref(code), ctypes.byref(step), buf, 256), "lbm_status")
        return code.value, step.value, buf.value.decode()

    def kernel_timing(self, on=True):
        self._check(lib().lbm_kernel_timing(self.h, 1 if on else 0), "lbm_kernel_timing")

    def get_kernel_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(lib().lbm_get_kernel_timing(self.h, ctypes.byref(ms), ctypes.byref(n)), "lbm_get_kernel_timing")
        return ms.value, n.value

    def set_kernel(self, variant, tile=0, zchunk=0):
        """variant KERNEL_FUSED (tile index, z chunk; 0 = default chunk) or
        KERNEL_TWOPASS."""
        self._check(lib().lbm_set_kernel(self.h, variant, tile, zchunk), "lbm_set_kernel")

    def get_kernel(self):
        v, t, z = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._check(lib().lbm_get_kernel(self.h, ctypes.byref(v), ctypes.byref(t), ctypes.byref(z)),
                    "lbm_get_kernel")
        return v.value, t.value, z.value

    def autotune(self, reps=5):
        """Time every valid (tile, z chunk) of the fused kernel and keep the
        fastest; returns (tile, zchunk, ms per launch)."""
        t, z, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        self._check(lib().lbm_autotune(self.h, reps, ctypes.byref(t), ctypes.byref(z), ctypes.byref(ms)),
                    "lbm_autotune")
        return t.value, z.value, ms.value

    def set_closure(self, block):
        """Enable the learned near-contact closure (block of
        MLP_CLOSURE_NFLOAT floats, see lbcuda_amd.closure) or disable it
        (None)."""
        if block is None:
            self._check(lib().lbm_set_closure(self.h, None), "lbm_set_closure")
            return
        b = np.ascontiguousarray(block, dtype=np.float32)
        self._check(lib().lbm_set_closure(self.h, _ptr(b)), "lbm_set_closure")

    def kernel_info(self):
        buf = ctypes.create_string_buffer(128)
        b = ctypes.c_double()
        self._check(lib().lbm_kernel_info(self.h, buf, 128, ctypes.byref(b)), "lbm_kernel_info")
        return buf.value.decode(), b.value


class Ensemble:
    """Independent boxes advanced together (lbm_ensemble_*): the fluid of all
    boxes in one persistent launch per step, each particle kernel once for
    the particles of all boxes.  boxes[i] is a Box view of box i (initialise
    and read it through the Box calls; it is freed with the ensemble)."""

    def __init__(self, cfgs):
        arr = (LbmConfig * len(cfgs))(*cfgs)
        self.h = ctypes.c_void_p()
        rc = lib().lbm_ensemble_create(len(cfgs), arr, ctypes.byref(self.h))
        if rc:
            raise LbmError(rc, "lbm_ensemble_create")
        self.boxes = []
        for i, c in enumerate(cfgs):
            bh = ctypes.c_void_p()
            rc = lib().lbm_ensemble_box(self.h, i, ctypes.byref(bh))
            if rc:
                raise LbmError(rc, "lbm_ensemble_box")
            self.boxes.append(Box(c, handle=bh))

    def step(self, n=1):
        rc = lib().lbm_ensemble_step(self.h, n)
        if rc in (1, 2, 3, 9, 10):
            raise LbmError(rc, "lbm_ensemble_step")
        return rc

    def set_tile(self, tile, zchunk):
        rc = lib().lbm_ensemble_set_tile(self.h, tile, zchunk)
        if rc:
            raise LbmError(rc, "lbm_ensemble_set_tile")

    def autotune(self, reps=3):
        t, z, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        rc = lib().lbm_ensemble_autotune(self.h, reps, ctypes.byref(t), ctypes.byref(z), ctypes.byref(ms))
        if rc:
            raise LbmError(rc, "lbm_ensemble_autotune")
        return t.value, z.value, ms.value

    def close(self):
        if self.h:
            for b in self.boxes:
                b.h = ctypes.c_void_p()
            lib().lbm_ensemble_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fill_random_rho(shape, seed, mean=0.5, amp=0.05):
    """Host-side seeded colour field of the library (no GPU work)."""
    nx, ny, nz = shape
    rr = np.zeros(shape, np.float32, order="F")
    rb = np.zeros(shape, np.float32, order="F")
    rc = lib().lbm_fill_random_rho(seed, mean, amp, nx, ny, nz, _ptr(rr), _ptr(rb))
    if rc:
        raise LbmError(rc, "lbm_fill_random_rho")
    return rr, rb


def tile_shapes():
    """The fused kernel's tile table: list of (tx, ty, wrap)."""
    L = lib()
    out = []
    for t in range(L.lbm_num_tiles()):
        a, b, w = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        if L.lbm_tile_shape(t, ctypes.byref(a), ctypes.byref(b), ctypes.byref(w)):
            raise LbmError(1, "lbm_tile_shape")
        out.append((a.value, b.value, w.value))
    return out


def copy_bandwidth(nbytes=2 << 30, reps=10):
    """Measured device copy bandwidth (GB/s, read + write) of a plain copy
    kernel over two buffers of nbytes."""
    g = ctypes.c_double()
    rc = lib().lbm_copy_bandwidth(nbytes, reps, ctypes.byref(g))
    if rc:
        raise LbmError(rc, "lbm_copy_bandwidth")
    return g.value


def stream_bandwidth(nodes=1 << 24, reps=10):
    """Measured bandwidth (GB/s, read + write) of the fused step's access
    pattern without its arithmetic (19 planes read, 19 written, x-shifted)."""
    g = ctypes.c_double()
    rc = lib().lbm_stream_bandwidth(nodes, reps, ctypes.byref(g))
    if rc:
        raise LbmError(rc, "lbm_stream_bandwidth")
    return g.value


MLP_NPARAM = 4675
MLP_CLOSURE_NFLOAT = MLP_NPARAM + 20


def mlp_forward(params, x):
    """GPU batch inference of the closure network: x (n,4) -> y (n,3)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros((xx.shape[0], 3), dtype=np.float32)
    rc = lib().lbm_mlp_forward(_ptr(p), _ptr(xx), _ptr(y), xx.shape[0])
    if rc:
        raise LbmError(rc, "lbm_mlp_forward")
    return y


def mlp_loss_grad(params, x, t):
    """Mean squared error and its gradient, computed by the GPU kernels."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    tt = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(MLP_NPARAM, dtype=np.float32)
    loss = ctypes.c_double()
    rc = lib().lbm_mlp_loss_grad(_ptr(p), _ptr(xx), _ptr(tt), xx.shape[0], ctypes.byref(loss), _ptr(g))
    if rc:
        raise LbmError(rc, "lbm_mlp_loss_grad")
    return loss.value, g


def mlp_train(params, x, t, epochs, batch=256, lr=1e-3, seed=0):
    """Adam training on the GPU; returns (trained params, final loss)."""
    p = np.array(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    tt = np.ascontiguousarray(t, dtype=np.float32)
    loss = ctypes.c_double()
    rc = lib().lbm_mlp_train(_ptr(p), _ptr(xx), _ptr(tt), xx.shape[0], epochs, batch, lr, seed, ctypes.byref(loss))
    if rc:
        raise LbmError(rc, "lbm_mlp_train")
    return p, loss.value


def version():
    return lib().lbm_version().decode()
