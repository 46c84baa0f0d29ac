"""ctypes mirror of include/lbcuda_amd.h (struct lbm_config and constants).

The reference has no FFI (kernels are launched with CUDA-Fortran chevrons from
main.CUF:127-241); this module describes the C ABI that replaces them.
"""
import ctypes

STORE_F32, STORE_F16, STORE_BF16 = 0, 1, 2
WET_NONE, WET_JANUS, WET_NEUTRAL = 0, 1, 2
KERNEL_FUSED, KERNEL_TWOPASS = 0, 1
INIT_DROPLET, INIT_ZPLANES, INIT_RHO, INIT_POPS, INIT_CYLINDER, INIT_RANDOM, INIT_VELFIELD = range(1, 8)

ERR_NAMES = {
    0: "LBM_OK", 1: "LBM_ERR_ARG", 2: "LBM_ERR_ALLOC", 3: "LBM_ERR_HIP",
    4: "LBM_ERR_RHO_RANGE", 5: "LBM_ERR_BIG_FORCE", 6: "LBM_ERR_EXTREME_VEL",
    7: "LBM_ERR_OUT_OF_BOX", 8: "LBM_ERR_PART_SETUP", 9: "LBM_ERR_STATE",
    10: "LBM_ERR_BATCH", 11: "LBM_ERR_LIST_OVERFLOW",
    12: "LBM_ERR_STORE_RANGE",
}


class LbmConfig(ctypes.Structure):
    """struct lbm_config of include/lbcuda_amd.h, field for field."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("tau_r", ctypes.c_float), ("tau_b", ctypes.c_float),
        ("sigma", ctypes.c_float), ("beta", ctypes.c_float),
        ("storage", ctypes.c_int32), ("store_rho0", ctypes.c_float),
        ("min_rho", ctypes.c_float), ("max_rho", ctypes.c_float),
        ("with_particles", ctypes.c_int32), ("n_particles", ctypes.c_int32),
        ("lrotate", ctypes.c_int32), ("radius", ctypes.c_float),
        ("rmass", ctypes.c_float), ("max_vel", ctypes.c_float),
        ("ext_force", ctypes.c_float * 3), ("ext_torque", ctypes.c_float * 3),
        ("wetting", ctypes.c_int32), ("hertz", ctypes.c_int32),
        ("hz_k", ctypes.c_float), ("hz_rmin", ctypes.c_float), ("hz_rcap", ctypes.c_float),
        ("lubrication", ctypes.c_int32),
        ("lub_k", ctypes.c_float), ("lub_rmin", ctypes.c_float), ("lub_rcap", ctypes.c_float),
        ("closure", ctypes.c_int32),
        ("diffdens", ctypes.c_int32), ("dens_r", ctypes.c_float), ("dens_b", ctypes.c_float),
        ("reserved", ctypes.c_int32 * 5),
    ]


def default_config(n, **kw):
    """Reference defaults (dimensions_m.CUF:50-84, 103-108; kernels_fluid.CUF:35)
    for an n^3 box, with keyword overrides.  Box edges may be overridden with
    nx/ny/nz."""
    c = LbmConfig()
    c.nx = c.ny = c.nz = n
    c.tau_r = c.tau_b = 1.0
    c.sigma, c.beta = 0.1, 0.999
    c.storage, c.store_rho0 = STORE_F32, 0.5
    c.min_rho, c.max_rho = -10.0, 10.0
    c.with_particles, c.n_particles, c.lrotate = 0, 0, 0
    c.radius, c.rmass, c.max_vel = 5.5, 1.0 / 1500.0, 0.0
    c.wetting = WET_JANUS
    c.hertz, c.hz_k, c.hz_rmin, c.hz_rcap = 1, 10.0, 10.5, 9.0
    c.lubrication, c.lub_k, c.lub_rmin, c.lub_rcap = 0, 10.0, 10.0, 9.0
    c.closure = 0
    c.diffdens, c.dens_r, c.dens_b = 0, 1.0, 1.0
    for k, v in kw.items():
        if k in ("ext_force", "ext_torque"):
            arr = getattr(c, k)
            for i in range(3):
                arr[i] = v[i]
        else:
            if not hasattr(c, k):
                raise AttributeError(f"lbm_config has no field {k}")
            setattr(c, k, v)
    return c
