"""MI355X-native LBcuda time step: C-ABI HIP library + Python bindings."""
from .abi import LbmConfig, default_config  # noqa: F401
