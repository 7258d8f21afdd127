"""sph_cuda_amd - MI355X-native anisotropic-SPH atmospheric step (gfx950 HIP kernels behind a
C ABI, include/asph.h). The compute path is libasph.so; this package binds it with ctypes."""
from . import _lib
from ._lib import AsphError, Config, ICSpec, RunStats
from .sim import (Sim, State, check_config, concat_states, ensemble_preset, ic_generate, preset,
                  reference_config)

__all__ = ["AsphError", "Config", "ICSpec", "RunStats", "Sim", "State", "check_config", "concat_states",
           "ensemble_preset",
           "ic_generate", "preset", "reference_config", "_lib"]
