"""fastatari -- MI355X batched Atari 2600 frame step.

`BatchAtariGpu` mirrors the reference's Rust API (src/cuda_env.rs:99-191):
new / set_sorted / reset / step / step_multi / download_states / num_envs,
over the C ABI in include/fastatari.h (libfastatari_hip.so, HIP kernels for
gfx950).  There is no CPU fallback: constructing a BatchAtariGpu without the
built library or without a GPU raises.

`BatchMixGpu` steps several ROM classes in one hipGraph launch (per-class
launch shapes from its autotuner); `EnvServer` serves step requests of many
clients with continuous batching; `BatchAtariGpuEnv` is the env adapter over
BatchAtariGpu (src/rl4burn_env.rs:268-380); `BatchAtariRender` steps the
rendering console (src/atari.rs:226-378) and returns 160 x 192 palette-index
frames drawn by the GPU's per-clock TIA.
"""
import os as _os

# The HIP runtime maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by
# default), and kernels of streams that share a queue run one after another.
# The server and the batch keep one stream per ROM class plus one more, so a
# process that has not initialised HIP yet gets 8 queues; an explicit setting
# is left as it is.
_os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

from .env import (STATE_DTYPE, BatchAtariGpu, Busy, FastAtariError, DeviceBuffer, lib_path, load_library)
from .batch import BatchMixGpu
from .serve import EnvServer
from .vecenv import BatchAtariGpuEnv, Step
from .render import BatchAtariRender

__all__ = ["BatchAtariGpu", "BatchMixGpu", "EnvServer", "BatchAtariGpuEnv", "Step", "BatchAtariRender", "Busy",
           "FastAtariError", "DeviceBuffer", "STATE_DTYPE", "lib_path", "load_library"]
