"""Python mirror of the MI355X PD-Flow library.

The product is the C-ABI in ``libpdflow_hip.so`` (include/pdflow_hip.h); this package binds it
with ctypes so tests and the benchmark can drive it.  ``Solver`` mirrors the reference's
``CSF_cuda`` host API stage by stage (pdflow_cudalib.h:125-174) and the pair evaluator's driver
(scene_flow_impair.cpp:80-168).  Loading fails loudly when the HIP library has not been built:
there is no CPU fallback.
"""
from .lib import (  # noqa: F401
    FIELDS,
    Batch,
    PYR_KINDS,
    PDFlowError,
    PinnedBuffer,
    Solver,
    default_params,
    device_count,
    hip_library,
    lib_path,
)
from . import png  # noqa: F401
from .synth import synth_pair, synth_pair_float, synth_pairs_float, synth_static_float, synth_sequence_float, synth_hash  # noqa: F401
from .schedule import ctf_levels_for_rows, num_max_iter, gaussian_mask, level_table  # noqa: F401
from .net import Net, mfma_probe  # noqa: F401
