"""petstrip: MI355X-native list-mode MLEM for the 2D strip PET scanner.

The compute path is libpetstrip_hip.so (hand-written HIP for gfx950 behind
the C ABI in include/petstrip.h); this package is its Python host side.
"""
from .petstrip import (  # noqa: F401
    ACCUM_COMPENSATED, ACCUM_FIXED64, ACCUM_FP32, KERNEL_ANALYTIC, KERNEL_GAUSSIAN,
    PATH_AUTO, PATH_PER_EVENT, UPDATE_CPU,
    UPDATE_SENSITIVITY_DENOMINATOR, PetstripError, Reconstruction, Scanner,
    device_count, load_library, phantom_events, phantom_generate, phantom_grid,
    phantom_tan_kernel_map, parse_text_floats, read_text_responses, REGION_ELLIPSE, REGION_POINT, REGION_RECTANGLE,
    PROBE_GLOBAL_F32, PROBE_GLOBAL_F64, PROBE_GLOBAL_U64, PROBE_LDS_U64, PROBE_LDS_U32,
    PROBE_LDS_F32, probe_atomics,
)
from .workloads import (WORKLOADS, workload_scanner, workload_events,  # noqa: F401
                        HYBRID_WORKLOADS, hybrid_workload, expected_image)
from . import hybrid_geometry  # noqa: F401
from .hybrid import HybridReconstruction  # noqa: F401
from .calibration import calibrate_sigma  # noqa: F401
