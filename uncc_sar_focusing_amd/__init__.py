"""MI355X (gfx950) SAR focusing engine: range compression, backprojection and
autofocus as hand-written HIP kernels behind a C ABI (include/sarf.h).

The C++ drop-in surface of the reference (SAR_Aperture,
SAR_ImageFormationParameters, focus_SAR_image / cuda_focus_SAR_image,
charlotte_sar_api callbacks) is in include/*.hpp; this package is the Python
host interface over the same library.
"""
from .scenes import (  # noqa: F401
    CONFIGS,
    Geometry,
    ImageGrid,
    SceneConfig,
    Targets,
    circular_aperture,
    point_targets,
)

__all__ = [
    "CONFIGS",
    "Geometry",
    "ImageGrid",
    "SceneConfig",
    "Targets",
    "circular_aperture",
    "point_targets",
]


def context(device: int = 0):
    """Create a GPU context (requires the built libsarf.so)."""
    from .api import Context

    return Context(device)
