"""scopyon_amd: an MI355X-native (gfx950 HIP) drop-in for scopyon's hot path.

Same public names as ``scopyon``: ``sample``, ``sample_inputs``,
``form_image``, ``generate_images``, ``create_simulator``,
``EPIFMSimulator``, ``Configuration``, ``DefaultConfiguration``, ``Image``,
``Video`` and ``constants``.  The compute runs in ``libscopyon_hip.so``
(built from ``scopyon_amd/csrc``); there is no CPU fallback.  Extras:
``PhiloxRNG`` (device random numbers) and ``BatchedSimulator`` (many fields
of view per launch).
"""
from .base import *  # noqa: F401,F403
from .config import *  # noqa: F401,F403
from .image import *  # noqa: F401,F403
from .sampling import *  # noqa: F401,F403
from .sampling2 import *  # noqa: F401,F403
from . import constants  # noqa: F401
from .rng import PhiloxRNG  # noqa: F401

__all__ = [
    "EnvironSettings", "EPIFMSimulator",
    "form_image", "generate_images", "create_simulator",
    "Configuration", "DefaultConfiguration",
    "Image", "Video",
    "sample_inputs",
    "sample",
    "constants",
    "PhiloxRNG",
]

__version__ = "1.0.0"
