"""MI355X-native GLUED enhanced-sampling step: CV evaluation, metadynamics /
OPES bias evaluation and deposition, and the chain-rule force scatter, as
hand-written HIP kernels for gfx950 behind a C-ABI (include/glued.h)."""
from . import bias_state, glued
from .bias_state import BiasStateMerger
from .glued import Force, GluedForce, multithermal_temperature_ladder

__all__ = ["bias_state", "glued", "BiasStateMerger", "Force", "GluedForce", "multithermal_temperature_ladder",
           "Context"]


def __getattr__(name):
    # Context imports torch; keep `import gluedmd_amd` light for spec-only use.
    if name == "Context":
        from .context import Context
        return Context
    raise AttributeError(name)
