"""MI355X-native FE² hot path of concrete_impact_ai.

Host classes mirror the reference's material-point, RVE and dynamics
interfaces; the arithmetic runs in the HIP library ``libcia_hip.so``
(gfx950) bound through ctypes (``concrete_impact_ai_amd._lib``).
"""

__all__ = ["materials", "rve", "solvers"]
