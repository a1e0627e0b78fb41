"""MI355X-native fluid step of Sprout (APIC transfer, MGPCG pressure, surface tension,
explicit viscosity) behind the reference's FluidSim interface. See fluid_sim.FluidSim."""
from .fluid_sim import FluidSim  # noqa: F401
from . import scene  # noqa: F401
