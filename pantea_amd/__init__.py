"""pantea_amd -- MI355X (gfx950) kernels for the Behler-Parrinello NNP hot path.

Drop-in host surface of pantea's NNP path (ACSF, NeuralNetworkPotential,
MDSimulator/System/simulate) over hand-written HIP kernels behind the C ABI
of include/pantea_hip.h.  The kernels run on the GPU only; there is no CPU
fallback.
"""

from .atoms import Box, ElementMap, Structure
from .descriptors import ACSF, G1, G2, G3, G9, CutoffFunction, DescriptorScaler, NeighborElements
from .models import NeuralNetworkModel, UniformInitializer
from .potential import NNP, AtomicPotential, NeuralNetworkPotential, NNPSettings
from .simulation import BrendsenThermostat, MDSimulator, System, simulate
from .units import units

__all__ = [
    "ACSF", "G1", "G2", "G3", "G9", "CutoffFunction", "DescriptorScaler", "NeighborElements",
    "Box", "ElementMap", "Structure", "NeuralNetworkModel", "UniformInitializer",
    "NNP", "AtomicPotential", "NeuralNetworkPotential", "NNPSettings",
    "BrendsenThermostat", "MDSimulator", "System", "simulate", "units",
]
