"""pyqg_amd: an MI355X-native (HIP / gfx950, FP64) time-stepping core for pyqg.

`PseudoSpectralKernel` is a drop-in for pyqg's kernel class; the model
front-ends (QGModel, LayeredModel, BTModel, SQGModel) sit on top of it, and
`Ensemble` steps many independent models in one fused launch sequence.
"""
from .kernel import (PseudoSpectralKernel, tendency_forward_euler, tendency_ab2,
                     tendency_ab3)
from .model import Model, QGModel, LayeredModel, BTModel, SQGModel
from .ensemble import Ensemble
from .parameterizations import (Parameterization, CompositeParameterization,
                                WeightedParameterization, UVParameterization, QParameterization,
                                Smagorinsky, BackscatterBiharmonic, ZannaBolton2020, RingForcing)
from . import diagnostic_tools

__all__ = ["PseudoSpectralKernel", "tendency_forward_euler", "tendency_ab2", "tendency_ab3",
           "Model", "QGModel", "LayeredModel", "BTModel", "SQGModel", "Ensemble",
           "Parameterization", "CompositeParameterization", "WeightedParameterization",
           "UVParameterization", "QParameterization", "Smagorinsky", "BackscatterBiharmonic",
           "ZannaBolton2020", "RingForcing", "diagnostic_tools"]
