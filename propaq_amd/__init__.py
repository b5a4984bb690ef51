"""MI355X-native engine for propaq's per-gate Pauli / Majorana term propagation.

The public names mirror ``propaq``'s numerical API so existing code switches by
changing the import.  The numerics run in ``libpropaq_amd.so`` (HIP kernels for
gfx950 behind the C ABI in ``include/propaq_amd.h``).
"""

from .circuits import (AbstractCircuit, AbstractRotation, MajoranaCircuit, MajoranaRotation,  # noqa: F401
                       PauliCircuit, PauliRotation)
from .datatypes import (MajoranaMonomial, MajoranaTermStreamer, MajoranaTermSum, PauliString,  # noqa: F401
                        PauliTermStreamer, PauliTermSum)
from .noise import GateNoiseModel, NativeNoiseModel, NoiseModel, UniformNoiseModel  # noqa: F401
from .plugins import NativeTruncator  # noqa: F401
from .adjoint import GradientRun, expectation, run_adjoint  # noqa: F401
from .events import LogParser, Logger  # noqa: F401
from .propagators import MajoranaPropagator, PauliPropagator, PropagationResult  # noqa: F401
from .surrogate import (FrequencyTruncationPolicy, MajoranaSurrogateModel, MajoranaSurrogatePropagator,  # noqa: F401
                        ParamSource, PauliSurrogateModel, PauliSurrogatePropagator, SurrogateMajoranaCircuit,
                        SurrogateMajoranaRotation, SurrogatePauliCircuit, SurrogateRotation,
                        VariationalSurrogateModel)
from .truncation import (CoefficientTruncator, FrequencyTruncator, Simplify, TermBudget,  # noqa: F401
                         TruncationPolicy, Truncator, WeightTruncator, resolve_truncation)

__all__ = [
    "AbstractCircuit", "AbstractRotation", "MajoranaCircuit", "MajoranaRotation", "PauliCircuit", "PauliRotation",
    "MajoranaMonomial", "MajoranaTermStreamer", "MajoranaTermSum", "PauliString", "PauliTermStreamer",
    "PauliTermSum", "GateNoiseModel", "NativeNoiseModel", "NoiseModel", "UniformNoiseModel", "NativeTruncator",
    "MajoranaPropagator", "PauliPropagator", "PropagationResult", "CoefficientTruncator", "FrequencyTruncator",
    "Simplify", "TermBudget", "TruncationPolicy", "Truncator", "WeightTruncator", "resolve_truncation",
    "GradientRun", "expectation", "run_adjoint", "Logger", "LogParser", "FrequencyTruncationPolicy",
    "MajoranaSurrogateModel", "MajoranaSurrogatePropagator", "PauliSurrogateModel", "PauliSurrogatePropagator",
    "SurrogateMajoranaCircuit", "SurrogateMajoranaRotation", "SurrogatePauliCircuit", "SurrogateRotation",
    "ParamSource", "VariationalSurrogateModel",
]
