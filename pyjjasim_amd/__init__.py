"""pyjjasim_amd: MI355X-native time evolution of Josephson-junction arrays."""
from .graph import *  # noqa: F401,F403
from .circuit import *  # noqa: F401,F403
from .time_evolution import *  # noqa: F401,F403
from .batch import TimeEvolutionBatch  # noqa: F401
from .static_problem import (StaticProblem, StaticConfiguration, NewtonIterInfo,  # noqa: F401
                             DEF_TOL, DEF_NEWTON_MAXITER)


def __getattr__(name):
    # the differentiable path needs torch; import it only when asked for
    if name == "DifferentiableEvolution":
        from .adjoint import DifferentiableEvolution
        return DifferentiableEvolution
    if name == "AnnealingProblem":
        from .annealing import AnnealingProblem
        return AnnealingProblem
    raise AttributeError(name)
