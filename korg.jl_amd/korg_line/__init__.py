"""MI355X line-opacity path of Korg.jl: host-side mirror over the gfx950 library.

    from korg_line import line_absorption_, Line, LineList, Wavelengths, Species

``line_absorption_`` is the drop-in for Korg's ``line_absorption!``
(``src/line_absorption.jl:33-129``); ``Context``/``DeviceProblem`` give the device-resident form.
"""
from .constants import *  # noqa: F401,F403
from .species import Species, species  # noqa: F401
from .wavelengths import Wavelengths  # noqa: F401
from .linelist import Line, LineList, approximate_gammas, approximate_radiative_gamma  # noqa: F401
from .problem import Problem, ContinuumGrid  # noqa: F401
from .line_absorption import (Context, DeviceProblem, FitTarget, FluxTarget, Plan,  # noqa: F401
                             line_absorption_, line_absorption_vjp_, line_absorption_jvp_)

__all__ = ["Species", "species", "Wavelengths", "Line", "LineList", "approximate_gammas",
           "approximate_radiative_gamma", "Problem", "ContinuumGrid", "Context", "DeviceProblem",
           "Plan", "FitTarget", "FluxTarget", "line_absorption_", "line_absorption_vjp_",
           "line_absorption_jvp_"]
