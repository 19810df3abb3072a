"""bdsim_amd -- bdsim's ChainIntegrator step on AMD Instinct MI355X.

The `_bdsim`-compatible surface (parameter types, Rng, integrate, forces) is in
`core`; the trajectory drivers with the `backend="gpu"` batch plug point are in
`ensemble` and `parallel`.
"""
from .core import (Spring, EV, Bending, DelSMethod, SpringParams, EVParams, BendingParams,
                   PhysParams, SimParams, Flow, Rng, ExternalForce, integrate, spring_force,
                   bending_force, ev_force, total_force, external_force)
from . import core
