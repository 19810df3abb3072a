"""bdsim_amd -- bdsim's ChainIntegrator step on AMD Instinct MI355X.

`core` presents the reference's `_bdsim` surface (parameter types, Rng,
integrate, forces) over the HIP library; `gpu_map` is the backend a
`parallel_map(..., backend="gpu")` dispatches to, batching the integrate calls
of per-trajectory functions; `resident` keeps ensembles in HBM between
segments; `serving` admits jobs into a running device batch.
"""
from .core import (Spring, EV, Bending, DelSMethod, SpringParams, EVParams, BendingParams,
                   PhysParams, SimParams, Flow, Rng, ExternalForce, integrate, spring_force,
                   bending_force, ev_force, total_force, external_force)
from . import core, flows, initial, properties, storage, resident, rheology
from ._abi import PRECISION_FP64, PRECISION_MIXED
from .gpu_map import gpu_map
from .properties import (end_to_end_sq, radius_of_gyration_sq, gyration_tensor, stretch,
                         kramers_stress, normal_stress_differences)
from .initial import (gaussian_chain, fene_fraenkel_chain, fene_fraenkel_chain_aligned_x,
                      equilibrium_chain, fene_fraenkel_bending_chain)
