"""bdsim_amd -- bdsim's ChainIntegrator step on AMD Instinct MI355X.

The `_bdsim`-compatible surface (parameter types, Rng, integrate, forces) is in
`core`; the trajectory drivers with the `backend="gpu"` batch plug point are in
`ensemble` and `parallel`.
"""
from .core import (Spring, EV, Bending, DelSMethod, SpringParams, EVParams, BendingParams,
                   PhysParams, SimParams, Flow, Rng, ExternalForce, integrate, spring_force,
                   bending_force, ev_force, total_force, external_force)
from . import core, flows, initial, properties, storage, parallel, ensemble, resident, rheology
from ._abi import PRECISION_FP64, PRECISION_MIXED
from .ensemble import (Initial, Output, INITIALIZERS, n_steps_of, simulate, run_ensemble,
                       trajectory_samples, final_state, sampled_states, mean_stderr,
                       MAX_FAILED_FRACTION)
from .parallel import parallel_map, Failure
from .properties import (end_to_end_sq, radius_of_gyration_sq, gyration_tensor, stretch,
                         kramers_stress, normal_stress_differences)
from .initial import (gaussian_chain, fene_fraenkel_chain, fene_fraenkel_chain_aligned_x,
                      equilibrium_chain, fene_fraenkel_bending_chain)
