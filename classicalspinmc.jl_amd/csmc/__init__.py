"""csmc -- MI355X (gfx950) engine for the sweep hot path of ClassicalSpinMC.jl.

The reference API is mirrored with Julia's `!` dropped from mutating names
(addBilinear! -> addBilinear, simulated_annealing! -> simulated_annealing).
Everything that touches spins runs in libcsmc on the GPU; there is no CPU
fallback.
"""
from ._lib import (CONE, CONE_ADAPTIVE, CONSTRAINT_ADAPTIVE, DETERMINISTIC, FP32, FP64, METROPOLIS, OVERRELAX,
                   BoundsError, Context, CsmcError, UndefError, default_context, device_count, lib)
from .batch import SpinBatch
from .hdf5io import (create_params_file, dump_metadata, h5open, overwrite_keys, read_lattice,
                     read_spin_configuration, read_unit_cell, write_MC_checkpoint)
from .lattice import (Lattice, compute_equal_time_correlations, energy, energy_density, get_local_field,
                      get_magnetization, get_spin, local_fields, random_spin_orientation, set_spin,
                      site_energies, total_energy)
from .reciprocal import get_allowed_wavevectors, get_k_path, get_k_plane, reciprocal
from .spin_correlations import compute_equal_time_structure_factor, runEqualTimeStructureFactor
from .montecarlo import (MCParamsBuffer, Metropolis, MetropolisAdaptive, MetropolisConstraint,
                         MetropolisConstraintAdaptive, MetropolisFixedCone, MonteCarlo,
                         SimulationParameters, SiteSumConstraint, deterministic_updates, overrelaxation,
                         parallel_tempering, simulated_annealing)
from .unitcell import (FCC, BreathingPyrochlore, Honeycomb, IMToMatrix, InteractionMatrix, Pyrochlore,
                       Square, Triangular, UnitCell, addBasisSite, addBilinear, addCubic, addOnSite,
                       addQuartic, addZeemanCoupling, transposeJ)

__all__ = [n for n in dir() if not n.startswith("_")]
