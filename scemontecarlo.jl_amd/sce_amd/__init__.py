"""sce_amd — MI355X-native device Metropolis sweep of SCEMonteCarlo.jl.

Host-side mirror of the reference's GPU API over the C ABI of libsce_hip.so
(include/sce_hip.h).  Import puts nothing on the GPU; the first device call
loads the library and fails loudly if it has not been built.
"""
from .models import (MultipoleTerm, Setup, config_model, lattice_model, random_config,  # noqa: F401
                     read_setup, setup_magmoms, setup_model)
from .hamiltonian import (TiledHamiltonian, ChainState, resolve_kt, KB_EV,  # noqa: F401
                          MU_B_EV_T)
from .gpu import (GPUTiledHamiltonian, GPUChainState, gpu_metropolis_sweep,  # noqa: F401
                  gpu_overrelaxation_sweep, gpu_run_sweeps, gpu_renormalize, to_host,
                  from_host, swap_payload, DEFAULT_LANES, gpu_energy_gradient,
                  gpu_chain_energy_gradient, gpu_time_gradient, GRADIENT_LANES, autotune_chain)
from .observables import (LogBinner, BinStore, jackknife, standard_stats,  # noqa: F401
                          effective_samples, tau_int_windowed)
from .batch import GPUBatch, gpu_run_pt, resume_pt, autotune_lanes, autotune_batch, run_concurrent  # noqa: F401
from .surrogate import Surrogate, gpu_da_run_sweeps  # noqa: F401

__all__ = [
    "MultipoleTerm", "lattice_model", "config_model", "random_config", "TiledHamiltonian",
    "ChainState", "resolve_kt", "KB_EV", "MU_B_EV_T", "GPUTiledHamiltonian", "GPUChainState",
    "gpu_metropolis_sweep", "gpu_overrelaxation_sweep", "gpu_run_sweeps", "gpu_renormalize",
    "to_host", "from_host", "swap_payload", "DEFAULT_LANES", "LogBinner", "BinStore",
    "jackknife", "standard_stats", "effective_samples", "tau_int_windowed", "GPUBatch", "gpu_run_pt", "resume_pt",
    "autotune_lanes", "autotune_batch", "run_concurrent", "autotune_chain", "Surrogate", "gpu_da_run_sweeps",
    "gpu_energy_gradient",
    "gpu_chain_energy_gradient", "gpu_time_gradient", "GRADIENT_LANES",
]
