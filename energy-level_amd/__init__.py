"""rydhip — MI355X-native matrix-free Rydberg Hamiltonian engine (host side).

H = Σ Ω_i/2 σx_i + Σ delta_i n_i + Σ_{i<j} V_ij n_i n_j on the full 2^N basis,
applied by hand-written gfx950 HIP kernels behind the C ABI of
``include/rydhip.h``; this package mirrors the reference's Python entry
points (tomisnot/energy-level src/core/manybody/) on top of that ABI.
"""
from . import _native
from .model import RydModel, from_calib, active_terms, chain_positions, square_positions, \
    from_positions, c6_coupling
from .plan import Plan, CsrPlan, cast
from .lanczos import hip_lanczos_ground_state, lanczos_run, ritz_vectors, rayleigh_residual, \
    site_occupations, hamiltonian_linear_operator_hip, ground_is_unique, locked_lowest
from .basis import fibonacci, chain_blockade_dim, is_chain_blockade_basis, sector_codes, \
    sector_rank, chain_edges, number_sector_basis, check_number_conserving, validate_edges, \
    graph_plan, independent_states_hip, independent_set_count_hip, blockade_edges, build_plan_hip
from .sector_scan import scan_ground_sectors_hip
from .evolve import CONV, cheb_coefficients, terms_for, PulseProblem, PulseEvolution, z2_index, \
    slice_bounds, segment_intervals, segment_drive, propagate_segments_hip
from .batch import BatchLanczos, BatchCsrLanczos, CsrInstance, disorder_ensemble, \
    batched_ground_state_hip, batched_ground_state_dense_hip, batched_ground_state_csr_hip
from .dispatch import EQ_BACKENDS_HIP, normalize_backend, solve_equilibrium_hip, \
    spectrum_lanczos_hip, quadrature, lorentz_spectrum, source_spectrum_hip, resolve_source
from .results import describe_result
from .observables import normalize_pauli_strings, PauliPlanHip, measure_pauli_strings_hip, \
    corr_matrix_hip, site_correlation_sums, loop_observables, link_loop_observables
from .manybody import evolve_manybody_hip, state_vector_np, product_label
from .kpm import kpm_moments_hip, spectral_bounds, jackson_kernel, thermal_from_moments, \
    dos_from_moments, probe_bits, kpm_moments_gpu_hip

__all__ = ['hamiltonian_linear_operator_hip', 'loop_observables', 'link_loop_observables', 'RydModel', 'from_calib', 'active_terms', 'chain_positions', 'square_positions',
           'from_positions', 'c6_coupling', 'Plan', 'CsrPlan', 'batched_ground_state_dense_hip', 'batched_ground_state_csr_hip', 'BatchCsrLanczos', 'CsrInstance', 'cast', 'kpm_moments_gpu_hip', 'hip_lanczos_ground_state',
           'lanczos_run', 'ritz_vectors', 'rayleigh_residual', 'site_occupations',
           'ground_is_unique', 'locked_lowest',
           'kpm_moments_hip', 'spectral_bounds', 'jackson_kernel', 'thermal_from_moments',
           'dos_from_moments', 'probe_bits', 'fibonacci', 'chain_blockade_dim',
           'is_chain_blockade_basis', 'sector_codes', 'sector_rank', 'chain_edges', 'CONV',
           'cheb_coefficients', 'terms_for', 'PulseProblem', 'PulseEvolution', 'z2_index',
           'slice_bounds', 'BatchLanczos', 'disorder_ensemble', 'batched_ground_state_hip',
           'EQ_BACKENDS_HIP', 'normalize_backend', 'solve_equilibrium_hip', 'spectrum_lanczos_hip',
           'quadrature', 'lorentz_spectrum', 'source_spectrum_hip', 'resolve_source',
           'describe_result', 'segment_intervals', 'segment_drive',
           'propagate_segments_hip', 'number_sector_basis', 'check_number_conserving',
           'scan_ground_sectors_hip', 'validate_edges', 'graph_plan', 'independent_states_hip',
           'independent_set_count_hip', 'blockade_edges', 'build_plan_hip',
           'normalize_pauli_strings', 'PauliPlanHip', 'measure_pauli_strings_hip',
           'corr_matrix_hip', 'site_correlation_sums', 'evolve_manybody_hip', 'state_vector_np',
           'product_label']
