"""twoldg: the MI355X-native field-evolution path of TreeWaveOLDG.jl.

The host side mirrors the reference's Julia API (src/TreeWaveOLDG.jl:53-85);
all arithmetic of the hot path runs in libtwoldg.so (HIP, gfx950) behind the C
ABI of include/twoldg.h.
"""
from .element import Element, order
from .forest import Forest, wave_forest, nested_cubes, refinement_regions, check_layout
from .problem import (WaveProblem, wave_rhs, truncate_state, truncate_stage, initial_state,
                      fill_at_nodes, wave_energy, wave_norm, quadrature_norm, oldg_timestep,
                      wave_integrator, Integrator, exchange_plan)
from .drivers import (wave_omega, wave_exact, reflect_omega, reflect_exact, gaussian_pulse,
                      PlaneWave, sevenpatch_exact, wave_errors, reflect_errors, sevenpatch_errors,
                      convergence_rate)
from .batch import Batch, gaussian_pulses
from .drivers import ball_forest, radiative_errors
from .geometry import SevenPatchBall, sevenpatch_forest
from .thirteenpatch import (ThirteenPatchBall, thirteenpatch_forest, RadiativeBoundary,
                            OutgoingPulse, ZeroData)
from .annulus import (PolarAnnulus, annulus_forest, annulus_wavenumber, annulus_exact,
                      annulus_errors)
from .regrid import (RegridFlag, Refine, Keep, Coarsen, level, doublefactorial, check_mode,
                     check_thresholds, modal_indicator, mode_amplitude, refine_mark, flag_blocks,
                     refine_flags, refinement_buffer, regrid_forest, transfer_map, regrid,
                     adapt_to_initial_data)
from .blast import (blast_initial, blast_reference, blast_exact, uniform_blast, adaptive_blast,
                    chunk_count, chunk_bounds)
from . import _lib

__all__ = [
    "Element", "order", "Forest", "wave_forest", "nested_cubes", "refinement_regions",
    "check_layout", "WaveProblem", "wave_rhs", "truncate_state", "truncate_stage",
    "initial_state", "fill_at_nodes", "wave_energy", "wave_norm", "quadrature_norm",
    "oldg_timestep", "wave_integrator", "Integrator", "exchange_plan", "wave_omega",
    "wave_exact", "reflect_omega", "reflect_exact", "gaussian_pulse", "PlaneWave",
    "sevenpatch_exact", "wave_errors", "reflect_errors", "sevenpatch_errors",
    "convergence_rate", "Batch", "PolarAnnulus", "annulus_forest", "annulus_wavenumber",
    "annulus_exact", "annulus_errors",
    "gaussian_pulses", "ball_forest", "radiative_errors", "SevenPatchBall", "sevenpatch_forest",
    "ThirteenPatchBall", "thirteenpatch_forest", "RadiativeBoundary", "OutgoingPulse",
    "ZeroData", "RegridFlag", "Refine", "Keep", "Coarsen", "level", "doublefactorial",
    "check_mode", "check_thresholds", "modal_indicator", "mode_amplitude", "refine_mark",
    "flag_blocks", "refine_flags", "refinement_buffer", "regrid_forest", "transfer_map", "regrid",
    "adapt_to_initial_data", "blast_initial", "blast_reference", "blast_exact", "uniform_blast",
    "adaptive_blast", "chunk_count", "chunk_bounds",
]
