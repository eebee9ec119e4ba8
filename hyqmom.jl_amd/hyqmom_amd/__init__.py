"""hyqmom_amd: host side of the MI355X HyQMOM.jl moment-solver step.

Mirrors the reference's Julia API for this path (simulation_runner(params),
run_simulation_with_snapshots) over the C ABI of libhyqmom_hip.so (include/hyqmom.h).
"""
from .moments import E35, NAMES35, maxwellian35
from .initial_conditions import CubicRegion, crossing_jets, initialize_moment_field, cell_centers
from .runner import (Solver, simulation_runner, run_simulation, run_simulation_with_snapshots,
                     initial_state, grid)
