"""egsolver_amd -- energy-game solver for AMD MI355X (gfx950).

Drop-in replacement of fvella/EGSolver's GPU path: the fixpoint of the energy lifting
operator runs as hand-written HIP kernels in ``lib/libegs.so`` behind a plain C-ABI
(``include/egs.h``); the reference's entry points (``include/egs_dev_solver.h``) and C
command-line driver sit on top.  This Python package is the host-side mirror used by
tests and the benchmark.
"""
from .games import (Game, GenParams, config_game, generate, family_params, read_mpg, write_mpg, parse_mpg,  # noqa: F401
                    save_csr, load_csr, TOP)
from .solver import Solver, solve, options, device_info  # noqa: F401
from .batch import Batch, batch_options  # noqa: F401,E402
from .serve import Server, Request, serve_options, run_served, run_plain, latency_stats  # noqa: F401,E402
