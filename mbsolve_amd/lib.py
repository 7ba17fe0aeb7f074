"""The mbsolve API (device, material, qm_description, scenario, record,
result, solver and writer factories), as in the reference's ``mbsolve.lib``
module.  Raises ImportError with a build hint if the extension is missing."""
try:
    from ._mbsolve import *  # noqa: F401,F403
    from ._mbsolve import numerics  # noqa: F401
    from ._mbsolve import (build_problem, create_hip_solver,  # noqa: F401
                           init_fdtd_simulation, solver_hip_batch)
except ImportError as exc:  # pragma: no cover - exercised without a build
    raise ImportError(
        "mbsolve_amd native extension not built; run "
        "'python -m mbsolve_amd.build' (%s)" % exc) from exc
