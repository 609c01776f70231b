"""pikoc_hip: the --target=HIP code generator for Piko pipes (the HIP
counterpart of pikoc's PTX and CPU backends, pikoc/src/Backend)."""
from .frontend import FrontendError, parse_pipe  # noqa: F401
from .hip_backend import BackendError, HIPBackend, compile_pipe  # noqa: F401
from .summary import PipeSummary, StageSummary  # noqa: F401
