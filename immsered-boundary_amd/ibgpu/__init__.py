"""ibgpu — host side of the MI355X residual engine (ctypes over the C ABI)."""
from . import _ffi  # noqa: F401
from .mesh import (CaseSpec, Mesh, config1, config2, config3, ensemble_cases, tiny3d_case,  # noqa: F401
                   tiny_case)
from .engine import Engine, IBGError, make_config  # noqa: F401,E402
from .closure import plain_engine, train_closure  # noqa: F401,E402
