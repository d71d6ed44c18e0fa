"""MI355X (gfx950) AVBD rigid-body step: Python binding of include/avbd.h."""
from ._lib import (CONTACTS_CPU, CONTACTS_GPU, GRAPHS_NONE, GRAPHS_SOLVE, GRAPHS_STEP,  # noqa: F401
                   PREC_EXACT, PREC_F64, PREC_MIXED, SYMBOLS, LIB_PATH, lib)
from .backend import AvbdError, HipBackend, scene_arrays  # noqa: F401
from . import scenes  # noqa: F401
from . import sceneio  # noqa: F401
