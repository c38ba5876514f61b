"""MI355X (gfx950) implementation of BEEP's stochastic simulation and
likelihood hot path.  The device code and its C ABI live in csrc/ and
include/beep_gpu.h; this package holds the ctypes binding (_lib), model specs
(spec, configs), the batched drivers and the serving loop."""
from . import spec, configs  # noqa: F401
