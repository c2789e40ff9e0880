"""vikshobha-mi355x: the 2D Gross-Pitaevskii pseudospectral step of
SourangshuGhosh/QuantumVortices on AMD MI355X (gfx950), behind a C ABI
(include/vikshobha.h)."""
from .gpe import GPE, params_from_namelist, nshell, forward, inverse  # noqa: F401
