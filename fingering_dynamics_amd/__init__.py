"""MI355X-native two-phase D2Q9 lattice-Boltzmann step of ebinan92/Fingering_dynamics.

Hot path: the per-iteration update of lattice_boltzmann/fingering_periodic.py
(collide, stream, half-way bounce-back, Zou-He inlet/outlet, moments) as one
fused pull-scheme HIP kernel for gfx950 behind the C ABI of include/fdlbm.h.

    from fingering_dynamics_amd import Compute, LbmParams, geometry
"""
from .params import LbmParams  # noqa: F401
from . import geometry  # noqa: F401


def __getattr__(name):
    # The solver class needs the native library; import it lazily so that the
    # geometry front-end and parameters work on machines without it.
    if name == "Compute":
        from .lbm import Compute
        return Compute
    raise AttributeError(name)
