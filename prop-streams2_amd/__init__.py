"""MI355X-native multi-species reacting RHS and stiff chemistry for the STREAmS-2 multideal solver.

The hot path (auxiliary thermo/transport, KEEP/WENO convective sweeps, viscous and
multicomponent diffusive fluxes, finite-rate chemistry, stiff per-cell integration, halo
copies) runs as hand-written gfx950 HIP kernels in ``libstreams_hip.so`` behind a C-ABI
(``include/streams_hip.h``).  This package holds the host side around it: mechanism
tables, seeded inputs, a device-resident field with the reference's RK3 orchestration,
the time loop of the reference's run with its input-file reader, the batched reactor
integrator and the continuous-batching reactor server.
"""
import importlib as _importlib

__all__ = ["mechanism", "capi", "grid", "cases", "solver", "roofline", "reactors", "serving", "build", "driver"]


def __getattr__(name):
    if name in __all__:
        return _importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
