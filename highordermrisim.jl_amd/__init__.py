"""MI355X-native BlochHighOrder free precession and multi-coil reception.

The directory name is not a Python identifier, so callers load it with
``importlib`` under the module name ``hom_amd`` (``load_package`` in
``__graft_entry__.py``, tests/conftest.py and bench.py do exactly that).

Layers:
- ``types``      host mirror of the reference's data types and flags;
- ``workloads``  seeded synthetic inputs (reference test workload, configs);
- ``runtime``    ctypes binding of the C ABI in ``libhom.so`` (include/hom_abi.h);
- ``simulate``   the reference's ``simulate`` / ``prepare_simulation`` surface
                 on top of the ABI;
- ``serving``    continuous-batching request loop over the ABI's server;
- ``trajectory`` ``get_kspace`` (native ``hom_kspace``), ``rf_flip_angles``;
- ``sequence``   block descriptions and ``discretize`` (native ``hom_discretize``).
"""
from .types import (GAMMA, BlochHighOrder, DimensionMismatch, HO_DiscreteSequence,  # noqa: F401
                    HO_Phantom, get_sim_ranges)
from . import workloads  # noqa: F401

__all__ = ["GAMMA", "BlochHighOrder", "DimensionMismatch", "HO_DiscreteSequence", "HO_Phantom",
           "get_sim_ranges", "workloads"]

from .simulate import (Mag, PreparedSimulation, default_sim_params, prepare_simulation,  # noqa: E402,F401
                       setup_simulation, simulate, simulation_spin_ranges)

__all__ += ["Mag", "PreparedSimulation", "default_sim_params", "prepare_simulation",
            "setup_simulation", "simulate", "simulation_spin_ranges"]

from . import serving  # noqa: E402,F401
from .trajectory import get_kspace, rf_flip_angles  # noqa: E402,F401
from .sequence import SeqBlock, discretize  # noqa: E402,F401

__all__ += ["serving", "get_kspace", "rf_flip_angles", "SeqBlock",
            "discretize"]
