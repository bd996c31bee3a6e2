"""MI355X growth machine for SCORE-style static-recrystallization cellular automata.

The GPU path is libscore_gm.so (HIP kernels for gfx950 behind the C ABI of
include/score_gm.h); this package holds its ctypes binding (`abi`), the Python
face of the growth-machine call (`growth.GrowthMachine`), synthetic inputs
(`synth`) and the in-tree build (`build`).
"""
from . import abi  # noqa: F401
from .synth import DomainArrays, ParamsSpec, make_voronoi_domain, workload  # noqa: F401

__all__ = ["abi", "DomainArrays", "ParamsSpec", "make_voronoi_domain", "workload", "GrowthMachine"]


def __getattr__(name):
    if name == "GrowthMachine":
        from .growth import GrowthMachine
        return GrowthMachine
    raise AttributeError(name)
