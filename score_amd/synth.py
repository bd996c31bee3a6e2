"""Synthetic growth-machine inputs: Voronoi deformed structures, pools, physics.

The reference cannot generate Voronoi structures (its POISSONVORONOI_DEFMS
switch is parsed at SCORE_Kernel.cpp:453 but never dispatched), so the
workloads are injected the way SCORE holds a domain after
solve_REPLACE_CA_STRUCTURE (SCORE_Kernel.cpp:6353-6445): a cell grid of
deformed-grain ids, an orientation pool, per-grain dislocation densities and
nucleus sites.  Everything is a pure function of the seed.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
SYNTH_LIB = os.path.join(HERE, "libsynth.so")

ECHARGE = 1.602176565e-19  # SCORE_Defs_Physics.h:53

_synth = None


def _lib() -> C.CDLL:
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_LIB):
            raise FileNotFoundError(f"{SYNTH_LIB} is missing: run `python -m score_amd.build`")
        lib = C.CDLL(SYNTH_LIB)
        u32p, f64p = C.POINTER(C.c_uint32), C.POINTER(C.c_double)
        lib.synth_quaternions.argtypes = [C.c_uint64, C.c_uint32, f64p]
        lib.synth_loguniform.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_double, f64p]
        lib.synth_lognormal.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_double, f64p]
        lib.synth_points.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, f64p]
        lib.synth_distinct_sites.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, u32p]
        lib.synth_distinct_sites.restype = C.c_int
        lib.synth_distinct_sites_gb.argtypes = [C.c_uint64, C.c_uint32, u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                                C.c_double, u32p]
        lib.synth_distinct_sites_gb.restype = C.c_int
        for fn in (lib.synth_voronoi, lib.synth_voronoi_bruteforce):
            fn.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, f64p, C.c_double, C.c_double,
                           C.c_double, C.c_int, u32p]
            fn.restype = C.c_int
        _synth = lib
    return _synth


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def substream(seed: int, k: int) -> int:
    """Independent 64-bit seeds per purpose (splitmix64 of seed and stream id)."""
    z = (int(seed) * 0x9E3779B97F4A7C15 + (k + 1) * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def quaternions(seed: int, n: int) -> np.ndarray:
    q = np.empty((n, 4), np.float64)
    if n:
        _lib().synth_quaternions(seed, n, _ptr(q, C.c_double))
    return q


def loguniform(seed: int, n: int, lo: float, hi: float) -> np.ndarray:
    out = np.empty(n, np.float64)
    if n:
        _lib().synth_loguniform(seed, n, lo, hi, _ptr(out, C.c_double))
    return out


def lognormal(seed: int, n: int, mu: float, sigma: float) -> np.ndarray:
    out = np.empty(n, np.float64)
    if n:
        _lib().synth_lognormal(seed, n, mu, sigma, _ptr(out, C.c_double))
    return out


def points(seed: int, n: int, nx: int, ny: int, nz: int) -> np.ndarray:
    xyz = np.empty((n, 3), np.float64)
    if n:
        _lib().synth_points(seed, n, nx, ny, nz, _ptr(xyz, C.c_double))
    return xyz


def distinct_sites(seed: int, n: int, N: int) -> np.ndarray:
    s = np.empty(n, np.uint32)
    if n and _lib().synth_distinct_sites(seed, n, N, _ptr(s, C.c_uint32)) != 0:
        raise ValueError("cannot draw that many distinct sites")
    return s


def distinct_sites_gb(seed: int, n: int, grid: np.ndarray, nx: int, ny: int, nz: int, bias: float) -> np.ndarray:
    """n distinct sites, cells next to a grain boundary (a 26-neighbour in another
    grain, periodic in x and y) `bias` times as likely as interior cells."""
    s = np.empty(n, np.uint32)
    g = np.ascontiguousarray(grid, np.uint32)
    if n and _lib().synth_distinct_sites_gb(seed, n, _ptr(g, C.c_uint32), nx, ny, nz, float(bias),
                                            _ptr(s, C.c_uint32)) != 0:
        raise ValueError("cannot draw that many distinct sites")
    return s


def on_boundary(grid: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    """Cells with a 26-neighbour in another grain (periodic x and y, open z)."""
    g = np.asarray(grid).reshape(nz, ny, nx)
    out = np.zeros(g.shape, bool)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if (dx, dy, dz) == (0, 0, 0):
                    continue
                sh = np.roll(np.roll(g, -dy, axis=1), -dx, axis=2)
                if dz:
                    diff = np.zeros(g.shape, bool)
                    if dz > 0:
                        diff[:-1] = g[:-1] != sh[1:]
                    else:
                        diff[1:] = g[1:] != sh[:-1]
                else:
                    diff = g != sh
                out |= diff
    return out.ravel()


def voronoi(nx: int, ny: int, nz: int, seeds_xyz: np.ndarray, stretch=(1.0, 1.0, 1.0), periodic_xy=True,
            brute_force=False) -> np.ndarray:
    """Nearest-seed labels of all voxel centres, x fastest (SCORE's cellgrid layout)."""
    seeds_xyz = np.ascontiguousarray(seeds_xyz, np.float64)
    grid = np.empty(nx * ny * nz, np.uint32)
    fn = _lib().synth_voronoi_bruteforce if brute_force else _lib().synth_voronoi
    rc = fn(nx, ny, nz, len(seeds_xyz), _ptr(seeds_xyz, C.c_double), float(stretch[0]), float(stretch[1]),
            float(stretch[2]), 1 if periodic_xy else 0, _ptr(grid, C.c_uint32))
    if rc != 0:
        raise MemoryError("voronoi labelling failed")
    return grid


@dataclasses.dataclass
class DomainArrays:
    """A domain as SCORE holds it after solve_REPLACE_CA_STRUCTURE."""
    nx: int
    ny: int
    nz: int
    cellsize: float
    cellgrid: np.ndarray   # uint32[N] deformed ids
    ori_quat: np.ndarray   # float64[n_ori,4]
    defg_ori: np.ndarray   # uint32[n_defg]
    defg_rho: np.ndarray   # float64[n_defg]
    rxg_ori: np.ndarray    # uint32[n_rxg]
    rxg_site: np.ndarray   # uint32[n_rxg]
    rxg_step: np.ndarray | None = None   # uint32[n_rxg] arrival step (0: site saturation); None: all 0

    @property
    def N(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def n_defg(self) -> int:
        return len(self.defg_ori)

    @property
    def n_rxg(self) -> int:
        return len(self.rxg_ori)

    def to_c(self):
        """ctypes Domain plus the arrays it points into (keep both alive)."""
        keep = [np.ascontiguousarray(self.cellgrid, np.uint32), np.ascontiguousarray(self.ori_quat, np.float64),
                np.ascontiguousarray(self.defg_ori, np.uint32), np.ascontiguousarray(self.defg_rho, np.float64),
                np.ascontiguousarray(self.rxg_ori, np.uint32), np.ascontiguousarray(self.rxg_site, np.uint32)]
        if self.rxg_step is not None:
            keep.append(np.ascontiguousarray(self.rxg_step, np.uint32))
        d = abi.Domain()
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        d.cellsize = self.cellsize
        d.cellgrid = _ptr(keep[0], C.c_uint32)
        d.n_ori = len(keep[1])
        d.ori_quat = _ptr(keep[1], C.c_double)
        d.n_defg = len(keep[2])
        d.defg_ori = _ptr(keep[2], C.c_uint32)
        d.defg_rho = _ptr(keep[3], C.c_double)
        d.n_rxg = len(keep[4])
        d.rxg_ori = _ptr(keep[4], C.c_uint32)
        d.rxg_site = _ptr(keep[5], C.c_uint32)
        d.rxg_step = _ptr(keep[6], C.c_uint32) if self.rxg_step is not None else C.cast(None, C.POINTER(C.c_uint32))
        return d, keep


@dataclasses.dataclass
class ParamsSpec:
    """Growth settings; defaults are the Fe / Humphreys set of the 100^3 and 512^3 workloads."""
    mobility_model: int = abi.MOB_ROLLETTHOLM
    rh_hagb_m0: float = 1e-1            # m^4/(J s)
    rh_hagb_hact_ev: float = 2.8        # eV
    rh_cut: float = 0.9
    rh_trans: float = 5.0
    rh_exponent: float = 9.0
    lagb_m0: float = 0.3
    lagb_hact_ev: float = 1.2
    hagb_m0: float = 3.0
    hagb_hact_ev: float = 1.2
    gs_m0: float = 3.0
    gs_hact_ev: float = 1.2
    G0: float = 6.92e10
    b0: float = 2.48e-10
    sched_time: tuple = (0.0, 3.6e5)
    sched_temp: tuple = (1000.0, 1000.0)
    maxfill: float = 0.1
    xmax: float = 0.99
    tmax: float = 3.6e5
    nmax: int = 100000
    rel_cell_caching: float = 0.2
    rel_cell_recaching: float = 1.0
    defrag_at: tuple = ()
    zener_mode: int = 0                 # 0 none, 1 constant, 2 time-dependent
    zener_alpha: float = 1.5
    zener_gamma: float = 0.324          # J/m^2
    disp_time: tuple = ()               # s
    disp_fr: tuple = ()                 # f/r, 1/m
    recovery_mode: int = 0              # 0 none, 1 Nes vacancy-core diffusion, 2 Nes solute drag, 3 Michalak-Paxton
    rec_vacancy_d0: float = 1.0
    rec_vacancy_hact_ev: float = 2.9
    rec_solute_d0: float = 1.0
    rec_solute_hact_ev: float = 1.6
    rec_solute_ls: float = 1.0
    rec_solute_conc: float = 1.0
    rec_alpha3: float = 1.0
    rec_kappa2: float = 0.1
    rec_c3: float = 1.0
    rec_c5: float = 0.65
    rec_tau: float = 0.0                # s, exponential recovery (recovery_mode 4)
    grainlog_at: tuple = ()             # grain-evolution log thresholds, in cells
    n_regions: int = 0                  # OMP_NUM_THREADS of the SCORE run to reproduce (0/1: one region)

    def to_c(self):
        keep = [np.ascontiguousarray(self.sched_time, np.float64), np.ascontiguousarray(self.sched_temp, np.float64),
                np.ascontiguousarray(self.defrag_at if len(self.defrag_at) else [0.0], np.float64),
                np.ascontiguousarray(self.disp_time if len(self.disp_time) else [0.0], np.float64),
                np.ascontiguousarray(self.disp_fr if len(self.disp_fr) else [0.0], np.float64),
                np.ascontiguousarray(self.grainlog_at if len(self.grainlog_at) else [0.0], np.float64)]
        p = abi.Params()
        p.mobility_model = self.mobility_model
        p.zener_mode = self.zener_mode
        p.recovery_mode = self.recovery_mode
        p.n_regions = self.n_regions
        p.rh_hagb_m0 = self.rh_hagb_m0
        p.rh_hagb_hact = self.rh_hagb_hact_ev * ECHARGE
        p.rh_cut, p.rh_trans, p.rh_exponent = self.rh_cut, self.rh_trans, self.rh_exponent
        p.lagb_m0, p.lagb_hact = self.lagb_m0, self.lagb_hact_ev * ECHARGE
        p.hagb_m0, p.hagb_hact = self.hagb_m0, self.hagb_hact_ev * ECHARGE
        p.gs_m0, p.gs_hact = self.gs_m0, self.gs_hact_ev * ECHARGE
        p.G0, p.b0 = self.G0, self.b0
        p.n_sched = len(keep[0])
        p.sched_time = _ptr(keep[0], C.c_double)
        p.sched_temp = _ptr(keep[1], C.c_double)
        p.maxfill, p.xmax, p.tmax, p.nmax = self.maxfill, self.xmax, self.tmax, self.nmax
        p.rel_cell_caching, p.rel_cell_recaching = self.rel_cell_caching, self.rel_cell_recaching
        p.n_defrag = len(self.defrag_at)
        p.defrag_at = _ptr(keep[2], C.c_double)
        p.zener_alpha, p.zener_gamma = self.zener_alpha, self.zener_gamma
        p.n_disp = len(self.disp_time)
        p.disp_time = _ptr(keep[3], C.c_double)
        p.disp_fr = _ptr(keep[4], C.c_double)
        p.rec_vacancy_d0, p.rec_vacancy_hact = self.rec_vacancy_d0, self.rec_vacancy_hact_ev * ECHARGE
        p.rec_solute_d0, p.rec_solute_hact = self.rec_solute_d0, self.rec_solute_hact_ev * ECHARGE
        p.rec_solute_ls, p.rec_solute_conc = self.rec_solute_ls, self.rec_solute_conc
        p.rec_alpha3, p.rec_kappa2, p.rec_c3, p.rec_c5 = self.rec_alpha3, self.rec_kappa2, self.rec_c3, self.rec_c5
        p.rec_tau = self.rec_tau
        p.n_grainlog = len(self.grainlog_at)
        p.grainlog_at = _ptr(keep[5], C.c_double)
        return p, keep


def make_voronoi_domain(nx: int, ny: int, nz: int, n_grains: int, n_nuclei: int, seed: int,
                        cellsize: float = 0.5e-6, log10_rho=(14.0, 15.0), stretch=(1.0, 1.0, 1.0),
                        rho_values: np.ndarray | None = None) -> DomainArrays:
    """Voronoi deformed grains (periodic in x and y), uniform orientations, rho
    log-uniform in 10^[lo,hi) m^-2, nuclei at distinct uniform voxels with uniform orientations."""
    N = nx * ny * nz
    xyz = points(substream(seed, 0), n_grains, nx, ny, nz)
    grid = voronoi(nx, ny, nz, xyz, stretch=stretch, periodic_xy=True)
    q_def = quaternions(substream(seed, 1), n_grains)
    rho = rho_values if rho_values is not None else loguniform(substream(seed, 2), n_grains, *log10_rho)
    sites = distinct_sites(substream(seed, 3), n_nuclei, N)
    q_rx = quaternions(substream(seed, 4), n_nuclei)
    return DomainArrays(
        nx=nx, ny=ny, nz=nz, cellsize=cellsize, cellgrid=grid,
        ori_quat=np.concatenate([q_def, q_rx]) if n_nuclei else q_def,
        defg_ori=np.arange(n_grains, dtype=np.uint32),
        defg_rho=np.ascontiguousarray(rho, np.float64),
        rxg_ori=(n_grains + np.arange(n_nuclei)).astype(np.uint32),
        rxg_site=sites)


def make_rolled_sheet_domain(nx: int = 1024, ny: int = 1024, nz: int = 256, n_grains: int = 200000,
                             n_nuclei: int = 1000000, arrival_steps: int = 60, gb_bias: float = 3.0, seed: int = 1,
                             cellsize: float = 0.5e-6) -> DomainArrays:
    """Rolled sheet: pancake Voronoi grains (the metric stretched 4 : 1 : 0.25 along
    RD : TD : ND), rho log-normal (median 5e14 m^-2, sigma 0.5 of ln rho), uniform
    orientations; the nuclei at distinct sites, cells next to a grain boundary
    `gb_bias` times as likely; the first half site-saturated (step 0), the second
    half arriving at a constant rate over steps 1..arrival_steps, in pool order."""
    N = nx * ny * nz
    xyz = points(substream(seed, 0), n_grains, nx, ny, nz)
    grid = voronoi(nx, ny, nz, xyz, stretch=(4.0, 1.0, 0.25), periodic_xy=True)
    q_def = quaternions(substream(seed, 1), n_grains)
    rho = lognormal(substream(seed, 2), n_grains, float(np.log(5e14)), 0.5)
    sites = distinct_sites_gb(substream(seed, 3), n_nuclei, grid, nx, ny, nz, gb_bias)
    q_rx = quaternions(substream(seed, 4), n_nuclei)
    half = n_nuclei // 2
    steps = np.zeros(n_nuclei, np.uint32)
    if arrival_steps > 0 and n_nuclei > half:
        j = np.arange(n_nuclei - half, dtype=np.int64)
        steps[half:] = (1 + (j * arrival_steps) // (n_nuclei - half)).astype(np.uint32)
    return DomainArrays(
        nx=nx, ny=ny, nz=nz, cellsize=cellsize, cellgrid=grid,
        ori_quat=np.concatenate([q_def, q_rx]) if n_nuclei else q_def,
        defg_ori=np.arange(n_grains, dtype=np.uint32), defg_rho=np.ascontiguousarray(rho, np.float64),
        rxg_ori=(n_grains + np.arange(n_nuclei)).astype(np.uint32), rxg_site=sites, rxg_step=steps)


# Workloads (sizes and distributions of the 100^3 CPU case and the 512^3 single-GPU case)
WORKLOADS = {
    "voronoi100": dict(nx=100, ny=100, nz=100, n_grains=50, n_nuclei=500),
    "voronoi512": dict(nx=512, ny=512, nz=512, n_grains=8000, n_nuclei=100000),
}


def workload(name: str, seed: int = 1, scale: float | None = None) -> DomainArrays:
    """Named workload; `scale` shrinks the edge and scales grain and nucleus counts by volume."""
    w = dict(WORKLOADS[name])
    if scale is not None:
        e = max(8, int(round(w["nx"] * scale)))
        f = (e / w["nx"]) ** 3
        w.update(nx=e, ny=e, nz=e, n_grains=max(1, int(round(w["n_grains"] * f))),
                 n_nuclei=max(1, int(round(w["n_nuclei"] * f))))
    return make_voronoi_domain(seed=seed, **w)
