"""Writes SCORE input decks in the current .uds layout (the block and key names
SCORE's ensembleHdl::init_parameter_* read, SCORE_Kernel.cpp:262-940) for the
setup tests: a cuboid-grain, CSR-nucleation iron deck with overridable values.
`python tests/golden/make_uds_deck.py` rewrites tests/golden/cuboid_fe.uds."""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

DEFAULTS = {
    "RuntimeControl": {"XMAX": "0.990", "TIMEMAX": "360000.0", "NMAX": "100000", "PERCOLATIONANALYSIS": "0",
                       "MAXFILLPERSTEP": "0.1", "RediscretizationSteps": "1000", "InitialCacheSizeRXCells": "0.20",
                       "ReCacheSizeRXCells": "1.0", "OnTheFlyDefragmentation": "1", "SolitaryUnitPostProcessing": "1",
                       "ReadEBSDMapping": "0", "ReadEBSDStepsize": "0.25E-06"},
    "SimulationOutput": {"RenderingOfMicrostructure": "0", "RenderingColorModel": "1", "RenderingFileFormat": "1",
                         "RenderingBoundaries": "0", "OutputLogBoundaries": "0", "OutputRXFrontStats": "0",
                         "OutputThreadProfiling": "1", "OutputSingleGrainStats": "1",
                         "OutputDAMASKGeometryFile": "0", "OutputArtificialSEMEBSD": "0"},
    "EnsembleDefinition": {"CAEnsembleSize": "4", "CellSize": "0.5", "3DCAEdgeLengthInCellsRD": "64",
                           "3DCAEdgeLengthInCellsTD": "60", "3DCAEdgeLengthInCellsND": "56",
                           "MeanDefGrainSizeinRD": "8.0", "MeanDefGrainSizeinTD": "7.0", "MeanDefGrainSizeinND": "6.0",
                           "DefStructureSynthesis": "1", "MeanDefSizePoissonDiameter": "10.0"},
    "NucleationModeling": {"NucleationDensityLocalCSR": "60", "CSRNucleation": "1", "ClusteredNucleation": "0",
                           "ClusteredNucleationNCluster": "10", "ClusteredNucleationLCluster": "10.0",
                           "ClusteredNucleationScaling": "100.0", "ClusteredNucleationExtendA": "5",
                           "ClusteredNucleationExtendB": "5", "ClusteredNucleationExtendC": "5", "GBNucleation": "0",
                           "GBNucDensity2Number": "1.0E-4", "GBNucRhoDiff2Density": "5E-14",
                           "GBNucMaximumScatter": "5.0", "IncubationTimeModel": "1", "IncubationTimeScatter": "100.0"},
    "MaterialProperties": {"ZeroKelvinShearModulusG0": "6.92e10", "FirstOrderdG0dT": "0.0",
                           "ZeroCelsiusBurgersVector": "2.48e-10", "AlloyConstantThermalExpCoeff": "1.0",
                           "FirstOrderThermalExpCoeff": "0.0", "SecondOrderThermalExpCoeff": "0.0",
                           "MeltingTemperature": "1538.0"},
    "BoundaryMobilityModel": {"MobilityModel": "2", "LAGBm0": "0.3", "LAGBHact": "1.2", "HAGBm0": "3.0",
                              "HAGBHact": "1.2", "GSm0": "3.0", "GSHact": "1.2", "RHModelHAGBm0": "0.1",
                              "RHModelHAGBHact": "2.8", "RHModelLAGBHAGBCut": "0.9", "RHModelLAGBHAGBTrans": "5.0",
                              "RHModelLAGBHAGBExponent": "9.0"},
    "RecoveryParameter": {"RecoveryConsider": "0", "RecoveryVacancyDiffGeometry": "1.0",
                          "RecoveryVacancyDiffPreexp": "1.0", "RecoveryVacancyDiffHact": "2.9",
                          "RecoverySoluteDiffPreexp": "1.0", "RecoverySoluteDiffHact": "1.6",
                          "RecoverySoluteLsPropFactor": "1.0", "RecoverySoluteConcentration": "1.0",
                          "RecoveryParameterAlpha3": "1.0", "RecoveryParameterKappa2": "0.1",
                          "RecoveryParameterC3": "1.0", "RecoveryParameterC4": "1.0", "RecoveryParameterC5": "0.65"},
    "DragParameter": {"ZenerConsider": "0", "ZenerAlpha": "1.5", "ZenerGamma": "0.324"},
}


def pool_rows(seed: int = 5, n: int = 24):
    """deformed grains: Bunge angles in degrees and rho0; rows 1 and 7 lie within
    0.3 degree of row 0, so the orientation pool merges them"""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        e = (rng.uniform(0, 360), rng.uniform(0, 180), rng.uniform(0, 360))
        rows.append((f"{e[0]:.4f}", f"{e[1]:.4f}", f"{e[2]:.4f}", f"{10 ** rng.uniform(14, 15.3):.4E}"))
    e0 = rows[0]
    rows[1] = (f"{float(e0[0]) + 0.2:.4f}", e0[1], e0[2], rows[1][3])
    rows[7] = (e0[0], f"{float(e0[1]) + 0.1:.4f}", e0[2], rows[7][3])
    return rows


def deck_text(overrides: dict | None = None, pool=None, schedule=((0.0, 726.85), (360000.0, 726.85)),
              defrag=(0.30, 0.60), rxpool=((0.0, 30.0, 45.0, 0.0), (55.0, 55.0, 45.0, 0.0)),
              ideals=(("Cube", 0.0, 0.0, 0.0, 10.0),)) -> str:
    blocks = {k: dict(v) for k, v in DEFAULTS.items()}
    for (blk, key), val in (overrides or {}).items():
        blocks[blk][key] = val
    out = []
    for name, kv in blocks.items():
        out.append(f"|| {name}\t\t|| *(ss)\t\t|| Parameter, Value\n")
        for k, v in kv.items():
            out.append(f'"{k}"\t\t\t"{v}"\t\t// {name}\n')
        out.append("\n")
    out.append("|| EvoMicrochemistry\t\t|| *(ff)\t\t|| time, fr\n0.0\t\t0.00E+0\n360000.0\t\t0.00E+0\n\n")
    out.append("|| IdealComponents\t\t|| *(sffff)\t\t|| Name, Bunge1, Bunge2, Bunge3, scatter\n")
    for name, b1, b2, b3, sc in ideals:
        out.append(f'"{name}"\t\t{b1}\t\t{b2}\t\t{b3}\t\t{sc}\n')
    out.append("\n")
    out.append("|| AnnealingSchedule\t\t|| *(ff)\t\t|| time, temperature\n")
    for t, T in schedule:
        out.append(f"{t}\t\t{T}\n")
    out.append("\n|| DeformedGrainsPool\t\t|| *(ffff)\t|| Bunge e1, Bunge e2, Bunge e3, rho0\n")
    for r in (pool if pool is not None else pool_rows()):
        out.append("\t\t".join(r) + "\n")
    out.append("\n|| RXGrainsPool\t\t|| *(ffff)\t\t|| Bunge e1, Bunge e2, Bunge e3, tincub\n")
    for r in rxpool:
        out.append("\t\t".join(str(x) for x in r) + "\n")
    out.append("\n|| HeuristicRXFrontListDefragmentation\t||\t*(sf)\t|| X\n")
    for x in defrag:
        out.append(f'"X"\t\t{x:.2f}\n')
    out.append("\n|| SingleGrainVolOverTime\t||  *(sf)\t\t|| X\n" + "".join(f'"X"\t\t{x / 10:.3f}\n' for x in range(11)))
    return "".join(out)


def write_deck(path: str, **kw) -> str:
    with open(path, "w") as f:
        f.write(deck_text(**kw))
    return path


if __name__ == "__main__":
    print(write_deck(os.path.join(HERE, "cuboid_fe.uds")))
