"""Configuration types mirroring Draugr.jl's (src/types.jl:1-584).

Type tags keep the reference's names, field meaning and defaults; the
configuration is handed to the C-ABI as the JSON document that
``draugr_amg_config_from_json`` accepts (src/cfunction_interface.jl:178-221),
extended with this backend's keys (``precision``, ``seed``, ``graphs``, ``gpu_setup``,
``initial_theta``, ``initial_interpolation``, ``sa_omega``, ``filtering``,
``filter_tol``, ``aggressive_base``).
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from typing import Optional, Union


# ── interpolation (src/types.jl:1-57) ────────────────────────────────────────
@dataclass(frozen=True)
class DirectInterpolation:
    trunc_factor: float = 0.0

    def to_json(self):
        return {"type": "direct", "trunc_factor": self.trunc_factor}


@dataclass(frozen=True)
class StandardInterpolation:
    trunc_factor: float = 0.0

    def to_json(self):
        return {"type": "standard", "trunc_factor": self.trunc_factor}


@dataclass(frozen=True)
class ExtendedIInterpolation:
    trunc_factor: float = 0.0
    max_elements: int = 0
    norm_p: int = 1
    rescale: bool = False

    def to_json(self):
        return {"type": "extended_i", "trunc_factor": self.trunc_factor, "max_elements": self.max_elements,
                "norm_p": self.norm_p, "rescale": self.rescale}


InterpolationType = Union[DirectInterpolation, StandardInterpolation, ExtendedIInterpolation]


# ── coarsening (src/types.jl:59-154) ─────────────────────────────────────────
@dataclass(frozen=True)
class AggregationCoarsening:
    θ: float = 0.25
    filtering: bool = False
    filter_tol: float = 0.1
    name = "aggregation"


@dataclass(frozen=True)
class PMISCoarsening:
    θ: float = 0.25
    interpolation: InterpolationType = field(default_factory=DirectInterpolation)
    name = "pmis"


@dataclass(frozen=True)
class HMISCoarsening:
    θ: float = 0.25
    interpolation: InterpolationType = field(default_factory=DirectInterpolation)
    name = "hmis"


@dataclass(frozen=True)
class RSCoarsening:
    θ: float = 0.25
    interpolation: InterpolationType = field(default_factory=DirectInterpolation)
    name = "rs"


@dataclass(frozen=True)
class AggressiveCoarsening:
    θ: float = 0.25
    base: str = "pmis"
    interpolation: InterpolationType = field(default_factory=ExtendedIInterpolation)

    @property
    def name(self):
        return "aggressive_hmis" if self.base == "hmis" else "aggressive_pmis"


@dataclass(frozen=True)
class SmoothedAggregationCoarsening:
    θ: float = 0.25
    ω: float = 2.0 / 3.0
    filtering: bool = False
    filter_tol: float = 0.1
    name = "smoothed_aggregation"


# ── strength / smoothers (src/types.jl:156-180) ──────────────────────────────
class AbsoluteStrength:
    name = "absolute"


class SignedStrength:
    name = "signed"


class JacobiSmootherType:
    name = "jacobi"


class ColoredGaussSeidelType:
    name = "colored_gs"


class L1JacobiSmootherType:
    name = "l1_jacobi"


class L1ColoredGaussSeidelType:
    name = "l1_colored_gs"


class SerialGaussSeidelType:
    name = "serial_gs"


class L1SerialGaussSeidelType:
    name = "l1_serial_gs"


class SPAI0SmootherType:
    name = "spai0"


class SPAI1SmootherType:
    name = "spai1"


class ChebyshevSmootherType:
    name = "chebyshev"


class ILU0SmootherType:
    name = "ilu0"


class SerialILU0SmootherType:
    name = "serial_ilu0"


# ── AMGConfig (src/types.jl:497-536) ─────────────────────────────────────────
@dataclass
class AMGConfig:
    coarsening: object = field(default_factory=lambda: HMISCoarsening(0.5, ExtendedIInterpolation()))
    smoother: object = field(default_factory=L1ColoredGaussSeidelType)
    max_levels: int = 20
    max_coarse_size: int = 50
    pre_smoothing_steps: int = 1
    post_smoothing_steps: int = 1
    jacobi_omega: float = 2.0 / 3.0
    verbose: int = 0
    initial_coarsening: Optional[object] = None
    initial_coarsening_levels: int = 0
    max_row_sum: float = 1.0
    cycle_type: str = "V"
    strength_type: object = field(default_factory=AbsoluteStrength)
    coarse_solve_on_cpu: bool = False
    # backend extensions
    precision: str = "fp64"
    seed: int = 0x5EED
    # True: replay the preconditioner cycle (batched: each PCG iteration) from a captured hipGraph; False:
    # eager launches; "auto": graphs for batches, eager launches for one system
    graphs: object = "auto"
    # sparse-pass kernel: "auto" (autotuned), "stream", "rowseq", "pipe", "stream16", "sorted", "stream16pipe",
    # "sortedpipe", "sortedquad", "vector" (FP32 / BF16 cycles; FP64 passes keep the stream kernel), "pattern"
    # (operators whose rows repeat a few column patterns relative to the row, e.g. a grid stencil; others take
    # "stream")
    kernel: str = "auto"
    fuse: str = "auto"  # zero-guess step + residual, one-pass down / up sweeps: "auto" (autotuned), "off", "on"
    gpu_setup: bool = True  # interpolation, coarse operators, triple maps and PMIS sweeps of the setup on the GPU
    gpu_setup_min_rows: int = 32768  # smaller levels interpolate on the host
    gs_symmetric: bool = False  # coloured GS post-sweeps in reverse colour order (a symmetric cycle, for PCG)
    batch_rhs_per_thread: int = 0  # batched passes: rhs per thread (0 autotuned, 1, 4, 8, 16; -1, -4 LDS-staged 1, 4)
    tail: str = "auto"  # small levels and the coarse solve in one launch: "auto" (autotuned), "off", "on" (deepest)

    def __post_init__(self):
        if str(self.cycle_type).upper() not in ("V", "W"):
            raise ValueError("cycle_type must be :V or :W")
        if isinstance(self.verbose, bool):
            self.verbose = int(self.verbose)
        if self.initial_coarsening is None:
            self.initial_coarsening = self.coarsening
        if self.precision not in ("fp64", "fp32", "bf16"):
            raise ValueError("precision must be 'fp64', 'fp32' or 'bf16'")

    def to_json(self) -> str:
        c, ic = self.coarsening, self.initial_coarsening
        d = {
            "coarsening": c.name,
            "theta": c.θ,
            "smoother": {"type": self.smoother.name, "omega": self.jacobi_omega},
            "strength": self.strength_type.name,
            "cycle": str(self.cycle_type).lower(),
            "max_levels": self.max_levels,
            "max_coarse_size": self.max_coarse_size,
            "pre_smoothing_steps": self.pre_smoothing_steps,
            "post_smoothing_steps": self.post_smoothing_steps,
            "verbose": self.verbose,
            "max_row_sum": self.max_row_sum,
            "coarse_solve_on_cpu": self.coarse_solve_on_cpu,
            "initial_coarsening": ic.name,
            "initial_coarsening_levels": self.initial_coarsening_levels,
            "initial_theta": ic.θ,
            "precision": self.precision,
            "seed": self.seed,
            "graphs": self.graphs,
            "kernel": self.kernel,
            "gpu_setup": self.gpu_setup,
            "fuse": self.fuse,
            "gpu_setup_min_rows": self.gpu_setup_min_rows,
            "gs_symmetric": self.gs_symmetric,
            "batch_rhs_per_thread": self.batch_rhs_per_thread,
            "tail": self.tail,
        }
        if hasattr(c, "interpolation"):
            d["interpolation"] = c.interpolation.to_json()
        if hasattr(ic, "interpolation"):
            d["initial_interpolation"] = ic.interpolation.to_json()
        if isinstance(c, (AggregationCoarsening, SmoothedAggregationCoarsening)):
            d["filtering"] = c.filtering
            d["filter_tol"] = c.filter_tol
        if isinstance(c, SmoothedAggregationCoarsening):
            d["sa_omega"] = c.ω
        for a in (c, ic):
            if isinstance(a, AggressiveCoarsening) and a.base not in ("pmis", "hmis"):
                d["aggressive_base"] = "aggregation"
        return json.dumps(d)


def hypre_default_config(θ: float = 0.5, agg_num_levels: int = 1, agg_trunc_factor: float = 0.3, **kwargs):
    """BoomerAMG-like configuration (src/types.jl:569-584): HMIS + ext+i with
    truncation, aggressive HMIS coarsening on the first level(s)."""
    interp = ExtendedIInterpolation(agg_trunc_factor)
    return AMGConfig(coarsening=HMISCoarsening(θ, interp),
                     initial_coarsening=AggressiveCoarsening(θ, "hmis", interp),
                     initial_coarsening_levels=agg_num_levels, **kwargs)
