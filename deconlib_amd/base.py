"""PDHG result type (base.py:14-34); the name is kept for drop-in use."""

from dataclasses import dataclass, field
from typing import Any, List


@dataclass
class MLXDeconvolutionResult:
    restored: Any
    iterations: int
    loss_history: List[float] = field(default_factory=list)
    converged: bool = False
    tau_history: List[float] = field(default_factory=list)
    sigma_history: List[float] = field(default_factory=list)
    metadata: dict = field(default_factory=dict)


DeconvolutionResult = MLXDeconvolutionResult
