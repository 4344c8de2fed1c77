"""MHDStep (reference: src/integrator.py:6-82).

The stages initialize / stage1 / stage2 run as the HIP kernels k_initialize,
k_stage1, k_stage2 and, inside multi-step runs, the fused
k_stage2_init_stage1 (stage2 of step n with initialize + stage1 of n+1)."""
from __future__ import annotations

from .base import IntegratorStep


class MHDStep(IntegratorStep):
    def __init__(self, gamma):
        assert gamma > 1.0
        self.gamma1 = gamma - 1.0
