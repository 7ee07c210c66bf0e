"""arctn_amd: MI355X (gfx950) executor for ArcTN contraction plans.

Public surface (mirrors the reference's arctn_py / arctn Python API):
``CompiledContraction``, ``contract_sliced``, ``arctn_contract``, saved
execution-plan v2 artifacts (``ExecutionPlan``) and the lower-level ``Plan``;
``contract`` (autograd) lives in ``arctn_amd.autograd``.
"""

from .api import (CompiledContraction, GroupedContraction, Plan, arctn_contract, contract_grouped,  # noqa: F401
                  contract_sliced)
from .execution_plan import ExecutionPlan, PlanValidationError  # noqa: F401

__version__ = "0.1.0"
__all__ = ["CompiledContraction", "ExecutionPlan", "GroupedContraction", "Plan", "PlanValidationError",
           "arctn_contract", "contract_grouped", "contract_sliced"]
