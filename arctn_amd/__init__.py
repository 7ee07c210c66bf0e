"""arctn_amd: MI355X (gfx950) executor for ArcTN contraction plans.

Public surface (mirrors the reference's arctn_py / arctn Python API):
``CompiledContraction``, ``contract_sliced``, ``arctn_contract`` and the
lower-level ``Plan``; ``contract`` (autograd) lives in ``arctn_amd.autograd``.
"""

from .api import (CompiledContraction, GroupedContraction, Plan, arctn_contract, contract_grouped,  # noqa: F401
                  contract_sliced)

__version__ = "0.1.0"
__all__ = ["CompiledContraction", "GroupedContraction", "Plan", "arctn_contract", "contract_grouped", "contract_sliced"]
