"""sap_warp_amd — MI355X-native SAP convex compliant-contact step.

Drop-in for the reference's ``SolverSAP`` / ``SapCollisionPipeline`` step API
(``sim/solver_sap.py``, ``sim/collision/pipeline.py``) with the hot path in
hand-written gfx950 HIP kernels behind the C ABI declared in ``include/sapx.h``.
"""
from .runtime import (  # noqa: F401
    SapContacts,
    SapControl,
    SapModel,
    SapState,
    allocate_contacts,
)

__all__ = ["SapModel", "SapState", "SapControl", "SapContacts", "allocate_contacts"]


def __getattr__(name):
    if name == "SolverSAP":
        from .solver_sap import SolverSAP

        return SolverSAP
    if name == "SapCollisionPipeline":
        from .collision import SapCollisionPipeline

        return SapCollisionPipeline
    raise AttributeError(name)
