"""MI355X (gfx950) implementation of MDFF's force/field evaluation.

The compute path is the HIP library ``libmdffgpu.so`` (mdff_amd/csrc), reached
through its C ABI (include/mdffgpu.h).  This package holds the host-side
mirror of the reference's interface: parameter set-up (``params``), the
engforce front-end (``field``), synthetic configurations (``configs``) and
the batched force-matching loss with its ``torch.autograd.Function``
(``batch.force_match_loss``).
"""
from .params import MdffEnergies, MdffParams  # noqa: F401
