"""sigpde.torch drop-in: `import sigpde_amd.torch as sig` replaces `import sigpde.torch as sig`
(stojl/SigPDE sigpde/torch/__init__.py:9-16)."""
from . import normalizers
from . import static_kernels as kernels
from .kernel import RobustSigPDE, SigPDE

__all__ = ["SigPDE", "RobustSigPDE", "kernels", "normalizers"]
