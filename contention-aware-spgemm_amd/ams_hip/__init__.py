"""MI355X hybrid-accumulation SpGEMM (C = A * B, CSR) behind a C ABI.

Host mirror of the reference's AMS hybacc path: ``csr`` (matrix container),
``gen`` (seeded synthetic inputs), ``device`` (torch-facing calls into
libams_hip.so), ``batched`` and ``autograd`` (the batched and differentiable
forms), ``sparse`` (torch.sparse_csr_tensor front end and the ams::
torch.library operators), ``mtx`` (MatrixMarket / PaRMAT text files through
the library's reader and writer), ``parmat`` (R-MAT edge lists with PaRMAT's
construction) and ``joinagg`` (join-aggregate incidence matrices and their
org/trs files).  Importing the package needs numpy only; the device modules
load the native library on first use and raise if it is missing.
"""
from .csr import Csr  # noqa: F401

__all__ = ["Csr"]
