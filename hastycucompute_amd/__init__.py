"""MI355X (gfx950) non-Cartesian reconstruction hot path.

Kaiser-Bessel NUFFT (type-1 spread, type-2 interp), Toeplitz-embedded SENSE
normal operator, and FP64-scalar CG, as hand-written HIP kernels behind the
C-ABI in ``include/hasty_hip.h`` (``libhasty_hip.so`` in this directory).
The Python modules mirror the reference's corelib operator API:

* ``hastycucompute_amd.fft`` -- ``nufft_plan``, ``toeplitz_kernel``,
  ``toeplitz_multiply``, ``toeplitz_multiplication`` ...
* ``hastycucompute_amd.mri`` -- ``NORMAL_T_T1_OP`` ... ``NORMAL_IDTW_T2_OP``,
  ``SenseOffresonance`` / ``sense_forward_offresonance`` / ``sense_adjoint_offresonance``
* ``hastycucompute_amd.op``  -- ``conjugate_gradient``, ``max_eig``
* ``hastycucompute_amd.offres`` -- time-segmentation interpolators
* ``hastycucompute_amd.io`` -- raw tensor files (``.htc``)
* ``hastycucompute_amd.hdf5`` -- HDF5 tensor files and the raw k-space loader
* ``hastycucompute_amd.generic_value`` -- the generic_value wire format
* ``hastycucompute_amd.ops`` -- ``torch.ops.hasty.*`` registrations
* ``hastycucompute_amd.autograd``, ``frames``, ``autotune``, ``trajectory``
"""
from ._lib import LIB_PATH, HastyError, lib  # noqa: F401

__all__ = ["LIB_PATH", "HastyError", "lib"]
