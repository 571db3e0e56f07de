"""dtmm_amd -- MI355X-native engine for the diffractive transfer-matrix path.

Drop-in for ``dtmm.transfer_field`` (4x4 method: reflection mode 1, any
diffraction and npass; 2x2 method: reflection modes 0 and 1) with the
reference's data model, plus ``transfer_batch`` for samples of uneven size and
``TransferGraph``, the same batch captured once as a HIP graph and replayed,
and the polarizing-microscope step after it (``pom``: specter and colour).
Propagation runs in hand-written HIP kernels for gfx950 (``dtmm_amd/csrc``)
reached through the C ABI of ``include/dtmmx.h``; there is no CPU fallback.
"""
from .data import (refind2eps, uniaxial_order, director2data, director2angles,  # noqa: F401
                   director2order, nematic_droplet_director, nematic_droplet_data,
                   sphere_mask, cholesteric_director, cholesteric_droplet_data,
                   validate_optical_block, validate_optical_data, effective_block,
                   effective_data, eig_symmetry, eps2epsva, epsva2eps)
from .field import (illumination_data, field2intensity, blackman, k0, betaphi,  # noqa: F401
                    planewave, aperture, illumination_aperture, aperture2rays,
                    illumination_rays)
from .transfer import (transfer_field, transfer_batch, TransferGraph, Plan, Stack,  # noqa: F401
                       DTMM_NORM_FFT, DTMM_NORM_REF, field2betaphi, mean_betaphi)
from .propagate import propagate_4x4_effective_1, propagate_2x2_effective_1  # noqa: F401
from .adjoint import AdjointTransfer, polarizer4x4, analyser_intensity_loss  # noqa: F401
from .pom import calculate_pom_specter, field2specter, specter2color  # noqa: F401
from . import pom  # noqa: F401
from . import tmm  # noqa: F401
from . import tmm2d  # noqa: F401
from .io import save_stack, load_stack, save_field, load_field  # noqa: F401

__version__ = "0.1.0"
