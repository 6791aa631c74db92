"""MI355X-native implementation of the self-absorption-corrected XRF
reconstruction hot path of gmysage/FL_correction: attenuation maps, the rotated
emission projector and its transpose, MLEM/OSL-Huber, the adjoint of the
attenuated forward model, and batched execution -- hand-written HIP kernels for
gfx950 behind the reference's Python API.

The package namespace re-exports the hot path's part of the reference's
top-level ``FL`` namespace (FL/__init__.py:7-36): the attenuation drivers, map
files, detector mask, torch helpers, MLEM wrappers and the op set.  Importing
it does not load the HIP library; the first op that runs does.
"""
__version__ = "0.2.0"

from .atten_util import (extract_em_cs_elem, extract_single_value_param, get_atten_coef,  # noqa: F401
                         make_param, update_dict_use_ref_at_angle)
from .core import (cal_atten_at_angles, cal_total_atten_at_angles, cal_total_atten_stack,  # noqa: F401
                   cal_xray_atten_at_angles, cal_xrf_atten_at_angles, re_projection_cuda)
from .cuda_lib.atten_cuda import (backward_emission, backward_emission_batch, cal_atten_cuda,  # noqa: F401
                                  forward_emission, forward_emission_batch, mlem_cuda, mlem_cuda_batch)
from .cuda_lib.autograd import forward_emission_autograd, forward_emission_batch_autograd  # noqa: F401
from .detector_mask import generate_detector_mask  # noqa: F401
from .file_io import (read_attenuation, read_attenuation_at_all_angle, read_attenuation_fl,  # noqa: F401
                      read_attenuation_xray, write_attenuation, write_attenuation_fl, write_attenuation_xray)
from .torch_functions import (torch_3D_forward_projection, torch_extract_3D_mu_xray,  # noqa: F401
                              torch_extract_3D_mu_xrf, torch_xray_atten_3D, torch_xrf_atten_3D)
from .torch_image_util import torch_rot90_3D, torch_rot90_4D, torch_rotate_image_nd  # noqa: F401
from .torch_recon import torch_mlem_recon, torch_mlem_recon_batch  # noqa: F401
