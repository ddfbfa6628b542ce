"""ffields: MI355X-native (gfx950 HIP) Factor Fields field evaluation and
volume rendering, behind the reference's FactorFields model API.

Every step of the field and render path runs in libffields.so (hand-written
HIP kernels behind a C ABI, include/ffields.h); PyTorch provides device memory
and streams.  The evaluation metrics (ffields.evaluation: PSNR, SSIM) are
computed with torch on the rendered images.
"""

from .config import AttrDict, load_config
from .field import FactorFields, MLPMixer
from .optim import FusedAdam
from .params import N_to_reso, cal_n_samples, dct_basis, setup_params

__all__ = ["AttrDict", "load_config", "FactorFields", "MLPMixer", "FusedAdam", "N_to_reso",
           "cal_n_samples", "dct_basis", "setup_params"]
