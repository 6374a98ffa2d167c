"""MI355X-native Behler-Parrinello energy/force/training path (HIP kernels behind a C-ABI)."""
from .model import Model, Network, Radial, Angular, tri_index  # noqa: F401
from .nnp import NNPDevice, PREC_FP32, PREC_BF16L1, PREC_FP64_BASIS  # noqa: F401
