"""ducc_amd: MI355X-native NUFFT (types 1 and 2, 1-3 D) behind the
ducc0.nufft call surface.  `ducc_amd.nufft` mirrors ducc0.nufft;
`ducc_amd.torch_nufft` wraps u2nu / nu2u in torch.autograd.Function;
`ducc_amd.fft.c2c` mirrors ducc0.fft.c2c on the same FFT kernels."""
from . import fft, nufft

__all__ = ["fft", "nufft"]
