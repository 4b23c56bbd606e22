"""tractor_amd: the Tractor's render-and-fit hot path on AMD MI355X.

Drop-in counterparts of the reference's public classes (Tractor, Image,
PointSource, ExpGalaxy, DevGalaxy, PixelizedPSF, GaussianMixturePSF, ...,
LsqrOptimizer, ConstrainedOptimizer, GpuOptimizer); all pixel work runs in
libtractor_hip (gfx950 HIP kernels) through a ctypes C ABI.  The batched
fitter for thousands of sources is tractor_amd.batch.BatchFitter.
"""
from .utils import (BaseParams, ScalarParam, NamedParams, ParamList, MultiParams,
                    GaussianPriorsMixin)
from .patch import Patch, ModelMask
from .calib import (NullWCS, PixelWCS, PixPos, Flux, Fluxes, NanoMaggies, NullPhotoCal,
                    LinearPhotoCal, NullSky, ConstantSky)
from .image import Image
from .ellipses import EllipseE, EllipseESoft
from .psf import (PixelizedPSF, GaussianMixturePSF, NCircularGaussianPSF, HybridPSF,
                  HybridPixelizedPSF, lanczos_shift_image)
from .pointsource import PointSource, SingleProfileSource, BasicSource
from .galaxy import (GalaxyShape, Galaxy, ProfileGalaxy, HoggGalaxy, GaussianGalaxy, ExpGalaxy,
                     DevGalaxy)
from .sersic import SersicGalaxy, SersicIndex, SersicMixture
from .mixture_profiles import MixtureOfGaussians, get_exp_mixture, get_dev_mixture
from .engine import Tractor, Images, Catalog, OptResult
from .optimize import Optimizer
from .lsqr_optimizer import LsqrOptimizer
from .constrained_optimizer import ConstrainedOptimizer, GpuOptimizer

__all__ = [n for n in dir() if not n.startswith('_')]
