"""jax_fli_amd: an MI355X-native particle-mesh engine with the public API of jax-fli's PM path.

Imported as ``jax_fli_amd`` (the directory is ``jax-fli_amd/``; ``__graft_entry__.load_package``
registers it). The compute path is libpmx.so (hand-written HIP for gfx950, C ABI in include/pmx.h);
PyTorch only provides device memory and streams. Names follow ``import jax_fli as jfli``.
"""

from . import _lib, autodiff, cosmology, engine, fields, initial, io, lensing, pm, power, prng, redshift, serving, utils
from .cosmology import Cosmology, Planck18
from .fields import (AbstractField, ConvergenceUnit, DensityField, DensityUnit, FieldStatus, FlatDensity,
                     FlatKappaField, ParticleField, PhysicalUnit, PositionUnit, SphericalDensity,
                     SphericalKappaField)
from .initial import gaussian_initial_conditions, interpolate_initial_conditions
from .lensing import born
from .redshift import redshift_distribution, tophat_z
from .painting import PaintingOptions
from .pm import (AbstractCorrection, AbstractInterp, AbstractNBodySolver, DriftInterp, EfficientDriftDoubleKick,
                 InterpTilerState, NBodyState, NoCorrection, NoInterp, OnionTiler, PGDKernel,
                 ReversibleDoubleKickDrift, SharpeningKernel, TelephotoInterp, lpt, nbody)
from .power import PowerSpectrum, coherence, power, transfer
from .utils import centers, compute_lightcone_shells, distances, edges

__version__ = "0.1.0"
