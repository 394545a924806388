"""MI355X-native implementation of StarFormationHistories.jl's two
data-parallel paths: template-mixture likelihood (composite!, loglikelihood,
∇loglikelihood(!), fg!, batched over walkers) and smoothed Hess-diagram
template deposition (bin_cmd_smooth / partial_cmd_smooth).

All numerical work runs in libsfh_hip.so (hand-written HIP for gfx950,
C ABI in include/sfh.h).  Importing this package loads that library and
raises if it has not been built; there is no CPU fallback.
"""
from . import _lib

_lib.load()

from ._lib import ArgumentError, DeviceError, DomainError, InexactError  # noqa: E402
from .core import (Context, HIPData, HIPTemplates, composite_, default_context, driver_clamps, fg_,  # noqa: E402
                   fg_batched, fg_batched_device, fg_device, fg_device_noalloc, grad_loglikelihood,
                   grad_loglikelihood_, loglikelihood, stack_models)
from .fitting import (HMCModel, LogTransformFTResult, MCMCModel, construct_x0,  # noqa: E402
                      fit_templates, fit_templates_fast, fit_templates_lbfgsb, mcmc_sample,
                      renormalize_x0)
from .multifit import MultiFit, StageInfo  # noqa: E402
from . import hierarchical  # noqa: E402
from . import mock  # noqa: E402
from .mock import (BinaryMassRatio, PowerLawIMF, Salpeter1955, bin_cmd, generate_stars_mag,  # noqa: E402
                   generate_stars_mag_composite, generate_stars_mass, generate_stars_mass_composite, ingest_mags,
                   mass_limits, model_cmd, recycling_fraction, recycling_mass_fraction, sort_ingested,
                   surviving_fraction, surviving_mass_fraction)
from .templates import (Histogram, NoBinaries, PointBatch, RandomBinaryPairs, Range,  # noqa: E402
                        binary_hess, binary_hess_batch, binary_mass_fraction, binary_number_fraction,
                        binary_pairs_hist, binary_pairs_hist_batch,
                        binary_system_fraction, bin_cmd_smooth, calculate_edges, flux2mag, mag2flux, trapz,
                        calculate_weights, deposit_batch, edges_range, exp_photerr,
                        histogram_data, histogram_pix, interpolate_mini, Martin2016_complete,
                        midpoints, mini_spacing, partial_cmd_smooth, partial_cmd_smooth_points,
                        templates_from_device, partial_cmd_smooth_batch, ExpPhotErr, Martin2016Completeness,
                        Constant, Tabulated, IsochroneBatch)

__all__ = [
    "ArgumentError", "DeviceError", "InexactError", "Context", "HIPData", "HIPTemplates",
    "composite_", "default_context", "driver_clamps", "fg_", "fg_batched", "fg_batched_device", "fg_device",
    "fg_device_noalloc",
    "grad_loglikelihood", "grad_loglikelihood_", "loglikelihood", "stack_models", "Histogram",
    "PointBatch", "Range", "bin_cmd_smooth", "calculate_edges", "calculate_weights",
    "deposit_batch", "edges_range", "exp_photerr", "histogram_data", "histogram_pix",
    "interpolate_mini", "Martin2016_complete", "midpoints", "mini_spacing", "partial_cmd_smooth",
    "partial_cmd_smooth_points", "templates_from_device", "HMCModel", "LogTransformFTResult",
    "MCMCModel", "construct_x0", "fit_templates", "fit_templates_fast", "fit_templates_lbfgsb",
    "mcmc_sample", "renormalize_x0", "MultiFit", "StageInfo", "NoBinaries", "RandomBinaryPairs",
    "binary_hess", "binary_hess_batch", "binary_pairs_hist_batch", "binary_mass_fraction", "binary_number_fraction", "binary_pairs_hist",
    "binary_system_fraction", "flux2mag", "mag2flux", "trapz", "hierarchical", "DomainError", "mock",
    "BinaryMassRatio", "PowerLawIMF", "Salpeter1955", "bin_cmd", "generate_stars_mag",
    "generate_stars_mag_composite", "generate_stars_mass", "generate_stars_mass_composite", "ingest_mags",
    "mass_limits", "model_cmd", "recycling_fraction", "recycling_mass_fraction", "sort_ingested",
    "surviving_fraction", "surviving_mass_fraction", "partial_cmd_smooth_batch", "ExpPhotErr",
    "Martin2016Completeness", "Constant", "Tabulated", "IsochroneBatch",
]
