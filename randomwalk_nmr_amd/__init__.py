"""MI355X-native random-walk NMR hot path (RandomWalk_NMR drop-in for the
CPMG / PFGSE walker kernels, echo reduction, Laplace inversion and the
relaxivity fit).  Compute runs in librwnmr_amd.so (HIP kernels for gfx950
behind the C ABI in include/rwnmr/); this package is the Python host mirror.
"""
from ._lib import RwnmrError, load  # noqa: F401
from .core import (BC, Context, Image, bc_code, bitblock_dims, collision_histogram, cpmg_batch_run,  # noqa: F401
                   CpmgGraph, cpmg_run_label_sets, cpmg_forward_jacobian,
                   cpmg_field_run, cpmg_histogram_simulate, cpmg_run, histogram_decay, histogram_list,
                   histogram_penalties, map_run,  # noqa: F401
                   pack_bitblocks, penalty, pfgse_run, pfgse_run_sampled, pfgse_sample_stats, seed_and_place,
                   setup_walkers, sphere_pack, time_framework,
                   time_step, voronoi_labels)
from .images import (UctImage, image_stats, internal_field_uniform, read_internal_field, read_png,  # noqa: F401
                     write_image_info, write_walkers_csv)

from .config import ConfigFile  # noqa: F401
from .drivers import cpmg_simulate, multitau_sequence, multitau_simulate, pfgse_simulate  # noqa: F401
from . import outputs  # noqa: F401

__version__ = "0.1.0"
