"""MI355X-native likelihood / EM / SQP engine with the API of OpenADMIXTURE.jl.

Compute runs in hand-written gfx950 HIP kernels (libadmixhip.so, C ABI in
include/admixhip.h); this package is the host-side mirror of the reference's
interface for that path:

  AdmixData                          src/structs.jl
  *_loop_skipmissing                 src/loops.jl (tile loops, += semantics)
  loglikelihood, em_q, em_p,         src/algorithms_inner.jl
  update_q, update_p
  init_em, admixture_qn              src/algorithms_outer.jl
  run_admixture, _filter_SKFR        src/driver.jl
  ImputedSnpMatrix, sparsekmeans1/2, SKFR AIM pre-selection (SparseKmeansFeatureRanking
  sparsekmeans_repeat/_path          calls of src/driver.jl:56-123)
  HIPAdmixData, hip_admixture_base   src/cuda/structs.jl (device mirror)
  cv_sweep, autotune_waves           K sweep with n-fold CV (no reference counterpart)
  MLP, Panel, warm_started_fit       learned warm start (no reference counterpart)

The directory name contains a dot, so load it by path, e.g.
``importlib.util.spec_from_file_location("openadmixture_jl_amd", ".../__init__.py",
submodule_search_locations=[...])``.
"""
from . import _native
from ._native import AdmixHIPError, device_count
from .algorithms import admixture_qn, em_p, em_q, init_em, loglikelihood, update_p, update_q
from .cv import best_k, cv_sweep, fit_on_device
from .driver import _admixture_base, _filter_SKFR, hip_admixture_base, run_admixture
from .hip import HIPAdmixData
from .loops import (UnitRange, em_p_loop, em_p_loop_skipmissing, em_q_loop, em_q_loop_skipmissing,
                    loglikelihood_loop, loglikelihood_loop_skipmissing, update_p_loop, update_p_loop_skipmissing,
                    update_q_loop, update_q_loop_skipmissing)
from .plink import SnpArray, SnpLinAlg, filter_plink, write_plink
from .skfr import ImputedSnpMatrix, sparsekmeans1, sparsekmeans2, sparsekmeans_path, sparsekmeans_repeat
from .structs import AdmixData
from .tune import autotune, autotune_waves
from .warmstart import MLP, Panel, drifted_populations, training_set, warm_started_fit, world_frequencies

# Fail loudly at import when the native library is absent.
_native.lib()

__all__ = [
    "AdmixData", "HIPAdmixData", "hip_admixture_base", "run_admixture", "_admixture_base",
    "init_em", "admixture_qn", "loglikelihood", "em_q", "em_p", "update_q", "update_p",
    "loglikelihood_loop_skipmissing", "em_q_loop_skipmissing", "em_p_loop_skipmissing",
    "update_q_loop_skipmissing", "update_p_loop_skipmissing", "UnitRange",
    "loglikelihood_loop", "em_q_loop", "em_p_loop", "update_q_loop", "update_p_loop",
    "SnpArray", "SnpLinAlg", "write_plink", "AdmixHIPError", "device_count",
    "cv_sweep", "best_k", "fit_on_device", "autotune", "autotune_waves",
    "MLP", "Panel", "training_set", "warm_started_fit", "world_frequencies", "drifted_populations",
    "_filter_SKFR", "filter_plink", "ImputedSnpMatrix", "sparsekmeans1", "sparsekmeans2", "sparsekmeans_repeat",
    "sparsekmeans_path",
]
