"""MI355X (gfx950) hot path of MCVSLAM: ORB extraction, Hamming matching,
projection-guided window search and multi-camera BA linearization, as HIP
kernels behind the C ABI in include/mcvslam_gpu.h."""
from .api import (KP_DTYPE, BALinearizer, BowVocabulary, dbow_match, Context, Matcher, MatchResKnn, ORB, RigBatch, default_context,
                  distinctive_descriptors, filter_orientation, kp_records, window_search, window_search_dev,
                  window_search_uv_dev)
from ._native import LIB_PATH, MCVError

__all__ = ["KP_DTYPE", "BALinearizer", "BowVocabulary", "dbow_match", "Context", "Matcher", "MatchResKnn", "ORB", "RigBatch", "default_context",
           "distinctive_descriptors", "filter_orientation", "kp_records", "window_search", "window_search_dev", "window_search_uv_dev",
           "LIB_PATH", "MCVError"]
