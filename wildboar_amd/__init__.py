"""wildboar_amd -- MI355X (gfx950) kernels for wildboar's distance hot path.

``wildboar_amd.distance`` mirrors ``wildboar.distance``'s ``pairwise_distance``,
``paired_distance``, ``argmin_distance`` and ``pairwise_subsequence_distance``
for ``metric`` in {"euclidean", "dtw", "wdtw"}; the arithmetic runs in
``lib/libwbx.so`` (HIP, C ABI in ``include/wbx.h``).
"""

__version__ = "0.1.0"

from .distance import (  # noqa: E402,F401
    argmin_distance,
    check_metric,
    paired_distance,
    pairwise_distance,
    pairwise_subsequence_distance,
)
from .variable_len import EOS, get_variable_length, is_end_of_series  # noqa: E402,F401
