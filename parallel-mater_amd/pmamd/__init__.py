# SPDX-License-Identifier: MIT
"""Host-side mirror of the ParallelMater rigid step over the HIP driver."""
from . import layouts  # noqa: F401
