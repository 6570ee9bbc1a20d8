"""riptide_amd: the riptide FFA periodogram (ffa_search and its kernels) on
AMD Instinct MI355X, with riptide's Python API.

Hand-written HIP kernels for gfx950 behind a C-ABI (include/riptide_amd.h),
loaded with ctypes; PyTorch-ROCm provides device buffers and streams.
"""
from .time_series import TimeSeries
from .periodogram import Periodogram
from .metadata import Metadata
from .search import ffa_search
from .running_medians import running_median, fast_running_median
from .libffa import ffa1, ffa2, ffafreq, ffaprd, generate_signal, downsample, boxcar_snr
from .ffautils import generate_width_trials
from .peak_detection import Peak, find_peaks
from .serialization import save_json, load_json
from ._lib import release_scratch

__version__ = "0.1.0"

__all__ = [
    "TimeSeries",
    "Periodogram",
    "Metadata",
    "ffa_search",
    "ffa1",
    "ffa2",
    "ffafreq",
    "ffaprd",
    "generate_signal",
    "downsample",
    "boxcar_snr",
    "running_median",
    "fast_running_median",
    "generate_width_trials",
    "Peak",
    "find_peaks",
    "save_json",
    "load_json",
    "release_scratch",
]
