"""Readers of dedispersed time series files: PRESTO .inf/.dat and SIGPROC
.tim (reading/presto.py and reading/sigproc.py of the reference)."""
from .presto import PrestoInf
from .sigproc import SigprocHeader

__all__ = ["PrestoInf", "SigprocHeader"]
