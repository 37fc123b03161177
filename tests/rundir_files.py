"""Synthetic run directories for the reader tests (the writer lives in the package)."""
from ocean_amd.rundir import EV2HARTREE as EV2H
from ocean_amd.rundir import write_synthetic_rundir as write_rundir

__all__ = ["EV2H", "write_rundir"]
