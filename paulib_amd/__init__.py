"""paulib_amd -- MI355X (gfx950) engine for PauLIB's PauliSum hot path.

The compute path is libpaulib_hip.so (HIP kernels behind the C ABI in
include/paulib_hip.h); this package is its Python host side: device
contexts and sums mirroring PauliSum<N,double>, host SoA containers, and
synthetic workload generators.
"""
from ._lib import EngineError, LIB_PATH, lib  # noqa: F401
from .device import (  # noqa: F401
    Context,
    PauliSum,
    TruncationOptions,
    apply_truncation,
    batch_multiply_canonical,
    batch_multiply_canonical_each,
    commutator,
    commutator_canonical,
    i_commutator,
    job_terms,
    propagate,
    propagate_expectation,
    run_program,
)
from .host import HostSum, words_for  # noqa: F401
