"""pico_amd: MI355X (gfx950) backend for the commitment hot path of Pico's
STARK prover over KoalaBear / BabyBear -- coset LDE, Poseidon2 mixed-height
Merkle commitment (MMCS), FRI folding, commit phase, PoW and openings.

The compute path is libpico_gpu.so (hand-written HIP kernels behind the C ABI
in include/pico_gpu.h).  Importing this package loads it and fails loudly if
it is missing; there is no CPU fallback.
"""
from ._lib import (BABYBEAR, KOALABEAR, AdmissionError, PgError, lib)  # noqa: F401
from .core import *  # noqa: F401,F403

lib()  # load now: a missing library is an import error, not a silent fallback
load_tuning()  # the measured LDE tile-shape table (host-side; no GPU touched)  # noqa: F405
