"""MI355X-native batched transaction engine (LTPMG's batched CC + execute path).

The product is native: libltpmg_hip.so (HIP kernels for gfx950 behind the C
ABI in include/ltpmg.h) and libltpmg_host.so (the C++ host API and the seeded
workload generators).  This package only binds them for tests and benches.
"""
from .engine import (MODE_A, MODE_O, OP_READ, OP_RMW_ADD, OP_UPDATE, Batch, Engine,
                     LtpmgError, Stream, StreamPlan, concat_batches, gen_kv, meta, zeta)

__all__ = ["Engine", "Batch", "LtpmgError", "Stream", "StreamPlan", "concat_batches", "gen_kv",
           "meta", "zeta", "MODE_A", "MODE_O", "OP_READ", "OP_UPDATE", "OP_RMW_ADD"]
