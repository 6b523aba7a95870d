"""sieve_amd: the MI355X search path of sieve (route -> 1-bit RaBitQ LUT scan -> int8 rerank).

The compute path is libsieve_gpu.so (hand-written HIP for gfx950 behind the C ABI in
include/sieve_gpu.h); this package binds it (abi, index), mirrors the reference's query
parameters (params) and builds synthetic workloads in the reference's index format (synth,
build).
"""
from .params import Params, Rerank, parse_config, ConfigError  # noqa: F401
