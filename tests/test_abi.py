"""The C ABI library loads and exports every entry point include/sieve_gpu.h declares; the
parameter mirror parses the reference's config strings. No GPU calls (runs on the CPU)."""
import ctypes
import math
import subprocess

import pytest

from sieve_amd import Params, Rerank, parse_config, ConfigError, abi


def test_library_exports_every_header_symbol():
    syms = abi.header_symbols()
    assert len(syms) >= 15
    L = abi.lib()
    for s in syms:
        assert hasattr(L, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(syms) <= exported


def test_last_error_is_empty_string_initially():
    assert abi.lib().sieve_last_error() == b""


def test_struct_layouts_match_header():
    # sieve_params: 7 x 4 bytes + f32 margin, then u64 budget at offset 32 -> 40 bytes.
    assert ctypes.sizeof(abi.Params) == 40 and abi.Params.budget.offset == 32
    assert ctypes.sizeof(abi.Timing) == 72
    assert ctypes.sizeof(abi.LiveOpts) == 16 and ctypes.sizeof(abi.LiveInfo) == 48
    assert ctypes.sizeof(abi.ServerOpts) == 24 and ctypes.sizeof(abi.ServerStats) == 56
    assert abi.IndexDesc.borrow_rerank.offset == 24 + 12 * 8  # after k, n, nblocks and 12 pointers


def test_parse_config_mirrors_search_cli():
    # src/bin/search.rs:180-218
    label, p = parse_config("32:500:full", 4096)
    assert p == Params(nprobe=32, cand=500, rerank=Rerank.FULL, k=100)
    _, p = parse_config("64:500:tiered200:b10x0.8:p0.12:s200k", 4096)
    assert p.rerank == ("tiered", 200) and p.prune == (10, 0.8) and p.margin == 0.12 and p.budget == 200000
    _, p = parse_config("all:100:ssd50:b5", 77)
    assert p.nprobe == 77 and p.rerank == ("ssd", 50) and p.prune == (5, 1.0)
    for bad in ("32:500", "32:500:fancy", "32:500:full:q3", "x:500:full"):
        with pytest.raises((ConfigError, ValueError)):
            parse_config(bad, 10)


def test_make_params_tags():
    c = abi.make_params(Params(nprobe=8, cand=100, rerank=Rerank.tiered(20), k=10, prune=(10, 0.5)))
    assert (c.rerank, c.rerank_r, c.prune_m, c.prune_w, c.nprobe, c.budget) == (3, 20, 10, 0.5, 8, 0)
    assert math.isnan(c.margin)


def test_stats_mirror_matches_header_field_order():
    """sieve_stats grows by appending fields; the ctypes mirror must list the header's fields in
    the header's order (every field is 8 bytes)."""
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "sieve_gpu.h")).read()
    body = re.search(r"typedef struct sieve_stats \{(.*?)\} sieve_stats;", hdr, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"(?:uint64_t|double)\s+(\w+);", body)
    assert names == [f for f, _ in abi.Stats._fields_]
    assert ctypes.sizeof(abi.Stats) == 8 * len(names)
