"""Query parameters: the reference's ivf::Params / Rerank and its config mini-language.

Params mirrors `ivf::Params` (src/ivf.rs:441-460) field for field; Rerank mirrors
`ivf::Rerank` (src/ivf.rs:425-439); `parse_config` accepts the `search` CLI's config strings
`nprobe:cand:rerank[:bM[xW]][:pD][:sN[k]]` (src/bin/search.rs:180-218) with the same errors.
"""
from dataclasses import dataclass, replace
from typing import Optional, Tuple

RERANK_TAGS = {"none": 0, "prefix": 1, "full": 2, "tiered": 3, "ssd": 4}


class Rerank:
    """Rerank modes (src/ivf.rs:425-439) as (tag, r) tuples, tags as in shard::Header."""
    NONE = ("none", 0)
    PREFIX = ("prefix", 0)
    FULL = ("full", 0)

    @staticmethod
    def tiered(r: int):
        return ("tiered", int(r))

    @staticmethod
    def ssd(r: int):
        return ("ssd", int(r))


@dataclass(frozen=True)
class Params:
    nprobe: int
    cand: int
    rerank: Tuple[str, int] = Rerank.FULL
    k: int = 10
    prune: Optional[Tuple[int, float]] = None
    margin: Optional[float] = None
    budget: Optional[int] = None

    def with_(self, **kw):
        return replace(self, **kw)


class ConfigError(ValueError):
    pass


def parse_config(s: str, nclusters: int, k: int = 100):
    """`nprobe:cand:rerank[:bM[xW]][:pD][:sN[k]]` -> (label, Params), as src/bin/search.rs:180-218.

    The reference fixes k at 100 (src/bin/search.rs:217); `k` defaults to that."""
    f = s.split(":")
    if len(f) < 3:
        raise ConfigError(f"config {s}: expected nprobe:cand:rerank[:bM[xW]][:pD][:sN[k]]")
    np_, cand, rr, opts = f[0], f[1], f[2], f[3:]
    prune = margin = budget = None
    for o in opts:
        head, v = o[:1], o[1:]
        try:
            if head == "b":
                m, _, w = v.partition("x")
                prune = (int(m), float(w) if w else 1.0)
            elif head == "p":
                margin = float(v)
            elif head == "s":
                mul = 1000 if v.endswith("k") else 1
                budget = int(v[:-1] if v.endswith("k") else v) * mul
            else:
                raise ConfigError(f"config option {o}: expected bM[xW], pD or sN[k]")
        except ValueError as e:
            if isinstance(e, ConfigError):
                raise
            raise ConfigError(f"config option {o}: {e}") from None
    nprobe = nclusters if np_ == "all" else int(np_)
    if rr == "none":
        rerank = Rerank.NONE
    elif rr == "prefix":
        rerank = Rerank.PREFIX
    elif rr == "full":
        rerank = Rerank.FULL
    elif rr.startswith("tiered"):
        rerank = Rerank.tiered(int(rr[len("tiered"):]))
    elif rr.startswith("ssd"):
        rerank = Rerank.ssd(int(rr[len("ssd"):]))
    else:
        raise ConfigError("rerank must be none, prefix, full, tieredN or ssdN")
    label = f"ivf {np_} probes, {cand} cand, rerank {rr}"
    if prune is not None:
        label += f", bound-pruned at rank {prune[0]}, width {prune[1]:g}"
    if margin is not None:
        label += f", probe margin {margin:g}"
    if budget is not None:
        label += f", scan budget {budget}"
    return label, Params(nprobe=nprobe, cand=int(cand), rerank=rerank, k=k, prune=prune, margin=margin,
                         budget=budget)
