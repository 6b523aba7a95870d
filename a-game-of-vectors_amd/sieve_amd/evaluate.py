"""Recall and latency percentiles, as the reference's harness defines them.

recall_at:      eval::recall_at (src/eval.rs:7-10)
recall10_ties:  tie-aware R@10 of the 1B harness (src/bin/spacev.rs:110-114)
pct:            sorted-index percentile (src/bin/search.rs:400-402)
"""
import numpy as np


def recall_at(k, got, truth):
    """|got[:k] ∩ truth[:k]| / min(k, |truth[:k]|)."""
    t = set(list(truth)[:k])
    hit = sum(1 for x in list(got)[:k] if x in t)
    return hit / max(min(k, len(t)), 1)


def mean_recall(k, got_rows, truth_rows):
    return float(np.mean([recall_at(k, g, t) for g, t in zip(got_rows, truth_rows)]))


def recall10_ties(got, truth_ids, truth_scores):
    """Counts a returned id as a hit when its exact score ties the 10th true neighbour's."""
    kth = truth_scores[9]
    t = {int(i) for i, s in zip(truth_ids, truth_scores) if s >= kth}
    return sum(1 for x in list(got)[:10] if int(x) in t) / 10.0


def pct(sorted_vals, p):
    """sorted[round((n - 1) * p)] (src/bin/search.rs:400-402)."""
    n = len(sorted_vals)
    return float(sorted_vals[int(round((n - 1) * p))])
