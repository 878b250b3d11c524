"""Distribution distances (perceval/utils/dist_metrics.py:34-52)."""
from __future__ import annotations


def tvd_dist(dist_lh: dict, dist_rh: dict) -> float:
    """Total variation distance 0.5 * sum |p - q| over the union of the keys;
    a key missing on one side counts as probability 0."""
    keys = set(dist_lh.keys()) | set(dist_rh.keys())
    return 0.5 * sum(abs(dist_lh.get(k, 0.0) - dist_rh.get(k, 0.0)) for k in keys)
