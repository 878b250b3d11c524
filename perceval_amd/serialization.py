"""BSSamples text wire format (perceval/serialization/_state_serialization.py:68-92,
perceval/serialization/serialize.py:199-205, deserialize.py:235-265).

A sample list is written as the distinct states in order of first appearance,
then the index of each sample's state:

    "|0,1>;|1,0>/0;1;0;0"          (samples |0,1>, |1,0>, |0,1>, |0,1>)

`serialize` wraps it as ":PCVL:BSSamples:<body>" and, by default, compresses
it (zlib, then base64) behind ":PCVL:zip:", as the reference's generic
serialize does for BSSamples.  The samplers return occupation arrays; the
array forms below number the distinct rows with one vectorised pass instead
of building a FockState per sample.
"""
from __future__ import annotations

import base64
import zlib

import numpy as np

from .utils.states import BSSamples, FockState

SEP = ":"
PCVL_PREFIX = f"{SEP}PCVL{SEP}"
ZIP_PREFIX = f"{PCVL_PREFIX}zip{SEP}"
BSS_TAG = "BSSamples"


def serialize_bssamples(samples) -> str:
    """Body of the wire format for a BSSamples list (or any list of states)."""
    index: dict = {}
    order = []
    for s in samples:
        s = FockState(s)
        k = index.get(s)
        if k is None:
            k = index[s] = len(index)
        order.append(k)
    return ";".join(str(s) for s in index) + "/" + ";".join(str(i) for i in order)


def serialize_bssamples_array(occ) -> str:
    """Same body from a (count, m) occupation array (e.g. a sampler's output),
    numbering the distinct rows by first appearance in one vectorised pass."""
    occ = np.ascontiguousarray(np.asarray(occ))
    if occ.ndim != 2:
        raise ValueError("occupation array must be 2-D (samples x modes)")
    if occ.shape[0] == 0:
        return "/"
    uniq, first, inverse = np.unique(occ, axis=0, return_index=True, return_inverse=True)
    rank = np.empty(len(uniq), dtype=np.int64)          # unique row -> first-appearance number
    order = np.argsort(first, kind="stable")
    rank[order] = np.arange(len(uniq))
    states = ";".join("|" + ",".join(map(str, uniq[u].tolist())) + ">" for u in order)
    return states + "/" + ";".join(map(str, rank[inverse.ravel()].tolist()))


def deserialize_bssamples(body: str) -> BSSamples:
    parts = body.split("/")
    assert len(parts) == 2, f"Bad serialized BSSamples: {body}"
    if not parts[0]:
        return BSSamples()
    states = [FockState(s) for s in parts[0].split(";")]
    return BSSamples(states[int(i)] for i in parts[1].split(";"))


def deserialize_bssamples_array(body: str) -> np.ndarray:
    """(count, m) int32 occupation array of a serialized sample list."""
    parts = body.split("/")
    assert len(parts) == 2, f"Bad serialized BSSamples: {body}"
    if not parts[0]:
        return np.zeros((0, 0), dtype=np.int32)
    table = np.array([FockState(s).tolist() for s in parts[0].split(";")], dtype=np.int32)
    idx = np.array(parts[1].split(";"), dtype=np.int64)
    return table[idx]


def _compress(text: str) -> str:
    return ZIP_PREFIX + base64.b64encode(zlib.compress(text.encode("utf-8"))).decode("utf-8")


def serialize(samples, compress: bool = True) -> str:
    """":PCVL:BSSamples:<body>", zlib + base64 behind ":PCVL:zip:" unless
    compress is False.  Accepts a BSSamples list or an occupation array."""
    if isinstance(samples, np.ndarray):
        body = serialize_bssamples_array(samples)
    else:
        body = serialize_bssamples(samples)
    text = f"{PCVL_PREFIX}{BSS_TAG}{SEP}" + body
    return _compress(text) if compress else text


def deserialize(text: str) -> BSSamples:
    """Inverse of serialize (compressed or not)."""
    if not text.startswith(PCVL_PREFIX):
        raise ValueError("not a Perceval serialization")
    if text.startswith(ZIP_PREFIX):
        text = zlib.decompress(base64.b64decode(text[len(ZIP_PREFIX):])).decode("utf-8")
    lp = len(PCVL_PREFIX)
    p = text[lp:].find(SEP)
    tag, body = text[lp:lp + p], text[lp + p + 1:]
    if tag != BSS_TAG:
        raise NotImplementedError(f"No serializer found for {tag}")
    return deserialize_bssamples(body)
