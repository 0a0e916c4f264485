"""Simulation checkpoints: one .npz holding the state and everything needed to resume it.

No reference counterpart (the reference has no on-disk state format).  Arrays: ``pos`` (M, 2)
float32, ``f`` (M, 7) float32, ``counts`` (instances,) int64, ``iteration`` int64, ``params`` a JSON
string (one GpuSimulator keyword dict per instance), ``lanes`` (2,) int32 the refine and gather
lanes per point and ``depth`` (2,) int32 their candidates in flight per lane (the tiles set the
order of fp32 partial sums, so a tuned run resumes bit for bit only with its own tiles), ``format``
the format version.  Written with numpy.savez and read with allow_pickle=False: the file holds no
executable content.  Format 1 files (no lanes) and format 2 files (no depths) load with the
defaults.
"""

from __future__ import annotations

import json

import numpy as np

FORMAT = 3
DEFAULT_LANES = (4, 4)
DEFAULT_DEPTH = (4, 4)


def save(path, pos, f, iteration, kws, counts, lanes=DEFAULT_LANES, depth=DEFAULT_DEPTH):
    pos = np.ascontiguousarray(np.asarray(pos, np.float32).reshape(-1, 2))
    f = np.ascontiguousarray(np.asarray(f, np.float32).reshape(-1, 7))
    counts = np.asarray(counts, np.int64)
    if len(pos) != len(f) or counts.sum() != len(pos) or len(counts) != len(kws):
        raise ValueError("inconsistent checkpoint arrays")
    with open(path, "wb") as fh:
        np.savez(fh, format=np.int64(FORMAT), pos=pos, f=f, counts=counts, iteration=np.int64(iteration),
                 params=np.array(json.dumps(list(kws))), lanes=np.asarray(lanes, np.int32).reshape(2),
                 depth=np.asarray(depth, np.int32).reshape(2))


def load(path):
    """(pos, f, iteration, [kw per instance], counts, (refine lanes, gather lanes),
    (refine depth, gather depth))."""
    with np.load(path, allow_pickle=False) as z:
        if int(z["format"]) not in (1, 2, FORMAT):
            raise ValueError(f"unsupported checkpoint format {int(z['format'])}")
        lanes = tuple(int(v) for v in z["lanes"]) if "lanes" in z.files else DEFAULT_LANES
        depth = tuple(int(v) for v in z["depth"]) if "depth" in z.files else DEFAULT_DEPTH
        kws = json.loads(str(z["params"]))
        for kw in kws:
            if kw.get("two_level") is not None:
                kw["two_level"] = tuple(kw["two_level"])
        return z["pos"], z["f"], int(z["iteration"]), kws, z["counts"], lanes, depth
