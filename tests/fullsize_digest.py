"""SHA-256 digests of a FeatureComplex too large to compare array by array: node
vertices (uint64, m+1 per node, ascending ids), lambda (float64 bits, m+1 per node),
track ids (uint64) and connectivity (uint32 index tuples), each hashed as one
little-endian byte stream read in ranges.  The full-size parity test hashes the GPU
result this way and compares with the digest the oracle's full-size mode produced
(tests/golden/make_cp3d_fullsize_digest.py)."""
import hashlib

import numpy as np

CHUNK = 1 << 23


def field_digest(comps):
    """One SHA-256 per component over its values in x-fastest order."""
    out = []
    for c in comps:
        h = hashlib.sha256()
        flat = np.asarray(c).reshape(-1, order="F")
        for i in range(0, flat.size, CHUNK * 4):
            h.update(np.ascontiguousarray(flat[i:i + CHUNK * 4]).tobytes())
        out.append(h.hexdigest())
    return out


def complex_digest(n_nodes, n_edges, read_nodes, read_edges):
    """read_nodes(begin, count) -> (verts u64 (count, m+1), lam f64 (count, m+1),
    track u64 (count,)); read_edges(begin, count) -> u32 (count, 2)."""
    hv, hl, ht, he = (hashlib.sha256() for _ in range(4))
    for b in range(0, n_nodes, CHUNK):
        v, lam, tr = read_nodes(b, min(CHUNK, n_nodes - b))
        hv.update(np.ascontiguousarray(v, dtype="<u8").tobytes())
        hl.update(np.ascontiguousarray(lam, dtype="<f8").tobytes())
        ht.update(np.ascontiguousarray(tr, dtype="<u8").tobytes())
    for b in range(0, n_edges, CHUNK):
        he.update(np.ascontiguousarray(read_edges(b, min(CHUNK, n_edges - b)), dtype="<u4").tobytes())
    return {"n_nodes": int(n_nodes), "n_edges": int(n_edges), "vertices": hv.hexdigest(),
            "lambda": hl.hexdigest(), "track": ht.hexdigest(), "edges": he.hexdigest()}
