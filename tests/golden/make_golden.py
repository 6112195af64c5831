"""Writes the golden fixtures under tests/golden/ (run from the repo root).

  selfjoin_r{1,2,5}.npz   self-join of test.idradec (the reference's sample data,
                          BoostXmatch Data/test.idradec, 10,000 SDSS DR7 objects) at
                          radius = zone height = 1, 2, 5 arcsec: sorted (idA, idB,
                          dist_arcsec).  The reference's debugger command lines
                          self-join exactly this file (Debug Switches.txt:2-4).
  small_patch_seed1.json  count, SHA-256 of the sorted (idA, idB) pairs and work
                          counts of the 1e5 x 1e5, 1 arcsec patch workload, seed 1.

Each set is produced by the oracle and accepted only if an independent all-pairs
evaluation (numpy sin/cos unit vectors, no zones, no windows) gives the same pairs
and distances within 1e-6 arcsec.  The reference itself ships no expected outputs.
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from xmatch_amd import synth  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def numpy_allpairs(a, b, radius_deg, h_deg, block=2048):
    """Independent check: numpy unit vectors, chord test on every pair within nz zones."""
    def uv(o):
        r, d = np.radians(o["ra"]), np.radians(o["dec"])
        return np.stack([np.cos(d) * np.cos(r), np.cos(d) * np.sin(r), np.sin(d)], 1)
    ua, ub = uv(a), uv(b)
    nzn = int(np.ceil(180.0 / h_deg))
    za = np.clip(np.floor((a["dec"] + 90.0) / h_deg), 0, nzn - 1).astype(np.int64)
    zb = np.clip(np.floor((b["dec"] + 90.0) / h_deg), 0, nzn - 1).astype(np.int64)
    nz = int(np.ceil(radius_deg / h_deg))
    thr = (2.0 * np.sin(np.radians(radius_deg) / 2.0)) ** 2
    out_a, out_b, out_d = [], [], []
    for s in range(0, len(a), block):
        d2 = ((ua[s:s + block, None, :] - ub[None, :, :]) ** 2).sum(-1)
        ok = (d2 < thr) & (np.abs(za[s:s + block, None] - zb[None, :]) <= nz)
        i, j = np.nonzero(ok)
        out_a.append(a["id"][s + i])
        out_b.append(b["id"][j])
        out_d.append(np.degrees(2.0 * np.arcsin(np.sqrt(d2[i, j]) / 2.0)) * 3600.0)
    ida, idb, dist = (np.concatenate(x) for x in (out_a, out_b, out_d))
    o = np.lexsort((idb, ida))
    return ida[o], idb[o], dist[o]


def pair_digest(ida, idb):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(ida, dtype="<i8").tobytes())
    h.update(np.ascontiguousarray(idb, dtype="<i8").tobytes())
    return h.hexdigest()


def main():
    a = synth.read_idradec(os.path.join(HERE, "test.idradec"))
    for r in (1, 2, 5):
        rd = r / 3600.0
        (ia, ib, d), cnt = O.match(a, a, rd, mode=0)
        na, nb, nd = numpy_allpairs(a, a, rd, rd)
        assert np.array_equal(ia, na) and np.array_equal(ib, nb), "oracle != all-pairs"
        assert np.abs(d - nd).max() <= 1e-6
        np.savez_compressed(os.path.join(HERE, "selfjoin_r%d.npz" % r), ida=ia, idb=ib,
                            dist_arcsec=d, w_ref=cnt["w_ref"])
        print("self-join r=%d\": %d pairs, w_ref %d" % (r, len(ia), cnt["w_ref"]))

    A, B, rd = synth.config_small_patch(seed=1)
    (ia, ib, d), cnt = O.match(A, B, rd, mode=0)
    # independent check on a deterministic 20% slice of A
    sl = A[A["ra"] < 2.0]
    na, nb, nd = numpy_allpairs(sl, B, rd, rd)
    keep = np.isin(ia, sl["id"])
    assert np.array_equal(ia[keep], na) and np.array_equal(ib[keep], nb)
    assert np.abs(d[keep] - nd).max() <= 1e-6
    summary = {"seed": 1, "n_a": int(len(A)), "n_b": int(len(B)), "radius_arcsec": 1.0,
               "count": int(len(ia)), "sha256_pairs": pair_digest(ia, ib),
               "w_ref": int(cnt["w_ref"]), "w_cand": int(cnt["w_cand"]),
               "first": [[int(ia[k]), int(ib[k]), float(d[k])] for k in range(5)],
               "last": [[int(ia[k]), int(ib[k]), float(d[k])] for k in range(-5, 0)]}
    with open(os.path.join(HERE, "small_patch_seed1.json"), "w") as f:
        json.dump(summary, f, indent=1)
    print("small patch seed 1: %d pairs" % len(ia))


if __name__ == "__main__":
    main()
