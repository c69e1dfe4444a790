"""Builds tests/golden/result_subset.fasta from the reference's shipped result.fasta.

The reference ships result.fasta (1,378 representatives written by its own run at the
default t=0.95; MPL-2.0, JuZhenCS/gene-sequences-de-redundancy). Every later
representative was unclustered while each earlier one was the representative, so
P(earlier, later) was false for every pair, and re-clustering any subset at t=0.95 must
return that subset unchanged (all records are representatives).

This script keeps the records that take part in the pairs closest to acceptance: pairs
that pass the composition and shared-word filters (so they reach the alignment), ranked
by LCS - ceil(L*t). It runs in this container only (it reads the reference's data file);
the fixture it writes travels with the repository.

    python tests/golden/make_result_subset.py [/root/reference/result.fasta]
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_ffi as O  # noqa: E402

T = 0.95
K = 6
TARGET = 200


def read(path):
    recs = O.read_fasta(path)
    return [(n, d) for n, d in recs]


def kmer_matrix(seqs):
    M = np.zeros((len(seqs), 4 ** K), dtype=np.int32)
    for i, s in enumerate(seqs):
        c = O.codes(s).astype(np.int64)
        valid = c < 4
        word = np.zeros(len(c), dtype=np.int64)
        ok = np.zeros(len(c), dtype=bool)
        for j in range(K - 1, len(c)):
            win = c[j - K + 1:j + 1]
            if valid[j - K + 1:j + 1].all():
                word[j] = int(sum(int(win[K - 1 - t]) << (2 * t) for t in range(K)))
                ok[j] = True
        np.add.at(M[i], word[ok], 1)
    return M


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/result.fasta"
    recs = read(src)
    seqs = [d for _, d in recs]
    n = len(seqs)
    L = np.array([len(s) for s in seqs])
    t32 = np.float32(T)
    bc = np.array([math.ceil(float(np.float32(x) * t32)) for x in L])
    om = np.float32(1.0) - t32
    wc = np.array([max(int(x) - K + 1 - math.ceil(float(np.float32(x) * om)) * K, 1) for x in L])
    comp = np.array([np.bincount(O.codes(s), minlength=5)[:4] for s in seqs])
    M = kmer_matrix(seqs)
    pairs = []
    order = np.argsort(-L, kind="stable")
    for a in range(n):
        r = order[a]
        for b in range(a + 1, n):
            q = order[b]
            if L[q] > L[r]:
                r, q = q, r
            if np.minimum(comp[r], comp[q]).sum() < bc[q]:
                continue
            if np.minimum(M[r], M[q]).sum() < wc[q]:
                continue
            lcs = O.lcs(seqs[r], seqs[q])
            pairs.append((int(lcs - bc[q]), int(r), int(q)))
            r, q = order[a], order[b]
    pairs.sort(reverse=True)
    keep = []
    for margin, r, q in pairs:
        for x in (r, q):
            if x not in keep:
                keep.append(x)
        if len(keep) >= TARGET:
            break
    keep.sort()
    out = os.path.join(HERE, "result_subset.fasta")
    with open(out, "wb") as f:
        for i in keep:
            f.write(recs[i][0] + b"\n" + recs[i][1] + b"\n")
    meta = {"source": "result.fasta of JuZhenCS/gene-sequences-de-redundancy (MPL-2.0)",
            "threshold": T, "records": len(keep), "aligned_pairs_in_source": len(pairs),
            "closest_margins": [p[0] for p in pairs[:10]],
            "expect": "every record is a representative (idempotence of the greedy output)"}
    with open(os.path.join(HERE, "result_subset.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
