"""Shared test helpers (random workloads, comparisons)."""
import numpy as np

import xtn

AA20 = "ACDEFGHIKLMNPQRSTVWY"


def random_seqs(rng, n, lmin=1, lmax=30, alphabet=AA20, unique=False):
    out, seen = [], set()
    while len(out) < n:
        L = int(rng.integers(lmin, lmax + 1))
        s = "".join(alphabet[i] for i in rng.integers(0, len(alphabet), L))
        if unique and s in seen:
            continue
        seen.add(s)
        out.append(s)
    return out


def mutate(rng, s, k, alphabet=AA20):
    s = list(s)
    for _ in range(k):
        t = int(rng.integers(0, 3))
        if t == 0 and s:
            s[int(rng.integers(0, len(s)))] = alphabet[int(rng.integers(0, len(alphabet)))]
        elif t == 1 or not s:
            s.insert(int(rng.integers(0, len(s) + 1)), alphabet[int(rng.integers(0, len(alphabet)))])
        else:
            del s[int(rng.integers(0, len(s)))]
    return "".join(s) if s else alphabet[0]


def clustered_seqs(rng, n, lmin=3, lmax=26, alphabet=AA20, edits=2):
    """sequences with many near relatives (and some exact duplicates)"""
    base = random_seqs(rng, max(1, n // 4), lmin, lmax, alphabet)
    out = list(base)
    while len(out) < n:
        s = out[int(rng.integers(0, len(out)))]
        out.append(mutate(rng, s, int(rng.integers(0, edits + 1)), alphabet))
    return out


def triples(pairs, dist):
    return np.concatenate([np.asarray(pairs, np.int64).reshape(-1, 2),
                           np.asarray(dist, np.int64).reshape(-1, 1)], axis=1)
