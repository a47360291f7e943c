"""The generator's coarse pass writes no record for a deletion set that
repeats another set's variant of the same sequence: a set is dropped when one
of its positions q continues a run of equal residues (residue q - 1 equals
residue q) while q - 1 is not deleted. This checks, by brute force over every
deletion set of size <= d on short strings with long runs, that the sets kept
by that bit test (run mask `runs`, deletion mask `dm`,
dropped iff dm & runs & ~(dm << 1)) still give every distinct variant, so the
candidate pairs (pairs of ids sharing a variant) are unchanged."""
import itertools
import random

import pytest

MAX_PREFIX = 18


def run_mask(s):
    return sum(1 << q for q in range(1, min(len(s), MAX_PREFIX)) if s[q - 1] == s[q])


def dropped(dset, runs):
    dm = sum(1 << q for q in dset)
    return (dm & runs & ~(dm << 1)) != 0


def variant_sets(s, d, drop):
    s = s[:MAX_PREFIX]
    runs = run_mask(s)
    out = []
    for k in range(d + 1):
        for dset in itertools.combinations(range(len(s)), k):
            if drop and dropped(dset, runs):
                continue
            out.append("".join(c for i, c in enumerate(s) if i not in dset))
    return out


@pytest.mark.parametrize("alphabet", ["AB", "ABC", "ACDEFGHIKLMNPQRSTVWY"])
def test_kept_sets_give_every_variant(alphabet):
    rng = random.Random(len(alphabet))
    for _ in range(400):
        s = "".join(rng.choice(alphabet) for _ in range(rng.randint(0, 13)))
        for d in (1, 2, 3):
            full = variant_sets(s, d, False)
            kept = variant_sets(s, d, True)
            assert set(kept) == set(full), (s, d)
            assert len(kept) <= len(full)


def test_runs_drop_repeats():
    # "CASSLGG": deleting either S, or either G, gives the same string
    full = variant_sets("CASSLGG", 1, False)
    kept = variant_sets("CASSLGG", 1, True)
    assert len(full) == 8 and len(set(full)) == 6 and len(kept) == 6
    # past the 18-residue prefix nothing counts
    assert run_mask("A" * 25) == sum(1 << q for q in range(1, MAX_PREFIX))
