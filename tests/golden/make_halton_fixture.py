"""Writes halton_fixture.json: the head of the Halton digit permutations and a
few radical inverses, computed here from their definitions.

Permutations: for each of the first 1024 primes p in order, start from the
identity 0..p-1 and for i = p-1 down to 1 swap entries i and r % (i+1), where
r is the next output of the Microsoft C runtime rand() (state = state *
214013 + 2531011 mod 2^32, output (state >> 16) & 0x7fff) seeded with 533799
(Samplers/HaltonSampler.cu:7-12,46-51). Radical inverse of index i in base b
with permutation P: sum over digits d_k of i (least significant first) of
P[d_k] / b^(k+1), accumulated as float r += f * P[d], f = f / b.
Run: python tests/golden/make_halton_fixture.py
"""
import json
import os

import numpy as np


def primes(n):
    out = []
    c = 2
    while len(out) < n:
        if all(c % p for p in out if p * p <= c):
            out.append(c)
        c += 1
    return out


def perms(ps, keep):
    state = 533799
    out = {}
    for p in ps:
        v = list(range(p))
        for i in range(p - 1, 0, -1):
            state = (state * 214013 + 2531011) & 0xFFFFFFFF
            r = (state >> 16) & 0x7FFF
            j = r % (i + 1)
            v[i], v[j] = v[j], v[i]
        if p in keep:
            out[p] = v
    return out


def radical_inverse(index, base, perm):
    r = np.float32(0)
    f = np.float32(1)
    b = np.float32(base)
    while index > 0:
        f = np.float32(f / b)
        r = np.float32(r + np.float32(f * np.float32(perm[index % base])))
        index //= base
    return float(r)


if __name__ == "__main__":
    ps = primes(1024)
    keep = {2, 3, 5, 7, 11, 13}
    pm = perms(ps, keep)
    dims = {2: 0, 3: 1, 5: 2, 7: 3, 11: 4, 13: 5}
    cases = []
    for idx in (0, 1, 2, 3, 7, 100, 12345, 999999, 4194303):
        for base in (2, 3, 5, 13):
            cases.append({"index": idx, "dim": dims[base], "value": radical_inverse(idx, base, pm[base])})
    out = {"prime_1023": ps[-1], "prime_sum": sum(ps),
           "perm_head": {str(p): pm[p] for p in sorted(keep)}, "radical_inverse": cases}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "halton_fixture.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print("wrote", path)
