"""Writes random-forest classifier files in CARE's .rf format for tests.

Format (include/forest.hpp:38-55, writer ml/care.py:198-282): u64 length +
descriptor string, u32 tree count, then per tree u32 internal-node count
followed by the nodes in preorder: feature u8, split f64, flag u8 (bit 1:
left child is a leaf, bit 0: right child is a leaf), children in order,
leaves as f32 probabilities of accepting the correction.
"""
import random
import struct

ANCHOR_DESC = "13 extract_anchor_v3"
CANDS_DESC = "14 extract_cands_v3"

# plausible value ranges of the v3 features (ratios, supports, coverage
# normalised by the estimated coverage)
RANGES = [(0, 1), (0, 1), (0, 1), (0, 1), (0.4, 1), (0, 1), (0, 1), (0, 2), (0, 2),
          (0.8, 1), (0.4, 1), (0, 2), (0, 1.5), (0, 1)]


def _tree(rng, depth, nfeat, out):
    """Emits one subtree; returns the number of internal nodes written."""
    att = rng.randrange(nfeat)
    lo, hi = RANGES[att]
    split = rng.uniform(lo, hi)
    left_leaf = depth <= 1 or rng.random() < 0.3
    right_leaf = depth <= 1 or rng.random() < 0.3
    flag = (2 if left_leaf else 0) + (1 if right_leaf else 0)
    out += struct.pack("<B", att) + struct.pack("<d", split) + struct.pack("<B", flag)
    n = 1
    for leaf in (left_leaf, right_leaf):
        if leaf:
            out += struct.pack("<f", rng.random())
        else:
            n += _tree(rng, depth - 1, nfeat, out)
    return n


def write_forest(path, desc, seed, trees=12, depth=5):
    rng = random.Random(seed)
    nfeat = int(desc.split()[0])
    body = bytearray()
    for _ in range(trees):
        t = bytearray()
        n = _tree(rng, depth, nfeat, t)
        body += struct.pack("<I", n) + t
    with open(path, "wb") as f:
        d = desc.encode()
        f.write(struct.pack("<Q", len(d)) + d + struct.pack("<I", trees) + bytes(body))
