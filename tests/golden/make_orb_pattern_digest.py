"""Writes orb_pattern_digest.json: the count, sum, sum of squares and SHA-256
(of the comma-joined integers) of the rotated-BRIEF pair table 'ref_pat' in
the reference's src/backend/cpu/kernel/orb.hpp:24-94, read as text. The test
compares the library's table (arrayfire_amd/csrc/kernels/orb_pattern.inc)
against these numbers.

    python tests/golden/make_orb_pattern_digest.py /root/reference
"""
import hashlib
import json
import os
import re
import sys


def main(ref_root: str) -> None:
    src = open(os.path.join(ref_root, "src/backend/cpu/kernel/orb.hpp")).read()
    body = src[src.index("ref_pat[REF_PAT_LENGTH]"):]
    body = body[body.index("{") + 1: body.index("};")]
    v = [int(t) for t in re.findall(r"-?\d+", body)]
    out = {"count": len(v), "sum": sum(v), "sum_sq": sum(x * x for x in v),
           "sha256": hashlib.sha256(",".join(map(str, v)).encode()).hexdigest(),
           "first_pairs": [v[0:4], v[4:8]], "last_pair": v[-4:]}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "orb_pattern_digest.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
