"""Writes tests/golden/legacy_loads.json: what the reference's own loader (bion_load_host,
src/bion_loader_host.cpp:26-425, compiled from the reference sources into oracle/_ref) makes of
every seeded BION file of tests/bion_legacy.py - versions 3-14, the v10+ record-size variants,
and the files it must refuse. Per file: its return code and, for files that load, the record
counts, the 100-byte header and the SHA-256 of each record array (neurons, segments,
connections, synapses, nuclei) with the bytes the loader leaves unwritten set to zero
(bion_legacy.defined). The files are rebuilt from their seeds wherever the fixture is checked.

Run where oracle/_ref is built:  python tests/golden/gen_legacy.py
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import bion_legacy as L  # noqa: E402
import harness as H  # noqa: E402
from test_bion_versions import _reference  # noqa: E402

OUT = os.path.join(HERE, "legacy_loads.json")
SIZES = (24, 3216, 16, 1568, 500)


def digests(blob, n, version):
    """SHA-256 of each record array of a loaded BionData (concatenated bytes), unwritten bytes zeroed."""
    cuts = np.cumsum([0] + [c * s for c, s in zip(n, SIZES)])
    parts = [np.frombuffer(blob, np.uint8)[cuts[i]:cuts[i + 1]] for i in range(5)]
    return [hashlib.sha256(np.where(m, a, 0).astype(np.uint8).tobytes()).hexdigest()
            for a, m in zip(parts, L.defined(version, parts))]


def main():
    if not H.ref_available():
        raise SystemExit("oracle/_ref is not built (the compiled reference)")
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        cases = L.cases(tmp)
        errs = L.error_cases(tmp)
        ref = _reference([p for p, _ in cases.values()] + list(errs.values()))
        for name, (p, v) in cases.items():
            r = ref[p]
            rec = {"version": v, "rc": r["rc"]}
            if r["rc"] == 0:
                blob = np.load(p + ".ref.npy").tobytes()
                rec.update(n=r["n"], header=bytes(r["header"]).hex(), sha256=digests(blob, r["n"], v))
            out[name] = rec
        for name, p in errs.items():
            out["error_" + name] = {"rc": ref[p]["rc"]}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {OUT}: {len(out)} files")


if __name__ == "__main__":
    main()
