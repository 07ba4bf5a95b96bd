"""Write tests/golden/reference_input_sha256.json: sha256 and row counts of the
dam-break and cavity input files shipped with the reference (input_file/input_*.txt).
Run in a checkout where the reference is available at REFERENCE_ROOT."""
import hashlib
import json
import os
from pathlib import Path

ROOT = Path(os.environ.get("REFERENCE_ROOT", "/root/reference"))
out = {}
for kind in ("SPH", "BND", "DEM", "DEMBND", "EISPH", "EISPHBND"):
    data = (ROOT / "input_file" / f"input_{kind}.txt").read_bytes()
    out[kind] = {"sha256": hashlib.sha256(data).hexdigest(), "rows": data.count(b"\n") - 1,
                 "bytes": len(data)}
Path(__file__).with_name("reference_input_sha256.json").write_text(json.dumps(out, indent=1) + "\n")
print(out)
