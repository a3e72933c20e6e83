"""Extract the field-level golden vectors the reference holds as data.

Reads crates/flock-core/src/field/phi8.rs (the PHI_8_TABLE literal values,
phi8.rs:19-1045) and the protocol-fixed F8 constants in
crates/flock-core/src/zerocheck/univariate_skip_optimized.rs:103-106, and
writes them as a JSON fixture.  Run in a checkout that has the reference
mounted at /root/reference; the fixture it writes is committed.
"""
import json
import os
import re
import sys

REF = os.environ.get("FLOCK_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "field_golden.json")


def main():
    src = open(os.path.join(REF, "crates/flock-core/src/field/phi8.rs")).read()
    body = src[src.index("PHI_8_TABLE"):src.index("];")]
    pairs = re.findall(r"lo:\s*0x([0-9a-fA-F]+),\s*hi:\s*0x([0-9a-fA-F]+)", body)
    assert len(pairs) == 256, len(pairs)
    table = [[lo.lower(), hi.lower()] for lo, hi in pairs]
    out = {"phi8_table": table}
    usk = open(os.path.join(REF, "crates/flock-core/src/zerocheck/univariate_skip_optimized.rs")).read()
    m = re.search(r"SMALL_CHAL_F8[^=]*=\s*\[([^\]]*)\]", usk)
    if m:
        out["small_chal_f8"] = [int(x, 0) for x in re.findall(r"0x[0-9a-fA-F]+", m.group(1))]
    m = re.search(r"C_S_F8[^=]*=\s*[^(]*\(\s*(0x[0-9a-fA-F]+)", usk)
    if m:
        out["c_s_f8"] = int(m.group(1), 0)
    json.dump(out, open(OUT, "w"), indent=0)
    print("wrote", OUT, {k: (len(v) if isinstance(v, list) else v) for k, v in out.items()})


if __name__ == "__main__":
    sys.exit(main())
