"""Write the reference's golden end-to-end checksums as a JSON fixture.

Reads the checksum tables the reference's own tests hold
(tests/constants/checksums/<d>d<Type>.h under ZFP_REF_SRC, default
/root/reference) and stores their (key1, key2, checksum) rows as data in
tests/golden/ref_checksums.json.  The rows are keyed as in
tests/utils/zfpChecksums.c:75-136 of the reference:

  key1 = ((test_type * 4 + subject) * 8 + mode) * 16 + param
  key2 = array extents packed like zfp_field_metadata, n0 most significant

Run once in a container that has the reference tree; the JSON is committed.
"""
import json
import os
import re
import sys

REF = os.environ.get("ZFP_REF_SRC", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    out = {}
    pat = re.compile(r"\{UINT64C\((0x[0-9a-fA-F]+)\),\s*UINT64C\((0x[0-9a-fA-F]+)\),\s*UINT64C\((0x[0-9a-fA-F]+)\)\}")
    for dims in (1, 2, 3, 4):
        for tname in ("Int32", "Int64", "Float", "Double"):
            path = os.path.join(REF, "tests", "constants", "checksums", f"{dims}d{tname}.h")
            rows = [[int(a, 16), int(b, 16), int(c, 16)] for a, b, c in pat.findall(open(path).read())]
            out[f"{dims}d{tname}"] = rows
    dst = os.path.join(HERE, "ref_checksums.json")
    with open(dst, "w") as f:
        json.dump({"source": "LLNL zfp 1.0.1 tests/constants/checksums", "tables": out}, f, indent=0)
    print(dst, sum(len(v) for v in out.values()), "rows")


if __name__ == "__main__":
    sys.exit(main())
