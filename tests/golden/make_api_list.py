"""Writes reference_api.json: the name of every function the reference's
class headers declare (src/Headers/*.h), as (header, line, name).  Run in a
container that holds the reference sources:

    python tests/golden/make_api_list.py /root/reference/src/Headers
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DECL = re.compile(r"^\s*(?:struct\s+\w+\s*\**|void|double|int)\s*\**\s*\(?\s*\**\s*(\w+)\s*\(")


def main(hdir):
    out = []
    for fn in sorted(os.listdir(hdir)):
        if not fn.endswith(".h"):
            continue
        with open(os.path.join(hdir, fn)) as f:
            for n, line in enumerate(f, 1):
                s = line.split("//")[0]
                m = DECL.match(s)
                if m and ";" in s and "#define" not in s:
                    out.append([fn, n, m.group(1)])
    with open(os.path.join(HERE, "reference_api.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(f"{len(out)} declarations")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/Headers")
