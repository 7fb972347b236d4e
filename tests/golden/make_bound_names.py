"""Writes bound_names.txt: the ma_* entry points the reference's python/map.py binds
with ctypes at import time (every `_lib.ma_<name>.argtypes` / `.restype` line), one per
line. Run in a container that holds the reference tree (text is read, nothing is run)."""
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/python/map.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bound_names.txt")

names = []
with open(REF) as f:
    for line in f:
        if line.lstrip().startswith("#"):
            continue
        m = re.match(r"\s*_lib\.(ma_\w+)\.(argtypes|restype)\s*=", line)
        if m and m.group(1) not in names:
            names.append(m.group(1))
with open(OUT, "w") as f:
    f.write("\n".join(names) + "\n")
print(len(names), "names ->", OUT)
