"""Writes tests/golden/border_tables.json: the border-copy known answers.

Reads the reference's Code/Maple/src/Shaders/DDGI/BorderUpdate.glsl as text
(lines 25-133 hold the literal `Offsets` tables for the 16x16 depth probes and
the 8x8 irradiance probes) and stores the integer tuples (srcX, srcY, dstX, dstY)
as data. Run in a container where /root/reference exists; the JSON is committed.
"""
import json
import os
import re
import sys

REF = "/root/reference/Code/Maple/src/Shaders/DDGI/BorderUpdate.glsl"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "border_tables.json")


def main():
    text = open(REF).read()
    tables = {}
    for m in re.finditer(r"const ivec4 Offsets\[(\d+)\] = ivec4\[\]\((.*?)\);", text, re.S):
        n = int(m.group(1))
        rows = [tuple(int(v) for v in t.split(",")) for t in re.findall(r"ivec4\(([^)]*)\)", m.group(2))]
        assert len(rows) == n, (n, len(rows))
        side = (n - 4) // 4
        tables[str(side)] = rows
    json.dump({"source": "Code/Maple/src/Shaders/DDGI/BorderUpdate.glsl:25-133", "tables": tables}, open(OUT, "w"))

    print("wrote", OUT, {k: len(v) for k, v in tables.items()})


if __name__ == "__main__":
    sys.exit(main())
