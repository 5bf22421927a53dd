"""Writes sample_head.ply: the header and first 48 records of the reference's
tour/sample.ply (15,888 Gaussians, SH0), with the vertex count set to 48.

Run in a tree that has the reference at /root/reference; the output is the
committed fixture (the reference itself does not travel).
"""
import os
import re

SRC = "/root/reference/tour/sample.ply"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sample_head.ply")
KEEP = 48

data = open(SRC, "rb").read()
end = data.index(b"end_header\n") + len(b"end_header\n")
header = data[:end].decode("latin1")
props = re.findall(r"property float (\w+)", header)
header = re.sub(r"element vertex \d+", f"element vertex {KEEP}", header)
open(OUT, "wb").write(header.encode("latin1") + data[end:end + KEEP * len(props) * 4])
print(OUT, KEEP, "records of", len(props), "floats")
