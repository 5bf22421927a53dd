"""Builds the N-API addon gsr.node next to this file (gcc, node headers).

The addon links libgsr.so from ../gsr with an $ORIGIN-relative rpath, so the
tree can move (it travels to the GPU host as is).  Skipped, with a message, when
the node headers are absent.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
NODE_INC = os.environ.get("GSR_NODE_INCLUDE", "/usr/include/node")


def build() -> str | None:
    if not os.path.exists(os.path.join(NODE_INC, "node_api.h")):
        print(f"node headers not found under {NODE_INC}; gsr.node not built", file=sys.stderr)
        return None
    out = os.path.join(HERE, "gsr.node")
    tmp = out + ".tmp"
    cmd = ["gcc", "-O2", "-shared", "-fPIC", "-std=c11", "-Wall", "-DNODE_GYP_MODULE_NAME=gsr",
           f"-I{NODE_INC}", f"-I{os.path.join(ROOT, 'include')}", os.path.join(HERE, "gsr_napi.c"),
           f"-L{os.path.join(PKG, 'gsr')}", "-lgsr", "-Wl,-rpath,$ORIGIN/../gsr", "-o", tmp]
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    print(f"built {out}")
    return out


if __name__ == "__main__":
    build()
