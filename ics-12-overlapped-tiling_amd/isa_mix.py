#!/usr/bin/env python3
"""Instruction mix of a kernel's steady-state loop in the gfx950 ISA of this
build.

Compiles one kernel source of csrc/ to assembly with the library's own flags
(hipcc -S --cuda-device-only, into a temporary directory), finds the named
kernel instantiation, takes the loop holding its largest basic block (the
unrolled steady-state blocks of the wave-pipelined kernels) and counts its
instructions by opcode, with the DPP forms counted apart.

  python3 ics-12-overlapped-tiling_amd/isa_mix.py                     # k3w<J3D, T=3, FR=12>
  python3 ics-12-overlapped-tiling_amd/isa_mix.py --kernel '_ZN2ot3k3w3k3wILi4ELi2ELi12ELb0EEEvNS_10LaunchGeomENS_9FieldPtrsEi'
"""
from __future__ import annotations

import argparse
import collections
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(HERE, "..", "include"), "-I" + os.path.join(HERE, "csrc"),
         "--cuda-device-only", "-S"]
DEFAULT_KERNEL = "_ZN2ot3k3w3k3wILi4ELi3ELi12ELb0EEEvNS_10LaunchGeomENS_9FieldPtrsEi"


def blocks_of(asm: list[str], kernel: str):
    """Basic blocks of the kernel: (label, loop header or None, instructions),
    the loop taken from the compiler's loop comments."""
    start = next(i for i, l in enumerate(asm) if l.startswith(kernel + ":"))
    end = next(i for i in range(start, len(asm)) if asm[i].strip().startswith("s_endpgm"))
    blocks, label, loop, ins = [], "entry", None, []
    for line in asm[start + 1:end + 1]:
        s = line.strip()
        m = re.match(r"^\.(LBB\w+):(.*)", s)
        if m:
            blocks.append((label, loop, ins))
            label, ins = m.group(1), []
            h = re.search(r"Header=(BB\w+)", m.group(2))
            loop = ("L" + h.group(1)) if h else (label if "Loop Header" in m.group(2) else None)
            continue
        if not s or s.startswith((".", ";")):
            continue
        ins.append(s.split(";")[0].strip())
    blocks.append((label, loop, ins))
    return blocks


def mix(ins: list[str]) -> collections.Counter:
    c = collections.Counter()
    for x in ins:
        op = x.split()[0]
        if op.startswith("v_") and ("row_" in x or "wave_sh" in x or "quad_perm" in x):
            op += " (dpp)"
        c[op] += 1
    return c


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--src", default=os.path.join(HERE, "csrc", "k3w_j3d.hip"))
    ap.add_argument("--kernel", default=DEFAULT_KERNEL, help="mangled kernel name")
    a = ap.parse_args()
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "k.s")
        subprocess.run([HIPCC] + FLAGS + [a.src, "-o", out], check=True, stderr=subprocess.DEVNULL)
        asm = open(out).read().split("\n")
    blocks = blocks_of(asm, a.kernel)
    loops = collections.defaultdict(list)
    for lab, loop, ins in blocks:
        if loop:
            loops[loop].append(ins)
    if not loops:
        sys.exit("no loop in " + a.kernel)
    # the steady-state loop: the one holding the largest basic block
    head, parts = max(loops.items(), key=lambda t: max(len(b) for b in t[1]))
    ins = [x for b in parts for x in b]
    c = mix(ins)
    valu = sum(v for k, v in c.items() if k.startswith("v_"))
    print(f"{a.kernel}: loop {head} ({len(parts)} basic blocks), {len(ins)} instructions, {valu} VALU")
    for k, v in c.most_common():
        print(f"{v:6d}  {k}")


if __name__ == "__main__":
    main()
