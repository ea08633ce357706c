"""The hand-written scalar-load blocks of the IDW kernel keep every load's
address out of the registers the block's earlier loads write.

A multi-instruction asm block issues its loads back to back; a destination
tuple that overlaps a later load's address register can be overwritten
before that load reads it, which sends the load to a garbage address.
Early-clobber outputs forbid the overlap; this test compiles the kernels as
the build does and checks the register assignment of every such block.
"""
import os
import re
import shutil
import subprocess

import pytest

from pargo_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def regs(tok):
    m = re.match(r"s\[(\d+):(\d+)\]", tok)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"s(\d+)$", tok)
    return {int(m.group(1))} if m else set()


def overlaps(asm_text):
    """(number of multi-load asm blocks, list of offending blocks)."""
    blocks = re.findall(r";;#ASMSTART(.*?);;#ASMEND", asm_text, re.S)
    n, bad = 0, []
    for b in blocks:
        loads = re.findall(r"s_load_dword\w*\s+(s\[\d+:\d+\]|s\d+),\s*(s\[\d+:\d+\])", b)
        if len(loads) < 2:
            continue
        n += 1
        for i, (dst, _) in enumerate(loads):
            if any(regs(dst) & regs(addr) for _, addr in loads[i + 1:]):
                bad.append(loads)
                break
    return n, bad


def test_checker_flags_an_overlap():
    clobbered = (";;#ASMSTART\n s_load_dwordx8 s[20:27], s[16:17], 0x0\n"
                 " s_load_dwordx16 s[68:83], s[20:21], 0x0\n;;#ASMEND\n")
    clean = (";;#ASMSTART\n s_load_dwordx8 s[20:27], s[16:17], 0x0\n"
             " s_load_dwordx16 s[68:83], s[28:29], 0x0\n;;#ASMEND\n")
    assert overlaps(clobbered) == (1, [[("s[20:27]", "s[16:17]"), ("s[68:83]", "s[20:21]")]])
    assert overlaps(clean) == (1, [])


@pytest.mark.skipif(not os.path.exists(build.HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_idw_scalar_load_blocks_keep_their_addresses(tmp_path):
    out = tmp_path / "idw.s"
    cmd = [build.HIPCC] + build.HIP_FLAGS + ["--cuda-device-only", "-S", "-I" + os.path.join(ROOT, "include"),
                                             "-I" + build.CSRC, os.path.join(build.CSRC, "idw.hip"), "-o", str(out)]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
    n, bad = overlaps(out.read_text())
    assert n >= 1, "no multi-load scalar asm block found"
    assert not bad, bad
