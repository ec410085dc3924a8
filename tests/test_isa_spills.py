"""The gfx950 code of the HIP sources has no register spill that is stored
under a loop's exit mask.

A value live across a loop whose threads leave it at different times must be
spilled with every lane that holds it active.  hipcc (ROCm 7.2) has been seen
to place such spill stores in the loop's exit block ahead of the instruction
that restores the execution mask; when every lane has left the loop that mask
is empty, the store writes nothing, and the reload after the loop returns
stale scratch (the SCF kernel once lost its repulsion and three-body energy
totals this way).  This test compiles each source to assembly with the
library's own flags and fails on the pattern: a `Folded Spill` scratch store
in a basic block before the block's first `s_or_b64 exec, exec, ...`.
Function entry blocks (callee-saved registers, stored under the caller's
mask or a whole-wave mask) are not checked.
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "tblite-cuda_amd", "csrc")
SOURCES = ["scf.hip", "scf_grid.hip", "gradient.hip", "setup.hip", "periodic.hip"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    return None


def masked_exit_spills(asm_text):
    """(function, spill line, restore line) for each block that stores a
    spill before restoring the execution mask."""
    hits = []
    fn, entry, block = None, False, []

    def check():
        if entry:
            return
        spill = None
        for ln, t in block:
            s = t.strip()
            if s.startswith("scratch_store") and "Folded Spill" in s and spill is None:
                spill = ln
            if re.match(r"s_or_b64 exec, exec, ", s):
                if spill is not None:
                    hits.append((fn, spill, ln))
                return

    for ln, t in enumerate(asm_text.splitlines(), 1):
        m = re.match(r"^([_A-Za-z][\w.$]*):", t)
        lbl = re.match(r"^\.LBB\w+:", t) or re.match(r"^; %bb\.\d+:", t)
        if m or lbl:
            if block:
                check()
            block = []
            if m:
                fn, entry = m.group(1), True
            elif re.match(r"^; %bb\.0:", t):
                entry = True
            else:
                entry = False
        block.append((ln, t))
    if block:
        check()
    return hits


def test_scanner_finds_the_pattern():
    """The scanner flags a spill ahead of the mask restore, and only there."""
    bad = "\n".join([
        "_Zk:", "; %bb.0:", "  s_mov_b32 s0, 0",
        ".LBB0_5:", "  scratch_store_dwordx2 off, v[2:3], off offset:8 ; 8-byte Folded Spill",
        "  s_or_b64 exec, exec, s[10:11]",
        ".LBB0_6:", "  s_or_b64 exec, exec, s[10:11]",
        "  scratch_store_dwordx2 off, v[2:3], off offset:16 ; 8-byte Folded Spill",
    ])
    hits = masked_exit_spills(bad)
    assert [(h[0], h[1]) for h in hits] == [("_Zk", 5)]


@pytest.mark.skipif(_hipcc() is None, reason="hipcc not found")
@pytest.mark.parametrize("src", SOURCES)
def test_no_spill_under_exit_mask(src):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tblite-cuda_amd"))
    import build

    flags = [f for f in build.FLAGS if f not in ("-fPIC",)]
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, src + ".s")
        cmd = [_hipcc()] + flags + ["--cuda-device-only", "-S", "-o", out,
                                    os.path.join(CSRC, src)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        with open(out) as fh:
            hits = masked_exit_spills(fh.read())
    assert not hits, "spills stored under a loop exit mask: " + "; ".join(
        f"{f} line {a} (restore at {b})" for f, a, b in hits[:10])
