"""The row-broadcast kernel reads lane data through DPP on inline-assembly instructions, which
the compiler does not check for the DPP hazard: a VALU write to the DPP source in the two
instructions before the read.  This compiles the kernels to gfx950 assembly (no GPU needed)
and scans every DPP instruction."""
import os
import re
import shutil
import subprocess

import pytest

from tests._helpers import ROOT


def _assembly(tmp_path):
    import importlib.util
    spec = importlib.util.spec_from_file_location("cctpy_amd_build", os.path.join(ROOT, "cctpy_amd", "_build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    try:
        hipcc = b.hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")
    flags = [f for f in b.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    out = tmp_path / "cctamd.s"
    src = os.path.join(ROOT, "cctpy_amd", "csrc", "cctamd.hip")
    res = subprocess.run([hipcc, *flags, "--cuda-device-only", "-S", "-o", str(out), src], capture_output=True,
                         text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    return out.read_text()


def _regs(op):
    op = op.strip()
    m = re.match(r"^v\[(\d+):(\d+)\]$", op)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"^v(\d+)$", op)
    return {int(m.group(1))} if m else set()


def dpp_hazards(asm: str):
    """(DPP instruction, earlier VALU writer) pairs closer than 2 wait states, over each
    function's instruction stream, also across loop back edges."""
    hazards, n_dpp = [], 0
    for body in re.split(r"\n(?=_Z\S*:)", asm):
        ins, labels = [], {}
        for line in body.split("\n"):
            t = line.strip()
            m = re.match(r"^(\.LBB\S+):", t)
            if m:
                labels[m.group(1)] = len(ins)
                continue
            if not t or t.startswith((";", ".")):
                continue
            ins.append(t.split(";")[0].strip())
        backs = {}
        for i, t in enumerate(ins):
            m = re.match(r"^s_(?:cbranch\w*|branch)\s+(\.LBB\S+)", t)
            if m and m.group(1) in labels and labels[m.group(1)] <= i:
                backs.setdefault(labels[m.group(1)], []).append(i)

        def writer(start, src0):
            waits, j = 0, start
            while j >= 0 and waits < 2:
                u = ins[j]
                if u.startswith("s_nop"):
                    waits += int(u.split()[1]) + 1
                elif u.startswith("v_"):
                    if not u.startswith("v_cmp") and len(u.split(None, 1)) > 1:
                        if _regs(u.split(None, 1)[1].split(",")[0]) & src0:
                            return u
                    waits += 1
                j -= 1
            return None

        for i, t in enumerate(ins):
            op = t.split()[0]
            if not op.startswith("v_") or "_dpp" not in op:
                continue
            n_dpp += 1
            src0 = _regs(t.split(None, 1)[1].split(",")[1].split()[0])
            hit = writer(i - 1, src0)
            for pos, brs in backs.items():
                if pos <= i < pos + 3:
                    for b in brs:
                        hit = hit or writer(b - 1, src0)
            if hit:
                hazards.append((t, hit))
    return hazards, n_dpp


def test_scanner_flags_a_planted_hazard():
    asm = "_Zk:\n\tv_mov_b32_e32 v3, v1\n\tv_fmac_f32_dpp v5, v3, v2 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
    hz, n = dpp_hazards(asm)
    assert n == 1 and len(hz) == 1
    ok = "_Zk:\n\tv_mov_b32_e32 v3, v1\n\ts_nop 1\n\tv_fmac_f32_dpp v5, v3, v2 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
    assert dpp_hazards(ok)[0] == []


def test_built_kernels_have_no_dpp_hazard(tmp_path):
    if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not available")
    hz, n = dpp_hazards(_assembly(tmp_path))
    assert n > 500, n  # the row kernel's broadcasts are there
    assert hz == [], hz[:5]
