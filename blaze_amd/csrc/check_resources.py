#!/usr/bin/env python3
"""Build-time guard of the pressure kernels' register allocation.

The Gauss-Seidel sweep kernels keep their loads in flight in a register
ring written by inline-asm buffer loads and consumed after explicit
s_waitcnt statements. The compiler does not see those loads as pending: if
it ever copies ring registers (to AGPRs or scratch, or between VGPRs when
the kernel outgrows the 256 architectural VGPRs) before the wait, the copy
reads a value that has not arrived yet and the solve silently diverges.
Reads the compiler's kernel-resource-usage remarks
(-Rpass-analysis=kernel-resource-usage) and fails when a sweep or residual
kernel uses AGPRs, scratch or spills, or more than 256 VGPRs.

usage: check_resources.py REMARKS_FILE
"""
import re
import sys

GUARDED = re.compile(r"k_rbgs_|k_residual")
VGPR_LIMIT = 256


def parse(text):
    """{mangled kernel name: {field: value}} from the remark lines."""
    out, cur = {}, None
    for line in text.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
            out[cur] = {}
            continue
        m = re.search(r"remark:\s+([A-Za-z][A-Za-z /\[\]]*?): (\S+) \[-Rpass-analysis", line)
        if m and cur is not None:
            out[cur][m.group(1).strip()] = m.group(2)
    return out


def violations(kernels):
    """(kernel, reason) for every guarded kernel that breaks a rule; raises
    ValueError when no guarded kernel is present (remarks missing or their
    format changed: the guard must not pass by seeing nothing)."""
    bad, seen = [], 0
    for name, f in kernels.items():
        if not GUARDED.search(name):
            continue
        seen += 1
        try:
            vgpr = int(f["VGPRs"])
            agpr = int(f["AGPRs"])
            scratch = int(f["ScratchSize [bytes/lane]"])
            spill = int(f["VGPRs Spill"])
        except (KeyError, ValueError):
            bad.append((name, "resource remark fields missing"))
            continue
        if agpr:
            bad.append((name, f"{agpr} AGPRs"))
        if scratch:
            bad.append((name, f"{scratch} bytes of scratch per lane"))
        if spill:
            bad.append((name, f"{spill} VGPR spills"))
        if vgpr > VGPR_LIMIT:
            bad.append((name, f"{vgpr} VGPRs > {VGPR_LIMIT}"))
    if not seen:
        raise ValueError("no k_rbgs_* / k_residual* kernel in the remarks")
    return bad


def main(argv):
    if len(argv) != 2:
        print(__doc__, file=sys.stderr)
        return 2
    with open(argv[1]) as f:
        kernels = parse(f.read())
    try:
        bad = violations(kernels)
    except ValueError as e:
        print(f"check_resources: {e}", file=sys.stderr)
        return 1
    for name, why in bad:
        print(f"check_resources: {name}: {why}", file=sys.stderr)
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
