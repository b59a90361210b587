"""Build check on the gfx950 assembly of this library's kernels.

Loop-carried MFMA accumulators must not be read by a VALU instruction at the
top of a loop too soon after the MFMA at the bottom of the loop wrote them.
The result of an XDL MFMA is only readable some wait states after issue (up to
19 for 16-pass shapes).  A read placed at the loop header, reached through the
back edge, was seen without those wait states and returned the previous value
of the last accumulator register.  This script scans the device assembly
(``hipcc --offload-device-only -S``) and fails when a loop header reads
(v_accvgpr_read / v_accvgpr_mov) an AGPR that an MFMA near the back-edge
branch wrote, with fewer than REQUIRED wait states between them.  Each
instruction in between counts one wait state, each MFMA at least four (its
issue occupies the matrix pipe for at least four cycles), each s_nop N N + 1.

    python hazard_lint.py kernel.s [more.s ...]     exit status 1 on a finding
"""
import re
import sys

REQUIRED = 19
MFMA_MIN_WAIT = 4


def agprs(tok):
    m = re.match(r"a\[(\d+):(\d+)\]", tok)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"a(\d+)$", tok)
    return {int(m.group(1))} if m else set()


def cost(ins):
    if ins.startswith("v_mfma"):
        return MFMA_MIN_WAIT
    m = re.match(r"s_nop\s+(\d+)", ins)
    return int(m.group(1)) + 1 if m else 1  # s_nop N: N + 1 wait states


def instructions(lines, start, stop, step):
    for j in range(start, stop, step):
        t = lines[j]
        if t and not t.startswith(";") and not t.startswith("."):
            yield j, t


def scan(path):
    text = open(path).read()
    findings = []
    for km in re.finditer(r"^(_Z\S+):[^\n]*\n(.*?)s_endpgm", text, re.S | re.M):
        name, body = km.group(1), km.group(2)
        lines = [l.strip() for l in body.split("\n")]
        labels = {}
        for k, l in enumerate(lines):
            m = re.match(r"(\.LBB\w+):", l)
            if m:
                labels[m.group(1)] = k
        for k, l in enumerate(lines):
            m = re.match(r"s_c?branch\w*\s+(\.LBB\w+)", l)
            if not m or labels.get(m.group(1), k + 1) > k:
                continue
            head = labels[m.group(1)]
            # wait states from each AGPR's last MFMA write to the back-edge branch
            written, waits = {}, 0
            for _, t in instructions(lines, k - 1, head, -1):
                if t.startswith("v_mfma"):
                    for r in agprs(t.split(None, 1)[1].split(",")[0].strip()):
                        written.setdefault(r, waits)
                waits += cost(t)
                if waits >= REQUIRED:
                    break
            # reads at the loop header before its first MFMA
            waits_head = 1
            for _, t in instructions(lines, head + 1, len(lines), 1):
                if t.startswith("v_mfma") or waits_head >= REQUIRED:
                    break
                if t.startswith("v_accvgpr_read") or t.startswith("v_accvgpr_mov"):
                    for r in agprs(t.split(",")[-1].strip()):
                        if r in written and written[r] + waits_head < REQUIRED:
                            findings.append(f"{path}: {name}: {m.group(1)}: '{t}' "
                                            f"{written[r] + waits_head} wait states after the MFMA write")
                waits_head += cost(t)
    return findings


def main(paths):
    bad = [f for p in paths for f in scan(p)]
    for f in bad:
        print("hazard:", f)
    if bad:
        print(f"{len(bad)} loop-header MFMA result hazard(s)")
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
