"""The explicit and adjoint kernels have no lane-divergent branch besides the node guard.

hipcc (ROCm 7.2, gfx950) has miscompiled lane-divergent if/else regions in these kernels: in a
join block it placed register copies (of a live state, of a per-node address) ahead of the
instruction that restores the exec mask, so lanes from the other arm kept stale values.  The
generated model code and the step helpers therefore select between precomputed values
(GOSS_SELECT / m_select in goss_model_support.h).  This test compiles one variant to device
assembly with the library's flags (``make -C goss_amd/csrc asm V=...``) and counts, per kernel,
the exec-mask saves that open a divergent region (``s_and_saveexec``, ``s_andn2_saveexec``,
``s_or_saveexec`` other than the whole-wave form, ``s_xor_b64 exec``).  The one allowed region is
``if (node < N)``.

The implicit, RKF32 and mixed-batch kernels branch per lane by design (Newton iterations, step
rejection, per-population code).  For them the test looks for the miscompile itself: in every
join block (a labelled block that restores exec with ``s_or_b64 exec, exec, ...``), a VGPR
written by a vector instruction before the restore and read after it within the block, before
any rewrite, would hand the lanes re-enabled by the restore a value they did not compute
(``v_readlane`` / ``v_writelane`` address one lane whatever exec holds and are exempt).  Uses in
later blocks are not followed, so a clean result bounds the pattern within join blocks only.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "goss_amd", "csrc")
VARIANT = "tp06_m_fp6"
CHECKED = ("explicit_kernel", "adjoint_record_kernel", "adjoint_recompute_kernel", "adjoint_reverse_kernel")
BRANCHING = ("implicit_kernel", "rkf32_kernel", "batch_kernel")
STORES = ("global_store", "buffer_store", "flat_store", "scratch_store", "ds_write")
LANE_OPS = ("v_writelane", "v_readlane", "v_readfirstlane")


def divergent_regions(asm_text):
    """kernel symbol -> number of exec-mask saves opening a divergent region."""
    counts, cur = {}, None
    for line in asm_text.split("\n"):
        m = re.match(r"^(_Z\w+):", line)
        if m:
            cur = m.group(1)
            counts[cur] = 0
            continue
        if cur is None:
            continue
        s = line.strip()
        if s.startswith(("s_and_saveexec", "s_andn2_saveexec", "s_xor_b64 exec")) or (
                s.startswith("s_or_saveexec") and not s.endswith("-1")):
            counts[cur] += 1
        elif s.startswith("s_endpgm"):
            cur = None
    return counts


def _vregs(op):
    m = re.match(r"^-?\|?([va])\[(\d+):(\d+)\]", op.strip())
    if m:
        return {f"{m.group(1)}{i}" for i in range(int(m.group(2)), int(m.group(3)) + 1)}
    m = re.match(r"^-?\|?([va])(\d+)\b", op.strip())
    return {f"{m.group(1)}{m.group(2)}"} if m else set()


def _split(line):
    s = line.split(";")[0].strip()
    if not s:
        return None, []
    parts = s.split(None, 1)
    return parts[0], [o.strip() for o in parts[1].split(",")] if len(parts) > 1 else []


def stale_lane_reads(asm_text):
    """kernel symbol -> [(block, write, read)] for VGPRs written before the exec restore of a
    join block and read after it within the block."""
    blocks, cur, blk = [], None, None
    for line in asm_text.split("\n"):
        m = re.match(r"^(_Z\w+):", line)
        if m:
            cur, blk = m.group(1), None
            continue
        if cur is None:
            continue
        s = line.strip()
        if re.match(r"^\.LBB\w+:", s):
            blk = (cur, s.split()[0], [])
            blocks.append(blk)
        elif s.startswith("s_endpgm"):
            cur = None
        elif blk is not None and s and not s.startswith((";", ".")):
            blk[2].append(s)
    found = {}
    for kern, label, ins in blocks:
        found.setdefault(kern, [])
        restore = None
        for i, s in enumerate(ins):
            if s.startswith("s_or_b64 exec, exec,"):
                restore = i
                break
            if s.startswith(("s_cbranch", "s_branch")):
                break
        if restore is None:
            continue
        written = {}
        for s in ins[:restore]:
            opc, ops = _split(s)
            if opc and opc.startswith("v_") and not opc.startswith(LANE_OPS) and ops:
                for r in _vregs(ops[0]):
                    written[r] = s
        for s in ins[restore + 1:]:
            opc, ops = _split(s)
            if not opc:
                continue
            if opc.startswith(("s_cbranch", "s_branch")):
                break
            is_store = opc.startswith(STORES)
            srcs = set().union(*[_vregs(o) for o in (ops if is_store or not opc.startswith("v_") else ops[1:])]) \
                if ops else set()
            for r in sorted(srcs & set(written)):
                found[kern].append((label, written.pop(r), s))
            if opc.startswith("v_") and ops and not is_store:
                for r in _vregs(ops[0]):
                    written.pop(r, None)
    return found


def test_stale_lane_read_finder_on_known_listings():
    bad = "\n".join([
        "_ZN8goss_hip14implicit_kernelIXE:",
        ".LBB0_3:",
        "\tv_mov_b32_e32 v7, v2",                 # copy made under the other arm's exec
        "\ts_or_b64 exec, exec, s[4:5]",
        "\tv_add_f32_e32 v8, v7, v1",             # read by every lane
        "\ts_endpgm",
    ])
    good = "\n".join([
        "_ZN8goss_hip14implicit_kernelIXE:",
        ".LBB0_3:",
        "\tv_writelane_b32 v9, s2, 0",            # one lane, whatever exec holds
        "\tv_mov_b32_e32 v7, v2",
        "\ts_or_b64 exec, exec, s[4:5]",
        "\tv_mov_b32_e32 v7, v3",                 # rewritten before any read
        "\tv_add_f32_e32 v8, v7, v9",
        "\tglobal_store_dword v[0:1], v8, off",
        "\ts_endpgm",
    ])
    hits = stale_lane_reads(bad)["_ZN8goss_hip14implicit_kernelIXE"]
    assert len(hits) == 1 and "v_mov_b32_e32 v7, v2" in hits[0][1]
    assert stale_lane_reads(good)["_ZN8goss_hip14implicit_kernelIXE"] == []


def test_divergent_region_counter_on_a_known_listing():
    text = "\n".join([
        "_ZN8goss_hip15explicit_kernelIXE:",
        "\tv_cmp_gt_u32_e32 vcc, s5, v0",
        "\ts_and_saveexec_b64 s[2:3], vcc",
        "\ts_or_saveexec_b64 s[100:101], -1",   # whole-wave mode for a spill: not a region
        "\ts_and_saveexec_b64 s[6:7], s[0:1]",
        "\ts_xor_b64 s[0:1], exec, s[2:3]",     # not a write of exec
        "\ts_endpgm",
    ])
    assert divergent_regions(text) == {"_ZN8goss_hip15explicit_kernelIXE": 2}


@pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs make and hipcc")
def test_explicit_and_adjoint_kernels_branch_only_on_the_node_guard():
    subprocess.check_call(["make", "-s", "-j2", "-C", CSRC, "asm", f"V={VARIANT}"])
    build = os.path.join(ROOT, "goss_amd", "_build")
    seen = 0
    for prec in ("f64", "f32"):
        with open(os.path.join(build, f"asm_{VARIANT}_{prec}.s")) as f:
            counts = divergent_regions(f.read())
        for name, n in counts.items():
            if any(k in name for k in CHECKED):
                seen += 1
                assert n == 1, (prec, name, n)
    # 7 explicit solvers x 3 policies (FP32 TU: plain and compensated; FP64 TU: one) + 3 adjoint
    # kernels (record, recompute, reverse) x 7 solvers per precision.
    assert seen >= 7 * 3 + 2 * 3 * 7


@pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs make and hipcc")
def test_branching_kernels_have_no_stale_lane_reads_in_join_blocks():
    subprocess.check_call(["make", "-s", "-j2", "-C", CSRC, "asm", f"V={VARIANT}"])
    subprocess.check_call(["make", "-s", "-C", CSRC, "asm-batch"])
    build = os.path.join(ROOT, "goss_amd", "_build")
    listings = [f"asm_{VARIANT}_f64.s", f"asm_{VARIANT}_f32.s", "asm_batch_f32.s"]
    kernels = 0
    for name in listings:
        with open(os.path.join(build, name)) as f:
            found = stale_lane_reads(f.read())
        for kern, hits in found.items():
            if any(k in kern for k in BRANCHING):
                kernels += 1
                assert hits == [], (name, kern, hits[:3])
    # FP64: implicit kernels (ImplicitEuler, ESDIRK23a, ThetaSolver x 2 Jacobian modes) and
    # RKF32; FP32: RKF32; the batch kernel.
    assert kernels >= 6 + 1 + 1 + 1
