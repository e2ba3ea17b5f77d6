"""The explicit and adjoint kernels have no lane-divergent branch besides the node guard.

hipcc (ROCm 7.2, gfx950) has miscompiled lane-divergent if/else regions in these kernels: in a
join block it placed register copies (of a live state, of a per-node address) ahead of the
instruction that restores the exec mask, so lanes from the other arm kept stale values.  The
generated model code and the step helpers therefore select between precomputed values
(GOSS_SELECT / m_select in goss_model_support.h).  This test compiles one variant to device
assembly with the library's flags (``make -C goss_amd/csrc asm V=...``) and counts, per kernel,
the exec-mask saves that open a divergent region (``s_and_saveexec``, ``s_andn2_saveexec``,
``s_or_saveexec`` other than the whole-wave form, ``s_xor_b64 exec``).  The one allowed region is
``if (node < N)``.  The implicit and RKF32 kernels branch per lane by design (Newton iterations,
step rejection) and are not checked.
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
