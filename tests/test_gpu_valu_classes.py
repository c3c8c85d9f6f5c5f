"""Issue cost of single VALU instruction forms on the GPU (tests/cpp/valu_classes,
built by __graft_entry__.build): the facts the force kernels' instruction
choices rest on. On gfx950 an FP32 FMA, add or multiply whose operands are
VGPRs and constants issues at about twice the rate of the same instruction
with an SGPR operand, and of every packed FP16 instruction; the
transcendentals take about twice as long again."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_valu_issue_classes():
    exe = os.path.join(ROOT, "tests", "cpp", "valu_classes")
    assert os.path.exists(exe), "tests/cpp/valu_classes is not built; run __graft_entry__.build()"
    out = subprocess.run([exe, "500"], check=True, capture_output=True, text=True, timeout=120).stdout
    print(out)
    ns = {}
    for line in out.splitlines()[1:]:
        name, t, _ = line.rsplit(None, 2)
        ns[name.strip()] = float(t)
    fma_v, fma_s = ns["v_fma_f32 (VGPR operands)"], ns["v_fma_f32 (one SGPR operand)"]
    assert all(t > 0 for t in ns.values())
    # the operand form decides the FP32 issue rate
    assert fma_s / fma_v > 1.3
    assert ns["v_add_f32 (SGPR)"] / ns["v_add_f32 (VGPR)"] > 1.3
    # a packed FP16 FMA costs about what the SGPR-operand FP32 FMA does
    assert 0.75 < ns["v_pk_fma_f16"] / fma_s < 1.3
    # transcendentals cost more than either
    assert ns["v_rcp_f32"] > 1.5 * fma_s
