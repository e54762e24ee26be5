"""CPU tests of the build-time register-allocation guard of the sweep and
residual kernels (blaze_amd/csrc/check_resources.py): its parser on remark
text in the compiler's format, each rule on a kernel that breaks it, and
the remarks of the current build."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "blaze_amd", "csrc")
spec = importlib.util.spec_from_file_location("check_resources", os.path.join(CSRC, "check_resources.py"))
CR = importlib.util.module_from_spec(spec)
spec.loader.exec_module(CR)


def remark(name, vgpr=183, agpr=0, scratch=0, spill=0):
    rows = [f"k_project.hip:905:1: remark: Function Name: {name} [-Rpass-analysis=kernel-resource-usage]",
            "  905 | __global__ void k(...) {",
            "k_project.hip:905:1: remark:     TotalSGPRs: 40 [-Rpass-analysis=kernel-resource-usage]",
            f"k_project.hip:905:1: remark:     VGPRs: {vgpr} [-Rpass-analysis=kernel-resource-usage]",
            f"k_project.hip:905:1: remark:     AGPRs: {agpr} [-Rpass-analysis=kernel-resource-usage]",
            f"k_project.hip:905:1: remark:     ScratchSize [bytes/lane]: {scratch} [-Rpass-analysis=kernel-resource-usage]",
            "k_project.hip:905:1: remark:     Occupancy [waves/SIMD]: 2 [-Rpass-analysis=kernel-resource-usage]",
            "k_project.hip:905:1: remark:     SGPRs Spill: 0 [-Rpass-analysis=kernel-resource-usage]",
            f"k_project.hip:905:1: remark:     VGPRs Spill: {spill} [-Rpass-analysis=kernel-resource-usage]",
            "k_project.hip:905:1: remark:     LDS Size [bytes/block]: 0 [-Rpass-analysis=kernel-resource-usage]"]
    return "\n".join(rows) + "\n"


PAIR = "_ZN3blz12_GLOBAL__N_111k_rbgs_pairILi4ELi1EEEvPfjjjjjjjiiiiifPKi"
RES = "_ZN3blz12_GLOBAL__N_110k_residualENS_2RBES1_NS_3RBLEiiidDv4_iDv2_iPdPKi"
OTHER = "_ZN3blz12_GLOBAL__N_112k_divergenceEPKfS2_S2_Pfiiif"


def test_parser_reads_every_field():
    k = CR.parse(remark(PAIR, vgpr=151, agpr=2, scratch=8, spill=3))
    f = k[PAIR]
    assert (f["VGPRs"], f["AGPRs"], f["ScratchSize [bytes/lane]"], f["VGPRs Spill"]) == ("151", "2", "8", "3")
    assert f["LDS Size [bytes/block]"] == "0"


def test_clean_kernels_pass():
    assert CR.violations(CR.parse(remark(PAIR) + remark(RES, vgpr=88))) == []


@pytest.mark.parametrize("kw,why", [(dict(agpr=16), "AGPRs"), (dict(scratch=64), "scratch"),
                                    (dict(spill=5), "spills"), (dict(vgpr=304), "VGPRs >")])
def test_each_rule_fails_a_sweep_kernel(kw, why):
    bad = CR.violations(CR.parse(remark(PAIR, **kw) + remark(RES)))
    assert len(bad) == 1 and bad[0][0] == PAIR and why in bad[0][1]


def test_unguarded_kernels_are_ignored_but_some_guarded_one_must_exist():
    assert CR.violations(CR.parse(remark(OTHER, agpr=8) + remark(RES))) == []
    with pytest.raises(ValueError):
        CR.violations(CR.parse(remark(OTHER)))
    with pytest.raises(ValueError):
        CR.violations(CR.parse("no remarks at all\n"))


def test_missing_fields_are_a_violation():
    text = "x.hip:1:1: remark: Function Name: " + PAIR + " [-Rpass-analysis=kernel-resource-usage]\n"
    bad = CR.violations(CR.parse(text))
    assert bad and "missing" in bad[0][1]


def test_current_build_passes_the_guard():
    path = os.path.join(CSRC, "_obj", "k_project.remarks")
    if not os.path.exists(path):
        pytest.skip("library not built here (no remarks file)")
    with open(path) as f:
        k = CR.parse(f.read())
    assert sum(1 for n in k if CR.GUARDED.search(n)) >= 20
    assert CR.violations(k) == []
