"""Code generation checks of the HIP kernels (hipcc cross-compiles for
gfx950 here; no GPU needed).

A buffer resource whose fields the compiler cannot prove uniform lives in
vector registers, and every buffer load through it then runs in a loop over
the distinct values among the lanes (v_readfirstlane / v_cmp_eq / exec
mask).  The large-tile cGEMM once lost 6-10% on its digit-mapped forms that
way; no kernel may contain such a loop."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

WATERFALL = re.compile(r"v_readfirstlane_b32[^\n]*\n(?:[^\n]*\n){0,6}?[^\n]*v_cmp_eq_u64")


@pytest.fixture(scope="module")
def cgemm_asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("asm") / "qfx_cgemm.s"
    cmd = [HIPCC, "-O3", "--offload-arch=gfx950", "-std=c++17", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"),
           "--cuda-device-only", "-S", os.path.join(ROOT, "qflex_amd", "csrc", "qfx_cgemm.hip"),
           "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text()


def kernels(asm):
    for name in re.findall(r"^(_Z\w+):", asm, re.M):
        i = asm.index(name + ":")
        yield name, asm[i:asm.find(".Lfunc_end", i)]


def test_no_kernel_loops_over_buffer_resources(cgemm_asm):
    found = {name: len(WATERFALL.findall(body)) for name, body in kernels(cgemm_asm)}
    assert len(found) > 20
    assert not {n: c for n, c in found.items() if c}, "per-lane loops over buffer resources"


def test_the_check_sees_such_a_loop():
    # the pattern as the compiler emits it around one buffer load
    body = """v_readfirstlane_b32 s16, v170
	v_readfirstlane_b32 s17, v171
	v_readfirstlane_b32 s18, v166
	v_readfirstlane_b32 s19, v167
	v_cmp_eq_u64_e64 s[8:9], s[16:17], v[170:171]
	s_and_saveexec_b64 s[8:9], s[8:9]
	buffer_load_dwordx4 v[142:145], v146, s[16:19], 0 offen
"""
    assert WATERFALL.findall(body)


def test_large_tile_kernels_do_not_spill(cgemm_asm):
    # the split-FP16 and split-BF16 large tiles keep everything in registers,
    # one tile per block or chained (every chained instance, whatever its mode)
    checked = 0
    for name, _ in kernels(cgemm_asm):
        if "cgemm_split_chain_kernel" in name or any(
                f"cgemm_split_big_kernelILi{mode}" in name for mode in (1, 2)):
            m = re.search(r"\.set " + re.escape(name) + r"\.private_seg_size, (\d+)", cgemm_asm)
            assert m, name
            assert int(m.group(1)) == 0, name
            checked += 1
    assert checked >= 20
