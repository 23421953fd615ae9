"""Register and LDS budgets of the step kernels, read from the gfx950 compiler's own kernel metadata
(device-only assembly of the library's sources, built here with the library's flags).

The fast grid kernel runs six workgroups per CU on 15x15 tiles: it is built for at most 72 VGPRs
per lane (room for seven waves per SIMD), and six LDS allocations fit in 160 KiB, each rounded up
to the 512-byte granule the hardware allocates in (the runtime's occupancy query rounds finer and
reports seven for 23,232 B, where six ran concurrently).  No step kernel may spill
vector registers to scratch memory: a spill costs HBM traffic per launch (each wave's scratch lines
are written back at the end of the kernel).  Scalar registers may spill: they go to lanes of a
vector register, not to memory (in the fast kernel only on the bounds-checked tiles' obstacle
search)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "feriphys_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
LDS_PER_CU = 160 * 1024
LDS_GRANULE = 512
VGPRS_PER_SIMD_LANE = 512

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")


def kernel_metadata(source, tmp_path):
    """{mangled kernel name: {field: int}} from the amdhsa metadata of `source` built for gfx950."""
    out = str(tmp_path / (os.path.basename(source) + ".s"))
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "--cuda-device-only", "-S", "-O3", "-std=c++17",
                    "-ffp-contract=off", "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", os.path.join(CSRC, source), "-o", out], check=True, cwd=CSRC,
                   stderr=subprocess.DEVNULL)
    text = open(out).read()
    meta = text[text.index("amdhsa.kernels:"):]
    kernels = {}
    # each kernel's metadata block starts at "  - .agpr_count:" and holds ".name:" among its fields
    for block in re.split(r"\n  - ", meta)[1:]:
        fields = dict(re.findall(r"\.(\w+):\s+(\S+)", block))
        if "name" in fields:
            kernels[fields["name"]] = {k: int(v) for k, v in fields.items() if v.isdigit()}
    return kernels


@pytest.fixture(scope="module")
def fast_kernels(tmp_path_factory):
    k = kernel_metadata("grid_fast.hip", tmp_path_factory.mktemp("res"))
    return {n: m for n, m in k.items() if "grid_fast_kernel" in n}


@pytest.fixture(scope="module")
def tiled_kernels(tmp_path_factory):
    k = kernel_metadata("grid_tiled.hip", tmp_path_factory.mktemp("res"))
    return {n: m for n, m in k.items() if "grid_tiled_kernel" in n}


def test_fast_kernel_instantiations_present(fast_kernels):
    # two tile shapes x two modes (step, forces)
    assert len(fast_kernels) == 4, sorted(fast_kernels)


def test_fast_kernel_never_spills(fast_kernels):
    for name, m in fast_kernels.items():
        assert m["vgpr_spill_count"] == 0, name
        assert m["private_segment_fixed_size"] == 0, name


def test_fast_kernel_15x15_fits_six_workgroups_per_cu(fast_kernels):
    k1515 = {n: m for n, m in fast_kernels.items() if "Li15ELi15E" in n}
    assert len(k1515) == 2
    for name, m in k1515.items():
        waves = VGPRS_PER_SIMD_LANE // (-(-m["vgpr_count"] // 8) * 8)
        assert waves >= 7, (name, m["vgpr_count"])
        lds = -(-m["group_segment_fixed_size"] // LDS_GRANULE) * LDS_GRANULE
        assert 6 * lds <= LDS_PER_CU, (name, m["group_segment_fixed_size"])


def test_exact_tiled_kernel_never_spills(tiled_kernels):
    assert tiled_kernels
    for name, m in tiled_kernels.items():
        assert m["vgpr_spill_count"] == 0 and m["private_segment_fixed_size"] == 0, name
