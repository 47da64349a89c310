"""Registers, LDS and scratch memory of the kernels on the default paths, read
from the gfx950 code objects inside the built libqcbflow.so (CPU only).

The stage kernels are bound by the rate at which a CU serves L1 misses.
Scratch memory and register spills add global traffic of their own, and each
LDS stage kernel is written for an occupancy (FP64 three, FP32-mixed six
wavefronts per SIMD) that its registers and LDS must fit. A source change that
breaks either still compiles without a warning, so the compiler's kernel
metadata is checked here. The library file is only read.
"""
import os
import re
import shutil
import struct
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "ecmc-qcb_amd", "libqcbflow.so")
READELF = "/opt/rocm/llvm/bin/llvm-readelf"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
TARGET = b"hipv4-amdgcn-amd-amdhsa--gfx950"

# per SIMD: 512 VGPRs (allocated in blocks of 8); per CU: 160 KiB of LDS
VGPRS_PER_SIMD, LDS_PER_CU = 512, 160 * 1024


def _code_objects(blob):
    """The gfx950 code objects of the uncompressed offload bundles in blob
    (bundle: magic, u64 entry count, then per entry u64 offset, u64 size,
    u64 triple length and the triple; offsets from the bundle start)."""
    out, i = [], blob.find(MAGIC)
    while i != -1:
        p = i + len(MAGIC)
        (n,) = struct.unpack_from("<Q", blob, p)
        p += 8
        for _ in range(n):
            off, size, tlen = struct.unpack_from("<QQQ", blob, p)
            p += 24
            triple = blob[p:p + tlen]
            p += tlen
            if triple == TARGET and size:
                out.append(blob[i + off:i + off + size])
        i = blob.find(MAGIC, i + 1)
    return out


def _demangle(names):
    r = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True,
                       check=True)
    short = []
    for d in r.stdout.splitlines():
        d = re.sub(r"\(qcb::storage.*|\(double2 const.*|\(HIP_vector.*|\(double const.*", "", d)
        short.append(d.replace("(qcb::prec)0", "f64").replace("(qcb::prec)1", "f32m")
                     .replace("void qcb::", "").replace("qcb::", ""))
    return short


@pytest.fixture(scope="module")
def kernels():
    if not os.path.exists(LIB):
        pytest.skip("libqcbflow.so is not built")
    if not os.path.exists(READELF) or shutil.which("c++filt") is None:
        pytest.skip("llvm-readelf or c++filt missing")
    with open(LIB, "rb") as fh:
        blob = fh.read()
    cos = _code_objects(blob)
    assert cos, "no gfx950 code object in libqcbflow.so"
    entries = []
    with tempfile.TemporaryDirectory() as d:
        for k, co in enumerate(cos):
            path = os.path.join(d, f"co{k}.o")
            with open(path, "wb") as fh:
                fh.write(co)
            notes = subprocess.run([READELF, "--notes", path], capture_output=True, text=True,
                                   check=True).stdout
            md = notes[notes.index("amdhsa.kernels:"):]
            for e in re.split(r"\n  - ", md)[1:]:
                f = dict(re.findall(r"^    \.([a-z_]+):\s+(\S+)", e, re.M))
                if "name" in f:  # (amdhsa.version's list items have none)
                    entries.append(f)
    names = _demangle([e["name"] for e in entries])
    out = {}
    for n, e in zip(names, entries):
        out[n] = {k: int(e[k]) for k in ("vgpr_count", "vgpr_spill_count", "sgpr_spill_count",
                                         "private_segment_fixed_size",
                                         "group_segment_fixed_size")}
    return out


def _family(kernels, prefix):
    ks = {n: v for n, v in kernels.items() if n.startswith(prefix)}
    assert ks, f"no {prefix} kernel found"
    return ks


# the kernels a default flow, ensemble, measurement or heatbath launches, and
# the ensemble stage kernels of every launch shape (the persistent fused
# measurement, launched only by the autotuner's trials, spills 2 dwords)
DEFAULT_PATH = ["k_rk3_stage_lds<", "k_rk3_stage<", "k_rk3_stage1_measure<", "k_batch_stage_lds<",
                "k_batch_stage<",
                "k_batch_stage1_measure<f64, false>", "k_batch_stage1_measure<f32m, false>",
                "k_batch_measure<f64, false>", "k_batch_measure<f32m, false>",
                "k_measure_clover<", "k_plaquette<", "k_reduce_measure", "k_batch_reduce",
                "k_heatbath", "k_aos_to_field<", "k_copy"]


@pytest.mark.parametrize("prefix", DEFAULT_PATH)
def test_no_scratch_or_spills_on_default_paths(kernels, prefix):
    for name, r in _family(kernels, prefix).items():
        # (SGPR spills go to VGPR lanes, not to memory, and are allowed)
        assert r["private_segment_fixed_size"] == 0, f"{name} uses scratch memory: {r}"
        assert r["vgpr_spill_count"] == 0, f"{name} spills VGPRs: {r}"


@pytest.mark.parametrize("prec,waves", [("f64", 3), ("f32m", 6)])
def test_lds_stage_kernels_fit_their_occupancy(kernels, prec, waves):
    # 4 wavefronts per workgroup: `waves` per SIMD is `waves` workgroups per CU
    vgpr_cap = (VGPRS_PER_SIMD // waves) // 8 * 8
    for prefix in (f"k_rk3_stage_lds<{prec},", f"k_batch_stage_lds<{prec},"):
        for name, r in _family(kernels, prefix).items():
            assert r["vgpr_count"] <= vgpr_cap, f"{name}: {r['vgpr_count']} VGPRs > {vgpr_cap}"
            assert r["group_segment_fixed_size"] * waves <= LDS_PER_CU, \
                f"{name}: {r['group_segment_fixed_size']} B of LDS for {waves} workgroups per CU"
