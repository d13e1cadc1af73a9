"""Resource checks on the built gfx950 code objects (no GPU needed).

The kernels keep their state in registers and LDS: a kernel that uses
scratch memory pays for it in HBM traffic on every work item (a runtime-
indexed parameter struct once cost the batched solve 3.4x its output
bytes).  The code objects embedded in libintp_amd.so are unbundled into a
temporary directory and their kernel metadata read with llvm-readelf.
"""
import glob
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "bsplineinterpolation_amd", "libintp_amd.so")
LLVM = "/opt/rocm/llvm/bin"

# kernels allowed to spill, and their scratch bound (bytes per lane): the
# one-line 1D solve runs at a 128-VGPR bound for four waves per SIMD
ALLOWED_SCRATCH = {"batch_line_kernel": 128}
# evaluation kernels of 4 to 6 dimensions walk the outer dimensions of the
# stencil in a rolled loop and index their weights at run time: they keep
# the weights in scratch (at most this many bytes per lane)
ND_KERNELS = ("eval_general_kernel", "eval_uniform_kernel", "proxy_apply_kernel")
ND_SCRATCH = 2048


def _dimension(mangled):
    m = re.search(r"GLOBAL__N_1\d+[A-Za-z_]+?ILi(\d+)E", mangled)
    return int(m.group(1)) if m else 0


def _kernel_metadata():
    objdump, readelf = os.path.join(LLVM, "llvm-objdump"), os.path.join(LLVM, "llvm-readelf")
    if not (os.path.exists(LIB) and os.path.exists(objdump) and os.path.exists(readelf)):
        pytest.skip("library or LLVM tools not available")
    tmp = tempfile.mkdtemp(prefix="intp_co_")
    try:
        so = os.path.join(tmp, "lib.so")
        shutil.copy(LIB, so)
        subprocess.run([objdump, "--offloading", so], cwd=tmp, check=True, capture_output=True, timeout=120)
        kernels = {}
        for co in sorted(glob.glob(so + ".*gfx950")):
            notes = subprocess.run([readelf, "--notes", co], check=True, capture_output=True, text=True,
                                   timeout=120).stdout
            name = None
            for line in notes.splitlines():
                m = re.match(r"\s+(?:- )?\.name:\s+(\S+)", line)
                if m:
                    name = m.group(1)
                    kernels.setdefault(name, {})
                    continue
                m = re.match(r"\s+\.(private_segment_fixed_size|vgpr_spill_count|sgpr_spill_count):\s+(\d+)", line)
                if m and name is not None:
                    kernels[name][m.group(1)] = int(m.group(2))
        return kernels
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def _short(mangled):
    m = re.search(r"GLOBAL__N_1\d+([A-Za-z_]+?)(?:I|E|$)", mangled)
    return m.group(1) if m else mangled


def test_kernels_use_no_scratch():
    kernels = _kernel_metadata()
    # the library's hot kernels are all there
    names = {_short(k) for k in kernels}
    for want in ("solve_tile_kernel", "solve_kernel", "eval_patch_kernel", "bin_scatter_kernel",
                 "band_matvec_kernel", "band_matvec_rows_kernel", "batch_solve_kernel", "batch_line_split_kernel",
                 "scan_apply_kernel"):
        assert any(n.endswith(want) for n in names), want
    bad = []
    for k, md in kernels.items():
        scratch = md.get("private_segment_fixed_size", 0)
        short = _short(k)
        limit = next((v for a, v in ALLOWED_SCRATCH.items() if short.endswith(a)), 0)
        if short.endswith(ND_KERNELS) and _dimension(k) >= 4:
            limit = ND_SCRATCH
        if scratch > limit:
            bad.append((short, scratch))
    assert not bad, bad
