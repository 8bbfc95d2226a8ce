"""Register budget of the gfx950 kernels, read from the compiler's resource
report (hipcc -Rpass-analysis=kernel-resource-usage, device code only; no GPU
needed).  The best-move tables are bound by instruction issue and by how many
workgroups a CU holds: a kernel that spills registers to scratch or drops a
wave per SIMD loses measurable time, so these limits are part of the build.

- no kernel of the library spills vector registers or uses scratch memory;
- the persistent exact-gain tables of the degree 257-1024 bucket run at 8
  waves per SIMD, on unit weights and on weighted levels;
- the one-wave-per-vertex tables of the degree 65-256 bucket run at 8."""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "louvain-fast-move-cuda_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# the library's own device flags (Makefile HIPFLAGS)
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "--cuda-device-only", "-Rpass-analysis=kernel-resource-usage"]
SOURCES = ["lv_kernels.hip", "lv_aggregate.hip", "lv_sort_sweep.hip", "lv_server.hip",
           "lv_ingest.hip", "lv_noalloc.hip"]


def _report(src, tmp_path):
    out = tmp_path / (src + ".o")
    p = subprocess.run([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", str(out)],
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-4000:]
    kernels = {}
    for block in p.stderr.split("remark: Function Name: ")[1:]:
        name = block.split()[0]
        fields = dict(re.findall(r"remark:\s+([A-Za-z \[\]/]+): (\S+)", block))
        kernels[name] = fields
    assert kernels, "no resource report"
    return kernels


def _demangled(names):
    p = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True)
    return dict(zip(names, p.stdout.split("\n")))


@pytest.fixture(scope="module")
def reports(tmp_path_factory):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    tmp = tmp_path_factory.mktemp("res")
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        res = list(ex.map(lambda src: _report(src, tmp), SOURCES))
    return dict(zip(SOURCES, res))


def test_no_kernel_spills(reports):
    bad = []
    for src, kernels in reports.items():
        for name, f in kernels.items():
            # (scalar registers spill to vector lanes, not to memory: allowed)
            if int(f.get("VGPRs Spill", "0")) or int(f.get("ScratchSize [bytes/lane]", "0")):
                bad.append((src, name, f.get("VGPRs Spill"), f.get("ScratchSize [bytes/lane]")))
    assert not bad, bad


def _occupancy(reports, pattern):
    kernels = reports["lv_kernels.hip"]
    dm = _demangled(list(kernels))
    hits = {dm[n]: int(kernels[n]["Occupancy [waves/SIMD]"]) for n in kernels
            if re.search(pattern, dm[n])}
    assert hits, pattern
    return hits


def test_bucket6_tables_waves_per_simd(reports):
    unit = _occupancy(reports, r"k_move_table<256, 1540, 0, false, true>")
    weighted = _occupancy(reports, r"k_move_table<256, 1540, 0, true, true>")
    assert all(o == 8 for o in unit.values()), unit
    assert all(o == 8 for o in weighted.values()), weighted


def test_bucket5_tables_waves_per_simd(reports):
    occ = _occupancy(reports, r"k_move_table<64, 388, 0, (true|false), true>")
    assert all(o == 8 for o in occ.values()), occ
