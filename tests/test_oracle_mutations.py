"""The oracle's pins catch plausible misreadings of the shaders: each mutant
below is a copy of the oracle with one such mistake, built on the side, and
the CPU suite (less the oracle-written golden fixtures and this file) must
fail against it. The mutants: the occlusion test's screen V axis not flipped
(occlusion_cull.wgsl:101-104), LOD thresholds compared with > instead of >=,
the sub-pixel rule on P00 instead of P11 (indirect_dispatch.wgsl:127-137), and
the VG Hi-Z footprint from clip.w instead of the nearest view depth
(vg_cull.wgsl:184-195)."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "oracle" / "hv_oracle.cpp"

MUTANTS = {
    "unchanged": [],  # control: the oracle as it is, built the same way, passes
    "screen_v_not_flipped": [("vv = ndy * -0.5f + 0.5f", "vv = ndy * 0.5f + 0.5f")],
    "lod_strict_compare": [("(d >= v->lod_distances[t])", "(d > v->lod_distances[t])")],
    "subpixel_on_p00": [("std::fabs(r * v->p11 / clip[3])", "std::fabs(r * v->p00 / clip[3])")],
    "vg_footprint_from_clip_w": [("std::fabs(wr * cam[kProj] / nvd), std::fabs(wr * cam[kProj + 5] / nvd)",
                                  "std::fabs(wr * cam[kProj] / clip[3]), std::fabs(wr * cam[kProj + 5] / clip[3])")],
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_fails_the_cpu_suite(name, tmp_path):
    src = SRC.read_text()
    for old, new in MUTANTS[name]:
        assert old in src, f"{name}: pattern not found in the oracle"
        src = src.replace(old, new)
    (tmp_path / "hv_oracle.cpp").write_text(src)
    (tmp_path / "hv_oracle.h").write_text((ROOT / "oracle" / "hv_oracle.h").read_text())
    lib = tmp_path / "libhv_oracle.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-pthread", "-shared",
                    "-o", str(lib), str(tmp_path / "hv_oracle.cpp")], check=True, capture_output=True, timeout=300)
    env = dict(os.environ, HV_ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests"), "-m", "not gpu", "-x", "-q",
                        "-p", "no:cacheprovider", "--deselect", str(ROOT / "tests" / "test_golden.py"),
                        "--ignore", str(ROOT / "tests" / "test_oracle_mutations.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    if name == "unchanged":
        assert r.returncode == 0, r.stdout[-1500:]
        return
    assert r.returncode != 0, f"{name}: every CPU test passed against the mutant\n{r.stdout[-1500:]}"
    assert "failed" in r.stdout, r.stdout[-1500:]
