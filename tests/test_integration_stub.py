"""The binding a PCAone maintainer adds (INTEGRATION.md's GpuRsvdOpData stub)
compiled against the reference's own headers (src/Halko.hpp, src/Data.hpp and
what they include) and this library's include/pcaone_gpu.h; every pcaone_*
symbol the object needs is exported by the built library.  Runs only where the
reference sources are present (the build container)."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")

pytestmark = pytest.mark.skipif(not (REF / "src" / "Halko.hpp").is_file() or shutil.which("g++") is None,
                                reason="reference sources or g++ not present")


def _stub() -> str:
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    stub = [b for b in blocks if "class GpuRsvdOpData" in b]
    assert len(stub) == 1
    return stub[0]


def test_stub_compiles_against_reference_headers(tmp_path):
    (tmp_path / "HalkoGpu.hpp").write_text(_stub())
    (tmp_path / "use.cpp").write_text(
        '#include "HalkoGpu.hpp"\n'
        "// one out-of-line use of every member so the object carries their code\n"
        "void use(GpuRsvdOpData& op, Mat2D& G, Mat2D& H) {\n"
        "  op.computeGandH(G, H, 0);\n  op.computeUSVGpu(3, 1e-4);\n"
        "  (void)op.rows(); (void)op.cols(); (void)op.ranks(); (void)op.oversamples();\n}\n"
        "GpuRsvdOpData* make(Data* d, const uint8_t* p) { return new GpuRsvdOpData(d, p, 10, 10, 2); }\n"
        "void drop(GpuRsvdOpData* op) { delete op; }\n")
    obj = tmp_path / "use.o"
    cmd = ["g++", "-std=c++17", "-O0", "-c", str(tmp_path / "use.cpp"), "-o", str(obj), f"-I{REF / 'src'}",
           f"-I{REF / 'external'}", f"-I{REF / 'external' / 'zstd' / 'lib'}", f"-I{ROOT / 'include'}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    und = subprocess.run(["nm", "-u", str(obj)], capture_output=True, text=True, check=True).stdout
    need = sorted(set(re.findall(r"\b(pcaone_\w+)", und)))
    assert "pcaone_gpu_rsvd_gandh" in need and "pcaone_gpu_rsvd_usv" in need
    lib = ROOT / "pcaone_amd" / "lib" / "libpcaone_amd.so"
    if not lib.exists():
        pytest.skip("library not built")
    dyn = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True, check=True).stdout
    have = set(re.findall(r"\bT (pcaone_\w+)", dyn))
    assert set(need) <= have, sorted(set(need) - have)
