"""Host code under the sanitizers (CPU only).

* AddressSanitizer + UndefinedBehaviorSanitizer: tests/sanitize_check.cpp
  drives the configuration parser (the reference's calibration.cfg, every
  prefix of it, broken files), the MAT-file writer and reader (a round trip,
  every truncation, flipped bytes; the reference's own data/*.mat sequences
  when present) and five frames of the FP64 oracle, built from the same
  sources as the library and the oracle.
* ThreadSanitizer: the device block cache (csrc/devmem.cpp) under eight
  threads allocating and releasing against a simulated device
  (tests/devmem_check.cpp), the check of its allocation floor and the
  allocation it admits being one critical section.

A sanitizer finding aborts the program, and the test fails with its report.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "scphd-camera-calibration_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")
REF_DATA = "/root/reference/data"


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)
    assert r.returncode == 0, f"{' '.join(cmd[:3])} ...: exit {r.returncode}\n{r.stdout[-3000:]}\n{r.stderr[-6000:]}"
    return r


def test_host_code_and_oracle_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_check")
    _run(["g++", "-O1", "-g", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-omit-frame-pointer",
          "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-I" + os.path.join(ROOT, "include"),
          "-I" + CSRC, os.path.join(HERE, "sanitize_check.cpp"), os.path.join(CSRC, "config.cpp"),
          os.path.join(CSRC, "matfile.cpp"), os.path.join(ORACLE, "scphd_oracle.cpp"), "-lz", "-o", exe])
    seqs = []
    if os.path.isdir(REF_DATA):
        seqs = [os.path.join(REF_DATA, f) for f in sorted(os.listdir(REF_DATA)) if f.endswith(".mat")]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = _run([exe, os.path.join(HERE, "golden", "calibration.cfg"), str(tmp_path)] + seqs, env=env)
    assert "0 failure(s)" in r.stdout


def test_device_block_cache_under_thread_sanitizer(tmp_path):
    exe = str(tmp_path / "devmem_tsan")
    main = tmp_path / "main.cpp"
    main.write_text('#include <cstdint>\n#include <cstdio>\n'
                    'extern "C" void devmem_stress(int, int, size_t, int, int64_t*);\n'
                    'int main() { int64_t o[5]; devmem_stress(8, 40, (size_t)64 << 30, 20, o);\n'
                    '  std::printf("violations %lld mallocs %lld\\n", (long long)o[0], (long long)o[1]);\n'
                    '  return o[0] == 0 && o[1] > 0 ? 0 : 1; }\n')
    _run(["g++", "-O1", "-g", "-std=c++17", "-pthread", "-fsanitize=thread", "-D__HIP_PLATFORM_AMD__",
          "-I/opt/rocm/include", "-I" + CSRC, str(main), os.path.join(HERE, "devmem_check.cpp"),
          os.path.join(CSRC, "devmem.cpp"), "-o", exe, "-L/opt/rocm/lib", "-lamdhip64",
          "-Wl,-rpath,/opt/rocm/lib"])
    r = _run([exe], env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    assert "violations 0" in r.stdout
