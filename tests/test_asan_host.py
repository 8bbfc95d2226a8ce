"""Host AddressSanitizer + UndefinedBehaviorSanitizer run of the library's
host code (graph_host.cpp: .edge parsing, CSR construction, COO; generators.cpp)
and of the oracle (louvain_oracle.c, every mode, gain, tie rule, chunk
emulation and refinement): tests/asan builds a g++/gcc driver with
-fsanitize=address,undefined and runs it; any sanitizer report fails it."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
DIR = os.path.join(HERE, "asan")


def test_host_code_under_asan_and_ubsan(fixture_path, tmp_path):
    subprocess.run(["make", "-s", "-j4", "-C", DIR], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    p = subprocess.run([os.path.join(DIR, "asan_driver"), fixture_path, str(tmp_path)],
                       capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failures" in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr
