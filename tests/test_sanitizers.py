"""The host code built with AddressSanitizer and UndefinedBehaviorSanitizer (g++, CPU only):
the library's topology set-up, grouped SpMV layout, GPS relabelling and scalar factorisations
(opendeformer_amd/csrc/od_setup.cpp, od_groups.cpp, od_relabel.cpp, od_precond_host.cpp) and the oracle
(oracle/odoracle.cpp), driven by tests/cpp/sanitize_check.cpp.  Any out-of-bounds access,
use after free, leak or undefined behaviour aborts the run."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    csrc = os.path.join(ROOT, "opendeformer_amd", "csrc")
    exe = str(tmp_path / "sanitize_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-ffp-contract=off", "-Wno-unknown-pragmas",
                    os.path.join(ROOT, "tests", "cpp", "sanitize_check.cpp"), os.path.join(csrc, "od_setup.cpp"),
                    os.path.join(csrc, "od_groups.cpp"), os.path.join(csrc, "od_relabel.cpp"), os.path.join(csrc, "od_precond_host.cpp"),
                    os.path.join(ROOT, "oracle", "odoracle.cpp"), "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitize_check: ok" in r.stdout
