"""The spektri-gpu command-line host (spektri_amd/host/main.cpp): option
handling that fails before any GPU call (CPU only).  The options follow the
reference's clap definition (spektri/src/main.rs:15-90)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "spektri_amd", "host", "spektri-gpu")
# the same host under AddressSanitizer and UBSan (host code only)
BIN_SAN = os.path.join(ROOT, "spektri_amd", "host", "spektri-gpu-san")

pytestmark = pytest.mark.skipif(not os.path.exists(BIN), reason="host binary not built (python -m spektri_amd.build)")
BINS = [BIN] + ([BIN_SAN] if os.path.exists(BIN_SAN) else [])


def run(*args, binary=BIN):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    return subprocess.run([binary, *args], capture_output=True, text=True, timeout=60, env=env)


@pytest.mark.parametrize("binary", BINS)
def test_help_lists_the_reference_options(binary):
    r = run("--help", binary=binary)
    assert r.returncode == 0
    for opt in ("--samplerate", "--centerfreq", "--fftsize", "--fftbuf", "--averages", "--inputformat",
                "--spectrumformat", "--filters", "--zmqbind"):
        assert opt in r.stderr


@pytest.mark.parametrize("args,msg", [
    (("-s", "1e6"), "--inputformat is required"),
    (("--inputformat=s16le",), "--samplerate is required"),
    (("-s", "1e6", "-I", "s17le"), "invalid input format"),
    (("-s", "1e6", "-I", "S16LE", "--spectrumformat=u32"), "invalid spectrum format"),
    (("-s", "1e6", "-I", "cf32le", "--filters", "fs=1000"), "needs fs and fc"),
    (("-s", "1e6", "-I", "cf32le", "--filters", "fs"), "is not key=value"),
    (("-s", "abc", "-I", "u8"), "invalid value"),
    (("-x", "1"), "unknown option"),
])
@pytest.mark.parametrize("binary", BINS)
def test_invalid_options_exit_with_status_2(args, msg, binary):
    r = run(*args, binary=binary)
    assert r.returncode == 2 and msg in r.stderr, (r.returncode, r.stderr)
