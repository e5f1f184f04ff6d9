"""The spektri-gpu command-line host (spektri_amd/host/main.cpp): option
handling that fails before any GPU call (CPU only).  The options follow the
reference's clap definition (spektri/src/main.rs:15-90)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "spektri_amd", "host", "spektri-gpu")

pytestmark = pytest.mark.skipif(not os.path.exists(BIN), reason="host binary not built (python -m spektri_amd.build)")


def run(*args):
    return subprocess.run([BIN, *args], capture_output=True, text=True, timeout=30)


def test_help_lists_the_reference_options():
    r = run("--help")
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
def test_invalid_options_exit_with_status_2(args, msg):
    r = run(*args)
    assert r.returncode == 2 and msg in r.stderr, (r.returncode, r.stderr)
